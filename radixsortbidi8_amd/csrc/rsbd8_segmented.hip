// rsbd8_segmented.hip -- many independent sorts in one captured launch sequence.
//
// The reference sorts a batch by calling radixsortcopy once per array (Radixsortbidi8.hpp:60468);
// here every segment of a CSR batch is sorted with the same mode/direction flags. Segments up to
// the LDS capacity are items of their own; longer ones are partitioned by their top digit first:
//   k_item_wave     items of at most 64 keys, one wave each: every lane ranks its key by counting
//                   the keys that precede it in the stable order;
//   k_item_sort     persistent blocks, one item (a key range sorted on its own) at a time, all in
//                   LDS: partition by (key - min) >> sh with LDS atomics, then each key's exact
//                   position in its partition by counting the partition's keys that precede it in
//                   the stable order (smaller key, or equal key read earlier);
//   k_item_lsd      items with a partition too long for that: stable LDS LSD passes;
//   k_seg_count1    per-tile counts of the top digit over the tiles of the long segments;
//   k_seg_scan1     one wave per long segment: bucket sizes and starts, tile offsets, and the
//                   segment's buckets packed into items of at most the LDS capacity (consecutive
//                   buckets of one segment may share an item: they are already in top-digit
//                   order, so a sort of the item by the whole key is exact); a bucket above the
//                   capacity sends the segment to the LSD passes below instead;
//   k_seg_scatter1  the partition of the long segments by the top digit;
//   k_seg_hist, k_seg_plan, k_seg_pass, k_seg_copy
//                   for those segments only: per-segment digit histograms, plan, onesweep passes
//                   with the look-back confined to the segment, and the copy when every digit is
//                   uniform;
//   k_seg_finish    resets the counters and advances the look-back epoch.
// When the key transform is injective (equal sort keys are equal keys) the order of ties cannot be
// observed, and the partition and the item sort use cheaper forms that do not keep it.
// A plan object owns the device tables (segment classes, tile map, items) and its workspace, and
// can capture the whole sequence (short segments on a side stream in parallel with the long-segment
// chain) into one hipGraph that is replayed with a single launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "rsbd8_engine.hpp"
#include "rsbd8_segmented.hpp"

namespace rsbd8hip {

namespace {

struct SegCtrl {
	uint32_t n_ftiles;   // tiles of the segments flagged for the LSD passes
	uint32_t tile_counter[kMaxPasses];
	uint32_t epoch;
	uint32_t q_small;    // queued short segments (long runs)
	uint32_t q_sub;      // queued items of long segments (long runs)
	uint32_t n_sub;      // items made from the buckets of the long segments
	uint32_t n_big;      // buckets above the item capacity, up to the largest tier: items of their own
	uint32_t q_big;      // queued big items (long runs)
	uint32_t pad[1];
};

// per long segment: sel holds, for pass p, bits 4p..4p+1 = source (0 input, 1 output, 2 buffer) and
// bits 4p+2..4p+3 = destination (all 32 bits for 8-byte keys); mask holds the executed passes in
// bits 0..7 and the two flags below
struct SegSel {
	uint32_t sel;
	uint32_t mask;
};
// one tile of a long segment: keys [y * tile, (y + 1) * tile) of segment L, which starts at sb and
// holds len keys (logical order; a physical range read back to front under reverse)
struct __align__(16) SegTile {
	uint64_t sb;
	uint32_t len;
	uint32_t L;
	uint32_t y;
	uint32_t pad[3];
};

constexpr uint32_t kSelCopy = 1u << 30;     // no pass executes: the segment is copied
constexpr uint32_t kSelReverse = 1u << 31;  // the first executed pass (or the copy) reads back to front

template <typename K>
__device__ __forceinline__ K *pick(uint32_t code, const K *in, K *out, K *buf)
{
	return code == 0 ? const_cast<K *>(in) : (code == 1 ? out : buf);
}

// ---------------------------------------------------------------------------------------------

// Words of the item sort. 2-byte keys pack (transformed key << 16) | position into one 32-bit word,
// whose order is then the stable order; 4- and 8-byte keys keep the transformed key as the word and
// the position in a separate 16-bit array. (1-byte keys go straight to k_item_lsd: one LDS pass
// ranks them by the whole key.)
template <typename K> struct ItemWord;
template <> struct ItemWord<uint16_t> { using W = uint32_t; static constexpr bool packed = true; static constexpr int kbits = 16; };
template <> struct ItemWord<uint32_t> { using W = uint32_t; static constexpr bool packed = false; static constexpr int kbits = 32; };
template <> struct ItemWord<uint64_t> { using W = uint64_t; static constexpr bool packed = false; static constexpr int kbits = 64; };

constexpr uint32_t kItemRunMax = 48;

// transformed key in the item word's type, bits above the key width cleared
template <typename K>
__device__ __forceinline__ typename ItemWord<K>::W item_key(K raw, const KeyXform &xf)
{
	using W = typename ItemWord<K>::W;
	constexpr int KB = ItemWord<K>::kbits;
	const W k = static_cast<W>(sort_key<K>(raw, xf));
	if constexpr (KB < 8 * static_cast<int>(sizeof(W))) return k & ((W(1) << KB) - 1u);
	else return k;
}

// Items i < *count (or count_value when count is null): keys [start, start + len) of src, read back
// to front under reverse, sorted into the same range of dst. Every block takes the items
// blockIdx.x, blockIdx.x + gridDim.x, ... and loads the keys of its next item while it sorts the
// current one (the loads are issued before the current item's stores, so waiting for them never
// waits for those). The keys of an item are partitioned with LDS atomics by (key - min) >> sh, sh
// chosen so that the item's key range spans the PB partition bits; each word's final position is
// its partition's start plus the words of the partition that precede it in the stable order
// (smaller key, or equal key read earlier), counted with independent LDS reads; the raw keys are
// staged in LDS at their final positions. An item whose keys are all equal is copied in order.
// Items with a partition of more than kItemRunMax keys are queued for k_item_lsd. With INJ (the
// transform only complements bits, so equal sort keys are equal keys and a key is its sort key xor
// E) the order among equal keys is unobservable: ties are broken by the index in the partitioned
// array, the ranks go to LDS as the inverse permutation (position -> word) and the output is
// gathered from the words in order, which needs no second word array.
template <typename K, int THREADS, int KPT, int PB, bool INJ>
__global__ __launch_bounds__(THREADS) void k_item_sort(const K *__restrict__ src, K *__restrict__ dst,
	const uint64_t *__restrict__ istart, const uint32_t *__restrict__ ilen, const uint32_t *count, uint32_t count_value,
	KeyXform xf, int reverse, uint32_t *__restrict__ queue, uint32_t *qcount, uint32_t qbase)
{
	constexpr int CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	using IW = ItemWord<K>;
	using W = typename IW::W;
	constexpr int KB = IW::kbits;
	constexpr int PBE = PB < KB ? PB : KB;
	constexpr uint32_t NB = 1u << PBE;
	constexpr uint32_t PER = (NB + THREADS - 1) / THREADS;
	static_assert(CAP <= 65536, "positions are held in 16 bits");
	static_assert(sizeof(K) <= sizeof(W), "the raw keys are staged in the word array");
	__shared__ __align__(16) W s_w[CAP];  // words, then (without INJ) the raw keys in final order
	__shared__ uint16_t s_l[(INJ || !IW::packed) ? CAP : 1];  // positions of unpacked words / the inverse permutation
	// partition counters, then cursors (after the scatter: ends): two 16-bit halves per word when
	// every thread of the scan owns whole words (counts and positions stay below CAP <= 65536)
	constexpr bool PACK = PER % 2 == 0;
	constexpr uint32_t NBW = PACK ? NB / 2 : NB;
	__shared__ uint32_t s_cur[NBW];
	auto cur_word = [](uint32_t b) { return PACK ? b >> 1 : b; };
	auto cur_shift = [](uint32_t b) { return PACK ? (b & 1u) * 16u : 0u; };
	auto cur_get = [&](uint32_t b) { return PACK ? (s_cur[b >> 1] >> ((b & 1u) * 16u)) & 0xFFFFu : s_cur[b]; };
	__shared__ uint32_t s_wsum[NW];
	__shared__ W s_min[NW], s_max[NW];
	__shared__ uint32_t s_big;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t nitems = count ? *count : count_value;
	K key[KPT];
	uint32_t len = 0;
	uint64_t start = 0;
	auto fetch = [&](uint32_t item, K (&kk)[KPT], uint32_t &ln, uint64_t &st) {
		ln = 0;
		st = 0;
		if (item < nitems) {
			ln = ilen[item];
			st = istart[item];
		}
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			kk[j] = l < ln ? src[st + (reverse ? ln - 1 - l : l)] : K(0);
		}
	};
	auto sort_item = [&](uint32_t it) {
		if (len == 0) return;
		if (len == 1) {
			if (t == 0) dst[start] = key[0];
			return;
		}
		W tk[KPT];
		W mn = ~W(0), mx = 0;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			tk[j] = item_key<K>(key[j], xf);
			if (l < len) {
				mn = min(mn, tk[j]);
				mx = max(mx, tk[j]);
			}
		}
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) {
			mn = min(mn, static_cast<W>(__shfl_xor(mn, o, 64)));
			mx = max(mx, static_cast<W>(__shfl_xor(mx, o, 64)));
		}
		for (uint32_t i = t; i < NBW; i += THREADS) s_cur[i] = 0;
		if (t == 0) s_big = 0;
		if (lane == 0) { s_min[w] = mn; s_max[w] = mx; }
		__syncthreads();
		mn = s_min[0];
		mx = s_max[0];
#pragma unroll
		for (int v = 1; v < NW; ++v) { mn = min(mn, s_min[v]); mx = max(mx, s_max[v]); }
		const W diff = mx - mn;
		if (diff == 0) {
			// every key is equal: the stable order is the order of reading
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				if (l < len) dst[start + l] = key[j];
			}
			return;
		}
		// highest bit of the key range; the partition digit is (key - min) >> sh, below NB
		int hb;
		if constexpr (sizeof(W) == 8) hb = 63 - __clzll(static_cast<long long>(diff));
		else hb = 31 - __clz(static_cast<int>(diff));
		const uint32_t sh = hb + 1 >= PBE ? static_cast<uint32_t>(hb + 1 - PBE) : 0u;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) {
				const uint32_t b = static_cast<uint32_t>((tk[j] - mn) >> sh);
				atomicAdd(&s_cur[cur_word(b)], 1u << cur_shift(b));
			}
		}
		__syncthreads();
		uint32_t c[PER], sum = 0, big = 0;
#pragma unroll
		for (uint32_t k = 0; k < PER; ++k) {
			const uint32_t b = t * PER + k;
			c[k] = b < NB ? cur_get(b) : 0u;
			sum += c[k];
			big |= c[k] > kItemRunMax;
		}
		const uint32_t incl = wave_inclusive_sum(sum);
		if (lane == 63) s_wsum[w] = incl;
		if (big) s_big = 1;
		__syncthreads();
		uint32_t run = incl - sum;
		for (uint32_t v = 0; v < w; ++v) run += s_wsum[v];
		if constexpr (PACK) {
#pragma unroll
			for (uint32_t k = 0; k < PER; k += 2) {
				const uint32_t b = t * PER + k;
				const uint32_t r0 = run, r1 = run + c[k];
				run = r1 + c[k + 1];
				if (b < NB) s_cur[b >> 1] = r0 | (r1 << 16);
			}
		} else {
#pragma unroll
			for (uint32_t k = 0; k < PER; ++k) {
				const uint32_t b = t * PER + k;
				if (b < NB) s_cur[b] = run;
				run += c[k];
			}
		}
		__syncthreads();
		if (s_big) {
			if (t == 0) queue[atomicAdd(qcount, 1u)] = qbase + it;  // an index into the whole item table
			return;
		}
		if constexpr (INJ) {
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				if (l < len) {
					const uint32_t b = static_cast<uint32_t>((tk[j] - mn) >> sh);
					s_w[(atomicAdd(&s_cur[cur_word(b)], 1u << cur_shift(b)) >> cur_shift(b)) & 0xFFFFu] = tk[j];
				}
			}
			__syncthreads();
			for (uint32_t i = t; i < len; i += THREADS) {
				const W x = s_w[i];
				const uint32_t d = static_cast<uint32_t>((x - mn) >> sh);
				const uint32_t lo = d ? cur_get(d - 1) : 0u, hi = cur_get(d);
				// ties go by the index in the partitioned array: the words before i count when not
				// larger, those after it when smaller
				uint32_t n_less = 0;
				for (uint32_t q = lo; q < i; ++q) n_less += static_cast<uint32_t>(s_w[q] <= x);
				for (uint32_t q = i + 1; q < hi; ++q) n_less += static_cast<uint32_t>(s_w[q] < x);
				s_l[lo + n_less] = static_cast<uint16_t>(i);
			}
			__syncthreads();
			const W e = static_cast<W>(xf.E);
			for (uint32_t i = t; i < len; i += THREADS) dst[start + i] = static_cast<K>(s_w[s_l[i]] ^ e);
			return;
		}
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) {
				const uint32_t b = static_cast<uint32_t>((tk[j] - mn) >> sh);
				const uint32_t p = (atomicAdd(&s_cur[cur_word(b)], 1u << cur_shift(b)) >> cur_shift(b)) & 0xFFFFu;
				if constexpr (IW::packed) {
					s_w[p] = (tk[j] << 16) | static_cast<W>(l);
				} else {
					s_w[p] = tk[j];
					s_l[p] = static_cast<uint16_t>(l);
				}
			}
		}
		__syncthreads();
		uint32_t fin[KPT], at[KPT];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t i = j * THREADS + t;
			at[j] = 0xFFFFFFFFu;
			if (i < len) {
				const W x = s_w[i];
				W kx;
				uint32_t lx;
				if constexpr (IW::packed) {
					kx = x >> 16;
					lx = static_cast<uint32_t>(x) & 0xFFFFu;
				} else {
					kx = x;
					lx = s_l[i];
				}
				// the cursors have advanced to the ends: partition d holds words [end of d - 1, end of d)
				const uint32_t d = static_cast<uint32_t>((kx - mn) >> sh);
				const uint32_t lo = d ? cur_get(d - 1) : 0u, hi = cur_get(d);
				uint32_t n_less = 0;
				if constexpr (IW::packed) {
					for (uint32_t q = lo; q < hi; ++q) n_less += static_cast<uint32_t>(s_w[q] < x);
				} else {
					for (uint32_t q = lo; q < hi; ++q) {
						const W y = s_w[q];
						n_less += static_cast<uint32_t>(y < x) | (static_cast<uint32_t>(y == x) & static_cast<uint32_t>(s_l[q] < lx));
					}
				}
				fin[j] = lo + n_less;
				at[j] = lx;
			}
		}
		__syncthreads();  // the positions are read: s_l holds the final positions from here
#pragma unroll
		for (int j = 0; j < KPT; ++j)
			if (at[j] != 0xFFFFFFFFu) s_l[at[j]] = static_cast<uint16_t>(fin[j]);
		__syncthreads();
		K *s_raw = reinterpret_cast<K *>(s_w);
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) s_raw[s_l[l]] = key[j];
		}
		__syncthreads();
		for (uint32_t i = t; i < len; i += THREADS) dst[start + i] = s_raw[i];
	};
	fetch(blockIdx.x, key, len, start);
	for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
		K nkey[KPT];
		uint32_t nlen;
		uint64_t nstart;
		fetch(it + gridDim.x, nkey, nlen, nstart);
		__syncthreads();  // the previous item's last reads of the shared arrays are done
		sort_item(it);
#pragma unroll
		for (int j = 0; j < KPT; ++j) key[j] = nkey[j];
		len = nlen;
		start = nstart;
	}
}

// Items of at most 64 keys, one wave each (no LDS, no block barrier): lane l holds the key read
// at position l and counts the keys that precede it in the stable order.
template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_item_wave(const K *__restrict__ src, K *__restrict__ dst,
	const uint64_t *__restrict__ istart, const uint32_t *__restrict__ ilen, uint32_t nitems, KeyXform xf, int reverse)
{
	using T = typename KeyTraits<K>::W;
	constexpr int KB = KeyTraits<K>::bits;
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t waves = gridDim.x * (THREADS / 64);
	for (uint32_t it = blockIdx.x * (THREADS / 64) + (threadIdx.x >> 6); it < nitems; it += waves) {
		const uint32_t len = ilen[it];
		const uint64_t start = istart[it];
		K key = K(0);
		T k = 0;
		if (lane < len) {
			key = src[start + (reverse ? len - 1 - lane : lane)];
			k = sort_key<K>(key, xf);
			if constexpr (KB < 8 * static_cast<int>(sizeof(T))) k &= (T(1) << KB) - 1u;
		}
		uint32_t rank = 0;
		for (uint32_t j = 0; j < len; ++j) {
			const T y = __shfl(k, static_cast<int>(j), 64);
			rank += static_cast<uint32_t>(y < k) | (static_cast<uint32_t>(y == k) & static_cast<uint32_t>(j < lane));
		}
		if (lane < len) dst[start + rank] = key;
	}
}

// The queued items (queue[i] for i < *qcount), or every item (i < *qcount, or qvalue when qcount is
// null) when queue is null: stable LDS LSD passes (the block ranking) over every digit.
template <typename K, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_item_lsd(const K *__restrict__ src, K *__restrict__ dst,
	const uint64_t *__restrict__ istart, const uint32_t *__restrict__ ilen, const uint32_t *__restrict__ queue,
	const uint32_t *qcount, uint32_t qvalue, KeyXform xf, int reverse)
{
	constexpr int CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_wsum[4];
	__shared__ __align__(16) unsigned char s_raw[max_bytes(CAP * sizeof(K), rank_peer_bytes(THREADS))];
	K *s_keys = reinterpret_cast<K *>(s_raw);
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_raw);
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t queued = qcount ? *qcount : qvalue;
	for (uint32_t qi = blockIdx.x; qi < queued; qi += gridDim.x) {
		const uint32_t it = queue ? queue[qi] : qi;
		const uint64_t b = istart[it];
		const uint32_t len = ilen[it];
		K keys[KPT];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			keys[j] = l < len ? src[b + (reverse ? len - 1 - l : l)] : K(0);
		}
		for (int p = 0; p < P; ++p) {
			uint32_t dig[KPT], pos[KPT], total;
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				dig[j] = l < len ? digit_of<K>(keys[j], xf, 8u * p) : 0xFFu;
			}
			block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_keys[pos[j]] = keys[j];
			__syncthreads();
#pragma unroll
			for (int j = 0; j < KPT; ++j) keys[j] = s_keys[w * 64u * KPT + j * 64u + lane];
			__syncthreads();
		}
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) dst[b + l] = keys[j];
		}
	}
}

// ---------------------------------------------------------------------------------------------
// long segments: one stable partition by the top digit

// per-tile counts of the top digit, one block per tile
template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_seg_count1(const K *__restrict__ in, const SegTile *__restrict__ tiles,
	uint32_t tile_keys, KeyXform xf, int reverse, uint32_t *__restrict__ counts)
{
	constexpr int P = sizeof(K);
	constexpr int C = count_copies(1);
	__shared__ uint32_t s_t[256 * C];
	const uint32_t g = blockIdx.x;
	for (int i = threadIdx.x; i < 256 * C; i += THREADS) s_t[i] = 0;
	const SegTile tl = tiles[g];
	const uint32_t lo = tl.y * tile_keys, hi = min(tl.len, lo + tile_keys);
	__syncthreads();
	for_each_key<K, THREADS>(in, tl.sb + (reverse ? tl.len - hi : lo), hi - lo,
		[&](K x) { wave_count<C>(s_t, digit_of<K>(x, xf, 8u * (P - 1))); });
	__syncthreads();
	for (int i = threadIdx.x; i < 256; i += THREADS) counts[static_cast<uint64_t>(g) * 256 + i] = copies_sum<C>(s_t, i);
}

// One wave per long segment (sixteen segments per block), four consecutive top digits per lane:
// bucket sizes and starts, the segment's tile counts turned into output offsets, and its buckets
// packed greedily into items of at most cap keys. A bucket above cap but within bigcap (the
// largest LDS tier) is an item of its own in the big-item table, sorted by that tier; a bucket
// above bigcap flags the segment for the LSD passes (and skips everything else). The items of a
// block's segments take their slots with one atomic per table and block: one per segment
// serialised on the item counter. Measured on the 14,278 long segments of 65,536 log-uniform
// arrays: 175 us with a block and an atomic per segment, 59 us with four segments per block, 37 us
// with sixteen.
constexpr uint32_t kScan1Segs = 16;  // segments (waves) per block
__global__ __launch_bounds__(64 * kScan1Segs) void k_seg_scan1(const uint64_t *__restrict__ seg,
	const uint32_t *__restrict__ large_ids, uint32_t nlarge, const uint32_t *__restrict__ tile_first,
	uint32_t *__restrict__ counts, uint32_t cap, uint32_t bigcap, uint32_t *__restrict__ flag, uint32_t *__restrict__ ftiles,
	SegCtrl *ctrl, uint64_t *__restrict__ istart, uint32_t *__restrict__ ilen, uint64_t *__restrict__ bstart,
	uint32_t *__restrict__ blen)
{
	__shared__ uint32_t s_pre[kScan1Segs][257];
	__shared__ uint16_t s_next[kScan1Segs][256];
	__shared__ uint32_t s_items[kScan1Segs], s_bigs[kScan1Segs];
	const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	const uint32_t L = blockIdx.x * kScan1Segs + w;
	const bool have = L < nlarge;  // wave-uniform
	uint32_t *pre = s_pre[w];
	uint16_t *next = s_next[w];
	bool scan = false;  // this wave's segment is partitioned (not flagged for the LSD passes)
	uint64_t sb = 0;
	if (have) {
		const uint32_t s = large_ids[L];
		sb = seg[s];
		const uint32_t t0 = tile_first[L], t1 = tile_first[L + 1];
		// digits 4 * lane .. 4 * lane + 3 of every tile row (rows of 256 words, 16-byte aligned)
		uint4 tot = make_uint4(0u, 0u, 0u, 0u);
		for (uint32_t tl = t0; tl < t1; ++tl) {
			const uint4 c = reinterpret_cast<const uint4 *>(counts + static_cast<uint64_t>(tl) * 256)[lane];
			tot.x += c.x; tot.y += c.y; tot.z += c.z; tot.w += c.w;
		}
		if (__ballot(tot.x > bigcap || tot.y > bigcap || tot.z > bigcap || tot.w > bigcap)) {
			// the segment's tiles join the list of the LSD passes as one run in tile order, so a
			// tile's look-back only meets tiles of its segment that were queued before it
			uint32_t at = 0;
			if (lane == 0) {
				flag[L] = 1;
				at = atomicAdd(&ctrl->n_ftiles, t1 - t0);
			}
			at = __shfl(at, 0, 64);
			for (uint32_t i = lane; i < t1 - t0; i += 64) ftiles[at + i] = t0 + i;
		} else {
			scan = true;
			const uint32_t sum = tot.x + tot.y + tot.z + tot.w;
			const uint32_t incl = wave_inclusive_sum(sum);
			uint4 run;  // bucket starts, relative to the segment start
			run.x = incl - sum;
			run.y = run.x + tot.x;
			run.z = run.y + tot.y;
			run.w = run.z + tot.z;
			pre[4 * lane] = run.x;
			pre[4 * lane + 1] = run.y;
			pre[4 * lane + 2] = run.z;
			pre[4 * lane + 3] = run.w;
			if (lane == 63) pre[256] = incl;
			for (uint32_t tl = t0; tl < t1; ++tl) {
				uint4 *row = reinterpret_cast<uint4 *>(counts + static_cast<uint64_t>(tl) * 256) + lane;
				const uint4 c = *row;
				*row = run;
				run.x += c.x; run.y += c.y; run.z += c.z; run.w += c.w;
			}
		}
	}
	__syncthreads();
	// The buckets are packed greedily, in digit order, into items of at most cap keys. An item that
	// starts at bucket b ends before the first bucket e past the largest prefix within cap:
	// next[b] = max{e : pre[e] - pre[b] <= cap}. A bucket above cap (at most bigcap here) is an item
	// of its own: next[b] = b + 1; the greedy packing of its predecessors stops before it.
	if (scan) {
#pragma unroll
		for (uint32_t i = 0; i < 4; ++i) {
			const uint32_t b = 4 * lane + i;
			const uint32_t base = pre[b];
			uint32_t e = b;
#pragma unroll
			for (uint32_t step = 256; step; step >>= 1)
				if (e + step <= 256u && pre[e + step] - base <= cap) e += step;
			next[b] = static_cast<uint16_t>(e > b ? e : b + 1u);
		}
	}
	__syncthreads();
	if (lane == 0) {
		uint32_t nitems = 0, nbig = 0;
		if (scan) {
			const uint32_t total = pre[256];
			for (uint32_t b = 0; b < 256u && pre[b] < total; b = next[b]) {
				const bool big = pre[b + 1] - pre[b] > cap;
				nbig += big ? 1u : 0u;
				nitems += big ? 0u : 1u;
			}
		}
		s_items[w] = nitems;
		s_bigs[w] = nbig;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t sum = 0, bsum = 0;
		for (uint32_t v = 0; v < kScan1Segs; ++v) {
			sum += s_items[v];
			bsum += s_bigs[v];
		}
		uint32_t at = sum ? atomicAdd(&ctrl->n_sub, sum) : 0u;
		uint32_t bat = bsum ? atomicAdd(&ctrl->n_big, bsum) : 0u;
		for (uint32_t v = 0; v < kScan1Segs; ++v) {
			const uint32_t c = s_items[v], bc = s_bigs[v];
			s_items[v] = at;
			s_bigs[v] = bat;
			at += c;
			bat += bc;
		}
	}
	__syncthreads();
	if (lane != 0 || !scan) return;
	const uint32_t total = pre[256];
	uint32_t it = s_items[w], bit = s_bigs[w];
	for (uint32_t b = 0; b < 256u && pre[b] < total; b = next[b]) {
		const uint32_t len = (next[b] < 256u ? pre[next[b]] : total) - pre[b];
		if (len > cap) {
			bstart[bit] = sb + pre[b];
			blen[bit] = len;
			++bit;
		} else {
			istart[it] = sb + pre[b];
			ilen[it] = len;
			++it;
		}
	}
}

// The partition of the long segments by the top digit, into the same ranges of dst. Persistent
// blocks; blocks are dealt to the XCDs round robin, so block b works through the tile range of XCD
// b % 8 (neighbouring tiles, whose digit runs are neighbours in dst, are written through one L2),
// loading its next tile before it writes the current one. With STABLE the keys of a digit keep
// their order (block ranking); without, a key's place within its digit's run of the tile comes from
// an LDS atomic, which is exact only up to the order of equal sort keys. That is enough when the key
// transform is injective (equal sort keys are equal keys), because the item sort that follows
// orders every bucket by the whole sort key.
template <typename K, int THREADS, int KPT, bool STABLE>
__global__ __launch_bounds__(THREADS) void k_seg_scatter1(const K *__restrict__ in, K *__restrict__ dst,
	const SegTile *__restrict__ tiles, uint32_t ntiles, const uint32_t *__restrict__ offsets,
	const uint32_t *__restrict__ flag, KeyXform xf, int reverse)
{
	constexpr int TILE = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	constexpr uint32_t kShift = 8u * (P - 1);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_glob[256];
	__shared__ uint32_t s_wsum[4];
	__shared__ __align__(16) unsigned char s_raw[STABLE ? max_bytes(TILE * sizeof(K), rank_peer_bytes(THREADS)) : TILE * sizeof(K)];
	K *s_keys = reinterpret_cast<K *>(s_raw);
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_raw);
	(void)s_peer;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	(void)lane;
	const uint32_t per = (ntiles + 7) / 8;
	const uint32_t x = blockIdx.x & 7u, step = gridDim.x >> 3;
	uint32_t k = blockIdx.x >> 3;
	uint32_t g = x * per + k;
	bool have = k < per && g < ntiles;
	SegTile tl{};
	uint32_t valid = 0, skip = 0;
	K keys[KPT];
	auto fetch = [&](uint32_t gg) {
		tl = tiles[gg];
		const uint32_t base = tl.y * TILE;
		valid = min(static_cast<uint32_t>(TILE), tl.len - base);
		skip = flag[tl.L];  // block-uniform: the segment takes the LSD passes
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			const uint32_t r = base + l;
			keys[j] = l < valid ? in[tl.sb + (reverse ? tl.len - 1 - r : r)] : K(0);
		}
	};
	if (have) fetch(g);
	while (have) {
		const uint64_t sb = tl.sb;
		const uint32_t cur = g, cvalid = valid, cskip = skip;
		if (!cskip) {
			uint32_t dig[KPT];
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				dig[j] = l < cvalid ? digit_of<K>(keys[j], xf, kShift) : 0xFFu;
			}
			if constexpr (STABLE) {
				uint32_t pos[KPT], total;
				block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
				if (t < 256) s_glob[t] = offsets[static_cast<uint64_t>(cur) * 256 + t] - s_digoff[t];
#pragma unroll
				for (int j = 0; j < KPT; ++j) s_keys[pos[j]] = keys[j];
			} else {
				uint32_t *s_cnt = s_whist;  // 256 counters, then cursors
				for (uint32_t i = t; i < 256; i += THREADS) s_cnt[i] = 0;
				__syncthreads();
#pragma unroll
				for (int j = 0; j < KPT; ++j) {
					const uint32_t l = w * 64u * KPT + j * 64u + lane;
					if (l < cvalid) wave_count(s_cnt, dig[j]);
				}
				__syncthreads();
				if (t < 256) {
					const uint32_t c = s_cnt[t];
					const uint32_t incl = wave_inclusive_sum(c);
					if ((t & 63u) == 63u) s_wsum[t >> 6] = incl;
					s_digoff[t] = incl - c;
				}
				__syncthreads();
				if (t < 256) {
					uint32_t before = 0;
					for (uint32_t v = 0; v < (t >> 6); ++v) before += s_wsum[v];
					const uint32_t start = s_digoff[t] + before;
					s_cnt[t] = start;
					s_glob[t] = offsets[static_cast<uint64_t>(cur) * 256 + t] - start;
				}
				__syncthreads();
#pragma unroll
				for (int j = 0; j < KPT; ++j) {
					const uint32_t l = w * 64u * KPT + j * 64u + lane;
					if (l < cvalid) s_keys[atomicAdd(&s_cnt[dig[j]], 1u)] = keys[j];
				}
			}
		}
		k += step;
		g = x * per + k;
		have = k < per && g < ntiles;
		if (have) fetch(g);  // the current keys are staged in LDS
		__syncthreads();
		if (!cskip) {
			K *dseg = dst + sb;
			for (uint32_t i = t; i < cvalid; i += THREADS) {
				const K key = s_keys[i];
				dseg[s_glob[digit_of<K>(key, xf, kShift)] + i] = key;
			}
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------------
// long segments with a bucket above the LDS capacity: LSD passes

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_seg_hist(const K *__restrict__ in, const SegTile *__restrict__ tiles,
	const uint32_t *__restrict__ ftiles, const SegCtrl *ctrl, uint32_t tile_keys, KeyXform xf, uint32_t *__restrict__ hist)
{
	constexpr int P = sizeof(K);
	constexpr int C = count_copies(P);
	__shared__ uint32_t s_h[P * 256 * C];
	const uint32_t nf = ctrl->n_ftiles;  // tiles of the segments flagged by k_seg_scan1
	for (uint32_t q = blockIdx.x; q < nf; q += gridDim.x) {
		const SegTile tl = tiles[ftiles[q]];
		for (int i = threadIdx.x; i < P * 256 * C; i += THREADS) s_h[i] = 0;
		__syncthreads();
		const uint32_t lo = tl.y * tile_keys, hi = min(tl.len, lo + tile_keys);
		for_each_key<K, THREADS>(in, tl.sb + lo, hi - lo, [&](K x) {
			const auto k = sort_key<K>(x, xf);
#pragma unroll
			for (int p = 0; p < P; ++p) wave_count<C>(s_h + p * 256 * C, static_cast<uint32_t>(k >> (8 * p)) & 0xFFu);
		});
		__syncthreads();
		uint32_t *h = hist + static_cast<uint64_t>(tl.L) * P * 256;
		for (int i = threadIdx.x; i < P * 256; i += THREADS) {
			const uint32_t v = copies_sum<C>(s_h, i);
			if (v) atomicAdd(&h[i], v);
		}
		__syncthreads();
	}
}

template <int P>
__global__ __launch_bounds__(256) void k_seg_plan(const uint64_t *__restrict__ seg, const uint32_t *__restrict__ large_ids,
	uint32_t *__restrict__ hist, uint32_t *__restrict__ bucket, SegSel *__restrict__ sel, int reverse,
	const uint32_t *__restrict__ flag)
{
	__shared__ uint32_t s_wave[4];
	__shared__ uint32_t s_uniform[P];
	const uint32_t L = blockIdx.x, t = threadIdx.x, lane = t & 63u, w = t >> 6;
	if (!flag[L]) {
		// partitioned by its top digit: no LSD pass and no copy for this segment
		if (t == 0) { sel[L].sel = 0; sel[L].mask = 0; }
		return;
	}
	const uint32_t s = large_ids[L];
	const uint64_t b = seg[s];
	const uint64_t len = seg[s + 1] - b;
	uint32_t *h = hist + static_cast<uint64_t>(L) * P * 256;
	uint32_t *bk = bucket + static_cast<uint64_t>(L) * P * 256;
	for (int p = 0; p < P; ++p) {
		const uint32_t c = h[p * 256 + t];
		h[p * 256 + t] = 0;
		const uint32_t incl = wave_inclusive_sum(c);
		if (lane == 63) s_wave[w] = incl;
		if (t == 0) s_uniform[p] = 0;
		__syncthreads();
		uint32_t before = 0;
		for (uint32_t v = 0; v < w; ++v) before += s_wave[v];
		bk[p * 256 + t] = before + incl - c;  // relative to the segment start
		if (static_cast<uint64_t>(c) == len) s_uniform[p] = 1;
		__syncthreads();
	}
	if (t != 0) return;
	int A = 0;
	uint32_t mask = 0;
	for (int p = 0; p < P; ++p)
		if (!s_uniform[p]) { ++A; mask |= 1u << p; }
	uint32_t code = 0;
	if (A == 0) {
		mask = kSelCopy;
	} else {
		uint32_t src = 0;  // input
		int j = 0;
		for (int p = 0; p < P; ++p) {
			if (!(mask >> p & 1u)) continue;
			const uint32_t dst = ((A - 1 - j) & 1) ? 2u : 1u;  // the last pass lands in the output
			code |= (src | dst << 2) << (4 * p);
			src = dst;
			++j;
		}
	}
	if (reverse) mask |= kSelReverse;
	sel[L].sel = code;
	sel[L].mask = mask;
}

template <typename K, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_seg_pass(const K *__restrict__ in, K *__restrict__ out, K *__restrict__ buf,
	const SegTile *__restrict__ tiles, const uint32_t *__restrict__ ftiles, const uint32_t *__restrict__ bucket,
	const SegSel *__restrict__ sel, uint64_t *__restrict__ lookback, SegCtrl *ctrl, KeyXform xf, int pass)
{
	constexpr int TILE = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_glob[256];
	__shared__ uint32_t s_wsum[4];
	__shared__ uint32_t s_tile;
	__shared__ __align__(16) unsigned char s_raw[max_bytes(TILE * sizeof(K), rank_peer_bytes(THREADS))];
	K *s_keys = reinterpret_cast<K *>(s_raw);
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_raw);
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t shift = 8u * static_cast<uint32_t>(pass);
	const uint32_t tag = ((ctrl->epoch * kMaxPasses + static_cast<uint32_t>(pass)) % 0x3FFFFFFFu) + 1u;
	const uint32_t num_tiles = ctrl->n_ftiles;
	for (;;) {
		if (t == 0) s_tile = atomicAdd(&ctrl->tile_counter[pass], 1u);
		__syncthreads();
		const uint32_t g = s_tile;  // position in the flagged list: the look-back row of the tile
		if (g >= num_tiles) break;
		const SegTile tl = tiles[ftiles[g]];
		const SegSel sl = sel[tl.L];
		if (!(sl.mask >> pass & 1u)) {
			__syncthreads();
			continue;  // this segment skips the digit: none of its tiles takes part
		}
		const uint32_t code = sl.sel >> (4 * pass);
		const K *src = pick<K>(code & 3u, in, out, buf);
		K *dst = pick<K>(code >> 2 & 3u, in, out, buf);
		const bool rev = (sl.mask & kSelReverse) && ((sl.mask & ((1u << pass) - 1u)) == 0);  // first executed pass
		const uint64_t sb = tl.sb;
		const uint32_t len = tl.len;
		const uint32_t base = tl.y * TILE;
		const uint32_t valid = min(static_cast<uint32_t>(TILE), len - base);
		K keys[KPT];
		uint32_t dig[KPT], pos[KPT], total = 0;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < valid) {
				const uint32_t r = base + l;
				keys[j] = src[sb + (rev ? len - 1 - r : r)];
				dig[j] = digit_of<K>(keys[j], xf, shift);
			} else {
				keys[j] = K(0);
				dig[j] = 0xFFu;
			}
		}
		block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
		if (t < 256) {
			const uint32_t pub = total - ((t == 255) ? static_cast<uint32_t>(TILE) - valid : 0u);
			uint64_t *lb = lookback + static_cast<uint64_t>(g) * 256;
			uint32_t excl = 0;
			if (tl.y == 0) {
				atomic_store_agent(&lb[t], lb_word(pub, kLbInclusive, tag));
			} else {
				atomic_store_agent(&lb[t], lb_word(pub, kLbAggregate, tag));
				excl = lookback_window<16>(lookback + t, static_cast<int64_t>(g), static_cast<int64_t>(g) - tl.y, tag);
				atomic_store_agent(&lb[t], lb_word(excl + pub, kLbInclusive, tag));
			}
			s_glob[t] = bucket[(static_cast<uint64_t>(tl.L) * P + pass) * 256 + t] + excl - s_digoff[t];
		}
#pragma unroll
		for (int j = 0; j < KPT; ++j) s_keys[pos[j]] = keys[j];
		__syncthreads();
		K *dseg = dst + sb;
		for (uint32_t i = t; i < valid; i += THREADS) {
			const K k = s_keys[i];
			dseg[s_glob[digit_of<K>(k, xf, shift)] + i] = k;
		}
		__syncthreads();
	}
}

template <typename K>
__global__ __launch_bounds__(256) void k_seg_copy(const K *__restrict__ in, K *__restrict__ out, const uint64_t *__restrict__ seg,
	const uint32_t *__restrict__ large_ids, const SegSel *__restrict__ sel, uint32_t nlarge)
{
	for (uint32_t L = blockIdx.x; L < nlarge; L += gridDim.x) {
		const SegSel sl = sel[L];
		if (!(sl.mask & kSelCopy)) continue;
		const uint32_t s = large_ids[L];
		const uint64_t b = seg[s], len = seg[s + 1] - b;
		const bool rev = (sl.mask & kSelReverse) != 0;
		for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) out[b + i] = in[b + (rev ? len - 1 - i : i)];
	}
}

__global__ void k_seg_finish(SegCtrl *ctrl, uint32_t *flag, uint32_t nlarge)
{
	for (uint32_t i = threadIdx.x; i < nlarge; i += blockDim.x) flag[i] = 0;
	if (threadIdx.x < kMaxPasses) ctrl->tile_counter[threadIdx.x] = 0;
	if (threadIdx.x == 0) {
		ctrl->n_ftiles = 0;
		ctrl->q_small = 0;
		ctrl->q_sub = 0;
		ctrl->n_sub = 0;
		ctrl->n_big = 0;
		ctrl->q_big = 0;
		ctrl->epoch = ctrl->epoch + 1u;
	}
}

__global__ void k_seg_zero(uint64_t *p, uint64_t words)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += step) p[i] = 0;
}

// ---------------------------------------------------------------------------------------------

inline size_t up256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

template <int THREADS, int KPT, typename F>
int occupancy_of(F kernel)
{
	int o = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, kernel, THREADS, 0) != hipSuccess || o < 1) o = 1;
	return o;
}

} // namespace

// tiles of the long segments' partition and LSD passes (threads x keys per thread), in the order of
// the tile_cfg argument of the segmented plan
constexpr int kNumSegTileCfgs = 5;
constexpr TileCfg kSegTileCfgs[kNumSegTileCfgs] = {{256, 8}, {256, 16}, {512, 8}, {512, 16}, {256, 24}};
// item sort capacity tiers (256x4, 256x16, 512x16)
constexpr int kNumSmallCaps = 3;
constexpr int kSmallCaps[kNumSmallCaps] = {256 * 4, 256 * 16, 512 * 16};
constexpr uint32_t kWaveItemMax = 64;  // short segments up to this length: one wave each

struct SegPlan {
	int key_bytes = 4;
	unsigned flags = 0;
	uint64_t nseg = 0, total = 0;
	int tile_cfg = 1;    // tile of the long segments' partition and LSD passes (kSegTileCfgs)
	int small_cfg = 1;   // LDS capacity tier of the item sort (kSmallCaps)
	uint32_t nsmall = 0, nlarge = 0, ntiles = 0, tile_keys = 0, small_cap = 0;
	uint32_t tier_count[3] = {0, 0, 0};  // short segments per capacity tier (the rest: at most 64 keys)
	char *mem = nullptr;  // one device allocation for every table and the workspace
	size_t mem_bytes = 0;
	SegCtrl *ctrl = nullptr;
	uint64_t *small_start = nullptr, *sub_start = nullptr, *big_start = nullptr;
	uint32_t *small_len = nullptr, *sub_len = nullptr, *q_small = nullptr, *q_sub = nullptr, *big_len = nullptr,
		*q_big = nullptr;
	uint32_t *large_ids = nullptr, *tile_first = nullptr, *counts1 = nullptr, *flag = nullptr, *ftiles = nullptr;
	uint32_t *hist = nullptr, *bucket = nullptr;
	SegTile *tiles = nullptr;
	SegSel *sel = nullptr;
	uint64_t *lookback = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipStream_t side = nullptr;
	hipEvent_t fork = nullptr, join = nullptr;
};


namespace {

// the transform keeps keys apart (at most a complement of some bits): equal sort keys are equal keys
inline bool xform_injective(const KeyXform &xf, int key_bytes)
{
	const uint64_t ones = key_bytes >= 8 ? ~0ull : ((1ull << (8 * key_bytes)) - 1ull);
	return xf.A == ones && xf.sh == 0 && xf.B == 0 && xf.C == 0 && xf.D == 0;
}

// One capacity tier of the item sort over items [0, count) of the given table (count null: count_value
// items); a long-run item is queued as qbase + its index. 1-byte keys take one LDS LSD pass instead.
template <typename K, int THREADS, int KPT, int PB>
hipError_t launch_sort_t(const void *src, void *dst, const uint64_t *istart, const uint32_t *ilen, const uint32_t *count,
	uint32_t count_value, const KeyXform &xf, int reverse, uint32_t *queue, uint32_t *qcount, uint32_t qbase, hipStream_t s)
{
	const uint32_t cus = static_cast<uint32_t>(device_cu_count());
	if constexpr (sizeof(K) == 1) {
		static const int occ = occupancy_of<THREADS, KPT>(k_item_lsd<K, THREADS, KPT>);
		const uint32_t full = static_cast<uint32_t>(occ) * cus;
		const uint32_t grid = count ? full : std::min<uint32_t>(full, std::max<uint32_t>(1u, count_value));
		hipLaunchKernelGGL((k_item_lsd<K, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(src),
			static_cast<K *>(dst), istart, ilen, nullptr, count, count_value, xf, reverse);
	} else if (xform_injective(xf, static_cast<int>(sizeof(K)))) {
		static const int occ = occupancy_of<THREADS, KPT>(k_item_sort<K, THREADS, KPT, PB, true>);
		const uint32_t full = static_cast<uint32_t>(occ) * cus;
		const uint32_t grid = count ? full : std::min<uint32_t>(full, std::max<uint32_t>(1u, count_value));
		hipLaunchKernelGGL((k_item_sort<K, THREADS, KPT, PB, true>), dim3(grid), dim3(THREADS), 0, s,
			static_cast<const K *>(src), static_cast<K *>(dst), istart, ilen, count, count_value, xf, reverse, queue, qcount, qbase);
	} else {
		static const int occ = occupancy_of<THREADS, KPT>(k_item_sort<K, THREADS, KPT, PB, false>);
		const uint32_t full = static_cast<uint32_t>(occ) * cus;
		const uint32_t grid = count ? full : std::min<uint32_t>(full, std::max<uint32_t>(1u, count_value));
		hipLaunchKernelGGL((k_item_sort<K, THREADS, KPT, PB, false>), dim3(grid), dim3(THREADS), 0, s,
			static_cast<const K *>(src), static_cast<K *>(dst), istart, ilen, count, count_value, xf, reverse, queue, qcount, qbase);
	}
	return hipGetLastError();
}

// the queued items of a table (no queue for 1-byte keys)
template <typename K, int THREADS, int KPT>
hipError_t launch_lsd_t(const void *src, void *dst, const uint64_t *istart, const uint32_t *ilen, const uint32_t *queue,
	const uint32_t *qcount, const KeyXform &xf, int reverse, hipStream_t s)
{
	if constexpr (sizeof(K) == 1) return hipSuccess;
	static const int occ = occupancy_of<THREADS, KPT>(k_item_lsd<K, THREADS, KPT>);
	hipLaunchKernelGGL((k_item_lsd<K, THREADS, KPT>), dim3(static_cast<uint32_t>(occ) * static_cast<uint32_t>(device_cu_count())),
		dim3(THREADS), 0, s, static_cast<const K *>(src), static_cast<K *>(dst), istart, ilen, queue, qcount, 0u, xf, reverse);
	return hipGetLastError();
}

// capacity tiers: 0 = 256x4 (1024 keys), 1 = 512x8 (4096), 2 = 512x16 (8192), partition bits chosen
// for about four keys per partition (two in the 4096 tier, whose counters take 16-bit halves: its
// LDS, 29 KB, lets four blocks of eight waves share a CU). Measured on 65,536 int32 segments of
// log-uniform lengths scaled to 2^28 keys, the 4096 tier's sort took 1.30-1.34 ms with 256x16
// whether partitions held four, two or one key (about 95 vector instructions per key); 512x8, with
// twice the waves per CU to hide the LDS latency of its phases, takes the whole graph from 2.36 to
// 2.25 ms. 512x2 for the 1024 tier measured slower (2.38 ms).
template <typename K>
hipError_t launch_sort_cfg(int cfg, const void *src, void *dst, const uint64_t *istart, const uint32_t *ilen,
	const uint32_t *count, uint32_t count_value, const KeyXform &xf, int reverse, uint32_t *queue, uint32_t *qcount,
	uint32_t qbase, hipStream_t s)
{
	// partition bits per tier: about one key per partition for 4-byte keys in the 4096- and 8192-key
	// tiers (C4 scaled to 2^28 keys: 2.24 -> 2.04-2.10 ms per graph, 13 bits in the 4096 tier 2.21; at
	// its stated lengths, where half the buckets of a 2^20-key segment take the 8192 tier, 46.3 ms per
	// graph warm, 50.4 in a fresh box's first process); 8-byte keys measured the same either way and
	// keep the smaller counter tables
	constexpr int PB1 = sizeof(K) == 4 ? 12 : 11, PB2 = sizeof(K) == 4 ? 13 : 11;
	switch (cfg) {
	case 0: return launch_sort_t<K, 256, 4, 8>(src, dst, istart, ilen, count, count_value, xf, reverse, queue, qcount, qbase, s);
	case 1: return launch_sort_t<K, 512, 8, PB1>(src, dst, istart, ilen, count, count_value, xf, reverse, queue, qcount, qbase, s);
	default: return launch_sort_t<K, 512, 16, PB2>(src, dst, istart, ilen, count, count_value, xf, reverse, queue, qcount, qbase, s);
	}
}

template <typename K>
hipError_t launch_lsd_cfg(int cfg, const void *src, void *dst, const uint64_t *istart, const uint32_t *ilen,
	const uint32_t *queue, const uint32_t *qcount, const KeyXform &xf, int reverse, hipStream_t s)
{
	switch (cfg) {
	case 0: return launch_lsd_t<K, 256, 4>(src, dst, istart, ilen, queue, qcount, xf, reverse, s);
	case 1: return launch_lsd_t<K, 256, 16>(src, dst, istart, ilen, queue, qcount, xf, reverse, s);
	default: return launch_lsd_t<K, 512, 16>(src, dst, istart, ilen, queue, qcount, xf, reverse, s);
	}
}

// The short segments, longest first: one launch per capacity tier down to 65 keys, one wave per
// segment of at most 64 keys, then the queued long-run segments.
template <typename K>
hipError_t launch_small(SegPlan &pl, const void *in, void *out, const KeyXform &xf, hipStream_t s)
{
	if (!pl.nsmall) return hipSuccess;
	const int rev = (pl.flags & kRev) ? 1 : 0;
	hipError_t e;
	uint32_t off = 0;
	for (int cfg = kNumSmallCaps - 1; cfg >= 0; --cfg) {
		const uint32_t n = pl.tier_count[cfg];
		if (!n) continue;
		if ((e = launch_sort_cfg<K>(cfg, in, out, pl.small_start + off, pl.small_len + off, nullptr, n, xf, rev, pl.q_small,
			&pl.ctrl->q_small, off, s)) != hipSuccess) return e;
		off += n;
	}
	if (off < pl.nsmall) {
		const uint32_t n = pl.nsmall - off;
		const uint32_t grid = std::min<uint32_t>((n + 3) / 4, 8u * static_cast<uint32_t>(device_cu_count()));
		hipLaunchKernelGGL((k_item_wave<K, 256>), dim3(grid), dim3(256), 0, s, static_cast<const K *>(in), static_cast<K *>(out),
			pl.small_start + off, pl.small_len + off, n, xf, rev);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	return launch_lsd_cfg<K>(pl.small_cfg, in, out, pl.small_start, pl.small_len, pl.q_small, &pl.ctrl->q_small, xf, rev, s);
}

template <typename K, int THREADS, int KPT>
hipError_t launch_seg_pass(SegPlan &pl, const void *in, void *out, void *buf, const uint64_t *seg, const KeyXform &xf,
	int pass, hipStream_t s)
{
	static int occ = occupancy_of<THREADS, KPT>(k_seg_pass<K, THREADS, KPT>);
	const uint32_t grid = std::min<uint32_t>(pl.ntiles, static_cast<uint32_t>(occ * device_cu_count()));
	hipLaunchKernelGGL((k_seg_pass<K, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(in),
		static_cast<K *>(out), static_cast<K *>(buf), pl.tiles, pl.ftiles, pl.bucket, pl.sel, pl.lookback, pl.ctrl, xf, pass);
	return hipGetLastError();
}

template <typename K, int THREADS, int KPT, bool STABLE>
hipError_t launch_scatter1_t(SegPlan &pl, const void *in, void *buf, const KeyXform &xf, int reverse, hipStream_t s)
{
	static int occ = occupancy_of<THREADS, KPT>(k_seg_scatter1<K, THREADS, KPT, STABLE>);
	const uint32_t grid = 8u * std::min<uint32_t>((pl.ntiles + 7) / 8, static_cast<uint32_t>(occ * device_cu_count()) / 8u);
	hipLaunchKernelGGL((k_seg_scatter1<K, THREADS, KPT, STABLE>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(in),
		static_cast<K *>(buf), pl.tiles, pl.ntiles, pl.counts1, pl.flag, xf, reverse);
	return hipGetLastError();
}

template <typename K, int THREADS, int KPT>
hipError_t launch_scatter1(SegPlan &pl, const void *in, void *buf, const uint64_t *, const KeyXform &xf, int reverse,
	hipStream_t s)
{
	if (xform_injective(xf, static_cast<int>(sizeof(K)))) return launch_scatter1_t<K, THREADS, KPT, false>(pl, in, buf, xf, reverse, s);
	return launch_scatter1_t<K, THREADS, KPT, true>(pl, in, buf, xf, reverse, s);
}

template <typename K>
hipError_t launch_large(SegPlan &pl, const void *in, void *out, void *buf, const uint64_t *seg, const KeyXform &xf, hipStream_t s)
{
	if (!pl.nlarge) return hipSuccess;
	constexpr int P = sizeof(K);
	const int rev = (pl.flags & kRev) ? 1 : 0;
	const uint32_t cus = static_cast<uint32_t>(device_cu_count());
	hipError_t e;
	// the partition by the top digit, into buf, then the items of its buckets into out
	hipLaunchKernelGGL((k_seg_count1<K, 256>), dim3(pl.ntiles), dim3(256), 0, s, static_cast<const K *>(in), pl.tiles,
		pl.tile_keys, xf, rev, pl.counts1);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	constexpr uint32_t kBigCap = static_cast<uint32_t>(kSmallCaps[kNumSmallCaps - 1]);
	hipLaunchKernelGGL(k_seg_scan1, dim3((pl.nlarge + kScan1Segs - 1) / kScan1Segs), dim3(64 * kScan1Segs), 0, s, seg,
		pl.large_ids, pl.nlarge, pl.tile_first, pl.counts1, pl.small_cap, kBigCap, pl.flag, pl.ftiles, pl.ctrl, pl.sub_start,
		pl.sub_len, pl.big_start, pl.big_len);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	switch (pl.tile_cfg) {
	case 0: e = launch_scatter1<K, 256, 8>(pl, in, buf, seg, xf, rev, s); break;
	case 1: e = launch_scatter1<K, 256, 16>(pl, in, buf, seg, xf, rev, s); break;
	case 2: e = launch_scatter1<K, 512, 8>(pl, in, buf, seg, xf, rev, s); break;
	case 3: e = launch_scatter1<K, 512, 16>(pl, in, buf, seg, xf, rev, s); break;
	default: e = launch_scatter1<K, 256, 24>(pl, in, buf, seg, xf, rev, s); break;
	}
	if (e != hipSuccess) return e;
	if ((e = launch_sort_cfg<K>(pl.small_cfg, buf, out, pl.sub_start, pl.sub_len, &pl.ctrl->n_sub, 0, xf, 0, pl.q_sub,
		&pl.ctrl->q_sub, 0, s)) != hipSuccess) return e;
	if ((e = launch_lsd_cfg<K>(pl.small_cfg, buf, out, pl.sub_start, pl.sub_len, pl.q_sub, &pl.ctrl->q_sub, xf, 0, s)) !=
		hipSuccess) return e;
	if (pl.small_cfg < kNumSmallCaps - 1) {
		// buckets above the item capacity: single-bucket items on the largest tier
		if ((e = launch_sort_cfg<K>(kNumSmallCaps - 1, buf, out, pl.big_start, pl.big_len, &pl.ctrl->n_big, 0, xf, 0, pl.q_big,
			&pl.ctrl->q_big, 0, s)) != hipSuccess) return e;
		if ((e = launch_lsd_cfg<K>(kNumSmallCaps - 1, buf, out, pl.big_start, pl.big_len, pl.q_big, &pl.ctrl->q_big, xf, 0, s)) !=
			hipSuccess) return e;
	}
	// segments with a bucket above the LDS capacity: the LSD passes (every launch skips the others)
	hipLaunchKernelGGL((k_seg_hist<K, 256>), dim3(std::min<uint32_t>(pl.ntiles, 8u * cus)), dim3(256), 0, s,
		static_cast<const K *>(in), pl.tiles, pl.ftiles, pl.ctrl, pl.tile_keys, xf, pl.hist);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	hipLaunchKernelGGL((k_seg_plan<P>), dim3(pl.nlarge), dim3(256), 0, s, seg, pl.large_ids, pl.hist, pl.bucket, pl.sel, rev,
		pl.flag);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	for (int p = 0; p < P; ++p) {
		switch (pl.tile_cfg) {
		case 0: e = launch_seg_pass<K, 256, 8>(pl, in, out, buf, seg, xf, p, s); break;
		case 1: e = launch_seg_pass<K, 256, 16>(pl, in, out, buf, seg, xf, p, s); break;
		case 2: e = launch_seg_pass<K, 512, 8>(pl, in, out, buf, seg, xf, p, s); break;
		case 3: e = launch_seg_pass<K, 512, 16>(pl, in, out, buf, seg, xf, p, s); break;
		default: e = launch_seg_pass<K, 256, 24>(pl, in, out, buf, seg, xf, p, s); break;
		}
		if (e != hipSuccess) return e;
	}
	const uint32_t cgrid = std::min<uint32_t>(pl.nlarge, 4096u);
	hipLaunchKernelGGL((k_seg_copy<K>), dim3(cgrid), dim3(256), 0, s, static_cast<const K *>(in), static_cast<K *>(out), seg,
		pl.large_ids, pl.sel, pl.nlarge);
	return hipGetLastError();
}

template <typename K>
hipError_t run_typed(SegPlan &pl, const void *in, void *out, void *buf, const uint64_t *seg, hipStream_t s)
{
	const KeyXform xf = make_xform(8 * sizeof(K), pl.flags);
	hipError_t e;
	// the short segments run on a side stream in parallel with the long-segment chain
	if ((e = hipEventRecord(pl.fork, s)) != hipSuccess) return e;
	if ((e = hipStreamWaitEvent(pl.side, pl.fork, 0)) != hipSuccess) return e;
	if ((e = launch_small<K>(pl, in, out, xf, pl.side)) != hipSuccess) return e;
	if ((e = hipEventRecord(pl.join, pl.side)) != hipSuccess) return e;
	if ((e = launch_large<K>(pl, in, out, buf, seg, xf, s)) != hipSuccess) return e;
	if ((e = hipStreamWaitEvent(s, pl.join, 0)) != hipSuccess) return e;
	hipLaunchKernelGGL(k_seg_finish, dim3(1), dim3(256), 0, s, pl.ctrl, pl.flag, pl.nlarge);
	return hipGetLastError();
}

} // namespace

int tile_keys_of(int cfg) { return kSegTileCfgs[cfg].threads * kSegTileCfgs[cfg].kpt; }

void segplan_default_cfg(int key_bytes, unsigned flags, int *tile_cfg, int *small_cfg)
{
	// winners of the sweep over (tile_cfg, small_cfg) on 65,536 segments with log-uniform lengths
	// 2^4..2^20 and 2^28 keys in total, per key width and transform kind (the tuned table ships with
	// the Python package)
	const bool inj = xform_injective(make_xform(8u * static_cast<unsigned>(key_bytes), flags), key_bytes);
	int tc = 3, sc = 1;
	switch (key_bytes) {
	case 1: tc = 1; sc = 1; break;
	case 2: tc = 4; sc = 1; break;
	case 4: tc = inj ? 3 : 2; sc = 1; break;
	default: tc = inj ? 3 : 2; sc = 0; break;
	}
	if (*tile_cfg < 0) *tile_cfg = tc;
	if (*small_cfg < 0) *small_cfg = sc;
}
int small_cap_of(int cfg) { return kSmallCaps[cfg]; }

hipError_t segplan_create(int key_bytes, unsigned flags, uint64_t nseg, const uint64_t *host_off, int tile_cfg, int small_cfg,
	SegPlan **out)
{
	*out = nullptr;
	if (key_bytes != 1 && key_bytes != 2 && key_bytes != 4 && key_bytes != 8) return hipErrorInvalidValue;
	if (tile_cfg < 0 || tile_cfg >= kNumSegTileCfgs || small_cfg < 0 || small_cfg >= kNumSmallCaps) return hipErrorInvalidValue;
	SegPlan *pl = new (std::nothrow) SegPlan;
	if (!pl) return hipErrorOutOfMemory;
	pl->key_bytes = key_bytes; pl->flags = flags; pl->nseg = nseg; pl->tile_cfg = tile_cfg; pl->small_cfg = small_cfg;
	pl->tile_keys = static_cast<uint32_t>(tile_keys_of(tile_cfg));
	pl->small_cap = static_cast<uint32_t>(kSmallCaps[small_cfg]);
	pl->total = nseg ? host_off[nseg] - host_off[0] : 0;
	std::vector<std::pair<uint64_t, uint32_t>> small;
	std::vector<uint32_t> large, tile_first;
	std::vector<SegTile> tiles;
	for (uint64_t s = 0; s < nseg; ++s) {
		if (host_off[s + 1] < host_off[s] || host_off[s + 1] - host_off[s] > 0xFFFFFFFFull) { delete pl; return hipErrorInvalidValue; }
		const uint64_t len = host_off[s + 1] - host_off[s];
		if (len == 0) continue;
		if (len <= pl->small_cap) {
			small.push_back({len, static_cast<uint32_t>(s)});  // one key is copied by the item sort
		} else {
			const uint32_t L = static_cast<uint32_t>(large.size());
			large.push_back(static_cast<uint32_t>(s));
			tile_first.push_back(static_cast<uint32_t>(tiles.size()));
			const uint32_t nt = static_cast<uint32_t>((len + pl->tile_keys - 1) / pl->tile_keys);
			for (uint32_t i = 0; i < nt; ++i) tiles.push_back(SegTile{host_off[s], static_cast<uint32_t>(len), L, i, {0, 0, 0}});
		}
	}
	tile_first.push_back(static_cast<uint32_t>(tiles.size()));
	// longest short segments first, so the tail of the item loop is made of quick ones
	std::stable_sort(small.begin(), small.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
	std::vector<uint64_t> small_start(small.size());
	std::vector<uint32_t> small_len(small.size());
	for (size_t i = 0; i < small.size(); ++i) {
		small_start[i] = host_off[small[i].second];
		small_len[i] = static_cast<uint32_t>(small[i].first);
	}
	pl->nsmall = static_cast<uint32_t>(small.size());
	for (const auto &it : small) {
		if (it.first <= kWaveItemMax) continue;
		int cfg = 0;
		while (cfg < small_cfg && it.first > static_cast<uint64_t>(kSmallCaps[cfg])) ++cfg;
		++pl->tier_count[cfg];
	}
	pl->nlarge = static_cast<uint32_t>(large.size());
	pl->ntiles = static_cast<uint32_t>(tiles.size());
	const int P = key_bytes;
	const size_t nsub = static_cast<size_t>(large.size()) * 256;
	size_t off = 0;
	const size_t o_ctrl = off; off += up256(sizeof(SegCtrl));
	const size_t o_sstart = off; off += up256(small.size() * 8 + 8);
	const size_t o_slen = off; off += up256(small.size() * 4 + 4);
	const size_t o_qsmall = off; off += up256(small.size() * 4 + 4);
	const size_t o_large = off; off += up256(large.size() * 4 + 4);
	const size_t o_tfirst = off; off += up256(tile_first.size() * 4 + 4);
	const size_t o_flag = off; off += up256(large.size() * 4 + 4);
	const size_t o_ftiles = off; off += up256(tiles.size() * 4 + 4);
	const size_t o_tiles = off; off += up256(tiles.size() * sizeof(SegTile) + 8);
	const size_t o_counts = off; off += up256(tiles.size() * 256 * 4 + 4);
	const size_t o_substart = off; off += up256(nsub * 8 + 8);
	const size_t o_sublen = off; off += up256(nsub * 4 + 4);
	const size_t o_qsub = off; off += up256(nsub * 4 + 4);
	// big items: buckets above the item capacity (none when the capacity is the largest tier)
	const size_t nbig = small_cfg < kNumSmallCaps - 1 ? nsub : 0;
	const size_t o_bigstart = off; off += up256(nbig * 8 + 8);
	const size_t o_biglen = off; off += up256(nbig * 4 + 4);
	const size_t o_qbig = off; off += up256(nbig * 4 + 4);
	const size_t o_sel = off; off += up256(large.size() * sizeof(SegSel) + 8);
	const size_t o_hist = off; off += up256(large.size() * P * 256 * 4 + 4);
	const size_t o_bucket = off; off += up256(large.size() * P * 256 * 4 + 4);
	const size_t o_lb = off; off += up256(static_cast<size_t>(tiles.size()) * 256 * 8 + 8);
	pl->mem_bytes = off;
	hipError_t e = hipMalloc(reinterpret_cast<void **>(&pl->mem), off);
	if (e != hipSuccess) { delete pl; return e; }
	pl->ctrl = reinterpret_cast<SegCtrl *>(pl->mem + o_ctrl);
	pl->small_start = reinterpret_cast<uint64_t *>(pl->mem + o_sstart);
	pl->small_len = reinterpret_cast<uint32_t *>(pl->mem + o_slen);
	pl->q_small = reinterpret_cast<uint32_t *>(pl->mem + o_qsmall);
	pl->large_ids = reinterpret_cast<uint32_t *>(pl->mem + o_large);
	pl->tile_first = reinterpret_cast<uint32_t *>(pl->mem + o_tfirst);
	pl->flag = reinterpret_cast<uint32_t *>(pl->mem + o_flag);
	pl->ftiles = reinterpret_cast<uint32_t *>(pl->mem + o_ftiles);
	pl->tiles = reinterpret_cast<SegTile *>(pl->mem + o_tiles);
	pl->counts1 = reinterpret_cast<uint32_t *>(pl->mem + o_counts);
	pl->sub_start = reinterpret_cast<uint64_t *>(pl->mem + o_substart);
	pl->sub_len = reinterpret_cast<uint32_t *>(pl->mem + o_sublen);
	pl->q_sub = reinterpret_cast<uint32_t *>(pl->mem + o_qsub);
	pl->big_start = reinterpret_cast<uint64_t *>(pl->mem + o_bigstart);
	pl->big_len = reinterpret_cast<uint32_t *>(pl->mem + o_biglen);
	pl->q_big = reinterpret_cast<uint32_t *>(pl->mem + o_qbig);
	pl->sel = reinterpret_cast<SegSel *>(pl->mem + o_sel);
	pl->hist = reinterpret_cast<uint32_t *>(pl->mem + o_hist);
	pl->bucket = reinterpret_cast<uint32_t *>(pl->mem + o_bucket);
	pl->lookback = reinterpret_cast<uint64_t *>(pl->mem + o_lb);
	hipLaunchKernelGGL(k_seg_zero, dim3(1024), dim3(256), 0, nullptr, reinterpret_cast<uint64_t *>(pl->mem), off / 8);
	e = hipGetLastError();
	auto put = [&](void *dst, const void *src, size_t bytes) {
		if (e == hipSuccess && bytes) e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
	};
	put(pl->small_start, small_start.data(), small_start.size() * 8);
	put(pl->small_len, small_len.data(), small_len.size() * 4);
	put(pl->large_ids, large.data(), large.size() * 4);
	put(pl->tile_first, tile_first.data(), tile_first.size() * 4);
	put(pl->tiles, tiles.data(), tiles.size() * sizeof(SegTile));
	if (e == hipSuccess) e = hipStreamCreateWithFlags(&pl->side, hipStreamNonBlocking);
	if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming);
	if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->join, hipEventDisableTiming);
	if (e == hipSuccess) e = hipDeviceSynchronize();
	if (e != hipSuccess) { segplan_destroy(pl); return e; }
	*out = pl;
	return hipSuccess;
}

void segplan_destroy(SegPlan *pl)
{
	if (!pl) return;
	if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
	if (pl->graph) (void)hipGraphDestroy(pl->graph);
	if (pl->fork) (void)hipEventDestroy(pl->fork);
	if (pl->join) (void)hipEventDestroy(pl->join);
	if (pl->side) (void)hipStreamDestroy(pl->side);
	if (pl->mem) (void)hipFree(pl->mem);
	delete pl;
}

hipError_t segplan_run(SegPlan *pl, const void *in, void *out, void *buf, const uint64_t *dev_off, hipStream_t s)
{
	if (!pl || !in || !out || (pl->nlarge && !buf) || !dev_off) return hipErrorInvalidValue;
	switch (pl->key_bytes) {
	case 1: return run_typed<uint8_t>(*pl, in, out, buf, dev_off, s);
	case 2: return run_typed<uint16_t>(*pl, in, out, buf, dev_off, s);
	case 4: return run_typed<uint32_t>(*pl, in, out, buf, dev_off, s);
	default: return run_typed<uint64_t>(*pl, in, out, buf, dev_off, s);
	}
}

hipError_t segplan_capture(SegPlan *pl, const void *in, void *out, void *buf, const uint64_t *dev_off, hipStream_t s)
{
	if (!pl) return hipErrorInvalidValue;
	if (pl->exec) { (void)hipGraphExecDestroy(pl->exec); pl->exec = nullptr; }
	if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
	hipStream_t cs = s;
	bool own = false;
	hipError_t e;
	if (!cs) {  // the null stream cannot be captured: capture on a private stream
		if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) return e;
		own = true;
	}
	if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) == hipSuccess) {
		const hipError_t er = segplan_run(pl, in, out, buf, dev_off, cs);
		e = hipStreamEndCapture(cs, &pl->graph);
		if (er != hipSuccess) e = er;
	}
	if (e == hipSuccess) e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
	if (own) (void)hipStreamDestroy(cs);
	return e;
}

hipError_t segplan_launch(SegPlan *pl, hipStream_t s)
{
	if (!pl || !pl->exec) return hipErrorInvalidValue;
	return hipGraphLaunch(pl->exec, s);
}

void segplan_info(const SegPlan *pl, uint64_t *info)
{
	info[0] = pl->nsmall; info[1] = pl->nlarge; info[2] = pl->ntiles; info[3] = pl->tile_keys; info[4] = pl->small_cap;
	info[5] = pl->total; info[6] = pl->mem_bytes;
}

} // namespace rsbd8hip
