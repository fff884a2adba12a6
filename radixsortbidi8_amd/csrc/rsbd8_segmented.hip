// rsbd8_segmented.hip -- many independent sorts in one captured launch sequence.
//
// The reference sorts a batch by calling radixsortcopy once per array (Radixsortbidi8.hpp:60468);
// here every segment of a CSR batch is sorted with the same mode/direction flags. Segments up to
// the LDS capacity are items of their own; longer ones are partitioned by their top digit first:
//   k_item_sort     persistent blocks, one item (a key range sorted on its own) at a time, all in
//                   LDS: partition by the top bits of the transformed key with LDS atomics, then
//                   each key's exact position in its run by counting the run's keys that precede it
//                   in the stable order (smaller key, or equal key at a lower position);
//   k_item_lsd      items with a run too long for that: stable LDS LSD passes (block ranking);
//   k_seg_count1    per-tile counts of the top digit over the tiles of the long segments;
//   k_seg_scan1     one block per long segment: bucket sizes and starts, tile offsets, and the
//                   segment's buckets packed into items of at most the LDS capacity (consecutive
//                   buckets of one segment may share an item: they are already in top-digit
//                   order, so a sort of the item by the whole key is exact); a bucket above the
//                   capacity sends the segment to the LSD passes below instead;
//   k_seg_scatter1  the stable partition of the long segments by the top digit;
//   k_seg_hist, k_seg_plan, k_seg_pass, k_seg_copy
//                   for those segments only: per-segment digit histograms, plan, onesweep passes
//                   with the look-back confined to the segment, and the copy when every digit is
//                   uniform;
//   k_seg_finish    resets the counters and advances the look-back epoch.
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
	uint32_t pad[3];
};

// per long segment: sel holds, for pass p, bits 4p..4p+1 = source (0 input, 1 output, 2 buffer) and
// bits 4p+2..4p+3 = destination (all 32 bits for 8-byte keys); mask holds the executed passes in
// bits 0..7 and the two flags below
struct SegSel {
	uint32_t sel;
	uint32_t mask;
};
constexpr uint32_t kSelCopy = 1u << 30;     // no pass executes: the segment is copied
constexpr uint32_t kSelReverse = 1u << 31;  // the first executed pass (or the copy) reads back to front

template <typename K>
__device__ __forceinline__ K *pick(uint32_t code, const K *in, K *out, K *buf)
{
	return code == 0 ? const_cast<K *>(in) : (code == 1 ? out : buf);
}

// ---------------------------------------------------------------------------------------------

// Words of the item sort: (transformed key << 16) | position where they fit in one integer (their
// order is then the stable order); 8-byte keys keep the position in a separate array.
template <typename K> struct ItemWord;
template <> struct ItemWord<uint8_t> { using W = uint32_t; static constexpr bool packed = true; static constexpr int kbits = 8; };
template <> struct ItemWord<uint16_t> { using W = uint32_t; static constexpr bool packed = true; static constexpr int kbits = 16; };
template <> struct ItemWord<uint32_t> { using W = uint64_t; static constexpr bool packed = true; static constexpr int kbits = 32; };
template <> struct ItemWord<uint64_t> { using W = uint64_t; static constexpr bool packed = false; static constexpr int kbits = 64; };

constexpr uint32_t kItemRunMax = 96;

// Items i < *count (or count_value when count is null): keys [start, start + len) of src, read back
// to front under reverse, sorted into the same range of dst. Items with a run of more than
// kItemRunMax keys sharing the partition bits are queued for k_item_lsd.
template <typename K, int THREADS, int KPT, int PB>
__global__ __launch_bounds__(THREADS) void k_item_sort(const K *__restrict__ src, K *__restrict__ dst,
	const uint64_t *__restrict__ istart, const uint32_t *__restrict__ ilen, const uint32_t *count, uint32_t count_value,
	KeyXform xf, int reverse, uint32_t *__restrict__ queue, uint32_t *qcount)
{
	constexpr int CAP = THREADS * KPT;
	using IW = ItemWord<K>;
	using W = typename IW::W;
	constexpr int KB = IW::kbits;
	constexpr int PBE = PB < KB ? PB : KB;
	constexpr uint32_t NB = 1u << PBE;
	constexpr uint32_t PER = (NB + THREADS - 1) / THREADS;
	constexpr uint32_t kFinWords = CAP / 2;
	constexpr uint32_t kUnionWords = NB > kFinWords ? NB : kFinWords;
	static_assert(CAP <= 65536, "positions are packed into 16 bits");
	__shared__ __align__(16) W s_w[CAP];                 // words, then the raw keys in final order
	__shared__ uint16_t s_l[IW::packed ? 1 : CAP];       // positions of unpacked words
	__shared__ uint32_t s_u[kUnionWords];                // partition counters/cursors, then final positions
	__shared__ uint16_t s_off[NB + 1];
	__shared__ uint32_t s_wsum[THREADS / 64];
	__shared__ uint32_t s_big;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t nitems = count ? *count : count_value;
	for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
		const uint32_t len = ilen[it];
		if (len == 0) continue;
		const uint64_t start = istart[it];
		if (len == 1) {
			if (t == 0) dst[start] = src[start];
			continue;
		}
		K key[KPT];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			key[j] = l < len ? src[start + (reverse ? len - 1 - l : l)] : K(0);
		}
		uint32_t *s_cur = s_u;
		__syncthreads();  // the previous item's last reads of s_u and s_w are done
		for (uint32_t i = t; i < NB; i += THREADS) s_cur[i] = 0;
		if (t == 0) s_big = 0;
		__syncthreads();
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) atomicAdd(&s_cur[static_cast<uint32_t>(sort_key<K>(key[j], xf) >> (KB - PBE))], 1u);
		}
		__syncthreads();
		uint32_t c[PER], sum = 0, big = 0;
#pragma unroll
		for (uint32_t k = 0; k < PER; ++k) {
			const uint32_t b = t * PER + k;
			c[k] = b < NB ? s_cur[b] : 0u;
			sum += c[k];
			big |= c[k] > kItemRunMax;
		}
		const uint32_t incl = wave_inclusive_sum(sum);
		if (lane == 63) s_wsum[w] = incl;
		if (big) s_big = 1;
		__syncthreads();
		uint32_t run = incl - sum;
		for (uint32_t v = 0; v < w; ++v) run += s_wsum[v];
#pragma unroll
		for (uint32_t k = 0; k < PER; ++k) {
			const uint32_t b = t * PER + k;
			if (b < NB) {
				s_off[b] = static_cast<uint16_t>(run);
				s_cur[b] = run;
			}
			run += c[k];
		}
		if (t == 0) s_off[NB] = static_cast<uint16_t>(len);
		__syncthreads();
		if (s_big) {
			if (t == 0) queue[atomicAdd(qcount, 1u)] = it;
			continue;
		}
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) {
				const auto tk = sort_key<K>(key[j], xf);
				const uint32_t p = atomicAdd(&s_cur[static_cast<uint32_t>(tk >> (KB - PBE))], 1u);
				if constexpr (IW::packed) {
					s_w[p] = (static_cast<W>(tk) << 16) | static_cast<W>(l);
				} else {
					s_w[p] = static_cast<W>(tk);
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
				uint32_t d, lx;
				if constexpr (IW::packed) {
					d = static_cast<uint32_t>(x >> (16 + KB - PBE));
					lx = static_cast<uint32_t>(x) & 0xFFFFu;
				} else {
					d = static_cast<uint32_t>(x >> (KB - PBE));
					lx = s_l[i];
				}
				const uint32_t lo = s_off[d], hi = s_off[d + 1];
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
		__syncthreads();  // the cursors are dead: s_u holds the final positions from here
		uint16_t *s_fin = reinterpret_cast<uint16_t *>(s_u);
#pragma unroll
		for (int j = 0; j < KPT; ++j)
			if (at[j] != 0xFFFFFFFFu) s_fin[at[j]] = static_cast<uint16_t>(fin[j]);
		__syncthreads();
		K *s_raw = reinterpret_cast<K *>(s_w);
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) s_raw[s_fin[l]] = key[j];
		}
		__syncthreads();
		for (uint32_t i = t; i < len; i += THREADS) dst[start + i] = s_raw[i];
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
	__shared__ K s_keys[CAP];
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
			block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total);
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

// per-tile counts of the top digit (a tile is keys [i*tile_keys, (i+1)*tile_keys) of its segment in
// logical order: a physical range read back to front under reverse)
template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_seg_count1(const K *__restrict__ in, const uint64_t *__restrict__ seg,
	const uint32_t *__restrict__ large_ids, const uint2 *__restrict__ tiles, uint32_t ntiles, uint32_t tile_keys,
	KeyXform xf, int reverse, uint32_t *__restrict__ counts)
{
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_t[256];
	for (uint32_t g = blockIdx.x; g < ntiles; g += gridDim.x) {
		for (int i = threadIdx.x; i < 256; i += THREADS) s_t[i] = 0;
		__syncthreads();
		const uint2 tl = tiles[g];
		const uint32_t s = large_ids[tl.x];
		const uint64_t sb = seg[s];
		const uint32_t len = static_cast<uint32_t>(seg[s + 1] - sb);
		const uint32_t lo = tl.y * tile_keys, hi = min(len, lo + tile_keys);
		const uint64_t plo = sb + (reverse ? len - hi : lo);
		for (uint32_t i = threadIdx.x; i < hi - lo; i += THREADS)
			wave_count(s_t, digit_of<K>(in[plo + i], xf, 8u * (P - 1)));
		__syncthreads();
		for (int i = threadIdx.x; i < 256; i += THREADS) counts[static_cast<uint64_t>(g) * 256 + i] = s_t[i];
		__syncthreads();
	}
}

// One block per long segment, one thread per top digit: bucket sizes and starts, the segment's
// tile counts turned into output offsets, and its buckets packed greedily into items of at most
// cap keys. A bucket above cap flags the segment for the LSD passes (and skips everything else).
__global__ __launch_bounds__(256) void k_seg_scan1(const uint64_t *__restrict__ seg, const uint32_t *__restrict__ large_ids,
	const uint32_t *__restrict__ tile_first, uint32_t *__restrict__ counts, uint32_t cap, uint32_t *__restrict__ flag,
	uint32_t *__restrict__ ftiles, SegCtrl *ctrl, uint64_t *__restrict__ istart, uint32_t *__restrict__ ilen)
{
	__shared__ uint32_t s_wave[4];
	__shared__ uint32_t s_tot[256];
	__shared__ uint32_t s_at;
	const uint32_t L = blockIdx.x, d = threadIdx.x, lane = d & 63u, w = d >> 6;
	const uint32_t s = large_ids[L];
	const uint64_t sb = seg[s];
	const uint32_t t0 = tile_first[L], t1 = tile_first[L + 1];
	uint32_t tot = 0;
	for (uint32_t tl = t0; tl < t1; tl += 8) {
		uint32_t c[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) c[j] = (tl + j < t1) ? counts[static_cast<uint64_t>(tl + j) * 256 + d] : 0u;
#pragma unroll
		for (int j = 0; j < 8; ++j) tot += c[j];
	}
	s_tot[d] = tot;
	const uint32_t incl = wave_inclusive_sum(tot);
	if (lane == 63) s_wave[w] = incl;
	if (__syncthreads_or(tot > cap)) {
		// the segment's tiles join the list of the LSD passes as one run in tile order, so a tile's
		// look-back only meets tiles of its segment that were queued before it
		if (d == 0) {
			flag[L] = 1;
			s_at = atomicAdd(&ctrl->n_ftiles, t1 - t0);
		}
		__syncthreads();
		for (uint32_t i = d; i < t1 - t0; i += 256) ftiles[s_at + i] = t0 + i;
		return;
	}
	uint32_t before = 0;
	for (uint32_t v = 0; v < w; ++v) before += s_wave[v];
	uint64_t run = sb + before + incl - tot;
	for (uint32_t tl = t0; tl < t1; tl += 8) {
		uint32_t c[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) c[j] = (tl + j < t1) ? counts[static_cast<uint64_t>(tl + j) * 256 + d] : 0u;
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			// offsets relative to the segment start (a segment holds at most 2^32 - 1 keys)
			if (tl + j < t1) counts[static_cast<uint64_t>(tl + j) * 256 + d] = static_cast<uint32_t>(run - sb);
			run += c[j];
		}
	}
	if (d != 0) return;
	// greedy packing of the buckets in digit order into items of at most cap keys
	uint32_t nitems = 0, acc = 0;
	for (int b = 0; b < 256; ++b) {
		const uint32_t n = s_tot[b];
		if (acc + n > cap) { ++nitems; acc = 0; }
		acc += n;
	}
	if (acc) ++nitems;
	uint32_t it = atomicAdd(&ctrl->n_sub, nitems);
	uint64_t first = sb;
	acc = 0;
	for (int b = 0; b < 256; ++b) {
		const uint32_t n = s_tot[b];
		if (acc + n > cap) {
			istart[it] = first; ilen[it] = acc; ++it;
			first += acc; acc = 0;
		}
		acc += n;
	}
	if (acc) { istart[it] = first; ilen[it] = acc; }
}

// the stable partition of the long segments by the top digit, into the same ranges of dst
template <typename K, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_seg_scatter1(const K *__restrict__ in, K *__restrict__ dst,
	const uint64_t *__restrict__ seg, const uint32_t *__restrict__ large_ids, const uint2 *__restrict__ tiles, uint32_t ntiles,
	const uint32_t *__restrict__ offsets, const uint32_t *__restrict__ flag, KeyXform xf, int reverse)
{
	constexpr int TILE = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	constexpr uint32_t kShift = 8u * (P - 1);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_glob[256];
	__shared__ uint32_t s_wsum[4];
	__shared__ K s_keys[TILE];
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	for (uint32_t g = blockIdx.x; g < ntiles; g += gridDim.x) {
		const uint2 tl = tiles[g];
		if (flag[tl.x]) continue;  // block-uniform: the segment takes the LSD passes
		const uint32_t s = large_ids[tl.x];
		const uint64_t sb = seg[s];
		const uint32_t len = static_cast<uint32_t>(seg[s + 1] - sb);
		const uint32_t base = tl.y * TILE;
		const uint32_t valid = min(static_cast<uint32_t>(TILE), len - base);
		K keys[KPT];
		uint32_t dig[KPT], pos[KPT], total;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < valid) {
				const uint32_t r = base + l;
				keys[j] = in[sb + (reverse ? len - 1 - r : r)];
				dig[j] = digit_of<K>(keys[j], xf, kShift);
			} else {
				keys[j] = K(0);
				dig[j] = 0xFFu;
			}
		}
		block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total);
		if (t < 256) s_glob[t] = offsets[static_cast<uint64_t>(g) * 256 + t] - s_digoff[t];
#pragma unroll
		for (int j = 0; j < KPT; ++j) s_keys[pos[j]] = keys[j];
		__syncthreads();
		K *dseg = dst + sb;
		for (uint32_t i = t; i < valid; i += THREADS) {
			const K k = s_keys[i];
			dseg[s_glob[digit_of<K>(k, xf, kShift)] + i] = k;
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------------
// long segments with a bucket above the LDS capacity: LSD passes

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_seg_hist(const K *__restrict__ in, const uint64_t *__restrict__ seg,
	const uint32_t *__restrict__ large_ids, const uint2 *__restrict__ tiles, const uint32_t *__restrict__ ftiles,
	const SegCtrl *ctrl, uint32_t tile_keys, KeyXform xf, uint32_t *__restrict__ hist)
{
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_h[P * 256];
	const uint32_t nf = ctrl->n_ftiles;  // tiles of the segments flagged by k_seg_scan1
	for (uint32_t q = blockIdx.x; q < nf; q += gridDim.x) {
		const uint2 tl = tiles[ftiles[q]];
		for (int i = threadIdx.x; i < P * 256; i += THREADS) s_h[i] = 0;
		__syncthreads();
		const uint32_t s = large_ids[tl.x];
		const uint64_t b = seg[s] + static_cast<uint64_t>(tl.y) * tile_keys;
		const uint64_t e = min(seg[s + 1], b + tile_keys);
		for (uint64_t i = b + threadIdx.x; i < e; i += THREADS) {
			const auto k = sort_key<K>(in[i], xf);
#pragma unroll
			for (int p = 0; p < P; ++p) wave_count(s_h + p * 256, static_cast<uint32_t>(k >> (8 * p)) & 0xFFu);
		}
		__syncthreads();
		uint32_t *h = hist + static_cast<uint64_t>(tl.x) * P * 256;
		for (int i = threadIdx.x; i < P * 256; i += THREADS)
			if (s_h[i]) atomicAdd(&h[i], s_h[i]);
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
	const uint64_t *__restrict__ seg, const uint32_t *__restrict__ large_ids, const uint2 *__restrict__ tiles,
	const uint32_t *__restrict__ ftiles, const uint32_t *__restrict__ bucket, const SegSel *__restrict__ sel,
	uint64_t *__restrict__ lookback, SegCtrl *ctrl, KeyXform xf, int pass)
{
	constexpr int TILE = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_glob[256];
	__shared__ uint32_t s_wsum[4];
	__shared__ uint32_t s_tile;
	__shared__ K s_keys[TILE];
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t shift = 8u * static_cast<uint32_t>(pass);
	const uint32_t tag = ((ctrl->epoch * kMaxPasses + static_cast<uint32_t>(pass)) % 0x3FFFFFFFu) + 1u;
	const uint32_t num_tiles = ctrl->n_ftiles;
	for (;;) {
		if (t == 0) s_tile = atomicAdd(&ctrl->tile_counter[pass], 1u);
		__syncthreads();
		const uint32_t g = s_tile;  // position in the flagged list: the look-back row of the tile
		if (g >= num_tiles) break;
		const uint2 tl = tiles[ftiles[g]];
		const SegSel sl = sel[tl.x];
		if (!(sl.mask >> pass & 1u)) {
			__syncthreads();
			continue;  // this segment skips the digit: none of its tiles takes part
		}
		const uint32_t code = sl.sel >> (4 * pass);
		const K *src = pick<K>(code & 3u, in, out, buf);
		K *dst = pick<K>(code >> 2 & 3u, in, out, buf);
		const bool rev = (sl.mask & kSelReverse) && ((sl.mask & ((1u << pass) - 1u)) == 0);  // first executed pass
		const uint32_t s = large_ids[tl.x];
		const uint64_t sb = seg[s];
		const uint32_t len = static_cast<uint32_t>(seg[s + 1] - sb);
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
		block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total);
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
			s_glob[t] = bucket[(static_cast<uint64_t>(tl.x) * P + pass) * 256 + t] + excl - s_digoff[t];
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

struct SegPlan {
	int key_bytes = 4;
	unsigned flags = 0;
	uint64_t nseg = 0, total = 0;
	int tile_cfg = 1;    // tile of the long segments' partition and LSD passes (kSegTileCfgs)
	int small_cfg = 1;   // LDS capacity tier of the item sort (kSmallCaps)
	uint32_t nsmall = 0, nlarge = 0, ntiles = 0, tile_keys = 0, small_cap = 0;
	char *mem = nullptr;  // one device allocation for every table and the workspace
	size_t mem_bytes = 0;
	SegCtrl *ctrl = nullptr;
	uint64_t *small_start = nullptr, *sub_start = nullptr;
	uint32_t *small_len = nullptr, *sub_len = nullptr, *q_small = nullptr, *q_sub = nullptr;
	uint32_t *large_ids = nullptr, *tile_first = nullptr, *counts1 = nullptr, *flag = nullptr, *ftiles = nullptr;
	uint32_t *hist = nullptr, *bucket = nullptr;
	uint2 *tiles = nullptr;
	SegSel *sel = nullptr;
	uint64_t *lookback = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipStream_t side = nullptr;
	hipEvent_t fork = nullptr, join = nullptr;
};

// tiles of the long segments' partition and LSD passes (threads x keys per thread), in the order of
// the tile_cfg argument of the segmented plan
constexpr int kNumSegTileCfgs = 5;
constexpr TileCfg kSegTileCfgs[kNumSegTileCfgs] = {{256, 8}, {256, 16}, {512, 8}, {512, 16}, {256, 24}};
// item sort capacities (256x4, 256x16, 512x16) and partition widths (about two keys per run)
constexpr int kNumSmallCaps = 3;
constexpr int kSmallCaps[kNumSmallCaps] = {256 * 4, 256 * 16, 512 * 16};

namespace {

// item sort + queued LSD items for one capacity tier; count null means count_value items
template <typename K, int THREADS, int KPT, int PB>
hipError_t launch_items_cfg(const void *src, void *dst, const uint64_t *istart, const uint32_t *ilen, const uint32_t *count,
	uint32_t count_value, const KeyXform &xf, int reverse, uint32_t *queue, uint32_t *qcount, hipStream_t s)
{
	static const int occ = occupancy_of<THREADS, KPT>(k_item_sort<K, THREADS, KPT, PB>);
	static const int occ_lsd = occupancy_of<THREADS, KPT>(k_item_lsd<K, THREADS, KPT>);
	const uint32_t cus = static_cast<uint32_t>(device_cu_count());
	hipError_t e;
	if constexpr (sizeof(K) == 1) {
		// one LSD pass ranks 8-bit keys stably by the whole key: every item goes there
		uint32_t grid = static_cast<uint32_t>(occ_lsd) * cus;
		if (!count) grid = std::min<uint32_t>(grid, std::max<uint32_t>(1u, count_value));
		hipLaunchKernelGGL((k_item_lsd<K, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(src),
			static_cast<K *>(dst), istart, ilen, nullptr, count, count_value, xf, reverse);
		return hipGetLastError();
	}
	uint32_t grid = static_cast<uint32_t>(occ) * cus;
	if (!count) grid = std::min<uint32_t>(grid, std::max<uint32_t>(1u, count_value));
	hipLaunchKernelGGL((k_item_sort<K, THREADS, KPT, PB>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(src),
		static_cast<K *>(dst), istart, ilen, count, count_value, xf, reverse, queue, qcount);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	hipLaunchKernelGGL((k_item_lsd<K, THREADS, KPT>), dim3(static_cast<uint32_t>(occ_lsd) * cus), dim3(THREADS), 0, s,
		static_cast<const K *>(src), static_cast<K *>(dst), istart, ilen, queue, qcount, 0u, xf, reverse);
	return hipGetLastError();
}

template <typename K>
hipError_t launch_items(const SegPlan &pl, const void *src, void *dst, const uint64_t *istart, const uint32_t *ilen,
	const uint32_t *count, uint32_t count_value, const KeyXform &xf, int reverse, uint32_t *queue, uint32_t *qcount,
	hipStream_t s)
{
	switch (pl.small_cfg) {
	case 0: return launch_items_cfg<K, 256, 4, 9>(src, dst, istart, ilen, count, count_value, xf, reverse, queue, qcount, s);
	case 1: return launch_items_cfg<K, 256, 16, 10>(src, dst, istart, ilen, count, count_value, xf, reverse, queue, qcount, s);
	default: return launch_items_cfg<K, 512, 16, 11>(src, dst, istart, ilen, count, count_value, xf, reverse, queue, qcount, s);
	}
}

template <typename K>
hipError_t launch_small(SegPlan &pl, const void *in, void *out, const KeyXform &xf, hipStream_t s)
{
	if (!pl.nsmall) return hipSuccess;
	return launch_items<K>(pl, in, out, pl.small_start, pl.small_len, nullptr, pl.nsmall, xf, (pl.flags & kRev) ? 1 : 0,
		pl.q_small, &pl.ctrl->q_small, s);
}

template <typename K, int THREADS, int KPT>
hipError_t launch_seg_pass(SegPlan &pl, const void *in, void *out, void *buf, const uint64_t *seg, const KeyXform &xf,
	int pass, hipStream_t s)
{
	static int occ = occupancy_of<THREADS, KPT>(k_seg_pass<K, THREADS, KPT>);
	const uint32_t grid = std::min<uint32_t>(pl.ntiles, static_cast<uint32_t>(occ * device_cu_count()));
	hipLaunchKernelGGL((k_seg_pass<K, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(in),
		static_cast<K *>(out), static_cast<K *>(buf), seg, pl.large_ids, pl.tiles, pl.ftiles, pl.bucket, pl.sel, pl.lookback,
		pl.ctrl, xf, pass);
	return hipGetLastError();
}

template <typename K, int THREADS, int KPT>
hipError_t launch_scatter1(SegPlan &pl, const void *in, void *buf, const uint64_t *seg, const KeyXform &xf, int reverse,
	hipStream_t s)
{
	static int occ = occupancy_of<THREADS, KPT>(k_seg_scatter1<K, THREADS, KPT>);
	const uint32_t grid = std::min<uint32_t>(pl.ntiles, static_cast<uint32_t>(occ * device_cu_count()));
	hipLaunchKernelGGL((k_seg_scatter1<K, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(in),
		static_cast<K *>(buf), seg, pl.large_ids, pl.tiles, pl.ntiles, pl.counts1, pl.flag, xf, reverse);
	return hipGetLastError();
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
	hipLaunchKernelGGL((k_seg_count1<K, 256>), dim3(std::min<uint32_t>(pl.ntiles, 8u * cus)), dim3(256), 0, s,
		static_cast<const K *>(in), seg, pl.large_ids, pl.tiles, pl.ntiles, pl.tile_keys, xf, rev, pl.counts1);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	hipLaunchKernelGGL(k_seg_scan1, dim3(pl.nlarge), dim3(256), 0, s, seg, pl.large_ids, pl.tile_first, pl.counts1,
		pl.small_cap, pl.flag, pl.ftiles, pl.ctrl, pl.sub_start, pl.sub_len);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	switch (pl.tile_cfg) {
	case 0: e = launch_scatter1<K, 256, 8>(pl, in, buf, seg, xf, rev, s); break;
	case 1: e = launch_scatter1<K, 256, 16>(pl, in, buf, seg, xf, rev, s); break;
	case 2: e = launch_scatter1<K, 512, 8>(pl, in, buf, seg, xf, rev, s); break;
	case 3: e = launch_scatter1<K, 512, 16>(pl, in, buf, seg, xf, rev, s); break;
	default: e = launch_scatter1<K, 256, 24>(pl, in, buf, seg, xf, rev, s); break;
	}
	if (e != hipSuccess) return e;
	if ((e = launch_items<K>(pl, buf, out, pl.sub_start, pl.sub_len, &pl.ctrl->n_sub, 0, xf, 0, pl.q_sub, &pl.ctrl->q_sub,
		s)) != hipSuccess) return e;
	// segments with a bucket above the LDS capacity: the LSD passes (every launch skips the others)
	hipLaunchKernelGGL((k_seg_hist<K, 256>), dim3(std::min<uint32_t>(pl.ntiles, 8u * cus)), dim3(256), 0, s,
		static_cast<const K *>(in), seg, pl.large_ids, pl.tiles, pl.ftiles, pl.ctrl, pl.tile_keys, xf, pl.hist);
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
	std::vector<uint2> tiles;
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
			for (uint32_t i = 0; i < nt; ++i) tiles.push_back(make_uint2(L, i));
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
	const size_t o_tiles = off; off += up256(tiles.size() * sizeof(uint2) + 8);
	const size_t o_counts = off; off += up256(tiles.size() * 256 * 4 + 4);
	const size_t o_substart = off; off += up256(nsub * 8 + 8);
	const size_t o_sublen = off; off += up256(nsub * 4 + 4);
	const size_t o_qsub = off; off += up256(nsub * 4 + 4);
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
	pl->tiles = reinterpret_cast<uint2 *>(pl->mem + o_tiles);
	pl->counts1 = reinterpret_cast<uint32_t *>(pl->mem + o_counts);
	pl->sub_start = reinterpret_cast<uint64_t *>(pl->mem + o_substart);
	pl->sub_len = reinterpret_cast<uint32_t *>(pl->mem + o_sublen);
	pl->q_sub = reinterpret_cast<uint32_t *>(pl->mem + o_qsub);
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
	put(pl->tiles, tiles.data(), tiles.size() * sizeof(uint2));
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
