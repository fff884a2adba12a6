// rsbd8_sort.hip -- gfx950 kernels of the stable 8-bit LSD radix sort.
//
// Replaces the reference's counting pass (Radixsortbidi8.hpp = H, H:32107-32473), its offsets
// transform with pass skipping (H:11626-12946) and its bidirectional scatter passes
// (H:31649-31982) with:
//   k_hist      one HBM read of the keys building every digit's 256-bin histogram at once in
//               LDS (wave-uniform digits are added with one atomic per wave);
//   k_plan      one block: per-digit exclusive scans (bucket starts), uniform-digit detection
//               (the pass-skip flag of H:11652) and the ping-pong buffer plan, all on device;
//   k_onesweep  one launch per digit: persistent blocks take tiles in order, rank digits stably
//               (LDS peer words, or ballots in skewed waves), publish per-tile digit counts and resolve the global
//               offsets with a decoupled look-back, stage the tile in LDS in digit order and
//               write coalesced runs; inactive (uniform-digit) passes exit at once;
//   k_copy      copies when no pass runs or an in-place sort needs a staging copy;
//   k_fill8     the 8-bit count fill of H:48769-48957 for bijective modes;
//   k_hist16, k_scan16, k_fill16
//               the same count fill for 16-bit keys (65536 counts) from 2^22 keys;
//   k_finish    resets the tile counters and advances the look-back epoch.
// The transform of every mode is fused into each load (rsbd8_device.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <type_traits>
#include <vector>

#include "rsbd8_engine.hpp"

namespace rsbd8hip {

namespace {

inline Control *ctrl_of(void *ws) { return reinterpret_cast<Control *>(ws); }
inline uint32_t *hist_of(void *ws) { return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kCtrlBytes); }
inline uint32_t *bucket_of(void *ws) { return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kCtrlBytes + kHistBytes); }
inline uint32_t *segtab_of(void *ws) { return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kCtrlBytes + 2 * kHistBytes); }
// look-back word (tile, digit d) at lookback_of(ws)[tile * kLbStride + d]; count (tile, d) at
// counts_of(ws)[tile * kCountStride + d] (the tile records of the workspace layout)
inline uint64_t *lookback_of(void *ws) { return reinterpret_cast<uint64_t *>(static_cast<char *>(ws) + kHeaderBytes); }
inline uint32_t *counts_of(void *ws)
{
	return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kHeaderBytes + kRadix * sizeof(uint64_t));
}

__device__ __forceinline__ uint32_t pass_tag(uint32_t epoch, int slot)
{
	return ((epoch * kTagSlots + static_cast<uint32_t>(slot)) % 0x3FFFFFFFu) + 1u;
}

constexpr uint32_t kJointBins = 65536;    // bins of the top-16-bit histogram
constexpr uint32_t kJointFlush = 32768;   // keys counted between two overflow sweeps of that histogram
constexpr uint32_t kRanges = 8;           // key ranges of the MSD plan's first partition (one per XCD)

// The MSD plan's tables (see kSegTableWords): tile_start[257], seg_start[256], seg_len[256],
// bucket_start[65536] (relative to the bucket's segment), bucket_len[65536], queue[65536],
// cursor[65536], ovf[65536], rcnt[8 * 256], bq[3 * 65536].
struct SegTables {
	uint32_t *tile_start;
	uint32_t *seg_start;
	uint32_t *seg_len;
	uint32_t *bucket_start;
	uint32_t *bucket_len;
	uint32_t *queue;
	uint32_t *cursor;
	uint32_t *ovf;
	uint32_t *rcnt;
	// the sample partition plan's bucket queues, entries {absolute start, length, bucket (segment *
	// 256 + id), equal keys}: [0, 65536) the buckets of at most kBucketCapA keys, [65536, 131072)
	// the others and the equality buckets, [131072, 196608) those the LDS sorts pass on
	uint4 *bq;
};
// the splitter tables (kSplitTables of kSplitTabWords words each): the first partition's, then one
// per first-partition segment
inline uint32_t *splittab_of(void *ws)
{
	return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kCtrlBytes + 2 * kHistBytes + kSegTableBytes);
}

// stages a splitter table in LDS (bounds narrowed to the transformed key's width); every thread of
// the block takes part
template <typename W, int THREADS>
__device__ __forceinline__ void split_load(const uint32_t *__restrict__ tab, SplitLds<W> &tb)
{
	const uint64_t *b = reinterpret_cast<const uint64_t *>(tab + kSplitBoundsWord);
	for (int i = threadIdx.x; i < kRadix; i += THREADS) {
		tb.b[i] = static_cast<W>(b[i]);
		tb.info[i] = tab[kSplitInfoWord + i];
	}
	const uint32_t *c = tab + kSplitCellWord;  // the cells, then the sub-tables in use
	uint32_t *dc = reinterpret_cast<uint32_t *>(tb.cell);
	uint32_t *ds = reinterpret_cast<uint32_t *>(tb.sub);
	const uint32_t nsub = min(tab[kSplitNsubWord], kSplitSubs);
	const int nsub_words = static_cast<int>(nsub * (kSplitSubCells / 2));
	for (int i = threadIdx.x; i < static_cast<int>(kSplitCells / 2); i += THREADS) dc[i] = c[i];
	for (int i = threadIdx.x; i < nsub_words; i += THREADS) ds[i] = c[kSplitCells / 2 + i];
	if (threadIdx.x == 0) {
		tb.lo = static_cast<W>(*reinterpret_cast<const uint64_t *>(tab + kSplitLoWord));
		tb.hi = static_cast<W>(*reinterpret_cast<const uint64_t *>(tab + kSplitHiWord));
		tb.sh = tab[kSplitShWord];
		tb.nsub = nsub;
	}
}

__host__ __device__ inline SegTables seg_tables(uint32_t *base)
{
	uint32_t *b = base + 257 + 512;
	uint32_t *q = b + 5 * kHybBuckets + 8 * kRadix;
	q += (4 - (reinterpret_cast<uintptr_t>(q) / 4) % 4) % 4;  // 16-byte entries (the base is 256-byte aligned)
	return SegTables{base, base + 257, base + 257 + 256, b, b + kHybBuckets, b + 2 * kHybBuckets, b + 3 * kHybBuckets,
		b + 4 * kHybBuckets, b + 5 * kHybBuckets, reinterpret_cast<uint4 *>(q)};
}

// 0: a real sort is needed; 1: the output is the input as it is; 2: the output is the input
// reversed (physical order). For isrevorder the logical input is the physical one read backwards.
// The presort check's verdict (check: bit 0 ran, bits 1-3 an ascent, a tie, a descent seen between
// physical neighbours, bit 4 the sort may take the run reversal): 1 the input is in order (a copy),
// 2 it is strictly in reverse order (a reversed copy), 3 it is in reverse order with ties (the runs
// of equal keys reversed as blocks, each in its own order: k_runs_*), 0 it needs sorting.
__device__ __forceinline__ uint32_t presort_result(uint32_t check, int reverse)
{
	if (!(check & 1u)) return 0;
	const bool lt = check & 2u, eq = check & 4u, gt = check & 8u, runs = check & 16u;
	if (!reverse) return !gt ? 1u : (!lt ? (!eq ? 2u : (runs ? 3u : 0u)) : 0u);
	return !lt ? 2u : (!gt ? (!eq ? 1u : (runs ? 3u : 0u)) : 0u);
}

// the verdict is "in order" (a copy) or "strictly reversed" (a reversed copy), and the check wrote
// that copy to dst while it read the input (every wave in that direction only)
__device__ __forceinline__ bool spec_done(const Control *ctrl, uint32_t pres, const void *dst, bool payload)
{
	const uint32_t dirs = ctrl->check & 96u;
	return ((pres == 1u && dirs == 32u) || (pres == 2u && dirs == 64u)) && ctrl->spec_dst == dst && !payload;
}

// ---------------------------------------------------------------------------------------------
// histogram of every digit in one read

template <typename K, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_hist(const K *__restrict__ in, uint64_t n, KeyXform xf,
	uint32_t *__restrict__ hist, const Control *ctrl, int reverse)
{
	if (presort_result(ctrl->check, reverse)) return;  // the input is already in order
	constexpr int P = sizeof(K);
	constexpr int C = count_copies(P);
	__shared__ uint32_t s_h[P * kRadix * C];
	for (int i = threadIdx.x; i < P * kRadix * C; i += THREADS) s_h[i] = 0;
	__syncthreads();
	// one contiguous, 64-key aligned range per block
	const uint64_t chunk = ((n + gridDim.x - 1) / gridDim.x + 63) & ~63ull;
	const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * chunk;
	if (lo < n) {
		const uint32_t len = static_cast<uint32_t>((n - lo) < chunk ? n - lo : chunk);
		for_each_key<K, THREADS>(in, lo, len, [&](K x) {
			const auto k = sort_key<K>(x, xf);
#pragma unroll
			for (int p = 0; p < P; ++p) wave_count<C>(s_h + p * kRadix * C, static_cast<uint32_t>(k >> (8 * p)) & 0xFFu);
		});
	}
	__syncthreads();
	for (int i = threadIdx.x; i < P * kRadix; i += THREADS) {
		const uint32_t v = copies_sum<C>(s_h, i);
		if (v) atomicAdd(&hist[i], v);
	}
}


// ---------------------------------------------------------------------------------------------
// presort check: is the input already in order, or in reverse order?

// The sort key as an ordered value: sort_key leaves bits above a 1- or 2-byte key's width undefined
// (the tiered modes shift the sign bit up there), which the digit passes never read but an order
// comparison would.
template <typename K>
__device__ __forceinline__ typename KeyTraits<K>::W order_key(K raw, const KeyXform &xf)
{
	using W = typename KeyTraits<K>::W;
	constexpr W mask = sizeof(K) >= sizeof(W) ? ~W(0) : static_cast<W>((W(1) << (8 * sizeof(K))) - 1u);
	return sort_key<K>(raw, xf) & mask;
}

// The run reversal (presort verdict 3: the input, read in its logical order, never ascends but has
// ties). The stable order is then the runs of equal keys in reverse order, each run in its own
// order: a key at physical position j of the run [a, b] goes to n - 1 - b + (j - a), or, when the
// input is read back to front, to a + (b - j). Tiles of kRunTile keys: the check records every
// wave's last run start and first run end (a wave of the check reads a contiguous quarter of a
// tile), k_runs_reduce and k_runs_carry carry them across tiles (the start of the run that enters
// a tile, the end of the run that leaves it), k_runs_copy places every key. The words live in the
// workspace's count rows, planar (word k of tile c at k * tiles + c, see run_word): eight wave
// words (start, end), the two carries, then two words per block of the carry kernels.
constexpr uint32_t kRunTile = 4096;
constexpr uint32_t kRunWaves = 4;          // waves of the check's blocks
constexpr uint32_t kRunWaveKeys = kRunTile / kRunWaves;
constexpr uint32_t kRunCarry = 2 * kRunWaves;
constexpr uint32_t kRunAgg = kRunCarry + 2;
constexpr uint32_t kRunBlock = 1024;       // tiles per block of the carry kernels

constexpr uint32_t kCheckSlots = 4096;   // the check's per-wave verdict words come first (at most 1024 blocks)

__device__ __forceinline__ uint32_t *count_word(uint32_t *counts, uint32_t v)
{
	return counts + static_cast<uint64_t>(v >> 8) * kCountStride + (v & 255u);
}
__device__ __forceinline__ uint32_t *run_word(uint32_t *counts, uint32_t v) { return count_word(counts, kCheckSlots + v); }

template <typename K>
__device__ __forceinline__ K lane0_of(K v)
{
	if constexpr (sizeof(K) == 8) {
		const uint64_t u = static_cast<uint64_t>(v);
		const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u));
		const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u >> 32));
		return static_cast<K>((static_cast<uint64_t>(hi) << 32) | lo);
	} else {
		return static_cast<K>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v)));
	}
}

// Tiles of kRunTile keys, blocks taking tiles in grid stride. Wave w reads the contiguous quarter w
// of a tile, key i of lane l at offset i * 64 + l (coalesced): the key after it is lane l + 1's, or
// lane 0's key i + 1, and for the wave's last key the first key after the quarter, loaded with the
// others (as are the key before the quarter and the check word, so every iteration waits for memory
// once). A wave stops once an ascent and a descent were seen anywhere (then no shortcut is possible,
// and no run word is needed), and publishes only what it had not published yet. With runs set,
// every wave records the last run start and the first run end of its quarter (a run starts right
// after an end), as bit masks over its keys: bounds kept in the loop by conditional updates lost the
// first iteration's update in the gfx950 build, so the run kernels use masks and selects only.
// With spec set (keys-only sorts to another array), a wave also copies its keys to spec, in order
// (spec[j]) while nothing it saw contradicts the verdict "in order", or reversed (spec[n - 1 - j])
// while nothing contradicts "strictly reversed" (read back to front: reversed with ties), deciding
// per tile from what it saw before and the tile's first row; its slot records the directions it
// wrote (bit 5 in order, bit 6 reversed). When the verdict is 1 or 2 and every wave wrote only that
// verdict's direction, every key is in place and the final copy is skipped; on unsorted input the
// waves stop copying within their first tile.
template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_check(const K *__restrict__ in, uint64_t n, KeyXform xf, Control *ctrl,
	uint32_t *__restrict__ words, int runs, uint32_t tiles, K *__restrict__ spec, int reverse)
{
	constexpr uint32_t PER = kRunWaveKeys / 64;
	static_assert(THREADS / 64 == kRunWaves && PER * 64 == kRunWaveKeys, "a wave reads a quarter tile");
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		if (spec) ctrl->spec_dst = spec;
		atomicOr(&ctrl->check, runs ? 17u : 1u);
	}
	if (n < 2) return;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint64_t pairs = n - 1;
	const uint32_t last = static_cast<uint32_t>(n - 1);
	uint32_t *slot = count_word(words, blockIdx.x * kRunWaves + w);
	uint32_t seen = 0;   // what this wave saw (wave-uniform), stored to its slot when it ends
	uint32_t wrote = 0;  // the directions this wave copied in (bits 5 and 6)
	uint32_t it = 0;
	for (uint64_t tb = static_cast<uint64_t>(blockIdx.x) * kRunTile; tb < n; tb += static_cast<uint64_t>(gridDim.x) * kRunTile, ++it) {
		const uint64_t base = tb + w * kRunWaveKeys;
		// another wave's stop, every 16 tiles (loads this CU's cache may serve a little late)
		uint32_t stop = 0;
		if ((it & 15u) == 0) {
			const uint32_t v = lane < kStopLines
				? __hip_atomic_load(&ctrl->check_stop[lane][0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0u;
			stop = __ballot(v != 0) ? 1u : 0u;
		}
		K k[PER];
#pragma unroll
		for (uint32_t i = 0; i < PER; ++i) {
			const uint64_t j = base + i * 64 + lane;
			k[i] = j < n ? in[j] : K(0);
		}
		const uint64_t qend = min(n, base + kRunWaveKeys);
		K tail = K(0), head = K(0);  // the key after the quarter (lane 63), the key before it (lane 0)
		if (lane == 63 && qend < n) tail = in[qend];
		if (lane == 0 && base > 0 && base < n) head = in[base - 1];
		uint32_t mine = 0, endm = 0, stm = 0;
		// the copy's direction for this tile, from the earlier tiles and the tile's first row (a tile
		// whose later rows contradict it is still copied: the sort that follows overwrites it), key by
		// key, so no key stays live past its compare
		bool fwd = false, bwd = false;
		if (spec && base < n) {  // (a quarter past the last key records no direction)
			const auto a0 = order_key<K>(k[0], xf), b0 = order_key<K>(static_cast<K>(__shfl_down(k[0], 1, 64)), xf);
			const bool pv = lane < 63 && base + lane + 1 < n;
			const uint32_t pre = seen | (__ballot(pv && a0 < b0) ? 2u : 0u) | (__ballot(pv && a0 == b0) ? 4u : 0u) |
				(__ballot(pv && a0 > b0) ? 8u : 0u);
			fwd = !(pre & 8u) && !(reverse && (pre & 4u));
			bwd = !fwd && !(pre & 2u) && !(!reverse && (pre & 4u));
			wrote |= (fwd ? 32u : 0u) | (bwd ? 64u : 0u);
		}
#pragma unroll
		for (uint32_t i = 0; i < PER; ++i) {
			const uint64_t j = base + i * 64 + lane;
			if ((fwd || bwd) && j < n) spec[fwd ? j : n - 1 - j] = k[i];
			// lane 0's key i + 1, read by the whole wave (the first active lane must be lane 0)
			const K down = __shfl_down(k[i], 1, 64), first = lane0_of<K>(k[i + 1 < PER ? i + 1 : i]);
			const K nx = lane < 63 ? down : (i + 1 < PER ? first : tail);
			const auto a = order_key<K>(k[i], xf), b = order_key<K>(nx, xf);
			const uint32_t r = (a < b ? 2u : 0u) | (a == b ? 4u : 0u) | (a > b ? 8u : 0u);
			mine |= j < pairs ? r : 0u;
			const bool e = (j < n) & ((j + 1 == n) | (a != b));
			endm |= static_cast<uint32_t>(e) << i;
			stm |= static_cast<uint32_t>(e & (j + 1 < qend)) << i;
		}
		seen |= (__ballot((mine & 2u) != 0) ? 2u : 0u) | (__ballot((mine & 4u) != 0) ? 4u : 0u) |
			(__ballot((mine & 8u) != 0) ? 8u : 0u);

		if (runs) {
			uint32_t en = endm ? static_cast<uint32_t>(base + (__ffs(endm) - 1u) * 64 + lane) : last;
			uint32_t st = stm ? static_cast<uint32_t>(base + (31u - __clz(stm)) * 64 + lane + 1u) : 0u;
			if (lane == 0 && base > 0 && base < n) {  // the quarter's first key starts a run when it differs from the one before
				const bool s0 = order_key<K>(head, xf) != order_key<K>(k[0], xf);
				st = s0 ? max(st, static_cast<uint32_t>(base)) : st;
			}
#pragma unroll
			for (int o = 32; o > 0; o >>= 1) {
				st = max(st, static_cast<uint32_t>(__shfl_xor(st, o, 64)));
				en = min(en, static_cast<uint32_t>(__shfl_xor(en, o, 64)));
			}
			const uint32_t tile = static_cast<uint32_t>(tb / kRunTile);
			if (lane == 0) {
				*run_word(words, (2 * w) * tiles + tile) = st;
				*run_word(words, (2 * w + 1) * tiles + tile) = en;
			}
		}
		if ((seen & 10u) == 10u) {  // no shortcut is possible
			if (lane == 0 && !stop)
				__hip_atomic_store(&ctrl->check_stop[blockIdx.x % kStopLines][0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			break;
		}
		if (stop) break;
	}
	if (lane == 0) *slot = seen | wrote;
}

// The verdict: every wave's slot of the check, or'd into the control block (one atomic).
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_check_fold(Control *ctrl, uint32_t *__restrict__ words, uint32_t slots)
{
	__shared__ uint32_t s_w[THREADS / 64];
	uint32_t v = 0;
	for (uint32_t q = threadIdx.x; q < slots; q += THREADS) v |= *count_word(words, q);
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v |= static_cast<uint32_t>(__shfl_xor(v, o, 64));
	if ((threadIdx.x & 63u) == 0) s_w[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int q = 1; q < THREADS / 64; ++q) v |= s_w[q];
		if (v) atomicOr(&ctrl->check, v);
	}
}

// Block-wide maximum / minimum over the 1024 tiles of a carry block, with every thread's value
// combined with the threads before it (INCL), or the reduction only.
template <int THREADS, bool MAX>
__device__ __forceinline__ uint32_t block_scan_mm(uint32_t v, uint32_t *s_w, uint32_t &total)
{
	constexpr int NW = THREADS / 64;
	const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	uint32_t incl = v;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t u = __shfl_up(incl, o, 64);
		incl = lane >= static_cast<uint32_t>(o) ? (MAX ? max(incl, u) : min(incl, u)) : incl;
	}
	if (lane == 63) s_w[w] = incl;
	__syncthreads();
	uint32_t before = MAX ? 0u : ~0u, all = before;
	for (int q = 0; q < NW; ++q) {
		const uint32_t x = s_w[q];
		before = q < static_cast<int>(w) ? (MAX ? max(before, x) : min(before, x)) : before;
		all = MAX ? max(all, x) : min(all, x);
	}
	__syncthreads();
	total = all;
	return MAX ? max(before, incl) : min(before, incl);
}

// Every carry block's largest run start and smallest run end (its tiles' wave words combined).
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_runs_reduce(const Control *ctrl, int reverse, uint64_t n, uint32_t *__restrict__ counts,
	uint32_t tiles)
{
	if (presort_result(ctrl->check, reverse) != 3u) return;
	__shared__ uint32_t s_w[THREADS / 64];
	const uint32_t c = blockIdx.x * THREADS + threadIdx.x;
	const uint32_t last = static_cast<uint32_t>(n - 1);
	uint32_t st = 0, en = last;
	if (c < tiles) {
#pragma unroll
		for (uint32_t q = 0; q < kRunWaves; ++q) {
			st = max(st, *run_word(counts, (2 * q) * tiles + c));
			en = min(en, *run_word(counts, (2 * q + 1) * tiles + c));
		}
	}
	uint32_t smax, emin;
	block_scan_mm<THREADS, true>(st, s_w, smax);
	block_scan_mm<THREADS, false>(en, s_w, emin);
	if (threadIdx.x == 0) {
		*run_word(counts, kRunAgg * tiles + blockIdx.x) = smax;
		*run_word(counts, kRunAgg * tiles + gridDim.x + blockIdx.x) = emin;
	}
}

// carry_start of tile c (word kRunCarry) = the largest run start of the tiles before c (0: position
// 0 starts a run), carry_end (word kRunCarry + 1) = the smallest run end of the tiles after c (n - 1
// ends one): the blocks before and after from the block words, then a block-wide scan each way.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_runs_carry(const Control *ctrl, int reverse, uint64_t n, uint32_t *__restrict__ counts,
	uint32_t tiles)
{
	if (presort_result(ctrl->check, reverse) != 3u) return;
	__shared__ uint32_t s_w[THREADS / 64];
	const uint32_t t = threadIdx.x, b = blockIdx.x, nb = gridDim.x;
	const uint32_t c = b * THREADS + t;
	const uint32_t last = static_cast<uint32_t>(n - 1);
	// the other blocks' words (nb <= THREADS * 4: up to four per thread)
	uint32_t pmax = 0, smin = last;
	for (uint32_t q = t; q < nb; q += THREADS) {
		const uint32_t sa = *run_word(counts, kRunAgg * tiles + q), ea = *run_word(counts, kRunAgg * tiles + nb + q);
		pmax = q < b ? max(pmax, sa) : pmax;
		smin = q > b ? min(smin, ea) : smin;
	}
	uint32_t bmax, bmin;
	block_scan_mm<THREADS, true>(pmax, s_w, bmax);
	block_scan_mm<THREADS, false>(smin, s_w, bmin);
	uint32_t st = 0, en = last;
	if (c < tiles) {
#pragma unroll
		for (uint32_t q = 0; q < kRunWaves; ++q) {
			st = max(st, *run_word(counts, (2 * q) * tiles + c));
			en = min(en, *run_word(counts, (2 * q + 1) * tiles + c));
		}
	}
	uint32_t unused;
	const uint32_t incl = block_scan_mm<THREADS, true>(st, s_w, unused);
	const uint32_t lane = t & 63u;
	// exclusive: the inclusive value of the thread before (a shuffle, or the wave before's total)
	__shared__ uint32_t s_last[THREADS / 64];
	if (lane == 63) s_last[t >> 6] = incl;
	__syncthreads();
	const uint32_t upv = __shfl_up(incl, 1, 64);
	const uint32_t prev = lane ? upv : ((t >> 6) ? s_last[(t >> 6) - 1] : 0u);
	// the suffix minimum: the same scan over the block's tiles in reverse order
	const uint32_t rc = b * THREADS + (THREADS - 1 - t);
	uint32_t ren = last;
	if (rc < tiles) {
#pragma unroll
		for (uint32_t q = 0; q < kRunWaves; ++q) ren = min(ren, *run_word(counts, (2 * q + 1) * tiles + rc));
	}
	const uint32_t rincl = block_scan_mm<THREADS, false>(ren, s_w, unused);
	__shared__ uint32_t s_rlast[THREADS / 64];
	if (lane == 63) s_rlast[t >> 6] = rincl;
	__syncthreads();
	const uint32_t rupv = __shfl_up(rincl, 1, 64);
	const uint32_t rprev = lane ? rupv : ((t >> 6) ? s_rlast[(t >> 6) - 1] : last);
	if (c < tiles) *run_word(counts, kRunCarry * tiles + c) = max(bmax, prev);
	if (rc < tiles) *run_word(counts, (kRunCarry + 1) * tiles + rc) = min(bmin, rprev);
}

__device__ __forceinline__ uint32_t pad16(uint32_t x) { return x + (x >> 4); }

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_runs_copy(const K *__restrict__ in, K *__restrict__ out, uint64_t n, KeyXform xf,
	const Control *ctrl, int reverse, uint32_t *__restrict__ counts, uint32_t tiles)
{
	if (presort_result(ctrl->check, reverse) != 3u) return;
	constexpr uint32_t PER = kRunTile / THREADS;
	constexpr int NW = THREADS / 64;
	// the tile, with the key before it and the key after it; then the output order. Index x lives at
	// x + x / 16 (pad): a thread's 16 consecutive keys and the threads' 16-key strides fall in
	// different banks.
	__shared__ K s_k[kRunTile + 2 + (kRunTile + 2) / 16];
	__shared__ uint32_t s_wmax[NW], s_wmin[NW];
	__shared__ uint32_t s_run[4];   // the runs of the tile's first and last key: a0, b0, a1, b1
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t last = static_cast<uint32_t>(n - 1);
	const uint32_t tile = blockIdx.x;
	const uint64_t base = static_cast<uint64_t>(tile) * kRunTile;
	const uint32_t tcount = static_cast<uint32_t>(min<uint64_t>(n - base, kRunTile));  // keys in the tile
	{
		// coalesced, every load in flight at once: s_k[i] = in[base - 1 + i]
		constexpr uint32_t LOADS = (kRunTile + 2 + THREADS - 1) / THREADS;
		K v[LOADS];
#pragma unroll
		for (uint32_t q = 0; q < LOADS; ++q) {
			const uint32_t i = q * THREADS + t;
			const uint64_t j = base + i;  // physical position + 1
			v[q] = (i < kRunTile + 2 && j >= 1 && j - 1 < n) ? in[j - 1] : K(0);
		}
#pragma unroll
		for (uint32_t q = 0; q < LOADS; ++q) {
			const uint32_t i = q * THREADS + t;
			if (i < kRunTile + 2) s_k[pad16(i)] = v[q];
		}
	}
	const uint32_t carry_st = *run_word(counts, kRunCarry * tiles + tile);
	const uint32_t carry_end = *run_word(counts, (kRunCarry + 1) * tiles + tile);
	__syncthreads();
	const uint64_t j0 = base + t * PER;
	// bit i: position j0 + i starts / ends a run, from the sort keys of positions j0 - 1 .. j0 + PER
	// (s_k[t * PER + i] holds position j0 - 1 + i); position 0 starts a run, n - 1 ends one, and
	// positions from n on do neither
	using W = typename KeyTraits<K>::W;
	W sk[PER + 2];
#pragma unroll
	for (uint32_t i = 0; i < PER + 2; ++i) sk[i] = order_key<K>(s_k[pad16(t * PER + i)], xf);
	uint32_t stf = 0, enf = 0;
#pragma unroll
	for (uint32_t i = 0; i < PER; ++i) {
		stf |= static_cast<uint32_t>(sk[i] != sk[i + 1]) << i;
		enf |= static_cast<uint32_t>(sk[i + 1] != sk[i + 2]) << i;
	}
	stf |= j0 == 0 ? 1u : 0u;
	const uint32_t valid = j0 >= n ? 0u : static_cast<uint32_t>(min<uint64_t>(n - j0, PER));  // keys below n
	const uint32_t vmask = valid >= 32u ? ~0u : (1u << valid) - 1u;
	stf &= vmask;
	enf = (enf | (j0 + valid == n && valid ? 1u << (valid - 1u) : 0u)) & vmask;
	const uint32_t mx = stf ? static_cast<uint32_t>(j0) + 31u - __clz(stf) : 0u;
	const uint32_t mn = enf ? static_cast<uint32_t>(j0) + __ffs(enf) - 1u : last;
	// the largest run start before this thread's keys, the smallest run end after them
	uint32_t imx = mx, imn = mn;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t a = __shfl_up(imx, o, 64), b = __shfl_down(imn, o, 64);
		imx = lane >= static_cast<uint32_t>(o) ? max(imx, a) : imx;
		imn = lane + o < 64u ? min(imn, b) : imn;
	}
	if (lane == 63) s_wmax[w] = imx;
	if (lane == 0) s_wmin[w] = imn;
	K kv[PER];
#pragma unroll
	for (uint32_t i = 0; i < PER; ++i) kv[i] = s_k[pad16(t * PER + i + 1)];
	__syncthreads();
	uint32_t before = carry_st, after = carry_end;
	for (int v = 0; v < NW; ++v) {
		before = v < static_cast<int>(w) ? max(before, s_wmax[v]) : before;
		after = v > static_cast<int>(w) ? min(after, s_wmin[v]) : after;
	}
	const uint32_t up = __shfl_up(imx, 1, 64), down = __shfl_down(imn, 1, 64);
	before = lane > 0 ? max(before, up) : before;
	after = lane < 63 ? min(after, down) : after;
	uint32_t endpos[PER];  // the ends of this thread's keys' runs, from the last key back
	{
		uint32_t e = after;
#pragma unroll
		for (int i = static_cast<int>(PER) - 1; i >= 0; --i) {
			e = (enf >> i) & 1u ? static_cast<uint32_t>(j0 + i) : e;
			endpos[i] = e;
		}
	}
	uint32_t startpos[PER];
	{
		uint32_t a = before;
#pragma unroll
		for (uint32_t i = 0; i < PER; ++i) {
			a = (stf >> i) & 1u ? static_cast<uint32_t>(j0 + i) : a;
			startpos[i] = a;
		}
	}
	const uint32_t lk = tcount - 1;  // the tile's last key (local index)
	if (t == 0) { s_run[0] = startpos[0]; s_run[1] = endpos[0]; }
	{
		uint32_t sa = 0, sb = 0;
#pragma unroll
		for (uint32_t i = 0; i < PER; ++i) {
			sa = i == lk % PER ? startpos[i] : sa;
			sb = i == lk % PER ? endpos[i] : sb;
		}
		if (t == lk / PER) { s_run[2] = sa; s_run[3] = sb; }
	}
	__syncthreads();
	// the three output ranges in output order (sizes may be 0): slot p of s_k holds output position
	// orig[r] + (p - off[r]) for range r with off[r] <= p < off[r + 1]
	const uint32_t a0 = s_run[0], b0 = s_run[1], a1 = s_run[2], b1 = s_run[3];
	const uint32_t tb = static_cast<uint32_t>(base), te = tb + lk;
	const bool multi = a1 > b0;  // the first and the last key lie in different runs
	const uint32_t lenE = min(b0, te) - tb + 1, lenL = multi ? te - a1 + 1 : 0u, lenI = multi ? a1 - 1 - b0 : 0u;
	uint32_t off1, off2, orig0, orig1, orig2;
	if (!reverse) {  // the last key's run first, then the inside, then the first key's run
		off1 = lenL; off2 = lenL + lenI;
		orig0 = last - b1; orig1 = last - a1 + 1; orig2 = last - b0 + (tb - a0);
	} else {         // runs keep their places: the first key's run, the inside, the last key's run
		off1 = lenE; off2 = lenE + lenI;
		orig0 = a0 + b0 - min(b0, te); orig1 = b0 + 1; orig2 = a1 + b1 - te;
	}
#pragma unroll
	for (uint32_t i = 0; i < PER; ++i) {
		const uint32_t jj = static_cast<uint32_t>(j0 + i), a = startpos[i], b = endpos[i];
		const uint32_t o = reverse ? a + (b - jj) : last - b + (jj - a);
		const bool inE = jj <= b0, inL = multi && jj >= a1;
		const uint32_t slot = inE ? (reverse ? o - orig0 : off2 + (o - orig2))
			: (inL ? (reverse ? off2 + (o - orig2) : o - orig0) : off1 + (o - orig1));
		if (i < valid && slot < kRunTile) s_k[pad16(slot)] = kv[i];
	}
	__syncthreads();
	for (uint32_t p = t; p < tcount; p += THREADS) {
		const uint32_t o = p < off1 ? orig0 + p : (p < off2 ? orig1 + (p - off1) : orig2 + (p - off2));
		if (o < n) out[o] = s_k[pad16(p)];  // (always: the ranges partition the output)
	}
}

// One block sorts the whole array in LDS (8-bit LSD passes with the block ranking). With PAY a u32
// payload (pin, or the source index when pin is null) moves with the keys. Every load happens
// before any store, so in place is fine for keys and payload alike.
template <typename K, int THREADS, int KPT, bool PAY>
__global__ __launch_bounds__(THREADS) void k_lds_sort(const K *in, K *out, uint32_t n, KeyXform xf, int reverse,
	const uint32_t *pin, uint32_t *pout)
{
	constexpr int CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_wsum[4];
	// the keys' exchange, the payload's, and during the ranking the peer words
	__shared__ __align__(16) unsigned char s_raw[max_bytes(max_bytes(CAP * sizeof(K), PAY ? CAP * 4 : 0), rank_peer_bytes(THREADS))];
	K *s_keys = reinterpret_cast<K *>(s_raw);
	uint32_t *s_pay = reinterpret_cast<uint32_t *>(s_raw);
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_raw);
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	K keys[KPT];
	uint32_t pay[PAY ? KPT : 1];
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		const uint32_t sidx = reverse ? n - 1 - l : l;
		keys[j] = l < n ? in[sidx] : K(0);
		if constexpr (PAY) pay[j] = l < n ? (pin ? pin[sidx] : sidx) : 0u;
	}
	// digits on which every sort key agrees need no pass (the pass skip of H:11652): the AND and
	// the OR of the block's sort keys agree on them
	using W = typename KeyTraits<K>::W;
	__shared__ W s_and[NW], s_or[NW];
	W k_and = ~W(0), k_or = 0;
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		if (l < n) {
			const W k = sort_key<K>(keys[j], xf);
			k_and &= k;
			k_or |= k;
		}
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		k_and &= static_cast<W>(__shfl_xor(k_and, o, 64));
		k_or |= static_cast<W>(__shfl_xor(k_or, o, 64));
	}
	if (lane == 0) { s_and[w] = k_and; s_or[w] = k_or; }
	__syncthreads();  // also: every load is done before any store (in place)
	// block-wide AND and OR first: a digit can be constant inside every wave's chunk and still
	// differ between chunks (waves with no key contribute the neutral ~0 and 0)
	W and_all = ~W(0), or_all = 0;
#pragma unroll
	for (int v = 0; v < NW; ++v) {
		and_all &= s_and[v];
		or_all |= s_or[v];
	}
	const W differ = and_all ^ or_all;
	for (int p = 0; p < P; ++p) {
		if (((differ >> (8 * p)) & 0xFFu) == 0) continue;  // block-uniform
		uint32_t dig[KPT], pos[KPT], total;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			dig[j] = l < n ? digit_of<K>(keys[j], xf, 8u * p) : 0xFFu;
		}
		block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
#pragma unroll
		for (int j = 0; j < KPT; ++j) s_keys[pos[j]] = keys[j];
		__syncthreads();
#pragma unroll
		for (int j = 0; j < KPT; ++j) keys[j] = s_keys[w * 64u * KPT + j * 64u + lane];
		__syncthreads();
		if constexpr (PAY) {
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_pay[pos[j]] = pay[j];
			__syncthreads();
#pragma unroll
			for (int j = 0; j < KPT; ++j) pay[j] = s_pay[w * 64u * KPT + j * 64u + lane];
			__syncthreads();
		}
	}
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		if (l < n) {
			out[l] = keys[j];
			if constexpr (PAY) pout[l] = pay[j];
		}
	}
}

// ---------------------------------------------------------------------------------------------
// plan: bucket starts, pass skipping and buffer assignment

struct PlanArgs {
	uint64_t n;
	int passes;
	const void *src;
	void *fin;
	void *other;
	const uint32_t *psrc;  // null: payload is the source index
	uint32_t *pfin;
	uint32_t *pother;
	int payload;
	int reverse;
	int keys_final_any;
	int counts0;
	int hybrid;         // the read was the two-level MSD plan's (k_hist_joint, k_joint_reduce: the
	                    // histograms of the top two digits and the exact bucket sizes), so that plan
	                    // may be chosen; the LSD plan is then incomplete, and its gated full read and
	                    // plan redo it when the MSD plan is not chosen
	uint32_t hyb_cap;   // LDS capacity (keys) of the per-bucket sort
	uint32_t tile2;     // tile size of the segmented second pass
	uint32_t *segtab;   // the MSD plan's tables
	const uint32_t *gate;  // non-null: plan only when it reads nonzero (the full plan after a fallback)
	int group;          // sample partition plan, keys only: equal sort keys are equal keys (1), or are
	                    // unless both +0.0 and -0.0 occur (2), so the partitions may group instead of
	                    // rank; 0: they rank
	const uint32_t *split0;  // sample partition plan: the first splitter table (its equality buckets skip the second level)
};

__global__ __launch_bounds__(256) void k_plan(PlanArgs a, Control *ctrl, uint32_t *hist, uint32_t *bucket)
{
	__shared__ uint32_t s_wave[4];
	__shared__ uint32_t s_uniform[kMaxPasses];
	if (a.gate && !*a.gate) return;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	uint32_t top_count = 0, top_start = 0;
	for (int p = 0; p < a.passes; ++p) {
		const uint32_t c = hist[p * kRadix + t];
		hist[p * kRadix + t] = 0;  // leave the histogram zeroed for the next sort
		const uint32_t incl = wave_inclusive_sum(c);
		if (lane == 63) s_wave[w] = incl;
		const bool full = static_cast<uint64_t>(c) == a.n;
		if (t == 0) s_uniform[p] = 0;
		__syncthreads();
		uint32_t before = 0;
		for (uint32_t v = 0; v < w; ++v) before += s_wave[v];
		bucket[p * kRadix + t] = before + incl - c;
		if (p == a.passes - 1) { top_count = c; top_start = before + incl - c; }
		if (full) s_uniform[p] = 1;
		__syncthreads();
	}
	if (a.hybrid) {
		// the MSD plan's segment map (used only if thread 0 below chooses that plan): tiles of the
		// second pass never straddle a top-digit segment
		const SegTables T = seg_tables(a.segtab);
		const uint32_t tiles = (top_count + a.tile2 - 1) / a.tile2;
		const uint32_t incl = wave_inclusive_sum(tiles);
		if (lane == 63) s_wave[w] = incl;
		__syncthreads();
		uint32_t before = 0;
		for (uint32_t v = 0; v < w; ++v) before += s_wave[v];
		T.tile_start[t] = before + incl - tiles;
		if (t == kRadix - 1) T.tile_start[kRadix] = before + incl;
		T.seg_start[t] = top_start;
		T.seg_len[t] = top_count;
		// the first partition's cursors: range x's keys of top digit t follow those of the ranges
		// before it (a counting table turned into its own prefix, one column per thread)
		uint32_t run = top_start;
		for (uint32_t x = 0; x < kRanges; ++x) {
			const uint32_t c = T.rcnt[x * kRadix + t];
			T.rcnt[x * kRadix + t] = run;
			run += c;
		}
	}
	if (t != 0) return;
	PassPlan none{};
	for (int h = 0; h < 3; ++h) ctrl->hyb[h] = none;
	const uint32_t pres = presort_result(ctrl->check, a.reverse);
	int A = 0;
	uint32_t mask = 0;
	for (int p = 0; p < a.passes && !pres; ++p)
		if (!s_uniform[p]) { ++A; mask |= 1u << p; }
	ctrl->num_active = static_cast<uint32_t>(A);
	ctrl->active_mask = mask;
	// the histogram read also counted digit 0 per tile (in the logical order of the first read)
	ctrl->counts_ready = (a.counts0 && (mask & 1u)) ? 1u : 0u;
	void *X[2] = {a.fin, a.other};
	uint32_t *PX[2] = {a.pfin, a.pother};
	ctrl->pre = none;
	ctrl->fin = none;
	for (int p = 0; p < kMaxPasses; ++p) ctrl->pass[p] = none;
	if (A == 0) {
		if (pres == 3u) return;  // the run reversal (k_runs_copy) writes the result
		PassPlan &f = ctrl->fin;
		f.src = a.src; f.dst = a.fin; f.psrc = a.psrc; f.pdst = a.pfin;
		f.iota = a.payload && !a.psrc;
		// already in order: the result is the input (reversed when it was strictly descending)
		f.reverse = pres ? (pres == 2u) : a.reverse;
		f.active = (a.src != a.fin) || f.reverse || (a.payload && (f.iota || a.psrc != a.pfin));
		if (spec_done(ctrl, pres, a.fin, a.payload)) f.active = 0;  // the check already copied it
		return;
	}
	const void *src = a.src;
	const uint32_t *psrc = a.psrc;
	int rev = a.reverse, iota = a.payload && !a.psrc;
	const int first = (A - 1) & 1;
	if (a.keys_final_any && src == X[first]) {
		// the caller only wants the payload: let the keys end in the other array
		void *t = X[0]; X[0] = X[1]; X[1] = t;
	}
	if (src == X[first] || (a.payload && !iota && psrc == PX[first])) {
		// an in-place sort would read and write the same array in its first pass: stage the
		// source in the other array first (the copy also applies the tie reversal)
		PassPlan &c = ctrl->pre;
		c.active = 1; c.src = src; c.dst = X[first ^ 1]; c.psrc = psrc; c.pdst = PX[first ^ 1];
		c.reverse = rev; c.iota = iota;
		src = X[first ^ 1]; psrc = PX[first ^ 1]; rev = 0; iota = 0;
	}
	int j = 0;
	for (int p = 0; p < a.passes; ++p) {
		if (!(mask >> p & 1u)) continue;
		PassPlan &pp = ctrl->pass[p];
		const int slot = (A - 1 - j) & 1;
		pp.active = 1; pp.src = src; pp.dst = X[slot]; pp.psrc = psrc; pp.pdst = PX[slot];
		pp.reverse = rev; pp.iota = iota;
		src = X[slot]; psrc = PX[slot]; rev = 0; iota = 0;
		++j;
	}
	// The two-level MSD plan: both top digits must vary, and the largest bucket of the top-16-bit
	// histogram must fit the per-bucket LDS sort.
	const int P = a.passes;
	bool msd = a.hybrid && P >= 4 && a.other && !a.payload && !s_uniform[P - 1] && !s_uniform[P - 2];
	if (msd) msd = ctrl->joint_max <= a.hyb_cap;
	if (!a.hybrid) return;
	// the LSD plan above is either replaced by the MSD plan or incomplete (the read counted only the
	// top two digits): the gated full read and plan redo it
	for (int p = 0; p < kMaxPasses; ++p) ctrl->pass[p] = none;
	ctrl->pre = none;
	ctrl->fin = none;
	ctrl->counts_ready = 0;
	ctrl->active_mask = 0;
	if (!msd) {
		ctrl->need_full = 1;
		return;
	}
	void *X1 = a.other != a.src ? a.other : a.fin;
	void *X2 = X1 == a.other ? a.fin : a.other;
	PassPlan h{};
	h.active = 1;
	h.src = a.src; h.dst = X1; h.reverse = a.reverse;
	ctrl->hyb[0] = h;
	h.src = X1; h.dst = X2; h.reverse = 0;
	ctrl->hyb[1] = h;
	h.src = X2; h.dst = a.fin;
	ctrl->hyb[2] = h;
}

// ---------------------------------------------------------------------------------------------
// copy (staging copy before the first pass, or the whole result when no pass runs)

template <typename K>
__global__ __launch_bounds__(256) void k_copy(const Control *ctrl, int which, uint64_t n, int payload)
{
	const PassPlan c = which ? ctrl->fin : ctrl->pre;
	if (!c.active) return;
	const K *src = static_cast<const K *>(c.src);
	K *dst = static_cast<K *>(c.dst);
	const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	if (c.reverse && c.src == c.dst) {
		// in-place reversal: swap mirrored pairs
		for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n / 2; i += stride) {
			const K a = dst[i], b = dst[n - 1 - i];
			dst[i] = b; dst[n - 1 - i] = a;
			if (payload) {
				if (c.iota) { c.pdst[i] = static_cast<uint32_t>(n - 1 - i); c.pdst[n - 1 - i] = static_cast<uint32_t>(i); }
				else { const uint32_t pa = c.pdst[i], pb = c.pdst[n - 1 - i]; c.pdst[i] = pb; c.pdst[n - 1 - i] = pa; }
			}
		}
		if (payload && c.iota && (n & 1) && blockIdx.x == 0 && threadIdx.x == 0) c.pdst[n / 2] = static_cast<uint32_t>(n / 2);
		return;
	}
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint64_t s = c.reverse ? n - 1 - i : i;
		if (c.src != c.dst) dst[i] = src[s];
		if (payload) c.pdst[i] = c.iota ? static_cast<uint32_t>(s) : c.psrc[s];
	}
}

// ---------------------------------------------------------------------------------------------
// onesweep scatter pass

template <typename K, bool PAY, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_onesweep(Control *ctrl, const uint32_t *__restrict__ bucket,
	uint64_t *__restrict__ lookback, uint64_t n, uint32_t num_tiles, KeyXform xf, int pass)
{
	constexpr int TILE = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	static_assert(THREADS >= 256 && THREADS % 64 == 0, "one thread per digit is needed");
	const PassPlan pp = ctrl->pass[pass];
	if (!pp.active) return;
	const K *__restrict__ src = static_cast<const K *>(pp.src);
	K *__restrict__ dst = static_cast<K *>(pp.dst);
	const uint32_t tag = pass_tag(ctrl->epoch, pass);
	const uint32_t shift = 8u * static_cast<uint32_t>(pass);

	__shared__ uint32_t s_whist[NW * kRadix];
	__shared__ uint32_t s_digoff[kRadix];
	__shared__ uint32_t s_glob[kRadix];
	__shared__ uint32_t s_wsum[4];
	__shared__ uint32_t s_tile[2];
	__shared__ uint8_t s_dig[PAY ? TILE : 1];
	constexpr int kStage = (PAY && sizeof(K) < 4) ? 4 : static_cast<int>(sizeof(K));
	// the tile staging, and before it the ranking's peer words
	__shared__ __align__(16) unsigned char s_raw[max_bytes(TILE * kStage, rank_peer_bytes(THREADS))];
	K *s_keys = reinterpret_cast<K *>(s_raw);
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_raw);

	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t *pbucket = bucket + pass * kRadix;
	auto valid_of = [&](uint32_t tl) {
		const uint64_t base = static_cast<uint64_t>(tl) * TILE;
		return (n - base) < static_cast<uint64_t>(TILE) ? static_cast<uint32_t>(n - base) : static_cast<uint32_t>(TILE);
	};
	K keys[KPT];
	uint32_t pay[PAY ? KPT : 1], npay[PAY ? KPT : 1];
	// the keys (and payload) of tile tl in logical order l = w*64*KPT + j*64 + lane
	auto load = [&](uint32_t tl, uint32_t (&py)[PAY ? KPT : 1]) {
		if (tl >= num_tiles) return;
		const uint64_t base = static_cast<uint64_t>(tl) * TILE;
		const uint32_t valid = valid_of(tl);
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			const uint64_t g = base + l;
			if (l < valid) {
				const uint64_t sidx = pp.reverse ? n - 1 - g : g;
				keys[j] = src[sidx];
				if constexpr (PAY) py[j] = pp.iota ? static_cast<uint32_t>(sidx) : pp.psrc[sidx];
			} else {
				keys[j] = K(0);
				if constexpr (PAY) py[j] = 0;
			}
		}
	};
	// tiles are taken in ticket order, one ticket ahead: the next tile's loads are issued before
	// this tile's look-back and stores, so that they overlap both
	if (t == 0) s_tile[0] = atomicAdd(&ctrl->tile_counter[pass], 1u);
	__syncthreads();
	uint32_t tile = s_tile[0];
	load(tile, pay);
	for (uint32_t it = 0; tile < num_tiles; ++it) {
		if (t == 0) s_tile[(it + 1) & 1u] = atomicAdd(&ctrl->tile_counter[pass], 1u);
		for (int i = t; i < NW * kRadix; i += THREADS) {
			s_whist[i] = 0;
			s_peer[i] = 0;
		}
		__syncthreads();
		const uint32_t next = s_tile[(it + 1) & 1u];
		const uint32_t valid = valid_of(tile);
		uint32_t dig[KPT];
		uint32_t rank[KPT];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			dig[j] = l < valid ? digit_of<K>(keys[j], xf, shift) : 0xFFu;  // past the end: ranked last, never written
		}
		// rank within the wave in logical order (j-major, lane-minor)
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint64_t peers = wave_peers(dig[j], s_peer + w * kRadix);
			const uint32_t below = lanes_below(peers);
			uint32_t *slot = &s_whist[w * kRadix + dig[j]];
			const uint32_t old = *slot;
			if (below == 0) *slot = old + static_cast<uint32_t>(__popcll(peers));
			rank[j] = old + below;
		}
		__syncthreads();
		uint32_t total = 0, pub = 0;
		if (t < kRadix) {
			for (int v = 0; v < NW; ++v) {
				const uint32_t c = s_whist[v * kRadix + t];
				s_whist[v * kRadix + t] = total;
				total += c;
			}
			pub = total - ((t == kRadix - 1) ? static_cast<uint32_t>(TILE) - valid : 0u);
			atomic_store_agent(&lookback[static_cast<uint64_t>(tile) * kLbStride + t],
				lb_word(pub, tile == 0 ? kLbInclusive : kLbAggregate, tag));
			const uint32_t incl = wave_inclusive_sum(total);
			if (lane == 63) s_wsum[w] = incl;
			s_digoff[t] = incl - total;
		}
		__syncthreads();
		if (t < kRadix) {
			uint32_t before = 0;
			for (uint32_t v = 0; v < w; ++v) before += s_wsum[v];
			s_digoff[t] += before;
		}
		__syncthreads();
		// stage the tile in LDS in digit order; the key registers are then free for the next tile
		uint32_t pos[KPT];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			pos[j] = s_digoff[dig[j]] + s_whist[w * kRadix + dig[j]] + rank[j];
			s_keys[pos[j]] = keys[j];
		}
		load(next, npay);
		if (t < kRadix) {
			// decoupled look-back over the preceding tiles of this digit
			uint32_t excl = 0;
			if (tile > 0) {
				excl = lookback_window<16, kLbStride>(lookback + t, tile, 0, tag);
				atomic_store_agent(&lookback[static_cast<uint64_t>(tile) * kLbStride + t], lb_word(excl + pub, kLbInclusive, tag));
			}
			s_glob[t] = pbucket[t] + excl - s_digoff[t];
		}
		__syncthreads();
		for (uint32_t i = t; i < valid; i += THREADS) {
			const K k = s_keys[i];
			const uint32_t d = digit_of<K>(k, xf, shift);
			if constexpr (PAY) s_dig[i] = static_cast<uint8_t>(d);
			dst[s_glob[d] + i] = k;
		}
		if constexpr (PAY) {
			__syncthreads();
			uint32_t *s_pay = reinterpret_cast<uint32_t *>(s_raw);
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_pay[pos[j]] = pay[j];
			__syncthreads();
			for (uint32_t i = t; i < valid; i += THREADS) pp.pdst[s_glob[s_dig[i]] + i] = s_pay[i];
#pragma unroll
			for (int j = 0; j < KPT; ++j) pay[j] = npay[j];
		}
		__syncthreads();
		tile = next;
	}
}


// ---------------------------------------------------------------------------------------------
// count + chained scan + scatter (no inter-block dependency inside the scatter)


// Tile-based read of every key: the histograms of all digits (as k_hist) and, when counts is not
// null, the per-tile counts of the digit at count_shift in the logical order of the first pass
// (tile t holds the logical positions [t*tile_keys, (t+1)*tile_keys), read back to front under
// isrevorder).
// PARTIAL (the two-level MSD plan's read): only the top two digits' histograms, plus the AND and
// OR of all transformed keys (a digit is uniform exactly when they agree on it). A non-null gate
// that reads zero skips the launch (the full read that only a fallback needs).
template <typename K, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_hist_tiles(const K *__restrict__ in, uint64_t n, KeyXform xf,
	uint32_t *__restrict__ hist, uint32_t *__restrict__ counts, uint32_t tile_keys, uint32_t num_tiles, int reverse,
	Control *ctrl, uint32_t count_shift, const uint32_t *gate)
{
	if (gate && !*gate) return;
	if (presort_result(ctrl->check, reverse)) return;  // the input is already in order
	constexpr int P = sizeof(K);
	// The top two digits (the skewed ones for floats and clustered integers) and the tile counts
	// keep C counter copies, the digits below them one; digit p's histogram starts at s_h[hoff(p)].
	constexpr int NLOW = P - 2 > 0 ? P - 2 : 0;
	constexpr int NTOP = P - NLOW;
	constexpr int C = count_copies(NTOP + 1, NLOW);
	constexpr int HWORDS = (NLOW + NTOP * C) * kRadix;
	auto hoff = [](int p) { return p < NLOW ? p * kRadix : (NLOW + (p - NLOW) * C) * kRadix; };
	auto hcop = [](int p) { return p < NLOW ? 1 : C; };
	__shared__ uint32_t s_h[HWORDS];
	__shared__ uint32_t s_t[kRadix * C];
	// the digit the tile counts follow is not counted twice: its histogram is the sum of its tile
	// counts, added in at the end of every tile
	const int pc = (counts && count_shift % 8u == 0 && static_cast<int>(count_shift / 8u) < P) ?
		static_cast<int>(count_shift / 8u) : -1;
	for (int i = threadIdx.x; i < HWORDS; i += THREADS) s_h[i] = 0;
	for (uint32_t tile = blockIdx.x; tile < num_tiles; tile += gridDim.x) {
		for (int i = threadIdx.x; i < kRadix * C; i += THREADS) s_t[i] = 0;
		__syncthreads();
		const uint64_t base = static_cast<uint64_t>(tile) * tile_keys;
		const uint32_t len = (n - base) < tile_keys ? static_cast<uint32_t>(n - base) : tile_keys;
		// the tile's keys as a physical range (the logical range read back to front under isrevorder)
		const uint64_t lo = reverse ? n - base - len : base;
		for_each_key<K, THREADS>(in, lo, len, [&](K x) {
			const auto k = sort_key<K>(x, xf);
#pragma unroll
			for (int p = 0; p < P; ++p) {
				if (p == pc) continue;
				const uint32_t d = static_cast<uint32_t>(k >> (8 * p)) & 0xFFu;
				if (p < NLOW) wave_count<1>(s_h + hoff(p), d);
				else wave_count<C>(s_h + hoff(p), d);
			}
			if (counts) wave_count<C>(s_t, static_cast<uint32_t>(k >> count_shift) & 0xFFu);
		});
		__syncthreads();
		if (counts)
			for (int i = threadIdx.x; i < kRadix; i += THREADS) {
				const uint32_t v = copies_sum<C>(s_t, i);
				counts[static_cast<uint64_t>(tile) * kCountStride + i] = v;
				if (pc >= 0) s_h[hoff(pc) + i * hcop(pc)] += v;  // one thread per digit: no race
			}
		__syncthreads();
	}
	__syncthreads();
	for (int i = threadIdx.x; i < P * kRadix; i += THREADS) {
		const int p = i >> 8;
		const uint32_t d = static_cast<uint32_t>(i) & 0xFFu;
		const uint32_t v = hcop(p) == 1 ? s_h[hoff(p) + d] : copies_sum<C>(s_h + hoff(p), d);
		if (v) atomicAdd(&hist[i], v);
	}
}

// The MSD plan's read when the key transform is injective: the histogram of the top 16 bits of
// every transformed key, which gives the exact sizes of the 65536 buckets the two partitions make.
// The keys are split into kRanges ranges of whole tiles of the first partition (range x: the tiles
// x * per .. x * per + per - 1 in logical order), and the blocks into as many groups, so that the
// reduction also yields every range's top-digit counts: the first partition's blocks take the
// output ranges of a range's tiles from that range's cursors. One block per CU holds the 65536
// counters as 16-bit halves of LDS words (128 KB). The block counts its keys kJointFlush at a time
// and then moves 32768 out of every counter that reached it into ovf (global; only a few dominant
// buckets make it happen), so no half ever exceeds 0xFFFF. The block's counts go to its own
// 65536-word slice of part; k_joint_reduce adds the slices.

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_hist_joint(const K *__restrict__ in, uint64_t n, KeyXform xf,
	uint64_t range_keys, int reverse, const Control *ctrl, uint32_t *__restrict__ part, uint32_t *__restrict__ ovf)
{
	if (presort_result(ctrl->check, reverse)) return;  // the input is already in order
	constexpr int P = sizeof(K);
	static_assert(P >= 4, "the two-level MSD plan sorts 4- and 8-byte keys");
	__shared__ uint32_t s_j[kJointBins / 2];  // bin b in the half (b & 1) of word b >> 1
	const uint32_t t = threadIdx.x;
	for (uint32_t i = t; i < kJointBins / 2; i += THREADS) s_j[i] = 0;
	// this block's share of its range (logical positions), as a physical range
	const uint32_t per_range = gridDim.x / kRanges;
	const uint32_t x = blockIdx.x / per_range, j = blockIdx.x % per_range;
	const uint64_t r0 = min(n, static_cast<uint64_t>(x) * range_keys), r1 = min(n, r0 + range_keys);
	const uint64_t share = (r1 - r0 + per_range - 1) / per_range;
	const uint64_t b0 = min(r1, r0 + j * share), b1 = min(r1, b0 + share);
	const uint64_t lo = reverse ? n - b1 : b0, hi = reverse ? n - b0 : b1;
	__syncthreads();
	for (uint64_t c = lo; c < hi; c += kJointFlush) {
		const uint32_t len = static_cast<uint32_t>(min<uint64_t>(kJointFlush, hi - c));
		for_each_key<K, THREADS>(in, c, len, [&](K key) {
			const uint32_t bin = static_cast<uint32_t>(sort_key<K>(key, xf) >> (8 * P - 16));
			// a wave whose keys share one bin adds them with one atomic (sorted input), else one per lane
			const uint64_t active = __ballot(1);
			const uint32_t bin0 = __builtin_amdgcn_readfirstlane(bin);
			const bool same = __ballot(bin == bin0) == active;
			if (!same || lanes_below(active) == 0)
				atomicAdd(&s_j[bin >> 1], (same ? static_cast<uint32_t>(__popcll(active)) : 1u) << ((bin & 1u) * 16u));
		});
		__syncthreads();
		// every half is below 0x8000 after the sweep, and at most kJointFlush keys are counted
		// before the next one
		for (uint32_t i = t; i < kJointBins / 2; i += THREADS) {
			const uint32_t v = s_j[i];
			if (v & 0x80008000u) {
				if (v & 0x8000u) atomicAdd(&ovf[2 * i], 0x8000u);
				if (v & 0x80000000u) atomicAdd(&ovf[2 * i + 1], 0x8000u);
				s_j[i] = v & 0x7FFF7FFFu;
			}
		}
		__syncthreads();
	}
	uint2 *slice = reinterpret_cast<uint2 *>(part + static_cast<uint64_t>(blockIdx.x) * kJointBins);
	for (uint32_t i = t; i < kJointBins / 2; i += THREADS) {
		const uint32_t v = s_j[i];
		slice[i] = make_uint2(v & 0xFFFFu, v >> 16);
	}
}

// One block per top digit, one thread per second digit: the bucket's count (the slices of
// k_hist_joint and the overflow, which is left zeroed), its start relative to its segment (also the
// second partition's cursor), the histograms of the top two digits, every range's count of the top
// digit (rcnt[x * 256 + top digit]) and the largest bucket.
template <int P>
__global__ __launch_bounds__(256) void k_joint_reduce(Control *ctrl, int reverse, const uint32_t *__restrict__ part,
	uint32_t slices, SegTables T, uint32_t *__restrict__ hist)
{
	if (presort_result(ctrl->check, reverse)) return;
	__shared__ uint32_t s_wave[4];
	__shared__ uint32_t s_r[kRanges][4];
	const uint32_t sg = blockIdx.x, d = threadIdx.x, lane = d & 63u, w = d >> 6;
	const uint32_t b = sg * kRadix + d;
	uint32_t v = T.ovf[b];
	if (v) T.ovf[b] = 0;
	const uint32_t per_range = slices / kRanges;
	uint32_t rv[kRanges] = {};
	// slice k of every range per step: eight independent loads, four steps in flight
#pragma unroll 4
	for (uint32_t k = 0; k < per_range; ++k) {
#pragma unroll
		for (uint32_t x = 0; x < kRanges; ++x) rv[x] += part[static_cast<uint64_t>(x * per_range + k) * kJointBins + b];
	}
#pragma unroll
	for (uint32_t x = 0; x < kRanges; ++x) v += rv[x];
	T.bucket_len[b] = v;
	const uint32_t incl = wave_inclusive_sum(v);
	uint32_t m = v;
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), o, 64)));
#pragma unroll
	for (uint32_t x = 0; x < kRanges; ++x) {
		uint32_t r = rv[x];
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) r += static_cast<uint32_t>(__shfl_xor(static_cast<int>(r), o, 64));
		if (lane == 0) s_r[x][w] = r;
	}
	if (lane == 63) s_wave[w] = incl;
	if (lane == 0) atomicMax(&ctrl->joint_max, m);
	__syncthreads();
	uint32_t before = 0;
	for (uint32_t u = 0; u < w; ++u) before += s_wave[u];
	const uint32_t rel = before + incl - v;
	T.bucket_start[b] = rel;
	T.cursor[b] = rel;
	if (d == kRadix - 1) hist[(P - 1) * kRadix + sg] = rel + v;
	if (d < kRanges) T.rcnt[d * kRadix + sg] = s_r[d][0] + s_r[d][1] + s_r[d][2] + s_r[d][3];
	if (v) atomicAdd(&hist[(P - 2) * kRadix + d], v);
}

// per-tile counts of one digit (passes after the first, or when the histogram read did not count)
template <typename K, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_count(const Control *ctrl, int pass, KeyXform xf, uint64_t n,
	uint32_t tile_keys, uint32_t num_tiles, uint32_t *__restrict__ counts)
{
	const PassPlan pp = ctrl->pass[pass];
	if (!pp.active || (ctrl->counts_ready >> pass & 1u)) return;
	const K *__restrict__ src = static_cast<const K *>(pp.src);
	const uint32_t shift = 8u * static_cast<uint32_t>(pass);
	constexpr int C = count_copies(1);
	__shared__ uint32_t s_t[kRadix * C];
	for (uint32_t tile = blockIdx.x; tile < num_tiles; tile += gridDim.x) {
		for (int i = threadIdx.x; i < kRadix * C; i += THREADS) s_t[i] = 0;
		__syncthreads();
		const uint64_t base = static_cast<uint64_t>(tile) * tile_keys;
		const uint32_t len = (n - base) < tile_keys ? static_cast<uint32_t>(n - base) : tile_keys;
		const uint64_t lo = pp.reverse ? n - base - len : base;
		for_each_key<K, THREADS>(src, lo, len, [&](K x) { wave_count<C>(s_t, digit_of<K>(x, xf, shift)); });
		__syncthreads();
		for (int i = threadIdx.x; i < kRadix; i += THREADS) counts[static_cast<uint64_t>(tile) * kCountStride + i] = copies_sum<C>(s_t, i);
		__syncthreads();
	}
}

// Chained scan with decoupled look-back over the tile-count matrix: every block takes the next
// chunk of kScanChunk tiles, scans it per digit, publishes the chunk total, resolves the prefix
// of the preceding chunks and turns the counts into global output offsets in place.
// pp: the pass (inactive: exit), counter: its chunk ticket, start_row: the 256 bucket starts of its
// digit, tag_slot: its look-back tag slot within the sort.
__global__ __launch_bounds__(256) void k_scan(const Control *ctrl, const PassPlan *pp, uint32_t *counter,
	const uint32_t *__restrict__ start_row, int tag_slot, uint32_t *__restrict__ counts, uint64_t *__restrict__ lookback,
	uint32_t num_tiles)
{
	if (!pp->active) return;
	const uint32_t nchunks = (num_tiles + kScanChunk - 1) / kScanChunk;
	const uint32_t tag = pass_tag(ctrl->epoch, tag_slot);
	const uint32_t d = threadIdx.x;
	const uint32_t start = start_row[d];
	__shared__ uint32_t s_chunk;
	for (;;) {
		if (d == 0) s_chunk = atomicAdd(counter, 1u);
		__syncthreads();
		const uint32_t chunk = s_chunk;
		__syncthreads();
		if (chunk >= nchunks) break;
		const uint32_t t0 = chunk * kScanChunk;
		const uint32_t t1 = min(t0 + kScanChunk, num_tiles);
		uint32_t run = 0;
		for (uint32_t t = t0; t < t1; t += 8) {
			uint32_t c[8];
#pragma unroll
			for (int j = 0; j < 8; ++j) c[j] = (t + j < t1) ? counts[static_cast<uint64_t>(t + j) * kCountStride + d] : 0u;
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				if (t + j < t1) counts[static_cast<uint64_t>(t + j) * kCountStride + d] = run;
				run += c[j];
			}
		}
		uint32_t excl = 0;
		uint64_t *lb = lookback + static_cast<uint64_t>(chunk) * kLbStride + d;
		if (chunk == 0) {
			atomic_store_agent(lb, lb_word(run, kLbInclusive, tag));
		} else {
			atomic_store_agent(lb, lb_word(run, kLbAggregate, tag));
			excl = lookback_window<8, kLbStride>(lookback + d, chunk, 0, tag);
			atomic_store_agent(lb, lb_word(excl + run, kLbInclusive, tag));
		}
		const uint32_t add = start + excl;
		for (uint32_t t = t0; t < t1; ++t) counts[static_cast<uint64_t>(t) * kCountStride + d] += add;
	}
}

// Scatter of one digit with known tile offsets. Blocks are dealt to the XCDs round robin, so
// block b works on the tile range of XCD b % 8: neighbouring tiles, whose digit runs are
// neighbours in the output, are written through the same L2.
// Tiles of the segmented MSD pass: segment s (a top-digit bucket) owns tiles
// [tile_start[s], tile_start[s+1]), the last of them partial; no tile straddles two segments.
struct SegMap {
	const uint32_t *tile_start;  // 257 entries, tile_start[256] = number of tiles
	const uint32_t *seg_start;
	const uint32_t *seg_len;
};

// first key and key count of tile tl of a segment map (s_ts: tile_start staged in LDS)
__device__ __forceinline__ void seg_tile(const uint32_t *s_ts, const SegMap &m, uint32_t TILE, uint32_t tl, uint64_t &base,
	uint32_t &valid, uint32_t &segment)
{
	// the last segment whose first tile is <= tl (empty segments share their successor's first tile)
	uint32_t sg = 0;
#pragma unroll
	for (uint32_t step = 128; step; step >>= 1)
		if (s_ts[sg + step] <= tl) sg += step;
	const uint32_t k = tl - s_ts[sg];
	base = static_cast<uint64_t>(m.seg_start[sg]) + static_cast<uint64_t>(k) * TILE;
	const uint32_t left = m.seg_len[sg] - k * static_cast<uint32_t>(TILE);
	valid = left < static_cast<uint32_t>(TILE) ? left : static_cast<uint32_t>(TILE);
	segment = sg;
}

// pp: the pass (inactive: exit); shift: its digit. With SEG, the tiles are those of the segment map
// and num_tiles is an upper bound (the exact count is tile_start[256]). With GROUP (keys only) the
// keys of a digit are not kept in order within the tile (LDS atomics instead of block ranking):
// only for the MSD plan's partitions when the key transform is injective, where the per-bucket
// sort that follows orders every bucket by the whole sort key and equal sort keys are equal keys.
// With CUR (GROUP only) the tiles take their output ranges from cursors (one atomic per digit and
// tile) instead of precomputed tile offsets, and fill a digit's run in any order, which is again
// invisible after the per-bucket sort: with SEG the cursors of the tile's segment (positions
// relative to it), else those of the tile's key range (range x holds the tiles x * per ..
// x * per + per - 1; absolute positions).
// With SPL the digit is not a byte of the sort key but its bucket id under a splitter table (the
// sample partition plan): SPL 1, the table at split for every tile; SPL 2 (with SEG), the table of
// the tile's segment, split + segment * kSplitTabWords. The digits are staged beside the keys, so
// the stores do not search again. With GROUP and SPL (keys only, tile offsets from counts) the
// per-bucket sort that follows orders each bucket by the whole sort key, and the sample plan takes
// these passes only when equal sort keys are equal keys. A non-null gsel selects between the two
// forms of a pass launched side by side: the grouping form runs when *gsel is 1, the ranking form
// when it is 0.
template <typename K, bool PAY, int THREADS, int KPT, bool SEG = false, bool GROUP = false, bool CUR = false, int SPL = 0>
__global__ __launch_bounds__(THREADS) void k_scatter(const PassPlan *ppp, uint32_t shift, KeyXform xf, uint64_t n,
	uint32_t num_tiles, const uint32_t *__restrict__ offsets, SegMap seg, uint32_t *__restrict__ cursor,
	const uint32_t *__restrict__ split, const uint32_t *gsel)
{
	static_assert(!CUR || GROUP, "cursors are for the grouping passes");
	static_assert(!(SPL && CUR) && (SPL != 2 || SEG), "splitter digits: counted passes; per-segment tables on segment tiles");
	using W = typename KeyTraits<K>::W;
	constexpr int TILE = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	const PassPlan pp = *ppp;
	if (!pp.active) return;
	if (gsel && *gsel != (GROUP ? 1u : 0u)) return;
	const K *__restrict__ src = static_cast<const K *>(pp.src);
	K *__restrict__ dst = static_cast<K *>(pp.dst);
	__shared__ uint32_t s_whist[NW * kRadix];
	__shared__ uint32_t s_digoff[kRadix];
	__shared__ uint32_t s_glob[kRadix];
	__shared__ uint32_t s_wsum[4];
	__shared__ uint32_t s_ts[SEG ? kRadix + 1 : 1];
	__shared__ uint8_t s_dig[(PAY || SPL) ? TILE : 1];
	__shared__ std::conditional_t<SPL != 0, SplitLds<W>, uint32_t> s_tab;  // the splitter table (SPL only)
	constexpr int kStage = (PAY && sizeof(K) < 4) ? 4 : static_cast<int>(sizeof(K));
	// the tile staging, and before it the ranking's peer words
	__shared__ __align__(16) unsigned char s_raw[GROUP ? TILE * kStage : max_bytes(TILE * kStage, rank_peer_bytes(THREADS))];
	K *s_keys = reinterpret_cast<K *>(s_raw);
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_raw);
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t per = (num_tiles + 7) / 8;
	const uint32_t x = blockIdx.x & 7u, step = gridDim.x >> 3;
	uint32_t real_tiles = num_tiles;
	if constexpr (SEG) {
		for (int i = t; i <= kRadix; i += THREADS) s_ts[i] = seg.tile_start[i];
		__syncthreads();
		real_tiles = s_ts[kRadix];
	}
	uint32_t tab_seg = 0xFFFFFFFFu;  // segment whose table is staged (SPL 2)
	// The first splitter partition writes the keys (and payload) of an equality bucket straight to
	// the sort's final arrays, at the same positions: the bucket is in its final order, and the second
	// level skips its segment. The final arrays are those of the plan's last pass (ppp[2]: hyb[2]).
	__shared__ uint32_t s_eqm[SPL == 1 ? 8 : 1];
	K *__restrict__ eq_dst = dst;
	uint32_t *eq_pdst = pp.pdst;
	if constexpr (SPL == 1) {
		split_load<W, THREADS>(split, s_tab);
		// (the same test as the plan's: not when the final arrays are the source, as in place)
		const PassPlan fp = ppp[2];
		const bool direct = fp.dst != pp.src && !(PAY && pp.psrc && pp.psrc == fp.pdst);
		if (threadIdx.x < 8) s_eqm[threadIdx.x] = direct ? split[kSplitEqWord + threadIdx.x] : 0u;
		eq_dst = static_cast<K *>(fp.dst);
		eq_pdst = fp.pdst;
		__syncthreads();
	}
	auto eq_bucket = [&](uint32_t d) {
		if constexpr (SPL == 1) return ((s_eqm[d >> 5] >> (d & 31u)) & 1u) != 0;
		else return false;
	};
	auto geom = [&](uint32_t tl, uint64_t &base, uint32_t &valid, uint32_t &sgm) {
		if constexpr (SEG) {
			seg_tile(s_ts, seg, TILE, tl, base, valid, sgm);
		} else {
			sgm = 0;
			base = static_cast<uint64_t>(tl) * TILE;
			valid = (n - base) < static_cast<uint64_t>(TILE) ? static_cast<uint32_t>(n - base) : static_cast<uint32_t>(TILE);
		}
	};
	K keys[KPT];
	uint32_t pay[PAY ? KPT : 1];
	// loads of a tile are issued before the stores of the previous one, so that waiting for them
	// (the memory counter is in order and counts stores too) never waits for those stores
	auto load_tile = [&](uint64_t base, uint32_t valid) {
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			const uint64_t g = base + l;
			if (l < valid) {
				const uint64_t sidx = pp.reverse ? n - 1 - g : g;
				keys[j] = src[sidx];
				if constexpr (PAY) pay[j] = pp.iota ? static_cast<uint32_t>(sidx) : pp.psrc[sidx];
			} else {
				keys[j] = K(0);
				if constexpr (PAY) pay[j] = 0;
			}
		}
	};
	uint32_t k = blockIdx.x >> 3;
	uint32_t tile = x * per + k;
	uint64_t base = 0;
	uint32_t valid = 0, sgm = 0;
	if (k < per && tile < real_tiles) {
		geom(tile, base, valid, sgm);
		load_tile(base, valid);
	}
	while (k < per && tile < real_tiles) {
		uint32_t dig[KPT], pos[KPT], total;
		if constexpr (SPL == 2) {
			if (sgm != tab_seg) {  // uniform across the block
				__syncthreads();
				split_load<W, THREADS>(split + static_cast<size_t>(sgm) * kSplitTabWords, s_tab);
				__syncthreads();
				tab_seg = sgm;
			}
		}
		if constexpr (SPL) {
			if (__builtin_amdgcn_readfirstlane(s_tab.nsub)) {
#pragma unroll
				for (int j = 0; j < KPT; ++j) {
					const uint32_t l = w * 64u * KPT + j * 64u + lane;
					dig[j] = l < valid ? split_id<W, true>(s_tab, sort_key<K>(keys[j], xf)) : 0xFFu;
				}
			} else {
#pragma unroll
				for (int j = 0; j < KPT; ++j) {
					const uint32_t l = w * 64u * KPT + j * 64u + lane;
					dig[j] = l < valid ? split_id<W, false>(s_tab, sort_key<K>(keys[j], xf)) : 0xFFu;
				}
			}
		} else {
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				dig[j] = l < valid ? digit_of<K>(keys[j], xf, shift) : 0xFFu;
			}
		}
		if constexpr (GROUP) {
			static_assert(!PAY, "grouping without order is for keys only");
			(void)total;
			(void)s_peer;
			block_group<THREADS, KPT>(dig, valid, pos, s_whist, s_digoff, s_wsum);
			if constexpr (CUR) {
				if (t < kRadix) {
					const uint32_t c = (t + 1 < kRadix ? s_digoff[t + 1] : valid) - s_digoff[t];
					if constexpr (SEG) {
						const uint32_t at = c ? atomicAdd(&cursor[sgm * kRadix + t], c) : 0u;
						s_glob[t] = seg.seg_start[sgm] + at - s_digoff[t];
					} else {
						const uint32_t at = c ? atomicAdd(&cursor[x * kRadix + t], c) : 0u;
						s_glob[t] = at - s_digoff[t];
					}
				}
			} else {
				if (t < kRadix) s_glob[t] = offsets[static_cast<uint64_t>(tile) * kCountStride + t] - s_digoff[t];
			}
#pragma unroll
			for (int j = 0; j < KPT; ++j)
				if (pos[j] != 0xFFFFFFFFu) s_keys[pos[j]] = keys[j];
			if constexpr (SPL) {
#pragma unroll
				for (int j = 0; j < KPT; ++j)
					if (pos[j] != 0xFFFFFFFFu) s_dig[pos[j]] = static_cast<uint8_t>(dig[j]);
			}
		} else {
			block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
			if (t < kRadix) s_glob[t] = offsets[static_cast<uint64_t>(tile) * kCountStride + t] - s_digoff[t];
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_keys[pos[j]] = keys[j];
			if constexpr (SPL) {
#pragma unroll
				for (int j = 0; j < KPT; ++j) s_dig[pos[j]] = static_cast<uint8_t>(dig[j]);
			}
		}
		uint32_t ppos[PAY ? KPT : 1];
		if constexpr (PAY) {
#pragma unroll
			for (int j = 0; j < KPT; ++j) ppos[j] = pos[j];
		}
		uint32_t pcur[PAY ? KPT : 1];
		if constexpr (PAY) {
#pragma unroll
			for (int j = 0; j < KPT; ++j) pcur[j] = pay[j];
		}
		// prefetch the next tile of this block
		const uint32_t k2 = k + step, tile2 = x * per + k2;
		uint64_t base2 = 0;
		uint32_t valid2 = 0, sgm2 = 0;
		if (k2 < per && tile2 < real_tiles) {
			geom(tile2, base2, valid2, sgm2);
			load_tile(base2, valid2);
		}
		__syncthreads();
		for (uint32_t i = t; i < valid; i += THREADS) {
			const K kk = s_keys[i];
			uint32_t dd;
			if constexpr (SPL) {
				dd = s_dig[i];
			} else {
				dd = digit_of<K>(kk, xf, shift);
				if constexpr (PAY) s_dig[i] = static_cast<uint8_t>(dd);
			}
			(eq_bucket(dd) ? eq_dst : dst)[s_glob[dd] + i] = kk;
		}
		if constexpr (PAY) {
			__syncthreads();
			uint32_t *s_pay = reinterpret_cast<uint32_t *>(s_raw);
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_pay[ppos[j]] = pcur[j];
			__syncthreads();
			for (uint32_t i = t; i < valid; i += THREADS) {
				const uint32_t dd = s_dig[i];
				(eq_bucket(dd) ? eq_pdst : pp.pdst)[s_glob[dd] + i] = s_pay[i];
			}
		}
		__syncthreads();
		k = k2;
		tile = tile2;
		base = base2;
		valid = valid2;
		sgm = sgm2;
	}
}

// ---------------------------------------------------------------------------------------------
// two-level MSD plan: segmented count and scan of the second digit, fallback switch, and the
// per-bucket LDS sort of the remaining digits

// The per-bucket sort, one block per bucket (a persistent form that loads the next bucket while
// sorting the current one measured slower: its registers halve the resident blocks). All keys of a
// bucket share the top two digits, so the transformed key shifted left by 16 keeps every digit
// below them. Each key becomes the word (transformed key << 16) | l, with l its position in the
// bucket: distinct words whose order is the stable order. The words are partitioned by their top
// PB bits with LDS atomics (in any order), then every word's final position inside its run of
// equal top bits is the number of smaller words in the run. The transform only complements bits
// (the plan runs for integer keys under an injective transform), so every word gives back its raw
// key: the bucket index (the top 16 bits), the word's key bits, the transform undone; the ranks go
// to LDS as the inverse permutation (position -> word) and the output is gathered from the words
// in order. A bucket with a run longer than kLocalRunMax is queued for k_local_lsd instead. A
// bucket is read whole before it is written, so source and destination may coincide.
constexpr uint32_t kLocalRunMax = 96;

template <typename K, int THREADS, int KPT, int PB>
__global__ __launch_bounds__(THREADS) void k_local(Control *ctrl, KeyXform xf, const uint32_t *__restrict__ seg_start,
	const uint32_t *__restrict__ bstart, const uint32_t *__restrict__ blen, uint32_t *__restrict__ queue)
{
	constexpr int CAP = THREADS * KPT;
	constexpr int P = sizeof(K);
	constexpr uint32_t NB = 1u << PB;            // partition buckets
	constexpr uint32_t PER = NB / THREADS;       // partition buckets per thread in the scan
	constexpr uint32_t kTop = 8u * P - PB;       // the partition bits' place in a packed word
	static_assert(P >= 4, "the two-level MSD plan sorts 4- and 8-byte keys");
	static_assert(CAP <= 65536, "positions are packed into 16 bits");
	static_assert(NB % THREADS == 0 && PB <= 16, "partition buckets per thread");
	using W = typename KeyTraits<K>::W;
	static_assert(sizeof(W) == sizeof(K), "transformed keys share the key width");
	// s_u: the NB partition counters, then cursors (after the scatter: the run ends, so that a run
	// starts where the previous one ends), the wave sums and the overflow flag above them; then the
	// final positions (u16) over all of it
	constexpr uint32_t kSmallWords = THREADS / 64 + 1;
	constexpr uint32_t kUnionWords = (NB + kSmallWords > CAP / 2) ? NB + kSmallWords : CAP / 2;
	__shared__ __align__(16) W s_t[CAP];          // packed words, then the raw keys in final order
	__shared__ uint32_t s_u[kUnionWords];
	uint32_t *s_wsum = s_u + NB, *s_big = s_wsum + THREADS / 64;
	const PassPlan pp = ctrl->hyb[2];
	if (!pp.active) return;
	const K *__restrict__ src_all = static_cast<const K *>(pp.src);
	K *__restrict__ dst_all = static_cast<K *>(pp.dst);
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	K key[KPT];
	uint32_t len = 0, start = 0;
	auto fetch = [&](uint32_t bk, K (&kk)[KPT], uint32_t &ln, uint32_t &st) {
		ln = 0;
		st = 0;
		if (bk < kHybBuckets) {
			ln = blen[bk];
			st = seg_start[bk >> 8] + bstart[bk];
		}
		const K *__restrict__ src = src_all + st;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			kk[j] = l < ln ? src[l] : K(0);
		}
	};
	const uint32_t bk = blockIdx.x;
	fetch(bk, key, len, start);
	{
		K *__restrict__ dst = dst_all + start;
		if (len == 1) {
			if (t == 0 && pp.src != pp.dst) dst[0] = key[0];
		} else if (len > 1) {
			// the top 16 bits of the sort key, which every key of the bucket shares: the bucket index
			const W pre = static_cast<W>(static_cast<W>(bk) << (8 * P - 16));
			uint32_t *s_cur = s_u;
			for (uint32_t i = t; i < NB; i += THREADS) s_cur[i] = 0;
			if (t == 0) *s_big = 0;
			__syncthreads();
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				if (l < len) atomicAdd(&s_cur[static_cast<uint32_t>(static_cast<W>(sort_key<K>(key[j], xf) << 16) >> kTop)], 1u);
			}
			__syncthreads();
			// exclusive scan of the counters, PER consecutive ones per thread
			uint32_t c[PER], sum = 0, big = 0;
#pragma unroll
			for (uint32_t k = 0; k < PER; ++k) {
				c[k] = s_cur[t * PER + k];
				sum += c[k];
				big |= c[k] > kLocalRunMax;
			}
			const uint32_t incl = wave_inclusive_sum(sum);
			if (lane == 63) s_wsum[w] = incl;
			if (big) *s_big = 1;
			__syncthreads();
			uint32_t run = incl - sum;
			for (uint32_t v = 0; v < w; ++v) run += s_wsum[v];
#pragma unroll
			for (uint32_t k = 0; k < PER; ++k) {
				s_cur[t * PER + k] = run;
				run += c[k];
			}
			__syncthreads();
			if (*s_big) {
				if (t == 0) queue[atomicAdd(&ctrl->hyb_queue, 1u)] = bk;
			} else {
#pragma unroll
				for (int j = 0; j < KPT; ++j) {
					const uint32_t l = w * 64u * KPT + j * 64u + lane;
					if (l < len) {
						const W word = static_cast<W>(sort_key<K>(key[j], xf) << 16) | static_cast<W>(l);
						s_t[atomicAdd(&s_cur[static_cast<uint32_t>(word >> kTop)], 1u)] = word;
					}
				}
				__syncthreads();
				// the final position of every word (its rank in its run, after the run's start)
				// Equal sort keys are equal keys (the transform only complements bits), so every
				// word gives back its raw key: the bucket's top 16 bits, the word's key bits, the
				// transform undone. The ranks go to LDS as the inverse permutation (position ->
				// word), and the output is gathered from the words in order.
				uint32_t fin[KPT];  // (position << 16) | word index
#pragma unroll
				for (int j = 0; j < KPT; ++j) {
					const uint32_t i = j * THREADS + t;
					fin[j] = 0xFFFFFFFFu;
					if (i < len) {
						const W x = s_t[i];
						const uint32_t d = static_cast<uint32_t>(x >> kTop);
						const uint32_t lo = d ? s_cur[d - 1] : 0u, hi = s_cur[d];
						uint32_t n_less = 0;  // the word itself is in the run and not smaller
						for (uint32_t q = lo; q < hi; ++q) n_less += static_cast<uint32_t>(s_t[q] < x);
						fin[j] = ((lo + n_less) << 16) | i;
					}
				}
				__syncthreads();  // every run bound has been read: s_u takes the inverse permutation
				uint16_t *s_inv = reinterpret_cast<uint16_t *>(s_u);
#pragma unroll
				for (int j = 0; j < KPT; ++j)
					if (fin[j] != 0xFFFFFFFFu) s_inv[fin[j] >> 16] = static_cast<uint16_t>(fin[j]);
				__syncthreads();
				for (uint32_t o = t; o < len; o += THREADS)
					dst[o] = static_cast<K>(((s_t[s_inv[o]] >> 16) | pre) ^ static_cast<W>(xf.E));
			}
		}
	}
}

// The queued buckets (a run of more than kLocalRunMax keys share the partition bits): stable LDS LSD
// passes over the digits below the top two, from the bucket's original order.
template <typename K, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_local_lsd(const Control *ctrl, KeyXform xf, const uint32_t *__restrict__ seg_start,
	const uint32_t *__restrict__ bstart, const uint32_t *__restrict__ blen, const uint32_t *__restrict__ queue)
{
	constexpr int CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_whist[NW * kRadix];
	__shared__ uint32_t s_digoff[kRadix];
	__shared__ uint32_t s_wsum[4];
	__shared__ __align__(16) unsigned char s_buf[max_bytes(CAP * sizeof(K), rank_peer_bytes(THREADS))];
	K *s_raw = reinterpret_cast<K *>(s_buf);
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_buf);
	const PassPlan pp = ctrl->hyb[2];
	if (!pp.active) return;
	const uint32_t queued = ctrl->hyb_queue;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	for (uint32_t qi = blockIdx.x; qi < queued; qi += gridDim.x) {
		const uint32_t b = queue[qi];
		const uint32_t len = blen[b], start = seg_start[b >> 8] + bstart[b];
		const K *__restrict__ src = static_cast<const K *>(pp.src) + start;
		K *__restrict__ dst = static_cast<K *>(pp.dst) + start;
		K key[KPT];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			key[j] = l < len ? src[l] : K(0);
		}
		uint32_t dig[KPT], pos[KPT], total;
		for (int p = 0; p <= P - 3; ++p) {
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				dig[j] = l < len ? digit_of<K>(key[j], xf, 8u * p) : 0xFFu;
			}
			block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_raw[pos[j]] = key[j];
			__syncthreads();
#pragma unroll
			for (int j = 0; j < KPT; ++j) key[j] = s_raw[w * 64u * KPT + j * 64u + lane];
			__syncthreads();
		}
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) dst[l] = key[j];
		}
	}
}

// ---------------------------------------------------------------------------------------------
// 8-bit count fill and finish

__global__ __launch_bounds__(256) void k_fill8(const uint32_t *__restrict__ bucket, KeyXform xf,
	uint8_t *__restrict__ out, uint64_t n)
{
	__shared__ uint32_t s_b[kRadix + 1];
	__shared__ uint8_t s_inv[kRadix];
	s_b[threadIdx.x] = bucket[threadIdx.x];
	if (threadIdx.x == 0) s_b[kRadix] = 0xFFFFFFFFu;  // past every position (n < 2^32)
	// the bijective 8-bit transforms are invertible: bucket f(x) holds only the pattern x
	s_inv[sort_key<uint8_t>(static_cast<uint8_t>(threadIdx.x), xf) & 0xFFu] = static_cast<uint8_t>(threadIdx.x);
	__syncthreads();
	// sixteen consecutive positions per thread, written with one 16-byte store
	const bool vec = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
	const uint64_t groups = (n + 15) / 16;
	const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < groups; g += stride) {
		const uint64_t i0 = g * 16;
		// last bucket whose start is <= i0, then onwards past the buckets that end at or before i
		uint32_t lo = 0;
#pragma unroll
		for (uint32_t step = 128; step; step >>= 1)
			if (s_b[lo + step] <= i0) lo += step;
		uint8_t v[16];
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			while (lo < kRadix - 1 && s_b[lo + 1] <= i0 + j) ++lo;
			v[j] = s_inv[lo];
		}
		if (vec && i0 + 16 <= n) {
			uint4 u;
			__builtin_memcpy(&u, v, 16);
			*reinterpret_cast<uint4 *>(out + i0) = u;
		} else {
			for (int j = 0; j < 16; ++j)
				if (i0 + j < n) out[i0 + j] = v[j];
		}
	}
}

// ---------------------------------------------------------------------------------------------
// 16-bit count fill (keys only, bijective modes): as in the 8-bit fill, the sorted output follows
// from the counts of the sort keys alone, here all 65536 of them. It replaces the two stable 8-bit
// passes the reference runs for 16-bit keys (digit plan H:1848, passes H:31774-31982); the output
// is the same sequence.

constexpr uint32_t kFill16Half = 32768;    // sort keys per LDS histogram (128 KB)
// smallest sort that takes the fill: measured on uniform keys, at 2^21 the fill and the two passes
// take the same time, at 2^22 the fill takes 0.75x and at 2^28 0.23x
constexpr uint64_t kFill16Min = 1u << 22;
constexpr uint32_t kFill16Words = 65536 + 65536 + 65536 / 2;  // hist, starts, inverse (u16); then the chunk map

// Block b counts the sort keys k with k >> 15 == b & 1 (one half of the key space) over chunk
// b >> 1 of the input, in LDS, and adds its nonzero counts to hist (65536 words, zero on entry).
// Every key is read by two blocks: reading twice costs less than counters that do not fit LDS.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_hist16(const uint16_t *__restrict__ in, uint64_t n, KeyXform xf,
	uint32_t *__restrict__ hist)
{
	__shared__ uint32_t s_h[kFill16Half];
	for (uint32_t i = threadIdx.x; i < kFill16Half; i += THREADS) s_h[i] = 0;
	__syncthreads();
	const uint32_t half = blockIdx.x & 1u, chunks = gridDim.x >> 1, c = blockIdx.x >> 1;
	const uint64_t per = ((n + chunks - 1) / chunks + 63) & ~63ull;
	const uint64_t lo = static_cast<uint64_t>(c) * per;
	if (lo < n) {
		const uint32_t len = static_cast<uint32_t>((n - lo) < per ? n - lo : per);
		for_each_key<uint16_t, THREADS>(in, lo, len, [&](uint16_t x) {
			const uint32_t k = sort_key<uint16_t>(x, xf) & 0xFFFFu;
			if ((k >> 15) == half) wave_count<1>(s_h, k & (kFill16Half - 1u));
		});
	}
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < kFill16Half; i += THREADS)
		if (s_h[i]) atomicAdd(&hist[half * kFill16Half + i], s_h[i]);
}

// One block: start[v] = the number of keys with a sort key below v, and inv[f(x)] = x (the
// bijective transform's inverse).
__global__ __launch_bounds__(1024) void k_scan16(const uint32_t *__restrict__ hist, uint32_t *__restrict__ start,
	uint16_t *__restrict__ inv, KeyXform xf)
{
	__shared__ uint32_t s_w[16];
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	uint32_t sum = 0;
	for (uint32_t j = 0; j < 64; ++j) sum += hist[t * 64u + j];
	const uint32_t incl = wave_inclusive_sum(sum);
	if (lane == 63) s_w[w] = incl;
	__syncthreads();
	uint32_t run = incl - sum;
	for (uint32_t v = 0; v < w; ++v) run += s_w[v];
	for (uint32_t j = 0; j < 64; ++j) {
		start[t * 64u + j] = run;
		run += hist[t * 64u + j];
	}
	for (uint32_t x = t; x < 65536u; x += 1024u) inv[sort_key<uint16_t>(static_cast<uint16_t>(x), xf) & 0xFFFFu] = static_cast<uint16_t>(x);
}

constexpr uint32_t kFill16Chunk = 256 * 8;  // output positions per block and step of k_fill16

// largest bucket v in [lo, hi) with start[v] <= i, given start[lo] <= i
__device__ __forceinline__ uint32_t fill16_bucket(const uint32_t *__restrict__ start, uint32_t lo, uint32_t hi, uint64_t i)
{
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (start[mid] <= i) lo = mid;
		else hi = mid;
	}
	return lo;
}

// map[c] = the bucket holding position c * kFill16Chunk for the chunks c < chunks, map[chunks] =
// 65535: chunk c's positions lie in the buckets map[c] .. map[c + 1]. One thread per chunk, so the
// searches' memory latencies overlap.
__global__ __launch_bounds__(256) void k_map16(const uint32_t *__restrict__ start, uint32_t *__restrict__ map, uint32_t chunks)
{
	const uint32_t c = blockIdx.x * 256u + threadIdx.x;
	if (c < chunks) map[c] = fill16_bucket(start, 0, 65536, static_cast<uint64_t>(c) * kFill16Chunk);
	else if (c == chunks) map[c] = 65535;
}

// out[i] = inv[v] for the bucket v that holds position i (the last v with start[v] <= i), eight
// consecutive positions per thread, searching only the buckets of the chunk.
__global__ __launch_bounds__(256) void k_fill16(const uint32_t *__restrict__ start, const uint32_t *__restrict__ map,
	const uint16_t *__restrict__ inv, uint16_t *__restrict__ out, uint64_t n)
{
	const uint64_t nchunks = (n + kFill16Chunk - 1) / kFill16Chunk;
	const bool vec = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
	for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
		const uint64_t i0 = ch * kFill16Chunk + threadIdx.x * 8ull;
		if (i0 >= n) continue;
		const uint32_t vlo = map[ch], vend = map[ch + 1] + 1;
		uint32_t v = fill16_bucket(start, vlo, vend, i0);
		uint16_t val[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			const uint64_t i = i0 + j;
			if (v + 1 < vend && start[v + 1] <= i) v = fill16_bucket(start, v + 1, vend, i);
			val[j] = inv[v];
		}
		if (vec && i0 + 8 <= n) {
			uint4 u;
			__builtin_memcpy(&u, val, 16);
			*reinterpret_cast<uint4 *>(out + i0) = u;
		} else {
			for (int j = 0; j < 8; ++j)
				if (i0 + j < n) out[i0 + j] = val[j];
		}
	}
}

__device__ __forceinline__ void reset_control(Control *ctrl)
{
	if (threadIdx.x < kMaxPasses) {
		ctrl->tile_counter[threadIdx.x] = 0;
		ctrl->scan_counter[threadIdx.x] = 0;
	}
	if (threadIdx.x < kStopLines) ctrl->check_stop[threadIdx.x][0] = 0;
	if (threadIdx.x == 0) {
		ctrl->epoch = ctrl->epoch + 1u;
		ctrl->check = 0;
		ctrl->hyb_scan_counter = 0;
		ctrl->hyb_queue = 0;
		ctrl->need_full = 0;
		ctrl->joint_max = 0;
		ctrl->bucket_count = 0;
		ctrl->bucket_count2 = 0;
		ctrl->zero_signs = 0;
	}
}

__global__ void k_finish(Control *ctrl) { reset_control(ctrl); }

// the histograms (which k_plan leaves zeroed), the overflow counts of the top-16-bit histogram
// (which k_joint_reduce leaves zeroed) and the control block, whatever state they are in
__global__ __launch_bounds__(256) void k_reset(Control *ctrl, uint32_t *hist, uint32_t *ovf)
{
	for (uint32_t i = threadIdx.x; i < kMaxPasses * kRadix; i += blockDim.x) hist[i] = 0;
	for (uint32_t i = threadIdx.x; i < kJointBins; i += blockDim.x) ovf[i] = 0;
	reset_control(ctrl);
}

__global__ void k_zero(uint64_t *p, uint64_t words)
{
	const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0;
}

// ---------------------------------------------------------------------------------------------
// host side

// Optional per-launch HIP-event timing on the launch stream (used by the benchmark).
struct ProfRec {
	int kind;  // 0 histogram, 1 scatter pass, 3 count + scan of a pass
	int pass;
	const void *ws;
	hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_prof_pool;  // events of earlier records, reused (creating them is slow)

hipEvent_t prof_event()
{
	if (!g_prof_pool.empty()) {
		hipEvent_t e = g_prof_pool.back();
		g_prof_pool.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

struct ProfScope {
	bool on;
	ProfRec rec;
	hipStream_t stream;
	ProfScope(int kind, int pass, const void *ws, hipStream_t s) : on(false), rec{kind, pass, ws, nullptr, nullptr}, stream(s)
	{
		std::lock_guard<std::mutex> lock(g_prof_mu);
		if (!g_prof_on) return;
		rec.a = prof_event();
		rec.b = prof_event();
		if (!rec.a || !rec.b) {
			if (rec.a) g_prof_pool.push_back(rec.a);
			if (rec.b) g_prof_pool.push_back(rec.b);
			return;
		}
		on = hipEventRecord(rec.a, s) == hipSuccess;
		if (!on) { g_prof_pool.push_back(rec.a); g_prof_pool.push_back(rec.b); }
	}
	~ProfScope()
	{
		if (!on) return;
		std::lock_guard<std::mutex> lock(g_prof_mu);
		if (hipEventRecord(rec.b, stream) == hipSuccess) {
			g_prof.push_back(rec);
		} else {
			(void)hipEventSynchronize(rec.a);
			g_prof_pool.push_back(rec.a);
			g_prof_pool.push_back(rec.b);
		}
	}
};

struct DevInfo {
	int cus = 256;
	int occ[2][4][kNumTileCfgs] = {};  // [payload][key size log2][cfg] resident blocks per CU
};

DevInfo &devinfo()
{
	static DevInfo info;
	static std::once_flag once;
	std::call_once(once, [] {
		int dev = 0;
		if (hipGetDevice(&dev) == hipSuccess) {
			hipDeviceProp_t prop;
			if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) info.cus = prop.multiProcessorCount;
		}
	});
	return info;
}

template <typename K, bool PAY, int THREADS, int KPT>
hipError_t launch_pass(Control *ctrl, const uint32_t *bucket, uint64_t *lookback, uint64_t n, const KeyXform &xf,
	int pass, hipStream_t stream)
{
	constexpr int TILE = THREADS * KPT;
	const uint32_t tiles = static_cast<uint32_t>((n + TILE - 1) / TILE);
	static int occ = -1;
	if (occ < 0) {
		int o = 0;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_onesweep<K, PAY, THREADS, KPT>, THREADS, 0) != hipSuccess || o < 1) o = 1;
		occ = o;
	}
	const uint32_t grid = std::min<uint32_t>(tiles, static_cast<uint32_t>(occ * devinfo().cus));
	hipLaunchKernelGGL((k_onesweep<K, PAY, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, stream, ctrl, bucket, lookback, n,
		tiles, xf, pass);
	return hipGetLastError();
}

template <typename K, bool PAY>
hipError_t launch_pass_cfg(int cfg, Control *ctrl, const uint32_t *bucket, uint64_t *lookback, uint64_t n,
	const KeyXform &xf, int pass, hipStream_t stream)
{
	switch (cfg) {
	case 0: return launch_pass<K, PAY, 256, 8>(ctrl, bucket, lookback, n, xf, pass, stream);
	case 1: return launch_pass<K, PAY, 256, 16>(ctrl, bucket, lookback, n, xf, pass, stream);
	case 2: return launch_pass<K, PAY, 512, 8>(ctrl, bucket, lookback, n, xf, pass, stream);
	case 3: return launch_pass<K, PAY, 1024, 12>(ctrl, bucket, lookback, n, xf, pass, stream);
	default: return launch_pass<K, PAY, 256, 24>(ctrl, bucket, lookback, n, xf, pass, stream);
	}
}

template <typename K, bool PAY, int THREADS, int KPT, bool SEG = false, bool GROUP = false, bool CUR = false, int SPL = 0>
hipError_t launch_scatter(const PassPlan *pp, uint32_t shift, const KeyXform &xf, uint64_t n, uint32_t tiles,
	const uint32_t *offsets, SegMap seg, hipStream_t stream, uint32_t *cursor = nullptr, const uint32_t *split = nullptr,
	const uint32_t *gsel = nullptr)
{
	static int occ = -1;
	if (occ < 0) {
		int o = 0;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_scatter<K, PAY, THREADS, KPT, SEG, GROUP, CUR, SPL>, THREADS, 0) !=
				hipSuccess || o < 1)
			o = 1;
		occ = o;
	}
	const uint32_t per = (tiles + 7) / 8;
	const uint32_t want = static_cast<uint32_t>(occ * devinfo().cus * 2 + 7) / 8;
	const uint32_t grid = 8u * std::max<uint32_t>(1u, std::min<uint32_t>(per, want));
	hipLaunchKernelGGL((k_scatter<K, PAY, THREADS, KPT, SEG, GROUP, CUR, SPL>), dim3(grid), dim3(THREADS), 0, stream, pp, shift, xf,
		n, tiles, offsets, seg, cursor, split, gsel);
	return hipGetLastError();
}

// the transform only complements bits: equal sort keys are equal keys
inline bool xform_injective(const KeyXform &xf, int key_bytes)
{
	const uint64_t ones = key_bytes >= 8 ? ~0ull : ((1ull << (8 * key_bytes)) - 1ull);
	return xf.A == ones && xf.sh == 0 && xf.B == 0 && xf.C == 0 && xf.D == 0;
}

template <typename K, bool PAY>
hipError_t launch_scatter_cfg(int cfg, Control *ctrl, int pass, const KeyXform &xf, uint64_t n, uint32_t tiles,
	const uint32_t *offsets, hipStream_t stream)
{
	const PassPlan *pp = &ctrl->pass[pass];
	const uint32_t sh = 8u * static_cast<uint32_t>(pass);
	const SegMap none{};
	switch (cfg) {
	case 0: return launch_scatter<K, PAY, 256, 8>(pp, sh, xf, n, tiles, offsets, none, stream);
	case 1: return launch_scatter<K, PAY, 256, 16>(pp, sh, xf, n, tiles, offsets, none, stream);
	case 2: return launch_scatter<K, PAY, 512, 8>(pp, sh, xf, n, tiles, offsets, none, stream);
	case 3: return launch_scatter<K, PAY, 1024, 12>(pp, sh, xf, n, tiles, offsets, none, stream);
	default: return launch_scatter<K, PAY, 256, 24>(pp, sh, xf, n, tiles, offsets, none, stream);
	}
}

// Capacities (keys) of the per-bucket LDS sort: 256x4, 256x8, 256x12 (512x6 for 8-byte keys), 512x9,
// 256x19, 512x16, 1024x12 (2^29 keys: u64 7.3 ms, u32 4.2, against 27.1 and 9.2 for the LSD
// passes). Eight waves per block where the LDS allows three blocks per CU: the block's
// phases between barriers are short, and more waves per block hide more of their LDS latency
// (512x9 sorts 2^28 u64 keys' buckets in 0.84 ms against 1.17 ms for 256x19).
constexpr int kNumLocalCfgs = 7;
constexpr uint32_t kLocalCaps[kNumLocalCfgs] = {1024, 2048, 3072, 4608, 4864, 8192, 12288};

// the smallest capacity that holds the largest of 65536 buckets of uniformly spread keys
// (mean + 6 standard deviations + slack); -1 when none does
int local_cfg_for(uint64_t n)
{
	const double mean = static_cast<double>(n) / kHybBuckets;
	const double need = mean + 6.0 * std::sqrt(mean) + 64.0;
	for (int c = 0; c < kNumLocalCfgs; ++c)
		if (need <= kLocalCaps[c]) return c;
	return -1;
}

template <typename K, int THREADS, int KPT, int PB>
hipError_t launch_local_cfg(Control *ctrl, const KeyXform &xf, const SegTables &T, hipStream_t stream)
{
	hipLaunchKernelGGL((k_local<K, THREADS, KPT, PB>), dim3(kHybBuckets), dim3(THREADS), 0, stream, ctrl, xf, T.seg_start,
		T.bucket_start, T.bucket_len, T.queue);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	const uint32_t lgrid = static_cast<uint32_t>(4 * devinfo().cus);
	hipLaunchKernelGGL((k_local_lsd<K, THREADS, KPT>), dim3(lgrid), dim3(THREADS), 0, stream, ctrl, xf, T.seg_start,
		T.bucket_start, T.bucket_len, T.queue);
	return hipGetLastError();
}

// capacity and partition width per configuration (kLocalCaps): one to three keys per partition
// run at the mean bucket size each capacity is chosen for (12 bits for 512x9: its LDS, with the run
// bounds read from the cursors, still lets three blocks share a CU; 0.83 against 0.85 ms at 2^28 u64)
template <typename K>
hipError_t launch_local(int lcfg, Control *ctrl, const KeyXform &xf, const SegTables &T, hipStream_t stream)
{
	constexpr bool kWide = sizeof(K) == 8;
	switch (lcfg) {
	case 0: return launch_local_cfg<K, 256, 4, 9>(ctrl, xf, T, stream);
	case 1: return launch_local_cfg<K, 256, 8, 10>(ctrl, xf, T, stream);
	case 2: return launch_local_cfg<K, kWide ? 512 : 256, kWide ? 6 : 12, 10>(ctrl, xf, T, stream);
	case 3: return launch_local_cfg<K, 512, 9, 12>(ctrl, xf, T, stream);
	case 4: return launch_local_cfg<K, 256, 19, 11>(ctrl, xf, T, stream);
	case 5: return launch_local_cfg<K, 512, 16, 11>(ctrl, xf, T, stream);
	default: return launch_local_cfg<K, 1024, 12, 12>(ctrl, xf, T, stream);
	}
}

// The two-level MSD plan's launches after k_plan (each exits at once when the plan did not choose
// it). The read was k_hist_joint: the bucket sizes are known and the plan checked them, so both
// partitions take their output ranges from cursors (the range cursors, then the bucket cursors)
// and need no count or scan of their own.
template <typename K, int THREADS, int KPT>
hipError_t launch_hybrid(Control *ctrl, const SegTables &T, const KeyXform &xf, uint64_t n, uint32_t tiles, int lcfg,
	void *ws, hipStream_t stream)
{
	constexpr int P = sizeof(K);
	const uint32_t tiles2 = tiles + kRadix;  // bound: one partial tile per segment
	const SegMap m{T.tile_start, T.seg_start, T.seg_len};
	hipError_t e;
	{
		ProfScope ps(5, 0, ws, stream);
		e = launch_scatter<K, false, THREADS, KPT, false, true, true>(&ctrl->hyb[0], 8u * (P - 1), xf, n, tiles, nullptr,
			SegMap{}, stream, T.rcnt);
		if (e != hipSuccess) return e;
	}
	{
		ProfScope ps(5, 1, ws, stream);
		e = launch_scatter<K, false, THREADS, KPT, true, true, true>(&ctrl->hyb[1], 8u * (P - 2), xf, n, tiles2, nullptr, m,
			stream, T.cursor);
		if (e != hipSuccess) return e;
	}
	ProfScope ps(6, 0, ws, stream);
	return launch_local<K>(lcfg, ctrl, xf, T, stream);
}

template <typename K>
hipError_t launch_hybrid_cfg(int cfg, Control *ctrl, const SegTables &T, const KeyXform &xf, uint64_t n, uint32_t tiles,
	int lcfg, void *ws, hipStream_t stream)
{
	switch (cfg) {
	case 0: return launch_hybrid<K, 256, 8>(ctrl, T, xf, n, tiles, lcfg, ws, stream);
	case 1: return launch_hybrid<K, 256, 16>(ctrl, T, xf, n, tiles, lcfg, ws, stream);
	case 2: return launch_hybrid<K, 512, 8>(ctrl, T, xf, n, tiles, lcfg, ws, stream);
	case 3: return launch_hybrid<K, 1024, 12>(ctrl, T, xf, n, tiles, lcfg, ws, stream);
	case 4: return launch_hybrid<K, 256, 24>(ctrl, T, xf, n, tiles, lcfg, ws, stream);
	default: return launch_hybrid<K, 1024, 16>(ctrl, T, xf, n, tiles, lcfg, ws, stream);
	}
}

#include "rsbd8_sample.hpp"

// Mean bucket length the sample partition plan aims its second partition at (half the capacity of
// its per-bucket sort, so that the sampling's spread rarely sends a bucket through HBM), unless a
// process-wide setting overrides it.
constexpr uint32_t kSampleTarget = 4096;
std::atomic<uint32_t> g_sample_target{0};

// The sample partition plan (rsbd8_sample.hpp) with THREADS x KPT partition tiles.
// whether a presort check may end in the run reversal: keys only (the plan entry point's sorts), and
// it writes out from in by positions, so not in place
inline bool runs_allowed(const SortJob &job)
{
	return job.in != job.out && !job.keys_final_any && !job.payload;
}

// the presort check and, when it finds the input in reverse order with ties, the run reversal (every
// later launch of the sort then exits at once)
template <typename K>
hipError_t launch_check(const SortJob &job, void *ws, const KeyXform &xf, hipStream_t stream)
{
	const uint64_t n = job.n;
	const int cus = devinfo().cus;
	Control *ctrl = ctrl_of(ws);
	const K *in = static_cast<const K *>(job.in);
	const int rev = (job.flags & kRev) ? 1 : 0;
	const bool runs = runs_allowed(job);
	const uint32_t tiles = static_cast<uint32_t>((n + kRunTile - 1) / kRunTile);
	// four blocks per CU (at most kCheckSlots / 4): each wave's verdict goes to a slot of its own
	const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(tiles, std::min(4u * cus, kCheckSlots / kRunWaves))));
	uint32_t *counts = counts_of(ws);
	hipLaunchKernelGGL((k_check<K, 256>), dim3(grid), dim3(256), 0, stream, in, n, xf, ctrl, counts, runs ? 1 : 0, tiles,
		runs ? static_cast<K *>(job.out) : nullptr, rev);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess || n < 2) return e;
	hipLaunchKernelGGL((k_check_fold<1024>), dim3(1), dim3(1024), 0, stream, ctrl, counts, grid * kRunWaves);
	if ((e = hipGetLastError()) != hipSuccess || !runs) return e;
	const uint32_t cblocks = (tiles + kRunBlock - 1) / kRunBlock;  // at most 1024 (n < 2^32)
	hipLaunchKernelGGL((k_runs_reduce<kRunBlock>), dim3(cblocks), dim3(kRunBlock), 0, stream, ctrl, rev, n, counts, tiles);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	hipLaunchKernelGGL((k_runs_carry<kRunBlock>), dim3(cblocks), dim3(kRunBlock), 0, stream, ctrl, rev, n, counts, tiles);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	// one block per tile (a tile loop in the block needs four times the registers): when the verdict is
	// another, the blocks exit at once (about 17 us for 2^16 tiles)
	hipLaunchKernelGGL((k_runs_copy<K, 256>), dim3(tiles), dim3(256), 0, stream, in, static_cast<K *>(job.out), n, xf, ctrl, rev,
		counts, tiles);
	return hipGetLastError();
}

// Whether the sample partition plan may group instead of rank (PlanArgs::group): keys-only sorts
// whose transform is one to one (unsigned, signed, the tiered absolute values and the special
// float order: equal sort keys are equal keys, so no order among them can be seen), and the float
// order, where +0.0 and -0.0 alone share a sort key (the first read checks that they do not both
// occur); never with a payload, nor for the absolute-value modes.
int sample_group_class(const SortJob &job)
{
	if (job.payload || job.keys_final_any) return 0;
	switch (job.flags & 7u) {
	case kAbs | kSign:
	case kAbs | kSign | kFltp: return 0;
	case kSign | kFltp: return 2;
	default: return 1;
	}
}

template <typename K, bool PAY, int THREADS, int KPT>
hipError_t launch_sample_cfg(const SortJob &job, void *ws, const KeyXform &xf, hipStream_t stream)
{
	constexpr uint32_t TILE = THREADS * KPT;
	const uint64_t n = job.n;
	Control *ctrl = ctrl_of(ws);
	uint32_t *hist = hist_of(ws), *bucket = bucket_of(ws), *counts = counts_of(ws), *split = splittab_of(ws);
	uint64_t *lookback = lookback_of(ws);
	const SegTables T = seg_tables(segtab_of(ws));
	const int cus = devinfo().cus;
	const int rev = (job.flags & kRev) ? 1 : 0;
	const uint32_t tiles = static_cast<uint32_t>((n + TILE - 1) / TILE);
	const K *in = static_cast<const K *>(job.in);
	const uint32_t st = g_sample_target.load();
	const uint32_t target = st ? st : kSampleTarget;
	// Two per-bucket sorts (kBucketCapA keys in 512-thread blocks, then the rest) for 8-byte keys,
	// whose 8192-key blocks hold one per CU (2^28 f64 N(0,1): 2.97 ms in the bucket sorts against
	// 3.63 with the large one only, when the smaller blocks had 256 threads; 2.2 ms with 512); 4-byte
	// keys keep the large one (2^28 f32 + index: 1.89 ms against 2.17 with both).
	constexpr bool kTiers = sizeof(K) == 8;
	const int group = sample_group_class(job);
	hipError_t e;
	if (job.presort_check && (e = launch_check<K>(job, ws, xf, stream)) != hipSuccess) return e;
	{
		ProfScope ps(0, 0, ws, stream);
		// chunks of at least kSampleMax keys each, a power of two (every chunk keeps kSampleMax / G samples)
		uint32_t G = 1;
		while (G < kL1Chunks && static_cast<uint64_t>(2 * G) * kSampleMax <= n) G *= 2;
		hipLaunchKernelGGL((k_samp_chunk<K, 1024>), dim3(G), dim3(1024), 0, stream, in, n, xf, ctrl, rev, split, nullptr);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		hipLaunchKernelGGL((k_samp_l1<K, 1024>), dim3(1), dim3(1024), 0, stream, ctrl, rev, split, nullptr);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(tiles, 4u * cus));
		if (group == 2)
			hipLaunchKernelGGL((k_split_count1<K, 256, true>), dim3(grid), dim3(256), 0, stream, in, n, xf, hist, counts, TILE, tiles,
				rev, ctrl, split, nullptr);
		else
			hipLaunchKernelGGL((k_split_count1<K, 256, false>), dim3(grid), dim3(256), 0, stream, in, n, xf, hist, counts, TILE,
				tiles, rev, ctrl, split, nullptr);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	PlanArgs pa{};
	pa.n = n; pa.passes = static_cast<int>(sizeof(K)); pa.src = job.in; pa.fin = job.out; pa.other = job.buf;
	pa.psrc = job.payload ? job.pin : nullptr;
	pa.pfin = job.payload ? job.pout : nullptr;
	pa.pother = job.payload ? job.pbuf : nullptr;
	pa.payload = job.payload;
	pa.reverse = rev;
	pa.keys_final_any = job.keys_final_any;
	pa.tile2 = TILE;
	pa.segtab = segtab_of(ws);
	pa.gate = nullptr;
	pa.group = group;
	pa.split0 = split;
	hipLaunchKernelGGL(k_plan_sample, dim3(1), dim3(256), 0, stream, pa, ctrl, hist, bucket);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	{
		ProfScope ps(7, 0, ws, stream);
		const uint32_t nchunks = (tiles + kScanChunk - 1) / kScanChunk;
		const uint32_t sgrid = std::max<uint32_t>(1u, std::min<uint32_t>(nchunks, 4u * cus));
		hipLaunchKernelGGL(k_scan, dim3(sgrid), dim3(256), 0, stream, ctrl, &ctrl->hyb[0], &ctrl->hyb_scan_counter, bucket,
			kMaxPasses, counts, lookback, tiles);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	// the partitions' forms: ranked (group 0), grouping (group 1), or both, each gated on the plan's
	// choice (group 2; 2^28 f64 N(0,1) keys: 6.98 ms grouped against 7.50 ranked, f32 4.80 against
	// 5.61)
	const uint32_t *gsel = group == 2 ? &ctrl->group_parts : nullptr;
	{
		ProfScope ps(5, 0, ws, stream);
		if constexpr (!PAY) {
			if (group) {
				e = launch_scatter<K, false, THREADS, KPT, false, true, false, 1>(&ctrl->hyb[0], 0, xf, n, tiles, counts, SegMap{},
					stream, nullptr, split, gsel);
				if (e != hipSuccess) return e;
			}
		}
		if (group != 1) {
			e = launch_scatter<K, PAY, THREADS, KPT, false, false, false, 1>(&ctrl->hyb[0], 0, xf, n, tiles, counts, SegMap{},
				stream, nullptr, split, gsel);
			if (e != hipSuccess) return e;
		}
	}
	const uint32_t tiles2 = tiles + kRadix;  // bound: one partial tile per segment
	const SegMap m{T.tile_start, T.seg_start, T.seg_len};
	{
		ProfScope ps(7, 1, ws, stream);
		hipLaunchKernelGGL((k_samp_l2<K, 1024>), dim3(kRadix), dim3(1024), 0, stream, ctrl, xf, T, split, target);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		const uint32_t cgrid = std::max<uint32_t>(1u, std::min<uint32_t>(tiles2, 8u * cus));
		hipLaunchKernelGGL((k_split_count2<K, 256>), dim3(cgrid), dim3(256), 0, stream, &ctrl->hyb[1], xf, TILE, m, counts,
			split + kSplitTabWords);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		hipLaunchKernelGGL(k_split_segscan, dim3(kRadix), dim3(256), 0, stream, ctrl, T, counts, split + kSplitTabWords,
			kTiers ? kBucketCapA : 0u);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	{
		ProfScope ps(5, 1, ws, stream);
		if constexpr (!PAY) {
			if (group) {
				e = launch_scatter<K, false, THREADS, KPT, true, true, false, 2>(&ctrl->hyb[1], 0, xf, n, tiles2, counts, m, stream,
					nullptr, split + kSplitTabWords, gsel);
				if (e != hipSuccess) return e;
			}
		}
		if (group != 1) {
			e = launch_scatter<K, PAY, THREADS, KPT, true, false, false, 2>(&ctrl->hyb[1], 0, xf, n, tiles2, counts, m, stream,
				nullptr, split + kSplitTabWords, gsel);
			if (e != hipSuccess) return e;
		}
	}
	{
		ProfScope ps(6, 0, ws, stream);
		// one block per queued bucket: a segment of len keys has at most 1.5 * ceil(len / target) + 1
		// non-empty buckets, so 2 n / target + 512 blocks cover the queue of the smaller ones, and
		// n / kBucketCapA + 1 the queue of the larger ones (the blocks past a queue exit at once)
		if constexpr (kTiers) {
			const uint32_t agrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(std::min<uint64_t>(n, kHybBuckets),
				2 * n / target + 512)));
			hipLaunchKernelGGL((k_bucket_sort<K, PAY, 512, kBucketCapA / 512, 11>), dim3(agrid), dim3(512), 0, stream, ctrl, xf, T,
				T.bq, &ctrl->bucket_count);
			if ((e = hipGetLastError()) != hipSuccess) return e;
		}
		const uint32_t bgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(std::min<uint64_t>(n, kHybBuckets),
			kTiers ? n / kBucketCapA + 1 : 2 * n / target + 512)));
		// above 2^28 keys the second level's 256 buckets per segment average 8192 keys or more: the
		// larger buckets then take 1024x12 (8-byte) or 1024x16 (4-byte) blocks, one per CU, rather
		// than LSD passes through HBM (2^29 u64 keys: 26.2 ms with 8192-key blocks)
		if (n > (1ull << 28)) {
			constexpr int KB = sizeof(K) == 8 ? 12 : 16;
			hipLaunchKernelGGL((k_bucket_sort<K, PAY, 1024, KB, 12>), dim3(bgrid), dim3(1024), 0, stream, ctrl, xf, T,
				T.bq + kHybBuckets, &ctrl->bucket_count2);
		} else {
			hipLaunchKernelGGL((k_bucket_sort<K, PAY, 1024, 8, 11>), dim3(bgrid), dim3(1024), 0, stream, ctrl, xf, T,
				T.bq + kHybBuckets, &ctrl->bucket_count2);
		}
		if ((e = hipGetLastError()) != hipSuccess) return e;
		hipLaunchKernelGGL((k_bucket_hard<K, PAY, 1024, 8>), dim3(2 * cus), dim3(1024), 0, stream, ctrl, xf, T);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	const uint32_t cgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8u * cus)));
	hipLaunchKernelGGL((k_copy<K>), dim3(cgrid), dim3(256), 0, stream, ctrl, 1, n, job.payload ? 1 : 0);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, stream, ctrl);
	return hipGetLastError();
}

// The partitions take 1024x8 tiles whatever the configuration (but see below for 8-byte keys that
// may group): with the splitter table in LDS, the ranked form spills registers in 1024x12 and
// larger tiles. Against 512x8, the larger tiles halve the tile count
// rows the counting reads write and the scans read (2^28 f32 with an index payload: 7.13 -> 6.50
// ms; u64 with a payload: 8.27 -> 7.44 ms; f64 N(0,1) keys, grouped: 7.33 -> 6.98 ms, though its
// ranked form, 104 KB of LDS per block, goes from 7.33 to 7.50).
template <typename K>
hipError_t launch_sample(int cfg, const SortJob &job, void *ws, hipStream_t stream)
{
	(void)cfg;
	const KeyXform xf = make_xform(8 * sizeof(K), job.flags);
	if (job.payload) return launch_sample_cfg<K, true, 1024, 8>(job, ws, xf, stream);
	if constexpr (sizeof(K) == 8) {
		// keys-only 8-byte sorts that may group take 1024x12 tiles, whose grouping form fits its
		// registers (2^28 f64 N(0,1): 6.13 ms against 6.35), though the ranked form, run when +0.0
		// and -0.0 both occur, then spills (7.17 against 6.91); 4-byte keys gain nothing from them
		if (sample_group_class(job) != 0) return launch_sample_cfg<K, false, 1024, 12>(job, ws, xf, stream);
	}
	return launch_sample_cfg<K, false, 1024, 8>(job, ws, xf, stream);
}

// One side stream per device (default priority, not synchronising with the null stream) and the
// events of the fork and the join. The events are recorded and waited for under a lock, so that
// every wait sees the record it was paired with.
struct SideStream {
	hipStream_t s = nullptr;
	hipEvent_t fork = nullptr, join = nullptr;
};
std::mutex g_side_mu;
SideStream g_side[64];

// side = a stream whose later work starts after everything enqueued on stream so far
hipError_t fork_side(hipStream_t stream, hipStream_t *side)
{
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess) return e;
	if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lock(g_side_mu);
	SideStream &ss = g_side[dev];
	if (!ss.s) {
		if ((e = hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking)) != hipSuccess) return e;
		if ((e = hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming)) != hipSuccess) return e;
		if ((e = hipEventCreateWithFlags(&ss.join, hipEventDisableTiming)) != hipSuccess) return e;
	}
	if ((e = hipEventRecord(ss.fork, stream)) != hipSuccess) return e;
	if ((e = hipStreamWaitEvent(ss.s, ss.fork, 0)) != hipSuccess) return e;
	*side = ss.s;
	return hipSuccess;
}

// stream's later work starts after everything enqueued on side so far
hipError_t join_side(hipStream_t side, hipStream_t stream)
{
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess) return e;
	std::lock_guard<std::mutex> lock(g_side_mu);
	SideStream &ss = g_side[dev];
	if ((e = hipEventRecord(ss.join, side)) != hipSuccess) return e;
	return hipStreamWaitEvent(stream, ss.join, 0);
}

template <typename K>
hipError_t run_sort(const SortJob &job, void *ws, hipStream_t stream)
{
	constexpr int P = sizeof(K);
	const uint64_t n = job.n;
	const KeyXform xf = make_xform(8 * sizeof(K), job.flags);
	Control *ctrl = ctrl_of(ws);
	uint32_t *hist = hist_of(ws), *bucket = bucket_of(ws);
	uint64_t *lookback = lookback_of(ws);
	const int cus = devinfo().cus;
	hipError_t e;

	constexpr int HITEMS = 16 / sizeof(K) < 4 ? 4 : 16 / sizeof(K) * 2;
	const bool bijective = (job.flags & 7u) != (kAbs | kSign) && (job.flags & 7u) != (kSign | kFltp) &&
		(job.flags & 7u) != (kAbs | kSign | kFltp);
	if constexpr (P == 2) {
		if (!job.payload && job.fill8 && bijective && n >= kFill16Min && (!job.algo || job.algo == kAlgoHybrid)) {
			// 16-bit count fill: the automatic plan, or algorithm 4 (the plan that is not a chain of
			// LSD passes; forcing onesweep or count + scan keeps those passes). Its tables borrow
			// the first tile records (look-back rows included), which are zeroed again afterwards,
			// as a fresh workspace has them.
			// tables plus the chunk map fit the records: 640 KB + 4 B per 2048 keys <= 1.5 B per key
			static_assert(kFill16Words * 4 + (kFill16Min / kFill16Chunk + 2) * 4 <= (kFill16Min / kMinTile) * kRecordBytes,
				"fill tables exceed the records");
			uint32_t *h16 = reinterpret_cast<uint32_t *>(lookback_of(ws));
			uint32_t *st16 = h16 + 65536;
			uint16_t *inv16 = reinterpret_cast<uint16_t *>(st16 + 65536);
			uint32_t *map16 = h16 + kFill16Words;
			const uint32_t chunks = static_cast<uint32_t>((n + kFill16Chunk - 1) / kFill16Chunk);
			const uint64_t used_words = kFill16Words + chunks + 1u;
			ProfScope ps(0, 0, ws, stream);
			hipLaunchKernelGGL(k_zero, dim3(64), dim3(256), 0, stream, reinterpret_cast<uint64_t *>(h16), uint64_t{65536 / 2});
			if ((e = hipGetLastError()) != hipSuccess) return e;
			const uint32_t hblocks = static_cast<uint32_t>(std::max(1, cus / 2)) * 2u;
			hipLaunchKernelGGL(k_hist16<1024>, dim3(hblocks), dim3(1024), 0, stream, static_cast<const uint16_t *>(job.in), n, xf,
				h16);
			if ((e = hipGetLastError()) != hipSuccess) return e;
			hipLaunchKernelGGL(k_scan16, dim3(1), dim3(1024), 0, stream, h16, st16, inv16, xf);
			if ((e = hipGetLastError()) != hipSuccess) return e;
			hipLaunchKernelGGL(k_map16, dim3((chunks + 1 + 255) / 256), dim3(256), 0, stream, st16, map16, chunks);
			if ((e = hipGetLastError()) != hipSuccess) return e;
			const uint32_t fgrid = std::max<uint32_t>(1u, std::min<uint32_t>(chunks, 8u * cus));
			hipLaunchKernelGGL(k_fill16, dim3(fgrid), dim3(256), 0, stream, st16, map16, inv16, static_cast<uint16_t *>(job.out), n);
			if ((e = hipGetLastError()) != hipSuccess) return e;
			hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, stream, reinterpret_cast<uint64_t *>(h16), (used_words + 1) / 2);
			if ((e = hipGetLastError()) != hipSuccess) return e;
			hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, stream, ctrl);
			return hipGetLastError();
		}
	}
	const bool float_keys = (job.flags & kFltp) != 0;
	const int cfg = job.cfg >= 0 ? job.cfg : default_cfg(P, n, job.payload, float_keys);
	const bool keys_only = !job.payload && !job.keys_final_any;
	int algo = job.algo ? job.algo : default_algo(P, n, keys_only, float_keys);
	if (algo == kAlgoSample) {
		// the sample partition plan whatever the keys: skewed and few-valued integer keys, which fail
		// the exact plan's bucket test, get equality buckets of their repeated values
		if constexpr (P >= 4) {
			if (job.buf) return launch_sample<K>(cfg, job, ws, stream);
		}
		algo = default_algo(P, n, keys_only, float_keys);
	}
	if (!job.algo) {
		// The automatic plan and the sample partition plan (measured on MI355X, uniform and N(0,1)
		// keys, against the LSD passes): 8-byte keys take it from 2^22 when they are floats (f64
		// N(0,1) keys only at 2^21 / 2^22 / 2^24 / 2^26 / 2^28: 0.37 / 0.43 / 0.75 / 2.12 / 7.3 ms
		// against 0.35 / 0.45 / 1.03 / 3.85 / 14.0; with a u32 index: 0.40 / 0.45 / 0.84 / 2.30
		// against 0.37 / 0.50 / 1.29 / 5.17), from 2^24 when keys-only under a transform that is not
		// injective (i64 absolute values at 2^23 / 2^24: 0.75 / 0.80 ms against 0.62 / 1.02), and
		// from 2^25 for integers with a payload (u64 + index at 2^24 / 2^25 / 2^27 / 2^28: 1.35 /
		// 1.66 / 3.76 / 7.4 ms against 1.30 / 2.68 / 9.68 / 19.1: below 2^25 the wide bucket ranges
		// of uniform 64-bit keys keep it behind). 4-byte keys keep the LSD passes: f32 keys-only
		// with grouping partitions are ahead only from 2^27 and only by a little (N(0,1) at 2^26 /
		// 2^27 / 2^28: 1.42 / 2.57 / 4.80 ms against 1.35 / 2.59 / 5.02), while with both +0.0 and
		// -0.0 among them the ranked partitions fall behind (5.62 against 5.02 ms at 2^28), which
		// the plan cannot know before it has read the keys; with a payload the two are level (f32
		// N(0,1) + index at 2^26 / 2^28: 1.82 / 6.49 ms against 1.76 / 6.48), and the absolute-value
		// modes behind (i32 at 2^28: 5.29 against 4.95). Algorithm 4 still takes the sample plan.
		const bool inj = xform_injective(xf, P);
		if constexpr (P == 8) {
			if (job.buf && ((keys_only && (float_keys || !inj)) || job.payload) &&
				n >= (float_keys ? (1ull << 22) : keys_only ? (1ull << 24) : (1ull << 25)))
				return launch_sample<K>(cfg, job, ws, stream);
		}
		if (P == 4 && algo == kAlgoHybrid && (float_keys || !inj)) algo = kAlgoCountScan;
	}
	// small sorts with a payload: one LDS block as well (up to kLdsPayMax keys)
	if (!job.algo && job.payload && n <= kLdsPayMax) algo = kAlgoLds;
	if (algo == kAlgoLds) {
		// one block holds the whole array: every load happens before any store, so in place is fine
		const int rev = (job.flags & kRev) ? 1 : 0;
		const uint32_t nn = static_cast<uint32_t>(n);
		const K *kin = static_cast<const K *>(job.in);
		K *kout = static_cast<K *>(job.out);
		if (!job.payload && n <= kLdsMax && !job.keys_final_any) {
			ProfScope ps(1, 0, ws, stream);
			if (n <= 1024)
				hipLaunchKernelGGL((k_lds_sort<K, 256, 4, false>), dim3(1), dim3(256), 0, stream, kin, kout, nn, xf, rev, nullptr, nullptr);
			else if (n <= 4096)
				hipLaunchKernelGGL((k_lds_sort<K, 256, 16, false>), dim3(1), dim3(256), 0, stream, kin, kout, nn, xf, rev, nullptr, nullptr);
			else if (n <= 8192)
				hipLaunchKernelGGL((k_lds_sort<K, 512, 16, false>), dim3(1), dim3(512), 0, stream, kin, kout, nn, xf, rev, nullptr, nullptr);
			else
				hipLaunchKernelGGL((k_lds_sort<K, 1024, 16, false>), dim3(1), dim3(1024), 0, stream, kin, kout, nn, xf, rev, nullptr, nullptr);
			return hipGetLastError();
		}
		if (job.payload && n <= kLdsPayMax) {
			ProfScope ps(1, 0, ws, stream);
			const uint32_t *pin = job.pin;
			uint32_t *pout = job.pout;
			if (n <= 1024)
				hipLaunchKernelGGL((k_lds_sort<K, 256, 4, true>), dim3(1), dim3(256), 0, stream, kin, kout, nn, xf, rev, pin, pout);
			else if (n <= 4096)
				hipLaunchKernelGGL((k_lds_sort<K, 256, 16, true>), dim3(1), dim3(256), 0, stream, kin, kout, nn, xf, rev, pin, pout);
			else
				hipLaunchKernelGGL((k_lds_sort<K, 512, 16, true>), dim3(1), dim3(512), 0, stream, kin, kout, nn, xf, rev, pin, pout);
			return hipGetLastError();
		}
		algo = default_algo(P, n, false, float_keys);
	}
	const bool inj_xf = xform_injective(xf, P);
	const uint32_t joint_slices = static_cast<uint32_t>(std::min<uint64_t>(static_cast<uint64_t>(cus),
		n * P / (kJointBins * sizeof(uint32_t)))) / kRanges * kRanges;  // the same number of blocks per key range
	int lcfg = -1;
	if (algo == kAlgoHybrid) {
		lcfg = local_cfg_for(n);
		if (P < 4 || !keys_only || !job.buf || lcfg < 0) lcfg = -1;
		// The MSD plan by the top two bytes needs exact bucket sizes (integer keys under an
		// injective transform, keys only); every other sort of 4- and 8-byte keys with a buffer
		// takes the sample partition plan.
		const bool exact = lcfg >= 0 && !float_keys && inj_xf && joint_slices >= kRanges;
		if constexpr (P >= 4) {
			if (!exact && job.buf) return launch_sample<K>(cfg, job, ws, stream);
		}
		if (lcfg < 0) algo = kAlgoCountScan;
	}
	const bool hybrid = lcfg >= 0;
	const uint32_t tile_keys = static_cast<uint32_t>(kTileCfgs[cfg].threads * kTileCfgs[cfg].kpt);
	const uint32_t tiles = static_cast<uint32_t>((n + tile_keys - 1) / tile_keys);
	// the MSD plan's partitions take 1024x16 tiles where the table picks 1024x12 (longer digit runs;
	// keys only, so the payload staging that rules this tile out elsewhere is absent): measured
	// 0.95 ms per u64 pass instead of 0.99, and 2.62 ms for the whole 2^28 u32 sort instead of
	// 2.70-2.78. Only the grouping form: the ranked form of that tile spills registers.
	const int hcfg = (hybrid && P >= 4 && cfg == 3 && xform_injective(xf, P)) ? 5 : cfg;
	const uint32_t htile_keys = hcfg == 5 ? 1024u * 16u : tile_keys;
	const uint32_t htiles = static_cast<uint32_t>((n + htile_keys - 1) / htile_keys);
	uint32_t *counts = counts_of(ws);
	const bool fill = (P == 1) && !job.payload && job.fill8 && bijective;
	// the histogram read also counts, per tile, the digit of the first pass: digit 0 for the LSD
	// count + scan plan, the top digit for the MSD plan
	const bool counts0 = algo == kAlgoCountScan && !fill;
	const bool count_tiles = counts0 || hybrid;
	// The MSD plan's read (k_hist_joint) leaves one 65536-word slice per block in the array the first
	// partition writes (X1, never the source), so the number of blocks is bounded by that array's
	// size (joint_slices).
	uint32_t *joint_part = static_cast<uint32_t *>(job.buf != job.in ? job.buf : job.out);
	if (job.presort_check && !fill && (e = launch_check<K>(job, ws, xf, stream)) != hipSuccess) return e;
	{
		ProfScope ps(0, 0, ws, stream);
		// four blocks per CU: measured 0.44 ms for the 2^28 u64 first read against 0.48 with eight
		const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(hybrid ? htiles : tiles, 4u * cus));
		if (hybrid) {
			// The MSD plan needs only the histograms of the top two digits and the exact bucket
			// sizes: the histogram of the top 16 bits (integer keys under an injective transform).
			if constexpr (P >= 4) {
				// the key ranges: whole first-partition tiles, as its blocks deal them to the XCDs
				const uint64_t range_keys = static_cast<uint64_t>((htiles + kRanges - 1) / kRanges) * htile_keys;
				hipLaunchKernelGGL((k_hist_joint<K, 1024>), dim3(joint_slices), dim3(1024), 0, stream,
					static_cast<const K *>(job.in), n, xf, range_keys, (job.flags & kRev) ? 1 : 0, ctrl, joint_part,
					seg_tables(segtab_of(ws)).ovf);
				if ((e = hipGetLastError()) != hipSuccess) return e;
				hipLaunchKernelGGL((k_joint_reduce<P>), dim3(kRadix), dim3(256), 0, stream, ctrl, (job.flags & kRev) ? 1 : 0,
					joint_part, joint_slices, seg_tables(segtab_of(ws)), hist);
			}
		} else if (count_tiles) {
			hipLaunchKernelGGL((k_hist_tiles<K, 256, HITEMS>), dim3(grid), dim3(256), 0, stream, static_cast<const K *>(job.in), n,
				xf, hist, counts, tile_keys, tiles, (job.flags & kRev) ? 1 : 0, ctrl, 0u, nullptr);
		} else {
			const uint32_t hgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((n + 256 * HITEMS - 1) / (256 * HITEMS), 8u * cus)));
			hipLaunchKernelGGL((k_hist<K, 256, HITEMS>), dim3(hgrid), dim3(256), 0, stream, static_cast<const K *>(job.in), n, xf, hist,
				ctrl, (job.flags & kRev) ? 1 : 0);
		}
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}

	PlanArgs pa;
	pa.n = n; pa.passes = P; pa.src = job.in; pa.fin = job.out; pa.other = job.buf;
	pa.psrc = job.payload ? job.pin : nullptr;
	pa.pfin = job.payload ? job.pout : nullptr;
	pa.pother = job.payload ? job.pbuf : nullptr;
	pa.payload = job.payload;
	pa.reverse = (job.flags & kRev) != 0;
	pa.keys_final_any = job.keys_final_any;
	pa.counts0 = counts0 ? 1 : 0;
	pa.hybrid = hybrid ? 1 : 0;
	pa.hyb_cap = hybrid ? kLocalCaps[lcfg] : 0;
	pa.tile2 = htile_keys;
	pa.segtab = segtab_of(ws);
	pa.gate = nullptr;
	pa.group = 0;
	pa.split0 = nullptr;

	hipLaunchKernelGGL(k_plan, dim3(1), dim3(256), 0, stream, pa, ctrl, hist, bucket);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	hipStream_t ls = stream;  // the stream of the LSD launches below
	if constexpr (P >= 4) {
		if (hybrid) {
			// With the exact bucket sizes the plan's choice is final once k_plan has run: the fallback
			// launches then go to a side stream that waits for k_plan and runs alongside the MSD
			// launches (each of the two sets exits at once unless it was chosen), so that the
			// fallback's launches, about 30 dispatches, cost no time after the bucket sort. The
			// stream waits for the side stream before the sort ends.
			if ((e = fork_side(stream, &ls)) != hipSuccess) return e;
			e = launch_hybrid_cfg<K>(hcfg, ctrl, seg_tables(segtab_of(ws)), xf, n, htiles, lcfg, ws, stream);
			if (e != hipSuccess) return e;
			// The LSD plan as the fallback: a full histogram read and a second plan, both gated on
			// need_full, then count + scan passes (inactive unless planned).
			algo = kAlgoCountScan;
			const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(tiles, 8u * cus));
			hipLaunchKernelGGL((k_hist_tiles<K, 256, HITEMS>), dim3(grid), dim3(256), 0, ls, static_cast<const K *>(job.in), n,
				xf, hist, counts, tile_keys, tiles, (job.flags & kRev) ? 1 : 0, ctrl, 0u, &ctrl->need_full);
			if ((e = hipGetLastError()) != hipSuccess) return e;
			pa.counts0 = 1;
			pa.hybrid = 0;
			pa.gate = &ctrl->need_full;
			hipLaunchKernelGGL(k_plan, dim3(1), dim3(256), 0, ls, pa, ctrl, hist, bucket);
			if ((e = hipGetLastError()) != hipSuccess) return e;
		}
	}

	const uint32_t cgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8u * cus)));
	if (fill) {
		// bijective 8-bit modes: the sorted output follows from the counts alone
		hipLaunchKernelGGL(k_fill8, dim3(cgrid), dim3(256), 0, ls, bucket, xf, static_cast<uint8_t *>(job.out), n);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	} else {
		hipLaunchKernelGGL((k_copy<K>), dim3(cgrid), dim3(256), 0, ls, ctrl, 0, n, job.payload ? 1 : 0);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		for (int p = 0; p < P; ++p) {
			if (algo == kAlgoOnesweep) {
				ProfScope ps(1, p, ws, ls);
				e = job.payload ? launch_pass_cfg<K, true>(cfg, ctrl, bucket, lookback, n, xf, p, ls)
				                : launch_pass_cfg<K, false>(cfg, ctrl, bucket, lookback, n, xf, p, ls);
				if (e != hipSuccess) return e;
				continue;
			}
			{
				ProfScope ps(3, p, ws, ls);
				const uint32_t cgrid = std::max<uint32_t>(1u, std::min<uint32_t>(tiles, 8u * cus));
				hipLaunchKernelGGL((k_count<K, 256, HITEMS>), dim3(cgrid), dim3(256), 0, ls, ctrl, p, xf, n, tile_keys, tiles, counts);
				if ((e = hipGetLastError()) != hipSuccess) return e;
				const uint32_t nchunks = (tiles + kScanChunk - 1) / kScanChunk;
				const uint32_t sgrid = std::max<uint32_t>(1u, std::min<uint32_t>(nchunks, 4u * cus));
				hipLaunchKernelGGL(k_scan, dim3(sgrid), dim3(256), 0, ls, ctrl, &ctrl->pass[p], &ctrl->scan_counter[p],
					bucket + p * kRadix, p, counts, lookback, tiles);
				if ((e = hipGetLastError()) != hipSuccess) return e;
			}
			ProfScope ps(1, p, ws, ls);
			e = job.payload ? launch_scatter_cfg<K, true>(cfg, ctrl, p, xf, n, tiles, counts, ls)
			                : launch_scatter_cfg<K, false>(cfg, ctrl, p, xf, n, tiles, counts, ls);
			if (e != hipSuccess) return e;
		}
		hipLaunchKernelGGL((k_copy<K>), dim3(cgrid), dim3(256), 0, ls, ctrl, 1, n, job.payload ? 1 : 0);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	if (ls != stream && (e = join_side(ls, stream)) != hipSuccess) return e;
	hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, stream, ctrl);
	return hipGetLastError();
}

} // namespace

int device_cu_count() { return devinfo().cus; }

uint32_t set_sample_target(uint32_t keys) { return g_sample_target.exchange(keys); }

void profile_begin()
{
	std::lock_guard<std::mutex> lock(g_prof_mu);
	for (auto &r : g_prof) { (void)hipEventSynchronize(r.b); g_prof_pool.push_back(r.a); g_prof_pool.push_back(r.b); }
	g_prof.clear();
	g_prof_on = true;
}

// Sums the recorded launch times per kind: 0 histogram, 1 scatter launches of executed passes,
// 2 scatter launches of skipped passes, 3 count+scan launches of executed passes, 4 count+scan
// launches of skipped passes, and for the two-level MSD plan (counted only when it ran) 5 its two
// partition scatters, 6 the per-bucket LDS sort, 7 its count/scan launches. A pass counts as active by the plan its workspace holds now (the
// plan of the last sort run on that workspace).
hipError_t profile_end(double *ms, unsigned *launches, int nkinds)
{
	std::vector<ProfRec> recs;
	{
		std::lock_guard<std::mutex> lock(g_prof_mu);
		g_prof_on = false;
		recs.swap(g_prof);
	}
	for (int k = 0; k < nkinds; ++k) { ms[k] = 0; launches[k] = 0; }
	hipError_t err = hipSuccess;
	for (auto &r : recs) {
		hipError_t e = hipEventSynchronize(r.b);
		float t = 0;
		if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
		int kind = r.kind;
		if (e == hipSuccess && (kind == 1 || kind == 3)) {
			uint32_t mask = 0;
			e = hipMemcpy(&mask, &static_cast<const Control *>(r.ws)->active_mask, sizeof(mask), hipMemcpyDeviceToHost);
			if (!(mask >> r.pass & 1u)) kind += 1;
		}
		if (e == hipSuccess && kind >= 5) {
			uint32_t on = 0;
			e = hipMemcpy(&on, &static_cast<const Control *>(r.ws)->hyb[2].active, sizeof(on), hipMemcpyDeviceToHost);
			if (!on) kind = nkinds;  // the MSD plan was not used: not counted
		}
		if (e != hipSuccess) err = e;
		else if (kind < nkinds) { ms[kind] += t; ++launches[kind]; }
	}
	std::lock_guard<std::mutex> lock(g_prof_mu);
	for (auto &r : recs) { g_prof_pool.push_back(r.a); g_prof_pool.push_back(r.b); }
	return err;
}

// measured on MI355X (uniform keys):
// - keys-only sorts take the two-level MSD plan from 2^18 keys for 8-byte integer keys (2^20 u64:
//   223 us against 296 for onesweep; 2^24: 427 against 1065 for count + scan) and from 2^22 for
//   4-byte integer keys (2^24 u32: 301 us against 411); 8-byte float keys from 2^22 and 4-byte
//   ones from 2^24 (algorithm 4 sorts them with the sample partition plan: 8-byte floats are ahead
//   of the LSD passes from 2^22; run_sort keeps 4-byte floats and non-injective 4-byte modes on
//   the LSD passes);
// - otherwise the count + scan path from 2^22 keys, the onesweep path below (fewer launches);
// - keys-only sorts of up to kLdsMax keys: one block sorts them in LDS, one launch instead of a dozen
//   (2^10 / 2^14 u32: 9 / 39 us against 49 / 66 for onesweep; u64: 16 / 85 against 87 / 124; u8 2^14:
//   12 against 19 for the count fill).
int default_algo(int key_bytes, uint64_t n, bool keys_only, bool float_keys)
{
	if (keys_only && n <= kLdsMax) return kAlgoLds;
	if (keys_only && key_bytes >= 4 && local_cfg_for(n) >= 0) {
		// 8-byte integer keys from 2^20: below it onesweep is faster (u64 2^18 / 2^19.5: 137 / 207 us
		// against 184 / 229 for the MSD plan; 2^20: 232 against 222)
		const uint64_t from = float_keys ? (key_bytes >= 8 ? (1ull << 22) : (1ull << 24))
		                                 : (key_bytes >= 8 ? (1ull << 20) : (1ull << 22));
		if (n >= from) return kAlgoHybrid;
	}
	return n >= (1u << 22) ? kAlgoCountScan : kAlgoOnesweep;
}

// Tile configuration of the automatic plan, measured on MI355X with uniform keys (times per sort).
int default_cfg(int key_bytes, uint64_t n, bool payload, bool float_keys)
{
	if (n <= (1u << 16)) return 0;  // 256x8: the autotuned table's choice up to 2^16 keys
	// MSD partitions: 512x8 tiles below 2^25 keys (2^24 u64: 427 us against 438 with 1024x12),
	// 1024x12 from it (2^25 u64: 661 against 695)
	if (!payload && default_algo(key_bytes, n, true, float_keys) == kAlgoHybrid) return n >= (1u << 25) ? 3 : 2;
	// measured with a u32 payload: at 2^28 keys of 1, 2 or 4 bytes, 1024x12 tiles (longer digit
	// runs per tile) take 0.74-0.92x the time of 256x16 tiles; at 2^24 every shape is within 6%
	if (payload && n >= (1u << 26)) return 3;
	// count + scan passes of 8-byte keys: 512x8 tiles below 2^25 keys (u64 keys only, 2^22 / 2^24:
	// 431 / 953 us against 504 / 1065 with 1024x12; with a u32 payload at 2^24: 1.30 ms against 1.47).
	// 4-byte keys only (float keys below 2^26, say): 256x8 (f32 2^24: 422 us against 437 with
	// 256x16); with a payload 256x16 (u32 2^24: 0.51 ms against 0.55 with 256x8).
	if (default_algo(key_bytes, n, false, float_keys) == kAlgoCountScan) {
		if (key_bytes >= 8) return n >= (1u << 25) ? 3 : 2;
		if (key_bytes == 4) return payload ? 1 : 0;
		return key_bytes == 2 ? 2 : 1;
	}
	// onesweep: 8-byte keys 512x8 (f64 at 2^17 / 2^21: 147 / 327 us against 230 / 377 with 256x24;
	// u64 with a u32 payload: 155 / 386 against 245 / 516); smaller keys 512x8, 1024x12 from 2^21
	// (u32 / f32 / u16 at 2^21: 149 / 153 / 78 us against 164 / 166 / 87 with 512x8; at 2^19 512x8
	// is ahead)
	if (key_bytes >= 8) return 2;
	return n >= (1u << 21) ? 3 : 2;
}

size_t sort_workspace_bytes(uint64_t n)
{
	// the header, then one record (look-back row + count row) per tile: max_tiles(n), plus one
	// partial tile per top-digit segment when the tiles of the segmented MSD pass are aligned to
	// the segments
	return workspace_bytes_for(n);
}

hipError_t workspace_reset(void *ws, hipStream_t stream)
{
	hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, stream, ctrl_of(ws), hist_of(ws), seg_tables(segtab_of(ws)).ovf);
	return hipGetLastError();
}

hipError_t workspace_init(void *ws, size_t bytes, hipStream_t stream)
{
	const uint64_t words = bytes / 8;
	const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((words + 255) / 256, 4096)));
	hipLaunchKernelGGL(k_zero, dim3(grid), dim3(256), 0, stream, static_cast<uint64_t *>(ws), words);
	return hipGetLastError();
}

hipError_t sort_launch(const SortJob &job, void *ws, size_t ws_bytes, hipStream_t stream)
{
	if (job.n == 0) return hipSuccess;
	if (job.n > 0xFFFFFFFFull) return hipErrorInvalidValue;
	if (ws_bytes < sort_workspace_bytes(job.n)) return hipErrorInvalidValue;
	switch (job.key_bytes) {
	case 1: return run_sort<uint8_t>(job, ws, stream);
	case 2: return run_sort<uint16_t>(job, ws, stream);
	case 4: return run_sort<uint32_t>(job, ws, stream);
	case 8: return run_sort<uint64_t>(job, ws, stream);
	default: return hipErrorInvalidValue;
	}
}

} // namespace rsbd8hip
