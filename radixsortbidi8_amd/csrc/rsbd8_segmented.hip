// rsbd8_segmented.hip -- many independent sorts in one captured launch sequence.
//
// The reference sorts a batch by calling radixsortcopy once per array (Radixsortbidi8.hpp:60468);
// here every segment of a CSR batch is sorted with the same mode/direction flags by:
//   k_seg_small   one persistent kernel: each block takes the next short segment from a queue
//                 and sorts it entirely in LDS (8-bit LSD passes with the block ranking of the
//                 onesweep kernel, so the order is the same stable order);
//   k_seg_hist    per-segment digit histograms of the long segments, one block per tile;
//   k_seg_plan    per long segment: bucket starts, uniform-digit skip mask, ping-pong selection;
//   k_seg_pass    one persistent onesweep launch per digit over the tiles of all long segments,
//                 with the decoupled look-back confined to each segment;
//   k_seg_copy    long segments whose digits are all uniform (copy, with tie reversal);
//   k_seg_finish  resets the queue/tile counters and advances the look-back epoch.
// A plan object owns the device tables (segment classes, tile map) and its workspace, and can
// capture the whole sequence (short-segment branch in parallel with the long-segment chain) into
// one hipGraph that is replayed with a single launch.
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
	uint32_t small_counter;
	uint32_t tile_counter[kMaxPasses];
	uint32_t epoch;
	uint32_t pad[6];
};

// per long segment: for pass p, bits 4p..4p+1 = source (0 input, 1 output, 2 buffer),
// bits 4p+2..4p+3 = destination; bit 31 = no pass executes (copy); bit 30 = first pass reads reversed
struct SegSel {
	uint32_t sel;
	uint32_t active_mask;
};

template <typename K>
__device__ __forceinline__ K *pick(uint32_t code, const K *in, K *out, K *buf)
{
	return code == 0 ? const_cast<K *>(in) : (code == 1 ? out : buf);
}

// ---------------------------------------------------------------------------------------------

template <typename K, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_seg_small(const K *__restrict__ in, K *__restrict__ out,
	const uint64_t *__restrict__ seg, const uint32_t *__restrict__ ids, uint32_t count, SegCtrl *ctrl, KeyXform xf,
	int reverse)
{
	constexpr int CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_wsum[4];
	__shared__ uint32_t s_item;
	__shared__ K s_keys[CAP];
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	for (;;) {
		if (t == 0) s_item = atomicAdd(&ctrl->small_counter, 1u);
		__syncthreads();
		const uint32_t item = s_item;
		if (item >= count) break;
		const uint32_t s = ids[item];
		const uint64_t b = seg[s];
		const uint32_t len = static_cast<uint32_t>(seg[s + 1] - b);
		K keys[KPT];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			keys[j] = l < len ? in[b + (reverse ? len - 1 - l : l)] : K(0);
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
			if (l < len) out[b + l] = keys[j];
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------------

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_seg_hist(const K *__restrict__ in, const uint64_t *__restrict__ seg,
	const uint32_t *__restrict__ large_ids, const uint2 *__restrict__ tiles, uint32_t tile_keys, KeyXform xf,
	uint32_t *__restrict__ hist)
{
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_h[P * 256];
	for (int i = threadIdx.x; i < P * 256; i += THREADS) s_h[i] = 0;
	__syncthreads();
	const uint2 tl = tiles[blockIdx.x];
	const uint32_t s = large_ids[tl.x];
	const uint64_t b = seg[s] + static_cast<uint64_t>(tl.y) * tile_keys;
	const uint64_t e = min(seg[s + 1], b + tile_keys);
	for (uint64_t i = b + threadIdx.x; i < e; i += THREADS) {
		const auto k = sort_key<K>(in[i], xf);
#pragma unroll
		for (int p = 0; p < P; ++p) atomicAdd(&s_h[p * 256 + (static_cast<uint32_t>(k >> (8 * p)) & 0xFFu)], 1u);
	}
	__syncthreads();
	uint32_t *h = hist + static_cast<uint64_t>(tl.x) * P * 256;
	for (int i = threadIdx.x; i < P * 256; i += THREADS)
		if (s_h[i]) atomicAdd(&h[i], s_h[i]);
}

template <int P>
__global__ __launch_bounds__(256) void k_seg_plan(const uint64_t *__restrict__ seg, const uint32_t *__restrict__ large_ids,
	uint32_t *__restrict__ hist, uint32_t *__restrict__ bucket, SegSel *__restrict__ sel, int reverse)
{
	__shared__ uint32_t s_wave[4];
	__shared__ uint32_t s_uniform[P];
	const uint32_t L = blockIdx.x, t = threadIdx.x, lane = t & 63u, w = t >> 6;
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
		code = 1u << 31;
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
		if (reverse) code |= 1u << 30;
	}
	sel[L].sel = code;
	sel[L].active_mask = mask;
}

template <typename K, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_seg_pass(const K *__restrict__ in, K *__restrict__ out, K *__restrict__ buf,
	const uint64_t *__restrict__ seg, const uint32_t *__restrict__ large_ids, const uint2 *__restrict__ tiles,
	uint32_t num_tiles, const uint32_t *__restrict__ bucket, const SegSel *__restrict__ sel, uint64_t *__restrict__ lookback,
	SegCtrl *ctrl, KeyXform xf, int pass)
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
	for (;;) {
		if (t == 0) s_tile = atomicAdd(&ctrl->tile_counter[pass], 1u);
		__syncthreads();
		const uint32_t g = s_tile;
		if (g >= num_tiles) break;
		const uint2 tl = tiles[g];
		const SegSel sl = sel[tl.x];
		if (!(sl.active_mask >> pass & 1u)) {
			__syncthreads();
			continue;  // this segment skips the digit: none of its tiles takes part
		}
		const uint32_t code = sl.sel >> (4 * pass);
		const K *src = pick<K>(code & 3u, in, out, buf);
		K *dst = pick<K>(code >> 2 & 3u, in, out, buf);
		const bool rev = (sl.sel >> 30 & 1u) && ((sl.active_mask & ((1u << pass) - 1u)) == 0);  // first executed pass
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
		if (!(sl.sel >> 31)) continue;
		const uint32_t s = large_ids[L];
		const uint64_t b = seg[s], len = seg[s + 1] - b;
		const bool rev = sl.sel >> 30 & 1u;
		for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) out[b + i] = in[b + (rev ? len - 1 - i : i)];
	}
}

__global__ void k_seg_finish(SegCtrl *ctrl)
{
	if (threadIdx.x < kMaxPasses) ctrl->tile_counter[threadIdx.x] = 0;
	if (threadIdx.x == 0) {
		ctrl->small_counter = 0;
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
	int tile_cfg = 1;    // onesweep tile of the long segments (kSegTileCfgs)
	int small_cfg = 1;   // LDS capacity tier (kSmallCaps)
	uint32_t nsmall = 0, nlarge = 0, ntiles = 0, tile_keys = 0, small_cap = 0;
	char *mem = nullptr;  // one device allocation for every table and the workspace
	size_t mem_bytes = 0;
	SegCtrl *ctrl = nullptr;
	uint32_t *small_ids = nullptr, *large_ids = nullptr, *hist = nullptr, *bucket = nullptr;
	uint2 *tiles = nullptr;
	SegSel *sel = nullptr;
	uint64_t *lookback = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipStream_t side = nullptr;
	hipEvent_t fork = nullptr, join = nullptr;
};

// tiles of the long segments' onesweep passes (threads x keys per thread), in the order of the
// tile_cfg argument of the segmented plan
constexpr int kNumSegTileCfgs = 5;
constexpr TileCfg kSegTileCfgs[kNumSegTileCfgs] = {{256, 8}, {256, 16}, {512, 8}, {512, 16}, {256, 24}};
constexpr int kNumSmallCaps = 3;
constexpr int kSmallCaps[kNumSmallCaps] = {256 * 4, 256 * 16, 512 * 16};

namespace {

template <typename K>
hipError_t launch_small(SegPlan &pl, const void *in, void *out, const uint64_t *seg, const KeyXform &xf, hipStream_t s)
{
	if (!pl.nsmall) return hipSuccess;
	const int cus = device_cu_count();
	const int rev = (pl.flags & kRev) != 0;
	auto go = [&](auto kernel, int threads, int occ) {
		const uint32_t grid = std::min<uint32_t>(pl.nsmall, static_cast<uint32_t>(occ * cus));
		hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, s, static_cast<const K *>(in), static_cast<K *>(out), seg,
			pl.small_ids, pl.nsmall, pl.ctrl, xf, rev);
	};
	switch (pl.small_cfg) {
	case 0: { static int o = occupancy_of<256, 4>(k_seg_small<K, 256, 4>); go(k_seg_small<K, 256, 4>, 256, o); break; }
	case 1: { static int o = occupancy_of<256, 16>(k_seg_small<K, 256, 16>); go(k_seg_small<K, 256, 16>, 256, o); break; }
	default: { static int o = occupancy_of<512, 16>(k_seg_small<K, 512, 16>); go(k_seg_small<K, 512, 16>, 512, o); break; }
	}
	return hipGetLastError();
}

template <typename K, int THREADS, int KPT>
hipError_t launch_seg_pass(SegPlan &pl, const void *in, void *out, void *buf, const uint64_t *seg, const KeyXform &xf,
	int pass, hipStream_t s)
{
	static int occ = occupancy_of<THREADS, KPT>(k_seg_pass<K, THREADS, KPT>);
	const uint32_t grid = std::min<uint32_t>(pl.ntiles, static_cast<uint32_t>(occ * device_cu_count()));
	hipLaunchKernelGGL((k_seg_pass<K, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, s, static_cast<const K *>(in),
		static_cast<K *>(out), static_cast<K *>(buf), seg, pl.large_ids, pl.tiles, pl.ntiles, pl.bucket, pl.sel, pl.lookback,
		pl.ctrl, xf, pass);
	return hipGetLastError();
}

template <typename K>
hipError_t launch_large(SegPlan &pl, const void *in, void *out, void *buf, const uint64_t *seg, const KeyXform &xf, hipStream_t s)
{
	if (!pl.nlarge) return hipSuccess;
	constexpr int P = sizeof(K);
	hipError_t e;
	hipLaunchKernelGGL((k_seg_hist<K, 256>), dim3(pl.ntiles), dim3(256), 0, s, static_cast<const K *>(in), seg, pl.large_ids,
		pl.tiles, pl.tile_keys, xf, pl.hist);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	hipLaunchKernelGGL((k_seg_plan<P>), dim3(pl.nlarge), dim3(256), 0, s, seg, pl.large_ids, pl.hist, pl.bucket, pl.sel,
		(pl.flags & kRev) ? 1 : 0);
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
	if ((e = launch_small<K>(pl, in, out, seg, xf, pl.side)) != hipSuccess) return e;
	if ((e = hipEventRecord(pl.join, pl.side)) != hipSuccess) return e;
	if ((e = launch_large<K>(pl, in, out, buf, seg, xf, s)) != hipSuccess) return e;
	if ((e = hipStreamWaitEvent(s, pl.join, 0)) != hipSuccess) return e;
	hipLaunchKernelGGL(k_seg_finish, dim3(1), dim3(64), 0, s, pl.ctrl);
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
	std::vector<uint32_t> large;
	std::vector<uint2> tiles;
	for (uint64_t s = 0; s < nseg; ++s) {
		if (host_off[s + 1] < host_off[s] || host_off[s + 1] - host_off[s] > 0xFFFFFFFFull) { delete pl; return hipErrorInvalidValue; }
		const uint64_t len = host_off[s + 1] - host_off[s];
		if (len < 2) {
			if (len == 1) small.push_back({len, static_cast<uint32_t>(s)});  // a single key is copied by the small tier
			continue;
		}
		if (len <= pl->small_cap) {
			small.push_back({len, static_cast<uint32_t>(s)});
		} else {
			const uint32_t L = static_cast<uint32_t>(large.size());
			large.push_back(static_cast<uint32_t>(s));
			const uint32_t nt = static_cast<uint32_t>((len + pl->tile_keys - 1) / pl->tile_keys);
			for (uint32_t i = 0; i < nt; ++i) tiles.push_back(make_uint2(L, i));
		}
	}
	// longest short segments first, so the tail of the queue is made of quick ones
	std::stable_sort(small.begin(), small.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
	std::vector<uint32_t> small_ids(small.size());
	for (size_t i = 0; i < small.size(); ++i) small_ids[i] = small[i].second;
	pl->nsmall = static_cast<uint32_t>(small_ids.size());
	pl->nlarge = static_cast<uint32_t>(large.size());
	pl->ntiles = static_cast<uint32_t>(tiles.size());
	const int P = key_bytes;
	size_t off = 0;
	const size_t o_ctrl = off; off += up256(sizeof(SegCtrl));
	const size_t o_small = off; off += up256(small_ids.size() * 4 + 4);
	const size_t o_large = off; off += up256(large.size() * 4 + 4);
	const size_t o_tiles = off; off += up256(tiles.size() * sizeof(uint2) + 8);
	const size_t o_sel = off; off += up256(large.size() * sizeof(SegSel) + 8);
	const size_t o_hist = off; off += up256(large.size() * P * 256 * 4 + 4);
	const size_t o_bucket = off; off += up256(large.size() * P * 256 * 4 + 4);
	const size_t o_lb = off; off += up256(static_cast<size_t>(tiles.size()) * 256 * 8 + 8);
	pl->mem_bytes = off;
	hipError_t e = hipMalloc(reinterpret_cast<void **>(&pl->mem), off);
	if (e != hipSuccess) { delete pl; return e; }
	pl->ctrl = reinterpret_cast<SegCtrl *>(pl->mem + o_ctrl);
	pl->small_ids = reinterpret_cast<uint32_t *>(pl->mem + o_small);
	pl->large_ids = reinterpret_cast<uint32_t *>(pl->mem + o_large);
	pl->tiles = reinterpret_cast<uint2 *>(pl->mem + o_tiles);
	pl->sel = reinterpret_cast<SegSel *>(pl->mem + o_sel);
	pl->hist = reinterpret_cast<uint32_t *>(pl->mem + o_hist);
	pl->bucket = reinterpret_cast<uint32_t *>(pl->mem + o_bucket);
	pl->lookback = reinterpret_cast<uint64_t *>(pl->mem + o_lb);
	hipLaunchKernelGGL(k_seg_zero, dim3(1024), dim3(256), 0, nullptr, reinterpret_cast<uint64_t *>(pl->mem), off / 8);
	if ((e = hipGetLastError()) == hipSuccess && !small_ids.empty())
		e = hipMemcpy(pl->small_ids, small_ids.data(), small_ids.size() * 4, hipMemcpyHostToDevice);
	if (e == hipSuccess && !large.empty()) e = hipMemcpy(pl->large_ids, large.data(), large.size() * 4, hipMemcpyHostToDevice);
	if (e == hipSuccess && !tiles.empty()) e = hipMemcpy(pl->tiles, tiles.data(), tiles.size() * sizeof(uint2), hipMemcpyHostToDevice);
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
