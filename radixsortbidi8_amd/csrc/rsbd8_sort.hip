// rsbd8_sort.hip -- gfx950 kernels of the stable 8-bit LSD radix sort.
//
// Replaces the reference's counting pass (Radixsortbidi8.hpp = H, H:32107-32473), its offsets
// transform with pass skipping (H:11626-12946) and its bidirectional scatter passes
// (H:31649-31982) with:
//   k_hist      one HBM read of the keys building every digit's 256-bin histogram at once in
//               LDS (wave-uniform digits are added with one atomic per wave);
//   k_plan      one block: per-digit exclusive scans (bucket starts), uniform-digit detection
//               (the pass-skip flag of H:11652) and the ping-pong buffer plan, all on device;
//   k_onesweep  one launch per digit: persistent blocks take tiles in order, rank digits with
//               wave64 ballot matching, publish per-tile digit counts and resolve the global
//               offsets with a decoupled look-back, stage the tile in LDS in digit order and
//               write coalesced runs; inactive (uniform-digit) passes exit at once;
//   k_copy      copies when no pass runs or an in-place sort needs a staging copy;
//   k_fill8     the 8-bit count fill of H:48769-48957 for bijective modes;
//   k_finish    resets the tile counters and advances the look-back epoch.
// The transform of every mode is fused into each load (rsbd8_device.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "rsbd8_engine.hpp"

namespace rsbd8hip {

namespace {

inline Control *ctrl_of(void *ws) { return reinterpret_cast<Control *>(ws); }
inline uint32_t *hist_of(void *ws) { return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kCtrlBytes); }
inline uint32_t *bucket_of(void *ws) { return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kCtrlBytes + kHistBytes); }
inline uint64_t *lookback_of(void *ws) { return reinterpret_cast<uint64_t *>(static_cast<char *>(ws) + kLookbackOffset); }
inline uint64_t max_tiles(uint64_t n) { return (n + kMinTile - 1) / kMinTile + 1; }
inline uint32_t *counts_of(void *ws, uint64_t n)
{
	return reinterpret_cast<uint32_t *>(static_cast<char *>(ws) + kLookbackOffset + max_tiles(n) * kRadix * sizeof(uint64_t));
}

__device__ __forceinline__ uint32_t pass_tag(uint32_t epoch, int pass)
{
	return ((epoch * kMaxPasses + static_cast<uint32_t>(pass)) % 0x3FFFFFFFu) + 1u;
}

// 0: a real sort is needed; 1: the output is the input as it is; 2: the output is the input
// reversed (physical order). For isrevorder the logical input is the physical one read backwards.
__device__ __forceinline__ uint32_t presort_result(uint32_t check, int reverse)
{
	if (!(check & 1u)) return 0;
	const bool lt = check & 2u, eq = check & 4u, gt = check & 8u;
	if (!reverse) return !gt ? 1u : ((!lt && !eq) ? 2u : 0u);
	return !lt ? 2u : ((!gt && !eq) ? 1u : 0u);
}

// ---------------------------------------------------------------------------------------------
// histogram of every digit in one read

// Calls f(key) for every key of the physical range [lo, lo + len) of p, with the keys of a block
// spread over its threads. 16-byte vector loads (four in flight per thread) when the range is
// 16-byte aligned; the order in which a thread sees its keys is unspecified (counting only).
template <typename K, int THREADS, typename F>
__device__ __forceinline__ void for_each_key(const K *__restrict__ p, uint64_t lo, uint32_t len, F &&f)
{
	constexpr uint32_t VEC = 16 / sizeof(K);
	const K *q = p + lo;
	uint32_t done = 0;
	if ((reinterpret_cast<uintptr_t>(q) & 15u) == 0) {
		const uint32_t nvec = len / VEC;
		const uint4 *qv = reinterpret_cast<const uint4 *>(q);
		for (uint32_t v = threadIdx.x; v < nvec; v += THREADS * 4) {
			uint4 u[4];
#pragma unroll
			for (int jj = 0; jj < 4; ++jj) {
				const uint32_t idx = v + jj * THREADS;
				if (idx < nvec) u[jj] = qv[idx];
			}
#pragma unroll
			for (int jj = 0; jj < 4; ++jj) {
				if (v + jj * THREADS >= nvec) break;
				K ks[VEC];
				__builtin_memcpy(ks, &u[jj], 16);
#pragma unroll
				for (uint32_t e = 0; e < VEC; ++e) f(ks[e]);
			}
		}
		done = nvec * VEC;
	}
	for (uint32_t i = done + threadIdx.x; i < len; i += THREADS) f(q[i]);
}

// adds 1 to s[d] for every lane, with one atomic for the whole wave when all lanes share d
__device__ __forceinline__ void wave_count(uint32_t *s, uint32_t d)
{
	const uint64_t active = __ballot(1);
	const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
	if (__ballot(d == d0) == active) {
		if (lanes_below(active) == 0) atomicAdd(&s[d0], static_cast<uint32_t>(__popcll(active)));
	} else {
		atomicAdd(&s[d], 1u);
	}
}

template <typename K, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_hist(const K *__restrict__ in, uint64_t n, KeyXform xf,
	uint32_t *__restrict__ hist, const Control *ctrl, int reverse)
{
	if (presort_result(ctrl->check, reverse)) return;  // the input is already in order
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_h[P * kRadix];
	for (int i = threadIdx.x; i < P * kRadix; i += THREADS) s_h[i] = 0;
	__syncthreads();
	// one contiguous, 64-key aligned range per block
	const uint64_t chunk = ((n + gridDim.x - 1) / gridDim.x + 63) & ~63ull;
	const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * chunk;
	if (lo < n) {
		const uint32_t len = static_cast<uint32_t>((n - lo) < chunk ? n - lo : chunk);
		for_each_key<K, THREADS>(in, lo, len, [&](K x) {
			const auto k = sort_key<K>(x, xf);
#pragma unroll
			for (int p = 0; p < P; ++p) wave_count(s_h + p * kRadix, static_cast<uint32_t>(k >> (8 * p)) & 0xFFu);
		});
	}
	__syncthreads();
	for (int i = threadIdx.x; i < P * kRadix; i += THREADS)
		if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}


// ---------------------------------------------------------------------------------------------
// presort check: is the input already in order, or strictly in reverse order?

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_check(const K *__restrict__ in, uint64_t n, KeyXform xf, Control *ctrl)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&ctrl->check, 1u);
	if (n < 2) return;
	const uint64_t pairs = n - 1;
	const uint64_t chunk = (pairs + gridDim.x - 1) / gridDim.x;
	const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * chunk;
	const uint64_t hi = min(pairs, lo + chunk);
	uint32_t mine = 0;
	uint64_t it = 0;
	for (uint64_t j = lo + threadIdx.x; j < hi; j += THREADS, ++it) {
		if ((it & 63u) == 0) {
			// stop once both an ascent and a descent were seen anywhere: no shortcut is possible
			const uint32_t seen = __hip_atomic_load(&ctrl->check, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) | mine;
			if ((seen & 10u) == 10u) break;
		}
		const auto a = sort_key<K>(in[j], xf), b = sort_key<K>(in[j + 1], xf);
		mine |= a < b ? 2u : (a == b ? 4u : 8u);
		if ((it & 63u) == 63u && mine) { atomicOr(&ctrl->check, mine); }
	}
	if (mine) atomicOr(&ctrl->check, mine);
}

// One block sorts the whole array in LDS (8-bit LSD passes with the block ranking), keys only.
template <typename K, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_lds_sort(const K *__restrict__ in, K *__restrict__ out, uint32_t n,
	KeyXform xf, int reverse)
{
	constexpr int CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_whist[NW * 256];
	__shared__ uint32_t s_digoff[256];
	__shared__ uint32_t s_wsum[4];
	__shared__ K s_keys[CAP];
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	K keys[KPT];
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		keys[j] = l < n ? in[reverse ? n - 1 - l : l] : K(0);
	}
	for (int p = 0; p < P; ++p) {
		uint32_t dig[KPT], pos[KPT], total;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			dig[j] = l < n ? digit_of<K>(keys[j], xf, 8u * p) : 0xFFu;
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
		if (l < n) out[l] = keys[j];
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
};

__global__ __launch_bounds__(256) void k_plan(PlanArgs a, Control *ctrl, uint32_t *hist, uint32_t *bucket)
{
	__shared__ uint32_t s_wave[4];
	__shared__ uint32_t s_uniform[kMaxPasses];
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	for (int p = 0; p < a.passes; ++p) {
		const uint32_t c = hist[p * kRadix + t];
		hist[p * kRadix + t] = 0;  // leave the histogram zeroed for the next sort
		const uint32_t incl = wave_inclusive_sum(c);
		if (lane == 63) s_wave[w] = incl;
		const bool full = static_cast<uint64_t>(c) == a.n;
		__syncthreads();
		uint32_t before = 0;
		for (uint32_t v = 0; v < w; ++v) before += s_wave[v];
		bucket[p * kRadix + t] = before + incl - c;
		if (t == 0) s_uniform[p] = 0;
		__syncthreads();
		if (full) s_uniform[p] = 1;
		__syncthreads();
	}
	if (t != 0) return;
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
	PassPlan none{};
	ctrl->pre = none;
	ctrl->fin = none;
	for (int p = 0; p < kMaxPasses; ++p) ctrl->pass[p] = none;
	if (A == 0) {
		PassPlan &f = ctrl->fin;
		f.src = a.src; f.dst = a.fin; f.psrc = a.psrc; f.pdst = a.pfin;
		f.iota = a.payload && !a.psrc;
		// already in order: the result is the input (reversed when it was strictly descending)
		f.reverse = pres ? (pres == 2u) : a.reverse;
		f.active = (a.src != a.fin) || f.reverse || (a.payload && (f.iota || a.psrc != a.pfin));
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
	__shared__ uint32_t s_tile;
	__shared__ uint8_t s_dig[PAY ? TILE : 1];
	constexpr int kStage = (PAY && sizeof(K) < 4) ? 4 : static_cast<int>(sizeof(K));
	__shared__ __align__(16) unsigned char s_raw[TILE * kStage];
	K *s_keys = reinterpret_cast<K *>(s_raw);

	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t *pbucket = bucket + pass * kRadix;

	for (;;) {
		if (t == 0) s_tile = atomicAdd(&ctrl->tile_counter[pass], 1u);
		for (int i = t; i < NW * kRadix; i += THREADS) s_whist[i] = 0;
		__syncthreads();
		const uint32_t tile = s_tile;
		if (tile >= num_tiles) break;
		const uint64_t base = static_cast<uint64_t>(tile) * TILE;
		const uint32_t valid = (n - base) < static_cast<uint64_t>(TILE) ? static_cast<uint32_t>(n - base) : static_cast<uint32_t>(TILE);

		K keys[KPT];
		uint32_t dig[KPT];
		uint32_t rank[KPT];
		uint32_t pay[PAY ? KPT : 1];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			const uint64_t g = base + l;
			if (l < valid) {
				const uint64_t s = pp.reverse ? n - 1 - g : g;
				keys[j] = src[s];
				dig[j] = digit_of<K>(keys[j], xf, shift);
				if constexpr (PAY) pay[j] = pp.iota ? static_cast<uint32_t>(s) : pp.psrc[s];
			} else {
				keys[j] = K(0);
				dig[j] = 0xFFu;  // past the end: ranked last, never written
				if constexpr (PAY) pay[j] = 0;
			}
		}
		// rank within the wave in logical order (j-major, lane-minor)
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint64_t peers = match_digit(dig[j]);
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
			atomic_store_agent(&lookback[static_cast<uint64_t>(tile) * kRadix + t],
				lb_word(pub, tile == 0 ? kLbInclusive : kLbAggregate, tag));
			const uint32_t incl = wave_inclusive_sum(total);
			if (lane == 63) s_wsum[w] = incl;
			s_digoff[t] = incl - total;
		}
		__syncthreads();
		if (t < kRadix) {
			uint32_t before = 0;
			for (uint32_t v = 0; v < w; ++v) before += s_wsum[v];
			const uint32_t digoff = s_digoff[t] + before;
			s_digoff[t] = digoff;
			// decoupled look-back over the preceding tiles of this digit
			uint32_t excl = 0;
			if (tile > 0) {
				excl = lookback_window<16>(lookback + t, tile, 0, tag);
				atomic_store_agent(&lookback[static_cast<uint64_t>(tile) * kRadix + t], lb_word(excl + pub, kLbInclusive, tag));
			}
			s_glob[t] = pbucket[t] + excl - digoff;
		}
		__syncthreads();
		// stage the tile in LDS in digit order
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t pos = s_digoff[dig[j]] + s_whist[w * kRadix + dig[j]] + rank[j];
			s_keys[pos] = keys[j];
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
			for (int j = 0; j < KPT; ++j) {
				const uint32_t pos = s_digoff[dig[j]] + s_whist[w * kRadix + dig[j]] + rank[j];
				s_pay[pos] = pay[j];
			}
			__syncthreads();
			for (uint32_t i = t; i < valid; i += THREADS) pp.pdst[s_glob[s_dig[i]] + i] = s_pay[i];
		}
		__syncthreads();
	}
}


// ---------------------------------------------------------------------------------------------
// count + chained scan + scatter (no inter-block dependency inside the scatter)


// Tile-based read of every key: the histograms of all digits (as k_hist) and, when counts is not
// null, the per-tile counts of digit 0 in the logical order of the first pass (tile t holds the
// logical positions [t*tile_keys, (t+1)*tile_keys), read back to front under isrevorder).
template <typename K, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_hist_tiles(const K *__restrict__ in, uint64_t n, KeyXform xf,
	uint32_t *__restrict__ hist, uint32_t *__restrict__ counts, uint32_t tile_keys, uint32_t num_tiles, int reverse,
	const Control *ctrl)
{
	if (presort_result(ctrl->check, reverse)) return;  // the input is already in order
	constexpr int P = sizeof(K);
	__shared__ uint32_t s_h[P * kRadix];
	__shared__ uint32_t s_t[kRadix];
	for (int i = threadIdx.x; i < P * kRadix; i += THREADS) s_h[i] = 0;
	for (uint32_t tile = blockIdx.x; tile < num_tiles; tile += gridDim.x) {
		for (int i = threadIdx.x; i < kRadix; i += THREADS) s_t[i] = 0;
		__syncthreads();
		const uint64_t base = static_cast<uint64_t>(tile) * tile_keys;
		const uint32_t len = (n - base) < tile_keys ? static_cast<uint32_t>(n - base) : tile_keys;
		// the tile's keys as a physical range (the logical range read back to front under isrevorder)
		const uint64_t lo = reverse ? n - base - len : base;
		for_each_key<K, THREADS>(in, lo, len, [&](K x) {
			const auto k = sort_key<K>(x, xf);
#pragma unroll
			for (int p = 0; p < P; ++p) wave_count(s_h + p * kRadix, static_cast<uint32_t>(k >> (8 * p)) & 0xFFu);
			if (counts) wave_count(s_t, static_cast<uint32_t>(k) & 0xFFu);
		});
		__syncthreads();
		if (counts)
			for (int i = threadIdx.x; i < kRadix; i += THREADS) counts[static_cast<uint64_t>(tile) * kRadix + i] = s_t[i];
		__syncthreads();
	}
	__syncthreads();
	for (int i = threadIdx.x; i < P * kRadix; i += THREADS)
		if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
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
	__shared__ uint32_t s_t[kRadix];
	for (uint32_t tile = blockIdx.x; tile < num_tiles; tile += gridDim.x) {
		for (int i = threadIdx.x; i < kRadix; i += THREADS) s_t[i] = 0;
		__syncthreads();
		const uint64_t base = static_cast<uint64_t>(tile) * tile_keys;
		const uint32_t len = (n - base) < tile_keys ? static_cast<uint32_t>(n - base) : tile_keys;
		const uint64_t lo = pp.reverse ? n - base - len : base;
		for_each_key<K, THREADS>(src, lo, len, [&](K x) { wave_count(s_t, digit_of<K>(x, xf, shift)); });
		__syncthreads();
		for (int i = threadIdx.x; i < kRadix; i += THREADS) counts[static_cast<uint64_t>(tile) * kRadix + i] = s_t[i];
		__syncthreads();
	}
}

// Chained scan with decoupled look-back over the tile-count matrix: every block takes the next
// chunk of kScanChunk tiles, scans it per digit, publishes the chunk total, resolves the prefix
// of the preceding chunks and turns the counts into global output offsets in place.
__global__ __launch_bounds__(256) void k_scan(Control *ctrl, int pass, const uint32_t *__restrict__ bucket,
	uint32_t *__restrict__ counts, uint64_t *__restrict__ lookback, uint32_t num_tiles)
{
	const PassPlan pp = ctrl->pass[pass];
	if (!pp.active) return;
	const uint32_t nchunks = (num_tiles + kScanChunk - 1) / kScanChunk;
	const uint32_t tag = pass_tag(ctrl->epoch, pass);
	const uint32_t d = threadIdx.x;
	const uint32_t start = bucket[pass * kRadix + d];
	__shared__ uint32_t s_chunk;
	for (;;) {
		if (d == 0) s_chunk = atomicAdd(&ctrl->scan_counter[pass], 1u);
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
			for (int j = 0; j < 8; ++j) c[j] = (t + j < t1) ? counts[static_cast<uint64_t>(t + j) * kRadix + d] : 0u;
#pragma unroll
			for (int j = 0; j < 8; ++j) {
				if (t + j < t1) counts[static_cast<uint64_t>(t + j) * kRadix + d] = run;
				run += c[j];
			}
		}
		uint32_t excl = 0;
		uint64_t *lb = lookback + static_cast<uint64_t>(chunk) * kRadix + d;
		if (chunk == 0) {
			atomic_store_agent(lb, lb_word(run, kLbInclusive, tag));
		} else {
			atomic_store_agent(lb, lb_word(run, kLbAggregate, tag));
			excl = lookback_window<8>(lookback + d, chunk, 0, tag);
			atomic_store_agent(lb, lb_word(excl + run, kLbInclusive, tag));
		}
		const uint32_t add = start + excl;
		for (uint32_t t = t0; t < t1; ++t) counts[static_cast<uint64_t>(t) * kRadix + d] += add;
	}
}

// Scatter of one digit with known tile offsets. Blocks are dealt to the XCDs round robin, so
// block b works on the tile range of XCD b % 8: neighbouring tiles, whose digit runs are
// neighbours in the output, are written through the same L2.
template <typename K, bool PAY, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_scatter(const Control *ctrl, int pass, KeyXform xf, uint64_t n,
	uint32_t num_tiles, const uint32_t *__restrict__ offsets)
{
	constexpr int TILE = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	const PassPlan pp = ctrl->pass[pass];
	if (!pp.active) return;
	const K *__restrict__ src = static_cast<const K *>(pp.src);
	K *__restrict__ dst = static_cast<K *>(pp.dst);
	const uint32_t shift = 8u * static_cast<uint32_t>(pass);
	__shared__ uint32_t s_whist[NW * kRadix];
	__shared__ uint32_t s_digoff[kRadix];
	__shared__ uint32_t s_glob[kRadix];
	__shared__ uint32_t s_wsum[4];
	__shared__ uint8_t s_dig[PAY ? TILE : 1];
	constexpr int kStage = (PAY && sizeof(K) < 4) ? 4 : static_cast<int>(sizeof(K));
	__shared__ __align__(16) unsigned char s_raw[TILE * kStage];
	K *s_keys = reinterpret_cast<K *>(s_raw);
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t per = (num_tiles + 7) / 8;
	const uint32_t x = blockIdx.x & 7u, step = gridDim.x >> 3;
	K keys[KPT];
	uint32_t pay[PAY ? KPT : 1];
	// loads of a tile are issued before the stores of the previous one, so that waiting for them
	// (the memory counter is in order and counts stores too) never waits for those stores
	auto load_tile = [&](uint32_t tl) {
		const uint64_t base = static_cast<uint64_t>(tl) * TILE;
		const uint32_t valid = (n - base) < static_cast<uint64_t>(TILE) ? static_cast<uint32_t>(n - base) : static_cast<uint32_t>(TILE);
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
	if (k < per && tile < num_tiles) load_tile(tile);
	while (k < per && tile < num_tiles) {
		const uint64_t base = static_cast<uint64_t>(tile) * TILE;
		const uint32_t valid = (n - base) < static_cast<uint64_t>(TILE) ? static_cast<uint32_t>(n - base) : static_cast<uint32_t>(TILE);
		uint32_t dig[KPT], pos[KPT], total;
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			dig[j] = l < valid ? digit_of<K>(keys[j], xf, shift) : 0xFFu;
		}
		block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total);
		if (t < kRadix) s_glob[t] = offsets[static_cast<uint64_t>(tile) * kRadix + t] - s_digoff[t];
#pragma unroll
		for (int j = 0; j < KPT; ++j) s_keys[pos[j]] = keys[j];
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
		if (k2 < per && tile2 < num_tiles) load_tile(tile2);
		__syncthreads();
		for (uint32_t i = t; i < valid; i += THREADS) {
			const K kk = s_keys[i];
			const uint32_t dd = digit_of<K>(kk, xf, shift);
			if constexpr (PAY) s_dig[i] = static_cast<uint8_t>(dd);
			dst[s_glob[dd] + i] = kk;
		}
		if constexpr (PAY) {
			__syncthreads();
			uint32_t *s_pay = reinterpret_cast<uint32_t *>(s_raw);
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_pay[ppos[j]] = pcur[j];
			__syncthreads();
			for (uint32_t i = t; i < valid; i += THREADS) pp.pdst[s_glob[s_dig[i]] + i] = s_pay[i];
		}
		__syncthreads();
		k = k2;
		tile = tile2;
	}
}

// ---------------------------------------------------------------------------------------------
// 8-bit count fill and finish

__global__ __launch_bounds__(256) void k_fill8(const uint32_t *__restrict__ bucket, KeyXform xf,
	uint8_t *__restrict__ out, uint64_t n)
{
	__shared__ uint32_t s_b[kRadix];
	__shared__ uint8_t s_inv[kRadix];
	s_b[threadIdx.x] = bucket[threadIdx.x];
	// the bijective 8-bit transforms are invertible: bucket f(x) holds only the pattern x
	s_inv[sort_key<uint8_t>(static_cast<uint8_t>(threadIdx.x), xf) & 0xFFu] = static_cast<uint8_t>(threadIdx.x);
	__syncthreads();
	const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		// last bucket whose start is <= i
		uint32_t lo = 0;
#pragma unroll
		for (uint32_t step = 128; step; step >>= 1)
			if (s_b[lo + step] <= i) lo += step;
		out[i] = s_inv[lo];
	}
}

__global__ void k_finish(Control *ctrl)
{
	if (threadIdx.x < kMaxPasses) {
		ctrl->tile_counter[threadIdx.x] = 0;
		ctrl->scan_counter[threadIdx.x] = 0;
	}
	if (threadIdx.x == 0) {
		ctrl->epoch = ctrl->epoch + 1u;
		ctrl->check = 0;
	}
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

struct ProfScope {
	bool on;
	ProfRec rec;
	hipStream_t stream;
	ProfScope(int kind, int pass, const void *ws, hipStream_t s) : on(false), rec{kind, pass, ws, nullptr, nullptr}, stream(s)
	{
		std::lock_guard<std::mutex> lock(g_prof_mu);
		if (!g_prof_on) return;
		if (hipEventCreate(&rec.a) != hipSuccess || hipEventCreate(&rec.b) != hipSuccess) return;
		on = hipEventRecord(rec.a, s) == hipSuccess;
	}
	~ProfScope()
	{
		if (!on) return;
		std::lock_guard<std::mutex> lock(g_prof_mu);
		if (hipEventRecord(rec.b, stream) == hipSuccess) g_prof.push_back(rec);
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
	case 3: return launch_pass<K, PAY, 512, 16>(ctrl, bucket, lookback, n, xf, pass, stream);
	default: return launch_pass<K, PAY, 256, 24>(ctrl, bucket, lookback, n, xf, pass, stream);
	}
}

template <typename K, bool PAY, int THREADS, int KPT>
hipError_t launch_scatter(Control *ctrl, int pass, const KeyXform &xf, uint64_t n, uint32_t tiles, const uint32_t *offsets,
	hipStream_t stream)
{
	static int occ = -1;
	if (occ < 0) {
		int o = 0;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_scatter<K, PAY, THREADS, KPT>, THREADS, 0) != hipSuccess || o < 1) o = 1;
		occ = o;
	}
	const uint32_t per = (tiles + 7) / 8;
	const uint32_t want = static_cast<uint32_t>(occ * devinfo().cus * 2 + 7) / 8;
	const uint32_t grid = 8u * std::max<uint32_t>(1u, std::min<uint32_t>(per, want));
	hipLaunchKernelGGL((k_scatter<K, PAY, THREADS, KPT>), dim3(grid), dim3(THREADS), 0, stream, ctrl, pass, xf, n, tiles, offsets);
	return hipGetLastError();
}

template <typename K, bool PAY>
hipError_t launch_scatter_cfg(int cfg, Control *ctrl, int pass, const KeyXform &xf, uint64_t n, uint32_t tiles,
	const uint32_t *offsets, hipStream_t stream)
{
	switch (cfg) {
	case 0: return launch_scatter<K, PAY, 256, 8>(ctrl, pass, xf, n, tiles, offsets, stream);
	case 1: return launch_scatter<K, PAY, 256, 16>(ctrl, pass, xf, n, tiles, offsets, stream);
	case 2: return launch_scatter<K, PAY, 512, 8>(ctrl, pass, xf, n, tiles, offsets, stream);
	case 3: return launch_scatter<K, PAY, 512, 16>(ctrl, pass, xf, n, tiles, offsets, stream);
	default: return launch_scatter<K, PAY, 256, 24>(ctrl, pass, xf, n, tiles, offsets, stream);
	}
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
	const int cfg = job.cfg >= 0 ? job.cfg : default_cfg(P, n, job.payload);
	int algo = job.algo ? job.algo : default_algo(P, n);
	if (algo == kAlgoLds) {
		if (!job.payload && n <= kLdsMax && !job.keys_final_any) {
			// one block holds the whole array: every load happens before any store, so in place is fine
			const int rev = (job.flags & kRev) ? 1 : 0;
			const uint32_t nn = static_cast<uint32_t>(n);
			ProfScope ps(1, 0, ws, stream);
			if (n <= 1024)
				hipLaunchKernelGGL((k_lds_sort<K, 256, 4>), dim3(1), dim3(256), 0, stream, static_cast<const K *>(job.in), static_cast<K *>(job.out), nn, xf, rev);
			else if (n <= 4096)
				hipLaunchKernelGGL((k_lds_sort<K, 256, 16>), dim3(1), dim3(256), 0, stream, static_cast<const K *>(job.in), static_cast<K *>(job.out), nn, xf, rev);
			else
				hipLaunchKernelGGL((k_lds_sort<K, 512, 16>), dim3(1), dim3(512), 0, stream, static_cast<const K *>(job.in), static_cast<K *>(job.out), nn, xf, rev);
			return hipGetLastError();
		}
		algo = default_algo(P, n);
	}
	const uint32_t tile_keys = static_cast<uint32_t>(kTileCfgs[cfg].threads * kTileCfgs[cfg].kpt);
	const uint32_t tiles = static_cast<uint32_t>((n + tile_keys - 1) / tile_keys);
	uint32_t *counts = counts_of(ws, n);
	const bool fill = (P == 1) && !job.payload && job.fill8 &&
		(job.flags & 7u) != (kAbs | kSign) && (job.flags & 7u) != (kSign | kFltp) && (job.flags & 7u) != (kAbs | kSign | kFltp);
	const bool counts0 = algo == kAlgoCountScan && !fill;
	if (job.presort_check && !fill) {
		const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((n + 4095) / 4096, 4u * cus)));
		hipLaunchKernelGGL((k_check<K, 256>), dim3(grid), dim3(256), 0, stream, static_cast<const K *>(job.in), n, xf, ctrl);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	{
		ProfScope ps(0, 0, ws, stream);
		if (counts0) {
			const uint32_t grid = std::max<uint32_t>(1u, std::min<uint32_t>(tiles, 8u * cus));
			hipLaunchKernelGGL((k_hist_tiles<K, 256, HITEMS>), dim3(grid), dim3(256), 0, stream, static_cast<const K *>(job.in), n,
				xf, hist, counts, tile_keys, tiles, (job.flags & kRev) ? 1 : 0, ctrl);
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

	hipLaunchKernelGGL(k_plan, dim3(1), dim3(256), 0, stream, pa, ctrl, hist, bucket);
	if ((e = hipGetLastError()) != hipSuccess) return e;

	const uint32_t cgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8u * cus)));
	if (fill) {
		// bijective 8-bit modes: the sorted output follows from the counts alone
		hipLaunchKernelGGL(k_fill8, dim3(cgrid), dim3(256), 0, stream, bucket, xf, static_cast<uint8_t *>(job.out), n);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	} else {
		hipLaunchKernelGGL((k_copy<K>), dim3(cgrid), dim3(256), 0, stream, ctrl, 0, n, job.payload ? 1 : 0);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		for (int p = 0; p < P; ++p) {
			if (algo == kAlgoOnesweep) {
				ProfScope ps(1, p, ws, stream);
				e = job.payload ? launch_pass_cfg<K, true>(cfg, ctrl, bucket, lookback, n, xf, p, stream)
				                : launch_pass_cfg<K, false>(cfg, ctrl, bucket, lookback, n, xf, p, stream);
				if (e != hipSuccess) return e;
				continue;
			}
			{
				ProfScope ps(3, p, ws, stream);
				const uint32_t cgrid = std::max<uint32_t>(1u, std::min<uint32_t>(tiles, 8u * cus));
				hipLaunchKernelGGL((k_count<K, 256, HITEMS>), dim3(cgrid), dim3(256), 0, stream, ctrl, p, xf, n, tile_keys, tiles, counts);
				if ((e = hipGetLastError()) != hipSuccess) return e;
				const uint32_t nchunks = (tiles + kScanChunk - 1) / kScanChunk;
				const uint32_t sgrid = std::max<uint32_t>(1u, std::min<uint32_t>(nchunks, 4u * cus));
				hipLaunchKernelGGL(k_scan, dim3(sgrid), dim3(256), 0, stream, ctrl, p, bucket, counts, lookback, tiles);
				if ((e = hipGetLastError()) != hipSuccess) return e;
			}
			ProfScope ps(1, p, ws, stream);
			e = job.payload ? launch_scatter_cfg<K, true>(cfg, ctrl, p, xf, n, tiles, counts, stream)
			                : launch_scatter_cfg<K, false>(cfg, ctrl, p, xf, n, tiles, counts, stream);
			if (e != hipSuccess) return e;
		}
		hipLaunchKernelGGL((k_copy<K>), dim3(cgrid), dim3(256), 0, stream, ctrl, 1, n, job.payload ? 1 : 0);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, stream, ctrl);
	return hipGetLastError();
}

} // namespace

int device_cu_count() { return devinfo().cus; }

void profile_begin()
{
	std::lock_guard<std::mutex> lock(g_prof_mu);
	for (auto &r : g_prof) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
	g_prof.clear();
	g_prof_on = true;
}

// Sums the recorded launch times per kind: 0 histogram, 1 scatter launches of executed passes,
// 2 scatter launches of skipped passes, 3 count+scan launches of executed passes, 4 count+scan
// launches of skipped passes. A pass counts as active by the plan its workspace holds now (the
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
		if (e != hipSuccess) err = e;
		else if (kind < nkinds) { ms[kind] += t; ++launches[kind]; }
		(void)hipEventDestroy(r.a);
		(void)hipEventDestroy(r.b);
	}
	return err;
}

// measured on MI355X (2^20..2^28 uniform keys): the count + scan path wins from about 2^22 keys,
// the onesweep path below that (fewer launches)
int default_algo(int key_bytes, uint64_t n)
{
	(void)key_bytes;
	return n >= (1u << 22) ? kAlgoCountScan : kAlgoOnesweep;
}

int default_cfg(int key_bytes, uint64_t n, bool payload)
{
	(void)payload;
	if (n <= (1u << 16)) return 0;
	if (default_algo(key_bytes, n) == kAlgoCountScan) return key_bytes >= 8 ? 4 : 1;
	return key_bytes >= 8 ? 4 : 2;
}

size_t sort_workspace_bytes(uint64_t n)
{
	// look-back words (one per tile and digit) followed by the tile-count matrix
	return kLookbackOffset + static_cast<size_t>(max_tiles(n)) * kRadix * (sizeof(uint64_t) + sizeof(uint32_t));
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
