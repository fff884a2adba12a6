// rsbd8_device.hpp -- device-side building blocks of the gfx950 rsbd8 radix sort.
//
// The sort is an 8-bit-digit LSD radix sort over a "sort key" derived from the raw key bits.
// Every mode of the reference (Radixsortbidi8.hpp = H, filterinput H:10000-10027, the virtual
// top-bit flips of the offsets transform H:11642-11801) reduces to one branch-free transform
//     k = (((((x & A) << sh) | (s & B)) ^ (smask & C)) + (s & D)) ^ E
// with s the sign bit of x and smask = -s, with descending order folded into E (complement).
// Arrays always hold the original bits; the transform is recomputed on every load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsbd8hip {

constexpr int kRadixBits = 8;
constexpr int kRadix = 256;
constexpr int kMaxPasses = 8;

// flag bits of the C-ABI (include/rsbd8_hip.h)
constexpr unsigned kAbs = 1u, kSign = 2u, kFltp = 4u, kDesc = 8u, kRev = 16u;

struct KeyXform {
	uint64_t A, B, C, D, E;
	uint32_t sh;
	uint32_t ident;  // the transform is the identity (unsigned keys, ascending): skip the arithmetic
};

// host+device: transform parameters for a W-bit key with the decoded flags
inline __host__ __device__ KeyXform make_xform(unsigned width_bits, unsigned flags)
{
	const uint64_t ones = width_bits >= 64 ? ~0ull : ((1ull << width_bits) - 1ull);
	const uint64_t msb = 1ull << (width_bits - 1);
	KeyXform t{ones, 0, 0, 0, 0, 0, 0};
	switch (flags & 7u) {
	case 0u: break;                                              // forceunsigned
	case kSign: t.E = msb; break;                                // forcesigned
	case kAbs | kSign: t.C = ones; t.D = 1; break;               // forceabssigned: -x for x < 0
	case kAbs: t.sh = 1; t.C = ones; break;                      // forcetieredabssigned
	case kSign | kFltp: t.A = ones & ~msb; t.C = ones; t.D = 1; t.E = msb; break; // forcefloatingp
	case kAbs | kSign | kFltp: t.A = ones & ~msb; break;         // forceabsfloatingp
	case kFltp: t.A = ones & ~msb; t.C = ones; break;            // specialfloatingp
	default: t.sh = 1; t.B = 1; t.E = 1; break;                  // forcetieredabsfloatingp
	}
	if (flags & kDesc) t.E ^= ones;
	t.ident = (t.A == ones && t.sh == 0 && t.B == 0 && t.C == 0 && t.D == 0 && t.E == 0) ? 1u : 0u;
	return t;
}

template <typename K> struct KeyTraits;
template <> struct KeyTraits<uint8_t> { using W = uint32_t; static constexpr int bits = 8; };
template <> struct KeyTraits<uint16_t> { using W = uint32_t; static constexpr int bits = 16; };
template <> struct KeyTraits<uint32_t> { using W = uint32_t; static constexpr int bits = 32; };
template <> struct KeyTraits<uint64_t> { using W = uint64_t; static constexpr int bits = 64; };

// the transform in the key's working width; bits above the key width are never looked at
template <typename K>
__device__ __forceinline__ typename KeyTraits<K>::W sort_key(K raw, const KeyXform &t)
{
	using W = typename KeyTraits<K>::W;
	const W x = static_cast<W>(raw);
	if (t.ident) return x;  // uniform across the launch: no divergence
	const W s = x >> (KeyTraits<K>::bits - 1);
	const W smask = W(0) - s;
	W k = (x & static_cast<W>(t.A)) << t.sh;
	k |= s & static_cast<W>(t.B);
	k ^= smask & static_cast<W>(t.C);
	k += s & static_cast<W>(t.D);
	return k ^ static_cast<W>(t.E);
}

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K raw, const KeyXform &t, uint32_t shift)
{
	return static_cast<uint32_t>(sort_key<K>(raw, t) >> shift) & 0xFFu;
}

// A splitter table of the sample partition plan staged in LDS: 256 sorted bounds (unused ones
// all-ones), per bound index t the open bucket id (bits 0-7) and whether bound t has an equality
// bucket (bit 8, id + 1), and a cell table over [lo, hi] (the lowest and highest real bound): cell
// c covers the keys lo + (c << sh) ... lo + ((c + 1) << sh) - 1 and holds the number of bounds below
// its first key (bits 0-7) and how many more lie below its last key (bits 8-15).
constexpr uint32_t kSplitCells = 4096;
constexpr uint32_t kSplitSubs = 16;      // sub-tables, for the first cells found with eight bounds or more
constexpr uint32_t kSplitSubCells = 16;  // sub-cells per sub-table
template <typename W>
struct SplitLds {
	W b[256];
	uint32_t info[256];
	uint16_t cell[kSplitCells];
	uint16_t sub[kSplitSubs * kSplitSubCells];
	W lo, hi;
	uint32_t sh;
	uint32_t nsub;  // sub-tables in use
};

// Bucket id of the transformed key k: t = the number of bounds below k (0 at or below lo, all of
// them above hi, else from the key's cell, refined by a binary search over the few bounds that
// fall inside the cell), then the open bucket below bound t, or bound t's equality bucket when k
// equals it. Ids grow with the key. A cell holding eight bounds or more (span field 0xFF) points
// to a sub-table of 16 sub-cells (index in its low byte), which take the search's place; SUB false
// is for tables with no sub-table (tb.nsub 0: keys spread over their range), whose lookups then
// take no second cell read (callers branch once on tb.nsub).
template <typename W, bool SUB = true>
__device__ __forceinline__ uint32_t split_id(const SplitLds<W> &tb, W k)
{
	uint32_t t;
	if (k <= tb.lo) {
		t = 0;
	} else if (k > tb.hi) {
		t = tb.info[255] >> 16;  // the number of real bounds
	} else {
		const W off = k - tb.lo;
		uint32_t e = tb.cell[static_cast<uint32_t>(off >> tb.sh)];
		if (SUB && (e >> 8) == 0xFFu)
			e = tb.sub[(e & 0xFFu) * kSplitSubCells + (static_cast<uint32_t>(off >> (tb.sh - 4u)) & (kSplitSubCells - 1u))];
		t = e & 0xFFu;
		uint32_t span = e >> 8;
		while (span) {
			const uint32_t half = span >> 1;
			if (tb.b[t + half] < k) {
				t += half + 1;
				span -= half + 1;
			} else {
				span = half;
			}
		}
	}
	const uint32_t inf = tb.info[t];
	return (inf & 0xFFu) + (((inf >> 8) & 1u) & static_cast<uint32_t>(tb.b[t] == k));
}

// lanes of the wave holding the same 8-bit digit (wave64 ballot match)
__device__ __forceinline__ uint64_t match_digit(uint32_t d)
{
	uint64_t m = ~0ull;
#pragma unroll
	for (int b = 0; b < 8; ++b) {
		const bool bit = (d >> b) & 1u;
		const uint64_t v = __ballot(bit);
		m &= bit ? v : ~v;
	}
	return m;
}

// number of set bits of m in lanes below this one
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
	return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
		__builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Adds 1 to digit d's count for every active lane, with one atomic for the whole wave when all
// lanes share d. The counts live in C interleaved copies (s[d * C + c], 256 * C words): lane l adds
// to copy l % C, so a digit that many lanes hold (skewed keys) takes C LDS words in C banks instead
// of serialising on one. A digit's count is the sum of its C words (copies_sum).
template <int C = 1>
__device__ __forceinline__ void wave_count(uint32_t *s, uint32_t d)
{
	static_assert(C >= 1 && C <= 64 && (C & (C - 1)) == 0, "copies: a power of two up to 64");
	const uint64_t active = __ballot(1);
	const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
	if (__ballot(d == d0) == active) {
		if (lanes_below(active) == 0) atomicAdd(&s[d0 * C], static_cast<uint32_t>(__popcll(active)));
	} else {
		atomicAdd(&s[d * C + (threadIdx.x & (C - 1))], 1u);
	}
}

template <int C>
__device__ __forceinline__ uint32_t copies_sum(const uint32_t *s, uint32_t d)
{
	uint32_t v = 0;
#pragma unroll
	for (int c = 0; c < C; ++c) v += s[d * C + c];
	return v;
}

// Counter copies for a counting kernel that keeps `copied` 256-word histograms with copies and
// `plain` ones without in LDS (1 KB per histogram and copy): as many copies as fit 19 KB, so that
// eight 256-thread blocks still fit one CU's 160 KB, and at most 8.
constexpr int count_copies(int copied, int plain = 0)
{
	return plain + copied * 8 <= 19 ? 8 : (plain + copied * 4 <= 19 ? 4 : (plain + copied * 2 <= 19 ? 2 : 1));
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// inclusive wave64 prefix sum
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
	const uint32_t lane = lane_id();
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t u = __shfl_up(v, o, 64);
		if (lane >= static_cast<uint32_t>(o)) v += u;
	}
	return v;
}

// look-back word: value in bits 0..31, status in bits 32..33, pass tag in bits 34..41
constexpr uint64_t kLbAggregate = 1ull << 32;
constexpr uint64_t kLbInclusive = 2ull << 32;
__device__ __forceinline__ uint64_t lb_word(uint32_t value, uint64_t status, uint32_t tag)
{
	return static_cast<uint64_t>(value) | status | (static_cast<uint64_t>(tag) << 34);
}

__device__ __forceinline__ uint64_t atomic_load_agent(const uint64_t *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_agent(uint64_t *p, uint64_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

} // namespace rsbd8hip

namespace rsbd8hip {

// Lanes of a full wave (all 64 lanes active) that hold the same 8-bit digit d. Three ways, chosen
// per wave by how many lanes share lane 0's digit:
// - all of them: the whole wave;
// - at least 16 (few distinct digits, so LDS atomics on one word would serialise): 8 ballots;
// - fewer: every lane ORs its bit into the wave's LDS word for d, reads the word back (a wave's
//   LDS operations execute in order, so every OR has landed), and the lowest peer clears it.
// s_peer: the wave's 256 words, zero on entry and left zero.
__device__ __forceinline__ uint64_t wave_peers(uint32_t d, uint64_t *s_peer)
{
	const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
	const uint64_t same = __ballot(d == d0);
	if (same == ~0ull) return same;
	if (__popcll(same) >= 16) return match_digit(d);
	uint64_t *word = s_peer + d;
	__hip_atomic_fetch_or(word, 1ull << (threadIdx.x & 63u), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
	const uint64_t peers = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
	if (lanes_below(peers) == 0) __hip_atomic_store(word, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
	return peers;
}

// LDS bytes of block_rank's peer words (s_peer): 256 words of 8 bytes per wave
constexpr size_t rank_peer_bytes(int threads) { return static_cast<size_t>(threads / 64) * 256u * 8u; }
constexpr size_t max_bytes(size_t a, size_t b) { return a > b ? a : b; }

// Block-wide stable rank of THREADS*KPT digits held in registers in logical order
// l = wave*64*KPT + j*64 + lane (wave-major, then slot j, then lane). Returns each key's
// position in the digit-sorted block and leaves, for threads t < 256, the digit's block total in
// *total and its exclusive digit offset in s_digoff[t]. The three __syncthreads() inside must be
// reached by every thread of the block. s_whist holds (THREADS/64)*256 words, s_wsum 4 words,
// s_peer rank_peer_bytes(THREADS) bytes (8-byte aligned) that are free from entry to return: they
// may share storage with the tile staging, which is written only after the ranking.
template <int THREADS, int KPT>
__device__ __forceinline__ void block_rank(const uint32_t (&dig)[KPT], uint32_t (&pos)[KPT], uint32_t *s_whist,
	uint32_t *s_digoff, uint32_t *s_wsum, uint32_t *total, uint64_t *s_peer)
{
	constexpr int NW = THREADS / 64;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	for (int i = t; i < NW * 256; i += THREADS) {
		s_whist[i] = 0;
		s_peer[i] = 0;
	}
	__syncthreads();
	uint32_t rank[KPT];
	uint64_t *s_wpeer = s_peer + w * 256;
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint64_t peers = wave_peers(dig[j], s_wpeer);
		const uint32_t below = lanes_below(peers);
		uint32_t *slot = &s_whist[w * 256 + dig[j]];
		const uint32_t old = *slot;
		if (below == 0) *slot = old + static_cast<uint32_t>(__popcll(peers));
		rank[j] = old + below;
	}
	__syncthreads();
	uint32_t tot = 0;
	if (t < 256) {
		for (int v = 0; v < NW; ++v) {
			const uint32_t c = s_whist[v * 256 + t];
			s_whist[v * 256 + t] = tot;
			tot += c;
		}
		const uint32_t incl = wave_inclusive_sum(tot);
		if (lane == 63) s_wsum[w] = incl;
		s_digoff[t] = incl - tot;
	}
	__syncthreads();
	if (t < 256) {
		uint32_t before = 0;
		for (uint32_t v = 0; v < w; ++v) before += s_wsum[v];
		s_digoff[t] += before;
		*total = tot;
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < KPT; ++j) pos[j] = s_digoff[dig[j]] + s_whist[w * 256 + dig[j]] + rank[j];
}

} // namespace rsbd8hip

namespace rsbd8hip {

// Claims one slot of s[d] per active lane and returns it: one atomic for the whole wave when all
// active lanes share d (slots in lane order), else one per lane (slots in no particular order).
__device__ __forceinline__ uint32_t wave_claim(uint32_t *s, uint32_t d)
{
	const uint64_t active = __ballot(1);
	const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
	if (__ballot(d == d0) == active) {
		const uint32_t below = lanes_below(active);
		uint32_t base = 0;
		if (below == 0) base = atomicAdd(&s[d0], static_cast<uint32_t>(__popcll(active)));
		return __builtin_amdgcn_readfirstlane(base) + below;
	}
	return atomicAdd(&s[d], 1u);
}

// Block-wide digit grouping of the `valid` first keys (logical order l = wave*64*KPT + j*64 +
// lane, as in block_rank): pos[j] is a place in the digit-sorted tile, inside its digit's run but
// in no particular order within it (LDS atomics), and s_digoff[d] the digit's exclusive offset.
// Keys at l >= valid get no place. s_cnt holds 256 words, s_wsum THREADS/64 words; every thread
// must reach the barriers.
template <int THREADS, int KPT>
__device__ __forceinline__ void block_group(const uint32_t (&dig)[KPT], uint32_t valid, uint32_t (&pos)[KPT],
	uint32_t *s_cnt, uint32_t *s_digoff, uint32_t *s_wsum)
{
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	for (uint32_t i = t; i < 256; i += THREADS) s_cnt[i] = 0;
	__syncthreads();
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		if (l < valid) wave_count(s_cnt, dig[j]);
	}
	__syncthreads();
	if (t < 256) {
		const uint32_t c = s_cnt[t];
		const uint32_t incl = wave_inclusive_sum(c);
		if (lane == 63) s_wsum[w] = incl;
		s_digoff[t] = incl - c;
	}
	__syncthreads();
	if (t < 256) {
		uint32_t before = 0;
		for (uint32_t v = 0; v < w; ++v) before += s_wsum[v];
		s_digoff[t] += before;
		s_cnt[t] = s_digoff[t];
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		pos[j] = l < valid ? wave_claim(s_cnt, dig[j]) : 0xFFFFFFFFu;
	}
}

} // namespace rsbd8hip

namespace rsbd8hip {

// Decoupled look-back for one digit: sums the published counts of the tiles before `tile` back to
// the nearest inclusive prefix, reading W predecessors per round so that W memory latencies overlap.
// `words` points at the look-back word of this digit for tile 0 (STRIDE words per tile). Every
// tile at or after `first` (the first tile of the sorted range) publishes before it looks back, and
// tile `first` publishes an inclusive prefix, so the walk always ends.
template <int W, int64_t STRIDE = 256>
__device__ __forceinline__ uint32_t lookback_window(const uint64_t *words, int64_t tile, int64_t first, uint32_t tag)
{
	uint32_t excl = 0;
	int64_t prev = tile - 1;
	for (;;) {
		uint64_t v[W];
#pragma unroll
		for (int i = 0; i < W; ++i) v[i] = (prev - i >= first) ? atomic_load_agent(words + (prev - i) * STRIDE) : 0ull;
		// consume the published words nearest first, up to the first inclusive one; stop at the
		// first word not published yet (read again from there). Predicated, so the window stays in
		// registers.
		int used = 0;
		bool done = false, stop = false;
#pragma unroll
		for (int i = 0; i < W; ++i) {
			const uint64_t x = v[i];
			const bool ok = !stop && (x >> 34) == tag && (x & (3ull << 32)) != 0;
			excl += ok ? static_cast<uint32_t>(x) : 0u;
			used += ok ? 1 : 0;
			done = done || (ok && (x & kLbInclusive) != 0);
			stop = stop || !ok || done;
		}
		if (done) return excl;
		prev -= used;
		if (used == 0) __builtin_amdgcn_s_sleep(1);
	}
}

} // namespace rsbd8hip

namespace rsbd8hip {

// Calls f(key) for every key of the physical range [lo, lo + len) of p, with the keys of a block
// spread over its threads: the keys before the first 16-byte boundary one per thread, then 16-byte
// vector loads (eight in flight per thread), then the rest one per thread. The order in which a
// thread sees its keys is unspecified (counting only). p must be aligned to the key size.
template <typename K, int THREADS, typename F>
__device__ __forceinline__ void for_each_key(const K *__restrict__ p, uint64_t lo, uint32_t len, F &&f)
{
	constexpr uint32_t VEC = 16 / sizeof(K);
	constexpr int INFLIGHT = 8;
	const K *q = p + lo;
	const uint32_t mis = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(q) & 15u) / static_cast<uint32_t>(sizeof(K));
	uint32_t head = mis ? VEC - mis : 0u;
	if (head > len) head = len;
	if (threadIdx.x < head) f(q[threadIdx.x]);
	const K *qa = q + head;
	const uint32_t body = len - head;
	const uint32_t nvec = body / VEC;
	const uint4 *qv = reinterpret_cast<const uint4 *>(qa);
	for (uint32_t v = threadIdx.x; v < nvec; v += THREADS * INFLIGHT) {
		uint4 u[INFLIGHT];
#pragma unroll
		for (int jj = 0; jj < INFLIGHT; ++jj) {
			const uint32_t idx = v + jj * THREADS;
			if (idx < nvec) u[jj] = qv[idx];
		}
#pragma unroll
		for (int jj = 0; jj < INFLIGHT; ++jj) {
			if (v + jj * THREADS >= nvec) break;
			K ks[VEC];
			__builtin_memcpy(ks, &u[jj], 16);
#pragma unroll
			for (uint32_t e = 0; e < VEC; ++e) f(ks[e]);
		}
	}
	for (uint32_t i = nvec * VEC + threadIdx.x; i < body; i += THREADS) f(qa[i]);
}

// bit 0 for +0.0, bit 1 for -0.0 (the raw key has no bit set below its top one), else 0
template <typename K>
__device__ __forceinline__ uint32_t zero_sign_bit(K raw)
{
	constexpr int B = 8 * static_cast<int>(sizeof(K));
	return (raw << 1) == K(0) ? (1u << static_cast<uint32_t>(raw >> (B - 1))) : 0u;
}

// Counts the splitter bucket id of every key of the physical range [lo, lo + len) of p into s_t
// (C counter copies), reading as for_each_key does; the ids of all the vectors a thread holds are
// searched before any is counted, so that the searches' LDS loads overlap instead of each waiting
// behind the previous key's counter update. With ZCHK, returns the zero_sign_bit of the thread's
// keys ORed together (else 0). SUB as for split_id.
template <typename K, int THREADS, int C, bool ZCHK = false, bool SUB = true>
__device__ __forceinline__ uint32_t count_split_ids(const K *__restrict__ p, uint64_t lo, uint32_t len, const KeyXform &xf,
	const SplitLds<typename KeyTraits<K>::W> &tb, uint32_t *s_t)
{
	using W = typename KeyTraits<K>::W;
	constexpr uint32_t VEC = 16 / sizeof(K);
	constexpr int INFLIGHT = ZCHK ? 4 : 8;
	const K *q = p + lo;
	uint32_t zf = 0;
	const uint32_t mis = static_cast<uint32_t>(reinterpret_cast<uintptr_t>(q) & 15u) / static_cast<uint32_t>(sizeof(K));
	uint32_t head = mis ? VEC - mis : 0u;
	if (head > len) head = len;
	if (threadIdx.x < head) {
		wave_count<C>(s_t, split_id<W, SUB>(tb, sort_key<K>(q[threadIdx.x], xf)));
		if constexpr (ZCHK) zf |= zero_sign_bit<K>(q[threadIdx.x]);
	}
	const K *qa = q + head;
	const uint32_t body = len - head;
	const uint32_t nvec = body / VEC;
	const uint4 *qv = reinterpret_cast<const uint4 *>(qa);
	for (uint32_t v = threadIdx.x; v < nvec; v += THREADS * INFLIGHT) {
		uint4 u[INFLIGHT];
#pragma unroll
		for (int jj = 0; jj < INFLIGHT; ++jj) {
			const uint32_t idx = v + jj * THREADS;
			u[jj] = idx < nvec ? qv[idx] : make_uint4(0, 0, 0, 0);
		}
		if constexpr (ZCHK) {
#pragma unroll
			for (int jj = 0; jj < INFLIGHT; ++jj) {
				K ks[VEC];
				__builtin_memcpy(ks, &u[jj], 16);
				uint32_t vz = 0;
#pragma unroll
				for (uint32_t e = 0; e < VEC; ++e) vz |= zero_sign_bit<K>(ks[e]);
				zf |= v + jj * THREADS < nvec ? vz : 0u;  // (the vectors past nvec are zero-filled)
			}
		}
		uint32_t id[INFLIGHT][VEC];
#pragma unroll
		for (int jj = 0; jj < INFLIGHT; ++jj) {
			K ks[VEC];
			__builtin_memcpy(ks, &u[jj], 16);
#pragma unroll
			for (uint32_t e = 0; e < VEC; ++e) id[jj][e] = split_id<W, SUB>(tb, sort_key<K>(ks[e], xf));
		}
#pragma unroll
		for (int jj = 0; jj < INFLIGHT; ++jj) {
			if (v + jj * THREADS >= nvec) break;
#pragma unroll
			for (uint32_t e = 0; e < VEC; ++e) wave_count<C>(s_t, id[jj][e]);
		}
	}
	for (uint32_t i = nvec * VEC + threadIdx.x; i < body; i += THREADS) {
		wave_count<C>(s_t, split_id<W, SUB>(tb, sort_key<K>(qa[i], xf)));
		if constexpr (ZCHK) zf |= zero_sign_bit<K>(qa[i]);
	}
	return zf;
}

} // namespace rsbd8hip
