// rsbd8_sample.hpp -- the sample partition plan of the gfx950 sort (included by rsbd8_sort.hip
// inside its internal namespace, after the kernels it shares).
//
// Two partitions and one per-bucket LDS sort, like the two-level MSD plan, but the buckets come
// from splitters drawn from the keys instead of from the top two bytes of the sort key, so that
// keys whose bytes cluster (floats, whose exponents crowd into a few top digits; skewed integers)
// still fall into buckets an LDS block can hold, and so that keys with a payload and modes whose
// ties must keep their source order take three moves through HBM instead of one per varying byte.
// The reference has no such plan: its scatter passes (Radixsortbidi8.hpp H:31649-31982) visit
// every varying digit; the output is the same stable order, so the plan is exact.
//   k_samp_chunk    one block per input chunk: a stratified sample at seeded positions, sorted
//                   in LDS, reduced to regular samples;
//   k_samp_l1       one block: sorts the regular samples and writes the first partition's
//                   splitter table;
//   k_split_count1  per-tile counts of the first bucket id and its histogram (the first read);
//   k_plan_sample   bucket starts, the segment map of the second partition, the pass plans;
//   k_scan, k_scatter<..., SPL 1>
//                   the first partition (ranked, so ties keep their order, or grouped when
//                   equal sort keys are equal keys);
//   k_samp_l2       one block per first-partition segment: its own splitter table from a sample
//                   of the segment, with as many buckets as its length needs;
//   k_split_count2, k_split_segscan, k_scatter<..., SEG, SPL 2>
//                   the second partition inside every segment;
//   k_bucket_sort   one block per bucket: orders the bucket by (sort key, position), a partition
//                   on the top bits of (key - min) and exact ranks within its runs; equality
//                   buckets are already in order;
//   k_bucket_hard   the buckets k_bucket_sort passes on: stable LSD passes in LDS, or through
//                   HBM above the LDS capacity.
// Splitter tables: a sorted sample, candidates at the ranks (j + 1) * S / f, j < f - 1, and their
// distinct values as bounds. A value that takes two or more candidate slots (at least about one
// bucket's worth of keys) gets an equality bucket of its own, so heavy repeats never overflow a
// bucket; the ids stay below 256 because every such value takes at least two of the 255 slots.

constexpr uint32_t kSampleMax = 4096;      // sample keys per splitter table
constexpr uint32_t kBucketRunMax = 96;     // longest run of one partition value the ranking takes
// With 8-byte keys, buckets of at most kBucketCapA keys take a 512-thread per-bucket sort (three
// blocks of eight waves per CU: 2^28 N(0,1) doubles sort 0.8 ms faster than with 256 threads),
// larger ones the 1024-thread one (8192 keys, two blocks per CU at that width); modelled at 2^28
// keys with the sampling below, about 80% of the buckets fit the smaller one. 4-byte keys take
// the 1024-thread one only (two blocks per CU).
constexpr uint32_t kBucketCapA = 512 * 10;

// seeded position in [lo, lo + width) for sample i (a 64-bit mix of i and the salt)
__device__ __forceinline__ uint64_t sample_pos(uint64_t lo, uint64_t width, uint32_t i, uint32_t salt)
{
	uint64_t h = (static_cast<uint64_t>(i) | (static_cast<uint64_t>(salt) << 32)) * 0x9E3779B97F4A7C15ull;
	h ^= h >> 31;
	h *= 0xD6E8FEB86659FD93ull;
	h ^= h >> 32;
	return lo + (width > 1 ? h % width : 0);
}

// Bitonic sort of s[0, sp) ascending in LDS (sp a power of two); every thread takes part.
template <typename W, int THREADS>
__device__ void lds_bitonic(W *s, uint32_t sp)
{
	for (uint32_t k = 2; k <= sp; k <<= 1)
		for (uint32_t j = k >> 1; j > 0; j >>= 1) {
			for (uint32_t i = threadIdx.x; i < sp / 2; i += THREADS) {
				const uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
				const uint32_t hi = lo + j;
				const bool up = (lo & k) == 0;
				const W a = s[lo], b = s[hi];
				if ((a > b) == up) {
					s[lo] = b;
					s[hi] = a;
				}
			}
			__syncthreads();
		}
}

// LDS scratch of the table builder
template <typename W>
struct SplitScratch {
	W c[kRadix];        // candidates
	W u[kRadix];        // distinct candidate values (the bounds)
	uint32_t x[kRadix]; // per bound: open id, equality flag in bit 8
	uint32_t eq[8];
	uint32_t wsum[4];
	uint32_t nsub;      // sub-tables handed out
};

// Writes the splitter table tab for f buckets from the sorted sample s[0, S) (S >= 1), or, with
// all_equal, a table with one equality bucket (every key of the range is equal). Every thread of
// the block (at least 256) takes part.
template <typename W, int THREADS>
__device__ void build_split_table(const W *s, uint32_t S, uint32_t f, bool all_equal, uint32_t *__restrict__ tab,
	SplitScratch<W> &sc)
{
	static_assert(THREADS >= 256, "one thread per candidate slot");
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t nc = (all_equal || f < 2) ? 0u : (f > kRadix ? kRadix - 1 : f - 1);
	if (t < kRadix) sc.c[t] = t < nc ? s[static_cast<uint32_t>((static_cast<uint64_t>(t + 1) * S) / (nc + 1))] : W(0);
	if (t < 8) sc.eq[t] = 0;
	if (t == 0) sc.nsub = 0;
	__syncthreads();
	uint32_t first = 0, heavy = 0;
	if (t < nc) {
		const W c = sc.c[t];
		first = (t == 0 || sc.c[t - 1] != c) ? 1u : 0u;
		heavy = (first && t + 1 < nc && sc.c[t + 1] == c) ? 1u : 0u;
	}
	// exclusive scan over the slots of (first, heavy) packed as first | heavy << 16
	const uint32_t v = first | (heavy << 16);
	uint32_t incl = 0;
	if (t < kRadix) {
		incl = wave_inclusive_sum(v);
		if (lane == 63) sc.wsum[w] = incl;
	}
	__syncthreads();
	const uint32_t tot = sc.wsum[0] + sc.wsum[1] + sc.wsum[2] + sc.wsum[3];
	const uint32_t m = tot & 0xFFFFu, H = tot >> 16;
	if (t < kRadix && first) {
		uint32_t before = 0;
		for (uint32_t q = 0; q < w; ++q) before += sc.wsum[q];
		const uint32_t ex = before + incl - v;
		const uint32_t ti = ex & 0xFFFFu, open = ti + (ex >> 16);
		sc.u[ti] = sc.c[t];
		sc.x[ti] = open | (heavy << 8);
		if (heavy) atomicOr(&sc.eq[(open + 1) >> 5], 1u << ((open + 1) & 31u));
	}
	__syncthreads();
	if (t < kRadix) {
		uint64_t *bounds = reinterpret_cast<uint64_t *>(tab + kSplitBoundsWord);
		bounds[t] = t < m ? static_cast<uint64_t>(sc.u[t]) : ~0ull;
		// keys above every bound: the last open bucket; info[255] (never a real bound's) also holds m
		tab[kSplitInfoWord + t] = (t < m ? sc.x[t] : m + H) | (t == kRadix - 1 ? m << 16 : 0u);
	}
	if (t < 8) tab[kSplitEqWord + t] = all_equal ? (t == 0 ? 1u : 0u) : sc.eq[t];
	// the cell table over [lo, hi]: kSplitCells cells of 2^sh keys each
	constexpr int WB = 8 * static_cast<int>(sizeof(W));
	const W lo = m ? sc.u[0] : ~W(0), hi = m ? sc.u[m - 1] : ~W(0);
	const W range = hi - lo;
	int bits = 0;
	if (range) bits = WB == 64 ? 64 - __clzll(static_cast<long long>(range)) : 32 - __clz(static_cast<int>(range));
	const uint32_t sh = bits > 12 ? static_cast<uint32_t>(bits - 12) : 0u;
	static_assert(kSplitCells == 4096, "cells cover 12 bits of the key range");
	if (t == 0) {
		tab[kSplitCountWord] = m;
		tab[kSplitIdsWord] = m + 1 + H;
		*reinterpret_cast<uint64_t *>(tab + kSplitLoWord) = static_cast<uint64_t>(lo);
		*reinterpret_cast<uint64_t *>(tab + kSplitHiWord) = static_cast<uint64_t>(hi);
		tab[kSplitShWord] = sh;
	}
	auto below = [&](W v) {  // the number of bounds below v
		uint32_t a = 0, left = m;
		while (left) {
			const uint32_t half = left >> 1;
			if (sc.u[a + half] < v) {
				a += half + 1;
				left -= half + 1;
			} else {
				left = half;
			}
		}
		return a;
	};
	// cells and sub-cells: (bounds below the cell's first key) | (bounds inside it) << 8. A cell with
	// eight bounds or more gets a sub-table of 16 sub-cells, while there are any left and the cells
	// are at least 16 keys wide, and the span field 0xFF with the sub-table's index (a cell that
	// holds all 255 bounds, the one span that could read as 0xFF, is then the only such cell and
	// gets the first). The first table of N(0,1) doubles has 10 such cells (8-41 bounds, a binary
	// search of up to six steps each), that of N(0,1) floats none (at most 7 bounds per cell: the
	// second read a sub-table costs is worth no more than the three steps it saves), that of keys
	// spread over their range none.
	uint16_t *cells = reinterpret_cast<uint16_t *>(tab + kSplitCellWord);
	uint16_t *subs = reinterpret_cast<uint16_t *>(tab + kSplitSubWord);
	auto cell_of = [&](uint64_t off, uint64_t wmask) -> uint32_t {
		if (!m || off > static_cast<uint64_t>(range)) return 0u;
		const uint64_t end = static_cast<uint64_t>(range) - off < wmask ? static_cast<uint64_t>(range) : off + wmask;
		const uint32_t t0 = below(lo + static_cast<W>(off)), t1 = below(lo + static_cast<W>(end));
		return t0 | ((t1 - t0) << 8);
	};
	const uint64_t mask = (1ull << sh) - 1ull;
	for (uint32_t c = t; c < kSplitCells; c += THREADS) {
		const uint64_t off = static_cast<uint64_t>(c) << sh;
		uint32_t e = cell_of(off, mask);
		if ((e >> 8) >= 8u && sh >= 4u) {
			const uint32_t si = atomicAdd(&sc.nsub, 1u);
			if (si < kSplitSubs) {
				for (uint32_t q = 0; q < kSplitSubCells; ++q)
					subs[si * kSplitSubCells + q] =
						static_cast<uint16_t>(cell_of(off + (static_cast<uint64_t>(q) << (sh - 4u)), (mask >> 4)));
				e = si | (0xFFu << 8);
			}
		}
		cells[c] = static_cast<uint16_t>(e);
	}
	__syncthreads();
	if (t == 0) tab[kSplitNsubWord] = min(sc.nsub, kSplitSubs);
}

// Gathers up to SMAX transformed keys of src[lo, lo + len) at seeded stratified positions (all of
// them when len <= SMAX) into s, pads to a power of two with all-ones and sorts them. Returns the
// sample size.
template <typename K, int THREADS, uint32_t SMAX>
__device__ uint32_t sample_sorted(const K *__restrict__ src, uint64_t lo, uint64_t len, const KeyXform &xf, uint32_t salt,
	typename KeyTraits<K>::W *s)
{
	using W = typename KeyTraits<K>::W;
	const uint32_t S = static_cast<uint32_t>(len < SMAX ? len : SMAX);
	uint32_t sp = 1;
	while (sp < S) sp <<= 1;
	for (uint32_t i = threadIdx.x; i < sp; i += THREADS) {
		W v = ~W(0);
		if (i < S) {
			const uint64_t a = static_cast<uint64_t>(i) * len / S, b = static_cast<uint64_t>(i + 1) * len / S;
			v = sort_key<K>(src[len <= SMAX ? lo + i : sample_pos(lo, b - a, i, salt) + a], xf);
		}
		s[i] = v;
	}
	__syncthreads();
	if (sp > 1) lds_bitonic<W, THREADS>(s, sp);
	return S;
}

// The first partition's splitters come from regular sampling: the input is cut into G chunks
// (kL1Chunks, fewer for short inputs); block g sorts a sample of kSampleMax keys of chunk g and
// keeps kSampleMax / G of them at evenly spaced ranks (u64, in the splitter tables of the segments,
// which are free until the second partition); one block then sorts those kSampleMax regular samples
// and takes the splitters from them. Modelled at 2^28 keys: the largest first-partition segment
// is 1.3x the mean, against 1.8x for splitters from one sample of kSampleMax keys.
constexpr uint32_t kL1Chunks = 16;

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_samp_chunk(const K *__restrict__ in, uint64_t n, KeyXform xf, const Control *ctrl,
	int reverse, uint32_t *__restrict__ split, const uint32_t *gate)
{
	using W = typename KeyTraits<K>::W;
	if (gate && !*gate) return;
	if (presort_result(ctrl->check, reverse)) return;  // the input is already in order
	__shared__ W s[kSampleMax];
	const uint32_t G = gridDim.x, g = blockIdx.x, R = kSampleMax / G;
	const uint64_t lo = g * n / G, hi = (g + 1) * n / G;
	const uint32_t S = sample_sorted<K, THREADS, kSampleMax>(in, lo, hi - lo, xf, 0x5A3Du + g, s);
	uint64_t *reg = reinterpret_cast<uint64_t *>(split + kSplitTabWords) + static_cast<size_t>(g) * R;
	for (uint32_t j = threadIdx.x; j < R; j += THREADS)
		reg[j] = S ? static_cast<uint64_t>(s[static_cast<uint32_t>((static_cast<uint64_t>(j) * S) / R + S / (2 * R))]) : ~0ull;
}

// The first partition's splitter table (split[0..kSplitTabWords)) from the G chunks' regular
// samples: kRadix buckets.
template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_samp_l1(const Control *ctrl, int reverse, uint32_t *__restrict__ split,
	const uint32_t *gate)
{
	using W = typename KeyTraits<K>::W;
	if (gate && !*gate) return;
	if (presort_result(ctrl->check, reverse)) return;
	__shared__ W s[kSampleMax];
	__shared__ SplitScratch<W> sc;
	const uint64_t *reg = reinterpret_cast<const uint64_t *>(split + kSplitTabWords);
	for (uint32_t i = threadIdx.x; i < kSampleMax; i += THREADS) s[i] = static_cast<W>(reg[i]);
	__syncthreads();
	lds_bitonic<W, THREADS>(s, kSampleMax);
	build_split_table<W, THREADS>(s, kSampleMax, kRadix, false, split, sc);
}

// The second partition's splitter tables, one block per first-partition segment (its keys in
// hyb[1].src at seg_start[sg], seg_len[sg] of them): ceil(len / target) buckets (at most 256),
// one equality bucket for a segment that is an equality bucket of the first partition, one bucket
// for a segment of at most target keys. The sample holds kSampleMax2 keys (a few buckets of a 2^28
// sort go above twice the target with kSampleMax).
constexpr uint32_t kSampleMax2 = 8192;

template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_samp_l2(const Control *ctrl, KeyXform xf, SegTables T, uint32_t *__restrict__ split,
	uint32_t target)
{
	using W = typename KeyTraits<K>::W;
	const PassPlan pp = ctrl->hyb[1];
	if (!pp.active) return;
	__shared__ W s[kSampleMax2];
	__shared__ SplitScratch<W> sc;
	const uint32_t sg = blockIdx.x;
	const uint32_t len = T.seg_len[sg];
	const bool eq = (split[kSplitEqWord + (sg >> 5)] >> (sg & 31u)) & 1u;
	uint32_t *tab = split + static_cast<size_t>(1 + sg) * kSplitTabWords;
	if (eq || len <= target) {
		build_split_table<W, THREADS>(s, 1, 1, eq, tab, sc);
		return;
	}
	const uint32_t f = static_cast<uint32_t>(min<uint64_t>(kRadix, (static_cast<uint64_t>(len) + target - 1) / target));
	const uint32_t S = sample_sorted<K, THREADS, kSampleMax2>(static_cast<const K *>(pp.src), T.seg_start[sg], len, xf, 0xB17Eu + sg, s);
	build_split_table<W, THREADS>(s, S, f, false, tab, sc);
}

// The first read: per-tile counts of the first bucket id (tiles in the logical order of the first
// partition) and the bucket histogram (hist[0..255], zero on entry); with ZCHK, the signs of the
// zero keys seen into ctrl->zero_signs (zero_sign_bit).
template <typename K, int THREADS, bool ZCHK>
__global__ __launch_bounds__(THREADS) void k_split_count1(const K *__restrict__ in, uint64_t n, KeyXform xf,
	uint32_t *__restrict__ hist, uint32_t *__restrict__ counts, uint32_t tile_keys, uint32_t num_tiles, int reverse,
	Control *ctrl, const uint32_t *__restrict__ split, const uint32_t *gate)
{
	static_assert(THREADS == kRadix, "one thread per bucket id");
	using W = typename KeyTraits<K>::W;
	if (gate && !*gate) return;
	if (presort_result(ctrl->check, reverse)) return;
	constexpr int C = count_copies(1);
	__shared__ uint32_t s_t[kRadix * C];
	__shared__ SplitLds<W> s_tab;
	split_load<W, THREADS>(split, s_tab);
	const uint32_t t = threadIdx.x;
	uint32_t mine = 0;  // this block's count of bucket id t
	uint32_t zf = 0;
	for (uint32_t tile = blockIdx.x; tile < num_tiles; tile += gridDim.x) {
		for (int i = t; i < kRadix * C; i += THREADS) s_t[i] = 0;
		__syncthreads();
		const uint64_t base = static_cast<uint64_t>(tile) * tile_keys;
		const uint32_t len = (n - base) < tile_keys ? static_cast<uint32_t>(n - base) : tile_keys;
		const uint64_t lo = reverse ? n - base - len : base;
		if (__builtin_amdgcn_readfirstlane(s_tab.nsub))
			zf |= count_split_ids<K, THREADS, C, ZCHK, true>(in, lo, len, xf, s_tab, s_t);
		else
			zf |= count_split_ids<K, THREADS, C, ZCHK, false>(in, lo, len, xf, s_tab, s_t);
		__syncthreads();
		const uint32_t v = copies_sum<C>(s_t, t);
		counts[static_cast<uint64_t>(tile) * kCountStride + t] = v;
		mine += v;
		__syncthreads();
	}
	if (mine) atomicAdd(&hist[t], mine);
	if constexpr (ZCHK) {
		const uint32_t wz = (__ballot((zf & 1u) != 0) ? 1u : 0u) | (__ballot((zf & 2u) != 0) ? 2u : 0u);
		if ((threadIdx.x & 63u) == 0 && wz) atomicOr(&ctrl->zero_signs, wz);
	}
}

// The second count: per-tile counts of the bucket id under the tile's segment table (split +
// segment * kSplitTabWords) over the tiles of the segment map, into the count rows.
template <typename K, int THREADS>
__global__ __launch_bounds__(THREADS) void k_split_count2(const PassPlan *ppp, KeyXform xf, uint32_t tile_keys, SegMap seg,
	uint32_t *__restrict__ counts, const uint32_t *__restrict__ split)
{
	using W = typename KeyTraits<K>::W;
	const PassPlan pp = *ppp;
	if (!pp.active) return;
	const K *__restrict__ src = static_cast<const K *>(pp.src);
	constexpr int C = count_copies(1);
	__shared__ uint32_t s_ts[kRadix + 1];
	__shared__ uint32_t s_t[kRadix * C];
	__shared__ SplitLds<W> s_tab;
	for (int i = threadIdx.x; i <= kRadix; i += THREADS) s_ts[i] = seg.tile_start[i];
	__syncthreads();
	const uint32_t num = s_ts[kRadix];
	uint32_t tab_seg = 0xFFFFFFFFu;
	for (uint32_t tile = blockIdx.x; tile < num; tile += gridDim.x) {
		for (int i = threadIdx.x; i < kRadix * C; i += THREADS) s_t[i] = 0;
		uint64_t base;
		uint32_t len, sg;
		seg_tile(s_ts, seg, tile_keys, tile, base, len, sg);
		if (sg != tab_seg) {  // uniform across the block; the last reads of the old table ended at a barrier
			split_load<W, THREADS>(split + static_cast<size_t>(sg) * kSplitTabWords, s_tab);
			tab_seg = sg;
		}
		__syncthreads();
		if (__builtin_amdgcn_readfirstlane(s_tab.nsub))
			count_split_ids<K, THREADS, C, false, true>(src, base, len, xf, s_tab, s_t);
		else
			count_split_ids<K, THREADS, C, false, false>(src, base, len, xf, s_tab, s_t);
		__syncthreads();
		for (int i = threadIdx.x; i < kRadix; i += THREADS) counts[static_cast<uint64_t>(tile) * kCountStride + i] = copies_sum<C>(s_t, i);
		__syncthreads();
	}
}

// One block per first-partition segment, one thread per bucket id: the bucket totals, their
// exclusive starts inside the segment (the bucket table), the tile counts turned into global output
// offsets, and every non-empty bucket queued for a per-bucket sort as the entry {absolute start,
// length, bucket, equal keys} (equal: the bucket is an equality bucket of the segment's table at
// split_l2 + segment * kSplitTabWords): those of at most qcap keys that are not equality buckets in
// T.bq[0, ...) (ctrl->bucket_count of them), the others in T.bq[65536, ...) (ctrl->bucket_count2);
// one atomic per block and queue.
__global__ __launch_bounds__(256) void k_split_segscan(Control *ctrl, SegTables T, uint32_t *__restrict__ counts,
	const uint32_t *__restrict__ split_l2, uint32_t qcap)
{
	if (!ctrl->hyb[1].active) return;
	__shared__ uint32_t s_wave[4];
	__shared__ uint32_t s_qw[2][4], s_qbase[2];
	const uint32_t sg = blockIdx.x, d = threadIdx.x, lane = d & 63u, w = d >> 6;
	const uint32_t t0 = T.tile_start[sg], t1 = T.tile_start[sg + 1];
	const uint32_t eq = (split_l2[static_cast<size_t>(sg) * kSplitTabWords + kSplitEqWord + (d >> 5)] >> (d & 31u)) & 1u;
	uint32_t tot = 0;
	for (uint32_t tl = t0; tl < t1; tl += 8) {
		uint32_t c[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) c[j] = (tl + j < t1) ? counts[static_cast<uint64_t>(tl + j) * kCountStride + d] : 0u;
#pragma unroll
		for (int j = 0; j < 8; ++j) tot += c[j];
	}
	const uint32_t incl = wave_inclusive_sum(tot);
	if (lane == 63) s_wave[w] = incl;
	const int q = (tot > qcap || eq) ? 1 : 0;
	const uint64_t m0 = __ballot(tot != 0 && q == 0), m1 = __ballot(tot != 0 && q == 1);
	if (lane == 0) {
		s_qw[0][w] = static_cast<uint32_t>(__popcll(m0));
		s_qw[1][w] = static_cast<uint32_t>(__popcll(m1));
	}
	__syncthreads();
	if (d < 2) {
		const uint32_t all = s_qw[d][0] + s_qw[d][1] + s_qw[d][2] + s_qw[d][3];
		s_qbase[d] = all ? atomicAdd(d ? &ctrl->bucket_count2 : &ctrl->bucket_count, all) : 0u;
	}
	uint32_t before = 0;
	for (uint32_t v = 0; v < w; ++v) before += s_wave[v];
	uint32_t run = T.seg_start[sg] + before + incl - tot;
	T.bucket_start[sg * kRadix + d] = before + incl - tot;
	T.bucket_len[sg * kRadix + d] = tot;
	__syncthreads();
	uint32_t at = s_qbase[q] + lanes_below(q ? m1 : m0);
	for (uint32_t v = 0; v < w; ++v) at += s_qw[q][v];
	if (tot) T.bq[(q ? kHybBuckets : 0u) + at] = make_uint4(run, tot, sg * kRadix + d, eq);
	for (uint32_t tl = t0; tl < t1; tl += 8) {
		uint32_t c[8];
#pragma unroll
		for (int j = 0; j < 8; ++j) c[j] = (tl + j < t1) ? counts[static_cast<uint64_t>(tl + j) * kCountStride + d] : 0u;
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			if (tl + j < t1) counts[static_cast<uint64_t>(tl + j) * kCountStride + d] = run;
			run += c[j];
		}
	}
}

// One block: the first partition's bucket starts (bucket[0..255]) and segment map, hist[0..255]
// zeroed for the next sort, and the pass plans: hyb[0] input -> X1, hyb[1] X1 -> X2, hyb[2] X2 ->
// the output, the payload arrays alongside. An input already in order becomes the final copy.
// ctrl->group_parts selects the partitions' form (see PlanArgs::group).
__global__ __launch_bounds__(256) void k_plan_sample(PlanArgs a, Control *ctrl, uint32_t *hist, uint32_t *bucket)
{
	__shared__ uint32_t s_wave[4], s_tw[4];
	if (a.gate && !*a.gate) return;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const uint32_t c = hist[t];
	hist[t] = 0;
	// an equality bucket of the first table is final after the first partition (which writes it to
	// the final array): its segment has no tiles and no length at the second level
	// (not when the final arrays are the source: an in-place sort reads them while it partitions)
	const bool direct = a.fin != a.src && !(a.payload && a.psrc && a.psrc == a.pfin);
	const bool eq = direct && a.split0 && ((a.split0[kSplitEqWord + (t >> 5)] >> (t & 31u)) & 1u);
	const uint32_t tiles = eq ? 0u : (c + a.tile2 - 1) / a.tile2;
	const uint32_t incl = wave_inclusive_sum(c), tincl = wave_inclusive_sum(tiles);
	if (lane == 63) {
		s_wave[w] = incl;
		s_tw[w] = tincl;
	}
	__syncthreads();
	uint32_t before = 0, tbefore = 0;
	for (uint32_t v = 0; v < w; ++v) {
		before += s_wave[v];
		tbefore += s_tw[v];
	}
	const SegTables T = seg_tables(a.segtab);
	const uint32_t start = before + incl - c;
	bucket[t] = start;
	T.tile_start[t] = tbefore + tincl - tiles;
	if (t == kRadix - 1) T.tile_start[kRadix] = tbefore + tincl;
	T.seg_start[t] = start;
	T.seg_len[t] = eq ? 0u : c;
	if (t != 0) return;
	PassPlan none{};
	for (int h = 0; h < 3; ++h) ctrl->hyb[h] = none;
	for (int p = 0; p < kMaxPasses; ++p) ctrl->pass[p] = none;
	ctrl->pre = none;
	ctrl->fin = none;
	ctrl->counts_ready = 0;
	ctrl->active_mask = 0;
	ctrl->num_active = 0;
	ctrl->group_parts = (a.group == 1 || (a.group == 2 && ctrl->zero_signs != 3u)) ? 1u : 0u;
	const uint32_t pres = presort_result(ctrl->check, a.reverse);
	if (pres == 3u) return;  // the run reversal (k_runs_copy) writes the result
	if (pres) {
		PassPlan &f = ctrl->fin;
		f.src = a.src; f.dst = a.fin; f.psrc = a.psrc; f.pdst = a.pfin;
		f.iota = a.payload && !a.psrc;
		f.reverse = (pres == 2u);
		f.active = (a.src != a.fin) || f.reverse || (a.payload && (f.iota || a.psrc != a.pfin));
		if (spec_done(ctrl, pres, a.fin, a.payload)) f.active = 0;  // the check already copied it
		return;
	}
	void *X1 = a.other != a.src ? a.other : a.fin;
	void *X2 = X1 == a.other ? a.fin : a.other;
	uint32_t *PX1 = nullptr, *PX2 = nullptr;
	if (a.payload) {
		PX1 = X1 == a.other ? a.pother : a.pfin;
		PX2 = X1 == a.other ? a.pfin : a.pother;
	}
	PassPlan h{};
	h.active = 1;
	h.src = a.src; h.dst = X1; h.psrc = a.payload ? a.psrc : nullptr; h.pdst = PX1;
	h.reverse = a.reverse;
	h.iota = a.payload && !a.psrc;
	ctrl->hyb[0] = h;
	h.src = X1; h.dst = X2; h.psrc = PX1; h.pdst = PX2; h.reverse = 0; h.iota = 0;
	ctrl->hyb[1] = h;
	h.src = X2; h.dst = a.fin; h.psrc = PX2; h.pdst = a.payload ? a.pfin : nullptr;
	ctrl->hyb[2] = h;
}

// Copies [0, len) of src to dst (and the payload) with the threads of the block, unless in place.
template <typename K, bool PAY, int THREADS>
__device__ __forceinline__ void block_copy(const K *src, K *dst, const uint32_t *psrc, uint32_t *pdst, uint32_t len)
{
	if (src != dst)
		for (uint32_t i = threadIdx.x; i < len; i += THREADS) dst[i] = src[i];
	if constexpr (PAY) {
		if (psrc != pdst)
			for (uint32_t i = threadIdx.x; i < len; i += THREADS) pdst[i] = psrc[i];
	}
}

// Stable LSD passes through HBM for one bucket of len keys (more than the block holds): the digits of
// (sort key - mn) below bit hb + 1, CAP keys at a time ranked in the block, between the bucket's range
// of a (its keys now) and of b (free), then the result copied to dst if it did not end there.
template <typename K, bool PAY, int THREADS, int KPT>
__device__ void bucket_lsd_global(K *a, K *b, K *dst, uint32_t *pa, uint32_t *pb, uint32_t *pdst, uint32_t len,
	const KeyXform &xf, typename KeyTraits<K>::W mn, int hb, uint32_t *s_whist, uint32_t *s_digoff, uint32_t *s_wsum,
	uint32_t *s_run, uint64_t *s_peer)
{
	constexpr int CAP = THREADS * KPT;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	const int passes = hb / 8 + 1;
	for (int p = 0; p < passes; ++p) {
		const uint32_t sh = 8u * static_cast<uint32_t>(p);
		// the digit counts of the whole bucket, then their exclusive offsets
		for (uint32_t i = t; i < kRadix; i += THREADS) s_run[i] = 0;
		__syncthreads();
		for (uint32_t i = t; i < len; i += THREADS)
			atomicAdd(&s_run[static_cast<uint32_t>((sort_key<K>(a[i], xf) - mn) >> sh) & 0xFFu], 1u);
		__syncthreads();
		uint32_t c = 0, incl = 0;
		if (t < kRadix) {  // whole waves
			c = s_run[t];
			incl = wave_inclusive_sum(c);
			if (lane == 63) s_wsum[w] = incl;
		}
		__syncthreads();
		if (t < kRadix) {
			uint32_t before = 0;
			for (uint32_t v = 0; v < w; ++v) before += s_wsum[v];
			s_run[t] = before + incl - c;
		}
		__syncthreads();
		for (uint32_t c0 = 0; c0 < len; c0 += CAP) {
			const uint32_t valid = min(static_cast<uint32_t>(CAP), len - c0);
			K key[KPT];
			uint32_t pay[PAY ? KPT : 1], dig[KPT], pos[KPT], total = 0;
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				key[j] = l < valid ? a[c0 + l] : K(0);
				if constexpr (PAY) pay[j] = l < valid ? pa[c0 + l] : 0u;
				dig[j] = l < valid ? static_cast<uint32_t>((sort_key<K>(key[j], xf) - mn) >> sh) & 0xFFu : 0xFFu;
			}
			block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				if (l < valid) {
					const uint32_t o = s_run[dig[j]] + pos[j] - s_digoff[dig[j]];
					b[o] = key[j];
					if constexpr (PAY) pb[o] = pay[j];
				}
			}
			__syncthreads();
			if (t < kRadix) s_run[t] += total;
			__syncthreads();
		}
		K *tk = a; a = b; b = tk;
		uint32_t *tp = pa; pa = pb; pb = tp;
	}
	if (a != dst) block_copy<K, false, THREADS>(a, dst, nullptr, nullptr, len);
	if constexpr (PAY) {
		if (pa != pdst) block_copy<uint32_t, false, THREADS>(pa, pdst, nullptr, nullptr, len);
	}
}

// The per-bucket sort of the sample partition plan: one block per bucket of its queue (sg = entry
// >> 8, id = entry & 255; blocks past the queue's *queued entries exit at once). A bucket's keys (and payload)
// are read whole into registers before anything is written, so in place is fine. Each key becomes
// the word ((sort key - min) << LB) | l, l < CAP = 2^LB its position in the bucket: distinct words
// whose order is the stable order. They are partitioned by the top PB bits of the key range with
// LDS atomics; every word's final position is its run's start plus the smaller words of its run.
// The raw keys and the payload reach their places through their owners. Equality buckets are
// copied in order. A bucket too long for the block, whose key range leaves no room for the
// position in a word, or with a run of more than kBucketRunMax keys is queued for k_bucket_hard.
template <typename K, bool PAY, int THREADS, int KPT, int PB>
__global__ __launch_bounds__(THREADS) void k_bucket_sort(Control *ctrl, KeyXform xf, SegTables T,
	const uint4 *__restrict__ queue, const uint32_t *queued)
{
	using W = typename KeyTraits<K>::W;
	constexpr uint32_t CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	constexpr int WB = 8 * static_cast<int>(sizeof(W));
	constexpr int LB = CAP <= 4096 ? 12 : (CAP <= 8192 ? 13 : (CAP <= 16384 ? 14 : 16));
	constexpr uint32_t NB = 1u << PB;
	constexpr uint32_t PER = NB / THREADS;
	static_assert(CAP <= (1u << LB) && CAP < 65536, "positions and run starts are held in 16 bits");
	static_assert(NB % THREADS == 0 && PB <= 16, "partition buckets per thread");
	// s_u: the NB cursors (after the scatter: the run ends) and, above them, the block's minimum and
	// maximum, wave sums and overflow flag; then the final positions (u16) over all of it. Nothing
	// else in LDS, so that two 1024x8 blocks of 8-byte keys share a CU (80 KiB each).
	constexpr uint32_t kSmallWords = NW * (2 * sizeof(W) / 4 + 1) + 1;
	constexpr uint32_t kUnionWords = NB + kSmallWords > CAP / 2 ? NB + kSmallWords : CAP / 2;
	__shared__ __align__(16) W s_t[CAP];        // words, then the raw keys and the payload in final order
	__shared__ __align__(16) uint32_t s_u[kUnionWords];
	W *s_min = reinterpret_cast<W *>(s_u + NB), *s_max = s_min + NW;
	uint32_t *s_wsum = reinterpret_cast<uint32_t *>(s_max + NW), *s_big = s_wsum + NW;
	const PassPlan pp = ctrl->hyb[2];
	if (!pp.active || blockIdx.x >= *queued) return;
	const uint4 entry = queue[blockIdx.x];
	const uint32_t start = entry.x, len = entry.y;
	const K *src = static_cast<const K *>(pp.src) + start;
	K *dst = static_cast<K *>(pp.dst) + start;
	const uint32_t *psrc = PAY ? pp.psrc + start : nullptr;
	uint32_t *pdst = PAY ? pp.pdst + start : nullptr;
	if (len == 1 || entry.w) {
		// equal keys: the partitions kept them in order
		block_copy<K, PAY, THREADS>(src, dst, psrc, pdst, len);
		return;
	}
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	if (len > CAP) {
		if (t == 0) T.bq[2 * kHybBuckets + atomicAdd(&ctrl->hyb_queue, 1u)] = entry;
		return;
	}
	K key[KPT];
	uint32_t pay[PAY ? KPT : 1];
	W mn = ~W(0), mx = 0;
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		key[j] = l < len ? src[l] : K(0);
		if constexpr (PAY) pay[j] = l < len ? psrc[l] : 0u;
		if (l < len) {
			const W k = sort_key<K>(key[j], xf);
			mn = min(mn, k);
			mx = max(mx, k);
		}
	}
	if (t == 0) *s_big = 0;
	for (uint32_t i = t; i < NB; i += THREADS) s_u[i] = 0;
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		mn = min(mn, static_cast<W>(__shfl_xor(mn, o, 64)));
		mx = max(mx, static_cast<W>(__shfl_xor(mx, o, 64)));
	}
	if (lane == 0) {
		s_min[w] = mn;
		s_max[w] = mx;
	}
	__syncthreads();
	mn = s_min[0];
	mx = s_max[0];
#pragma unroll
	for (int v = 1; v < NW; ++v) {
		mn = min(mn, s_min[v]);
		mx = max(mx, s_max[v]);
	}
	const W diff = mx - mn;
	if (diff == 0) {
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) {
				dst[l] = key[j];
				if constexpr (PAY) pdst[l] = pay[j];
			}
		}
		return;
	}
	const int hb = sizeof(W) == 8 ? 63 - __clzll(static_cast<long long>(diff)) : 31 - __clz(static_cast<int>(diff));
	if (hb >= WB - LB) {  // (key - min) and the position do not fit one word
		if (t == 0) T.bq[2 * kHybBuckets + atomicAdd(&ctrl->hyb_queue, 1u)] = entry;
		return;
	}
	const uint32_t sh = hb + 1 >= PB ? static_cast<uint32_t>(hb + 1 - PB) : 0u;
	uint32_t *s_cur = s_u;
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		if (l < len) atomicAdd(&s_cur[static_cast<uint32_t>((sort_key<K>(key[j], xf) - mn) >> sh)], 1u);
	}
	__syncthreads();
	uint32_t c[PER], sum = 0, big = 0;
#pragma unroll
	for (uint32_t k = 0; k < PER; ++k) {
		c[k] = s_cur[t * PER + k];
		sum += c[k];
		big |= c[k] > kBucketRunMax;
	}
	const uint32_t incl = wave_inclusive_sum(sum);
	if (lane == 63) s_wsum[w] = incl;
	if (big) *s_big = 1;
	__syncthreads();
	if (*s_big) {
		if (t == 0) T.bq[2 * kHybBuckets + atomicAdd(&ctrl->hyb_queue, 1u)] = entry;
		return;
	}
	uint32_t run = incl - sum;
	for (uint32_t v = 0; v < w; ++v) run += s_wsum[v];
#pragma unroll
	for (uint32_t k = 0; k < PER; ++k) {
		s_cur[t * PER + k] = run;
		run += c[k];
	}
	__syncthreads();
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		if (l < len) {
			const W word = static_cast<W>((sort_key<K>(key[j], xf) - mn) << LB) | static_cast<W>(l);
			s_t[atomicAdd(&s_cur[static_cast<uint32_t>((word >> LB) >> sh)], 1u)] = word;
		}
	}
	__syncthreads();
	// the final place of every word: its run's start (the previous run's end) plus the smaller words
	// of its run
	uint32_t fin[KPT];
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t i = j * THREADS + t;
		fin[j] = 0xFFFFFFFFu;
		if (i < len) {
			const W x = s_t[i];
			const uint32_t d = static_cast<uint32_t>((x >> LB) >> sh);
			const uint32_t lo = d ? s_cur[d - 1] : 0u, hi = s_cur[d];
			uint32_t n_less = 0;
			for (uint32_t q = lo; q < hi; ++q) n_less += static_cast<uint32_t>(s_t[q] < x);
			fin[j] = ((static_cast<uint32_t>(x) & ((1u << LB) - 1u)) << 16) | (lo + n_less);
		}
	}
	__syncthreads();  // every run has been read: s_u takes the final positions
	uint16_t *s_fin = reinterpret_cast<uint16_t *>(s_u);
#pragma unroll
	for (int j = 0; j < KPT; ++j)
		if (fin[j] != 0xFFFFFFFFu) s_fin[fin[j] >> 16] = static_cast<uint16_t>(fin[j]);
	__syncthreads();
	K *s_out = reinterpret_cast<K *>(s_t);
#pragma unroll
	for (int j = 0; j < KPT; ++j) {
		const uint32_t l = w * 64u * KPT + j * 64u + lane;
		if (l < len) s_out[s_fin[l]] = key[j];
	}
	__syncthreads();
	for (uint32_t i = t; i < len; i += THREADS) dst[i] = s_out[i];
	if constexpr (PAY) {
		__syncthreads();
		uint32_t *s_pout = reinterpret_cast<uint32_t *>(s_t);
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) s_pout[s_fin[l]] = pay[j];
		}
		__syncthreads();
		for (uint32_t i = t; i < len; i += THREADS) pdst[i] = s_pout[i];
	}
}

// The buckets k_bucket_sort passed on (T.bq[131072 + i], i < ctrl->hyb_queue), persistent blocks: a bucket
// that fits the block takes stable LDS LSD passes over the digits of (sort key - min) that vary
// (block ranking, payload alongside); a longer one takes the same passes through HBM, CAP keys at
// a time, between its range and the range of the other partition array (free since the second
// partition read it) or of the destination.
template <typename K, bool PAY, int THREADS, int KPT>
__global__ __launch_bounds__(THREADS) void k_bucket_hard(const Control *ctrl, KeyXform xf, SegTables T)
{
	using W = typename KeyTraits<K>::W;
	constexpr int CAP = THREADS * KPT;
	constexpr int NW = THREADS / 64;
	static_assert(CAP * sizeof(K) >= rank_peer_bytes(THREADS), "the peer words share the staging array");
	__shared__ __align__(16) K s_k[CAP];        // peer words, then staged keys or payload
	__shared__ uint32_t s_whist[NW * kRadix];
	__shared__ uint32_t s_digoff[kRadix], s_run[kRadix];
	__shared__ uint32_t s_wsum[NW > 4 ? NW : 4];
	__shared__ W s_min[NW], s_max[NW];
	const PassPlan pp = ctrl->hyb[2], p1 = ctrl->hyb[1];
	if (!pp.active) return;
	const uint32_t queued = ctrl->hyb_queue;
	const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
	uint64_t *s_peer = reinterpret_cast<uint64_t *>(s_k);
	uint32_t *s_p = reinterpret_cast<uint32_t *>(s_k);
	for (uint32_t qi = blockIdx.x; qi < queued; qi += gridDim.x) {
		__syncthreads();  // the previous bucket's last reads of the shared arrays are done
		const uint4 entry = T.bq[2 * kHybBuckets + qi];
		const uint32_t start = entry.x, len = entry.y;
		const K *src = static_cast<const K *>(pp.src) + start;
		K *dst = static_cast<K *>(pp.dst) + start;
		const uint32_t *psrc = PAY ? pp.psrc + start : nullptr;
		uint32_t *pdst = PAY ? pp.pdst + start : nullptr;
		W mn = ~W(0), mx = 0;
		for (uint32_t i = t; i < len; i += THREADS) {
			const W k = sort_key<K>(src[i], xf);
			mn = min(mn, k);
			mx = max(mx, k);
		}
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) {
			mn = min(mn, static_cast<W>(__shfl_xor(mn, o, 64)));
			mx = max(mx, static_cast<W>(__shfl_xor(mx, o, 64)));
		}
		if (lane == 0) {
			s_min[w] = mn;
			s_max[w] = mx;
		}
		__syncthreads();
		mn = s_min[0];
		mx = s_max[0];
		for (int v = 1; v < NW; ++v) {
			mn = min(mn, s_min[v]);
			mx = max(mx, s_max[v]);
		}
		if (mn == mx) {
			block_copy<K, PAY, THREADS>(src, dst, psrc, pdst, len);
			continue;
		}
		const W diff = mx - mn;
		const int hb = sizeof(W) == 8 ? 63 - __clzll(static_cast<long long>(diff)) : 31 - __clz(static_cast<int>(diff));
		if (len > static_cast<uint32_t>(CAP)) {
			K *a = const_cast<K *>(src);
			K *b = pp.src != pp.dst ? dst : static_cast<K *>(const_cast<void *>(p1.src)) + start;
			uint32_t *pa = const_cast<uint32_t *>(psrc);
			uint32_t *pb = nullptr;
			if constexpr (PAY) pb = pp.psrc != pp.pdst ? pdst : const_cast<uint32_t *>(p1.psrc) + start;
			bucket_lsd_global<K, PAY, THREADS, KPT>(a, b, dst, pa, pb, pdst, len, xf, mn, hb, s_whist, s_digoff, s_wsum, s_run,
				s_peer);
			continue;
		}
		K key[KPT];
		uint32_t pay[PAY ? KPT : 1];
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			key[j] = l < len ? src[l] : K(0);
			if constexpr (PAY) pay[j] = l < len ? psrc[l] : 0u;
		}
		__syncthreads();  // every thread has read the bucket (in place: the writes below may overlap it)
		const int passes = hb / 8 + 1;
		for (int p = 0; p < passes; ++p) {
			uint32_t dig[KPT], pos[KPT], total;
#pragma unroll
			for (int j = 0; j < KPT; ++j) {
				const uint32_t l = w * 64u * KPT + j * 64u + lane;
				dig[j] = l < len ? static_cast<uint32_t>((sort_key<K>(key[j], xf) - mn) >> (8 * p)) & 0xFFu : 0xFFu;
			}
			block_rank<THREADS, KPT>(dig, pos, s_whist, s_digoff, s_wsum, &total, s_peer);
#pragma unroll
			for (int j = 0; j < KPT; ++j) s_k[pos[j]] = key[j];
			__syncthreads();
#pragma unroll
			for (int j = 0; j < KPT; ++j) key[j] = s_k[w * 64u * KPT + j * 64u + lane];
			if constexpr (PAY) {
				__syncthreads();
#pragma unroll
				for (int j = 0; j < KPT; ++j) s_p[pos[j]] = pay[j];
				__syncthreads();
#pragma unroll
				for (int j = 0; j < KPT; ++j) pay[j] = s_p[w * 64u * KPT + j * 64u + lane];
			}
			__syncthreads();
		}
#pragma unroll
		for (int j = 0; j < KPT; ++j) {
			const uint32_t l = w * 64u * KPT + j * 64u + lane;
			if (l < len) {
				dst[l] = key[j];
				if constexpr (PAY) pdst[l] = pay[j];
			}
		}
	}
}
