/*
 * rsbd8_oracle.c -- CPU restatement of the Radixsortbidi8 (rsbd8) stable LSD radix sort.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product, loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product.
 *
 * Pinning (what checks each function; all run in tests/test_oracle.py):
 *   rsbd8o_decode       -- the reference's mode table read as text (H:59778-59904) restated as an
 *                          independent Python table, plus the known-answer tests below.
 *   rsbd8o_sortkeys     -- equality with rsbd8o_sortkey element by element.
 *   rsbd8o_sortkey      -- brute force over all 2^8 and 2^16 patterns against the ordering
 *                          documented at H:224-248 and H:315-319 (IEEE order by float value for
 *                          forcefloatingp, |x| order for the absolute modes, the listed tiered and
 *                          inside-out sequences) and against numpy's float ordering.
 *   rsbd8o_radixsort    -- multiplatform/main.cpp:10-19, 182-215 known answers; equality with
 *                          rsbd8o_mergesort (a different algorithm) on random and adversarial
 *                          inputs; numpy.argsort(kind="stable") on transformed keys; permutation,
 *                          monotonicity and stability invariants.
 *   rsbd8o_mergesort    -- the same known answers and numpy.argsort(kind="stable").
 *   rsbd8o_sort_indirect, rsbd8o_sort_segmented, rsbd8o_histogram8, rsbd8o_radixsort_mt --
 *                          equality with rsbd8o_radixsort / numpy on the same inputs.
 *   rsbd8o_sortkey_wide, rsbd8o_radixsort_wide (keys wider than 64 bits; tests/test_oracle_wide.py) --
 *                          the known answers of multiplatform/main.cpp:21-178 (long double in 10-,
 *                          12- and 16-byte records, the 128-bit test records, ascending and
 *                          dscrevorder, direct and through pointers); numpy.longdouble ordering of
 *                          canonical 80-bit patterns (zeros of both signs, subnormals, normals,
 *                          infinities) in every float mode; Python integer ordering of 128-bit
 *                          patterns in every integer mode; padding bytes never reaching the key.
 *
 * Citations: H = Radixsortbidi8.hpp in the reference repository.
 */
#include "rsbd8_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- mode decoding: H:59778-59904 ---------------------------------------------------------- */

unsigned rsbd8o_decode(int keytype, unsigned mode, unsigned direction, int indirect)
{
	int issigned = keytype >= RSBD8O_I8;           /* std::is_signed_v<T>: true for floats too */
	int isfloat = keytype >= RSBD8O_F32;           /* std::is_floating_point_v<T> */
	unsigned abs, sign, fltp, flags;
	if (keytype >= RSBD8O_LD80) {                   /* the 80-bit layouts are long double (H:2849-2856) */
		isfloat = keytype <= RSBD8O_LD128;
		issigned = isfloat || keytype == RSBD8O_I128;
	}
	if (mode >= 64u) {                              /* native, nativeabs, nativetieredabs */
		abs = mode >= 65u && issigned;          /* H:59779-59780 */
		sign = mode <= 65u && issigned;         /* H:59809-59811 */
		fltp = isfloat != 0;                    /* H:59840-59842 */
	} else {                                        /* explicit modes: bits 0/1/2 of the value */
		abs = mode & 1u;                        /* H:59804 */
		sign = (mode >> 1) & 1u;                /* H:59835 */
		fltp = (mode >> 2) & 1u;                /* H:59865 */
	}
	flags = abs * RSBD8O_ABS | sign * RSBD8O_SIGN | fltp * RSBD8O_FLTP;
	if (direction & 1u) flags |= RSBD8O_DESC;       /* isdescsort H:59867-59870 */
	if (direction & 2u) {
		/* isrevorderwithindirection H:59872-59875; without indirection the tie reversal is only
		 * kept where distinct bit patterns share a key (H:59877-59904) */
		if (indirect || (abs && sign) || (!abs && sign && fltp)) flags |= RSBD8O_REV;
	}
	return flags;
}

/* ---- closed-form keys ---------------------------------------------------------------------- */

static inline uint64_t wmask(unsigned w) { return w >= 64u ? ~(uint64_t)0 : (((uint64_t)1 << w) - 1u); }

uint64_t rsbd8o_sortkey(unsigned w, unsigned flags, uint64_t x)
{
	uint64_t mask = wmask(w), msb = (uint64_t)1 << (w - 1u);
	uint64_t s, smask, m, k;
	x &= mask;
	s = x >> (w - 1u);
	smask = (0u - s) & mask;
	m = x & ~msb;
	switch (flags & 7u) {
	case 0u:                                   /* forceunsigned: no filter */
		k = x; break;
	case RSBD8O_SIGN:                          /* forcesigned: virtual MSB flip H:11649-11680 */
		k = x ^ msb; break;
	case RSBD8O_ABS | RSBD8O_SIGN:             /* forceabssigned: (x ^ q) - q, H:10006-10022 */
		k = ((x ^ smask) - smask) & mask; break;
	case RSBD8O_ABS:                           /* forcetieredabssigned: (x << 1) ^ q, H:10009-10022 */
		k = ((x << 1) ^ smask) & mask; break;
	case RSBD8O_SIGN | RSBD8O_FLTP:            /* forcefloatingp: s ? -m : m, then MSB flip */
		k = (((m ^ smask) - smask) & mask) ^ msb; break;
	case RSBD8O_ABS | RSBD8O_SIGN | RSBD8O_FLTP: /* forceabsfloatingp: m, H:10024 */
		k = m; break;
	case RSBD8O_FLTP:                          /* specialfloatingp: m ^ q, H:3376-3396 */
		k = m ^ smask; break;
	default:                                   /* forcetieredabsfloatingp: rotl 1, LSB flip H:10025, H:11682-11701 */
		k = (((x << 1) | s) & mask) ^ 1u; break;
	}
	return k;
}

/* ---- the reference algorithm --------------------------------------------------------------- */

static uint64_t load_key(const uint8_t *p, unsigned bytes);

void rsbd8o_sortkeys(unsigned w, unsigned flags, size_t n, const void *in, uint64_t *out)
{
	const unsigned bytes = w / 8u;
	for (size_t i = 0; i < n; ++i) out[i] = rsbd8o_sortkey(w, flags, load_key((const uint8_t *)in + i * bytes, bytes));
}

/* digit geometry on 64-bit hosts: typeradix H:1848, offsetsloopcount H:1864,
 * typeradixremainder H:1869 */
static unsigned geometry(unsigned w, unsigned *shifts, unsigned *bits)
{
	unsigned radix = (w == 32u || w == 64u) ? 11u : 8u, np = 0u, sh = 0u;
	while (sh < w) {
		shifts[np] = sh;
		bits[np] = (w - sh < radix) ? (w - sh) : radix;
		sh += bits[np];
		++np;
	}
	return np;
}

static inline uint64_t load_key(const uint8_t *p, unsigned bytes)
{
	switch (bytes) {
	case 1u: return *p;
	case 2u: { uint16_t v; memcpy(&v, p, 2); return v; }
	case 4u: { uint32_t v; memcpy(&v, p, 4); return v; }
	default: { uint64_t v; memcpy(&v, p, 8); return v; }
	}
}

static inline void store_key(uint8_t *p, unsigned bytes, uint64_t v)
{
	switch (bytes) {
	case 1u: *p = (uint8_t)v; break;
	case 2u: { uint16_t t = (uint16_t)v; memcpy(p, &t, 2); } break;
	case 4u: { uint32_t t = (uint32_t)v; memcpy(p, &t, 4); } break;
	default: memcpy(p, &v, 8); break;
	}
}

/* counts -> forward offsets (first slot of a bucket) and backward offsets (last slot of a bucket),
 * walking buckets high-to-low for a descending sort, as generateoffsetshelpernotshared does
 * (H:11626-11809). Returns 1 when one bucket holds all n items (the pass-skip flag "b"). */
static int make_offsets(size_t n, unsigned nb, int desc, const size_t *count, size_t *fwd, size_t *bwd)
{
	size_t off = 0;
	int uniform = 0;
	for (unsigned j = 0; j < nb; ++j) {
		unsigned d = desc ? nb - 1u - j : j;
		if (count[d] == n) uniform = 1;
		fwd[d] = off;
		off += count[d];
		bwd[d] = off - 1u; /* wraps for empty buckets, never used then */
	}
	return uniform;
}

int rsbd8o_radixsort(unsigned w, unsigned flags, size_t n, const void *in, void *out,
	const uint32_t *pin, uint32_t *pout)
{
	unsigned shifts[8], bits[8], np, bytes = w / 8u;
	int desc = (flags & RSBD8O_DESC) != 0, rev = (flags & RSBD8O_REV) != 0, executed = 0;
	uint8_t *a = NULL, *b = NULL;
	uint32_t *pa = NULL, *pb = NULL;
	size_t *counts = NULL, *fwd = NULL, *bwd = NULL;
	int withpayload = pout != NULL;
	if (w != 8u && w != 16u && w != 32u && w != 64u) return -1;
	if (!n) return 0;
	np = geometry(w, shifts, bits);
	a = (uint8_t *)malloc(n * bytes);
	b = (uint8_t *)malloc(n * bytes);
	counts = (size_t *)calloc((size_t)np << 11, sizeof(size_t));
	fwd = (size_t *)malloc(sizeof(size_t) << 11);
	bwd = (size_t *)malloc(sizeof(size_t) << 11);
	if (withpayload) {
		pa = (uint32_t *)malloc(n * 4u);
		pb = (uint32_t *)malloc(n * 4u);
	}
	if (!a || !b || !counts || !fwd || !bwd || (withpayload && (!pa || !pb))) {
		executed = -1;
		goto done;
	}
	/* counting pass: one read of every key, all digit histograms at once, writing the logical
	 * (reversed under isrevorder) copy as the reference's count loops do (H:32107-32473) */
	for (size_t i = 0; i < n; ++i) {
		size_t src = rev ? n - 1u - i : i;
		uint64_t x = load_key((const uint8_t *)in + src * bytes, bytes);
		uint64_t k = rsbd8o_sortkey(w, flags, x);
		store_key(a + i * bytes, bytes, x);
		if (withpayload) pa[i] = pin ? pin[src] : (uint32_t)src;
		for (unsigned p = 0; p < np; ++p) ++counts[((size_t)p << 11) + ((k >> shifts[p]) & wmask(bits[p]))];
	}
	if (w == 8u && !withpayload && (flags & 7u) != (RSBD8O_ABS | RSBD8O_SIGN) &&
		(flags & 7u) != (RSBD8O_SIGN | RSBD8O_FLTP) && (flags & 7u) != (RSBD8O_ABS | RSBD8O_SIGN | RSBD8O_FLTP)) {
		/* 8-bit bijective modes: the count fill of radixsortnoallocsinglesimplesortmain
		 * (H:48769-48957), one run of the unique bit pattern per bucket */
		uint8_t inv[256];
		size_t pos = 0;
		for (unsigned x = 0; x < 256u; ++x) inv[rsbd8o_sortkey(8u, flags, x)] = (uint8_t)x;
		for (unsigned j = 0; j < 256u; ++j) {
			unsigned d = desc ? 255u - j : j;
			memset((uint8_t *)out + pos, inv[d], counts[d]);
			pos += counts[d];
		}
		executed = 1;
		goto done;
	}
	/* scatter passes, low digit first; a pass whose digit is uniform is skipped (H:11652) */
	for (unsigned p = 0; p < np; ++p) {
		size_t nb = (size_t)1 << bits[p];
		size_t *cnt = counts + ((size_t)p << 11);
		size_t lo, hi;
		if (make_offsets(n, (unsigned)nb, desc, cnt, fwd, bwd)) continue;
		/* bidirectional fill: the low half walks forward from the bucket starts and the high half
		 * walks backward from the bucket ends, the two-ended scatter of the main and companion
		 * sort loops (H:31823-31851 and H:31688-31706); items of one bucket keep source order */
		for (lo = 0, hi = n; lo < hi;) {
			uint64_t x = load_key(a + lo * bytes, bytes);
			size_t d = (rsbd8o_sortkey(w, flags, x) >> shifts[p]) & wmask(bits[p]);
			size_t dst = fwd[d]++;
			store_key(b + dst * bytes, bytes, x);
			if (withpayload) pb[dst] = pa[lo];
			++lo;
			if (lo < hi) {
				--hi;
				x = load_key(a + hi * bytes, bytes);
				d = (rsbd8o_sortkey(w, flags, x) >> shifts[p]) & wmask(bits[p]);
				dst = bwd[d]--;
				store_key(b + dst * bytes, bytes, x);
				if (withpayload) pb[dst] = pa[hi];
			}
		}
		{ uint8_t *t = a; a = b; b = t; }
		if (withpayload) { uint32_t *t = pa; pa = pb; pb = t; }
		++executed;
	}
	memcpy(out, a, n * bytes);
	if (withpayload) memcpy(pout, pa, n * 4u);
done:
	free(a); free(b); free(pa); free(pb); free(counts); free(fwd); free(bwd);
	return executed;
}

/* ---- independent formulation: stable merge sort on (key, logical index) --------------------- */

struct msctx { const uint64_t *key; int desc; };

static int before(const struct msctx *c, uint32_t i, uint32_t j)
{
	/* item i precedes item j when its key is strictly smaller (larger when descending);
	 * equal keys keep logical order, which the caller arranges as ascending i */
	uint64_t ki = c->key[i], kj = c->key[j];
	if (ki != kj) return c->desc ? ki > kj : ki < kj;
	return i < j;
}

static void msort(const struct msctx *c, uint32_t *v, uint32_t *tmp, size_t n)
{
	size_t h, i, j, k;
	if (n < 2u) return;
	h = n / 2u;
	msort(c, v, tmp, h);
	msort(c, v + h, tmp, n - h);
	for (i = 0, j = h, k = 0; i < h && j < n;) tmp[k++] = before(c, v[j], v[i]) ? v[j++] : v[i++];
	while (i < h) tmp[k++] = v[i++];
	while (j < n) tmp[k++] = v[j++];
	memcpy(v, tmp, n * sizeof(uint32_t));
}

int rsbd8o_mergesort(unsigned w, unsigned flags, size_t n, const void *in, void *out, uint32_t *pout)
{
	unsigned bytes = w / 8u;
	uint64_t *key;
	uint32_t *v, *tmp;
	struct msctx c;
	if (w != 8u && w != 16u && w != 32u && w != 64u) return -1;
	if (!n) return 0;
	key = (uint64_t *)malloc(n * sizeof(uint64_t));
	v = (uint32_t *)malloc(n * sizeof(uint32_t));
	tmp = (uint32_t *)malloc(n * sizeof(uint32_t));
	if (!key || !v || !tmp) { free(key); free(v); free(tmp); return -1; }
	/* logical index i names source element rev ? n-1-i : i (H:252-275: isrevorder reverses the
	 * order of equal items) */
	for (size_t i = 0; i < n; ++i) {
		size_t src = (flags & RSBD8O_REV) ? n - 1u - i : i;
		key[i] = rsbd8o_sortkey(w, flags, load_key((const uint8_t *)in + src * bytes, bytes));
		v[i] = (uint32_t)i;
	}
	c.key = key;
	c.desc = (flags & RSBD8O_DESC) != 0;
	msort(&c, v, tmp, n);
	for (size_t i = 0; i < n; ++i) {
		size_t src = (flags & RSBD8O_REV) ? n - 1u - v[i] : v[i];
		if (out) store_key((uint8_t *)out + i * bytes, bytes, load_key((const uint8_t *)in + src * bytes, bytes));
		if (pout) pout[i] = (uint32_t)src;
	}
	free(key); free(v); free(tmp);
	return 0;
}

/* ---- indirection, segments, histograms ----------------------------------------------------- */

int rsbd8o_sort_indirect(unsigned w, unsigned flags, size_t n, const uint8_t *base, size_t stride,
	size_t offset, const uint32_t *idx, uint32_t *idx_out)
{
	unsigned bytes = w / 8u;
	uint8_t *keys, *sorted;
	int r;
	if (!n) return 0;
	keys = (uint8_t *)malloc(n * bytes);
	sorted = (uint8_t *)malloc(n * bytes);
	if (!keys || !sorted) { free(keys); free(sorted); return -1; }
	/* every pass of the reference re-reads key = *(T *)(p + indirection1) (H:2745-2821,
	 * H:39765-39768); one read gives the same keys */
	for (size_t i = 0; i < n; ++i) memcpy(keys + i * bytes, base + (size_t)idx[i] * stride + offset, bytes);
	r = rsbd8o_radixsort(w, flags, n, keys, sorted, idx, idx_out);
	free(keys); free(sorted);
	return r < 0 ? r : 0;
}

int rsbd8o_sort_segmented(unsigned w, unsigned flags, size_t nseg, const uint64_t *seg, const void *in, void *out)
{
	unsigned bytes = w / 8u;
	for (size_t s = 0; s < nseg; ++s) {
		size_t b = (size_t)seg[s], e = (size_t)seg[s + 1u];
		if (e < b) return -1;
		if (rsbd8o_radixsort(w, flags, e - b, (const uint8_t *)in + b * bytes, (uint8_t *)out + b * bytes, NULL, NULL) < 0)
			return -1;
	}
	return 0;
}

void rsbd8o_histogram8(unsigned w, unsigned flags, size_t n, const void *in, uint64_t *counts)
{
	unsigned bytes = w / 8u, np = w / 8u;
	uint64_t flip = (flags & RSBD8O_DESC) ? wmask(w) : 0u;
	memset(counts, 0, (size_t)np * 256u * sizeof(uint64_t));
	for (size_t i = 0; i < n; ++i) {
		uint64_t k = rsbd8o_sortkey(w, flags, load_key((const uint8_t *)in + i * bytes, bytes)) ^ flip;
		for (unsigned p = 0; p < np; ++p) ++counts[p * 256u + ((k >> (8u * p)) & 255u)];
	}
}

/* ---- multithreaded variant (the timed CPU baseline) ------------------------------------------ */

struct mtjob {
	unsigned w, flags, np, t, nt;
	unsigned shifts[8], bits[8];
	size_t n;
	const uint8_t *src;
	size_t *counts;   /* [nt][np][2048] */
	size_t *offsets;  /* [nt][2048] */
	int rev, executed;
	pthread_barrier_t *bar;
	uint8_t *bufa, *bufb, *out;
};

static void *mt_worker(void *arg)
{
	struct mtjob *j = (struct mtjob *)arg;
	unsigned bytes = j->w / 8u, np = j->np;
	size_t n = j->n, lo = n * j->t / j->nt, hi = n * (j->t + 1u) / j->nt;
	size_t *cnt = j->counts + (size_t)j->t * np * 2048u;
	size_t *off = j->offsets + (size_t)j->t * 2048u;
	uint8_t *a = j->bufa, *b = j->bufb;
	int desc = (j->flags & RSBD8O_DESC) != 0, uniform[8];
	/* sliced counting pass (radixsortnoallocmultiinitmt H:28919): each thread counts its own
	 * slice into its own histograms and writes its part of the logical copy */
	for (size_t i = lo; i < hi; ++i) {
		size_t s = j->rev ? n - 1u - i : i;
		uint64_t x = load_key(j->src + s * bytes, bytes);
		uint64_t k = rsbd8o_sortkey(j->w, j->flags, x);
		store_key(a + i * bytes, bytes, x);
		for (unsigned p = 0; p < np; ++p) ++cnt[(size_t)p * 2048u + ((k >> j->shifts[p]) & wmask(j->bits[p]))];
	}
	pthread_barrier_wait(j->bar);
	j->executed = 0;
	/* the pass-skip flags come from the merged counts of the counting pass (H:11652); they are
	 * all taken before any thread recounts a digit below */
	for (unsigned p = 0; p < np; ++p) {
		uniform[p] = 0;
		for (size_t d = 0; d < ((size_t)1 << j->bits[p]); ++d) {
			size_t tot = 0;
			for (unsigned t = 0; t < j->nt; ++t) tot += j->counts[((size_t)t * np + p) * 2048u + d];
			if (tot == n) uniform[p] = 1;
		}
	}
	for (unsigned p = 0; p < np; ++p) {
		size_t nb = (size_t)1 << j->bits[p], base = 0;
		if (uniform[p]) continue;
		if (j->executed) {
			/* after a scatter the slice holds other keys: count this digit again */
			memset(cnt + (size_t)p * 2048u, 0, 2048u * sizeof(size_t));
			for (size_t i = lo; i < hi; ++i) {
				uint64_t k = rsbd8o_sortkey(j->w, j->flags, load_key(a + i * bytes, bytes));
				++cnt[(size_t)p * 2048u + ((k >> j->shifts[p]) & wmask(j->bits[p]))];
			}
			pthread_barrier_wait(j->bar);
		}
		/* per-thread bucket starts: every thread's items of earlier buckets, then the items of
		 * earlier slices in the same bucket, which keeps the scatter stable across slices */
		for (size_t q = 0; q < nb; ++q) {
			size_t d = desc ? nb - 1u - q : q, tot = 0;
			for (unsigned t = 0; t < j->nt; ++t) {
				size_t c = j->counts[((size_t)t * np + p) * 2048u + d];
				if (t == j->t) off[d] = base + tot;
				tot += c;
			}
			base += tot;
		}
		for (size_t i = lo; i < hi; ++i) {
			uint64_t x = load_key(a + i * bytes, bytes);
			size_t d = (rsbd8o_sortkey(j->w, j->flags, x) >> j->shifts[p]) & wmask(j->bits[p]);
			store_key(b + (off[d]++) * bytes, bytes, x);
		}
		{ uint8_t *t = a; a = b; b = t; }
		++j->executed;
		pthread_barrier_wait(j->bar);
	}
	memcpy(j->out + lo * bytes, a + lo * bytes, (hi - lo) * bytes);
	return NULL;
}

int rsbd8o_radixsort_mt(unsigned w, unsigned flags, size_t n, const void *in, void *out, unsigned threads)
{
	unsigned shifts[8], bits[8], np, bytes = w / 8u, nt = threads ? threads : 1u;
	struct mtjob *jobs;
	pthread_t *tid;
	pthread_barrier_t bar;
	size_t *counts, *offsets;
	uint8_t *a, *b;
	int executed = 0;
	if (w == 8u || !n || nt == 1u || n < (size_t)nt * 4096u)
		return rsbd8o_radixsort(w, flags, n, in, out, NULL, NULL);
	if (w != 16u && w != 32u && w != 64u) return -1;
	np = geometry(w, shifts, bits);
	jobs = (struct mtjob *)calloc(nt, sizeof(*jobs));
	tid = (pthread_t *)calloc(nt, sizeof(*tid));
	counts = (size_t *)calloc((size_t)nt * np * 2048u, sizeof(size_t));
	offsets = (size_t *)calloc((size_t)nt * 2048u, sizeof(size_t));
	a = (uint8_t *)malloc(n * bytes);
	b = (uint8_t *)malloc(n * bytes);
	if (!jobs || !tid || !counts || !offsets || !a || !b) { executed = -1; goto done; }
	pthread_barrier_init(&bar, NULL, nt);
	for (unsigned t = 0; t < nt; ++t) {
		struct mtjob *j = &jobs[t];
		j->w = w; j->flags = flags; j->np = np; j->t = t; j->nt = nt; j->n = n;
		memcpy(j->shifts, shifts, sizeof(shifts));
		memcpy(j->bits, bits, sizeof(bits));
		j->src = (const uint8_t *)in; j->counts = counts; j->offsets = offsets;
		j->rev = (flags & RSBD8O_REV) != 0; j->bar = &bar;
		j->bufa = a; j->bufb = b; j->out = (uint8_t *)out;
	}
	for (unsigned t = 1; t < nt; ++t) pthread_create(&tid[t], NULL, mt_worker, &jobs[t]);
	mt_worker(&jobs[0]);
	for (unsigned t = 1; t < nt; ++t) pthread_join(tid[t], NULL);
	pthread_barrier_destroy(&bar);
	executed = jobs[0].executed;
done:
	free(jobs); free(tid); free(counts); free(offsets); free(a); free(b);
	return executed;
}

/* ---- keys wider than 64 bits: H:1061-1804 (layouts), H:7479-7530 and H:9047-9150 (filters) ---- */

typedef unsigned __int128 u128o;

static int wide_layout(int keytype, unsigned *rec_bytes, unsigned *value_bits)
{
	switch (keytype) {
	case RSBD8O_LD80: *rec_bytes = 10u; *value_bits = 80u; return 1;
	case RSBD8O_LD96: *rec_bytes = 12u; *value_bits = 80u; return 1;
	case RSBD8O_LD128: *rec_bytes = 16u; *value_bits = 80u; return 1;
	case RSBD8O_U128: case RSBD8O_I128: *rec_bytes = 16u; *value_bits = 128u; return 1;
	default: return 0;
	}
}

/* the closed forms of rsbd8o_sortkey (Appendix-A table of the narrow keys) for W = 80 or 128 */
static u128o wide_key(unsigned w, unsigned flags, u128o x)
{
	const u128o ones = w >= 128u ? ~(u128o)0 : (((u128o)1 << w) - 1u), msb = (u128o)1 << (w - 1u);
	u128o s, smask, m, k;
	x &= ones;
	s = x >> (w - 1u);
	smask = ((u128o)0 - s) & ones;
	m = x & ~msb;
	switch (flags & 7u) {
	case 0u: k = x; break;
	case RSBD8O_SIGN: k = x ^ msb; break;
	case RSBD8O_ABS | RSBD8O_SIGN: k = ((x ^ smask) - smask) & ones; break;
	case RSBD8O_ABS: k = ((x << 1) ^ smask) & ones; break;
	case RSBD8O_SIGN | RSBD8O_FLTP: k = (((m ^ smask) - smask) & ones) ^ msb; break;
	case RSBD8O_ABS | RSBD8O_SIGN | RSBD8O_FLTP: k = m; break;
	case RSBD8O_FLTP: k = m ^ smask; break;
	default: k = (((x << 1) | s) & ones) ^ 1u; break;
	}
	return (flags & RSBD8O_DESC) ? k ^ ones : k;
}

static u128o wide_value(const uint8_t *p, unsigned value_bits)
{
	uint64_t lo, hi = 0;
	memcpy(&lo, p, 8);
	memcpy(&hi, p + 8, value_bits == 80u ? 2u : 8u);
	return ((u128o)hi << 64) | lo;
}

void rsbd8o_sortkey_wide(int keytype, unsigned flags, const void *rec, uint64_t *lo, uint64_t *hi)
{
	unsigned rb, vb;
	u128o k;
	if (!wide_layout(keytype, &rb, &vb)) { *lo = *hi = 0; return; }
	k = wide_key(vb, flags, wide_value((const uint8_t *)rec, vb));
	*lo = (uint64_t)k;
	*hi = (uint64_t)(k >> 64);
}

struct wsctx {
	const u128o *key;
	int rev;
};

/* i before j in the stable order: smaller key, or an equal key earlier in the logical order */
static int wide_before(const struct wsctx *c, uint32_t i, uint32_t j)
{
	if (c->key[i] != c->key[j]) return c->key[i] < c->key[j];
	return c->rev ? i > j : i < j;
}

static void wide_msort(const struct wsctx *c, uint32_t *v, uint32_t *tmp, size_t n)
{
	size_t h, a, b, o;
	if (n < 2u) return;
	h = n / 2u;
	wide_msort(c, v, tmp, h);
	wide_msort(c, v + h, tmp, n - h);
	a = 0; b = h; o = 0;
	while (a < h && b < n) tmp[o++] = wide_before(c, v[b], v[a]) ? v[b++] : v[a++];
	while (a < h) tmp[o++] = v[a++];
	while (b < n) tmp[o++] = v[b++];
	memcpy(v, tmp, n * sizeof(uint32_t));
}

int rsbd8o_radixsort_wide(int keytype, unsigned flags, size_t n, const void *in, size_t stride, uint32_t *perm)
{
	unsigned rb, vb;
	u128o *key;
	uint32_t *tmp;
	struct wsctx c;
	if (!wide_layout(keytype, &rb, &vb) || stride < rb || n > 0xFFFFFFFFu) return -1;
	key = (u128o *)malloc((n ? n : 1) * sizeof(u128o));
	tmp = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
	if (!key || !tmp) { free(key); free(tmp); return -1; }
	for (size_t i = 0; i < n; ++i) {
		key[i] = wide_key(vb, flags, wide_value((const uint8_t *)in + i * stride, vb));
		perm[i] = (uint32_t)i;
	}
	c.key = key;
	c.rev = (flags & RSBD8O_REV) != 0;
	wide_msort(&c, perm, tmp, n);
	free(key);
	free(tmp);
	return 0;
}
