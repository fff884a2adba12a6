/*
 * rsbd8_oracle.h -- CPU restatement of the Radixsortbidi8 (rsbd8) stable LSD radix sort.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product in
 * radixsortbidi8_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it. Nothing in the product links it, and the product never falls back to it.
 *
 * Citations are file:line into the reference repository (Radixsortbidi8.hpp = "H").
 *
 * Flag bits (the five compile-time booleans the reference derives from (mode, T, direction),
 * H:59778-59904):
 *   RSBD8O_ABS  = isabsvalue     RSBD8O_SIGN = issignmode     RSBD8O_FLTP = isfltpmode
 *   RSBD8O_DESC = isdescsort     RSBD8O_REV  = isrevorder (already filtered through
 *                                isrevorderwithoutindirection / isrevorderwithindirection)
 */
#ifndef RSBD8_ORACLE_H
#define RSBD8_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	RSBD8O_ABS = 1u,
	RSBD8O_SIGN = 2u,
	RSBD8O_FLTP = 4u,
	RSBD8O_DESC = 8u,
	RSBD8O_REV = 16u
};

/* key type codes, shared with the product's C-ABI (include/rsbd8_hip.h) */
enum {
	RSBD8O_U8 = 0, RSBD8O_U16 = 1, RSBD8O_U32 = 2, RSBD8O_U64 = 3,
	RSBD8O_I8 = 4, RSBD8O_I16 = 5, RSBD8O_I32 = 6, RSBD8O_I64 = 7,
	RSBD8O_F32 = 8, RSBD8O_F64 = 9,
	/* wider than 64 bits: the 80-bit extended float in 10-, 12- and 16-byte records
	 * (longdoubletest80/96/128, H:1061-1804) and 128-bit integers (test128, H:2186-2404) */
	RSBD8O_LD80 = 10, RSBD8O_LD96 = 11, RSBD8O_LD128 = 12, RSBD8O_U128 = 13, RSBD8O_I128 = 14
};

/* mode decoding of H:59778-59904 for a plain key type code (enums pass their unsigned
 * carrier, as std::is_signed_v is false for enums). indirect selects
 * isrevorderwithindirection (H:59872) over isrevorderwithoutindirection (H:59877). */
unsigned rsbd8o_decode(int keytype, unsigned mode, unsigned direction, int indirect);

/* closed-form sortable key of a W-bit pattern x (W in {8,16,32,64}); the descending bit is
 * not applied. Follows filterinput H:10000-10027 plus the virtual top-bit flip of the offsets
 * transform H:11649-11701. */
uint64_t rsbd8o_sortkey(unsigned width_bits, unsigned flags, uint64_t x);

/* rsbd8o_sortkey over an array of n width_bits/8-byte keys, into n 64-bit words. */
void rsbd8o_sortkeys(unsigned width_bits, unsigned flags, size_t n, const void *in, uint64_t *out);

/* The reference's single-threaded algorithm: one counting pass over all digits with the
 * reference's digit geometry (H:1848-1869), per-digit bidirectional offsets with pass skip
 * (H:11626-11809), and stable bidirectional scatter passes (H:31649-31982). keys: width_bits/8
 * bytes each. pin may be NULL (then the payload is the source index when pout != NULL).
 * Returns the number of executed scatter passes, or -1 on bad arguments/allocation failure. */
int rsbd8o_radixsort(unsigned width_bits, unsigned flags, size_t n, const void *in, void *out,
	const uint32_t *pin, uint32_t *pout);

/* Independent formulation used to pin rsbd8o_radixsort: stable top-down merge sort on
 * (sortkey, logical index) with the same direction rules (H:252-275). */
int rsbd8o_mergesort(unsigned width_bits, unsigned flags, size_t n, const void *in, void *out,
	uint32_t *pout);

/* Indirection through records: key = *(T*)(base + idx[i]*stride + offset) (the type+offset
 * wrapper H:61360-61393 with pointers base + idx*stride). Writes the sorted indices. */
int rsbd8o_sort_indirect(unsigned width_bits, unsigned flags, size_t n, const uint8_t *base,
	size_t stride, size_t offset, const uint32_t *idx, uint32_t *idx_out);

/* Independent segmented sort: every segment [seg[s], seg[s+1]) sorted on its own with
 * rsbd8o_radixsort (the reference runs one radixsortcopy per segment, H:60468). */
int rsbd8o_sort_segmented(unsigned width_bits, unsigned flags, size_t nseg, const uint64_t *seg,
	const void *in, void *out);

/* 8-bit-digit histograms (digit d of pass p = (sortkey >> 8p) & 255, descending applied as a
 * complement), counts[p*256 + d]; used to check a device histogram. */
void rsbd8o_histogram8(unsigned width_bits, unsigned flags, size_t n, const void *in, uint64_t *counts);

/* Keys wider than 64 bits. The value of a record is its first 80 bits (the 80-bit layouts: 64-bit
 * mantissa, then sign and exponent in the next 16 bits; padding beyond them is ignored, as the
 * reference's filters read only signexponent's low 16 bits, H:7479-7530) or its 128 bits (low
 * word first, H:9047-9150), and its sort key the closed form of rsbd8o_sortkey with W = 80 or 128
 * (sign at bit W-1). rsbd8o_sortkey_wide writes the key's low and high 64-bit words, descending
 * applied as a complement of the W bits. */
void rsbd8o_sortkey_wide(int keytype, unsigned flags, const void *rec, uint64_t *lo, uint64_t *hi);

/* Stable sort of n wide-key records (stride bytes apart) by their sort keys, ties in source order
 * (reversed under RSBD8O_REV): perm[i] = source index of the i-th record of the result. A merge
 * sort on the 128-bit keys, independent of the device's two-pass method. */
int rsbd8o_radixsort_wide(int keytype, unsigned flags, size_t n, const void *in, size_t stride, uint32_t *perm);

/* Multithreaded variant of rsbd8o_radixsort (pthreads; counting and scatter sliced across
 * threads with per-thread offsets, as the reference's split counting H:28919 plus slice
 * scatters); used as the timed CPU baseline. */
int rsbd8o_radixsort_mt(unsigned width_bits, unsigned flags, size_t n, const void *in, void *out,
	unsigned threads);

#ifdef __cplusplus
}
#endif
#endif
