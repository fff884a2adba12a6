/*
 * rsbd8_hip.h -- C ABI of the MI355X (gfx950) rsbd8 radix sort.
 *
 * Every entry point is a drop-in for one of the reference's sorting templates (Radixsortbidi8.hpp,
 * cited as H:line) with the compile-time template arguments turned into run-time values:
 *   - the key type T becomes a key type code (an enum passes its unsigned underlying type, as
 *     the reference sorts enums by their unsigned bit pattern, H:2893 and H:59779);
 *   - sortingmode/sortingdirection become the five decoded flags of H:59778-59904
 *     (rsbd8_hip_decode performs that decoding from the raw enum values);
 *   - allowedthreads and pfuturesplaceholder have no counterpart (the GPU grid replaces the CPU
 *     threads), and pzeroedindices becomes a device workspace of rsbd8_hip_workspace_size bytes
 *     that must be zero before its first use (rsbd8_hip_workspace_init) and is left ready for
 *     reuse by every sort.
 * Array arguments are device pointers (hipMalloc'd, or managed memory) unless stated: every entry
 * point that launches kernels on the caller's arrays checks them on the host first and returns
 * RSBD8_HIP_EINVAL, before any launch, for an array that is host memory (pageable or pinned). Only
 * rsbd8_hip_radixsort and rsbd8_hip_radixsortcopy take host arrays (they stage them).
 * All sorts are stream-ordered and asynchronous unless stated; calls on one workspace must not
 * overlap. Return value: 0 on success, a negative RSBD8_HIP_E* code otherwise
 * (rsbd8_hip_last_error gives the message).
 */
#ifndef RSBD8_HIP_H
#define RSBD8_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *rsbd8_hip_stream; /* a hipStream_t; NULL is the null stream */

enum rsbd8_hip_keytype {
	RSBD8_HIP_U8 = 0, RSBD8_HIP_U16 = 1, RSBD8_HIP_U32 = 2, RSBD8_HIP_U64 = 3,
	RSBD8_HIP_I8 = 4, RSBD8_HIP_I16 = 5, RSBD8_HIP_I32 = 6, RSBD8_HIP_I64 = 7,
	RSBD8_HIP_F32 = 8, RSBD8_HIP_F64 = 9,
	/* Keys wider than 64 bits (H:1061-1804, H:2186-2404), little-endian records: the 80-bit
	 * extended float (64-bit mantissa, then 16 bits of sign and exponent) in its packed 10-byte
	 * layout (longdoubletest80), its 12-byte layout (longdoubletest96) and its 16-byte layout
	 * (longdoubletest128, x86-64 long double), whose padding bytes are moved but never compared;
	 * and 128-bit integers (test128, unsigned __int128 / __int128), low word first. They decode as
	 * long double (signed, floating point) and as the 128-bit integers. Taken by the noalloc,
	 * allocating, sort_pairs and _indirect entry points. */
	RSBD8_HIP_LD80 = 10, RSBD8_HIP_LD96 = 11, RSBD8_HIP_LD128 = 12, RSBD8_HIP_U128 = 13, RSBD8_HIP_I128 = 14
};

/* decoded flags: isabsvalue, issignmode, isfltpmode, isdescsort, isrevorder */
enum {
	RSBD8_HIP_ABS = 1, RSBD8_HIP_SIGN = 2, RSBD8_HIP_FLTP = 4, RSBD8_HIP_DESC = 8, RSBD8_HIP_REV = 16
};

/* raw enum values of rsbd8::sortingmode (H:212-250) and rsbd8::sortingdirection (H:252-265) */
enum {
	RSBD8_HIP_MODE_NATIVE = 64, RSBD8_HIP_MODE_NATIVEABS = 65, RSBD8_HIP_MODE_NATIVETIEREDABS = 66,
	RSBD8_HIP_MODE_FORCEUNSIGNED = 0, RSBD8_HIP_MODE_FORCESIGNED = 2, RSBD8_HIP_MODE_FORCEABSSIGNED = 3,
	RSBD8_HIP_MODE_FORCEFLOATINGP = 6, RSBD8_HIP_MODE_FORCEABSFLOATINGP = 7,
	RSBD8_HIP_MODE_SPECIALFLOATINGP = 4, RSBD8_HIP_MODE_FORCETIEREDABSFLOATINGP = 5,
	RSBD8_HIP_MODE_FORCETIEREDABSSIGNED = 1,
	RSBD8_HIP_ASCFWDORDER = 0, RSBD8_HIP_DSCFWDORDER = 1, RSBD8_HIP_ASCREVORDER = 2, RSBD8_HIP_DSCREVORDER = 3
};

/* workspace layouts for rsbd8_hip_workspace_size */
enum { RSBD8_HIP_LAYOUT_KEYS = 0, RSBD8_HIP_LAYOUT_PAIRS = 1, RSBD8_HIP_LAYOUT_INDIRECT = 2 };

enum {
	RSBD8_HIP_OK = 0, RSBD8_HIP_EINVAL = -1, RSBD8_HIP_ENOMEM = -2, RSBD8_HIP_EDEVICE = -3
};

/* helper::isabsvalue, issignmode, isfltpmode (H:59778-59865), isdescsort (H:59867-59870),
 * isrevorderwithindirection (H:59872-59875) when indirect, else isrevorderwithoutindirection
 * (H:59877-59904). mode and direction are the raw enum values. */
int rsbd8_hip_decode(int keytype, unsigned mode, unsigned direction, int indirect, unsigned *flags);

/* The pzeroedindices equivalent (allocatearray H:59921-60065 sizes it on the CPU side). */
size_t rsbd8_hip_workspace_size(int keytype, size_t count, int layout);
int rsbd8_hip_workspace_init(void *workspace, size_t workspace_bytes, rsbd8_hip_stream stream);

/* rsbd8::radixsortcopynoalloc without indirection: multi-part H:60121-60153, single-part
 * H:60186-60242 (8-bit keys: buffer may be NULL, as the reference's dummy-buffer overload
 * ignores it). input is not written. */
int rsbd8_hip_radixsortcopynoalloc(int keytype, unsigned flags, size_t count, void *workspace,
	size_t workspace_bytes, const void *input, void *output, void *buffer, rsbd8_hip_stream stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): rsbd8_hip_radixsortcopynoalloc
 * without the workspace argument. Its workspace is a file-scope device array of the library, one
 * per GPU, sized for RSBD8_HIP_STATIC_MAX_COUNT keys (33.4 MB, held by every GPU the library
 * runs on until the process ends); a larger count returns RSBD8_HIP_EINVAL. The call allocates
 * nothing (the first call on a GPU looks the array's address up once). input, output and buffer
 * are the caller's device arrays as in rsbd8_hip_radixsortcopynoalloc. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams, other host threads) race
 * and must be serialized by the caller. */
#define RSBD8_HIP_STATIC_MAX_COUNT (1u << 24)
int rsbd8_hip_radixsortcopynoalloc_static(int keytype, unsigned flags, size_t count, const void *input, void *output,
	void *buffer, rsbd8_hip_stream stream);

/* rsbd8::radixsortnoalloc without indirection: multi-part H:60155-60184, single-part
 * H:60244-60395. The result lands in input, or in buffer when movetobuffer is set; the other
 * array holds garbage afterwards (H:188). buffer may be NULL for 8-bit keys in the bijective
 * modes, as in the reference's buffer-less single-part overload. */
int rsbd8_hip_radixsortnoalloc(int keytype, unsigned flags, size_t count, void *workspace,
	size_t workspace_bytes, void *input, void *buffer, int movetobuffer, rsbd8_hip_stream stream);

/* rsbd8::radixsortcopy without indirection (H:60468-60527, H:60529-60588): allocates and frees
 * its own workspace and buffer; input/output may be host or device memory (host arrays are
 * staged through device memory). Synchronous. Returns RSBD8_HIP_ENOMEM where the reference
 * returns false (H:60463). */
int rsbd8_hip_radixsortcopy(int keytype, unsigned flags, size_t count, const void *input, void *output,
	rsbd8_hip_stream stream);

/* rsbd8::radixsort without indirection (H:60397-60466), in place; host or device memory.
 * Synchronous. */
int rsbd8_hip_radixsort(int keytype, unsigned flags, size_t count, void *input, rsbd8_hip_stream stream);

/* Indirection over an array of device pointers, the T** and type+offset overloads
 * (H:60590-60981, H:61359-61738) with member pointers expressed as byte offsets
 * (memberobjectgenerator H:656-668, getoffsetof H:711-720):
 *   first level:  key = *(T *)((char *)input[i] + indirection1)
 *   second level: key = ((T *)(*(char **)((char *)input[i] + indirection1) + indirection2))[index2]
 * (indirectinput1/indirectinput2 H:2745-2821). Keys are read once; the pointers are sorted.
 * flags should carry RSBD8_HIP_REV as isrevorderwithindirection decodes it. */
int rsbd8_hip_radixsortcopynoalloc_indirect(int keytype, unsigned flags, size_t count, void *workspace,
	size_t workspace_bytes, const void *const *input, void **output, ptrdiff_t indirection1, int second_level,
	ptrdiff_t indirection2, size_t index2, rsbd8_hip_stream stream);

/* In-place form of the above (radixsortnoalloc with indirection, H:61017-61045, H:61395-61427):
 * the sorted pointers land in input, or in buffer when movetobuffer is set. */
int rsbd8_hip_radixsortnoalloc_indirect(int keytype, unsigned flags, size_t count, void *workspace,
	size_t workspace_bytes, void **input, void **buffer, int movetobuffer, ptrdiff_t indirection1,
	int second_level, ptrdiff_t indirection2, size_t index2, rsbd8_hip_stream stream);

/* Records addressed by 32-bit indices: the type+offset overload (H:61360-61393) applied to the
 * pointers base + idx_in[i] * stride, returning the sorted indices instead of pointers. */
int rsbd8_hip_sort_indices(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *base, size_t stride, size_t key_offset, const uint32_t *idx_in, uint32_t *idx_out,
	rsbd8_hip_stream stream);

/* Keys with a 32-bit payload (the engine under the indirect sorts): a stable sort of keys_in
 * that moves payload_in along (payload_in NULL: the payload is the source index).
 * keys_buffer/payload_buffer are count-element scratch; no array may alias another. */
int rsbd8_hip_sort_pairs(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *keys_in, void *keys_out, void *keys_buffer, const uint32_t *payload_in, uint32_t *payload_out,
	uint32_t *payload_buffer, rsbd8_hip_stream stream);

/* Segmented (batched) sort: nseg independent arrays in one CSR batch, each sorted stably with the
 * same flags (the reference runs one radixsortcopy per array, H:60468-60527; no counterpart for the
 * batch itself). A plan is made once from the host copy of the nseg+1 offsets. Segments up to the
 * LDS capacity of small_cfg (0: 1024, 1: 4096, 2: 8192 keys) are sorted in LDS, one block per
 * segment (one wave per segment of at most 64 keys); longer ones are partitioned by their top digit
 * in tiles of tile_cfg (threads x keys per thread: 256x8, 256x16, 512x8, 512x16, 256x24) and each
 * run of buckets of at most that capacity is then sorted in LDS the same way; a segment with a
 * larger bucket takes per-digit onesweep passes over its tiles instead. -1 selects the tuned choice
 * for the key width and mode. The plan owns its device tables and workspace. run issues the kernels
 * on a stream; capture records them into a hipGraph kept in the plan, and launch_graph replays that
 * graph with one launch (on the arrays given to capture). buffer holds total elements; input,
 * output and buffer must not alias. dev_offsets is the device copy of the same offsets. info
 * receives {short segments, long segments, tiles, tile keys, LDS capacity, total keys, plan bytes}. */
typedef struct rsbd8_hip_segplan rsbd8_hip_segplan;
int rsbd8_hip_segmented_plan_create(int keytype, unsigned flags, size_t nseg, const uint64_t *host_offsets, int tile_cfg,
	int small_cfg, rsbd8_hip_segplan **plan);
/* The automatic (tuned) tile_cfg and small_cfg of a segmented plan for this key type and flags;
 * host only, no GPU call. */
int rsbd8_hip_segmented_default_config(int keytype, unsigned flags, int *tile_cfg, int *small_cfg);
int rsbd8_hip_segmented_run(rsbd8_hip_segplan *plan, const void *input, void *output, void *buffer,
	const uint64_t *dev_offsets, rsbd8_hip_stream stream);
int rsbd8_hip_segmented_capture(rsbd8_hip_segplan *plan, const void *input, void *output, void *buffer,
	const uint64_t *dev_offsets, rsbd8_hip_stream stream);
int rsbd8_hip_segmented_launch_graph(rsbd8_hip_segplan *plan, rsbd8_hip_stream stream);
int rsbd8_hip_segmented_plan_info(const rsbd8_hip_segplan *plan, uint64_t info[7]);
void rsbd8_hip_segmented_plan_destroy(rsbd8_hip_segplan *plan);

/* Process-wide tile configuration of the scatter passes: 0..4 selects one of
 * {256x8, 256x16, 512x8, 1024x12, 256x24} threads x keys per thread, -1 restores the automatic
 * choice. Returns the previous setting. (No reference counterpart: the reference's static
 * thread thresholds H:12948-13152 are its closest analogue.) */
int rsbd8_hip_set_tile_config(int cfg);

/* Process-wide pass algorithm: 1 onesweep (each scatter pass resolves its tile offsets with a
 * decoupled look-back over the preceding tiles), 2 count + scan + scatter (per-tile digit counts,
 * a decoupled look-back chained scan of the count matrix, then a scatter with no inter-block
 * dependency), 4 two MSD partitions followed by one LDS sort per resulting bucket, 0 automatic.
 * Algorithm 4 partitions keys-only integer sorts under an injective transform (unsigned and
 * signed modes) by their top two digits, with exact bucket sizes from one histogram read (the
 * device falls back to 2 when those buckets are too large for LDS); every other sort of 4- and
 * 8-byte keys with a buffer (floats, the absolute-value modes, sorts with a payload) is
 * partitioned by splitters sampled from the keys; 1- and 2-byte keys use 2. Returns the previous
 * setting. */
int rsbd8_hip_set_algorithm(int algo);

/* Process-wide mean bucket length the splitter-sampled partitions of algorithm 4 aim at (keys per
 * bucket after the second partition; at most 256 buckets per first-partition segment, so large
 * sorts may exceed it): 0 restores the default (4096). Buckets above the per-bucket LDS sort's
 * capacity (8192 keys) are sorted through HBM, so a large target exercises that path. Returns the
 * previous setting (0 when it was the default). No reference counterpart. */
int rsbd8_hip_set_sample_target(int keys);

/* A sort plan chosen per call (no reference counterpart; the reference picks its plan from
 * compile-time thresholds, H:12948-13152): algo 0 automatic, 1 onesweep, 2 count + scan + scatter,
 * 3 a single block sorting everything in LDS (keys only, count <= 16384; otherwise automatic),
 * 4 the two-level MSD plan (as for rsbd8_hip_set_algorithm), 5 the two-level MSD plan by splitters
 * whatever the keys (4- and 8-byte keys with a buffer, else automatic: integer keys with few values
 * or a skewed distribution get equality buckets instead of failing the exact plan's bucket test);
 * tile_cfg -1 automatic or 0..4 as for rsbd8_hip_set_tile_config; presort_check first tests
 * whether the input is already in order or strictly reversed (then the sort is one copy), or in
 * reverse order with ties (then one pass moves the runs of equal keys, each in its own order).
 * For keys-only sorts to another array the check writes output while it reads (the copy the
 * verdict will ask for, where the input allows it) and the later work overwrites what it must: the
 * output array holds no input-dependent data the caller may rely on until the sort completes. */
typedef struct rsbd8_hip_plan {
	int algo;
	int tile_cfg;
	int presort_check;
	int reserved;
} rsbd8_hip_plan;

/* rsbd8_hip_radixsortcopynoalloc with an explicit plan. */
int rsbd8_hip_radixsortcopynoalloc_plan(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *input, void *output, void *buffer, const rsbd8_hip_plan *plan, rsbd8_hip_stream stream);

/* The candidate plans, indexed in the order of the selector network's outputs (0 <= index <
 * rsbd8_hip_plan_count()): onesweep, count + scan, then the two-level MSD plan, each over the tile
 * configurations, each without and with the presort check; then algorithm 5 (4- and 8-byte keys, count
 * >= 65536); the last is the single-block LDS sort (count <= 16384). */
int rsbd8_hip_plan_count(void);
int rsbd8_hip_plan_at(int index, rsbd8_hip_plan *plan);

/* Learned plan selection (no reference counterpart; the reference's analogue is its compile-time
 * thresholds, Radixsortbidi8.hpp:12948-13152, and count-based branches, H:56797-57052). One block
 * samples the keys, evaluates the selector network (params: 24 -> hidden -> hidden ->
 * rsbd8_hip_plan_count() outputs, hidden <= 64, as trained by the mlp_* calls) and picks the plan with
 * the lowest predicted time among those valid for count. select_plan blocks until the choice is
 * known (it synchronizes the stream; the first call on a thread allocates one page of pinned host
 * memory, kept until the process ends). radixsortcopynoalloc_learned sorts with the chosen plan
 * exactly as rsbd8_hip_radixsortcopynoalloc_plan does; below min_count keys it skips the selection
 * (and its round trip) and uses the plan below (NULL: the automatic plan). plan_index (may be NULL)
 * receives the choice, or -1 when the selection was skipped. */
int rsbd8_hip_select_plan(int keytype, unsigned flags, size_t count, const void *input, const float *params, int hidden,
	int *plan_index, rsbd8_hip_stream stream);
int rsbd8_hip_radixsortcopynoalloc_learned(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *input, void *output, void *buffer, const float *params, int hidden, size_t min_count,
	const rsbd8_hip_plan *below, int *plan_index, rsbd8_hip_stream stream);

/* The plan selector's inputs and network (no reference counterpart). rsbd8_hip_features writes 24
 * floats describing a key array from a sample of its keys (size, width, flags, per-digit byte
 * entropy, neighbour order statistics); the mlp_* calls run a 3-layer ReLU network
 * features -> hidden -> hidden -> outputs (hidden <= 64) on device arrays: forward (acts, 2*hidden
 * floats per sample, may be NULL), backward of the masked mean squared error (per_sample holds
 * batch * param_count floats, grad param_count floats) and an Adam step. */
int rsbd8_hip_features(int keytype, unsigned flags, size_t count, const void *input, float *features, rsbd8_hip_stream stream);
int rsbd8_hip_mlp_param_count(int features, int hidden, int outputs);
int rsbd8_hip_mlp_forward(int features, int hidden, int outputs, const float *params, const float *x, int batch,
	float *acts, float *y, rsbd8_hip_stream stream);
int rsbd8_hip_mlp_backward(int features, int hidden, int outputs, const float *params, const float *x, int batch,
	const float *acts, const float *y, const float *target, const float *mask, float mask_sum, float *per_sample,
	float *grad, rsbd8_hip_stream stream);
int rsbd8_hip_mlp_adam(int count, float *params, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
	float eps, int step, float decay, rsbd8_hip_stream stream);

/* Per-launch HIP-event timing of the sorts issued between begin and end, on their own streams
 * (no reference counterpart; the reference's benchmark times whole calls with rdtsc,
 * seqptests/WindowsProject1/WindowsProject1.cpp:912-937). end synchronizes and returns, per kind,
 * the summed milliseconds and launch counts: 0 histogram, 1 scatter launches of executed passes,
 * 2 scatter launches of skipped (uniform-digit) passes, 3 count + scan launches of executed passes,
 * 4 count + scan launches of skipped passes. */
int rsbd8_hip_profile_begin(void);
int rsbd8_hip_profile_end(double *ms, unsigned *launches, int nkinds);

/* Message of the last failure on this thread ("" if none). */
const char *rsbd8_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
