// rsbd8_engine.hpp -- host-side interface of the gfx950 sort engine (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "rsbd8_device.hpp"

namespace rsbd8hip {

// Tile configurations of the onesweep scatter kernel: threads per block x keys per thread.
struct TileCfg {
	int threads;
	int kpt;
};
constexpr int kNumTileCfgs = 5;
constexpr TileCfg kTileCfgs[kNumTileCfgs] = {{256, 8}, {256, 16}, {512, 8}, {1024, 12}, {256, 24}};
constexpr int kMinTile = 256 * 8;

// One planned pass as the device sees it (written by the plan kernel, read by the scatter kernel).
struct PassPlan {
	const void *src;
	void *dst;
	const uint32_t *psrc;
	uint32_t *pdst;
	uint32_t active;
	uint32_t reverse;  // read the source back to front (isrevorder on the first read)
	uint32_t iota;     // payload is the physical source index (no payload array to read)
	uint32_t pad;
};

// Device control block at the start of every workspace.
struct Control {
	uint32_t tile_counter[kMaxPasses];
	uint32_t scan_counter[kMaxPasses];
	uint32_t epoch;        // advanced by every sort; tags the look-back words of each pass
	uint32_t num_active;
	uint32_t active_mask;
	uint32_t counts_ready; // bit p: the tile counts of pass p were produced by the histogram read
	uint32_t check;        // presort check: bit 0 ran, bit 1 some k[j] < k[j+1], bit 2 equal, bit 3 greater,
	                       // bit 4 the run reversal allowed, bits 5 / 6 some wave copied the input to spec_dst
	                       // in order / reversed
	uint32_t pad[3];
	PassPlan pass[kMaxPasses];
	PassPlan pre;          // copy before the first pass (in-place sorts that would overlap)
	PassPlan fin;          // copy when every digit is uniform (no pass executes)
	// two-level MSD plans (4- and 8-byte keys): hyb[0] is the first partition, hyb[1] the second
	// inside each first-partition segment, hyb[2] the per-bucket sort. The plan by the top two
	// digits reads only the top-16-bit histogram; when the LSD plan is needed after all, need_full
	// gates a full histogram read and a second plan.
	PassPlan hyb[3];
	uint32_t need_full;
	uint32_t hyb_scan_counter;
	uint32_t hyb_queue;    // buckets queued for the LDS LSD passes (long runs, or no room for positions)
	uint32_t joint_max;    // largest bucket of the top-16-bit histogram (the exact MSD plan test)
	uint32_t bucket_count; // sample partition plan: buckets queued for the smaller per-bucket sort
	uint32_t bucket_count2;// ... and for the larger one
	uint32_t zero_signs;   // sample partition plan's first read: bit 0 a +0.0 key seen, bit 1 a -0.0 key
	uint32_t group_parts;  // sample partition plan: 1 when its partitions group (unstable), 0 when they rank
	void *spec_dst;        // presort check: the array it copied the input to (check bits 5 and 6)
	// presort check: some wave saw an ascent and a descent (the others may stop): word 0 of 16 cache
	// lines, block b setting line b % 16 (stores to one word from every wave of the grid queue on one
	// L2 channel)
	alignas(128) uint32_t check_stop[16][32];
};
constexpr uint32_t kStopLines = 16;

constexpr size_t kCtrlBytes = 4096;
constexpr size_t kHistBytes = kMaxPasses * kRadix * sizeof(uint32_t);
constexpr int kScanChunk = 64;  // tiles per block of the chained scan over the tile counts
constexpr int kTagSlots = 16;   // look-back tag slots per sort: LSD passes 0..7, the MSD plan's scan 8
constexpr uint32_t kHybBuckets = kRadix * kRadix;  // buckets after the two MSD partitions
// tables of the two-level MSD plan after the tile counts: first tile of every top-digit segment
// (257 words), segment starts and lengths (256 + 256), bucket starts (relative to their segment)
// and lengths (2 x 65536), the queue of buckets for the LDS LSD passes (65536), the bucket cursors
// of the second partition (65536), the overflow counts of the top-16-bit histogram (65536, zero
// between sorts), the top-digit counts of the first partition's key ranges, then its cursors
// (8 x 256), and (16-byte aligned) the sample partition plan's bucket queues: three of 65536
// four-word entries (its smaller buckets, its larger ones, and those its LDS sorts pass on)
constexpr size_t kSegTableWords = 257 + 256 + 256 + 5 * kHybBuckets + 8 * 256 + 3 + 3 * 4 * kHybBuckets;
// Splitter tables of the sample partition plan (algorithm 4 for keys it cannot bucket by bits):
// one for the first partition and one per first-partition segment for the second. A table maps a
// transformed key to a bucket id by the number of its bounds below the key (see split_id):
// bounds as 256 u64 (unused ones all-ones), 256 info words (open bucket id in bits 0-7, bit 8: the
// bound has an equality bucket of its own, whose id is the open id + 1), the number of real bounds,
// the number of ids, an 8-word mask of the equality buckets, the number of sub-tables in use, then
// the cell table that finds a key's place among the bounds: the lowest and highest bound (u64
// each), the cell shift, kSplitCells u16 cells and kSplitSubs sub-tables of kSplitSubCells u16
// sub-cells (see SplitLds, split_id).
constexpr uint32_t kSplitBoundsWord = 0, kSplitInfoWord = 512, kSplitCountWord = 768, kSplitIdsWord = 769,
	kSplitEqWord = 770, kSplitNsubWord = 778, kSplitLoWord = 784, kSplitHiWord = 786, kSplitShWord = 788,
	kSplitCellWord = 792, kSplitSubWord = kSplitCellWord + kSplitCells / 2;
constexpr size_t kSplitTabWords = kSplitSubWord + kSplitSubs * kSplitSubCells / 2 + 8;
constexpr size_t kSplitTables = 1 + kRadix;
static_assert(sizeof(Control) <= kCtrlBytes, "control block too large");

// Workspace layout: [control][histograms][bucket starts][MSD tables][splitter tables][tile records ...]. Tile record
// t holds the 256 look-back words of tile (or scan chunk) t followed by the 256 digit counts of
// tile t. Every region sits at the same address for every sort on a workspace, whatever its size,
// so a look-back word is only ever overwritten by another (epoch-tagged) look-back word: counts,
// tables or keys left by a sort of another size can never pass for a published one.
constexpr size_t kSegTableBytes = ((kSegTableWords * 4 + 255) / 256) * 256;
constexpr size_t kSplitTableBytes = ((kSplitTables * kSplitTabWords * 4 + 255) / 256) * 256;
constexpr size_t kHeaderBytes = kCtrlBytes + 2 * kHistBytes + kSegTableBytes + kSplitTableBytes;
constexpr size_t kRecordBytes = kRadix * (sizeof(uint64_t) + sizeof(uint32_t));
constexpr size_t kLbStride = kRecordBytes / sizeof(uint64_t);     // u64 words from one tile's look-back row to the next
constexpr size_t kCountStride = kRecordBytes / sizeof(uint32_t);  // u32 words from one tile's count row to the next

struct SortJob {
	int key_bytes = 4;             // 1, 2, 4 or 8
	unsigned flags = 0;            // kAbs | kSign | kFltp | kDesc | kRev (already decoded)
	uint64_t n = 0;
	const void *in = nullptr;      // source keys (may equal out or buf for in-place sorts)
	void *out = nullptr;           // final keys
	void *buf = nullptr;           // the other ping-pong array (may be null when never needed)
	bool payload = false;          // carry a u32 payload along
	const uint32_t *pin = nullptr; // source payload; null means "source index"
	uint32_t *pout = nullptr;
	uint32_t *pbuf = nullptr;
	int cfg = -1;                  // tile configuration index, -1 picks the default
	bool fill8 = true;             // 8-bit keys-only bijective modes may use the count fill
	bool keys_final_any = false;   // only the payload is wanted: keys may end in out or buf
	int algo = 0;                  // 0 automatic, 1 onesweep (look-back in the scatter), 2 count + chained scan + scatter,
	                               // 3 one block sorts everything in LDS (keys only, n <= kLdsMax),
	                               // 4 two MSD partitions + per-bucket LDS sort (4/8-byte keys: by the top two
	                               //   digits for keys-only integer sorts under an injective transform, the
	                               //   device falling back to algorithm 2 for skewed keys; else by splitters),
	                               // 5 the same by splitters whatever the keys (4/8-byte keys with a buffer; else automatic)
	bool presort_check = false;    // first test whether the input is already in order (or strictly reversed)
};

constexpr int kAlgoAuto = 0, kAlgoOnesweep = 1, kAlgoCountScan = 2, kAlgoLds = 3, kAlgoHybrid = 4, kAlgoSample = 5;
constexpr uint32_t kLdsMax = 1024 * 16;    // keys-only sorts one LDS block takes
constexpr uint32_t kLdsPayMax = 512 * 16;  // the same with a u32 payload

// bytes of a keys-only workspace for n keys: the header, then one tile record per kMinTile keys,
// one more, and one partial tile per top-digit segment of the MSD plan
constexpr size_t workspace_bytes_for(uint64_t n)
{
	return kHeaderBytes + static_cast<size_t>((n + kMinTile - 1) / kMinTile + 1 + kRadix) * kRecordBytes;
}

size_t sort_workspace_bytes(uint64_t n);
hipError_t workspace_init(void *ws, size_t bytes, hipStream_t stream);
// Brings a workspace that another sort used (or left half-done after an error) to the state every
// sort expects: histograms zero, tile and scan counters and flags zero, a fresh look-back epoch.
// One small kernel (no memset node), so a caller may capture it.
hipError_t workspace_reset(void *ws, hipStream_t stream);
hipError_t sort_launch(const SortJob &job, void *ws, size_t ws_bytes, hipStream_t stream);
int default_cfg(int key_bytes, uint64_t n, bool payload, bool float_keys);
// the sample partition plan's mean bucket target (0: the default); returns the previous setting
uint32_t set_sample_target(uint32_t keys);
int default_algo(int key_bytes, uint64_t n, bool keys_only, bool float_keys);

// Keys wider than 64 bits (rsbd8_wide.hip): the 80-bit extended float in 10-, 12- or 16-byte
// records and 128-bit integers, sorted by their 128-bit sort key through two stable pair sorts.
struct WideJob {
	int rec_bytes = 16;                  // bytes per key record: 10, 12 or 16
	int value_bits = 80;                 // 80 (extended float layouts) or 128 (integers)
	unsigned flags = 0;                  // decoded flags
	uint64_t n = 0;
	const void *recs = nullptr;          // the key records, or (when ptrs is set)
	const void *const *ptrs = nullptr;   // pointers to them: first level at ptrs[i] + off1, second level
	int64_t off1 = 0;                    // at *(ptrs[i] + off1) + off2 + index2 * rec_bytes
	int second_level = 0;
	int64_t off2 = 0;
	uint64_t index2 = 0;
	void *rec_out = nullptr;             // the records in sorted order (may be null)
	void **ptr_out = nullptr;            // the pointers in sorted order (may be null)
	const uint32_t *pin = nullptr;       // a payload moved along into pout (null: the source index)
	uint32_t *pout = nullptr;            // (may be null)
	int cfg = -1, algo = 0;              // plan of the two pair sorts
};
size_t wide_array_bytes(uint64_t n);
size_t wide_workspace_bytes(uint64_t n);
hipError_t wide_sort(const WideJob &job, void *ws, size_t ws_bytes, hipStream_t stream);

// Gather keys through records or pointers into a dense key array (indirection layouts).
hipError_t gather_by_index(int key_bytes, uint64_t n, const void *base, uint64_t stride, uint64_t offset,
	const uint32_t *idx, void *keys_out, hipStream_t stream);
hipError_t gather_by_pointer(int key_bytes, uint64_t n, const void *const *ptrs, int64_t off1, int second_level,
	int64_t off2, uint64_t index2, void *keys_out, hipStream_t stream);
hipError_t permute_pointers(uint64_t n, const void *const *ptrs, const uint32_t *perm, void **out, hipStream_t stream);
hipError_t permute_u32(uint64_t n, const uint32_t *src, const uint32_t *perm, uint32_t *out, hipStream_t stream);

int device_cu_count();
void profile_begin();
hipError_t profile_end(double *ms, unsigned *launches, int nkinds);

} // namespace rsbd8hip
