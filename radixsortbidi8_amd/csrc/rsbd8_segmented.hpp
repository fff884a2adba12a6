// rsbd8_segmented.hpp -- host interface of the segmented (batched) sort (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsbd8hip {

struct SegPlan;
constexpr int kNumSmallCapsPublic = 3;

hipError_t segplan_create(int key_bytes, unsigned flags, uint64_t nseg, const uint64_t *host_off, int tile_cfg, int small_cfg,
	SegPlan **out);
void segplan_destroy(SegPlan *pl);
hipError_t segplan_run(SegPlan *pl, const void *in, void *out, void *buf, const uint64_t *dev_off, hipStream_t s);
hipError_t segplan_capture(SegPlan *pl, const void *in, void *out, void *buf, const uint64_t *dev_off, hipStream_t s);
hipError_t segplan_launch(SegPlan *pl, hipStream_t s);
void segplan_info(const SegPlan *pl, uint64_t *info);
int tile_keys_of(int cfg);
// the tuned automatic choice for tile_cfg / small_cfg arguments below zero
void segplan_default_cfg(int key_bytes, unsigned flags, int *tile_cfg, int *small_cfg);
int small_cap_of(int cfg);

} // namespace rsbd8hip
