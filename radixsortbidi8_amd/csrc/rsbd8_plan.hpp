// rsbd8_plan.hpp -- host interface of the plan selector's device kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsbd8hip {

constexpr int kNumFeatures = 24;

hipError_t features_launch(int key_bytes, unsigned flags, uint64_t n, const void *in, float *out, hipStream_t s);
// One block: features, network forward (hidden <= 64, outputs <= 32) and the index of the lowest
// output among the bits set in valid, written to *result (device or host-pinned memory).
hipError_t select_launch(int key_bytes, unsigned flags, uint64_t n, const void *in, int H, int O, const float *prm,
	uint32_t valid, int *result, hipStream_t s);
int mlp_param_count(int F, int H, int O);
hipError_t mlp_forward(int F, int H, int O, const float *prm, const float *x, int B, float *acts, float *y, hipStream_t s);
hipError_t mlp_backward(int F, int H, int O, const float *prm, const float *x, int B, const float *acts, const float *y,
	const float *target, const float *mask, float msum, float *per_sample, float *grad, hipStream_t s);
hipError_t mlp_adam(int count, float *prm, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
	float eps, int step, float decay, hipStream_t s);

} // namespace rsbd8hip
