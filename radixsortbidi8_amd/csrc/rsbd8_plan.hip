// rsbd8_plan.hip -- the learned plan selector's device side: a sampled-key feature kernel and the
// forward, backward and Adam kernels of a small fully connected network (float32).
//
// No reference counterpart: the reference chooses its threading plan from compile-time thresholds
// (Radixsortbidi8.hpp:12948-13152) and run-time branches on the count (H:56797-57052). Here a network
// reads features of a sample of the keys and predicts the time of every candidate plan.
#include <hip/hip_runtime.h>

#include <math.h>

#include "rsbd8_engine.hpp"
#include "rsbd8_plan.hpp"

namespace rsbd8hip {

namespace {

// ---------------------------------------------------------------------------------------------
// features of a key array, from 4096 sampled keys and 2048 sampled neighbour pairs:
//  0      log2(n) / 32
//  1..4   key width one-hot (1, 2, 4, 8 bytes)
//  5..9   decoded flags abs, sign, fltp, desc, rev
//  10..17 normalised byte entropy of digits 0..7 of the sort key (0 past the key width)
//  18..20 fractions of sampled logical neighbour pairs with k[i] < k[i+1], ==, >
//  21     fraction of digits whose sampled entropy is zero
//  22     normalised entropy of the top 16 bits over 4096 buckets (log2 of distinct-ish spread)
//  23     1 (bias-like constant)
// Run by all 256 threads of a block; thread 0 writes the 24 features to out (global or LDS).
template <typename K>
__device__ void block_features(const K *__restrict__ in, uint64_t n, KeyXform xf, int reverse, unsigned flags,
	float *__restrict__ out)
{
	constexpr int P = sizeof(K);
	constexpr int S = 4096, PAIRS = 2048;
	__shared__ uint32_t s_h[8 * 256];
	__shared__ uint32_t s_top[4096];
	__shared__ uint32_t s_pair[3];
	__shared__ float s_ent[8];
	__shared__ float s_top_ent;
	const uint32_t t = threadIdx.x;
	for (int i = t; i < 8 * 256; i += 256) s_h[i] = 0;
	for (int i = t; i < 4096; i += 256) s_top[i] = 0;
	if (t < 3) s_pair[t] = 0;
	__syncthreads();
	if (n > 0) {
		for (int s = t; s < S; s += 256) {
			const uint64_t i = static_cast<uint64_t>(s) * n / S;
			const auto k = sort_key<K>(in[i], xf);
#pragma unroll
			for (int p = 0; p < P; ++p) atomicAdd(&s_h[p * 256 + (static_cast<uint32_t>(k >> (8 * p)) & 0xFFu)], 1u);
			const uint32_t top = static_cast<uint32_t>(k >> (8 * P - (P >= 2 ? 12 : 8))) & 0xFFFu;
			atomicAdd(&s_top[top], 1u);
		}
		if (n > 1) {
			for (int s = t; s < PAIRS; s += 256) {
				const uint64_t i = static_cast<uint64_t>(s) * (n - 1) / PAIRS;
				// logical neighbours: physical i and i + 1, swapped when the input is read backwards
				auto a = sort_key<K>(in[i], xf), b = sort_key<K>(in[i + 1], xf);
				if (reverse) { const auto c = a; a = b; b = c; }
				atomicAdd(&s_pair[a < b ? 0 : (a == b ? 1 : 2)], 1u);
			}
		}
	}
	__syncthreads();
	// entropies: wave w < P handles digit w; the top-bit entropy uses all threads afterwards
	if (t < 64 * 4) {
		const uint32_t w = t >> 6, lane = t & 63u;
		for (uint32_t p = w; p < 8; p += 4) {
			float h = 0.f;
			if (static_cast<int>(p) < P) {
				for (uint32_t b = lane; b < 256; b += 64) {
					const float c = static_cast<float>(s_h[p * 256 + b]);
					if (c > 0.f) {
						const float q = c / S;
						h -= q * log2f(q);
					}
				}
			}
#pragma unroll
			for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
			if (lane == 0) s_ent[p] = h / 8.f;
		}
	}
	__syncthreads();
	if (t < 64) {
		float h = 0.f;
		for (uint32_t b = t; b < 4096; b += 64) {
			const float c = static_cast<float>(s_top[b]);
			if (c > 0.f) {
				const float q = c / S;
				h -= q * log2f(q);
			}
		}
#pragma unroll
		for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
		if (t == 0) s_top_ent = h / 12.f;
	}
	__syncthreads();
	if (t == 0) {
		const float pairs = n > 1 ? static_cast<float>(PAIRS) : 1.f;
		out[0] = n > 0 ? log2f(static_cast<float>(n)) / 32.f : 0.f;
		out[1] = P == 1; out[2] = P == 2; out[3] = P == 4; out[4] = P == 8;
		out[5] = (flags & kAbs) != 0; out[6] = (flags & kSign) != 0; out[7] = (flags & kFltp) != 0;
		out[8] = (flags & kDesc) != 0; out[9] = (flags & kRev) != 0;
		int zero = 0;
		for (int p = 0; p < 8; ++p) {
			out[10 + p] = s_ent[p];
			zero += (p < P && s_ent[p] == 0.f);
		}
		out[18] = s_pair[0] / pairs; out[19] = s_pair[1] / pairs; out[20] = s_pair[2] / pairs;
		out[21] = zero / 8.f;
		out[22] = s_top_ent;
		out[23] = 1.f;
	}
}

template <typename K>
__global__ __launch_bounds__(256) void k_features(const K *__restrict__ in, uint64_t n, KeyXform xf, int reverse,
	unsigned flags, float *__restrict__ out)
{
	block_features<K>(in, n, xf, reverse, flags, out);
}

// ---------------------------------------------------------------------------------------------
// network: x[F] -> relu(W1 x + b1)[H] -> relu(W2 h1 + b2)[H] -> W3 h2 + b3 [O]
// parameters packed as W1[H][F], b1[H], W2[H][H], b2[H], W3[O][H], b3[O]

struct Dims {
	int F, H, O;
	__host__ __device__ int w1() const { return 0; }
	__host__ __device__ int b1() const { return H * F; }
	__host__ __device__ int w2() const { return H * F + H; }
	__host__ __device__ int b2() const { return H * F + H + H * H; }
	__host__ __device__ int w3() const { return H * F + H + H * H + H; }
	__host__ __device__ int b3() const { return H * F + H + H * H + H + O * H; }
	__host__ __device__ int count() const { return H * F + H + H * H + H + O * H + O; }
};

// one block per sample; acts holds h1 and h2 (post-activation) per sample for the backward pass
__global__ __launch_bounds__(64) void k_mlp_forward(Dims d, const float *__restrict__ prm, const float *__restrict__ x, int B,
	float *__restrict__ acts, float *__restrict__ y)
{
	extern __shared__ float sm[];
	float *sx = sm, *s1 = sm + d.F, *s2 = s1 + d.H;
	const int b = blockIdx.x, t = threadIdx.x;
	if (b >= B) return;
	for (int i = t; i < d.F; i += blockDim.x) sx[i] = x[static_cast<size_t>(b) * d.F + i];
	__syncthreads();
	for (int j = t; j < d.H; j += blockDim.x) {
		float a = prm[d.b1() + j];
		for (int i = 0; i < d.F; ++i) a += prm[d.w1() + j * d.F + i] * sx[i];
		s1[j] = a > 0.f ? a : 0.f;
	}
	__syncthreads();
	for (int j = t; j < d.H; j += blockDim.x) {
		float a = prm[d.b2() + j];
		for (int i = 0; i < d.H; ++i) a += prm[d.w2() + j * d.H + i] * s1[i];
		s2[j] = a > 0.f ? a : 0.f;
	}
	__syncthreads();
	for (int j = t; j < d.O; j += blockDim.x) {
		float a = prm[d.b3() + j];
		for (int i = 0; i < d.H; ++i) a += prm[d.w3() + j * d.H + i] * s2[i];
		y[static_cast<size_t>(b) * d.O + j] = a;
	}
	if (acts) {
		for (int j = t; j < d.H; j += blockDim.x) {
			acts[static_cast<size_t>(b) * 2 * d.H + j] = s1[j];
			acts[static_cast<size_t>(b) * 2 * d.H + d.H + j] = s2[j];
		}
	}
}

// Per-sample gradients of the masked mean squared error
//   L = sum_b sum_o m[b,o] (y[b,o] - target[b,o])^2 / sum m
// written to g[b][count]; k_mlp_reduce sums them over the batch in a fixed order.
__global__ __launch_bounds__(64) void k_mlp_backward(Dims d, const float *__restrict__ prm, const float *__restrict__ x,
	int B, const float *__restrict__ acts, const float *__restrict__ y, const float *__restrict__ target,
	const float *__restrict__ mask, float inv_msum, float *__restrict__ g)
{
	extern __shared__ float sm[];
	float *sx = sm, *s1 = sm + d.F, *s2 = s1 + d.H, *dy = s2 + d.H, *d2 = dy + d.O, *d1 = d2 + d.H;
	const int b = blockIdx.x, t = threadIdx.x;
	if (b >= B) return;
	float *gb = g + static_cast<size_t>(b) * d.count();
	for (int i = t; i < d.F; i += blockDim.x) sx[i] = x[static_cast<size_t>(b) * d.F + i];
	for (int j = t; j < d.H; j += blockDim.x) {
		s1[j] = acts[static_cast<size_t>(b) * 2 * d.H + j];
		s2[j] = acts[static_cast<size_t>(b) * 2 * d.H + d.H + j];
	}
	for (int o = t; o < d.O; o += blockDim.x) {
		const size_t k = static_cast<size_t>(b) * d.O + o;
		dy[o] = 2.f * mask[k] * (y[k] - target[k]) * inv_msum;
	}
	__syncthreads();
	// layer 3
	for (int o = t; o < d.O; o += blockDim.x) {
		for (int i = 0; i < d.H; ++i) gb[d.w3() + o * d.H + i] = dy[o] * s2[i];
		gb[d.b3() + o] = dy[o];
	}
	for (int i = t; i < d.H; i += blockDim.x) {
		float a = 0.f;
		for (int o = 0; o < d.O; ++o) a += prm[d.w3() + o * d.H + i] * dy[o];
		d2[i] = s2[i] > 0.f ? a : 0.f;
	}
	__syncthreads();
	// layer 2
	for (int j = t; j < d.H; j += blockDim.x) {
		for (int i = 0; i < d.H; ++i) gb[d.w2() + j * d.H + i] = d2[j] * s1[i];
		gb[d.b2() + j] = d2[j];
	}
	for (int i = t; i < d.H; i += blockDim.x) {
		float a = 0.f;
		for (int j = 0; j < d.H; ++j) a += prm[d.w2() + j * d.H + i] * d2[j];
		d1[i] = s1[i] > 0.f ? a : 0.f;
	}
	__syncthreads();
	// layer 1
	for (int j = t; j < d.H; j += blockDim.x) {
		for (int i = 0; i < d.F; ++i) gb[d.w1() + j * d.F + i] = d1[j] * sx[i];
		gb[d.b1() + j] = d1[j];
	}
}

// Features, forward pass and choice in one block: the index of the lowest predicted time among the
// plans whose bit is set in valid (outputs <= 32, hidden <= 64); ties go to the lower index.
template <typename K>
__global__ __launch_bounds__(256) void k_select(const K *__restrict__ in, uint64_t n, KeyXform xf, int reverse,
	unsigned flags, Dims d, const float *__restrict__ prm, uint32_t valid, int *__restrict__ result)
{
	__shared__ float sx[kNumFeatures], s1[64], s2[64], sy[32];
	block_features<K>(in, n, xf, reverse, flags, sx);
	__syncthreads();
	const int t = threadIdx.x;
	if (t < d.H) {
		float a = prm[d.b1() + t];
		for (int i = 0; i < d.F; ++i) a += prm[d.w1() + t * d.F + i] * sx[i];
		s1[t] = a > 0.f ? a : 0.f;
	}
	__syncthreads();
	if (t < d.H) {
		float a = prm[d.b2() + t];
		for (int i = 0; i < d.H; ++i) a += prm[d.w2() + t * d.H + i] * s1[i];
		s2[t] = a > 0.f ? a : 0.f;
	}
	__syncthreads();
	if (t < d.O) {
		float a = prm[d.b3() + t];
		for (int i = 0; i < d.H; ++i) a += prm[d.w3() + t * d.H + i] * s2[i];
		sy[t] = a;
	}
	__syncthreads();
	if (t == 0) {
		int best = -1;
		float bv = 0.f;
		for (int o = 0; o < d.O; ++o) {
			if (!((valid >> o) & 1u)) continue;
			if (best < 0 || sy[o] < bv) { best = o; bv = sy[o]; }
		}
		*result = best;
	}
}

__global__ __launch_bounds__(256) void k_mlp_reduce(int count, int B, const float *__restrict__ g, float *__restrict__ out)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= count) return;
	float a = 0.f;
	for (int b = 0; b < B; ++b) a += g[static_cast<size_t>(b) * count + k];
	out[k] = a;
}

__global__ __launch_bounds__(256) void k_adam(int count, float *__restrict__ prm, const float *__restrict__ grad,
	float *__restrict__ m, float *__restrict__ v, float lr, float beta1, float beta2, float eps, float bc1, float bc2,
	float decay)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= count) return;
	const float gk = grad[k] + decay * prm[k];
	const float mk = beta1 * m[k] + (1.f - beta1) * gk;
	const float vk = beta2 * v[k] + (1.f - beta2) * gk * gk;
	m[k] = mk;
	v[k] = vk;
	prm[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
}

} // namespace

hipError_t features_launch(int key_bytes, unsigned flags, uint64_t n, const void *in, float *out, hipStream_t s)
{
	const KeyXform xf = make_xform(8 * key_bytes, flags);
	const int rev = (flags & kRev) ? 1 : 0;
	switch (key_bytes) {
	case 1: hipLaunchKernelGGL(k_features<uint8_t>, dim3(1), dim3(256), 0, s, static_cast<const uint8_t *>(in), n, xf, rev, flags, out); break;
	case 2: hipLaunchKernelGGL(k_features<uint16_t>, dim3(1), dim3(256), 0, s, static_cast<const uint16_t *>(in), n, xf, rev, flags, out); break;
	case 4: hipLaunchKernelGGL(k_features<uint32_t>, dim3(1), dim3(256), 0, s, static_cast<const uint32_t *>(in), n, xf, rev, flags, out); break;
	case 8: hipLaunchKernelGGL(k_features<uint64_t>, dim3(1), dim3(256), 0, s, static_cast<const uint64_t *>(in), n, xf, rev, flags, out); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

hipError_t select_launch(int key_bytes, unsigned flags, uint64_t n, const void *in, int H, int O, const float *prm,
	uint32_t valid, int *result, hipStream_t s)
{
	if (H <= 0 || H > 64 || O <= 0 || O > 32) return hipErrorInvalidValue;
	const KeyXform xf = make_xform(8 * key_bytes, flags);
	const int rev = (flags & kRev) ? 1 : 0;
	const Dims d{kNumFeatures, H, O};
	switch (key_bytes) {
	case 1: hipLaunchKernelGGL(k_select<uint8_t>, dim3(1), dim3(256), 0, s, static_cast<const uint8_t *>(in), n, xf, rev, flags, d, prm, valid, result); break;
	case 2: hipLaunchKernelGGL(k_select<uint16_t>, dim3(1), dim3(256), 0, s, static_cast<const uint16_t *>(in), n, xf, rev, flags, d, prm, valid, result); break;
	case 4: hipLaunchKernelGGL(k_select<uint32_t>, dim3(1), dim3(256), 0, s, static_cast<const uint32_t *>(in), n, xf, rev, flags, d, prm, valid, result); break;
	case 8: hipLaunchKernelGGL(k_select<uint64_t>, dim3(1), dim3(256), 0, s, static_cast<const uint64_t *>(in), n, xf, rev, flags, d, prm, valid, result); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

int mlp_param_count(int F, int H, int O) { return Dims{F, H, O}.count(); }

hipError_t mlp_forward(int F, int H, int O, const float *prm, const float *x, int B, float *acts, float *y, hipStream_t s)
{
	if (B <= 0) return hipSuccess;
	const Dims d{F, H, O};
	const size_t sh = static_cast<size_t>(F + 2 * H) * sizeof(float);
	hipLaunchKernelGGL(k_mlp_forward, dim3(B), dim3(64), sh, s, d, prm, x, B, acts, y);
	return hipGetLastError();
}

hipError_t mlp_backward(int F, int H, int O, const float *prm, const float *x, int B, const float *acts, const float *y,
	const float *target, const float *mask, float msum, float *per_sample, float *grad, hipStream_t s)
{
	if (B <= 0) return hipSuccess;
	const Dims d{F, H, O};
	const size_t sh = static_cast<size_t>(F + 4 * H + O) * sizeof(float);
	hipLaunchKernelGGL(k_mlp_backward, dim3(B), dim3(64), sh, s, d, prm, x, B, acts, y, target, mask,
		msum > 0.f ? 1.f / msum : 0.f, per_sample);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	const int c = d.count();
	hipLaunchKernelGGL(k_mlp_reduce, dim3((c + 255) / 256), dim3(256), 0, s, c, B, per_sample, grad);
	return hipGetLastError();
}

hipError_t mlp_adam(int count, float *prm, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
	float eps, int step, float decay, hipStream_t s)
{
	const float bc1 = 1.f - powf(beta1, static_cast<float>(step));
	const float bc2 = 1.f - powf(beta2, static_cast<float>(step));
	hipLaunchKernelGGL(k_adam, dim3((count + 255) / 256), dim3(256), 0, s, count, prm, grad, m, v, lr, beta1, beta2, eps,
		bc1, bc2, decay);
	return hipGetLastError();
}

} // namespace rsbd8hip
