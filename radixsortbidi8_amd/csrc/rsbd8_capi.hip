// rsbd8_capi.hip -- the extern "C" boundary (include/rsbd8_hip.h) over the gfx950 sort engine,
// plus the gather/permute kernels of the indirection layouts.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <string>

#include "../../include/rsbd8_hip.h"
#include "rsbd8_engine.hpp"
#include "rsbd8_segmented.hpp"
#include "rsbd8_plan.hpp"

namespace rsbd8hip {

namespace {

// ---------------------------------------------------------------------------------------------
// gather / permute kernels

template <typename K>
__device__ __forceinline__ K load_unaligned(const unsigned char *p)
{
	K v = 0;
#pragma unroll
	for (int b = 0; b < static_cast<int>(sizeof(K)); ++b) v |= static_cast<K>(p[b]) << (8 * b);
	return v;
}

template <typename K>
__global__ __launch_bounds__(256) void k_gather_index(const unsigned char *__restrict__ base, uint64_t stride,
	uint64_t offset, const uint32_t *__restrict__ idx, K *__restrict__ keys, uint64_t n, int aligned)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
		const unsigned char *p = base + static_cast<uint64_t>(idx[i]) * stride + offset;
		keys[i] = aligned ? *reinterpret_cast<const K *>(p) : load_unaligned<K>(p);
	}
}

template <typename K>
__global__ __launch_bounds__(256) void k_gather_pointer(const unsigned char *const *__restrict__ ptrs, int64_t off1,
	int second_level, int64_t off2, uint64_t index2, K *__restrict__ keys, uint64_t n)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
		const unsigned char *p = ptrs[i] + off1;
		if (second_level) {
			const unsigned char *q = (reinterpret_cast<uintptr_t>(p) & 7u)
				? reinterpret_cast<const unsigned char *>(load_unaligned<uint64_t>(p))
				: *reinterpret_cast<const unsigned char *const *>(p);
			p = q + off2 + index2 * sizeof(K);
		}
		keys[i] = (reinterpret_cast<uintptr_t>(p) & (sizeof(K) - 1)) ? load_unaligned<K>(p) : *reinterpret_cast<const K *>(p);
	}
}

__global__ __launch_bounds__(256) void k_permute_ptr(const void *const *__restrict__ src, const uint32_t *__restrict__ perm,
	void **__restrict__ dst, uint64_t n)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step)
		dst[i] = const_cast<void *>(src[perm[i]]);
}

__global__ __launch_bounds__(256) void k_permute_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ perm,
	uint32_t *__restrict__ dst, uint64_t n)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) dst[i] = src[perm[i]];
}

inline dim3 grid_for(uint64_t n)
{
	const uint64_t g = (n + 255) / 256;
	const uint64_t cap = 8ull * device_cu_count();
	return dim3(static_cast<unsigned>(g < 1 ? 1 : (g > cap ? cap : g)));
}

} // namespace

hipError_t gather_by_index(int kb, uint64_t n, const void *base, uint64_t stride, uint64_t offset, const uint32_t *idx,
	void *keys, hipStream_t s)
{
	const auto *b = static_cast<const unsigned char *>(base);
	const int aligned = ((reinterpret_cast<uintptr_t>(b) + offset) % kb == 0) && (stride % kb == 0);
	switch (kb) {
	case 1: hipLaunchKernelGGL(k_gather_index<uint8_t>, grid_for(n), dim3(256), 0, s, b, stride, offset, idx, static_cast<uint8_t *>(keys), n, aligned); break;
	case 2: hipLaunchKernelGGL(k_gather_index<uint16_t>, grid_for(n), dim3(256), 0, s, b, stride, offset, idx, static_cast<uint16_t *>(keys), n, aligned); break;
	case 4: hipLaunchKernelGGL(k_gather_index<uint32_t>, grid_for(n), dim3(256), 0, s, b, stride, offset, idx, static_cast<uint32_t *>(keys), n, aligned); break;
	case 8: hipLaunchKernelGGL(k_gather_index<uint64_t>, grid_for(n), dim3(256), 0, s, b, stride, offset, idx, static_cast<uint64_t *>(keys), n, aligned); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

hipError_t gather_by_pointer(int kb, uint64_t n, const void *const *ptrs, int64_t off1, int second_level, int64_t off2,
	uint64_t index2, void *keys, hipStream_t s)
{
	const auto *p = reinterpret_cast<const unsigned char *const *>(ptrs);
	switch (kb) {
	case 1: hipLaunchKernelGGL(k_gather_pointer<uint8_t>, grid_for(n), dim3(256), 0, s, p, off1, second_level, off2, index2, static_cast<uint8_t *>(keys), n); break;
	case 2: hipLaunchKernelGGL(k_gather_pointer<uint16_t>, grid_for(n), dim3(256), 0, s, p, off1, second_level, off2, index2, static_cast<uint16_t *>(keys), n); break;
	case 4: hipLaunchKernelGGL(k_gather_pointer<uint32_t>, grid_for(n), dim3(256), 0, s, p, off1, second_level, off2, index2, static_cast<uint32_t *>(keys), n); break;
	case 8: hipLaunchKernelGGL(k_gather_pointer<uint64_t>, grid_for(n), dim3(256), 0, s, p, off1, second_level, off2, index2, static_cast<uint64_t *>(keys), n); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

hipError_t permute_pointers(uint64_t n, const void *const *ptrs, const uint32_t *perm, void **out, hipStream_t s)
{
	hipLaunchKernelGGL(k_permute_ptr, grid_for(n), dim3(256), 0, s, ptrs, perm, out, n);
	return hipGetLastError();
}

hipError_t permute_u32(uint64_t n, const uint32_t *src, const uint32_t *perm, uint32_t *out, hipStream_t s)
{
	hipLaunchKernelGGL(k_permute_u32, grid_for(n), dim3(256), 0, s, src, perm, out, n);
	return hipGetLastError();
}

} // namespace rsbd8hip

using namespace rsbd8hip;

namespace {

thread_local std::string g_last_error;
std::atomic<int> g_tile_cfg{-1};
std::atomic<int> g_algo{0};

int fail(int code, const char *what)
{
	g_last_error = what;
	return code;
}

int fail_hip(hipError_t e, const char *what)
{
	g_last_error = std::string(what) + ": " + hipGetErrorString(e);
	return e == hipErrorOutOfMemory ? RSBD8_HIP_ENOMEM : (e == hipErrorInvalidValue ? RSBD8_HIP_EINVAL : RSBD8_HIP_EDEVICE);
}

int key_bytes_of(int keytype)
{
	switch (keytype) {
	case RSBD8_HIP_U8: case RSBD8_HIP_I8: return 1;
	case RSBD8_HIP_U16: case RSBD8_HIP_I16: return 2;
	case RSBD8_HIP_U32: case RSBD8_HIP_I32: case RSBD8_HIP_F32: return 4;
	case RSBD8_HIP_U64: case RSBD8_HIP_I64: case RSBD8_HIP_F64: return 8;
	default: return 0;
	}
}

// keys wider than 64 bits: bytes per record (0 for the narrow types) and value bits
int wide_bytes_of(int keytype)
{
	switch (keytype) {
	case RSBD8_HIP_LD80: return 10;
	case RSBD8_HIP_LD96: return 12;
	case RSBD8_HIP_LD128: case RSBD8_HIP_U128: case RSBD8_HIP_I128: return 16;
	default: return 0;
	}
}
int wide_bits_of(int keytype) { return keytype == RSBD8_HIP_U128 || keytype == RSBD8_HIP_I128 ? 128 : 80; }
// bytes per key record, narrow or wide
int record_bytes_of(int keytype) { return key_bytes_of(keytype) ? key_bytes_of(keytype) : wide_bytes_of(keytype); }

inline size_t align_up(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

struct IndirectWs {
	void *keys_a, *keys_b;
	uint32_t *perm_a, *perm_b;
};

// The gathered keys and the permutations sit at the end of the workspace: a larger sort's tile
// records (which grow from the start) never reach where a smaller sort left its arrays, since the
// workspace holds both the records and the arrays of its largest count.
IndirectWs carve_indirect(void *ws, size_t ws_bytes, size_t count, int kb)
{
	const size_t arrays = 2 * align_up(count * kb) + 2 * align_up(count * 4);
	char *p = static_cast<char *>(ws) + ((ws_bytes - arrays) & ~static_cast<size_t>(255));
	IndirectWs r;
	r.keys_a = p; p += align_up(count * kb);
	r.keys_b = p; p += align_up(count * kb);
	r.perm_a = reinterpret_cast<uint32_t *>(p); p += align_up(count * 4);
	r.perm_b = reinterpret_cast<uint32_t *>(p);
	return r;
}

bool valid_flags(unsigned flags) { return flags < 32u; }

bool is_device_memory(const void *p)
{
	hipPointerAttribute_t attr;
	if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
		(void)hipGetLastError();
		return false;
	}
	return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// Every array a noalloc entry point hands to its kernels must be GPU memory (hipMalloc'd, or
// managed): a host array would reach the kernels and fault the GPU. Checked on the host before any
// launch; null pointers are left to the callers' own checks.
int check_device_arrays(std::initializer_list<const void *> arrays, const char *what)
{
	for (const void *p : arrays)
		if (p && !is_device_memory(p))
			return fail(RSBD8_HIP_EINVAL, (std::string(what) + ": an array is not device memory (host arrays are accepted only by "
				"rsbd8_hip_radixsort and rsbd8_hip_radixsortcopy)").c_str());
	return RSBD8_HIP_OK;
}

int check_common(int keytype, unsigned flags, size_t count, size_t ws_bytes, int layout)
{
	if (!record_bytes_of(keytype)) return fail(RSBD8_HIP_EINVAL, "unknown key type");
	if (!valid_flags(flags)) return fail(RSBD8_HIP_EINVAL, "unknown flag bits");
	if (count > 0xFFFFFFFFull) return fail(RSBD8_HIP_EINVAL, "count above 2^32-1 is not supported");
	if (ws_bytes < rsbd8_hip_workspace_size(keytype, count, layout)) return fail(RSBD8_HIP_EINVAL, "workspace too small");
	return RSBD8_HIP_OK;
}

bool needs_buffer(int kb, unsigned flags)
{
	if (kb != 1) return true;
	const unsigned m = flags & 7u;
	return m == (RSBD8_HIP_ABS | RSBD8_HIP_SIGN) || m == (RSBD8_HIP_SIGN | RSBD8_HIP_FLTP) ||
		m == (RSBD8_HIP_ABS | RSBD8_HIP_SIGN | RSBD8_HIP_FLTP);
}

// ---------------------------------------------------------------------------------------------
// The workspace of rsbd8_hip_radixsortcopynoalloc_static: a file-scope device array, one copy per
// GPU, zero when the code object loads and held until the process ends. Host code never passes
// the array's name to a kernel (on the host it is only a stand-in for the symbol): it uses the
// device address hipGetSymbolAddress returns, looked up once per device with that device current.

constexpr uint64_t kStaticMaxCount = RSBD8_HIP_STATIC_MAX_COUNT;
constexpr size_t kStaticWorkspaceBytes = workspace_bytes_for(kStaticMaxCount);
static_assert(kStaticWorkspaceBytes % 256 == 0, "the static workspace keeps 256-byte alignment");

static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

constexpr int kMaxStaticDevices = 64;
std::mutex g_static_mu;
void *g_static_ws[kMaxStaticDevices];

// the device address of scratch on the GPU that runs stream (the current device for the null stream)
int static_workspace(hipStream_t stream, void **ptr)
{
	int cur = 0;
	hipError_t e = hipGetDevice(&cur);
	if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
	int dev = cur;
	if (stream && (e = hipStreamGetDevice(stream, &dev)) != hipSuccess) return fail_hip(e, "hipStreamGetDevice");
	if (dev < 0 || dev >= kMaxStaticDevices) return fail(RSBD8_HIP_EINVAL, "device index beyond the static workspace table");
	std::lock_guard<std::mutex> lock(g_static_mu);
	if (!g_static_ws[dev]) {
		if (dev != cur && (e = hipSetDevice(dev)) != hipSuccess) return fail_hip(e, "hipSetDevice");
		void *p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (dev != cur) (void)hipSetDevice(cur);
		if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
		if (!p) return fail(RSBD8_HIP_EDEVICE, "hipGetSymbolAddress(scratch) returned no address");
		g_static_ws[dev] = p;
	}
	*ptr = g_static_ws[dev];
	return RSBD8_HIP_OK;
}

} // namespace

extern "C" {

int rsbd8_hip_decode(int keytype, unsigned mode, unsigned direction, int indirect, unsigned *flags)
{
	if (!record_bytes_of(keytype) || !flags) return fail(RSBD8_HIP_EINVAL, "bad key type or flags pointer");
	if (!(mode <= 7u || (mode >= 64u && mode <= 66u))) return fail(RSBD8_HIP_EINVAL, "not a sortingmode value");
	if (direction > 3u) return fail(RSBD8_HIP_EINVAL, "not a sortingdirection value");
	// std::is_signed_v (true for the floating-point types) and std::is_floating_point_v; the 80-bit
	// layouts are long double, the 128-bit integers __int128 and unsigned __int128
	const bool wide = wide_bytes_of(keytype) != 0;
	const bool isfloat = wide ? wide_bits_of(keytype) == 80 : keytype >= RSBD8_HIP_F32;
	const bool issigned = wide ? (isfloat || keytype == RSBD8_HIP_I128) : keytype >= RSBD8_HIP_I8;
	bool abs, sign, fltp;
	if (mode >= 64u) {
		abs = mode >= 65u && issigned;    // H:59779-59780
		sign = mode <= 65u && issigned;   // H:59809-59811
		fltp = isfloat;                   // H:59840-59842
	} else {
		abs = mode & 1u;                  // H:59804
		sign = (mode >> 1) & 1u;          // H:59835
		fltp = (mode >> 2) & 1u;          // H:59865
	}
	unsigned f = (abs ? RSBD8_HIP_ABS : 0u) | (sign ? RSBD8_HIP_SIGN : 0u) | (fltp ? RSBD8_HIP_FLTP : 0u);
	if (direction & 1u) f |= RSBD8_HIP_DESC;  // H:59867-59870
	// tie reversal: always with indirection (H:59872-59875), otherwise only where distinct bit
	// patterns can share a key (H:59877-59904)
	if ((direction & 2u) && (indirect || (abs && sign) || (!abs && sign && fltp))) f |= RSBD8_HIP_REV;
	*flags = f;
	return RSBD8_HIP_OK;
}

size_t rsbd8_hip_workspace_size(int keytype, size_t count, int layout)
{
	const int kb = key_bytes_of(keytype);
	if (wide_bytes_of(keytype)) return wide_workspace_bytes(count);  // every layout
	size_t s = sort_workspace_bytes(count);
	if (layout == RSBD8_HIP_LAYOUT_INDIRECT) s = align_up(s) + 2 * align_up(count * (kb ? kb : 8)) + 2 * align_up(count * 4);
	return s;
}

int rsbd8_hip_workspace_init(void *workspace, size_t bytes, rsbd8_hip_stream stream)
{
	if (!workspace) return fail(RSBD8_HIP_EINVAL, "null workspace");
	const hipError_t e = workspace_init(workspace, bytes, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "workspace_init");
}

int rsbd8_hip_set_tile_config(int cfg)
{
	if (cfg < -1 || cfg >= kNumTileCfgs) return fail(RSBD8_HIP_EINVAL, "tile config out of range");
	return g_tile_cfg.exchange(cfg);
}

int rsbd8_hip_set_algorithm(int algo)
{
	if (algo < 0 || algo > 4 || algo == 3) return fail(RSBD8_HIP_EINVAL, "algorithm out of range");
	return g_algo.exchange(algo);
}

int rsbd8_hip_set_sample_target(int keys)
{
	if (keys < 0) return fail(RSBD8_HIP_EINVAL, "sample target out of range");
	return static_cast<int>(set_sample_target(static_cast<uint32_t>(keys)));
}

const char *rsbd8_hip_last_error(void) { return g_last_error.c_str(); }

int rsbd8_hip_profile_begin(void)
{
	profile_begin();
	return RSBD8_HIP_OK;
}

int rsbd8_hip_profile_end(double *ms, unsigned *launches, int nkinds)
{
	if (!ms || !launches || nkinds < 0) return fail(RSBD8_HIP_EINVAL, "null output");
	const hipError_t e = profile_end(ms, launches, nkinds);
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "profile_end");
}

// Keys wider than 64 bits (the 80-bit layouts and 128-bit integers): one wide sort on the workspace.
static int run_wide(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes, WideJob job,
	hipStream_t s, const char *what)
{
	job.rec_bytes = wide_bytes_of(keytype);
	job.value_bits = wide_bits_of(keytype);
	job.flags = flags;
	job.n = count;
	job.cfg = g_tile_cfg.load();
	job.algo = g_algo.load();
	const hipError_t e = wide_sort(job, workspace, workspace_bytes, s);
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, what);
}

int rsbd8_hip_radixsortcopynoalloc(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *input, void *output, void *buffer, rsbd8_hip_stream stream)
{
	int r = check_common(keytype, flags, count, workspace_bytes, RSBD8_HIP_LAYOUT_KEYS);
	if (r) return r;
	const int kb = key_bytes_of(keytype);
	if (!count) return RSBD8_HIP_OK;
	if (!input || !output || !workspace) return fail(RSBD8_HIP_EINVAL, "null array");
	if (input == output || (buffer && (buffer == input || buffer == output))) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if (wide_bytes_of(keytype)) {  // no buffer needed: the records are gathered once
		if ((r = check_device_arrays({workspace, input, output, buffer}, "radixsortcopynoalloc"))) return r;
		WideJob j;
		j.recs = input;
		j.rec_out = output;
		return run_wide(keytype, flags, count, workspace, workspace_bytes, j, static_cast<hipStream_t>(stream), "radixsortcopynoalloc");
	}
	if (!buffer && needs_buffer(kb, flags) && kb != 1) return fail(RSBD8_HIP_EINVAL, "buffer required");
	if ((r = check_device_arrays({workspace, input, output, buffer}, "radixsortcopynoalloc"))) return r;
	SortJob job;
	job.key_bytes = kb; job.flags = flags; job.n = count;
	job.in = input; job.out = output; job.buf = buffer;
	job.cfg = g_tile_cfg.load();
	job.algo = g_algo.load();
	const hipError_t e = sort_launch(job, workspace, workspace_bytes, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "radixsortcopynoalloc");
}

int rsbd8_hip_radixsortcopynoalloc_static(int keytype, unsigned flags, size_t count, const void *input, void *output,
	void *buffer, rsbd8_hip_stream stream)
{
	const int kb = key_bytes_of(keytype);
	if (!kb) return fail(RSBD8_HIP_EINVAL, "unknown key type");
	if (!valid_flags(flags)) return fail(RSBD8_HIP_EINVAL, "unknown flag bits");
	if (count > kStaticMaxCount)
		return fail(RSBD8_HIP_EINVAL, "radixsortcopynoalloc_static: count above RSBD8_HIP_STATIC_MAX_COUNT (2^24 keys), the "
			"capacity of the static workspace; use rsbd8_hip_radixsortcopynoalloc with a workspace argument");
	if (!count) return RSBD8_HIP_OK;
	if (!input || !output) return fail(RSBD8_HIP_EINVAL, "null array");
	if (input == output || (buffer && (buffer == input || buffer == output))) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if (!buffer && needs_buffer(kb, flags) && kb != 1) return fail(RSBD8_HIP_EINVAL, "buffer required");
	int r = check_device_arrays({input, output, buffer}, "radixsortcopynoalloc_static");
	if (r) return r;
	const hipStream_t s = static_cast<hipStream_t>(stream);
	void *ws = nullptr;
	if ((r = static_workspace(s, &ws))) return r;
	// the previous call on this GPU left the array in whatever state it ended in
	hipError_t e = workspace_reset(ws, s);
	if (e != hipSuccess) return fail_hip(e, "radixsortcopynoalloc_static");
	SortJob job;
	job.key_bytes = kb; job.flags = flags; job.n = count;
	job.in = input; job.out = output; job.buf = buffer;
	job.cfg = g_tile_cfg.load();
	job.algo = g_algo.load();
	e = sort_launch(job, ws, kStaticWorkspaceBytes, s);
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "radixsortcopynoalloc_static");
}

int rsbd8_hip_radixsortnoalloc(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	void *input, void *buffer, int movetobuffer, rsbd8_hip_stream stream)
{
	int r = check_common(keytype, flags, count, workspace_bytes, RSBD8_HIP_LAYOUT_KEYS);
	if (r) return r;
	const int kb = key_bytes_of(keytype);
	if (!count) return RSBD8_HIP_OK;
	if (!input || !workspace) return fail(RSBD8_HIP_EINVAL, "null array");
	if (buffer == input) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if (!buffer && (needs_buffer(kb, flags) || movetobuffer)) return fail(RSBD8_HIP_EINVAL, "buffer required");
	if ((r = check_device_arrays({workspace, input, buffer}, "radixsortnoalloc"))) return r;
	if (const int wb = wide_bytes_of(keytype)) {  // gathered into buffer, then copied back unless wanted there
		const hipStream_t s = static_cast<hipStream_t>(stream);
		WideJob j;
		j.recs = input;
		j.rec_out = buffer;
		if ((r = run_wide(keytype, flags, count, workspace, workspace_bytes, j, s, "radixsortnoalloc")) || movetobuffer) return r;
		const hipError_t e = hipMemcpyAsync(input, buffer, count * static_cast<size_t>(wb), hipMemcpyDeviceToDevice, s);
		return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "radixsortnoalloc copy back");
	}
	SortJob job;
	job.key_bytes = kb; job.flags = flags; job.n = count;
	job.in = input;
	job.out = movetobuffer ? buffer : input;
	job.buf = movetobuffer ? input : buffer;
	job.cfg = g_tile_cfg.load();
	job.algo = g_algo.load();
	const hipError_t e = sort_launch(job, workspace, workspace_bytes, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "radixsortnoalloc");
}

int rsbd8_hip_sort_pairs(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *keys_in, void *keys_out, void *keys_buffer, const uint32_t *payload_in, uint32_t *payload_out,
	uint32_t *payload_buffer, rsbd8_hip_stream stream)
{
	int r = check_common(keytype, flags, count, workspace_bytes, RSBD8_HIP_LAYOUT_PAIRS);
	if (r) return r;
	if (!count) return RSBD8_HIP_OK;
	if (!keys_in || !keys_out || !keys_buffer || !payload_out || !payload_buffer || !workspace)
		return fail(RSBD8_HIP_EINVAL, "null array");
	if (keys_in == keys_out || keys_in == keys_buffer || keys_out == keys_buffer || payload_out == payload_buffer ||
		(payload_in && (payload_in == payload_out || payload_in == payload_buffer)))
		return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if ((r = check_device_arrays({workspace, keys_in, keys_out, keys_buffer, payload_in, payload_out, payload_buffer},
		"sort_pairs")))
		return r;
	if (wide_bytes_of(keytype)) {
		WideJob j;
		j.recs = keys_in;
		j.rec_out = keys_out;
		j.pin = payload_in;
		j.pout = payload_out;
		return run_wide(keytype, flags, count, workspace, workspace_bytes, j, static_cast<hipStream_t>(stream), "sort_pairs");
	}
	SortJob job;
	job.key_bytes = key_bytes_of(keytype); job.flags = flags; job.n = count;
	job.in = keys_in; job.out = keys_out; job.buf = keys_buffer;
	job.payload = true; job.pin = payload_in; job.pout = payload_out; job.pbuf = payload_buffer;
	job.cfg = g_tile_cfg.load();
	job.algo = g_algo.load();
	const hipError_t e = sort_launch(job, workspace, workspace_bytes, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "sort_pairs");
}

int rsbd8_hip_sort_indices(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *base, size_t stride, size_t key_offset, const uint32_t *idx_in, uint32_t *idx_out, rsbd8_hip_stream stream)
{
	int r = check_common(keytype, flags, count, workspace_bytes, RSBD8_HIP_LAYOUT_INDIRECT);
	if (r) return r;
	if (!count) return RSBD8_HIP_OK;
	if (!base || !idx_in || !idx_out || !workspace) return fail(RSBD8_HIP_EINVAL, "null array");
	if (idx_in == idx_out) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if ((r = check_device_arrays({workspace, base, idx_in, idx_out}, "sort_indices"))) return r;
	const int kb = key_bytes_of(keytype);
	if (!kb) return fail(RSBD8_HIP_EINVAL, "sort_indices takes keys of at most 64 bits");
	hipStream_t s = static_cast<hipStream_t>(stream);
	IndirectWs w = carve_indirect(workspace, workspace_bytes, count, kb);
	hipError_t e = gather_by_index(kb, count, base, stride, key_offset, idx_in, w.keys_a, s);
	if (e != hipSuccess) return fail_hip(e, "gather_by_index");
	SortJob job;
	job.key_bytes = kb; job.flags = flags; job.n = count;
	job.in = w.keys_a; job.out = w.keys_b; job.buf = w.keys_a;  // keys end wherever the parity puts them
	job.keys_final_any = true;
	job.payload = true; job.pin = idx_in; job.pout = idx_out; job.pbuf = w.perm_a;
	job.cfg = g_tile_cfg.load();
	job.algo = g_algo.load();
	e = sort_launch(job, workspace, workspace_bytes, s);
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "sort_indices");
}

static int sort_pointers(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *const *input, void **output, ptrdiff_t ind1, int second, ptrdiff_t ind2, size_t index2, hipStream_t s)
{
	if (wide_bytes_of(keytype)) {
		WideJob j;
		j.ptrs = input;
		j.off1 = ind1; j.second_level = second; j.off2 = ind2; j.index2 = index2;
		j.ptr_out = output;
		return run_wide(keytype, flags, count, workspace, workspace_bytes, j, s, "sort_pointers");
	}
	const int kb = key_bytes_of(keytype);
	IndirectWs w = carve_indirect(workspace, workspace_bytes, count, kb);
	hipError_t e = gather_by_pointer(kb, count, input, ind1, second, ind2, index2, w.keys_a, s);
	if (e != hipSuccess) return fail_hip(e, "gather_by_pointer");
	SortJob job;
	job.key_bytes = kb; job.flags = flags; job.n = count;
	job.in = w.keys_a; job.out = w.keys_b; job.buf = w.keys_a;
	job.keys_final_any = true;
	job.payload = true; job.pin = nullptr; job.pout = w.perm_b; job.pbuf = w.perm_a;
	job.cfg = g_tile_cfg.load();
	job.algo = g_algo.load();
	if ((e = sort_launch(job, workspace, workspace_bytes, s)) != hipSuccess) return fail_hip(e, "sort_pointers");
	if ((e = permute_pointers(count, input, w.perm_b, output, s)) != hipSuccess) return fail_hip(e, "permute_pointers");
	return RSBD8_HIP_OK;
}

int rsbd8_hip_radixsortcopynoalloc_indirect(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *const *input, void **output, ptrdiff_t indirection1, int second_level, ptrdiff_t indirection2, size_t index2,
	rsbd8_hip_stream stream)
{
	int r = check_common(keytype, flags, count, workspace_bytes, RSBD8_HIP_LAYOUT_INDIRECT);
	if (r) return r;
	if (!count) return RSBD8_HIP_OK;
	if (!input || !output || !workspace) return fail(RSBD8_HIP_EINVAL, "null array");
	if (static_cast<const void *>(input) == static_cast<const void *>(output)) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if ((r = check_device_arrays({workspace, input, output}, "radixsortcopynoalloc_indirect"))) return r;
	return sort_pointers(keytype, flags, count, workspace, workspace_bytes, input, output, indirection1, second_level,
		indirection2, index2, static_cast<hipStream_t>(stream));
}

int rsbd8_hip_radixsortnoalloc_indirect(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	void **input, void **buffer, int movetobuffer, ptrdiff_t indirection1, int second_level, ptrdiff_t indirection2,
	size_t index2, rsbd8_hip_stream stream)
{
	int r = check_common(keytype, flags, count, workspace_bytes, RSBD8_HIP_LAYOUT_INDIRECT);
	if (r) return r;
	if (!count) return RSBD8_HIP_OK;
	if (!input || !buffer || !workspace) return fail(RSBD8_HIP_EINVAL, "null array");
	if (input == buffer) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if ((r = check_device_arrays({workspace, input, buffer}, "radixsortnoalloc_indirect"))) return r;
	hipStream_t s = static_cast<hipStream_t>(stream);
	// permute into buffer, then copy back unless the result is wanted in buffer
	r = sort_pointers(keytype, flags, count, workspace, workspace_bytes, input, buffer, indirection1, second_level,
		indirection2, index2, s);
	if (r || movetobuffer) return r;
	const hipError_t e = hipMemcpyAsync(input, buffer, count * sizeof(void *), hipMemcpyDeviceToDevice, s);
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "copy back");
}

static int sort_allocating(int keytype, unsigned flags, size_t count, const void *input, void *output, bool inplace,
	hipStream_t s)
{
	const int kb = record_bytes_of(keytype);
	const size_t bytes = count * kb;
	const size_t wsb = rsbd8_hip_workspace_size(keytype, count, RSBD8_HIP_LAYOUT_KEYS);
	const bool dev_in = is_device_memory(input), dev_out = inplace ? dev_in : is_device_memory(output);
	char *mem = nullptr;
	// one allocation: workspace, buffer, and staging arrays for host-side input/output
	const size_t need = align_up(wsb) + align_up(bytes) + (dev_in ? 0 : align_up(bytes)) + (dev_out ? 0 : align_up(bytes));
	hipError_t e = hipMallocAsync(reinterpret_cast<void **>(&mem), need, s);
	if (e != hipSuccess) return fail_hip(e, "allocation");
	char *ws = mem, *buf = mem + align_up(wsb), *p = buf + align_up(bytes);
	const void *din = input;
	void *dout = output;
	if (!dev_in) { din = p; p += align_up(bytes); }
	if (!dev_out) dout = inplace ? const_cast<void *>(din) : p;
	int r = RSBD8_HIP_OK;
	if ((e = workspace_init(ws, wsb, s)) != hipSuccess) { r = fail_hip(e, "workspace_init"); goto out; }
	if (!dev_in && (e = hipMemcpyAsync(const_cast<void *>(din), input, bytes, hipMemcpyHostToDevice, s)) != hipSuccess) {
		r = fail_hip(e, "upload");
		goto out;
	}
	if (wide_bytes_of(keytype)) {
		// the records are gathered into buf (never onto their own source), then copied to dout
		WideJob j;
		j.recs = din;
		j.rec_out = buf;
		if ((r = run_wide(keytype, flags, count, ws, wsb, j, s, "sort"))) goto out;
		if ((e = hipMemcpyAsync(dout, buf, bytes, hipMemcpyDeviceToDevice, s)) != hipSuccess) { r = fail_hip(e, "sort"); goto out; }
	} else {
		SortJob job;
		job.key_bytes = kb; job.flags = flags; job.n = count;
		job.in = din; job.out = dout; job.buf = buf;
		job.cfg = g_tile_cfg.load();
		job.algo = g_algo.load();
		if ((e = sort_launch(job, ws, wsb, s)) != hipSuccess) { r = fail_hip(e, "sort"); goto out; }
	}
	if (!dev_out && (e = hipMemcpyAsync(output, dout, bytes, hipMemcpyDeviceToHost, s)) != hipSuccess) {
		r = fail_hip(e, "download");
		goto out;
	}
out:
	(void)hipFreeAsync(mem, s);
	if ((e = hipStreamSynchronize(s)) != hipSuccess && r == RSBD8_HIP_OK) r = fail_hip(e, "synchronize");
	return r;
}

int rsbd8_hip_radixsortcopy(int keytype, unsigned flags, size_t count, const void *input, void *output, rsbd8_hip_stream stream)
{
	if (!record_bytes_of(keytype) || !valid_flags(flags)) return fail(RSBD8_HIP_EINVAL, "bad key type or flags");
	if (count > 0xFFFFFFFFull) return fail(RSBD8_HIP_EINVAL, "count above 2^32-1 is not supported");
	if (!input || !output) return fail(RSBD8_HIP_EINVAL, "null array");
	if (input == output) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if (count == 0) return RSBD8_HIP_OK;
	if (count == 1) {  // the reference copies the single element (H:60525)
		const hipError_t e = hipMemcpyAsync(output, input, record_bytes_of(keytype), hipMemcpyDefault, static_cast<hipStream_t>(stream));
		if (e != hipSuccess) return fail_hip(e, "copy");
		const hipError_t e2 = hipStreamSynchronize(static_cast<hipStream_t>(stream));
		return e2 == hipSuccess ? RSBD8_HIP_OK : fail_hip(e2, "synchronize");
	}
	return sort_allocating(keytype, flags, count, input, output, false, static_cast<hipStream_t>(stream));
}

int rsbd8_hip_radixsort(int keytype, unsigned flags, size_t count, void *input, rsbd8_hip_stream stream)
{
	if (!record_bytes_of(keytype) || !valid_flags(flags)) return fail(RSBD8_HIP_EINVAL, "bad key type or flags");
	if (count > 0xFFFFFFFFull) return fail(RSBD8_HIP_EINVAL, "count above 2^32-1 is not supported");
	if (!input) return fail(RSBD8_HIP_EINVAL, "null array");
	if (count <= 1) return RSBD8_HIP_OK;  // already sorted (H:60465)
	return sort_allocating(keytype, flags, count, input, input, true, static_cast<hipStream_t>(stream));
}

struct rsbd8_hip_segplan {
	rsbd8hip::SegPlan *impl;
};

int rsbd8_hip_segmented_plan_create(int keytype, unsigned flags, size_t nseg, const uint64_t *host_offsets, int tile_cfg,
	int small_cfg, rsbd8_hip_segplan **plan)
{
	if (!plan || (!host_offsets && nseg)) return fail(RSBD8_HIP_EINVAL, "null argument");
	*plan = nullptr;
	const int kb = key_bytes_of(keytype);
	if (!kb || !valid_flags(flags)) return fail(RSBD8_HIP_EINVAL, "bad key type or flags");
	rsbd8hip::segplan_default_cfg(kb, flags, &tile_cfg, &small_cfg);
	rsbd8hip::SegPlan *impl = nullptr;
	const hipError_t e = segplan_create(kb, flags, nseg, host_offsets, tile_cfg, small_cfg, &impl);
	if (e != hipSuccess) return fail_hip(e, "segmented_plan_create");
	*plan = new (std::nothrow) rsbd8_hip_segplan{impl};
	if (!*plan) { segplan_destroy(impl); return fail(RSBD8_HIP_ENOMEM, "plan allocation"); }
	return RSBD8_HIP_OK;
}

int rsbd8_hip_segmented_default_config(int keytype, unsigned flags, int *tile_cfg, int *small_cfg)
{
	if (!tile_cfg || !small_cfg) return fail(RSBD8_HIP_EINVAL, "null argument");
	const int kb = key_bytes_of(keytype);
	if (!kb || !valid_flags(flags)) return fail(RSBD8_HIP_EINVAL, "bad key type or flags");
	*tile_cfg = -1;
	*small_cfg = -1;
	rsbd8hip::segplan_default_cfg(kb, flags, tile_cfg, small_cfg);
	return RSBD8_HIP_OK;
}

int rsbd8_hip_segmented_run(rsbd8_hip_segplan *plan, const void *input, void *output, void *buffer, const uint64_t *dev_offsets,
	rsbd8_hip_stream stream)
{
	if (!plan) return fail(RSBD8_HIP_EINVAL, "null plan");
	if (input == output || (buffer && (buffer == input || buffer == output))) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if (int r = check_device_arrays({input, output, buffer, dev_offsets}, "segmented_run")) return r;
	const hipError_t e = segplan_run(plan->impl, input, output, buffer, dev_offsets, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "segmented_run");
}

int rsbd8_hip_segmented_capture(rsbd8_hip_segplan *plan, const void *input, void *output, void *buffer,
	const uint64_t *dev_offsets, rsbd8_hip_stream stream)
{
	if (!plan) return fail(RSBD8_HIP_EINVAL, "null plan");
	if (input == output || (buffer && (buffer == input || buffer == output))) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if (int r = check_device_arrays({input, output, buffer, dev_offsets}, "segmented_capture")) return r;
	const hipError_t e = segplan_capture(plan->impl, input, output, buffer, dev_offsets, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "segmented_capture");
}

int rsbd8_hip_segmented_launch_graph(rsbd8_hip_segplan *plan, rsbd8_hip_stream stream)
{
	if (!plan) return fail(RSBD8_HIP_EINVAL, "null plan");
	const hipError_t e = segplan_launch(plan->impl, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "segmented_launch_graph");
}

int rsbd8_hip_segmented_plan_info(const rsbd8_hip_segplan *plan, uint64_t info[7])
{
	if (!plan || !info) return fail(RSBD8_HIP_EINVAL, "null argument");
	segplan_info(plan->impl, info);
	return RSBD8_HIP_OK;
}

void rsbd8_hip_segmented_plan_destroy(rsbd8_hip_segplan *plan)
{
	if (!plan) return;
	segplan_destroy(plan->impl);
	delete plan;
}

namespace {

// The candidate plans in the order of the selector network's outputs: (onesweep, count + scan,
// two-level MSD) x tile configuration x (no presort check, presort check), then the sample
// partition plan (algorithm 5, its own tiles) and the single-block LDS sort.
constexpr int kPlanAlgos[3] = {kAlgoOnesweep, kAlgoCountScan, kAlgoHybrid};
constexpr int kNumGridPlans = 3 * kNumTileCfgs * 2;
constexpr int kNumPlans = kNumGridPlans + 2;
constexpr int kPlanSample = kNumGridPlans, kPlanLds = kNumGridPlans + 1;
constexpr size_t kSampleMinCount = size_t{1} << 16;
static_assert(kNumPlans <= 32, "plans are selected through a 32-bit validity mask");

rsbd8_hip_plan plan_at(int i)
{
	if (i == kPlanLds) return rsbd8_hip_plan{kAlgoLds, -1, 0, 0};
	if (i == kPlanSample) return rsbd8_hip_plan{kAlgoSample, -1, 0, 0};
	return rsbd8_hip_plan{kPlanAlgos[i / (2 * kNumTileCfgs)], (i / 2) % kNumTileCfgs, i & 1, 0};
}

// plans that can run on count keys of key_bytes bytes, as a bit mask over plan indices
uint32_t valid_plans(size_t count, int key_bytes)
{
	uint32_t m = (1u << kNumGridPlans) - 1u;
	if (key_bytes >= 4 && count >= kSampleMinCount) m |= 1u << kPlanSample;
	if (count <= kLdsMax) m |= 1u << kPlanLds;
	return m;
}

int sort_with_plan(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes, const void *input,
	void *output, void *buffer, const rsbd8_hip_plan &plan, hipStream_t stream, const char *what)
{
	int r = check_common(keytype, flags, count, workspace_bytes, RSBD8_HIP_LAYOUT_KEYS);
	if (r) return r;
	const int kb = key_bytes_of(keytype);
	if (!count) return RSBD8_HIP_OK;
	if (!input || !output || !workspace) return fail(RSBD8_HIP_EINVAL, "null argument");
	if (input == output || (buffer && (buffer == input || buffer == output))) return fail(RSBD8_HIP_EINVAL, "arrays alias");
	if (!buffer && needs_buffer(kb, flags) && kb != 1) return fail(RSBD8_HIP_EINVAL, "buffer required");
	if (plan.algo < 0 || plan.algo > kAlgoSample || plan.tile_cfg < -1 || plan.tile_cfg >= kNumTileCfgs)
		return fail(RSBD8_HIP_EINVAL, "plan out of range");
	if ((r = check_device_arrays({workspace, input, output, buffer}, what))) return r;
	SortJob job;
	job.key_bytes = kb; job.flags = flags; job.n = count;
	job.in = input; job.out = output; job.buf = buffer;
	job.cfg = plan.tile_cfg;
	job.algo = plan.algo;
	job.presort_check = plan.presort_check != 0;
	const hipError_t e = sort_launch(job, workspace, workspace_bytes, stream);
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, what);
}

// One page of pinned host memory per calling thread, allocated on its first selection and kept
// for the life of the process: the selection kernel writes its answer there.
int *selection_slot()
{
	thread_local int *slot = nullptr;
	if (!slot) {
		void *p = nullptr;
		if (hipHostMalloc(&p, sizeof(int), hipHostMallocDefault) != hipSuccess) return nullptr;
		slot = static_cast<int *>(p);
	}
	return slot;
}

} // namespace

int rsbd8_hip_radixsortcopynoalloc_plan(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *input, void *output, void *buffer, const rsbd8_hip_plan *plan, rsbd8_hip_stream stream)
{
	if (!plan) return fail(RSBD8_HIP_EINVAL, "null plan");
	return sort_with_plan(keytype, flags, count, workspace, workspace_bytes, input, output, buffer, *plan,
		static_cast<hipStream_t>(stream), "radixsortcopynoalloc_plan");
}

int rsbd8_hip_plan_count(void) { return kNumPlans; }

int rsbd8_hip_plan_at(int index, rsbd8_hip_plan *plan)
{
	if (index < 0 || index >= kNumPlans || !plan) return fail(RSBD8_HIP_EINVAL, "plan index out of range");
	*plan = plan_at(index);
	return RSBD8_HIP_OK;
}

int rsbd8_hip_select_plan(int keytype, unsigned flags, size_t count, const void *input, const float *params, int hidden,
	int *plan_index, rsbd8_hip_stream stream)
{
	const int kb = key_bytes_of(keytype);
	if (!kb || !valid_flags(flags) || !params || !plan_index || (count && !input) || hidden <= 0 || hidden > 64)
		return fail(RSBD8_HIP_EINVAL, "bad argument");
	if (count)
		if (int r = check_device_arrays({input, params}, "select_plan")) return r;
	int *slot = selection_slot();
	if (!slot) return fail(RSBD8_HIP_ENOMEM, "pinned selection slot");
	*slot = -1;
	const hipStream_t s = static_cast<hipStream_t>(stream);
	hipError_t e = select_launch(kb, flags, count, input, hidden, kNumPlans, params, valid_plans(count, kb), slot, s);
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	if (e != hipSuccess) return fail_hip(e, "select_plan");
	const int idx = __atomic_load_n(slot, __ATOMIC_ACQUIRE);
	if (idx < 0 || idx >= kNumPlans) return fail(RSBD8_HIP_EINVAL, "selection produced no plan");
	*plan_index = idx;
	return RSBD8_HIP_OK;
}

int rsbd8_hip_radixsortcopynoalloc_learned(int keytype, unsigned flags, size_t count, void *workspace, size_t workspace_bytes,
	const void *input, void *output, void *buffer, const float *params, int hidden, size_t min_count,
	const rsbd8_hip_plan *below, int *plan_index, rsbd8_hip_stream stream)
{
	if (count < min_count || count < 2) {
		// too small for the selection's round trip to pay: the given plan (or the automatic one)
		if (plan_index) *plan_index = -1;
		const rsbd8_hip_plan autoplan{0, -1, 0, 0};
		return sort_with_plan(keytype, flags, count, workspace, workspace_bytes, input, output, buffer,
			below ? *below : autoplan, static_cast<hipStream_t>(stream), "radixsortcopynoalloc_learned");
	}
	int idx = 0;
	const int r = rsbd8_hip_select_plan(keytype, flags, count, input, params, hidden, &idx, stream);
	if (r) return r;
	if (plan_index) *plan_index = idx;
	return sort_with_plan(keytype, flags, count, workspace, workspace_bytes, input, output, buffer, plan_at(idx),
		static_cast<hipStream_t>(stream), "radixsortcopynoalloc_learned");
}

int rsbd8_hip_features(int keytype, unsigned flags, size_t count, const void *input, float *features, rsbd8_hip_stream stream)
{
	const int kb = key_bytes_of(keytype);
	if (!kb || !valid_flags(flags) || !features || (count && !input)) return fail(RSBD8_HIP_EINVAL, "bad argument");
	if (count)
		if (int r = check_device_arrays({input, features}, "features")) return r;
	const hipError_t e = features_launch(kb, flags, count, input, features, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "features");
}

int rsbd8_hip_mlp_param_count(int features, int hidden, int outputs)
{
	if (features <= 0 || hidden <= 0 || outputs <= 0) return fail(RSBD8_HIP_EINVAL, "bad dimensions");
	return mlp_param_count(features, hidden, outputs);
}

int rsbd8_hip_mlp_forward(int features, int hidden, int outputs, const float *params, const float *x, int batch,
	float *acts, float *y, rsbd8_hip_stream stream)
{
	if (features <= 0 || hidden <= 0 || hidden > 64 || outputs <= 0 || !params || !x || !y || batch < 0)
		return fail(RSBD8_HIP_EINVAL, "bad argument");
	const hipError_t e = mlp_forward(features, hidden, outputs, params, x, batch, acts, y, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "mlp_forward");
}

int rsbd8_hip_mlp_backward(int features, int hidden, int outputs, const float *params, const float *x, int batch,
	const float *acts, const float *y, const float *target, const float *mask, float mask_sum, float *per_sample,
	float *grad, rsbd8_hip_stream stream)
{
	if (features <= 0 || hidden <= 0 || hidden > 64 || outputs <= 0 || !params || !x || !acts || !y || !target || !mask ||
		!per_sample || !grad || batch < 0)
		return fail(RSBD8_HIP_EINVAL, "bad argument");
	const hipError_t e = mlp_backward(features, hidden, outputs, params, x, batch, acts, y, target, mask, mask_sum, per_sample,
		grad, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "mlp_backward");
}

int rsbd8_hip_mlp_adam(int count, float *params, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
	float eps, int step, float decay, rsbd8_hip_stream stream)
{
	if (count <= 0 || !params || !grad || !m || !v || step < 1) return fail(RSBD8_HIP_EINVAL, "bad argument");
	const hipError_t e = mlp_adam(count, params, grad, m, v, lr, beta1, beta2, eps, step, decay, static_cast<hipStream_t>(stream));
	return e == hipSuccess ? RSBD8_HIP_OK : fail_hip(e, "mlp_adam");
}

} // extern "C"
