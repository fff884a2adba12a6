// rsbd8_wide.hip -- keys wider than 64 bits: the 80-bit extended float in its 10-, 12- and 16-byte
// layouts (the reference's longdoubletest80/96/128 and x86 long double, Radixsortbidi8.hpp = H,
// H:1061-1804) and 128-bit integers (test128, H:2186-2404).
//
// The reference sorts these with dedicated two-register filters and scatter passes
// (filterinput H:7479-7530 for the 80-bit layouts, H:9047-9150 for test128; engines
// H:13154-23397). Here every key is turned once into its 128-bit sort key (the same closed forms
// as the narrow keys, with the value width W = 80 or 128 bits: the sign at bit W-1, padding above
// the value ignored), split into a low and a high 64-bit word, and ordered by two stable sorts of
// (word, u32 index) pairs on the narrow engine: first by the low word (source order, or reversed
// under isrevorder), then by the high word carrying the first permutation. For the 80-bit
// layouts the high word holds 16 bits, so the second sort runs two digit passes (the plan skips
// the uniform ones). The records (or pointers, or payloads) are then gathered in that order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rsbd8_engine.hpp"

namespace rsbd8hip {

namespace {

using u128 = unsigned __int128;

// the sort key of a W-bit pattern x (W = 80 or 128), descending applied as a complement
__device__ __forceinline__ u128 wide_sortkey(unsigned w, unsigned flags, u128 x)
{
	const u128 ones = w >= 128u ? ~u128(0) : ((u128(1) << w) - 1u);
	const u128 msb = u128(1) << (w - 1u);
	x &= ones;
	const bool s = (x & msb) != 0;
	const u128 m = x & ~msb;
	u128 k;
	switch (flags & 7u) {
	case 0u: k = x; break;                                               // forceunsigned
	case kSign: k = x ^ msb; break;                                      // forcesigned
	case kAbs | kSign: k = (s ? u128(0) - x : x) & ones; break;          // forceabssigned
	case kAbs: k = ((x << 1) ^ (s ? ones : u128(0))) & ones; break;      // forcetieredabssigned
	case kSign | kFltp: k = (msb + (s ? u128(0) - m : m)) & ones; break; // forcefloatingp
	case kAbs | kSign | kFltp: k = m; break;                             // forceabsfloatingp
	case kFltp: k = s ? ones - m : m; break;                             // specialfloatingp
	default: k = ((m << 1) | (s ? u128(0) : u128(1))) & ones; break;     // forcetieredabsfloatingp
	}
	return (flags & kDesc) ? k ^ ones : k;
}

// the value bits of one key record at p: bytes 0..7 the low word, then 2 bytes (80-bit layouts)
// or 8 bytes (128-bit integers) of the high word; records may be unaligned
template <int W>
__device__ __forceinline__ u128 load_wide(const unsigned char *p)
{
	uint64_t lo;
	__builtin_memcpy(&lo, p, 8);
	if constexpr (W == 80) {
		uint16_t hi;
		__builtin_memcpy(&hi, p + 8, 2);
		return (u128(hi) << 64) | lo;
	} else {
		uint64_t hi;
		__builtin_memcpy(&hi, p + 8, 8);
		return (u128(hi) << 64) | lo;
	}
}

// key i's record: recs + i * stride, or through ptrs[i] (+ off1, or at the second level
// *(ptrs[i] + off1) + off2 + index2 * stride)
template <int W, bool PTR>
__global__ __launch_bounds__(256) void k_wide_split(const unsigned char *__restrict__ recs, uint64_t stride,
	const unsigned char *const *__restrict__ ptrs, int64_t off1, int second, int64_t off2, uint64_t index2, uint64_t n,
	unsigned flags, uint64_t *__restrict__ lo, uint64_t *__restrict__ hi)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
		const unsigned char *p;
		if constexpr (PTR) {
			p = ptrs[i] + off1;
			if (second) {
				const unsigned char *q;
				__builtin_memcpy(&q, p, sizeof(q));
				p = q + off2 + index2 * stride;
			}
		} else {
			p = recs + i * stride;
		}
		const u128 k = wide_sortkey(W, flags, load_wide<W>(p));
		lo[i] = static_cast<uint64_t>(k);
		hi[i] = static_cast<uint64_t>(k >> 64);
	}
}

__global__ __launch_bounds__(256) void k_gather_u64(const uint64_t *__restrict__ src, const uint32_t *__restrict__ perm,
	uint64_t *__restrict__ dst, uint64_t n)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) dst[i] = src[perm[i]];
}

// out record i = in record perm[i], S bytes each (16-byte records as one vector when aligned)
template <int S>
__global__ __launch_bounds__(256) void k_gather_records(const unsigned char *__restrict__ in, const uint32_t *__restrict__ perm,
	unsigned char *__restrict__ out, uint64_t n, int aligned16)
{
	const uint64_t step = static_cast<uint64_t>(gridDim.x) * blockDim.x;
	for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
		const unsigned char *s = in + static_cast<uint64_t>(perm[i]) * S;
		unsigned char *d = out + i * S;
		if (S == 16 && aligned16) {
			*reinterpret_cast<uint4 *>(d) = *reinterpret_cast<const uint4 *>(s);
		} else {
			unsigned char b[S];
			__builtin_memcpy(b, s, S);
			__builtin_memcpy(d, b, S);
		}
	}
}

inline dim3 wide_grid(uint64_t n)
{
	const uint64_t g = (n + 255) / 256;
	const uint64_t cap = 8ull * static_cast<uint64_t>(device_cu_count());
	return dim3(static_cast<unsigned>(g < 1 ? 1 : (g > cap ? cap : g)));
}

inline size_t up256(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }

// the wide-key arrays, at the end of the workspace (a larger sort's tile records, which grow from
// the start, never reach where a smaller sort left them)
struct WideArrays {
	uint64_t *lo, *hi, *kb;
	uint32_t *pa, *pb, *pc;
};

WideArrays carve_wide(void *ws, size_t ws_bytes, uint64_t n)
{
	const size_t arrays = wide_array_bytes(n);
	char *p = static_cast<char *>(ws) + ((ws_bytes - arrays) & ~static_cast<size_t>(255));
	WideArrays a;
	a.lo = reinterpret_cast<uint64_t *>(p); p += up256(n * 8);
	a.hi = reinterpret_cast<uint64_t *>(p); p += up256(n * 8);
	a.kb = reinterpret_cast<uint64_t *>(p); p += up256(n * 8);
	a.pa = reinterpret_cast<uint32_t *>(p); p += up256(n * 4);
	a.pb = reinterpret_cast<uint32_t *>(p); p += up256(n * 4);
	a.pc = reinterpret_cast<uint32_t *>(p);
	return a;
}

} // namespace

size_t wide_array_bytes(uint64_t n) { return 3 * up256(n * 8) + 3 * up256(n * 4) + 256; }

size_t wide_workspace_bytes(uint64_t n) { return up256(sort_workspace_bytes(n)) + wide_array_bytes(n); }

hipError_t wide_sort(const WideJob &job, void *ws, size_t ws_bytes, hipStream_t s)
{
	const uint64_t n = job.n;
	if (n == 0) return hipSuccess;
	if (n > 0xFFFFFFFFull || ws_bytes < wide_workspace_bytes(n)) return hipErrorInvalidValue;
	if (job.value_bits != 80 && job.value_bits != 128) return hipErrorInvalidValue;
	const WideArrays a = carve_wide(ws, ws_bytes, n);
	const auto *recs = static_cast<const unsigned char *>(job.recs);
	const auto *ptrs = reinterpret_cast<const unsigned char *const *>(job.ptrs);
	const unsigned kf = job.flags & (kAbs | kSign | kFltp | kDesc);
	if (job.value_bits == 80) {
		if (ptrs) hipLaunchKernelGGL((k_wide_split<80, true>), wide_grid(n), dim3(256), 0, s, nullptr, job.rec_bytes, ptrs, job.off1,
			job.second_level, job.off2, job.index2, n, kf, a.lo, a.hi);
		else hipLaunchKernelGGL((k_wide_split<80, false>), wide_grid(n), dim3(256), 0, s, recs, job.rec_bytes, nullptr, 0, 0, 0, 0, n,
			kf, a.lo, a.hi);
	} else {
		if (ptrs) hipLaunchKernelGGL((k_wide_split<128, true>), wide_grid(n), dim3(256), 0, s, nullptr, job.rec_bytes, ptrs, job.off1,
			job.second_level, job.off2, job.index2, n, kf, a.lo, a.hi);
		else hipLaunchKernelGGL((k_wide_split<128, false>), wide_grid(n), dim3(256), 0, s, recs, job.rec_bytes, nullptr, 0, 0, 0, 0, n,
			kf, a.lo, a.hi);
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	// stable by the low word: source order, or reversed under isrevorder; payload = source index
	SortJob p1;
	p1.key_bytes = 8; p1.flags = job.flags & kRev; p1.n = n;
	p1.in = a.lo; p1.out = a.kb; p1.buf = a.lo; p1.keys_final_any = true;
	p1.payload = true; p1.pin = nullptr; p1.pout = a.pa; p1.pbuf = a.pb;
	p1.cfg = job.cfg; p1.algo = job.algo;
	if ((e = sort_launch(p1, ws, ws_bytes - wide_array_bytes(n), s)) != hipSuccess) return e;
	hipLaunchKernelGGL(k_gather_u64, wide_grid(n), dim3(256), 0, s, a.hi, a.pa, a.kb, n);
	if ((e = hipGetLastError()) != hipSuccess) return e;
	// stable by the high word, carrying the first permutation
	SortJob p2 = p1;
	p2.flags = 0;
	p2.in = a.kb; p2.out = a.lo; p2.buf = a.kb;
	p2.pin = a.pa; p2.pout = a.pc; p2.pbuf = a.pb;
	if ((e = sort_launch(p2, ws, ws_bytes - wide_array_bytes(n), s)) != hipSuccess) return e;
	const uint32_t *perm = a.pc;
	if (job.rec_out) {
		const int al = ((reinterpret_cast<uintptr_t>(job.recs) | reinterpret_cast<uintptr_t>(job.rec_out)) & 15u) == 0;
		switch (job.rec_bytes) {
		case 10: hipLaunchKernelGGL(k_gather_records<10>, wide_grid(n), dim3(256), 0, s, recs, perm,
			static_cast<unsigned char *>(job.rec_out), n, al); break;
		case 12: hipLaunchKernelGGL(k_gather_records<12>, wide_grid(n), dim3(256), 0, s, recs, perm,
			static_cast<unsigned char *>(job.rec_out), n, al); break;
		case 16: hipLaunchKernelGGL(k_gather_records<16>, wide_grid(n), dim3(256), 0, s, recs, perm,
			static_cast<unsigned char *>(job.rec_out), n, al); break;
		default: return hipErrorInvalidValue;
		}
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	if (job.ptr_out && (e = permute_pointers(n, job.ptrs, perm, job.ptr_out, s)) != hipSuccess) return e;
	if (job.pout) {
		if (job.pin) e = permute_u32(n, job.pin, perm, job.pout, s);
		else e = hipMemcpyAsync(job.pout, perm, n * 4, hipMemcpyDeviceToDevice, s);
		if (e != hipSuccess) return e;
	}
	return hipSuccess;
}

} // namespace rsbd8hip
