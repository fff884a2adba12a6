// rsbd8_hip.hpp -- drop-in C++ front end of the MI355X rsbd8 radix sort.
//
// Same namespace, template names, template parameters and argument lists as the reference
// header's public surface (Radixsortbidi8.hpp = H):
//   rsbd8::sortingmode / rsbd8::sortingdirection                         H:212-265
//   rsbd8::radixsort, radixsortcopy (allocating, return false on failure)  H:60397-60588
//   rsbd8::radixsortnoalloc, radixsortcopynoalloc (caller-provided memory) H:60121-60395
//   indirection: T** (H:60590-60981), member object / getter function
//   (H:60983-61357) and type+offset (H:61359-61738) overloads
//   rsbd8::allocatearray / deallocatearray / buffermemorywrapper           H:59921-60117
//   rsbd8::getoffsetof                                                     H:711-720
// Sorting runs on the GPU through the C ABI of include/rsbd8_hip.h.
//
// Memory: the allocating forms accept host or device arrays (host arrays are staged through
// device memory). The noalloc forms take the arrays where they live: device arrays are sorted
// in place on the GPU; pzeroedindices must come from rsbd8::allocatearray (a zeroed device
// workspace sized for the count). allowedthreads and pfuturesplaceholder are accepted and
// ignored (the GPU grid replaces the CPU threads). Indirection over host pointer arrays reads
// every key once on the host (this is also how getter functions are evaluated), sorts the
// (key, index) pairs on the GPU and permutes the pointers; device pointer arrays with member
// objects or type+offset keys are gathered on the GPU.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include "rsbd8_hip.h"

namespace rsbd8 {

enum struct sortingmode : unsigned char {
	native = 64u,
	nativeabs = 65u,
	nativetieredabs = 66u,
	forceunsigned = 0u,
	specialsigned = forceunsigned,
	forcesigned = 1u << 1,
	forceabssigned = 1u | 1u << 1,
	forcefloatingp = 1u << 1 | 1u << 2,
	forceabsfloatingp = 1u | 1u << 1 | 1u << 2,
	specialunsigned = forceabsfloatingp,
	specialfloatingp = 1u << 2,
	forcetieredabsfloatingp = 1u | 1u << 2,
	forcetieredabssigned = 1u
};

enum struct sortingdirection : unsigned char {
	ascfwdorder = 0u,
	dscrevorder = 1u | 1u << 1,
	dscfwdorder = 1u,
	ascrevorder = 1u << 1
};

namespace helper {

// unsigned carrier and key type code of T (enums and classes sort by their unsigned bit pattern)
template<typename T>
constexpr int keytypeof()
{
	using S = std::remove_cv_t<T>;
	if constexpr (std::is_floating_point_v<S>) {
		static_assert(sizeof(S) == 4 || sizeof(S) == 8, "only 32- and 64-bit floating-point keys are supported");
		return sizeof(S) == 4 ? RSBD8_HIP_F32 : RSBD8_HIP_F64;
	} else if constexpr (std::is_signed_v<S>) {
		return sizeof(S) == 1 ? RSBD8_HIP_I8 : sizeof(S) == 2 ? RSBD8_HIP_I16 : sizeof(S) == 4 ? RSBD8_HIP_I32 : RSBD8_HIP_I64;
	} else {
		static_assert(sizeof(S) == 1 || sizeof(S) == 2 || sizeof(S) == 4 || sizeof(S) == 8, "key size must be 1, 2, 4 or 8 bytes");
		return sizeof(S) == 1 ? RSBD8_HIP_U8 : sizeof(S) == 2 ? RSBD8_HIP_U16 : sizeof(S) == 4 ? RSBD8_HIP_U32 : RSBD8_HIP_U64;
	}
}

// compile-time mode decoding for the standard key types (H:59778-59904)
template<sortingmode mode, typename T>
constexpr bool isabsvalue{(sortingmode::nativeabs <= mode && std::is_signed_v<T>) ||
	(sortingmode::native > mode && static_cast<bool>(1u & static_cast<unsigned char>(mode)))};
template<sortingmode mode, typename T>
constexpr bool issignmode{(sortingmode::native <= mode && sortingmode::nativeabs >= mode && std::is_signed_v<T>) ||
	(sortingmode::native > mode && static_cast<bool>(1u << 1 & static_cast<unsigned char>(mode)))};
template<sortingmode mode, typename T>
constexpr bool isfltpmode{(sortingmode::native <= mode && std::is_floating_point_v<T>) ||
	(sortingmode::native > mode && static_cast<bool>(1u << 2 & static_cast<unsigned char>(mode)))};
template<sortingdirection direction>
constexpr bool isdescsort{static_cast<bool>(1u & static_cast<unsigned char>(direction))};
template<sortingdirection direction>
constexpr bool isrevorderwithindirection{static_cast<bool>(1u << 1 & static_cast<unsigned char>(direction))};
template<sortingdirection direction, bool abs, bool sign, bool fltp>
constexpr bool isrevorderwithoutindirection{static_cast<bool>(1u << 1 & static_cast<unsigned char>(direction)) &&
	((abs && sign) || (!abs && sign && fltp))};

template<sortingdirection direction, sortingmode mode, typename T, bool indirect>
constexpr unsigned flags()
{
	using S = std::remove_cv_t<T>;
	constexpr bool a{isabsvalue<mode, S>}, s{issignmode<mode, S>}, f{isfltpmode<mode, S>};
	constexpr bool rev{indirect ? isrevorderwithindirection<direction> : isrevorderwithoutindirection<direction, a, s, f>};
	return (a ? RSBD8_HIP_ABS : 0u) | (s ? RSBD8_HIP_SIGN : 0u) | (f ? RSBD8_HIP_FLTP : 0u) |
		(isdescsort<direction> ? RSBD8_HIP_DESC : 0u) | (rev ? RSBD8_HIP_REV : 0u);
}

template<typename T>
using carrier = std::conditional_t<sizeof(T) == 1, std::uint8_t, std::conditional_t<sizeof(T) == 2, std::uint16_t,
	std::conditional_t<sizeof(T) == 4, std::uint32_t, std::uint64_t>>>;

inline bool isdevice(void const *p)
{
	hipPointerAttribute_t attr;
	if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
		static_cast<void>(hipGetLastError());
		return false;
	}
	return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

inline int &lasterrorcode()
{
	static thread_local int code{};
	return code;
}

// Sorts host keys (already extracted) on the GPU and returns the stable source permutation.
template<typename T>
bool sortperm(std::size_t count, unsigned flags, carrier<T> const *keys, std::uint32_t *perm)
{
	int const kt{keytypeof<T>()};
	std::size_t const ws{rsbd8_hip_workspace_size(kt, count, RSBD8_HIP_LAYOUT_PAIRS)};
	std::size_t const kb{count * sizeof(T)}, pb{count * sizeof(std::uint32_t)};
	char *mem{};
	if (hipMalloc(reinterpret_cast<void **>(&mem), ws + 3 * kb + 2 * pb + 1024) != hipSuccess) return false;
	char *pws{mem}, *k0{mem + ((ws + 255) & ~std::size_t{255})};
	char *k1{k0 + ((kb + 255) & ~std::size_t{255})}, *k2{k1 + ((kb + 255) & ~std::size_t{255})};
	char *p0{k2 + ((kb + 255) & ~std::size_t{255})}, *p1{p0 + ((pb + 255) & ~std::size_t{255})};
	bool ok{hipMemcpy(k0, keys, kb, hipMemcpyHostToDevice) == hipSuccess &&
		rsbd8_hip_workspace_init(pws, ws, nullptr) == RSBD8_HIP_OK &&
		rsbd8_hip_sort_pairs(kt, flags, count, pws, ws, k0, k1, k2, nullptr, reinterpret_cast<std::uint32_t *>(p0),
			reinterpret_cast<std::uint32_t *>(p1), nullptr) == RSBD8_HIP_OK &&
		hipMemcpy(perm, p0, pb, hipMemcpyDeviceToHost) == hipSuccess};
	static_cast<void>(hipFree(mem));
	return ok;
}

// key extraction mirroring indirectinput1/indirectinput2 (H:2745-2821) for the host path
template<typename T, std::ptrdiff_t indirection2, bool isindexed2>
carrier<T> secondlevel(void const *intermediate, std::size_t index2)
{
	carrier<T> v;
	std::memcpy(&v, static_cast<std::byte const *>(intermediate) + indirection2 + (isindexed2 ? index2 * sizeof(T) : 0u), sizeof(T));
	return v;
}

} // namespace helper

// last failure of a void noalloc call on this thread (0 if none), as an RSBD8_HIP_E* code
inline int lasterror() noexcept { return helper::lasterrorcode(); }

// ---- allocation (H:59907-60117) --------------------------------------------------------------

template<typename T>
struct combinedallocation {
	T *pbuffer;
	void *pfuturesplaceholder;
	void *pzeroedindices;
};

// one device allocation: the count-element buffer, then the zeroed workspace; pbuffer is null
// when the allocation failed
template<bool isabsvalue, bool issignmode, bool isfltpmode, typename T, bool isbufferpointers, bool isbufferallocated = true>
[[nodiscard]] inline std::pair<combinedallocation<std::conditional_t<isbufferpointers, T *, T>>, std::size_t>
allocatearray(std::size_t count, unsigned allowedthreads = 0u, int mmapflags = 0)
{
	static_cast<void>(allowedthreads);
	static_cast<void>(mmapflags);
	using E = std::conditional_t<isbufferpointers, T *, T>;
	int const kt{helper::keytypeof<std::conditional_t<isbufferpointers, std::remove_pointer_t<T>, T>>()};
	std::size_t const ws{rsbd8_hip_workspace_size(kt, count, isbufferpointers ? RSBD8_HIP_LAYOUT_INDIRECT : RSBD8_HIP_LAYOUT_KEYS)};
	std::size_t const bufbytes{isbufferallocated ? ((count * sizeof(E) + 255) & ~std::size_t{255}) : 256u};
	std::size_t const total{bufbytes + ws};
	char *mem{};
	if (hipMalloc(reinterpret_cast<void **>(&mem), total) != hipSuccess) return {{nullptr, nullptr, nullptr}, 0u};
	if (rsbd8_hip_workspace_init(mem + bufbytes, ws, nullptr) != RSBD8_HIP_OK || hipDeviceSynchronize() != hipSuccess) {
		static_cast<void>(hipFree(mem));
		return {{nullptr, nullptr, nullptr}, 0u};
	}
	return {{reinterpret_cast<E *>(mem), nullptr, mem + bufbytes}, total};
}

inline void deallocatearray(void *buffer, std::size_t allocsize = 0u) noexcept
{
	static_cast<void>(allocsize);
	static_cast<void>(hipFree(buffer));
}

struct buffermemorywrapper {
	void *ptr;
	std::size_t size;
	buffermemorywrapper(buffermemorywrapper const &) = delete;
	buffermemorywrapper &operator=(buffermemorywrapper const &) = delete;
	buffermemorywrapper(void *ptrmem, std::size_t sizemem = 0u) noexcept : ptr{ptrmem}, size{sizemem} {}
	~buffermemorywrapper() noexcept { deallocatearray(ptr, size); }
};

namespace helper {
template<typename C, typename M>
C classof(M C::*);

template<auto memberptr>
std::size_t offsetofmember() noexcept
{
	using C = decltype(classof(memberptr));
	alignas(C) static unsigned char storage[sizeof(C)];
	C const *c{reinterpret_cast<C const *>(storage)};
	return static_cast<std::size_t>(reinterpret_cast<unsigned char const *>(&(c->*memberptr)) - storage);
}
} // namespace helper

// byte offset of a data member (getoffsetof H:711-720); a run-time constant here, where the
// reference evaluates it at compile time
template<auto memberptr>
inline std::size_t const getoffsetof{helper::offsetofmember<memberptr>()};

// ---- the four templates without indirection ------------------------------------------------

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native, typename T>
inline void radixsortcopynoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder,
	void *pzeroedindices, T const *input, T *output, T *buffer = nullptr) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	int const kt{helper::keytypeof<T>()};
	helper::lasterrorcode() = rsbd8_hip_radixsortcopynoalloc(kt, helper::flags<direction, mode, T, false>(), count,
		pzeroedindices, rsbd8_hip_workspace_size(kt, count, RSBD8_HIP_LAYOUT_KEYS), input, output, buffer, nullptr);
	if (!helper::lasterrorcode() && hipStreamSynchronize(nullptr) != hipSuccess) helper::lasterrorcode() = RSBD8_HIP_EDEVICE;
}

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native, typename T>
inline void radixsortnoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices,
	T *input, T *buffer, bool movetobuffer = false) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	int const kt{helper::keytypeof<T>()};
	helper::lasterrorcode() = rsbd8_hip_radixsortnoalloc(kt, helper::flags<direction, mode, T, false>(), count, pzeroedindices,
		rsbd8_hip_workspace_size(kt, count, RSBD8_HIP_LAYOUT_KEYS), input, buffer, movetobuffer, nullptr);
	if (!helper::lasterrorcode() && hipStreamSynchronize(nullptr) != hipSuccess) helper::lasterrorcode() = RSBD8_HIP_EDEVICE;
}

// 8-bit keys without a buffer (the reference's buffer-less single-part overload, H:60272-60298)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native, typename T>
inline std::enable_if_t<1u == sizeof(T), void> radixsortnoalloc(std::size_t count, unsigned allowedthreads,
	void *pfuturesplaceholder, void *pzeroedindices, T *input) noexcept
{
	radixsortnoalloc<direction, mode, T>(count, allowedthreads, pfuturesplaceholder, pzeroedindices, input, nullptr, false);
}

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native, typename T>
[[nodiscard]] inline std::enable_if_t<!std::is_pointer_v<T>, bool> radixsort(std::size_t count, T *input, int mmapflags = 0) noexcept
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == rsbd8_hip_radixsort(helper::keytypeof<T>(), helper::flags<direction, mode, T, false>(), count, input, nullptr);
}

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native, typename T>
[[nodiscard]] inline std::enable_if_t<!std::is_pointer_v<T>, bool> radixsortcopy(std::size_t count, T const *input, T *output,
	int mmapflags = 0) noexcept
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == rsbd8_hip_radixsortcopy(helper::keytypeof<T>(), helper::flags<direction, mode, T, false>(), count,
		input, output, nullptr);
}

// ---- indirection ----------------------------------------------------------------------------

namespace helper {

// sort count pointers by extracted keys: the device path for device pointer arrays with a byte
// offset, the host path (one key read per element on the host) otherwise
template<typename K, unsigned fl, typename V, typename Extract>
bool sortpointers(std::size_t count, V *const *input, V **output, Extract &&extract, bool deviceok, std::ptrdiff_t off1,
	bool secondlevel, std::ptrdiff_t off2, std::size_t index2)
{
	if (count < 2u) {
		if (count == 1u && output != input) {
			// a device pointer array is never dereferenced on the host
			if (isdevice(input) || isdevice(output))
				return hipMemcpy(output, input, sizeof(V *), hipMemcpyDefault) == hipSuccess;
			*output = *input;
		}
		return true;
	}
	if (deviceok && isdevice(input)) {
		int const kt{keytypeof<K>()};
		std::size_t const ws{rsbd8_hip_workspace_size(kt, count, RSBD8_HIP_LAYOUT_INDIRECT)};
		void *mem{};
		if (hipMalloc(&mem, ws) != hipSuccess) return false;
		bool ok{rsbd8_hip_workspace_init(mem, ws, nullptr) == RSBD8_HIP_OK &&
			rsbd8_hip_radixsortcopynoalloc_indirect(kt, fl, count, mem, ws, static_cast<void const *const *>(static_cast<void const *>(input)),
				static_cast<void **>(static_cast<void *>(output)), off1, secondlevel, off2, index2, nullptr) == RSBD8_HIP_OK &&
			hipDeviceSynchronize() == hipSuccess};
		static_cast<void>(hipFree(mem));
		return ok;
	}
	std::vector<carrier<K>> keys(count);
	std::vector<std::uint32_t> perm(count);
	for (std::size_t i{}; i != count; ++i) keys[i] = extract(input[i]);
	if (!sortperm<K>(count, fl, keys.data(), perm.data())) return false;
	std::vector<V *> src(input, input + count);
	for (std::size_t i{}; i != count; ++i) output[i] = src[perm[i]];
	return true;
}

template<typename T>
carrier<T> bitsof(T const &v)
{
	carrier<T> u;
	std::memcpy(&u, &v, sizeof(T));
	return u;
}

} // namespace helper

// T** (second-level pointers to the keys themselves, H:60590-60981): key = input[i][index]
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_pointer_v<T>, bool> radixsortcopy(std::size_t count, T *const *input, T **output,
	int mmapflags = 0, vararguments... index) noexcept
{
	static_cast<void>(mmapflags);
	static_assert(sizeof...(index) <= 1u, "T** indirection takes at most one index argument");
	std::size_t idx{};
	((idx = static_cast<std::size_t>(index)), ...);
	using K = std::remove_cv_t<T>;
	return helper::sortpointers<K, helper::flags<direction, mode, K, true>()>(count, input, output,
		[idx](T *p) { return helper::bitsof<K>(p[idx]); }, true, static_cast<std::ptrdiff_t>(idx * sizeof(T)), false, 0, 0u);
}

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_pointer_v<T>, bool> radixsort(std::size_t count, T **input, int mmapflags = 0,
	vararguments... index) noexcept
{
	return radixsortcopy<direction, mode, indirection2, isindexed2>(count, static_cast<T *const *>(input), input, mmapflags, index...);
}

// member object or getter function (H:60983-61357)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)>, bool> radixsortcopy(std::size_t count,
	V *const *input, V **output, int mmapflags = 0, vararguments... args)
{
	static_cast<void>(mmapflags);
	if constexpr (std::is_member_object_pointer_v<decltype(indirection1)>) {
		using M = std::remove_cv_t<std::remove_reference_t<decltype(std::declval<V &>().*indirection1)>>;
		std::ptrdiff_t const off{static_cast<std::ptrdiff_t>(helper::offsetofmember<indirection1>())};
		if constexpr (!std::is_pointer_v<M>) {
			using K = M;
			return helper::sortpointers<K, helper::flags<direction, mode, K, true>()>(count, input, output,
				[](V *p) { return helper::bitsof<K>(p->*indirection1); }, true, off, false, 0, 0u);
		} else {
			using K = std::remove_cv_t<std::remove_pointer_t<M>>;
			std::size_t idx{};
			((idx = static_cast<std::size_t>(args)), ...);
			return helper::sortpointers<K, helper::flags<direction, mode, K, true>()>(count, input, output,
				[idx](V *p) { return helper::secondlevel<K, indirection2, isindexed2>(p->*indirection1, idx); }, true, off, true,
				indirection2, isindexed2 ? idx : 0u);
		}
	} else {
		using R = std::remove_cv_t<std::remove_reference_t<decltype((std::declval<V &>().*indirection1)(args...))>>;
		if constexpr (!std::is_pointer_v<R>) {
			return helper::sortpointers<R, helper::flags<direction, mode, R, true>()>(count, input, output,
				[&args...](V *p) { return helper::bitsof<R>((p->*indirection1)(args...)); }, false, 0, false, 0, 0u);
		} else {
			using K = std::remove_cv_t<std::remove_pointer_t<R>>;
			return helper::sortpointers<K, helper::flags<direction, mode, K, true>()>(count, input, output,
				[&args...](V *p) { return helper::secondlevel<K, indirection2, false>((p->*indirection1)(args...), 0u); }, false,
				0, false, 0, 0u);
		}
	}
}

template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)>, bool> radixsort(std::size_t count,
	V **input, int mmapflags = 0, vararguments... args)
{
	return radixsortcopy<indirection1, direction, mode, indirection2, isindexed2>(count, static_cast<V *const *>(input), input,
		mmapflags, args...);
}

// type + byte offset (H:61359-61738): key = *(T *)((std::byte *)input[i] + indirection1); with T a
// pointer type the key is ((U *)(*(std::byte **)(...) + indirection2))[index2]
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_member_pointer_v<T>, bool> radixsortcopy(std::size_t count, V *const *input,
	V **output, int mmapflags = 0, vararguments... index) noexcept
{
	static_cast<void>(mmapflags);
	std::size_t idx{};
	((idx = static_cast<std::size_t>(index)), ...);
	if constexpr (!std::is_pointer_v<T>) {
		using K = std::remove_cv_t<T>;
		return helper::sortpointers<K, helper::flags<direction, mode, K, true>()>(count, input, output,
			[](V *p) { return helper::secondlevel<K, indirection1, false>(p, 0u); }, true, indirection1, false, 0, 0u);
	} else {
		using K = std::remove_cv_t<std::remove_pointer_t<T>>;
		return helper::sortpointers<K, helper::flags<direction, mode, K, true>()>(count, input, output,
			[idx](V *p) {
				void const *q;
				std::memcpy(&q, reinterpret_cast<std::byte const *>(p) + indirection1, sizeof(q));
				return helper::secondlevel<K, indirection2, isindexed2>(q, idx);
			},
			true, indirection1, true, indirection2, isindexed2 ? idx : 0u);
	}
}

template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_member_pointer_v<T>, bool> radixsort(std::size_t count, V **input,
	int mmapflags = 0, vararguments... index) noexcept
{
	return radixsortcopy<T, indirection1, direction, mode, indirection2, isindexed2>(count, static_cast<V *const *>(input), input,
		mmapflags, index...);
}

} // namespace rsbd8
