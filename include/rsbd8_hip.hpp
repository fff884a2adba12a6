// rsbd8_hip.hpp -- drop-in C++ front end of the MI355X rsbd8 radix sort.
//
// Same namespace, template names, template parameters and argument lists as the reference
// header's public surface (Radixsortbidi8.hpp = H):
//   rsbd8::sortingmode / rsbd8::sortingdirection                         H:212-265
//   rsbd8::radixsort, radixsortcopy (allocating, return false on failure)  H:60397-60588
//   rsbd8::radixsortnoalloc, radixsortcopynoalloc (caller-provided memory) H:60121-60395
//   indirection, allocating and noalloc forms: T** (H:60590-60981), member object / getter
//   function (H:60983-61357) and type+offset (H:61359-61738)
//   rsbd8::allocatearray / deallocatearray / buffermemorywrapper           H:59921-60117
//   rsbd8::memberobjectgenerator / getoffsetof (compile time)              H:656-720
// Sorting runs on the GPU through the C ABI of include/rsbd8_hip.h.
//
// Memory: the allocating forms accept host or device arrays (host arrays are staged through
// device memory). The noalloc forms without indirection take device arrays only (a host array
// sets lasterror() to RSBD8_HIP_EINVAL and nothing runs); pzeroedindices must come from
// rsbd8::allocatearray (a zeroed device workspace sized for the count). allowedthreads and
// pfuturesplaceholder are accepted and ignored (the GPU grid replaces the CPU threads).
// Indirection: device pointer arrays whose keys are member objects, type+offset keys or T**
// elements are gathered on the GPU, and the noalloc forms then allocate nothing. Host pointer
// arrays, and getter functions (which only the host can call), read every key once on the host,
// sort the (key, index) pairs on the GPU through a temporary device allocation and permute the
// pointers on the host; a getter over a device pointer array is rejected (RSBD8_HIP_EINVAL).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
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

// A packed record holding a T at byte offset indirection1 (H:656-668): the type+offset forms sort by
// the member `object` of this record, and users name offsets with it.
#pragma pack(push, 1)
template<typename T, std::ptrdiff_t indirection1>
struct memberobjectgenerator;
template<typename T>
struct memberobjectgenerator<T, 0> {
	T object;
};
template<typename T, std::ptrdiff_t indirection1>
struct memberobjectgenerator {
	std::byte padding[static_cast<std::size_t>(indirection1)];
	T object;
};
#pragma pack(pop)

namespace helper {

// The reference's record types for keys wider than 64 bits (H:1061-1804, H:2186-2404): the 80-bit
// extended float as a 64-bit mantissa plus sign and exponent in the low 16 bits of signexponent
// (10-, 12- and 16-byte layouts), and a 128-bit integer, low word first. The template arguments
// are the isabsvalue, issignmode and isfltpmode that sorting them implies in the native modes.
template<bool isabsvalue, bool issignmode, bool isfltpmode>
struct longdoubletest128 {
	std::uint_least64_t mantissa;
	std::uint_least64_t signexponent;
};
#pragma pack(push, 1)
template<bool isabsvalue, bool issignmode, bool isfltpmode>
struct longdoubletest96 {
	std::uint_least64_t mantissa;
	std::uint_least32_t signexponent;
};
template<bool isabsvalue, bool issignmode, bool isfltpmode>
struct longdoubletest80 {
	std::uint_least64_t mantissa;
	std::uint_least16_t signexponent;
};
#pragma pack(pop)
template<bool isabsvalue, bool issignmode, bool isfltpmode>
struct test128 {
	std::uint_least64_t data[2];
};

// the wide key kind of T: its key type code and the flags its record type implies (0 code: not wide)
template<typename T>
struct widekind {
	static constexpr int code{0};
	static constexpr bool abs{false}, sign{false}, fltp{false};
};
template<template<bool, bool, bool> class R, int C, bool a, bool s, bool f>
struct widekindof {
	static constexpr int code{C};
	static constexpr bool abs{a}, sign{s}, fltp{f};
};
template<bool a, bool s, bool f>
struct widekind<longdoubletest80<a, s, f>> : widekindof<longdoubletest80, RSBD8_HIP_LD80, a, s, f> {};
template<bool a, bool s, bool f>
struct widekind<longdoubletest96<a, s, f>> : widekindof<longdoubletest96, RSBD8_HIP_LD96, a, s, f> {};
template<bool a, bool s, bool f>
struct widekind<longdoubletest128<a, s, f>> : widekindof<longdoubletest128, RSBD8_HIP_LD128, a, s, f> {};
template<bool a, bool s, bool f>
struct widekind<test128<a, s, f>> : widekindof<test128, RSBD8_HIP_U128, a, s, f> {};
// x86 long double: the 80-bit extended float in 16 bytes (signed, floating point by the type traits)
template<>
struct widekind<long double> {
	static constexpr int code{sizeof(long double) == 16 && 64 == std::numeric_limits<long double>::digits ? RSBD8_HIP_LD128 : 0};
	static constexpr bool abs{false}, sign{false}, fltp{false};
};
#ifdef __SIZEOF_INT128__
template<>
struct widekind<unsigned __int128> {
	static constexpr int code{RSBD8_HIP_U128};
	static constexpr bool abs{false}, sign{false}, fltp{false};
};
template<>
struct widekind<__int128> {
	static constexpr int code{RSBD8_HIP_I128};
	static constexpr bool abs{false}, sign{false}, fltp{false};
};
#endif

// std::is_signed_v, also for __int128 where the library leaves it out (strict standard modes)
template<typename T>
constexpr bool issigned{std::is_signed_v<T> || widekind<T>::code == RSBD8_HIP_I128};

// unsigned carrier and key type code of T (enums and classes sort by their unsigned bit pattern)
template<typename T>
constexpr int keytypeof()
{
	using S = std::remove_cv_t<T>;
	if constexpr (widekind<S>::code != 0) {
		return widekind<S>::code;
	} else if constexpr (std::is_floating_point_v<S>) {
		static_assert(sizeof(S) == 4 || sizeof(S) == 8, "floating-point keys: 32 or 64 bits, or the x86 80-bit long double");
		return sizeof(S) == 4 ? RSBD8_HIP_F32 : RSBD8_HIP_F64;
	} else if constexpr (std::is_signed_v<S>) {
		return sizeof(S) == 1 ? RSBD8_HIP_I8 : sizeof(S) == 2 ? RSBD8_HIP_I16 : sizeof(S) == 4 ? RSBD8_HIP_I32 : RSBD8_HIP_I64;
	} else {
		static_assert(sizeof(S) == 1 || sizeof(S) == 2 || sizeof(S) == 4 || sizeof(S) == 8,
			"key size must be 1, 2, 4 or 8 bytes, or one of the 80-bit and 128-bit types");
		return sizeof(S) == 1 ? RSBD8_HIP_U8 : sizeof(S) == 2 ? RSBD8_HIP_U16 : sizeof(S) == 4 ? RSBD8_HIP_U32 : RSBD8_HIP_U64;
	}
}

// compile-time mode decoding (H:59778-59904); the wide record types add the flags they carry
template<sortingmode mode, typename T>
constexpr bool isabsvalue{(sortingmode::nativeabs <= mode && issigned<T>) || widekind<T>::abs ||
	(sortingmode::native > mode && static_cast<bool>(1u & static_cast<unsigned char>(mode)))};
template<sortingmode mode, typename T>
constexpr bool issignmode{(sortingmode::native <= mode && sortingmode::nativeabs >= mode && issigned<T>) || widekind<T>::sign ||
	(sortingmode::native > mode && static_cast<bool>(1u << 1 & static_cast<unsigned char>(mode)))};
template<sortingmode mode, typename T>
constexpr bool isfltpmode{(sortingmode::native <= mode && std::is_floating_point_v<T>) || widekind<T>::fltp ||
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

// the bytes of a key as the C ABI takes them: an unsigned integer of the key's size, or the record
// bytes of a wide key
template<std::size_t N>
struct widebytes {
	unsigned char b[N];
};
template<typename T>
using carrier = std::conditional_t<(sizeof(T) > 8), widebytes<sizeof(T)>,
	std::conditional_t<sizeof(T) == 1, std::uint8_t, std::conditional_t<sizeof(T) == 2, std::uint16_t,
	std::conditional_t<sizeof(T) == 4, std::uint32_t, std::uint64_t>>>>;

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

// Workspaces handed out by allocatearray, with their sizes: the indirect sorts place their gathered
// keys at the end of the workspace, so they need its real size, which the noalloc signatures do not
// carry.
struct wsrecord {
	void const *base;
	void const *workspace;
	std::size_t bytes;
};
inline std::mutex &wsmutex()
{
	static std::mutex m;
	return m;
}
inline std::vector<wsrecord> &wsrecords()
{
	static std::vector<wsrecord> v;
	return v;
}
inline bool wsremember(void const *base, void const *workspace, std::size_t bytes) noexcept
{
	try {
		std::lock_guard<std::mutex> lock(wsmutex());
		wsrecords().push_back({base, workspace, bytes});
		return true;
	} catch (...) {
		return false;
	}
}
inline void wsforget(void const *base) noexcept
{
	std::lock_guard<std::mutex> lock(wsmutex());
	auto &v{wsrecords()};
	for (std::size_t i{}; i < v.size(); ++i)
		if (v[i].base == base) {
			v[i] = v.back();
			v.pop_back();
			return;
		}
}
// the size allocatearray gave workspace, else the size for this count
inline std::size_t wsbytes(void const *workspace, std::size_t forcount) noexcept
{
	std::lock_guard<std::mutex> lock(wsmutex());
	for (auto const &r : wsrecords())
		if (r.workspace == workspace) return r.bytes;
	return forcount;
}

// the byte offset of a data member at compile time: a union lays an M at byte offset `at` over a C,
// and the member's address inside the C compares below, equal to or above the M's, so a binary
// search over the offsets [0, sizeof(C)) finds it
template<typename C, typename M>
C classof(M C::*);
template<typename C, typename M>
M memberof(M C::*);

template<typename C, typename M, std::size_t at>
struct offsetprobe {
	union overlay {
		char none;
		C object;
		memberobjectgenerator<M, static_cast<std::ptrdiff_t>(at)> placed;
	};
	static constexpr overlay probe{};
};

template<auto memberptr, typename C, typename M, std::size_t lo, std::size_t hi>
constexpr std::size_t offsetsearch() noexcept
{
	if constexpr (hi - lo <= 1u) {
		return lo;
	} else {
		constexpr std::size_t mid{lo + (hi - lo) / 2u};
		using P = offsetprobe<C, M, mid>;
		if constexpr (&(P::probe.object.*memberptr) < &P::probe.placed.object) return offsetsearch<memberptr, C, M, lo, mid>();
		else return offsetsearch<memberptr, C, M, mid, hi>();
	}
}

} // namespace helper

// byte offset of a data member, usable in constant expressions (getoffsetof H:711-720)
template<auto memberptr>
constexpr std::size_t getoffsetof{helper::offsetsearch<memberptr, decltype(helper::classof(memberptr)),
	std::remove_cv_t<decltype(helper::memberof(memberptr))>, 0u, sizeof(decltype(helper::classof(memberptr)))>()};

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
	if (rsbd8_hip_workspace_init(mem + bufbytes, ws, nullptr) != RSBD8_HIP_OK || hipDeviceSynchronize() != hipSuccess ||
		!helper::wsremember(mem, mem + bufbytes, ws)) {
		static_cast<void>(hipFree(mem));
		return {{nullptr, nullptr, nullptr}, 0u};
	}
	return {{reinterpret_cast<E *>(mem), nullptr, mem + bufbytes}, total};
}

inline void deallocatearray(void *buffer, std::size_t allocsize = 0u) noexcept
{
	static_cast<void>(allocsize);
	helper::wsforget(buffer);
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

// ---- the four templates without indirection ------------------------------------------------

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native, typename T>
inline std::enable_if_t<!std::is_pointer_v<T>, void> radixsortcopynoalloc(std::size_t count, unsigned allowedthreads,
	void *pfuturesplaceholder, void *pzeroedindices, T const *input, T *output, T *buffer = nullptr) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	int const kt{helper::keytypeof<T>()};
	helper::lasterrorcode() = rsbd8_hip_radixsortcopynoalloc(kt, helper::flags<direction, mode, T, false>(), count,
		pzeroedindices, rsbd8_hip_workspace_size(kt, count, RSBD8_HIP_LAYOUT_KEYS), input, output, buffer, nullptr);
	if (!helper::lasterrorcode() && hipStreamSynchronize(nullptr) != hipSuccess) helper::lasterrorcode() = RSBD8_HIP_EDEVICE;
}

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native, typename T>
inline std::enable_if_t<!std::is_pointer_v<T>, void> radixsortnoalloc(std::size_t count, unsigned allowedthreads,
	void *pfuturesplaceholder, void *pzeroedindices, T *input, T *buffer, bool movetobuffer = false) noexcept
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
inline std::enable_if_t<!std::is_pointer_v<T> && 1u == sizeof(T), void> radixsortnoalloc(std::size_t count,
	unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, T *input) noexcept
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
//
// Where the key of element p (a V *) sits, as the reference reads it (indirectinput1/2 and
// splitget, H:2405-2821), with off(m) the member's byte offset:
//   member object m of a non-pointer type M: *(M *)((byte *)p + off(m) + index1 * sizeof(M))
//   member object m of a pointer type U *:   q = *(U **)((byte *)p + off(m) + index1 * sizeof(U *)),
//                                           key = ((U *)((byte *)q + indirection2))[index2]
//     (no argument: index1 = index2 = 0; one: index2 when isindexed2, else index1; two: index1, index2)
//   getter f returning a value:              (p->*f)(args...)
//   getter f returning U * or U &:           ((U *)((byte *)q + indirection2))[index2] with q the returned
//                                           pointer (or the referenced object's address); when isindexed2
//                                           the first argument is index2 and the rest go to the getter
//   type+offset <T, indirection1>:          the member `object` of memberobjectgenerator<T, indirection1>
//   T**:                                     the member `object` of memberobjectgenerator<T, 0>: p[index1]
// The key type is that of the value read; isrevorder is honoured as with every indirection
// (isrevorderwithindirection H:59872-59875).

namespace helper {

enum class sortop { copy, inplace, tobuffer };

// a parameter that takes part in no template argument deduction (the pointer arrays may be nullptr)
template<typename T>
struct identity {
	using type = T;
};
template<typename T>
using nodeduce = typename identity<T>::type;

// the GPU gather of one key: first level at p + off1, second level at *(p + off1) + off2 + index2 * sizeof(K)
struct deviceaccess {
	std::ptrdiff_t off1;
	bool second;
	std::ptrdiff_t off2;
	std::size_t index2;
};

template<typename T>
carrier<T> bitsof(T const &v) noexcept
{
	carrier<T> u;
	std::memcpy(&u, &v, sizeof(T));
	return u;
}

// the K at byte address p (records may be packed: no alignment assumed)
template<typename K>
carrier<K> loadkey(void const *p) noexcept
{
	carrier<K> u;
	std::memcpy(&u, p, sizeof(K));
	return u;
}

// the same access as the GPU gather, on the host
template<typename K>
carrier<K> readaccess(void const *p, deviceaccess const &a) noexcept
{
	std::byte const *b{static_cast<std::byte const *>(p) + a.off1};
	if (!a.second) return loadkey<K>(b);
	std::byte const *q;
	std::memcpy(&q, b, sizeof(q));
	return loadkey<K>(q + a.off2 + static_cast<std::ptrdiff_t>(a.index2 * sizeof(K)));
}

// byte offset of a data member for any class (the constant-expression getoffsetof needs a literal
// class): the member's address in an unused, suitably aligned storage area
template<auto memberptr>
std::ptrdiff_t memberoffset() noexcept
{
	using C = decltype(classof(memberptr));
	alignas(C) static unsigned char storage[sizeof(C)];
	C const *c{reinterpret_cast<C const *>(storage)};
	return reinterpret_cast<unsigned char const *>(&(c->*memberptr)) - storage;
}

// Sorts host keys (already extracted) on the GPU and returns the stable source permutation.
template<typename K>
bool sortperm(std::size_t count, unsigned flags, carrier<K> const *keys, std::uint32_t *perm) noexcept
{
	int const kt{keytypeof<K>()};
	std::size_t const ws{rsbd8_hip_workspace_size(kt, count, RSBD8_HIP_LAYOUT_PAIRS)};
	std::size_t const kb{count * sizeof(K)}, pb{count * sizeof(std::uint32_t)};
	auto up{[](std::size_t v) { return (v + 255) & ~std::size_t{255}; }};
	char *mem{};
	if (hipMalloc(reinterpret_cast<void **>(&mem), up(ws) + 3 * up(kb) + 2 * up(pb)) != hipSuccess) return false;
	char *pws{mem}, *k0{mem + up(ws)};
	char *k1{k0 + up(kb)}, *k2{k1 + up(kb)};
	char *p0{k2 + up(kb)}, *p1{p0 + up(pb)};
	bool ok{hipMemcpy(k0, keys, kb, hipMemcpyHostToDevice) == hipSuccess &&
		rsbd8_hip_workspace_init(pws, ws, nullptr) == RSBD8_HIP_OK &&
		rsbd8_hip_sort_pairs(kt, flags, count, pws, ws, k0, k1, k2, nullptr, reinterpret_cast<std::uint32_t *>(p0),
			reinterpret_cast<std::uint32_t *>(p1), nullptr) == RSBD8_HIP_OK &&
		hipMemcpy(perm, p0, pb, hipMemcpyDeviceToHost) == hipSuccess};
	static_cast<void>(hipFree(mem));
	return ok;
}

// Sorts the count pointers of input by their keys: into output (copy), back into input (inplace,
// through buffer when one is given) or into buffer (tobuffer). pzeroed null: the allocating forms
// (a temporary workspace); otherwise the caller's allocatearray workspace. access null: only the
// host can read the keys (getters). Returns an RSBD8_HIP_* code.
template<typename K, unsigned fl, typename V, typename X>
int sortpointers(sortop op, std::size_t count, void *pzeroed, V *const *input, nodeduce<V **> output, nodeduce<V **> buffer, X const &extract,
	deviceaccess const *access) noexcept
{
	V **dst{op == sortop::copy ? output : (op == sortop::tobuffer ? buffer : const_cast<V **>(input))};
	if (!input || !dst) return RSBD8_HIP_EINVAL;
	bool const devin{isdevice(input)};
	if (count < 2u) {
		if (count == 1u && dst != input) {
			// a device pointer array is never dereferenced on the host
			if (devin || isdevice(dst))
				return hipMemcpy(dst, input, sizeof(V *), hipMemcpyDefault) == hipSuccess ? RSBD8_HIP_OK : RSBD8_HIP_EDEVICE;
			*dst = *input;
		}
		return RSBD8_HIP_OK;
	}
	int const kt{keytypeof<K>()};
	if (devin) {
		if (!access) return RSBD8_HIP_EINVAL;  // a getter cannot run on device memory
		std::size_t const need{rsbd8_hip_workspace_size(kt, count, RSBD8_HIP_LAYOUT_INDIRECT)};
		void *ws{pzeroed};
		std::size_t const wsb{pzeroed ? wsbytes(pzeroed, need) : need};
		V **tmp{buffer};
		char *mem{};
		if (!pzeroed || (op == sortop::inplace && !tmp)) {
			// the allocating forms: a workspace, and for in-place sorts a pointer buffer
			std::size_t const wpart{pzeroed ? 0u : ((need + 255) & ~std::size_t{255})};
			if (hipMalloc(reinterpret_cast<void **>(&mem), wpart + count * sizeof(V *)) != hipSuccess) return RSBD8_HIP_ENOMEM;
			if (!pzeroed) {
				ws = mem;
				if (rsbd8_hip_workspace_init(ws, wsb, nullptr) != RSBD8_HIP_OK) {
					static_cast<void>(hipFree(mem));
					return RSBD8_HIP_EDEVICE;
				}
			}
			if (!tmp) tmp = reinterpret_cast<V **>(mem + wpart);
		}
		int r;
		if (op == sortop::copy)
			r = rsbd8_hip_radixsortcopynoalloc_indirect(kt, fl, count, ws, wsb,
				static_cast<void const *const *>(static_cast<void const *>(input)), static_cast<void **>(static_cast<void *>(dst)),
				access->off1, access->second, access->off2, access->index2, nullptr);
		else
			r = rsbd8_hip_radixsortnoalloc_indirect(kt, fl, count, ws, wsb,
				static_cast<void **>(static_cast<void *>(const_cast<V **>(input))), static_cast<void **>(static_cast<void *>(tmp)),
				op == sortop::tobuffer, access->off1, access->second, access->off2, access->index2, nullptr);
		if (r == RSBD8_HIP_OK && hipStreamSynchronize(nullptr) != hipSuccess) r = RSBD8_HIP_EDEVICE;
		if (mem) static_cast<void>(hipFree(mem));
		return r;
	}
	try {
		std::vector<carrier<K>> keys(count);
		std::vector<std::uint32_t> perm(count);
		for (std::size_t i{}; i != count; ++i) keys[i] = extract(input[i]);
		if (!sortperm<K>(count, fl, keys.data(), perm.data())) return RSBD8_HIP_EDEVICE;
		if (op == sortop::inplace) {
			std::vector<V *> src(input, input + count);
			for (std::size_t i{}; i != count; ++i) dst[i] = src[perm[i]];
		} else {
			for (std::size_t i{}; i != count; ++i) dst[i] = input[perm[i]];
		}
	} catch (...) {
		return RSBD8_HIP_ENOMEM;
	}
	return RSBD8_HIP_OK;
}

// ---- key types (alias templates and declarations only, so that an unsuitable argument list is a
// substitution failure that removes the overload)

template<typename R>
using valueorpointee = std::remove_cv_t<std::conditional_t<std::is_pointer_v<R>, std::remove_pointer_t<R>, std::remove_reference_t<R>>>;
template<auto m, typename V>
using membertype = std::remove_cv_t<std::remove_reference_t<decltype(std::declval<V &>().*m)>>;

template<auto m, bool isindexed2, typename V, typename... A>
auto memberkeyof(A &&...) -> std::enable_if_t<std::is_member_object_pointer_v<decltype(m)>, valueorpointee<membertype<m, V>>>;
template<auto g, bool isindexed2, typename V, typename... A>
auto memberkeyof(A &&...) -> std::enable_if_t<std::is_member_function_pointer_v<decltype(g)> && !isindexed2,
	valueorpointee<std::invoke_result_t<decltype(g), V *, A...>>>;
template<auto g, bool isindexed2, typename V, typename A0, typename... A>
auto memberkeyof(A0 &&, A &&...) -> std::enable_if_t<std::is_member_function_pointer_v<decltype(g)> && isindexed2,
	valueorpointee<std::invoke_result_t<decltype(g), V *, A...>>>;

// the key type of the member object / getter forms for these trailing arguments
template<auto indirection1, bool isindexed2, typename V, typename... A>
using memberkey = decltype(memberkeyof<indirection1, isindexed2, V>(std::declval<A>()...));

template<typename... A>
struct firstarg {
	using type = void;
};
template<typename A0, typename... A>
struct firstarg<A0, A...> {
	using type = std::decay_t<A0>;
};
template<typename... A>
using firstarg_t = typename firstarg<A...>::type;

// ---- the three families

// a data member of type M at byte offset off (member objects, type+offset keys)
template<typename M, std::ptrdiff_t indirection2, bool isindexed2, typename... A>
deviceaccess memberaccess(std::ptrdiff_t off, A const &...a) noexcept
{
	constexpr std::size_t n{sizeof...(A)};
	static_assert(!std::is_pointer_v<std::remove_pointer_t<M>>, "third level indirection is not supported");
	static_assert(n <= 2u, "at most two index arguments (first and second level)");
	static_assert(std::is_pointer_v<M> || (!isindexed2 && n <= 1u),
		"only first-level indirection used: one index at most, and isindexed2 does not apply");
	static_assert(n != 2u || isindexed2, "two index arguments need isindexed2");
	static_assert(n != 0u || !isindexed2, "isindexed2 needs an index argument");
	std::ptrdiff_t const v[n + 1u]{static_cast<std::ptrdiff_t>(a)..., 0};
	// the first-level index displaces the record in units of the member's type (H:2755-2758)
	std::ptrdiff_t const i1{(n == 2u || (n == 1u && !isindexed2)) ? v[0] : 0};
	std::size_t const i2{isindexed2 ? static_cast<std::size_t>(v[n - 1u]) : 0u};
	return deviceaccess{off + i1 * static_cast<std::ptrdiff_t>(sizeof(M)), std::is_pointer_v<M>, indirection2, i2};
}

template<typename M, sortingdirection direction, sortingmode mode, std::ptrdiff_t indirection2, bool isindexed2,
	typename V, typename... A>
int sortbymember(sortop op, std::size_t count, void *pzeroed, V *const *input, nodeduce<V **> output, nodeduce<V **> buffer, std::ptrdiff_t off,
	A const &...a) noexcept
{
	using K = valueorpointee<M>;
	deviceaccess const acc{memberaccess<M, indirection2, isindexed2>(off, a...)};
	return sortpointers<K, flags<direction, mode, K, true>()>(op, count, pzeroed, input, output, buffer,
		[&acc](V *p) { return readaccess<K>(p, acc); }, &acc);
}

template<auto g, std::ptrdiff_t indirection2, typename K, typename V, typename... A>
carrier<K> getterkey(V *p, std::size_t index2, A const &...a)
{
	using R = decltype((p->*g)(a...));
	if constexpr (!std::is_pointer_v<R> && !std::is_lvalue_reference_v<R>) {
		return bitsof<K>((p->*g)(a...));
	} else {
		void const *q;
		if constexpr (std::is_pointer_v<R>) q = (p->*g)(a...);
		else q = &(p->*g)(a...);  // a reference is read as a pointer to its object
		return loadkey<K>(static_cast<std::byte const *>(q) + indirection2 + static_cast<std::ptrdiff_t>(index2 * sizeof(K)));
	}
}

// isindexed2: the first argument is the second-level index, the others go to the getter
template<auto g, sortingdirection direction, sortingmode mode, std::ptrdiff_t indirection2, typename V, typename I,
	typename... A>
int sortbygetterindexed(sortop op, std::size_t count, void *pzeroed, V *const *input, nodeduce<V **> output, nodeduce<V **> buffer, I const &index2,
	A const &...a)
{
	using R = std::invoke_result_t<decltype(g), V *, A const &...>;
	static_assert(std::is_pointer_v<R> || std::is_lvalue_reference_v<R>,
		"a getter returning a value takes no second-level index (isindexed2)");
	using K = valueorpointee<R>;
	std::size_t const i2{static_cast<std::size_t>(static_cast<std::ptrdiff_t>(index2))};
	return sortpointers<K, flags<direction, mode, K, true>()>(op, count, pzeroed, input, output, buffer,
		[&](V *p) { return getterkey<g, indirection2, K>(p, i2, a...); }, nullptr);
}

template<auto g, sortingdirection direction, sortingmode mode, std::ptrdiff_t indirection2, bool isindexed2, typename V,
	typename... A>
int sortbygetter(sortop op, std::size_t count, void *pzeroed, V *const *input, nodeduce<V **> output, nodeduce<V **> buffer, A const &...a)
{
	if constexpr (isindexed2) {
		static_assert(0u < sizeof...(A), "isindexed2 needs an index argument");
		return sortbygetterindexed<g, direction, mode, indirection2>(op, count, pzeroed, input, output, buffer, a...);
	} else {
		using K = valueorpointee<std::invoke_result_t<decltype(g), V *, A const &...>>;
		return sortpointers<K, flags<direction, mode, K, true>()>(op, count, pzeroed, input, output, buffer,
			[&](V *p) { return getterkey<g, indirection2, K>(p, 0u, a...); }, nullptr);
	}
}

template<auto indirection1, sortingdirection direction, sortingmode mode, std::ptrdiff_t indirection2, bool isindexed2,
	typename V, typename... A>
int sortbymemberpointer(sortop op, std::size_t count, void *pzeroed, V *const *input, nodeduce<V **> output, nodeduce<V **> buffer, A const &...a)
{
	if constexpr (std::is_member_object_pointer_v<decltype(indirection1)>)
		return sortbymember<membertype<indirection1, V>, direction, mode, indirection2, isindexed2>(op, count, pzeroed, input,
			output, buffer, memberoffset<indirection1>(), a...);
	else
		return sortbygetter<indirection1, direction, mode, indirection2, isindexed2>(op, count, pzeroed, input, output, buffer,
			a...);
}

inline void setlast(int r) noexcept { lasterrorcode() = r; }

} // namespace helper

// ---- T** (H:60590-60981): key = input[i][index] ----------------------------------------------

// multi-part radixsortcopynoalloc with simple second-level indirection (H:60590-60626)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
inline std::enable_if_t<!std::is_pointer_v<T> && 1u < sizeof(T), void> radixsortcopynoalloc(std::size_t count,
	unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, T *const *input, T **output, T **buffer,
	vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	static_cast<void>(buffer);  // the gather needs no pointer buffer for a copy
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::copy, count,
		pzeroedindices, input, output, nullptr, 0, varparameters...));
}

// multi-part radixsortnoalloc with simple second-level indirection (H:60628-60661)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
inline std::enable_if_t<!std::is_pointer_v<T> && 1u < sizeof(T), void> radixsortnoalloc(std::size_t count,
	unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, T **input, T **buffer, bool movetobuffer = false,
	vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(
		movetobuffer ? helper::sortop::tobuffer : helper::sortop::inplace, count, pzeroedindices, input, nullptr, buffer, 0,
		varparameters...));
}

// single-part radixsortcopynoalloc with simple second-level indirection (H:60663-60695)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
inline std::enable_if_t<!std::is_pointer_v<T> && 1u == sizeof(T) && !std::is_same_v<helper::firstarg_t<vararguments...>, T **>,
	void>
radixsortcopynoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices,
	T *const *input, T **output, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::copy, count,
		pzeroedindices, input, output, nullptr, 0, varparameters...));
}

// single-part radixsortcopynoalloc with a dummy buffer argument (H:60697-60723)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
inline std::enable_if_t<!std::is_pointer_v<T> && 1u == sizeof(T), void> radixsortcopynoalloc(std::size_t count,
	unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, T *const *input, T **output, T **buffer,
	vararguments &&...varparameters) noexcept
{
	static_cast<void>(buffer);  // the single-part version never needs an extra buffer
	radixsortcopynoalloc<direction, mode, indirection2, isindexed2, T>(count, allowedthreads, pfuturesplaceholder,
		pzeroedindices, input, output, varparameters...);
}

// single-part radixsortnoalloc with simple second-level indirection (H:60725-60757)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
inline std::enable_if_t<!std::is_pointer_v<T> && 1u == sizeof(T) && !std::is_same_v<helper::firstarg_t<vararguments...>, bool>,
	void>
radixsortnoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, T **input,
	T **buffer, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::inplace, count,
		pzeroedindices, input, nullptr, buffer, 0, varparameters...));
}

// single-part radixsortnoalloc with an explicit movetobuffer argument (H:60759-60798)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
inline std::enable_if_t<!std::is_pointer_v<T> && 1u == sizeof(T), void> radixsortnoalloc(std::size_t count,
	unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, T **input, T **buffer, bool movetobuffer,
	vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(
		movetobuffer ? helper::sortop::tobuffer : helper::sortop::inplace, count, pzeroedindices, input, nullptr, buffer, 0,
		varparameters...));
}

// radixsortcopy / radixsort with simple second-level indirection (H:60800-60981)
template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_pointer_v<T>, bool> radixsortcopy(std::size_t count, T *const *input, T **output,
	int mmapflags = 0, vararguments &&...varparameters) noexcept
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::copy, count, nullptr,
		input, output, nullptr, 0, varparameters...);
}

template<sortingdirection direction = sortingdirection::ascfwdorder, sortingmode mode = sortingmode::native,
	std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename T, typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_pointer_v<T>, bool> radixsort(std::size_t count, T **input, int mmapflags = 0,
	vararguments &&...varparameters) noexcept
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::inplace, count,
		nullptr, input, nullptr, nullptr, 0, varparameters...);
}

// ---- member object or getter function (H:60983-61357) ----------------------------------------

#define RSBD8_HIP_MEMBERKEY helper::memberkey<indirection1, isindexed2, V, vararguments...>
#define RSBD8_HIP_MEMBERNOEXCEPT noexcept(std::is_member_object_pointer_v<decltype(indirection1)>)

// multi-part radixsortcopynoalloc with indirection (H:60983-61015)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> && 1u < sizeof(RSBD8_HIP_MEMBERKEY), void>
radixsortcopynoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices,
	V *const *input, V **output, V **buffer, vararguments &&...varparameters) RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	static_cast<void>(buffer);
	helper::setlast(helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(helper::sortop::copy,
		count, pzeroedindices, input, output, nullptr, varparameters...));
}

// multi-part radixsortnoalloc with indirection (H:61016-61045)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> && 1u < sizeof(RSBD8_HIP_MEMBERKEY), void>
radixsortnoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V **input,
	V **buffer, bool movetobuffer = false, vararguments &&...varparameters) RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(
		movetobuffer ? helper::sortop::tobuffer : helper::sortop::inplace, count, pzeroedindices, input, nullptr, buffer,
		varparameters...));
}

// single-part radixsortcopynoalloc with indirection (H:61046-61079)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> &&
		!std::is_same_v<helper::firstarg_t<vararguments...>, V **> && 1u == sizeof(RSBD8_HIP_MEMBERKEY), void>
radixsortcopynoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices,
	V *const *input, V **output, vararguments &&...varparameters) RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(helper::sortop::copy,
		count, pzeroedindices, input, output, nullptr, varparameters...));
}

// single-part radixsortcopynoalloc with indirection and a dummy buffer argument (H:61080-61105)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> && 1u == sizeof(RSBD8_HIP_MEMBERKEY), void>
radixsortcopynoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices,
	V *const *input, V **output, V **buffer, vararguments &&...varparameters) RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	static_cast<void>(buffer);  // the single-part version never needs an extra buffer
	helper::setlast(helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(helper::sortop::copy,
		count, pzeroedindices, input, output, nullptr, varparameters...));
}

// single-part radixsortnoalloc with indirection (H:61106-61140)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> &&
		!std::is_same_v<helper::firstarg_t<vararguments...>, bool> && 1u == sizeof(RSBD8_HIP_MEMBERKEY), void>
radixsortnoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V **input,
	V **buffer, vararguments &&...varparameters) RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(helper::sortop::inplace,
		count, pzeroedindices, input, nullptr, buffer, varparameters...));
}

// single-part radixsortnoalloc with indirection and an explicit movetobuffer argument (H:61141-61177)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> && 1u == sizeof(RSBD8_HIP_MEMBERKEY), void>
radixsortnoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V **input,
	V **buffer, bool movetobuffer, vararguments &&...varparameters) RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(
		movetobuffer ? helper::sortop::tobuffer : helper::sortop::inplace, count, pzeroedindices, input, nullptr, buffer,
		varparameters...));
}

// radixsortcopy / radixsort with indirection (H:61179-61357)
template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> && 0u < sizeof(RSBD8_HIP_MEMBERKEY), bool>
radixsortcopy(std::size_t count, V *const *input, V **output, int mmapflags = 0, vararguments &&...varparameters)
	RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(
		helper::sortop::copy, count, nullptr, input, output, nullptr, varparameters...);
}

template<auto indirection1, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<std::is_member_pointer_v<decltype(indirection1)> && 0u < sizeof(RSBD8_HIP_MEMBERKEY), bool>
radixsort(std::size_t count, V **input, int mmapflags = 0, vararguments &&...varparameters) RSBD8_HIP_MEMBERNOEXCEPT
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == helper::sortbymemberpointer<indirection1, direction, mode, indirection2, isindexed2>(
		helper::sortop::inplace, count, nullptr, input, nullptr, nullptr, varparameters...);
}

#undef RSBD8_HIP_MEMBERKEY
#undef RSBD8_HIP_MEMBERNOEXCEPT

// ---- type + byte offset (H:61359-61738): key = the member `object` of ---------------------------
// memberobjectgenerator<T, indirection1>, i.e. *(T *)((std::byte *)input[i] + indirection1); with T a
// pointer type U *, the key is the U that pointer leads to (second level)

// multi-part radixsortcopynoalloc with type and offset indirection (H:61359-61394)
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<!std::is_member_pointer_v<T> && 1u < sizeof(helper::valueorpointee<T>), void> radixsortcopynoalloc(
	std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V *const *input, V **output,
	V **buffer, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	static_cast<void>(buffer);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::copy, count,
		pzeroedindices, input, output, nullptr, indirection1, varparameters...));
}

// multi-part radixsortnoalloc with type and offset indirection (H:61395-61427)
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<!std::is_member_pointer_v<T> && 1u < sizeof(helper::valueorpointee<T>), void> radixsortnoalloc(
	std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V **input, V **buffer,
	bool movetobuffer = false, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(
		movetobuffer ? helper::sortop::tobuffer : helper::sortop::inplace, count, pzeroedindices, input, nullptr, buffer,
		indirection1, varparameters...));
}

// single-part radixsortcopynoalloc with type and offset indirection (H:61428-61464)
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<!std::is_member_pointer_v<T> && 1u == sizeof(helper::valueorpointee<T>) &&
		!std::is_same_v<helper::firstarg_t<vararguments...>, V **>, void>
radixsortcopynoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices,
	V *const *input, V **output, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::copy, count,
		pzeroedindices, input, output, nullptr, indirection1, varparameters...));
}

// single-part radixsortcopynoalloc with type and offset indirection and a dummy buffer (H:61465-61489)
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<!std::is_member_pointer_v<T> && 1u == sizeof(helper::valueorpointee<T>), void> radixsortcopynoalloc(
	std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V *const *input, V **output,
	V **buffer, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	static_cast<void>(buffer);  // the single-part version never needs an extra buffer
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::copy, count,
		pzeroedindices, input, output, nullptr, indirection1, varparameters...));
}

// single-part radixsortnoalloc with type and offset indirection (H:61490-61527)
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<!std::is_member_pointer_v<T> && 1u == sizeof(helper::valueorpointee<T>) &&
		!std::is_same_v<helper::firstarg_t<vararguments...>, bool>, void>
radixsortnoalloc(std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V **input,
	V **buffer, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::inplace, count,
		pzeroedindices, input, nullptr, buffer, indirection1, varparameters...));
}

// single-part radixsortnoalloc with type and offset indirection and an explicit movetobuffer (H:61528-61564)
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
inline std::enable_if_t<!std::is_member_pointer_v<T> && 1u == sizeof(helper::valueorpointee<T>), void> radixsortnoalloc(
	std::size_t count, unsigned allowedthreads, void *pfuturesplaceholder, void *pzeroedindices, V **input, V **buffer,
	bool movetobuffer, vararguments &&...varparameters) noexcept
{
	static_cast<void>(allowedthreads);
	static_cast<void>(pfuturesplaceholder);
	helper::setlast(helper::sortbymember<T, direction, mode, indirection2, isindexed2>(
		movetobuffer ? helper::sortop::tobuffer : helper::sortop::inplace, count, pzeroedindices, input, nullptr, buffer,
		indirection1, varparameters...));
}

// radixsortcopy / radixsort with type and offset indirection (H:61566-61738)
template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_member_pointer_v<T>, bool> radixsortcopy(std::size_t count, V *const *input,
	V **output, int mmapflags = 0, vararguments &&...varparameters) noexcept
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::copy, count, nullptr,
		input, output, nullptr, indirection1, varparameters...);
}

template<typename T, std::ptrdiff_t indirection1 = 0, sortingdirection direction = sortingdirection::ascfwdorder,
	sortingmode mode = sortingmode::native, std::ptrdiff_t indirection2 = 0, bool isindexed2 = false, typename V,
	typename... vararguments>
[[nodiscard]] inline std::enable_if_t<!std::is_member_pointer_v<T>, bool> radixsort(std::size_t count, V **input,
	int mmapflags = 0, vararguments &&...varparameters) noexcept
{
	static_cast<void>(mmapflags);
	return RSBD8_HIP_OK == helper::sortbymember<T, direction, mode, indirection2, isindexed2>(helper::sortop::inplace, count,
		nullptr, input, nullptr, nullptr, indirection1, varparameters...);
}

} // namespace rsbd8
