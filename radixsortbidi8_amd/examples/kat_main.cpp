// kat_main.cpp -- the reference's portable known-answer tests (multiplatform/main.cpp:10-304) and
// the documented direction example (Radixsortbidi8.hpp:265-274), written against the drop-in
// header include/rsbd8_hip.hpp, plus:
//   - the compile-time offsets of main.cpp:235-238 (static_assert on getoffsetof);
//   - every noalloc form, with and without indirection, on device arrays and allocatearray
//     workspaces (T**, member object, getter, type+offset; copy, in place, to buffer);
//   - indexed first-level and second-level indirection (one and two index arguments), getters
//     returning a pointer or a reference, with and without isindexed2, in all four directions;
//   - the wide keys: main.cpp:21-178 for the three 80-bit layouts (directly, through pointers,
//     and through the noalloc forms on device arrays), x86 long double and the 128-bit integers.
// Expected orders come from std::stable_sort over the keys read on the host (ties in source order,
// or reversed for the rev directions). Exits non-zero on the first mismatch and prints which check
// failed.
#include <algorithm>
#include <cmath>
#include <limits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/rsbd8_hip.hpp"

#pragma pack(push, 1)
class Testmeclass {
	std::uint64_t wasted{};
	char misalignoffset{};

public:
	std::uint64_t co;
	std::int64_t sco;
	constexpr std::uint64_t get() const noexcept { return co; }
	constexpr std::uint64_t getwparam(int testfor8in) const noexcept { return testfor8in == 8 ? co : 0; }
	constexpr std::int64_t sget() const noexcept { return sco; }
	constexpr std::uint64_t bget() noexcept { return co; }
	constexpr Testmeclass(std::uint64_t input) noexcept : co{input}, sco{static_cast<std::int64_t>(input) - 1} {}
};
#pragma pack(pop)

struct myclass {
	int keyorder;
	char const *name;
};

// a record for the indexed forms: four 32-bit keys (first-level index), a pointer to 16-bit values
// (second-level index), and getters onto both
struct Rec {
	std::uint32_t k[4];
	std::uint16_t const *vals;
	std::uint16_t const *second() const noexcept { return vals; }
	std::uint32_t const &third() const noexcept { return k[3]; }
	std::uint16_t const *offsetvals(int add) const noexcept { return vals + add; }
};

static int failures;
#define CHECK(cond)                                                                    \
	do {                                                                               \
		if (!(cond)) {                                                                 \
			std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);               \
			++failures;                                                                \
		}                                                                              \
	} while (0)

// the compile-time offsets of main.cpp:233-238
static std::size_t constexpr sizecontainer{sizeof(Testmeclass)};
static_assert(25 == sizecontainer);
static std::size_t constexpr offsetco{rsbd8::getoffsetof<&Testmeclass::co>};
static_assert(9 == offsetco);
static std::size_t constexpr offsetsco{rsbd8::getoffsetof<&Testmeclass::sco>};
static_assert(17 == offsetsco);
static_assert(0 == rsbd8::getoffsetof<&myclass::keyorder> && 8 == rsbd8::getoffsetof<&myclass::name>);
static_assert(16 == rsbd8::getoffsetof<&Rec::vals>);
static_assert(5 == rsbd8::getoffsetof<&rsbd8::memberobjectgenerator<std::uint32_t, 5>::object>);
static_assert(0 == rsbd8::getoffsetof<&rsbd8::memberobjectgenerator<double, 0>::object>);

// the stable order of keys (ascending, or descending), ties in source order or reversed
template<typename K>
std::vector<std::size_t> stableorder(std::vector<K> const &keys, bool desc, bool rev)
{
	std::vector<std::size_t> idx(keys.size());
	std::iota(idx.begin(), idx.end(), std::size_t{});
	if (rev) std::reverse(idx.begin(), idx.end());
	std::stable_sort(idx.begin(), idx.end(), [&](std::size_t a, std::size_t b) { return desc ? keys[b] < keys[a] : keys[a] < keys[b]; });
	return idx;
}

template<typename T>
T *todevice(std::vector<T> const &h)
{
	T *d{};
	if (hipMalloc(reinterpret_cast<void **>(&d), h.size() * sizeof(T) + 8) != hipSuccess) return nullptr;
	if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
	return d;
}

template<typename T>
std::vector<T> tohost(T const *d, std::size_t n)
{
	std::vector<T> h(n);
	if (hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) h.clear();
	return h;
}

static std::uint64_t rng_state{0x9E3779B97F4A7C15u};
static std::uint64_t next()
{
	rng_state ^= rng_state << 13;
	rng_state ^= rng_state >> 7;
	rng_state ^= rng_state << 17;
	return rng_state;
}

// pointer results as indices into a record array
template<typename V>
std::vector<std::size_t> indices(std::vector<V *> const &ptrs, V *base)
{
	std::vector<std::size_t> r(ptrs.size());
	for (std::size_t i{}; i != ptrs.size(); ++i) r[i] = static_cast<std::size_t>(ptrs[i] - base);
	return r;
}

// multiplatform/main.cpp:21-178 for one 80-bit layout R: -inf, +inf, the machine QNaN, max normal,
// min normal, max and min subnormal (arbitrary padding above the 16-bit sign/exponent), sorted
// ascending and dscrevorder, directly and through pointers; then the same through the noalloc
// forms on device arrays
template<typename R, typename SE>
void widekat(SE const (&se)[7])
{
	std::uint64_t const mant[7]{0, 0, 0x8000000000000000u, 0xFFFFFFFFFFFFFFFFu, 0, 0xFFFFFFFFFFFFFFFFu, 1u};
	R in[7], out[7];
	for (int i = 0; i < 7; ++i) {
		in[i].mantissa = mant[i];
		in[i].signexponent = se[i];
	}
	int const asc[7]{2, 0, 6, 5, 4, 3, 1}, dsc[7]{1, 3, 4, 5, 6, 0, 2};
	auto same = [&](R const *o, int const *w) {
		for (int i = 0; i < 7; ++i)
			if (std::memcmp(&o[i], &in[w[i]], sizeof(R))) return false;
		return true;
	};
	CHECK(rsbd8::radixsortcopy(7, in, out) && same(out, asc));
	CHECK(rsbd8::radixsortcopy<rsbd8::sortingdirection::dscrevorder>(7, in, out) && same(out, dsc));
	R *pin[7], *pout[7];
	for (int i = 0; i < 7; ++i) pin[i] = in + i;
	auto psame = [&](int const *w) {
		for (int i = 0; i < 7; ++i)
			if (pout[i] != in + w[i]) return false;
		return true;
	};
	CHECK(rsbd8::radixsortcopy(7, pin, pout) && psame(asc));
	CHECK(rsbd8::radixsortcopy<rsbd8::sortingdirection::dscrevorder>(7, pin, pout) && psame(dsc));
	auto [alloc, size] = rsbd8::allocatearray<false, true, true, R, false>(7);
	CHECK(alloc.pbuffer != nullptr);
	rsbd8::buffermemorywrapper guard{alloc.pbuffer, size};
	R *din{todevice(std::vector<R>(in, in + 7))}, *dout{};
	CHECK(hipMalloc(reinterpret_cast<void **>(&dout), sizeof(in)) == hipSuccess);
	rsbd8::radixsortcopynoalloc(7, 0u, nullptr, alloc.pzeroedindices, din, dout, alloc.pbuffer);
	CHECK(rsbd8::lasterror() == 0 && same(tohost(dout, 7).data(), asc));
	rsbd8::radixsortnoalloc<rsbd8::sortingdirection::dscrevorder>(7, 0u, nullptr, alloc.pzeroedindices, din, alloc.pbuffer, true);
	CHECK(rsbd8::lasterror() == 0 && same(tohost(alloc.pbuffer, 7).data(), dsc));
	static_cast<void>(hipFree(din));
	static_cast<void>(hipFree(dout));
}

// every indirection form of one direction over n records, host and device arrays
template<rsbd8::sortingdirection dir>
void indirection_forms(std::size_t n)
{
	constexpr bool desc{rsbd8::helper::isdescsort<dir>}, rev{rsbd8::helper::isrevorderwithindirection<dir>};
	// records on the host, and the same records (with device value arrays) on the device
	std::vector<std::uint16_t> vals(n * 3);
	for (auto &v : vals) v = static_cast<std::uint16_t>(next() % 97);  // many ties
	std::uint16_t *dvals{todevice(vals)};
	std::vector<Rec> hrec(n), drec_h(n);
	for (std::size_t i{}; i != n; ++i) {
		for (auto &k : hrec[i].k) k = static_cast<std::uint32_t>(next() % 1000u);
		hrec[i].vals = vals.data() + 3 * i;
		drec_h[i] = hrec[i];
		drec_h[i].vals = dvals + 3 * i;
	}
	Rec *drec{todevice(drec_h)};
	std::vector<Rec *> hin(n), din_h(n);
	for (std::size_t i{}; i != n; ++i) {
		hin[i] = hrec.data() + (i * 7919u) % n;
		din_h[i] = drec + (i * 7919u) % n;
	}
	Rec **din{todevice(din_h)};
	Rec **dout{}, **dbuf{};
	CHECK(hipMalloc(reinterpret_cast<void **>(&dout), n * sizeof(Rec *)) == hipSuccess);
	CHECK(hipMalloc(reinterpret_cast<void **>(&dbuf), n * sizeof(Rec *)) == hipSuccess);
	auto [alloc, size] = rsbd8::allocatearray<false, false, false, std::uint32_t, true>(n);
	CHECK(alloc.pbuffer != nullptr);
	rsbd8::buffermemorywrapper guard{alloc.pbuffer, size};
	void *z{alloc.pzeroedindices};
	// the source order of the pointer arrays, as record indices
	std::vector<std::size_t> src(n);
	for (std::size_t i{}; i != n; ++i) src[i] = (i * 7919u) % n;
	auto expect{[&](auto keyof) {
		using K = decltype(keyof(std::size_t{}));
		std::vector<K> keys(n);
		for (std::size_t i{}; i != n; ++i) keys[i] = keyof(src[i]);
		std::vector<std::size_t> order{stableorder(keys, desc, rev)};
		for (auto &o : order) o = src[o];
		return order;
	}};
	auto dres{[&](Rec **d) { return indices(tohost(d, n), drec); }};
	std::vector<Rec *> hout(n), hbuf(n);

	// first level, indexed: key = k[2] through the member k[0] displaced by 2 elements (H:2755-2758)
	{
		auto const want{expect([&](std::size_t r) { return hrec[r].k[2]; })};
		CHECK((rsbd8::radixsortcopy<std::uint32_t, 0, dir>(n, hin.data(), hout.data(), 0, 2)) && indices(hout, hrec.data()) == want);
		rsbd8::radixsortcopynoalloc<std::uint32_t, 0, dir>(n, 0u, nullptr, z, din, dout, dbuf, 2);
		CHECK(rsbd8::lasterror() == 0 && dres(dout) == want);
		CHECK(hipMemcpy(dbuf, din, n * sizeof(Rec *), hipMemcpyDeviceToDevice) == hipSuccess);
		rsbd8::radixsortnoalloc<std::uint32_t, 0, dir>(n, 0u, nullptr, z, dbuf, dout, false, 2);
		CHECK(rsbd8::lasterror() == 0 && dres(dbuf) == want);
		rsbd8::radixsortnoalloc<std::uint32_t, 0, dir>(n, 0u, nullptr, z, din, dout, true, 2);
		CHECK(rsbd8::lasterror() == 0 && dres(dout) == want);
	}
	// type+offset with a byte offset: k[1] at +4
	{
		auto const want{expect([&](std::size_t r) { return hrec[r].k[1]; })};
		CHECK((rsbd8::radixsortcopy<std::uint32_t, 4, dir>(n, hin.data(), hout.data())) && indices(hout, hrec.data()) == want);
		rsbd8::radixsortcopynoalloc<std::uint32_t, 4, dir>(n, 0u, nullptr, z, din, dout, dbuf);
		CHECK(rsbd8::lasterror() == 0 && dres(dout) == want);
		hbuf = hin;
		CHECK((rsbd8::radixsort<std::uint32_t, 4, dir>(n, hbuf.data())) && indices(hbuf, hrec.data()) == want);
	}
	// second level through the pointer member, no index: vals[0]; indexed: vals[2]; and both
	// indices: the pointer member of the record displaced by one pointer... (records are not
	// contiguous pointers, so the two-index form is taken through a type+offset at the pointer)
	{
		auto const want0{expect([&](std::size_t r) { return vals[3 * r]; })};
		CHECK((rsbd8::radixsortcopy<&Rec::vals, dir>(n, hin.data(), hout.data())) && indices(hout, hrec.data()) == want0);
		rsbd8::radixsortcopynoalloc<&Rec::vals, dir>(n, 0u, nullptr, z, din, dout, dbuf);
		CHECK(rsbd8::lasterror() == 0 && dres(dout) == want0);
		auto const want2{expect([&](std::size_t r) { return vals[3 * r + 2]; })};
		CHECK((rsbd8::radixsortcopy<&Rec::vals, dir, rsbd8::sortingmode::native, 0, true>(n, hin.data(), hout.data(), 0, 2)) &&
			indices(hout, hrec.data()) == want2);
		rsbd8::radixsortcopynoalloc<&Rec::vals, dir, rsbd8::sortingmode::native, 0, true>(n, 0u, nullptr, z, din, dout, dbuf, 2);
		CHECK(rsbd8::lasterror() == 0 && dres(dout) == want2);
		// indirection2 in bytes: vals + 2 bytes, index 1 -> vals[2]
		rsbd8::radixsortcopynoalloc<&Rec::vals, dir, rsbd8::sortingmode::native, 2, true>(n, 0u, nullptr, z, din, dout, dbuf, 1);
		CHECK(rsbd8::lasterror() == 0 && dres(dout) == want2);
		CHECK(hipMemcpy(dbuf, din, n * sizeof(Rec *), hipMemcpyDeviceToDevice) == hipSuccess);
		rsbd8::radixsortnoalloc<&Rec::vals, dir, rsbd8::sortingmode::native, 0, true>(n, 0u, nullptr, z, dbuf, dout, false, 2);
		CHECK(rsbd8::lasterror() == 0 && dres(dbuf) == want2);
		// type+offset of a pointer type with two indices: the first displaces the record by whole
		// pointers (16 + 0 * 8: the vals member), the second indexes the values
		CHECK((rsbd8::radixsortcopy<std::uint16_t *, 16, dir, rsbd8::sortingmode::native, 0, true>(n, hin.data(), hout.data(), 0, 0, 1)) &&
			indices(hout, hrec.data()) == expect([&](std::size_t r) { return vals[3 * r + 1]; }));
		rsbd8::radixsortcopynoalloc<std::uint16_t *, 8, dir, rsbd8::sortingmode::native, 0, true>(n, 0u, nullptr, z, din, dout, dbuf, 1, 1);
		CHECK(rsbd8::lasterror() == 0 && dres(dout) == expect([&](std::size_t r) { return vals[3 * r + 1]; }));
	}
	// getters (host arrays only): a pointer with isindexed2 (the first argument is the index), a
	// pointer with an argument for the getter, and a reference (read as a pointer to its object)
	{
		CHECK((rsbd8::radixsortcopy<&Rec::second, dir, rsbd8::sortingmode::native, 0, true>(n, hin.data(), hout.data(), 0, 1)) &&
			indices(hout, hrec.data()) == expect([&](std::size_t r) { return vals[3 * r + 1]; }));
		CHECK((rsbd8::radixsortcopy<&Rec::offsetvals, dir, rsbd8::sortingmode::native, 0, true>(n, hin.data(), hout.data(), 0, 1, 1)) &&
			indices(hout, hrec.data()) == expect([&](std::size_t r) { return vals[3 * r + 2]; }));
		CHECK((rsbd8::radixsortcopy<&Rec::third, dir>(n, hin.data(), hout.data())) &&
			indices(hout, hrec.data()) == expect([&](std::size_t r) { return hrec[r].k[3]; }));
		rsbd8::radixsortcopynoalloc<&Rec::third, dir, rsbd8::sortingmode::native, -4>(n, 0u, nullptr, z, hin.data(), hout.data(), hbuf.data());
		CHECK(rsbd8::lasterror() == 0 && indices(hout, hrec.data()) == expect([&](std::size_t r) { return hrec[r].k[2]; }));
		hbuf = hin;
		rsbd8::radixsortnoalloc<&Rec::second, dir, rsbd8::sortingmode::native, 0, true>(n, 0u, nullptr, z, hbuf.data(), hout.data(), false, 2);
		CHECK(rsbd8::lasterror() == 0 && indices(hbuf, hrec.data()) == expect([&](std::size_t r) { return vals[3 * r + 2]; }));
		// a getter cannot run on device memory
		rsbd8::radixsortcopynoalloc<&Rec::second, dir, rsbd8::sortingmode::native, 0, true>(n, 0u, nullptr, z, din, dout, dbuf, 1);
		CHECK(rsbd8::lasterror() == RSBD8_HIP_EINVAL);
	}
	// T**: key = p[index], with the pointers into the device value array
	{
		std::vector<std::uint16_t const *> hp(n);
		std::vector<std::uint16_t *> dp_h(n);
		for (std::size_t i{}; i != n; ++i) {
			hp[i] = vals.data() + 3 * src[i];
			dp_h[i] = dvals + 3 * src[i];
		}
		std::vector<std::uint16_t const *> hpo(n);
		auto const want1{expect([&](std::size_t r) { return vals[3 * r + 1]; })};
		CHECK(rsbd8::radixsortcopy<dir>(n, hp.data(), hpo.data(), 0, 1));
		std::vector<std::size_t> got(n);
		for (std::size_t i{}; i != n; ++i) got[i] = static_cast<std::size_t>(hpo[i] - vals.data()) / 3;
		CHECK(got == want1);
		std::uint16_t **dp{todevice(dp_h)};
		std::uint16_t **dpo{}, **dpb{};
		CHECK(hipMalloc(reinterpret_cast<void **>(&dpo), n * sizeof(void *)) == hipSuccess);
		CHECK(hipMalloc(reinterpret_cast<void **>(&dpb), n * sizeof(void *)) == hipSuccess);
		rsbd8::radixsortcopynoalloc<dir>(n, 0u, nullptr, z, dp, dpo, dpb, 1);
		CHECK(rsbd8::lasterror() == 0);
		auto const res{tohost(dpo, n)};
		for (std::size_t i{}; i != n; ++i) got[i] = static_cast<std::size_t>(res[i] - dvals) / 3;
		CHECK(got == want1);
		rsbd8::radixsortnoalloc<dir>(n, 0u, nullptr, z, dp, dpb, true);
		CHECK(rsbd8::lasterror() == 0);
		auto const res0{tohost(dpb, n)};
		for (std::size_t i{}; i != n; ++i) got[i] = static_cast<std::size_t>(res0[i] - dvals) / 3;
		CHECK(got == expect([&](std::size_t r) { return vals[3 * r]; }));
		static_cast<void>(hipFree(dp));
		static_cast<void>(hipFree(dpo));
		static_cast<void>(hipFree(dpb));
	}
	static_cast<void>(hipFree(dvals));
	static_cast<void>(hipFree(drec));
	static_cast<void>(hipFree(din));
	static_cast<void>(hipFree(dout));
	static_cast<void>(hipFree(dbuf));
}

int main()
{
	{ // main.cpp:10-19, in-place plain sort
		std::uint32_t simple[]{18, 33, 21, 0, 5, 1, 13};
		CHECK(rsbd8::radixsort(7, simple));
		std::uint32_t const want[]{0, 1, 5, 13, 18, 21, 33};
		CHECK(0 == std::memcmp(simple, want, sizeof(want)));
	}
	{ // main.cpp:182-198, single-byte enum, descending with reversed ties, then ascending
		enum code64 : std::uint8_t { c0 = 0, ci = 18, cq = 26, cB = 37, cE = 40, cU = 56, cunder = 63 };
		code64 const in[7]{c0, cB, cunder, ci, cq, cU, cE};
		code64 out[7];
		CHECK(rsbd8::radixsortcopy<rsbd8::sortingdirection::dscrevorder>(7, in, out));
		code64 const dsc[7]{cunder, cU, cE, cB, cq, ci, c0};
		CHECK(0 == std::memcmp(out, dsc, 7));
		CHECK(rsbd8::radixsortcopy<rsbd8::sortingdirection::ascfwdorder>(7, in, out));
		code64 const asc[7]{c0, ci, cq, cB, cE, cU, cunder};
		CHECK(0 == std::memcmp(out, asc, 7));
	}
	std::uint32_t outm[7];
	{ // main.cpp:200-205, float bit patterns with negative subnormals and ties
		std::uint32_t inm[7]{8, 0, 3, 1u << 31 | 2, 3, 1u << 31 | 18, 1u << 31 | 2};
		CHECK(rsbd8::radixsortcopy(7, reinterpret_cast<float *>(inm), reinterpret_cast<float *>(outm)));
		std::uint32_t const want[7]{1u << 31 | 18, 1u << 31 | 2, 1u << 31 | 2, 0, 3, 3, 8};
		CHECK(0 == std::memcmp(outm, want, sizeof(want)));
	}
	{ // main.cpp:207-215, the same through pointers: dscrevorder pins the reversed tie order
		std::uint32_t const *inim[7]{outm + 6, outm + 3, outm + 4, outm + 1, outm + 5, outm, outm + 2};
		std::uint32_t const *outim[7];
		CHECK(rsbd8::radixsortcopy<rsbd8::sortingdirection::dscrevorder>(7, reinterpret_cast<float const *const *>(inim),
			reinterpret_cast<float const **>(outim)));
		CHECK(outim[0] == inim[0] && outim[1] == inim[4] && outim[2] == inim[2] && outim[3] == inim[1] &&
			outim[4] == inim[6] && outim[5] == inim[3] && outim[6] == inim[5]);
		CHECK(rsbd8::radixsortcopy<rsbd8::sortingdirection::ascfwdorder>(7, reinterpret_cast<float const *const *>(inim),
			reinterpret_cast<float const **>(outim)));
		CHECK(*outim[0] == (1u << 31 | 18) && *outim[1] == (1u << 31 | 2) && *outim[2] == (1u << 31 | 2) && *outim[3] == 0 &&
			*outim[4] == 3 && *outim[5] == 3 && *outim[6] == 8);
	}
	Testmeclass cin[]{8, 0, 6, 4, 0, 2, 6};
	// ascending by co with ties in source order: 0(1) 0(4) 2(5) 4(3) 6(2) 6(6) 8(0)
	int const want[7]{1, 4, 5, 3, 2, 6, 0};
	{ // main.cpp:217-304, getter, member and type+offset indirection on a packed 25-byte class
		Testmeclass const *fin[7]{cin, cin + 1, cin + 2, cin + 3, cin + 4, cin + 5, cin + 6};
		Testmeclass const *fout[7];
		auto same = [&] {
			for (int i = 0; i < 7; ++i)
				if (fout[i] != fin[want[i]]) return false;
			return true;
		};
		CHECK(rsbd8::radixsortcopy<&Testmeclass::get>(7, fin, fout) && same());
		CHECK(rsbd8::radixsortcopy<&Testmeclass::getwparam>(7, fin, fout, 0, 8) && same());
		CHECK(rsbd8::radixsortcopy<&Testmeclass::co>(7, fin, fout) && same());
		CHECK((rsbd8::radixsortcopy<std::uint64_t, 9>(7, fin, fout)) && same());
		CHECK(rsbd8::radixsortcopy<&Testmeclass::sget>(7, fin, fout) && same());
		CHECK(rsbd8::radixsortcopy<&Testmeclass::sco>(7, fin, fout) && same());
		CHECK((rsbd8::radixsortcopy<std::int64_t, 17>(7, fin, fout)) && same());
		Testmeclass *yin[7]{cin, cin + 1, cin + 2, cin + 3, cin + 4, cin + 5, cin + 6};
		Testmeclass *yout[7];
		auto ysame = [&] {
			for (int i = 0; i < 7; ++i)
				if (yout[i] != yin[want[i]]) return false;
			return true;
		};
		// main.cpp:286-304: non-const getters work on non-const records
		CHECK(rsbd8::radixsortcopy<&Testmeclass::bget>(7, yin, yout) && ysame());
		CHECK(rsbd8::radixsortcopy<&Testmeclass::getwparam>(7, yin, yout, 0, 8) && ysame());
	}
	{ // Radixsortbidi8.hpp:265-274: dscfwdorder keeps tie order, dscrevorder reverses it
		myclass collA[]{{1, "first"}, {1, "second"}, {-5, "third"}, {2, "fourth"}};
		myclass *pcollA[]{collA, collA + 1, collA + 2, collA + 3};
		CHECK((rsbd8::radixsort<&myclass::keyorder, rsbd8::sortingdirection::dscfwdorder>(4, pcollA)));
		CHECK(pcollA[0] == collA + 3 && pcollA[1] == collA && pcollA[2] == collA + 1 && pcollA[3] == collA + 2);
		myclass *pcollB[]{collA, collA + 1, collA + 2, collA + 3};
		CHECK((rsbd8::radixsort<&myclass::keyorder, rsbd8::sortingdirection::dscrevorder>(4, pcollB)));
		CHECK(pcollB[0] == collA + 3 && pcollB[1] == collA + 1 && pcollB[2] == collA && pcollB[3] == collA + 2);
	}
	{ // device arrays through the noalloc forms with an allocatearray workspace
		std::size_t const n{1u << 20};
		std::vector<std::uint64_t> h(n);
		for (auto &v : h) v = next();
		auto [alloc, size] = rsbd8::allocatearray<false, false, false, std::int64_t, false>(n);
		CHECK(alloc.pbuffer != nullptr);
		rsbd8::buffermemorywrapper guard{alloc.pbuffer, size};
		std::int64_t *din{}, *dout{};
		CHECK(hipMalloc(reinterpret_cast<void **>(&din), n * 8) == hipSuccess);
		CHECK(hipMalloc(reinterpret_cast<void **>(&dout), n * 8) == hipSuccess);
		CHECK(hipMemcpy(din, h.data(), n * 8, hipMemcpyHostToDevice) == hipSuccess);
		rsbd8::radixsortcopynoalloc<rsbd8::sortingdirection::dscfwdorder>(n, 0u, nullptr, alloc.pzeroedindices, din, dout, alloc.pbuffer);
		CHECK(rsbd8::lasterror() == 0);
		std::vector<std::int64_t> got(n);
		CHECK(hipMemcpy(got.data(), dout, n * 8, hipMemcpyDeviceToHost) == hipSuccess);
		bool ok{true};
		for (std::size_t i = 1; i < n; ++i) ok &= got[i - 1] >= got[i];  // signed descending
		CHECK(ok);
		rsbd8::radixsortnoalloc(n, 0u, nullptr, alloc.pzeroedindices, din, alloc.pbuffer, false);
		CHECK(rsbd8::lasterror() == 0);
		CHECK(hipMemcpy(got.data(), din, n * 8, hipMemcpyDeviceToHost) == hipSuccess);
		ok = true;
		for (std::size_t i = 1; i < n; ++i) ok &= got[i - 1] <= got[i];
		CHECK(ok);
		// a host array handed to a noalloc form is rejected before any launch
		std::vector<std::int64_t> hostarr(n);
		rsbd8::radixsortcopynoalloc(n, 0u, nullptr, alloc.pzeroedindices, reinterpret_cast<std::int64_t const *>(h.data()),
			hostarr.data(), alloc.pbuffer);
		CHECK(rsbd8::lasterror() == RSBD8_HIP_EINVAL);
		static_cast<void>(hipFree(din));
		static_cast<void>(hipFree(dout));
	}
	{ // the packed class on the device: member, type+offset, in place and to buffer, single pointer
		std::vector<Testmeclass> hc(cin, cin + 7);
		Testmeclass *dc{todevice(hc)};
		std::vector<Testmeclass *> dptr(7);
		for (int i = 0; i < 7; ++i) dptr[i] = dc + i;
		Testmeclass **din{todevice(dptr)}, **dout{}, **dbuf{};
		CHECK(hipMalloc(reinterpret_cast<void **>(&dout), 7 * sizeof(void *)) == hipSuccess);
		CHECK(hipMalloc(reinterpret_cast<void **>(&dbuf), 7 * sizeof(void *)) == hipSuccess);
		auto [alloc, size] = rsbd8::allocatearray<false, false, false, std::uint64_t, true>(7);
		CHECK(alloc.pbuffer != nullptr);
		rsbd8::buffermemorywrapper guard{alloc.pbuffer, size};
		auto dsame = [&](Testmeclass **d) {
			auto const r{tohost(d, 7)};
			for (int i = 0; i < 7; ++i)
				if (r[i] != dc + want[i]) return false;
			return true;
		};
		rsbd8::radixsortcopynoalloc<&Testmeclass::co>(7, 0u, nullptr, alloc.pzeroedindices, din, dout, dbuf);
		CHECK(rsbd8::lasterror() == 0 && dsame(dout));
		rsbd8::radixsortcopynoalloc<std::int64_t, 17>(7, 0u, nullptr, alloc.pzeroedindices, din, dout, dbuf);
		CHECK(rsbd8::lasterror() == 0 && dsame(dout));
		rsbd8::radixsortnoalloc<&Testmeclass::sco>(7, 0u, nullptr, alloc.pzeroedindices, din, dbuf, true);
		CHECK(rsbd8::lasterror() == 0 && dsame(dbuf));
		rsbd8::radixsortnoalloc<std::uint64_t, 9>(7, 0u, nullptr, alloc.pzeroedindices, din, dbuf);
		CHECK(rsbd8::lasterror() == 0 && dsame(din));
		CHECK(rsbd8::radixsortcopy<&Testmeclass::co>(7, din, dout) && dsame(dout));
		CHECK((rsbd8::radixsort<std::uint64_t, 9>(7, din)) && dsame(din));
		// count 1 with device pointer arrays: the element is copied on the device
		rsbd8::radixsortcopynoalloc<&Testmeclass::co>(1, 0u, nullptr, alloc.pzeroedindices, din + 3, dout, dbuf);
		CHECK(rsbd8::lasterror() == 0 && tohost(dout, 1)[0] == dc + want[3]);
		static_cast<void>(hipFree(dc));
		static_cast<void>(hipFree(din));
		static_cast<void>(hipFree(dout));
		static_cast<void>(hipFree(dbuf));
	}
	{ // keys wider than 64 bits: the reference's 80-bit known answers, x86 long double, 128-bit integers
		std::uint16_t const se80[7]{0xFFFFu, 0x7FFFu, 0xFFFFu, 0x7FFEu, 1u, 0, 0};
		std::uint32_t const se96[7]{0xABABFFFFu, 0xD0017FFFu, 0xEEEEFFFFu, 0x01017FFEu, 0xFFF80001u, 0xC7C80000u, 0xB3710000u};
		std::uint64_t const se128[7]{0xBBBBAAAAABABFFFFu, 0x22221111D0017FFFu, 0x44443333EEEEFFFFu, 0x6666555501017FFEu,
			0x88887777FFF80001u, 0xCCCC9999C7C80000u, 0xFFFFDDDDB3710000u};
		widekat<rsbd8::helper::longdoubletest80<false, true, true>>(se80);
		widekat<rsbd8::helper::longdoubletest96<false, true, true>>(se96);
		widekat<rsbd8::helper::longdoubletest128<false, true, true>>(se128);
		static_assert(10 == sizeof(rsbd8::helper::longdoubletest80<false, true, true>) &&
			12 == sizeof(rsbd8::helper::longdoubletest96<false, true, true>));
		if constexpr (sizeof(long double) == 16 && std::numeric_limits<long double>::digits == 64) {
			long double ld[6]{-1.5L, 0.0L, -0.0L, 2.0L, 1e300L, -1e-4000L};
			long double lo[6];
			CHECK(rsbd8::radixsortcopy(6, ld, lo));
			// -1.5 < -1e-4000 < 0.0 == -0.0 (ties in source order) < 2 < 1e300
			CHECK(lo[0] == -1.5L && lo[1] == -1e-4000L && !std::signbit(lo[2]) && std::signbit(lo[3]) && lo[3] == 0.0L &&
				lo[4] == 2.0L && lo[5] == 1e300L);
			CHECK(rsbd8::radixsortcopy<rsbd8::sortingdirection::dscrevorder>(6, ld, lo));
			CHECK(lo[0] == 1e300L && lo[1] == 2.0L && std::signbit(lo[2]) && !std::signbit(lo[3]) && lo[4] == -1e-4000L &&
				lo[5] == -1.5L);
			CHECK((rsbd8::radixsort<rsbd8::sortingdirection::ascfwdorder, rsbd8::sortingmode::nativeabs>(6, ld)));
			CHECK(ld[0] == 0.0L && ld[1] == 0.0L && ld[2] == -1e-4000L && ld[3] == -1.5L && ld[4] == 2.0L && ld[5] == 1e300L);
		}
#ifdef __SIZEOF_INT128__
		__int128 const big{static_cast<__int128>(1) << 100};
		__int128 iv[6]{-3, 5, 0, -big, big * 1024, -1};
		__int128 io[6];
		CHECK(rsbd8::radixsortcopy(6, iv, io));
		CHECK(io[0] == -big && io[1] == -3 && io[2] == -1 && io[3] == 0 && io[4] == 5 && io[5] == big * 1024);
		unsigned __int128 uv[4]{static_cast<unsigned __int128>(-1), 7, static_cast<unsigned __int128>(big), 0};
		CHECK((rsbd8::radixsort<rsbd8::sortingdirection::dscfwdorder>(4, uv)));
		CHECK(uv[0] == static_cast<unsigned __int128>(-1) && uv[1] == static_cast<unsigned __int128>(big) && uv[2] == 7 && uv[3] == 0);
#endif
	}
	for (std::size_t n : {std::size_t{1000}, std::size_t{70001}}) {
		indirection_forms<rsbd8::sortingdirection::ascfwdorder>(n);
		indirection_forms<rsbd8::sortingdirection::dscfwdorder>(n);
		indirection_forms<rsbd8::sortingdirection::ascrevorder>(n);
		indirection_forms<rsbd8::sortingdirection::dscrevorder>(n);
	}
	std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
	return failures ? 1 : 0;
}
