// kat_main.cpp -- the reference's portable known-answer tests (multiplatform/main.cpp:10-304) and
// the documented direction example (Radixsortbidi8.hpp:265-274), written against the drop-in
// header include/rsbd8_hip.hpp, plus device-array runs of the noalloc forms. Exits non-zero on the
// first mismatch and prints which check failed.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rsbd8_hip.hpp"

#pragma pack(push, 1)
class Testmeclass {
	std::uint64_t wasted{};
	char misalignoffset{};

public:
	std::uint64_t co;
	std::int64_t sco;
	std::uint64_t get() const noexcept { return co; }
	std::uint64_t getwparam(int testfor8in) const noexcept { return testfor8in == 8 ? co : 0; }
	std::int64_t sget() const noexcept { return sco; }
	Testmeclass(std::uint64_t input) noexcept : co{input}, sco{static_cast<std::int64_t>(input) - 1} {}
};
#pragma pack(pop)

struct myclass {
	int keyorder;
	char const *name;
};

static int failures;
#define CHECK(cond)                                                                    \
	do {                                                                               \
		if (!(cond)) {                                                                 \
			std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);               \
			++failures;                                                                \
		}                                                                              \
	} while (0)

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
	{ // main.cpp:217-304, getter, member and type+offset indirection on a packed 25-byte class
		static_assert(25 == sizeof(Testmeclass));
		CHECK(9 == rsbd8::getoffsetof<&Testmeclass::co>);
		CHECK(17 == rsbd8::getoffsetof<&Testmeclass::sco>);
		Testmeclass cin[]{8, 0, 6, 4, 0, 2, 6};
		Testmeclass const *fin[7]{cin, cin + 1, cin + 2, cin + 3, cin + 4, cin + 5, cin + 6};
		Testmeclass const *fout[7];
		// ascending by co with ties in source order: 0(1) 0(4) 2(5) 4(3) 6(2) 6(6) 8(0)
		int const want[7]{1, 4, 5, 3, 2, 6, 0};
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
		std::uint64_t x{0x9E3779B97F4A7C15u};
		for (auto &v : h) {
			x ^= x << 13; x ^= x >> 7; x ^= x << 17;
			v = x;
		}
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
		static_cast<void>(hipFree(din));
		static_cast<void>(hipFree(dout));
	}
	std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
	return failures ? 1 : 0;
}
