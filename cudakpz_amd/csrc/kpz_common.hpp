// Shared host/device definitions of the KPZ octahedron-model engine.
//
// Lattice encodings are those of kpz/systemSize.h:94-165 (bit-identical):
//  * ENC_LOCAL: 4x4 sites per u32, word ((y>>2)<<(lx-2)) + (x>>2), bit
//    ((x&3)<<1)|((y&3)<<3); bit+0 = x slope, bit+1 = y slope.
//  * ENC_CBBIT: four quadrants of sizeW/4 u32 (sub0 X, sub0 Y, sub1 X,
//    sub1 Y), sub = (x^y)&1, word (y<<(lx-6)) + (x>>6), bit (x>>1)&31.
// A slope bit of 1 means +1 (height rises towards +x / +y).
#pragma once

#include <cstdint>

#ifndef KPZ_HD
#if defined(__HIPCC__)
#define KPZ_HD __host__ __device__
#else
#define KPZ_HD
#endif
#endif

namespace kpz {

// SLCG64 base step, inc/kmcRandom.h:59-60
constexpr uint64_t SLCG_A = 2862933555777941757ull;
constexpr uint64_t SLCG_C = 1442695040888963407ull;

// x -> a*x + c (mod 2^64)
struct Affine {
	uint64_t a, c;
	KPZ_HD uint64_t operator()(uint64_t x) const { return a * x + c; }
};

// g o f
KPZ_HD inline Affine compose(const Affine& g, const Affine& f) { return Affine{g.a * f.a, g.a * f.c + g.c}; }

// f applied k times
KPZ_HD inline Affine affine_pow(Affine f, uint64_t k)
{
	Affine r{1, 0};
	while(k)
	{
		if(k & 1)
			r = compose(f, r);
		f = compose(f, f);
		k >>= 1;
	}
	return r;
}

// Reference skip generator (kpz/kpzConst.h:85-87, inc/kmcRandom.h:74-117):
// A_S = A^S, C_S = C * sum_{k=1..S} A^k.  Note C_S = A * (true S-step c).
inline Affine slcg_skip(uint64_t S)
{
	const Affine j = affine_pow(Affine{SLCG_A, SLCG_C}, S);
	return Affine{j.a, SLCG_A * j.c};
}

// Skip length for T streams: the reference constant 30721 when T < 30721
// (kpz/GPURandom.cu:44 asserts T < S), else the smallest odd S > T so that
// A_S = 5 mod 8 and C_S is odd (full period 2^64).
inline uint64_t default_skip(uint64_t T) { return T < 30721 ? 30721 : T + 1 + (T & 1); }

// output fold of kpz/kpzRandom.cu:62-63
KPZ_HD inline uint32_t slcg_fold(uint64_t s) { return (uint32_t)(s ^ (s >> 32)); }

enum Encoding { ENC_LOCAL = 0, ENC_CBBIT = 1 };

struct Size {
	int lx, ly;
	KPZ_HD uint64_t sites() const { return 1ull << (lx + ly); }
	KPZ_HD uint64_t sizeW() const { return 1ull << (lx + ly - 4); }
	KPZ_HD int dimX() const { return 1 << lx; }
	KPZ_HD int dimY() const { return 1 << ly; }
};

// randomFloat() < p  <=>  rf < thresh(p), exactly: randomFloat = rf * 2^-bits
// compared in double (kpz/schedulerCPUBaseThreadedRng.inl:34-44); 23 bits for
// SLCG64 ((lo ^ hi) & 0x7FFFFF), 24 for TinyMT32 (tempered >> 8)
uint32_t thresh23(double p);
uint32_t thresh24(double p);

} // namespace kpz
