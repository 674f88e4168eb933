// TinyMT32 jump ahead.  The state transition of TinyMT32
// (inc/tinyMT/tinymt32.h:70-85) is linear over GF(2) on the 128 status bits
// (status[0]'s top bit is dropped by the next step), so the state n steps on
// is p(A) v for the polynomial p = x^n mod m, m the minimal polynomial of the
// orbit of v: v_n = sum_i p_i v_i with v_i the states 0..deg m - 1 steps on.
// m is found once per stream from its first 129 states by elimination over
// GF(2) (it stays valid along the orbit); x^n mod m by square and multiply.
// The tempering (not linear) is applied to jumped states as to stepped ones.
#pragma once

#include "kpz_common.hpp"

#include <cstdint>

namespace kpz {

// polynomial over GF(2) of degree < 128, bit i = coefficient of x^i
struct Poly128 {
	uint32_t w[4];
};

// a stream's orbit polynomial m = x^deg + low (deg <= 128)
struct TmtOrbit {
	uint32_t low[4];
	uint32_t deg;
};

KPZ_HD inline void tmt_step4(uint32_t s[4], uint32_t mat1, uint32_t mat2)
{
	uint32_t y = s[3];
	uint32_t x = (s[0] & 0x7FFFFFFFu) ^ s[1] ^ s[2];
	x ^= x << 1;
	y ^= (y >> 1) ^ x;
	const uint32_t m = 0u - (y & 1u);
	s[0] = s[1];
	s[1] = s[2] ^ (m & mat1);
	s[2] = x ^ (y << 10) ^ (m & mat2);
	s[3] = y;
}

KPZ_HD inline bool poly_bit(const uint32_t p[4], uint32_t i) { return (p[i >> 5] >> (i & 31)) & 1u; }

// r = a * b mod m (a, b of degree < deg)
KPZ_HD inline Poly128 poly_mulmod(const Poly128& a, const Poly128& b, const TmtOrbit& m)
{
	Poly128 r{{0, 0, 0, 0}};
	if(m.deg == 0)
		return r;
	const uint32_t top = m.deg - 1;
	for(int i = (int)top; i >= 0; --i)
	{
		// r = r * x mod m
		const bool carry = poly_bit(r.w, top);
		r.w[3] = (r.w[3] << 1) | (r.w[2] >> 31);
		r.w[2] = (r.w[2] << 1) | (r.w[1] >> 31);
		r.w[1] = (r.w[1] << 1) | (r.w[0] >> 31);
		r.w[0] <<= 1;
		if(top < 127) // drop x^deg (for deg = 128 it left the 128 bits)
			r.w[(top + 1) >> 5] &= ~(1u << ((top + 1) & 31));
		if(carry)
			for(int k = 0; k < 4; ++k)
				r.w[k] ^= m.low[k];
		if(poly_bit(a.w, (uint32_t)i))
			for(int k = 0; k < 4; ++k)
				r.w[k] ^= b.w[k];
	}
	return r;
}

// x^n mod m
KPZ_HD inline Poly128 poly_xpow(uint64_t n, const TmtOrbit& m)
{
	Poly128 r{{0, 0, 0, 0}}, b{{0, 0, 0, 0}};
	if(m.deg == 0)
		return r;
	// r = 1, b = x (both reduced: deg 1 means x = m.low)
	r.w[0] = 1;
	if(m.deg == 1)
	{
		b = Poly128{{m.low[0], m.low[1], m.low[2], m.low[3]}};
		r.w[0] = 1;
	}
	else
		b.w[0] = 2;
	while(n)
	{
		if(n & 1)
			r = poly_mulmod(r, b, m);
		n >>= 1;
		if(n)
			b = poly_mulmod(b, b, m);
	}
	return r;
}

// v <- p(A) v
KPZ_HD inline void tmt_jump4(uint32_t s[4], const Poly128& p, const TmtOrbit& m, uint32_t mat1, uint32_t mat2)
{
	uint32_t acc[4] = {0, 0, 0, 0};
	for(uint32_t i = 0; i < m.deg; ++i)
	{
		if(poly_bit(p.w, i))
			for(int k = 0; k < 4; ++k)
				acc[k] ^= s[k];
		tmt_step4(s, mat1, mat2);
	}
	for(int k = 0; k < 4; ++k)
		s[k] = acc[k];
}

// the minimal polynomial of the orbit of s (host)
TmtOrbit tmt_orbit(const uint32_t s[4], uint32_t mat1, uint32_t mat2);

} // namespace kpz
