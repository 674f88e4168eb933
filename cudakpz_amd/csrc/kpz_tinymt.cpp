#include "kpz_tinymt.hpp"

namespace kpz {

// Krylov sequence v_0, v_1, ... of the 128-bit status; the first v_k in the
// span of the earlier ones gives m = x^k + sum_{i<k} a_i x^i with
// v_k = sum a_i v_i.  Each reduced vector carries its combination of the v_i.
TmtOrbit tmt_orbit(const uint32_t s0[4], uint32_t mat1, uint32_t mat2)
{
	struct Row {
		uint64_t v[2];
		uint64_t c[3];
		bool used;
	};
	Row piv[128] = {};
	uint32_t s[4] = {s0[0], s0[1], s0[2], s0[3]};
	for(int k = 0; k <= 128; ++k)
	{
		uint64_t v[2] = {s[0] | ((uint64_t)s[1] << 32), s[2] | ((uint64_t)s[3] << 32)};
		uint64_t c[3] = {0, 0, 0};
		c[k >> 6] |= 1ull << (k & 63);
		while(v[0] | v[1])
		{
			const int p = v[1] ? 127 - __builtin_clzll(v[1]) : 63 - __builtin_clzll(v[0]);
			if(!piv[p].used)
			{
				piv[p] = Row{{v[0], v[1]}, {c[0], c[1], c[2]}, true};
				break;
			}
			for(int j = 0; j < 2; ++j)
				v[j] ^= piv[p].v[j];
			for(int j = 0; j < 3; ++j)
				c[j] ^= piv[p].c[j];
		}
		if(!(v[0] | v[1]))
		{
			// sum over the set bits i of c of v_i = 0; bit k is set
			TmtOrbit m{};
			m.deg = (uint32_t)k;
			for(int i = 0; i < k; ++i)
				if((c[i >> 6] >> (i & 63)) & 1)
					m.low[i >> 5] |= 1u << (i & 31);
			return m;
		}
		tmt_step4(s, mat1, mat2);
	}
	return TmtOrbit{}; // not reached: 129 vectors of 128 bits are dependent
}

} // namespace kpz
