#include "kpz_master.hpp"
#include "kpz_common.hpp"

#include <cmath>

namespace kpz {

namespace {
// MEXP = 19937 parameters (inc/dSFMT/dSFMT-params19937.h:6-16)
constexpr int POS1 = 117;
constexpr int SL1 = 19;
constexpr int SR = 12;
constexpr uint64_t MSK[2] = {0x000ffafffffffb3full, 0x000ffdfffc90fffdull};
constexpr uint64_t FIX[2] = {0x90014964b32f4329ull, 0x3b8d12ac548a7c7aull};
constexpr uint64_t PCV[2] = {0x3d84e1ac0dc82880ull, 0x0000000000000001ull};
constexpr uint64_t MANT = 0x000FFFFFFFFFFFFFull;
constexpr uint64_t ONE = 0x3FF0000000000000ull;
} // namespace

void MasterRng::init(uint32_t seed)
{
	// 32-bit initialisation of the whole state (lung included)
	uint32_t w32[(N + 1) * 4];
	w32[0] = seed;
	for(uint32_t i = 1; i < (N + 1) * 4; ++i)
		w32[i] = 1812433253u * (w32[i - 1] ^ (w32[i - 1] >> 30)) + i;
	std::memcpy(w_, w32, sizeof w32);
	// map to doubles in [1, 2)
	for(int i = 0; i < N64; ++i)
		w_[i] = (w_[i] & MANT) | ONE;
	// period certification of the lung
	uint64_t inner = ((w_[N64] ^ FIX[0]) & PCV[0]) ^ ((w_[N64 + 1] ^ FIX[1]) & PCV[1]);
	for(int sh = 32; sh; sh >>= 1)
		inner ^= inner >> sh;
	if(!(inner & 1))
		w_[N64 + 1] ^= 1;
	idx_ = N64;
}

void MasterRng::refill()
{
	uint64_t l0 = w_[N64], l1 = w_[N64 + 1];
	for(int i = 0; i < N; ++i)
	{
		const int j = (i + POS1 < N) ? i + POS1 : i + POS1 - N;
		const uint64_t t0 = w_[2 * i], t1 = w_[2 * i + 1];
		const uint64_t n0 = (t0 << SL1) ^ (l1 >> 32) ^ (l1 << 32) ^ w_[2 * j];
		const uint64_t n1 = (t1 << SL1) ^ (l0 >> 32) ^ (l0 << 32) ^ w_[2 * j + 1];
		l0 = n0;
		l1 = n1;
		w_[2 * i] = (n0 >> SR) ^ (n0 & MSK[0]) ^ t0;
		w_[2 * i + 1] = (n1 >> SR) ^ (n1 & MSK[1]) ^ t1;
	}
	w_[N64] = l0;
	w_[N64 + 1] = l1;
}

void MasterRng::shuffle4(int seq[4])
{
	for(int a = 0; a < 4; ++a)
		seq[a] = a;
	uint32_t bits = (uint32_t)(close_open() * 4294967296.0);
	for(int a = 0; a < 4; ++a, bits >>= 2)
	{
		const int i = bits & 3;
		const int v = seq[i];
		seq[i] = seq[a];
		seq[a] = v;
	}
}

void MasterRng::save(uint8_t* out) const
{
	std::memcpy(out, w_, sizeof w_);
	std::memcpy(out + sizeof w_, &idx_, sizeof idx_);
}

void MasterRng::load(const uint8_t* in)
{
	std::memcpy(w_, in, sizeof w_);
	std::memcpy(&idx_, in + sizeof w_, sizeof idx_);
}

static uint32_t thresh_bits(double p, int bits)
{
	if(!(p > 0.0))
		return 0;
	if(p >= 1.0)
		return 1u << bits;
	return (uint32_t)std::ceil(p * (double)(1u << bits));
}

uint32_t thresh23(double p) { return thresh_bits(p, 23); }
uint32_t thresh24(double p) { return thresh_bits(p, 24); }

} // namespace kpz
