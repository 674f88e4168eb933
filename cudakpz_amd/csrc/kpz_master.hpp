// Host-side master generator and helpers.
//
// MasterRng has the semantics of the dSFMT-19937 instance every
// SchedulerService owns (kpz/schedulerService.h:102, seeded by reSeed at
// kpz/schedulerService.cpp:460-470): the published double-precision SFMT
// recursion with the MEXP=19937 parameter set (inc/dSFMT/dSFMT-params19937.h),
// drawing doubles in [0, 1) as close_open().
#pragma once

#include <cstdint>
#include <cstring>

namespace kpz {

class MasterRng {
public:
	static constexpr int N = 191;       // 128-bit words of state
	static constexpr int N64 = 2 * N;   // doubles per refill

	MasterRng() { init(0); }

	void init(uint32_t seed);
	// SchedulerService::reSeed: init then discard 100000 draws
	void reseed(uint32_t seed)
	{
		init(seed);
		for(int a = 0; a < 100000; ++a)
			close_open();
	}
	double close_open()
	{
		if(idx_ >= N64)
		{
			refill();
			idx_ = 0;
		}
		double r;
		std::memcpy(&r, &w_[idx_++], sizeof r);
		return r - 1.0;
	}
	// Kmc::Shuffle<4> (inc/shuffle.h:76-88): one draw, four RC4-style swaps
	void shuffle4(int seq[4]);

	// raw state, for checkpoints and parity checks
	static constexpr size_t STATE_BYTES = sizeof(uint64_t) * (N64 + 2) + sizeof(int);
	void save(uint8_t* out) const;
	void load(const uint8_t* in);

private:
	void refill();
	uint64_t w_[N64 + 2]; // N 128-bit words + the "lung"
	int idx_;
};

// (contract) 64-bit user seed -> 32-bit dSFMT seed; identity below 2^32
inline uint32_t seed32(uint64_t seed) { return (uint32_t)(seed ^ (seed >> 32)); }

} // namespace kpz
