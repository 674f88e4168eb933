// Batched replicas: many independent lattices of uneven sizes advanced
// together.  Each replica is bit-identical to a stand-alone SchedulerDT /
// SchedulerDT(random origin) / SchedulerSCABit with the same size, seed,
// (p, q) and default worker count: its streams depend only on its seed and
// size, never on its slot in the batch.
//
// One MCS of the whole batch is 4 DT-phase launches (every DT replica's
// domain set d) and 2 SCABit half-sweep launches, on two streams; chunks of
// MCS are captured once into a hipGraph and replayed.  The per-MCS domain
// orders (and DTr origins) come from each replica's own master dSFMT on the
// host, double-buffered so the next chunk is drawn while the GPU runs.
#pragma once

#include "kpz_kernels.hpp"
#include "kpz_master.hpp"
#include "kpz_scheduler.hpp"

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

namespace kpz {

struct ReplicaSpec {
	int kind;       // KIND_DT, KIND_DTR, KIND_SCABIT
	int lx, ly;
	double p, q;    // DT: (1, 0) = mcs (pqOneUpdate), else mcsPQSyncRng
	uint64_t seed;  // supplySeed(seed)
};

// performance knobs; none of them changes any replica's result
struct BatchTuning {
	int maxLanesLog2 = 3;  // SCABit: lanes per row = min(words per row, 2^this)
	int pqWords64 = 1;     // SCABit per-site replicas: 64-bit (1) or 32-bit (0) words
	int chunk = 50;        // MCS per captured graph
	int concurrent = 2;    // 0: one stream; 1: DT and SCABit chains on two streams;
	                       // 2: also the DT and the SCABit replicas each in two groups of
	                       // equal site count on streams of their own, so one group's
	                       // launch tail overlaps the other's next launch
	int sortBySize = 1;    // large replicas first in the launch order
};

class Batch {
public:
	// quickSeed skips the 100000-draw warm-up of the masters (timing trials only)
	Batch(const std::vector<ReplicaSpec>& specs, const BatchTuning& tune, int device = -1, bool quickSeed = false);
	~Batch();
	Batch(const Batch&) = delete;
	Batch& operator=(const Batch&) = delete;

	int size() const { return (int)reps_.size(); }
	const BatchTuning& tuning() const { return tune_; }
	void initHomogeneous();
	void mcs(int n);
	void roughness(int64_t* out); // 3 int64 per replica: sum h, sum h^2 lo, hi
	void download(int r, uint32_t* host);
	void upload(int r, const uint32_t* host);
	void getStreams(int r, uint64_t* out);
	int64_t threads(int r) const { return reps_.at(r).T; }
	uint64_t sizeWords(int r) const { return 1ull << (reps_.at(r).spec.lx + reps_.at(r).spec.ly - 4); }
	uint64_t sites() const { return sites_; }
	unsigned int mcsCount() const { return mcs_; }
	hipStream_t stream() const { return s0_; }
	void sync();

private:
	struct Rep {
		ReplicaSpec spec;
		uint64_t latOff = 0, stOff = 0;
		int64_t T = 0;
		MasterRng master;
		BatchRep dev{};
	};
	struct Chunk {
		hipGraphExec_t exec = nullptr;
		hipGraph_t graph = nullptr;
	};
	void buildTables();
	void drawChunk(int b, int len);
	Chunk& graphFor(int b, int len);

	std::vector<Rep> reps_;
	BatchTuning tune_;
	int device_ = 0;
	// s0_: DT group 0; s2_: DT group 1; s1_: SCABit group 0; s3_: SCABit group 1
	hipStream_t s0_ = nullptr, s1_ = nullptr, s2_ = nullptr, s3_ = nullptr;
	hipEvent_t fork_ = nullptr, join_[3] = {nullptr, nullptr, nullptr};
	DevBuf latArena_, stArena_, tabArena_, repTable_, laneTable_, waveTable_, rowTable_, rowScratch_, outBuf_;
	DevBuf orderBuf_[2], shiftBuf_[2];
	uint8_t* orderHost_[2] = {nullptr, nullptr};
	uint16_t* shiftHost_[2] = {nullptr, nullptr};
	hipEvent_t copied_[2] = {nullptr, nullptr};
	// replica groups (the second ones are empty unless concurrent == 2): DT
	// lanes [dtOff_[g], dtOff_[g] + dtLanes_[g]) of the lane table, static DT
	// first; SCABit waves [cbOff_[g], cbOff_[g] + cbWaves_[g]) of the wave table
	int64_t dtOff_[2] = {0, 0}, dtLanes_[2] = {0, 0}, dtStatic_[2] = {0, 0};
	int64_t cbOff_[2] = {0, 0}, cbWaves_[2] = {0, 0};
	int64_t nRows_ = 0, nRowWaves_ = 0; // site rows of all replicas, row-statistics waves
	bool hasDt_ = false, hasCb_ = false;
	uint64_t sites_ = 0;
	unsigned int mcs_ = 0;
	int cur_ = 0;
	std::map<std::pair<int, int>, Chunk> graphs_;
};

// Chooses the BatchTuning with the lowest time for `mcs` MCS on batches of
// the given shapes (coordinate descent over chunk, concurrency, lanes per row
// and SCABit word width).  Returns the best time in ms per MCS.
double autotuneBatch(const std::vector<ReplicaSpec>& specs, int mcs, int device, BatchTuning& best);

} // namespace kpz
