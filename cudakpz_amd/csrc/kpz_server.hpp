// Serving loop: simulation jobs arrive at any time, are admitted into slots
// as HBM frees up (FIFO with backfill, by footprint), and every active job
// advances in the same batched launches (continuous batching).  Lattices and
// RNG state stay resident in one device arena managed by a buddy allocator,
// so admission and retirement never call hipMalloc/hipFree.  Each job is
// bit-identical to a stand-alone scheduler of its kind, size, seed and (p, q)
// (the batch kernels' guarantee); it returns its exact w^2(t) series at the
// log-spaced times ceil(exp(log(t+10)+0.1)) (kpz/sampling.h:84-100) and at
// its final MCS.
#pragma once

#include "kpz_batch.hpp"

#include <chrono>
#include <deque>
#include <unordered_map>
#include <vector>

namespace kpz {

// power-of-two buddy allocator over one device buffer
class BuddyArena {
public:
	void init(size_t bytes, size_t minBlock);
	bool alloc(size_t bytes, uint64_t& off);
	void release(uint64_t off);
	size_t used() const { return used_; }
	size_t capacity() const { return cap_; }
	char* base() const { return buf_.as<char>(); }

private:
	DevBuf buf_;
	int minOrder_ = 12, maxOrder_ = 12;
	size_t cap_ = 0, used_ = 0;
	std::vector<std::vector<uint64_t>> free_;
	std::unordered_map<uint64_t, int> live_;
};

struct JobSpec {
	int kind;        // KIND_DT, KIND_DTR, KIND_SCABIT
	int lx, ly;
	int mcs;         // MCS to run
	double p, q;
	uint64_t seed;
	int keepLattice; // return the final lattice
};

struct JobSample {
	uint32_t mcs;
	int64_t sumH;
	__int128 sumH2;
};

struct JobResult {
	uint64_t id = 0;
	JobSpec spec{};
	std::vector<JobSample> series;
	std::vector<uint32_t> lattice;
	double tSubmit = 0, tStart = 0, tEnd = 0; // seconds since server start
};

// pinned host memory handed out by a bump pointer; blocks are kept (and
// reused after reset()) so that the steady state allocates nothing
class PinnedStage {
public:
	PinnedStage() = default;
	~PinnedStage();
	PinnedStage(const PinnedStage&) = delete;
	PinnedStage& operator=(const PinnedStage&) = delete;
	void reset();
	void* take(size_t bytes);

private:
	struct Block {
		char* p;
		size_t size, used;
	};
	std::vector<Block> blocks_;
	size_t cur_ = 0;
};

class Server {
public:
	Server(size_t hbmBudget, int maxSlots, int maxTickMcs = 16, int device = -1);
	~Server();
	Server(const Server&) = delete;
	Server& operator=(const Server&) = delete;

	uint64_t submit(const JobSpec& j);
	// one tick, pipelined: plan and enqueue the next tick (admission, tables,
	// master draws, launches) while the tick enqueued by the previous call is
	// still running, then wait for that previous tick and finish it (samples,
	// retirement).  Returns the number of jobs still queued or active.
	size_t step();
	bool pop(JobResult& out);
	size_t queued() const { return queue_.size(); }
	size_t active() const { return nActive_; }
	size_t hbmUsed() const { return arena_.used(); }
	size_t hbmCapacity() const { return arena_.capacity(); }
	uint64_t siteUpdates() const { return siteUpdates_; }
	// device time spent in ticks and the span from the first tick's start to
	// the last finished tick's end, in seconds (hipEvents on s0_, into which
	// each tick's second chain joins)
	void busy(double& gpuSeconds, double& spanSeconds) const;
	double now() const;

private:
	struct Job {
		uint64_t id = 0;
		JobSpec spec{};
		int slot = -1;
		uint64_t off[3] = {0, 0, 0}; // lattice, streams, lane table
		int64_t T = 0;
		MasterRng master;
		uint32_t planned = 0;        // MCS enqueued
		uint32_t done = 0;           // MCS finished (device work complete)
		std::vector<uint32_t> samples;
		size_t nextPlan = 0;         // next sample time to plan
		bool inTable = false;        // in the current device tables
		int group = 0;               // launch chain (stream) in the current tables
		JobResult res;
		BatchRep dev{};
	};
	struct Sample {
		int m;      // MCS index inside its tick
		Job* job;
	};
	// one tick in flight: its pinned staging, sample plan, per-job MCS and events
	struct Tick {
		bool live = false;
		PinnedStage stage;
		std::vector<Sample> plan;
		int64_t* out = nullptr; // pinned: 3 int64 per planned sample
		std::vector<std::pair<Job*, int>> ran; // (job, MCS it runs in this tick)
		hipEvent_t begin = nullptr, end = nullptr;
	};
	static size_t footprint(const JobSpec& s, size_t* parts);
	bool admit(Job& j, PinnedStage& stage);
	void rebuild(PinnedStage& stage);
	void ensure(DevBuf& b, size_t bytes);
	bool enqueue(Tick& t);
	void finish(Tick& t);

	BuddyArena arena_;
	int maxSlots_, maxTick_;
	int device_ = 0;
	// a tick's launches run in two chains of jobs of about equal site count,
	// s0_ and s1_, forked after the tick's uploads and joined before its
	// downloads, so one chain's launch tails overlap the other's next launch
	hipStream_t s0_ = nullptr, s1_ = nullptr;
	hipEvent_t fork_ = nullptr, join_ = nullptr;
	std::deque<Job*> queue_;
	std::vector<Job*> slots_;
	std::deque<JobResult> done_;
	size_t nActive_ = 0;
	bool dirty_ = true;
	uint64_t nextId_ = 1, siteUpdates_ = 0;
	// tables grow geometrically (rare hipMalloc); per-job memory comes from arena_
	DevBuf repTable_, laneTable_, waveTable_, rowTable_, rowScratch_, outBuf_, orderBuf_, shiftBuf_, identity_, sampleReps_,
			limitBuf_;
	std::vector<BatchRep> hostReps_;
	// per chain: DT lanes [laneOff_[g], + nLanes_[g]) (static DT first,
	// nStatic_[g] of them) and SCABit waves [waveOff_[g], + nWaves_[g])
	int64_t laneOff_[2] = {0, 0}, nLanes_[2] = {0, 0}, nStatic_[2] = {0, 0};
	int64_t waveOff_[2] = {0, 0}, nWaves_[2] = {0, 0};
	Tick ticks_[2];
	int cur_ = 0;
	double gpuMs_ = 0, spanMs_ = 0;
	bool started_ = false;
	hipEvent_t first_ = nullptr;
	std::chrono::steady_clock::time_point t0_;
};

} // namespace kpz
