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

class Server {
public:
	Server(size_t hbmBudget, int maxSlots, int maxTickMcs = 16, int device = -1);
	~Server();
	Server(const Server&) = delete;
	Server& operator=(const Server&) = delete;

	uint64_t submit(const JobSpec& j);
	// one tick: admit, advance every active job up to its next sample, sample,
	// retire; returns the number of jobs still queued or active
	size_t step();
	bool pop(JobResult& out);
	size_t queued() const { return queue_.size(); }
	size_t active() const { return nActive_; }
	size_t hbmUsed() const { return arena_.used(); }
	size_t hbmCapacity() const { return arena_.capacity(); }
	uint64_t siteUpdates() const { return siteUpdates_; }
	double now() const;

private:
	struct Job {
		uint64_t id = 0;
		JobSpec spec{};
		int slot = -1;
		uint64_t off[3] = {0, 0, 0}; // lattice, streams, lane table
		int64_t T = 0;
		MasterRng master;
		uint32_t done = 0;
		std::vector<uint32_t> samples;
		size_t next = 0;
		JobResult res;
		BatchRep dev{};
	};
	static size_t footprint(const JobSpec& s, size_t* parts);
	bool admit(Job& j);
	void rebuild();
	void ensure(DevBuf& b, size_t bytes);

	BuddyArena arena_;
	int maxSlots_, maxTick_;
	int device_ = 0;
	hipStream_t s0_ = nullptr;
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
	std::vector<int> limitHost_;
	int64_t nLanes_ = 0, nWaves_ = 0;
	std::vector<uint8_t> orderHost_;
	std::vector<uint16_t> shiftHost_;
	std::chrono::steady_clock::time_point t0_;
};

} // namespace kpz
