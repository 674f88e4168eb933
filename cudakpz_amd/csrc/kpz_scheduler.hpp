// C++ host side of the engine: SchedulerService-style classes that own a
// device lattice, the per-worker SLCG streams and the master dSFMT, and
// enqueue the gfx950 kernels.  Names and semantics follow
// kpz/schedulerService.h:73-255 (mcs, mcsPQSyncRng, setPQ, initHomogeneous,
// pushSystem/popSystem, roughness, correlateTag/correlate, mcsScale).
#pragma once

#include "kpz_common.hpp"
#include "kpz_kernels.hpp"
#include "kpz_master.hpp"

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace kpz {

struct Error : std::runtime_error {
	int code;
	Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Kind { KIND_DT = 0, KIND_DTR = 1, KIND_SCABIT = 2, KIND_SCA = 3 };
// worker stream generators: SLCG64 (skip form) or TinyMT32, the generator the
// reference's frontend gives SchedulerCPUDTr / CB / CBBit
// (kpz/kpzFrontend.cpp:308-331)
enum Rng { RNG_SLCG64 = 0, RNG_TINYMT32 = 1 };

// device buffer owned by a scheduler
class DevBuf {
public:
	DevBuf() = default;
	explicit DevBuf(size_t bytes) { alloc(bytes); }
	~DevBuf();
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	void alloc(size_t bytes);
	void* get() const { return p_; }
	size_t bytes() const { return n_; }
	template<class T> T* as() const { return (T*)p_; }

private:
	void* p_ = nullptr;
	size_t n_ = 0;
};

void check(hipError_t e, const char* what);

struct Params {
	int64_t threads = 0;  // streams T = worker count of the CPU counterpart (0: engine default)
	int layout_x = 4;     // log2 DT block width  (SchedulerCPUDT::setLayout)
	int layout_y = 4;     // log2 DT block height
	uint64_t skip = 0;    // SLCG skip S (0: default_skip(T))
	int device = -1;      // HIP device (-1: current)
	int rng = RNG_SLCG64; // worker stream generator
};

// exact reductions: sums over all N sites
struct Sums {
	int64_t sumH = 0;
	__int128 sumH2 = 0;
	int64_t sumHs = 0;      // snapshot heights (correlate)
	__int128 sumHHs = 0;    // sum h * h_snapshot
	int64_t tcs = 0;        // slope overlap count (kpz/correlator.cpp:127-164)
};

class SchedulerService {
public:
	SchedulerService(int lx, int ly, Encoding enc, const Params& p);
	virtual ~SchedulerService();

	const Size& size() const { return size_; }
	Encoding encoding() const { return enc_; }
	int64_t threads() const { return T_; }
	uint64_t skip() const { return S_; }
	hipStream_t stream() const { return stream_; }

	// ---- state set-up (kpz/schedulerService.cpp:156-175, 460-495)
	virtual void initHomogeneous();          // flat initial condition, on device
	void supplySeed(uint64_t seed);          // reSeed master + randomize()
	virtual void randomize();                // stream states from two master draws
	virtual bool setPQ(double p, double q);
	double rateP() const { return p_; }
	double rateQ() const { return q_; }

	// ---- Disorder2 (kpz/schedulerService.cpp:48-69, kpz/schedulerService.h:185-197)
	// rates of disorder class n (0..3; class 0 = setPQ(p, q)); a class whose
	// rates were never set uses class 0's (contract: the reference leaves them
	// uninitialised)
	bool setPQ(double p, double q, unsigned n);
	double rateP(unsigned n) const { return n == 0 ? p_ : (cset_[n & 3] ? cp_[n & 3] : p_); }
	double rateQ(unsigned n) const { return n == 0 ? q_ : (cset_[n & 3] ? cq_[n & 3] : q_); }
	// generateDisorder2: per word 16 master draws, bit 2k set when draw k < fill
	void generateDisorder2(double fill);
	void setDisorder(const uint32_t* host);  // map in the lattice's layout
	void getDisorder(uint32_t* host);
	bool hasDisorder() const { return dis_.get() != nullptr; }
	// disorder2SyncRngUpdate dynamics (per-site class rates)
	virtual void mcsDisorder2SyncRng(int n);
	// SyncRngEW (kpz/kpzSimulationPrimitivesTemplates.cu:199-234): p = q = 1
	// without an acceptance draw; DT schedulers only
	virtual void mcsEW(int n);

	// ---- dynamics
	virtual void mcs(int n) = 0;             // default dynamics of the scheduler
	virtual void mcsSyncRng(int n) { mcs(n); }
	virtual void mcsPQSyncRng(int n) = 0;    // general p, q
	void mcsPQ(int n) { mcsPQSyncRng(n); }
	virtual double mcsScale() const { return 1.0; }
	unsigned int mcsCount() const { return mcs_; }
	void setMcs(unsigned int m) { mcs_ = m; }

	// ---- host copies (kpz/schedulerBase.cu:35-63)
	void pushSystem(const uint32_t* host);   // H2D of a reference-layout lattice
	void popSystem(uint32_t* host);          // D2H
	void sync();
	// SchedulerService::copySystem (kpz/schedulerService.cpp:497-504): device copy
	void copySystem(const SchedulerService& other);

	// ---- measurements (exact, on device)
	Sums roughness();
	// height moments <h^k>, k = 1..7, as Roughness::h[] after normAndSetW2
	// (kpz/roughness.cpp:38-43): per-row double sums, rows added in long double
	void moments(long double out[7]);
	// correlateTag(stop) (kpz/schedulerService.cpp:266-271): a device snapshot
	// kept while mcs <= stop; several tags may be live at once
	void correlateTag(unsigned int stop = 0xFFFFFFFFu);
	bool hasTag() const { return !tags_.empty(); }
	Sums correlate();                        // against the most recent tag
	struct TagResult {
		unsigned int tagMcs;
		Sums sums;
	};
	// every live tag (threadCorrelate, kpz/schedulerService.cpp:216-251):
	// tags whose stop time has passed are dropped first
	std::vector<TagResult> correlateAll();

	// SchedulerService::consitencyCheck (kpz/schedulerService.cpp:800-854):
	// out[0] rows whose x slopes, out[1] columns whose y slopes do not sum to
	// zero, out[2] loops that do not close; with full, every live correlator
	// tag's snapshot is checked too and its counts added in.  True if all are 0.
	bool consistencyCheck(bool full, int64_t out[3]);

	// ---- RNG state access (parity checks, checkpoints)
	int rng() const { return params_.rng; }
	bool tinyMT() const { return params_.rng == RNG_TINYMT32; }
	void getStreams(uint64_t* out);          // SLCG64 states (T)
	void setStreams(const uint64_t* in);
	// TinyMT32: parameter set t of stream t (the tinyMTmat1/2/tmat tables of
	// tinymt32_param.h, absent from the reference tree, so the caller's);
	// needed before randomize()
	void setTinyMTParams(const uint32_t* mat1, const uint32_t* mat2, const uint32_t* tmat);
	bool hasTinyMTParams() const { return tmtParams_; }
	// the TinyMT32 streams as 7 planes of T words: status[0..3], mat1, mat2, tmat
	void getTinyMT(uint32_t* out);
	void setTinyMT(const uint32_t* in);
	MasterRng& master() { return master_; }

	// ---- checkpoint / resume (kpz_state.cpp): everything needed to continue
	// bit for bit; aux is 32 bits of caller data stored with it
	int kind() const { return kind_; }
	// live correlator tags as host data, oldest first, and the way back
	struct TagImage {
		unsigned int mcs, stop;
		std::vector<uint32_t> snap; // reference-layout lattice at the tag's MCS
		int64_t sumH;               // sum of the snapshot's heights (exact)
	};
	std::vector<TagImage> exportTags();
	void importTag(unsigned int mcs, unsigned int stop, const uint32_t* snap);
	void saveState(const char* path, uint32_t aux);
	static std::unique_ptr<SchedulerService> loadState(const char* path, int device, uint32_t* aux);

	uint32_t* deviceLattice() const { return lat_.as<uint32_t>(); }

protected:
	void allocStreams(bool twoBuffers);
	uint32_t thresh(double p) const { return tinyMT() ? thresh24(p) : thresh23(p); }
	uint64_t* stIn() const { return st_[cur_].as<uint64_t>(); }
	uint64_t* stOut() const { return st_[cur_ ^ 1].as<uint64_t>(); }
	void swapStreams() { cur_ ^= 1; }
	const Affine* upload(std::vector<Affine>& host, DevBuf& buf);

	friend std::unique_ptr<SchedulerService> makeScheduler(int kind, int lx, int ly, const Params& p);
	int kind_ = -1;
	Params params_;
	Size size_;
	Encoding enc_;
	int64_t T_;
	uint64_t S_;
	Affine step_;
	int device_;
	hipStream_t stream_ = nullptr;
	struct Tag {
		std::unique_ptr<DevBuf> snap;
		unsigned int mcs, stop;
	};
	Sums correlateWith(const DevBuf& snap);
	Sums totals(const uint32_t* snap, int64_t* hs);
	DevBuf lat_, rows_, totals_, rowStart_, rowMoments_, check_;
	void consistencyOf(const uint32_t* lat, int64_t out[3]);
	std::vector<Tag> tags_;
	DevBuf st_[2];
	int cur_ = 0;
	// TinyMT32 streams: two sets of 7 planes (the current one and the one a
	// split launch writes the states after it into); orbit polynomials and
	// jump tables (x^n mod m per lane) for launches that start a stream
	// mid-sequence, both rebuilt when the streams are set
	DevBuf tmt_[2];
	int tmtCur_ = 0;
	bool tmtParams_ = false;
	uint32_t* tmtIn() const { return tmt_[tmtCur_].as<uint32_t>(); }
	uint32_t* tmtOut() const { return tmt_[tmtCur_ ^ 1].as<uint32_t>(); }
	DevBuf orbit_;
	bool orbitValid_ = false;
	std::map<uint64_t, std::unique_ptr<DevBuf>> tmtTab_;
	DevBuf rndBuf_;
	const TmtOrbit* tmtOrbits();
	void tmtStreamsChanged()
	{
		orbitValid_ = false;
		tmtTab_.clear();
	}
	MasterRng master_;
	double p_ = 1.0, q_ = 0.0;
	double cp_[4] = {0, 0, 0, 0}, cq_[4] = {0, 0, 0, 0};
	bool cset_[4] = {false, false, false, false};
	DevBuf dis_;
	void classThresholds(uint32_t tp[4], uint32_t tq[4]) const;
	unsigned int mcs_ = 0;
};

// Random-sequential double tiling: SchedulerCPUDT (static origin) and
// SchedulerCPUDTr (random origin) semantics on ENC_LOCAL.
class SchedulerDT : public SchedulerService {
public:
	SchedulerDT(int lx, int ly, const Params& p, bool randomOrigin);
	void mcs(int n) override;                // p=1, q=0 (pqOneUpdate)
	void mcsPQSyncRng(int n) override;       // per-update pq rule
	void mcsDisorder2SyncRng(int n) override;
	void mcsEW(int n) override;
	int layoutX() const { return lbx_; }
	int layoutY() const { return lby_; }
	int64_t blocksPerDomainSet() const { return nb_; }

private:
	void run(int n, int upd);
	const Affine* jumps(int du);
	int lbx_, lby_;
	bool ro_;
	int64_t nb_;
	std::map<int, std::unique_ptr<DevBuf>> jumpBuf_;
};

// Checkerboard SCA on ENC_CBBIT: SchedulerCPUCBBit semantics.
class SchedulerSCABit : public SchedulerService {
public:
	SchedulerSCABit(int lx, int ly, const Params& p);
	void mcs(int n) override;                // p=.5, q=0 bit-parallel (mcsSyncRng)
	void mcsPQSyncRng(int n) override;       // bit-parallel if (.5, 0), else per site
	void mcsDisorder2SyncRng(int n) override;
	double mcsScale() const override { return p_ + q_; }

private:
	struct Tables {
		DevBuf row, lane;
		Affine stepJ, fullBig, fullSmall;
	};
	void run(int n, int mode);
	Tables& tables(int mode);
	int per_, rem_, lG_;
	std::map<int, std::unique_ptr<Tables>> tab_;
};

// Checkerboard SCA on ENC_LOCAL: SchedulerCPUCB semantics (the CPU
// counterpart of the GPU SchedulerSCA).
class SchedulerSCA : public SchedulerService {
public:
	SchedulerSCA(int lx, int ly, const Params& p);
	void mcs(int n) override;                // SchedulerCPUCB::mcs: p=.95, q=0
	void mcsSyncRng(int n) override { mcsPQSyncRng(n); }
	void mcsPQSyncRng(int n) override;
	void mcsDisorder2SyncRng(int n) override;

private:
	void run(int n, bool disorder);
	CbLocalArgs args_{};
	DevBuf rowJ_, baseJ_, cellJ_, edgeR_, edgeU_;
};

std::unique_ptr<SchedulerService> makeScheduler(int kind, int lx, int ly, const Params& p);

} // namespace kpz
