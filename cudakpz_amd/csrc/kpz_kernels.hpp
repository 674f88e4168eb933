// Launchers of the gfx950 kernels (kpz_kernels.hip).  All pointers are
// device pointers; every launcher enqueues on `stream` and returns the
// hipError_t of the launch.
#pragma once

#include "kpz_common.hpp"
#include "kpz_tinymt.hpp"

#include <hip/hip_runtime.h>

namespace kpz {

// ---- DT: one double-tiling domain set (phase) --------------------------
struct DtPhaseArgs {
	uint32_t* lat;          // ENC_LOCAL lattice
	int lx, ly, lbx, lby;   // log2 system and block sizes
	const uint64_t* st_in;  // stream states (T)
	uint64_t* st_out;       // == st_in when every stream owns <= 1 block
	int64_t T;              // streams (CPU worker count of the counterpart)
	const Affine* jumps;    // jumps[j]: advance to the j-th block of a stream
	Affine step;            // per-draw SLCG step (skip form)
	Affine blockJ;          // J(draws per block): a stream's state after one block
	int ds;                 // domain set 0..3
	int sx, sy;             // DT origin shift (0 for the static DT)
	int upd;                // 0 = KPZ (pqOneUpdate), 1 = pq (pqSyncRngUpdate),
	                        // 2 = disorder2SyncRngUpdate, 3 = EW (p = q = 1, no acceptance draw)
	bool random_origin;     // DTr coordinate/mask convention
	uint32_t tp, tq;        // 23-bit thresholds
	const uint32_t* dis;    // upd 2: disorder map (ENC_LOCAL layout, 2 bits = class per site)
	uint32_t tpc[4], tqc[4]; // upd 2: thresholds per class
	uint32_t* tmt;          // TinyMT32 streams (7 planes of T words; random origin only;
	                        // thresholds then on the 2^24 scale), or null for SLCG64
	// TinyMT32 with fewer streams than blocks: the states after the phase go
	// to tmtOut; lane c runs the tmtChain consecutive blocks t + T (k C + m),
	// m < C = tmtChain, of worker t = c % T, k = c / T, with the stream jumped
	// by tmtJ[c] = x^(k C draws per block) (launch_tmt_block_polys)
	uint32_t* tmtOut;
	const Poly128* tmtJ;
	const TmtOrbit* tmtOrb;
	int tmtChain;
};
// lanes of a TinyMT32 DT phase with T < nb streams and chains of C blocks
inline int64_t dt_tmt_lanes(int64_t nb, int64_t T, int C) { return T * (((nb + T - 1) / T + C - 1) / C); }
hipError_t launch_dt_phase(const DtPhaseArgs& a, hipStream_t stream);
// tmtJ for a DT phase with T < nb streams: entry c < n -> x^((c / T) * drawsPerEntry)
// mod the orbit polynomial of stream c % T (drawsPerEntry: the draws of a chain)
hipError_t launch_tmt_block_polys(const TmtOrbit* orb, int64_t T, int64_t n, uint64_t drawsPerEntry, Poly128* out,
		hipStream_t stream);

// ---- TinyMT32 stream sets: tinymt32_init of every stream from the seed in
// plane 0 and the parameters in planes 4..6 (7 planes of T words)
hipError_t launch_tmt_seed(uint32_t* g, int64_t T, hipStream_t stream);
bool dt_layout_supported(int lbx, int lby);

// ---- SCABit: one checkerboard half-sweep on ENC_CBBIT ------------------
struct CbHalfArgs {
	uint32_t* lat;
	int lx, ly, signum;
	int mode;               // 0 = bit-parallel (VL3), 1 = per-site pq
	const uint64_t* st_in;
	uint64_t* st_out;
	int64_t T;
	int per, rem;           // rows per worker, workers with one extra row
	const Affine* rowJ;     // rowJ[r]: skip r rows of draws
	const Affine* laneJ;    // laneJ[g]: skip g words of draws
	Affine stepJ;           // J(draws per word * G): a lane's word start to its next word start
	Affine step;            // per-draw step
	Affine fullBig, fullSmall; // whole half-sweep of a worker (per+1 / per rows)
	uint32_t tp, tq;
	int lG;                 // log2 lanes per row
	// mode 2 (disorder2SyncRngUpdate, kpz/schedulerCPUCBBit_nested.inl:298-331):
	// class = x-slope bit of the site in the map (ENC_CBBIT layout), class-1 thresholds
	const uint32_t* dis;
	uint32_t tp1, tq1;
	// modes 3 (VL3 random words) and 4 (per-site acceptance masks, minima then
	// maxima, Q words each) take the half-sweep's random part from this
	// buffer (the TinyMT32 workers, launch_cbbit_tmt_gen)
	const void* rnd;
};
hipError_t launch_cbbit_half(const CbHalfArgs& a, hipStream_t stream);
// TinyMT32 workers of SCABit: the random part of one half-sweep (mode 0 VL3,
// 1 per-site pq, 2 disorder) into `out`, one lane per 1/2^lC of a row;
// streams read from gin, the state after the half-sweep written to gout (7
// planes of T words each); rowJ (per lane) / orb (per worker) are needed when
// a lane starts mid-stream (lC > 0 or a worker with several rows)
hipError_t launch_cbbit_tmt_gen(const CbHalfArgs& a, int mode, int lC, const uint32_t* gin, uint32_t* gout,
		const Poly128* rowJ, const TmtOrbit* orb, void* out, hipStream_t stream);
// rowJ for launch_cbbit_tmt_gen: x^(r * drawsPerRow + c * drawsPerRow / 2^lC)
// mod the orbit polynomial of row y's worker, r the row's index among the
// worker's rows, for lane (y << lC) + c
hipError_t launch_tmt_row_polys(const TmtOrbit* orb, int64_t nrows, int lC, int per, int rem, uint64_t drawsPerRow,
		Poly128* out, hipStream_t stream);

// ---- SCA on ENC_LOCAL: one checkerboard half-sweep (SchedulerCPUCB) --------
// Worker t owns cell columns [c0_t, c0_t + w_t), w_t = pc (+1 for t < rem);
// its strip is split into halves of w/2 and w - w/2 cells, walked one after
// the other, each block row (4 site rows) by block row, cell by cell, 8 draws
// per cell.  V[] lists the distinct half widths (cells); rowJ[k*nyb + yb] =
// J(yb*V[k]*8), baseJ[k] = J(nyb*V[k]*8), cellJ[i] = J(8 i).
struct CbLocalArgs {
	uint32_t* lat;
	int lx, ly, signum;
	const uint64_t* st_in;
	uint64_t* st_out;
	int pc, rem, nyb;
	int V[4], nV;
	const Affine* rowJ;
	const Affine* baseJ;
	const Affine* cellJ;
	Affine fullBig, fullSmall; // J(nyb * w * 8) for w = pc+1 / pc
	Affine step;
	uint32_t tp, tq;
	int lG;                 // log2 cell columns per lane group (<= 6, <= lx - 2)
	int lNB;                // log2 cell rows per band (<= ly - 2)
	uint32_t* edgeR;        // [nyb][groups per row]: flips leaving a group's right edge
	uint32_t* edgeU;        // [bands][cells per row]: flips leaving a band's top row
	// disorder2SyncRngUpdate (kpz/schedulerCPUCB.cpp:103-108): map in ENC_LOCAL
	// layout (null: plain pq) and thresholds per class
	const uint32_t* dis;
	uint32_t tpc[4], tqc[4];
	// acceptance words from a buffer instead of drawn (the TinyMT32 workers,
	// launch_cb_local_tmt_gen), or null
	const uint32_t* acc;
	int64_t T;
};
// one half-sweep: the band kernel plus the edge fix-up (two launches)
hipError_t launch_cb_local_half(const CbLocalArgs& a, hipStream_t stream);
// TinyMT32 workers of SCA: the acceptance words of one half-sweep into `acc`
// (lattice layout; thresholds on the 2^24 scale), one lane per chunk of a
// worker's walk; streams read from gin, the states after the half-sweep
// written to gout; chunkJ / orb needed when a walk has several chunks
hipError_t launch_cb_local_tmt_gen(const CbLocalArgs& a, const uint32_t* gin, uint32_t* gout, const Poly128* chunkJ,
		const TmtOrbit* orb, uint32_t* acc, hipStream_t stream);
// chunks per worker for launch_cb_local_tmt_gen, and its jump table:
// out[t * chunks + k] = x^(k * drawsPerChunk) mod the orbit polynomial of worker t
int cb_tmt_chunks(int lx, int ly, int64_t T);
constexpr uint64_t CB_TMT_CHUNK_DRAWS = 256 * 8;
hipError_t launch_tmt_chunk_polys(const TmtOrbit* orb, int64_t T, int chunks, uint64_t drawsPerChunk, Poly128* out,
		hipStream_t stream);

// ---- reductions ----------------------------------------------------------
// Per site row y, relative to the row's start height (see kpz_kernels.hip):
// out[y*ROW_STATS + k] = {S1, S2, D, sy0, S1', D', sy0', S11, tcs, 0}, the
// primed values from `snap` (zero when snap is null).
constexpr int ROW_STATS = 10;
// out[y*7 + k-1] = sum over row y of h^k (double), k = 1..7, given the exact
// start height hs[y] of every row
hipError_t launch_row_moments(const uint32_t* lat, int enc, int lx, int ly, const int64_t* hs, double* out,
		hipStream_t stream);
// consistency of a lattice (SchedulerService::consitencyCheck): rowOnes[y] +=
// ones among row y's x slopes, colOnes[x] = ones among column x's y slopes,
// badLoops += loops that do not close (rowOnes and badLoops cleared by the
// caller); part: consistency_part_words(lx, ly) u16 of per-band column counts
int64_t consistency_part_words(int lx, int ly);
hipError_t launch_consistency(const uint32_t* lat, int enc, int lx, int ly, int* rowOnes, int* colOnes,
		uint16_t* part, unsigned long long* badLoops, hipStream_t stream);
hipError_t launch_row_stats(const uint32_t* lat, const uint32_t* snap, int enc, int lx, int ly,
		int64_t* out, hipStream_t stream);
// exact totals from the row statistics: res[7] = {sum h, sum h^2 lo, hi,
// sum h_snap, sum h h_snap lo, hi, tcs} (the last four with corr only); hs
// (may be null): every row's start height; part: row_combine_part_words(ly)
// int64 of per-workgroup partial sums
int64_t row_combine_part_words(int ly);
hipError_t launch_row_combine(const int64_t* rows, int lx, int ly, bool corr, int64_t* hs, int64_t* part,
		int64_t* res, hipStream_t stream);
// Roughness totals {sum h, sum h^2 lo, hi} of a lattice in caller-owned device
// memory into out (device, 3 int64), queued on stream, with every intermediate
// in a fixed per-device workspace of noalloc_workspace_bytes() (a static device
// array): no allocation.  Rows up to 2^NOALLOC_MAX_LY.  Calls that overlap on
// one device share the workspace and race.
constexpr int NOALLOC_MAX_LY = 16;
int64_t noalloc_workspace_bytes();
hipError_t launch_roughness_noalloc(const uint32_t* lat, int enc, int lx, int ly, int64_t* out, hipStream_t stream);

// ---- batched replicas ------------------------------------------------------
// One device-table entry per replica of a batch.  DT replicas use one stream
// per block (in place); SCABit replicas one stream per row (ping-pong st[0/1]).
enum BatchKind { BK_DT = 0, BK_DTR = 1, BK_SCABIT = 2 };
struct BatchRep {
	uint32_t* lat;
	uint64_t* st[2];
	const Affine* laneJ;    // SCABit: laneJ[g] = J(g * draws per word)
	const Affine* rowJ;     // SCABit: {identity} (one stream per row)
	Affine step;            // per-draw step (skip form)
	Affine stepJ;           // SCABit: J(draws per word * G)
	Affine full;            // SCABit: J(draws per row); DT: J(draws per block)
	int lx, ly;
	int kind;               // BatchKind
	int mode;               // DT: 0 KPZ (pqOneUpdate), 1 pq; SCABit: 0 bit-parallel, 1 per-site
	int lG;                 // SCABit: log2 lanes per row
	int wide;               // SCABit: 64-bit words (else 32-bit)
	uint32_t tp, tq;
	int64_t rowBase;        // first row of this replica in batch row tables
};

struct BatchDtArgs {
	const BatchRep* reps;
	const uint2* lanes;     // per lane {replica, block}; replica ~0u = idle lane
	const uint8_t* order;   // [m * nreps + r]: domain order, 2 bits per position
	const uint16_t* shift;  // [m * nreps + r]: DTr origin shift, sx | sy << 8
	int nreps;
	int64_t nlanes;
	int64_t nstatic;        // lanes [0, nstatic) hold static DT replicas, the rest DTr
	                        // (the tables put static DT first, padded to whole waves)
	int m, d;               // MCS within the chunk, position in the domain order
	const int* limit;       // optional [r]: replica r only runs MCS m < limit[r]
};
hipError_t launch_batch_dt(const BatchDtArgs& a, hipStream_t stream);

struct BatchCbArgs {
	const BatchRep* reps;
	const uint2* waves;     // per wave {replica, first row}
	int64_t nwaves;
	int signum;
	int buf;                // read st[buf], write st[buf ^ 1]
	int m;                  // MCS within the chunk
	const int* limit;       // optional [r]: replica r only runs MCS m < limit[r]
};
hipError_t launch_batch_cbbit(const BatchCbArgs& a, hipStream_t stream);

// lanes per site row in the batched row statistics: one per 64-site chunk, at
// most 64; a wave takes 64 / batch_row_lanes(lx) consecutive rows of a replica
inline int batch_row_lanes(int lx) { return lx >= 12 ? 64 : (lx >= 6 ? 1 << (lx - 6) : 1); }

// exact roughness of every replica: waves[] = per-wave {replica, first row},
// covering every row of every replica (64 / batch_row_lanes(lx) rows per
// wave), scratch = 4 int64 per row, out[3 r..] = {sum h, sum h^2 lo, hi}
hipError_t launch_batch_roughness(const BatchRep* reps, int nreps, const uint2* waves, int64_t nwaves,
		int64_t* scratch, int64_t* out, hipStream_t stream);

// ---- utilities -------------------------------------------------------------
hipError_t launch_fill_flat(uint32_t* lat, int enc, int lx, int ly, hipStream_t stream);
// dst <- src, 16-byte vector stream copy (bytes must be a multiple of 16)
hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t stream);
// integer VALU issue-rate probe: blocks x 256 lanes x iters x 8 v_xad_u32
// (fast: v_xor_b32); out needs blocks * 256 u32
hipError_t launch_valu_probe(uint32_t* out, unsigned blocks, int iters, bool fast, hipStream_t stream);
// blocks of 8 independent instructions, `slow` (0, 2, 3, 4, 5 or 8) of them
// v_mul_lo_u32 and the rest v_xor_b32, 8 chains per lane
hipError_t launch_valu_mix(uint32_t* out, unsigned blocks, int iters, int slow, hipStream_t stream);

} // namespace kpz
