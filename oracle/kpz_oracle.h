/*
 * ORACLE -- test infrastructure only.
 *
 * Plain-C restatement of the CPU counterparts of the CudaKpz hot path, used
 * as the checker for the HIP engine.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library; the product
 * (cudakpz_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Where the
 * reference is ambiguous the contract choice is marked "(contract)".
 *
 * Pinning status (how each function is checked; see tests/test_oracle_*.py):
 *   kor_dsfmt_*          golden vectors from the reference's own dSFMT.cpp
 *                        compiled into oracle/_ref (tests/golden/dsfmt.json)
 *   kor_shuffle4         golden vectors from the reference's Kmc::Shuffle<4>
 *   kor_dt_block_offset  golden vectors from the reference's
 *                        Kmc::DoubleTilingDD<2>::setBlock
 *   kor_slcg_skip        known-answer constants of inc/helper/slcgInst.h and
 *                        the reference's own SLCGskipA/SLCGskipC
 *   kor_slcg_jump        closed form vs. step-by-step iteration
 *   kor_slcg_draw fold,  bit counts and histograms of 2^30 draws printed by
 *   23-bit randomFloat   the reference's own inc/slcgProfile.cpp, built
 *                        unchanged into oracle/_ref
 *                        (tests/golden/slcg_profile.txt)
 *   update rules         brute-force explicit-height octahedron simulator
 *   half-sweeps          per-site scalar equivalence, order independence
 *   heights/roughness    closed forms (flat IC), brute force on tiny lattices
 *   kor_tmt_*            golden vectors from the reference's vendored
 *                        inc/tinyMT/tinymt32.c compiled into oracle/_ref
 *                        (tests/golden/tinymt32.json)
 *   DT/DTr/CBBit/CB/DB   composition of the pinned pieces above plus
 *                        invariants (consistency, height-change accounting);
 *                        the TinyMT32 variants share every line with the
 *                        SLCG64 ones except the stream cursor
 *
 * Parity unpinned (no reference-built code, fixture or closed form reaches
 * them in this image; each restates its cited lines and is tested only
 * against that reading):
 *   kor_streams_randomize  kpz/GPURandom.cu:42-54 (CUDA source)
 *   kor_reseed's 100000    kpz/schedulerService.cpp:460-470 (needs libSplash)
 *   discarded draws
 *   kor_generate_disorder2 kpz/schedulerService.cpp:54-69 (needs libSplash)
 */
#ifndef KPZ_ORACLE_H
#define KPZ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- master RNG: dSFMT-19937 ---------------- */
#define KOR_DSFMT_N 191
typedef struct kor_dsfmt {
	union {
		uint64_t u64[(KOR_DSFMT_N + 1) * 2];
		uint32_t u32[(KOR_DSFMT_N + 1) * 4];
	} st;
	int idx;
} kor_dsfmt;

void kor_dsfmt_init(kor_dsfmt* d, uint32_t seed);
double kor_dsfmt_close_open(kor_dsfmt* d);
/* SchedulerService::reSeed: init + 100000 discarded draws */
void kor_reseed(kor_dsfmt* d, uint32_t seed);
/* (contract) 64-bit seed -> 32-bit dSFMT seed: lo ^ hi */
uint32_t kor_seed32(uint64_t seed);
void kor_shuffle4(kor_dsfmt* d, int seq[4]);
int kor_dsfmt_size(void);

/* ---------------- SLCG64 ---------------- */
#define KOR_SLCG_A 2862933555777941757ull
#define KOR_SLCG_C 1442695040888963407ull
void kor_slcg_skip(uint64_t S, uint64_t* a, uint64_t* c);
void kor_slcg_jump(uint64_t a, uint64_t c, uint64_t k, uint64_t* ak, uint64_t* ck);
uint64_t kor_default_skip(uint64_t T);
void kor_streams_randomize(kor_dsfmt* master, uint64_t* states, int64_t T);
uint32_t kor_slcg_draw(uint64_t* s, uint64_t a, uint64_t c);
/* bit counts and histograms of n folded base-generator draws from `state`,
 * as inc/slcgProfile.cpp:41-60 prints them */
void kor_slcg_profile(uint64_t state, uint64_t n, int nthreads, uint64_t bits[32], uint64_t j[12], uint64_t j2[12]);

/* ---------------- TinyMT32 ---------------- */
/* tinymt32_t field order (inc/tinyMT/tinymt32.h:80-85) */
typedef struct kor_tmt {
	uint32_t status[4];
	uint32_t mat1, mat2, tmat;
} kor_tmt;
/* tinymt32_init: mat1, mat2, tmat must be set */
void kor_tmt_init(kor_tmt* g, uint32_t seed);
uint32_t kor_tmt_u32(kor_tmt* g);     /* tinymt32_generate_uint32 */
float kor_tmt_float(kor_tmt* g);      /* tinymt32_generate_float */
/* RngIface::TinyMT32::randomize: generator t seeded with
 * (unsigned)(close_open() * UINT_MAX), t = 0..T-1 */
void kor_tmt_randomize(kor_dsfmt* master, kor_tmt* g, int64_t T);

/* ---------------- lattice ---------------- */
void kor_init_flat_local(uint32_t* lat, int lx, int ly);
void kor_init_flat_cbbit(uint32_t* lat, int lx, int ly);
void kor_local_to_cbbit(const uint32_t* in, uint32_t* out, int lx, int ly);
void kor_cbbit_to_local(const uint32_t* in, uint32_t* out, int lx, int ly);
int kor_get_slope(const uint32_t* lat, int enc, int lx, int ly, int x, int y, int dir);

/* heights following Correlator::set; h has 2^(lx+ly) entries, row-major */
void kor_heights(const uint32_t* lat, int enc, int lx, int ly, int32_t* h);
/* exact sums; out = {sum h, sum h^2 (lo), sum h^2 (hi)} as int64 words */
void kor_roughness(const uint32_t* lat, int enc, int lx, int ly, int64_t out[3]);
/* out = {sum h_sys, sum h_snap, sum h_sys^2 lo, hi, sum h_sys*h_snap lo, hi, tcs} */
void kor_correlate(const uint32_t* sys, const uint32_t* snap, int enc, int lx, int ly, int64_t out[7]);
/* the same sums row by row on nthreads threads, without lattice-sized height
 * arrays (any lattice size) */
void kor_roughness_rows(const uint32_t* lat, int enc, int lx, int ly, int nthreads, int64_t out[3]);
void kor_correlate_rows(const uint32_t* sys, const uint32_t* snap, int enc, int lx, int ly, int nthreads, int64_t out[7]);
/* 1 if every row/column slope sum is zero and every plaquette closes */
int kor_consistent(const uint32_t* lat, int enc, int lx, int ly);

/* ---------------- schedulers ---------------- */
/* update kinds */
#define KOR_UPD_KPZ 0   /* pqOneUpdate: p=1, q=0, two draws per update */
#define KOR_UPD_PQ  1   /* pqSyncRngUpdate: one extra randomFloat per update */
#define KOR_UPD_DIS 2   /* disorder2SyncRngUpdate: as PQ with the rates of the site's class */
#define KOR_UPD_EW  3   /* SyncRngEW (GPU fast path, p = q = 1): every minimum and maximum
                           flips, no acceptance draw (contract: no CPU counterpart) */

/* rates per disorder class (m_disorderP/Q, kpz/schedulerService.h:117,
 * 185-197) and the disorder map (m_disorder, laid out like the lattice it
 * goes with; NULL = class 0 everywhere) */
typedef struct kor_rates {
	double p[4], q[4];
	const uint32_t* disorder;
} kor_rates;
/* SchedulerService::generateDisorder2, kpz/schedulerService.cpp:54-69:
 * per word 16 master draws, bit 2k set when draw k < fill */
void kor_generate_disorder2(kor_dsfmt* master, uint32_t* dis, uint64_t size_w, double fill);

/* SchedulerCPUDT (origin_shift=0) and SchedulerCPUDTr (origin_shift=1) */
int kor_dt_mcs(uint32_t* lat, int lx, int ly, int lbx, int lby,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int random_origin, int upd, double p, double q);
void kor_dt_block_offset(int lx, int ly, int lbx, int lby, int domain_set,
		const unsigned shift[2], int i, int off[2]);
/* same, with nthreads pthreads per phase (timing baseline) */
int kor_dt_mcs_threaded(uint32_t* lat, int lx, int ly, int lbx, int lby,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int random_origin, int upd, double p, double q, int nthreads);

/* SchedulerCPUDB (dead border, dynamic origin; pqOneUpdate only) with an
 * explicit block layout (lbx, lby >= 2) and origin restriction lorest;
 * T streams <= number of blocks; nthreads pthreads per MCS (1 = inline) */
int kor_db_mcs(uint32_t* lat, int lx, int ly, int lbx, int lby, int lorest,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int nthreads);

/* SchedulerCPUCBBit: mode 0 = pSyncRngUpdateVL3 (p=.5, q=0), 1 = per-site pq */
int kor_cbbit_mcs(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int mode, double p, double q);
/* one half-sweep with an explicit per-site mask source, for equivalence tests:
 * mask bits are taken from `masks` (one u64 per u64 word, row-major) */
void kor_cbbit_halfsweep_masked(uint32_t* lat, int lx, int ly, int signum, const uint64_t* masks);
/* SchedulerCPUCB (ENC_LOCAL checkerboard, per-site pq) */
int kor_cb_mcs(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, double p, double q);
/* the same two on nthreads pthreads (the sequential result: a half-sweep's
 * updates own disjoint bits; threads never write one word at once) */
int kor_cbbit_mcs_threaded(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int mode, double p, double q, int nthreads);
int kor_cb_mcs_threaded(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, double p, double q, int nthreads);

/* the schedulers the reference's frontend builds with TinyMT32 workers
 * (kpz/kpzFrontend.cpp:308-331): SchedulerCPUDTr (upd KPZ / PQ / EW with p,
 * q; DIS with r), SchedulerCPUCBBit (mode 0 VL3, 1 per-site pq, 2 disorder),
 * SchedulerCPUCB (upd PQ / DIS); randomFloat() has 24 bits */
int kor_dtr_mcs_tmt(uint32_t* lat, int lx, int ly, int lbx, int lby, kor_tmt* g, int64_t T,
		kor_dsfmt* master, int n, int upd, double p, double q, const kor_rates* r);
int kor_cbbit_mcs_tmt(uint32_t* lat, int lx, int ly, kor_tmt* g, int64_t T,
		int n, int mode, double p, double q, const kor_rates* r);
int kor_cb_mcs_tmt(uint32_t* lat, int lx, int ly, kor_tmt* g, int64_t T,
		int n, int upd, double p, double q, const kor_rates* r);

/* single updates (for brute-force checks) */
void kor_update_local(uint32_t* lat, int lx, int ly, int x, int y, int upd, float rnd, double p, double q);
void kor_update_cbbit(uint32_t* lat, int lx, int ly, int x, int y, float rnd, double p, double q);

/* the same schedulers with per-class rates and a disorder map
 * (mcsDisorder2SyncRng: kpz/schedulerCPUDTr.cpp:115-119,
 * kpz/schedulerCPUCBBit.cpp:105-109, kpz/schedulerCPUCB.cpp:103-108) */
int kor_dt_mcs_rates(uint32_t* lat, int lx, int ly, int lbx, int lby,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int random_origin, int upd, const kor_rates* r);
/* mode 1 = per-site pq (class 0 rates), 2 = disorder2SyncRngUpdate */
int kor_cbbit_mcs_rates(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int mode, const kor_rates* r);
int kor_cb_mcs_rates(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int upd, const kor_rates* r);
void kor_update_local_rates(uint32_t* lat, int lx, int ly, int x, int y, int upd, float rnd, const kor_rates* r);

/* plain random-sequential SchedulerCPU on ENC_LOCAL, all draws from the
 * master dSFMT (kpz/schedulerCPU.cpp:35-179, kpz/schedulerCPU.inl:21-90):
 *   KOR_RS_MCS        mcs: x, y per update, pqOneUpdate
 *   KOR_RS_PQ         mcsPQ: x, y; a third draw only at a minimum / maximum
 *                     (pqUpdate; p = q = 1: pqOneOneUpdate, no third draw)
 *   KOR_RS_SYNC       mcsSyncRng: RngPackage of 3 draws per update, pqOneUpdate
 *   KOR_RS_PQ_SYNC    mcsPQSyncRng: RngPackage, rnd[2] against p / q
 *   KOR_RS_DIS        mcsDisorder2: as KOR_RS_PQ with the rates of the site's class
 *   KOR_RS_DIS_SYNC   mcsDisorder2SyncRng: as KOR_RS_PQ_SYNC with class rates
 * rates are compared in double (no 23-bit mantissa here). */
#define KOR_RS_MCS 0
#define KOR_RS_PQ 1
#define KOR_RS_SYNC 2
#define KOR_RS_PQ_SYNC 3
#define KOR_RS_DIS 4
#define KOR_RS_DIS_SYNC 5
int kor_rs_mcs(uint32_t* lat, int lx, int ly, kor_dsfmt* master, int n, int mode, const kor_rates* r);
void kor_update_cbbit_rates(uint32_t* lat, int lx, int ly, int x, int y, float rnd, const kor_rates* r);

#ifdef __cplusplus
}
#endif
#endif
