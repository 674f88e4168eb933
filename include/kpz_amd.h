/*
 * kpz_amd.h -- C ABI of the MI355X-native KPZ octahedron-model engine.
 *
 * Drop-in boundary for the scheduler layer of CudaKpz: one opaque handle is
 * one Kpz::SchedulerService (kpz/schedulerService.h:73-255) with its device
 * lattice, per-worker SLCG64 streams and master dSFMT.  Plain pointers and
 * sizes only; no exceptions cross the ABI: every call returns KPZ_OK (0) or
 * a negative KPZ_E* code and kpz_last_error() describes the failure
 * (thread-local).  One handle must be used by one host thread at a time.
 *
 * Lattices passed through kpz_upload/kpz_download use the reference layout
 * bit for bit: ENC_LOCAL (4x4 sites per u32) for DT/DTr, ENC_CBBIT (four
 * sublattice bit-plane quadrants) for SCABit, 2^(lx+ly-4) u32 words.
 */
#ifndef KPZ_AMD_H
#define KPZ_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kpz_scheduler kpz_scheduler;

/* scheduler kinds; the CPU counterpart each one is bit-exact against */
enum {
	KPZ_DT = 0,     /* SchedulerCPUDT<Rng> (static DT)  kpz/schedulerCPUDT.cpp:65-90 */
	KPZ_DTR = 1,    /* SchedulerCPUDTr<Rng> (random-origin DT) kpz/schedulerCPUDTr.cpp:66-98 */
	KPZ_SCABIT = 2, /* SchedulerCPUCBBit<Rng> (checkerboard, bit vectorised) kpz/schedulerCPUCBBit.cpp:43-108 */
	KPZ_SCA = 3     /* SchedulerCPUCB<Rng> (checkerboard on ENC_LOCAL; GPU SchedulerSCA) kpz/schedulerCPUCB.cpp:43-170 */
};

enum { KPZ_ENC_LOCAL = 0, KPZ_ENC_CBBIT = 1 };

enum {
	KPZ_OK = 0,
	KPZ_EINVAL = -1,  /* bad argument or unsupported configuration */
	KPZ_EDEVICE = -2, /* HIP runtime error */
	KPZ_ENOMEM = -3,  /* device allocation failed */
	KPZ_ESTATE = -4   /* call out of order (e.g. correlate before corr_tag) */
};

typedef struct kpz_params {
	int64_t threads;   /* RNG streams T = worker threads of the CPU counterpart
	                      (SchedulerCPUpThreadRng(lx, ly, threads)); 0 = one per
	                      DT block / one per SCABit row */
	int layout_x;      /* log2 DT block width  (SchedulerCPUDT::setLayout), 2..4; 0 = 4 */
	int layout_y;      /* log2 DT block height, 2..4; 0 = 4 */
	uint64_t skip;     /* SLCG skip length S (kpz/kpzConst.h:85); 0 = 30721 if T < 30721
	                      else the smallest odd S > T */
	int device;        /* HIP device ordinal; -1 = current device */
	int rng;           /* worker stream generator: KPZ_RNG_SLCG64 (0) or KPZ_RNG_TINYMT32,
	                      the Rng the reference's frontend gives SchedulerCPUDTr, CB and
	                      CBBit (kpz/kpzFrontend.cpp:308-331); TinyMT32 for KPZ_DTR,
	                      KPZ_SCA, KPZ_SCABIT */
} kpz_params;

enum {
	KPZ_RNG_SLCG64 = 0,   /* SLCG64 in the skip form, inc/kmcRandom.h:59-124 */
	KPZ_RNG_TINYMT32 = 1  /* TinyMT32, inc/tinyMT/tinymt32.h; RngIface::TinyMT32 kpz/CPURngIface.h:71-109 */
};

/* thread-local description of the last failure */
const char* kpz_last_error(void);

/* SchedulerConfig::mkScheduler(x, y) + the scheduler constructor
 * (kpz/kpzFrontend.cpp:234-250): lx, ly are log2 lattice sizes */
int kpz_create(int kind, int lx, int ly, const kpz_params* params, kpz_scheduler** out);
int kpz_destroy(kpz_scheduler* h);
/* the KPZ_* kind the scheduler was created as (>= 0), or a negative error */
int kpz_kind(const kpz_scheduler* h);
/* log2 lattice sizes (SchedulerService::size(), kpz/systemSize.h) */
int kpz_size(const kpz_scheduler* h, int* lx, int* ly);

/* Checkpoint / resume, the role of SchedulerService::collectH5 / retrieveH5
 * (kpz/schedulerService.cpp:583-676) and of --saveH5 / --loadH5
 * (kpz/kpzScaling.cpp:118-139, 217-222).  The file holds kind, sizes,
 * parameters, MCS count, (p, q), the master dSFMT, every SLCG stream state,
 * the lattice and all live correlator tags, ending with an FNV-1a 64
 * checksum; a loaded scheduler continues bit for bit.  aux: 32 bits of
 * caller data kept in the file (kpz_scaling: the last measurement time, the
 * reference's "prevMeasureOverride" attribute).  kpz_load_state validates
 * the whole file before it touches a device. */
int kpz_save_state(kpz_scheduler* h, const char* path, uint32_t aux);
int kpz_load_state(const char* path, int device, kpz_scheduler** out, uint32_t* aux);

/* SchedulerService::supplySeed + randomize (kpz/schedulerService.cpp:460-482,
 * kpz/GPURandom.cu:42-54): master dSFMT reSeed(seed ^ seed>>32), then the T
 * stream states from two master draws */
int kpz_set_seed(kpz_scheduler* h, uint64_t seed);
/* randomize(): re-seed the T streams from two draws of the engine's master
 * (SLCG64::randomize, kpz/GPURandom.cu:42-54; SchedulerDT::randomize
 * kpz/schedulerDT.h:98) */
int kpz_randomize(kpz_scheduler* h);
/* SchedulerService::setPQ (kpz/schedulerService.h:190-197) */
int kpz_set_pq(kpz_scheduler* h, double p, double q);
/* initHomogeneous (kpz/schedulerService.cpp:156-175), on the device */
int kpz_init_flat(kpz_scheduler* h);
/* pushSystem / popSystem with a caller-owned host lattice (kpz/schedulerBase.cu:35-63) */
int kpz_upload(kpz_scheduler* h, const uint32_t* lattice);
int kpz_download(kpz_scheduler* h, uint32_t* lattice);
/* mcs(n): the scheduler's default dynamics (DT: p=1,q=0; SCABit: p=.5,q=0) */
int kpz_mcs(kpz_scheduler* h, int n);
/* Disorder2 (kpz/schedulerService.cpp:48-69, kpz/schedulerService.h:185-197):
 * kpz_set_pq_class = setPQ(p, q, n) for class n = 0..3 (class 0 is
 * kpz_set_pq; a class never set uses class 0's rates); the map has the
 * lattice's size and layout and holds the class of every site (ENC_LOCAL:
 * the site's 2 bits; ENC_CBBIT: its x-slope bit); kpz_generate_disorder2
 * draws it from the master dSFMT exactly as generateDisorder2(fill).
 * kpz_mcs_disorder2 = mcsDisorder2SyncRng (kpz/schedulerCPUDTr.cpp:115-119,
 * kpz/schedulerCPUCBBit.cpp:105-109, kpz/schedulerCPUCB.cpp:103-108). */
int kpz_set_pq_class(kpz_scheduler* h, double p, double q, unsigned n);
/* rateP(n) / rateQ(n) (kpz/schedulerService.h:193-194): the rates class n
 * uses (p, q may each be NULL) */
int kpz_get_pq_class(const kpz_scheduler* h, unsigned n, double* p, double* q);
int kpz_generate_disorder2(kpz_scheduler* h, double fill);
int kpz_set_disorder(kpz_scheduler* h, const uint32_t* map);
int kpz_get_disorder(kpz_scheduler* h, uint32_t* map);
int kpz_mcs_disorder2(kpz_scheduler* h, int n);
/* EW fast path of the GPU DT scheduler (SyncRngEW,
 * kpz/kpzSimulationPrimitivesTemplates.cu:199-234, selected at kpz/kpz.cu:221-228
 * for p = q = 1): every minimum and maximum flips and no acceptance number is
 * drawn (two draws per update); KPZ_DT / KPZ_DTR only.  kpz_mcs_pq with
 * p = q = 1 keeps the CPU counterparts' acceptance draw. */
int kpz_mcs_ew(kpz_scheduler* h, int n);
/* mcsPQSyncRng(n): dynamics with the current p, q */
int kpz_mcs_pq(kpz_scheduler* h, int n);
int kpz_sync(kpz_scheduler* h);
/* SchedulerService::copySystem (kpz/schedulerService.cpp:497-504): device-side
 * copy of src's lattice into dst (same size and encoding), queued on dst's stream */
int kpz_copy_system(kpz_scheduler* dst, const kpz_scheduler* src);

/* exact height sums over all sites, heights as Correlator::set
 * (kpz/correlator.cpp:37-106): out = {sum h, sum h^2 low 64, sum h^2 high 64} */
int kpz_roughness(kpz_scheduler* h, int64_t out[3]);
/* No counterpart in the reference: the variant of kpz_roughness for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same sums, of a lattice of
 * 2^(lx+ly-4) u32 words in `encoding` (KPZ_ENC_*) in the caller's device
 * memory on the current device; writes {sum h, sum h^2 low 64, high 64} to
 * `out` (device memory, 3 int64).  Only queues work on `stream` (a
 * hipStream_t; NULL = the null stream) and returns: no allocation, no
 * synchronisation.  Intermediate sums use a fixed per-device workspace of
 * the library (a static device array of about 5.3 MB, held by each GPU the
 * library runs on until the process ends), resolved on the first call on a
 * device (make that first call outside stream capture), so at most 2^16 site
 * rows (ly <= 16; larger ly is rejected with KPZ_EINVAL).  Calls that overlap
 * on one device share the workspace and race: serialise them (one stream,
 * or events). */
int kpz_roughness_noalloc(const uint32_t* lattice, int lx, int ly, int encoding, int64_t* out, void* stream);
/* height moments <h^k>, k = 1..7 (Roughness::h[] after normAndSetW2,
 * kpz/roughness.cpp:38-43, kpz/schedulerService.h:52-71): per-row double
 * sums on the device, rows added in long double; the 7 values the
 * reference prints after w^2 (operator<< of Roughness,
 * kpz/schedulerService.cpp:73-79) */
int kpz_moments(kpz_scheduler* h, double out[7]);
/* correlateTag: device snapshot of the current lattice, kept while the MCS
 * count is <= stop (kpz/schedulerService.cpp:266-271); kpz_corr_tag keeps it
 * for ever.  Several tags may be live. */
int kpz_corr_tag(kpz_scheduler* h);
int kpz_corr_tag_stop(kpz_scheduler* h, uint32_t stop);
/* correlate against the most recent tag (kpz/correlator.cpp:127-164, 450-455):
 * out = {sum h, sum h_snap, sum h^2 lo, hi, sum h*h_snap lo, hi, slope overlap} */
int kpz_correlate(kpz_scheduler* h, int64_t out[7]);
/* every live tag, oldest first, after dropping expired ones
 * (kpz/schedulerService.cpp:216-251): 8 int64 per tag = {tag MCS, then the 7
 * values of kpz_correlate}; returns the number of tags written (<= max_tags)
 * or a negative error */
int kpz_correlate_tags(kpz_scheduler* h, int64_t* out, int max_tags);
/* SchedulerService::consitencyCheck(full, out) (kpz/schedulerService.cpp:721-854,
 * declared at kpz/schedulerService.h:148), on the device: out = {rows whose x
 * slopes do not sum to zero, columns whose y slopes do not, loops
 * sX(x,y) + sY(x-1,y) != sY(x,y) + sX(x,y-1)}; full != 0 adds the counts of
 * every live correlator tag's snapshot.  Returns 1 if all three are zero, 0
 * if not, or a negative error. */
int kpz_consistency_check(kpz_scheduler* h, int full, int64_t out[3]);

/* RNG state: stream states (T u64) and the master dSFMT.  The master state
 * is kpz_master_bytes() bytes laid out as the reference's dsfmt_t
 * (inc/dSFMT/dSFMT.h: 192 x 128-bit status words, then int idx). */
int64_t kpz_threads(const kpz_scheduler* h);
int kpz_get_streams(kpz_scheduler* h, uint64_t* states);
int kpz_set_streams(kpz_scheduler* h, const uint64_t* states);
/* the worker stream generator (KPZ_RNG_*) */
int kpz_rng(const kpz_scheduler* h);
/* TinyMT32 streams.  Parameter set t of stream t, as RngIface::TinyMT32::setMat
 * (kpz/CPURngIface.cpp:37-42) and TinyMT32::initCUDA (kpz/GPUTinyMT.cu:38-61)
 * take them from tinymt32_param.h (generated by inc/tinyMT/genParam.sh; the
 * table is not in the reference tree): T words each; required before
 * kpz_set_seed / kpz_randomize.  State: 7 planes of T words, status[0..3],
 * mat1, mat2, tmat (the layout of TinyMT32::d_Random and of its `.tinyMT`
 * dataset, kpz/GPUTinyMT.cu:113-144). */
int kpz_set_tinymt_params(kpz_scheduler* h, const uint32_t* mat1, const uint32_t* mat2, const uint32_t* tmat);
/* Host utility, no counterpart in the reference: jump one TinyMT32 generator
 * (gen = status[0..3], mat1, mat2, tmat) n steps ahead with the polynomial of
 * its orbit, as the engine's TinyMT32 kernels do; writes the 4 status words
 * and returns the orbit polynomial's degree (<= 128).  Needs no device. */
int kpz_tinymt_jump(const uint32_t gen[7], uint64_t n, uint32_t status[4]);
int kpz_get_tinymt(kpz_scheduler* h, uint32_t* planes);
int kpz_set_tinymt(kpz_scheduler* h, const uint32_t* planes);
int kpz_master_bytes(void);
int kpz_get_master(kpz_scheduler* h, uint8_t* state);
int kpz_set_master(kpz_scheduler* h, const uint8_t* state);
/* next master draw, as SchedulerService::genrand_close_open (schedulerService.h:222) */
double kpz_master_draw(kpz_scheduler* h);

/* counters and geometry */
uint32_t kpz_mcs_count(const kpz_scheduler* h);
double kpz_mcs_scale(const kpz_scheduler* h);     /* SchedulerSCABit::mcsScale = p + q */
int kpz_encoding(const kpz_scheduler* h);
uint64_t kpz_size_words(const kpz_scheduler* h);
uint64_t kpz_skip(const kpz_scheduler* h);
/* device pointer of the lattice and the HIP stream all work is queued on */
void* kpz_device_lattice(const kpz_scheduler* h);
void* kpz_stream(const kpz_scheduler* h);

/* Measurement utility, no counterpart in the reference: the integer VALU
 * issue rates of the device, the roofline that bounds the DT and per-site
 * SCA kernels, in lane operations per second (best of `reps` launches of 8
 * independent chains per lane on 32 x 256-lane blocks per CU).  slow_class:
 * v_xad_u32, the rate of left shifts, multiplies, add3, bfe, perm and
 * alignbit on gfx950; fast_class (may be NULL): v_xor_b32, the rate of plain
 * xor, add, bitop3 and right shifts, about twice as high. */
int kpz_valu_peak(int device, int reps, double* slow_class, double* fast_class);
/* Measurement utility, no counterpart in the reference: the issue rate of an
 * instruction mix, lane operations per second, of blocks of 8 independent
 * VALU instructions of which slow_of_8 (0, 2, 3, 4, 5 or 8) are v_mul_lo_u32
 * (slow class) and the rest v_xor_b32 (fast class), 8 waves per SIMD.  On
 * gfx950 a fast-class instruction next to slow-class ones issues at about the
 * slow rate, so this, at a kernel's own slow fraction, is the VALU roofline
 * of that kernel. */
int kpz_valu_mix_rate(int device, int reps, int slow_of_8, double* lane_ops_per_s);

/* ---------------------------------------------------------------------
 * Batched replicas (no counterpart in the reference: CudaKpz runs one
 * lattice per process).  Each replica is bit-identical to a stand-alone
 * scheduler of the same kind, size, seed and (p, q) with default threads
 * and 16x16 DT blocks: DT/DTr with (p, q) = (1, 0) run kpz_mcs, every other
 * replica kpz_mcs_pq.  All replicas advance together; chunks of MCS are
 * replayed from captured hipGraphs.
 * ------------------------------------------------------------------- */
typedef struct kpz_batch kpz_batch;

typedef struct kpz_replica {
	int kind;          /* KPZ_DT, KPZ_DTR or KPZ_SCABIT */
	int lx, ly;        /* log2 sizes */
	double p, q;
	uint64_t seed;     /* as kpz_set_seed */
} kpz_replica;

/* performance knobs; no knob changes any result */
typedef struct kpz_batch_tuning {
	int max_lanes_log2; /* SCABit lanes per row = min(words per row, 2^this), 0..6 */
	int pq_words64;     /* SCABit per-site replicas on 64-bit (1) or 32-bit (0) words */
	int chunk;          /* MCS per captured graph */
	int concurrent;     /* 0: one stream; 1: DT and SCABit work on two streams; 2: also the
	                       DT replicas in two groups of equal lane count on two streams */
	int sort_by_size;   /* large replicas first */
} kpz_batch_tuning;

/* tuning = NULL: defaults {3, 1, 50, 2, 1} */
int kpz_batch_create(const kpz_replica* reps, int n, const kpz_batch_tuning* tuning, int device, kpz_batch** out);
int kpz_batch_destroy(kpz_batch* b);
int kpz_batch_init_flat(kpz_batch* b);
int kpz_batch_mcs(kpz_batch* b, int n);
/* out: 3 int64 per replica = {sum h, sum h^2 low 64, high 64} */
int kpz_batch_roughness(kpz_batch* b, int64_t* out);
int kpz_batch_download(kpz_batch* b, int replica, uint32_t* lattice);
int kpz_batch_get_streams(kpz_batch* b, int replica, uint64_t* states);
int64_t kpz_batch_threads(const kpz_batch* b, int replica);
uint64_t kpz_batch_sites(const kpz_batch* b);
int kpz_batch_sync(kpz_batch* b);
void* kpz_batch_stream(const kpz_batch* b);
/* autotuner: times `mcs` MCS per candidate on throwaway batches of the same
 * shapes; *tuning is the start point and receives the best knobs */
int kpz_batch_autotune(const kpz_replica* reps, int n, int mcs, int device, kpz_batch_tuning* tuning, double* ms_per_mcs);

/* ---------------------------------------------------------------------
 * Serving loop (no counterpart in the reference).  Jobs are admitted into
 * slots as HBM frees up (FIFO with backfill by footprint) and advance
 * together in batched launches; lattices and RNG state live in one device
 * arena (buddy allocator).  A job is bit-identical to the stand-alone
 * scheduler with its kind, size, seed and (p, q) (DT/DTr (1,0): kpz_mcs,
 * else kpz_mcs_pq) and reports exact sums at the log-spaced times
 * ceil(exp(log(t+10)+0.1)) (kpz/sampling.h:84-100) and at its last MCS.
 * ------------------------------------------------------------------- */
typedef struct kpz_server kpz_server;

typedef struct kpz_job {
	int kind;          /* KPZ_DT, KPZ_DTR, KPZ_SCABIT */
	int lx, ly;
	int mcs;           /* MCS to run */
	double p, q;
	uint64_t seed;
	int keep_lattice;  /* return the final lattice with the result */
} kpz_job;

int kpz_server_create(uint64_t hbm_budget_bytes, int max_slots, int max_tick_mcs, int device, kpz_server** out);
int kpz_server_destroy(kpz_server* s);
int kpz_server_submit(kpz_server* s, const kpz_job* job, uint64_t* id);
/* one tick: queues the next tick's work, then waits for the tick queued by
 * the previous call and finishes it (samples, retirement); returns queued +
 * active jobs (>= 0) or a negative error code */
int64_t kpz_server_step(kpz_server* s);
/* 1 if a finished job is waiting: its sample count and lattice words (0 if not kept) */
int kpz_server_peek(kpz_server* s, int* nsamples, uint64_t* lattice_words);
/* take the finished job: times = {submit, start, end} in s since server
 * start; per sample mcs[i] and sums[3i..3i+2] = {sum h, sum h^2 lo, hi};
 * lattice may be NULL */
int kpz_server_pop(kpz_server* s, uint64_t* id, double times[3], uint32_t* mcs, int64_t* sums, uint32_t* lattice);
/* {queued, active, hbm bytes in use, hbm arena bytes, site updates so far} */
int kpz_server_stats(kpz_server* s, uint64_t out[5]);
/* the server's clock (s since creation; monotonic) */
double kpz_server_now(const kpz_server* s);
/* device time of the server: gpu_seconds = time the device spent in ticks
 * (hipEvents around each tick's work on the server's main stream; a tick's
 * second launch chain forks from and joins into it), span_seconds =
 * from the first tick's start to the end of the last finished tick; their
 * ratio is the fraction of time the device was busy.  Ticks are pipelined:
 * kpz_server_step queues the next tick before it waits for the previous one. */
int kpz_server_busy(const kpz_server* s, double* gpu_seconds, double* span_seconds);

#ifdef __cplusplus
}
#endif

#endif
