/*
 * ORACLE -- test infrastructure only (see kpz_oracle.h for the pinning
 * status of every function).  Never linked into the product.
 *
 * Plain C restatement of the reference's CPU hot path.  Reference paths are
 * relative to the CudaKpz source tree.
 */
#include "kpz_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * dSFMT-19937, restated from the published algorithm (Saito & Matsumoto)
 * with the parameters of inc/dSFMT/dSFMT-params19937.h:6-16 and the
 * generic recursion of inc/dSFMT/dSFMT.cpp:153-167 (non-SIMD variant),
 * gen_rand_all :506-521, init_gen_rand :622-639, initial_mask :428-436,
 * period_certification :442-476, genrand_close_open dSFMT.h:230-265.
 * ====================================================================== */
#define DS_N 191
#define DS_N64 (DS_N * 2)
#define DS_POS1 117
#define DS_SL1 19
#define DS_SR 12
#define DS_MSK1 0x000ffafffffffb3full
#define DS_MSK2 0x000ffdfffc90fffdull
#define DS_FIX1 0x90014964b32f4329ull
#define DS_FIX2 0x3b8d12ac548a7c7aull
#define DS_PCV1 0x3d84e1ac0dc82880ull
#define DS_PCV2 0x0000000000000001ull
#define DS_LOW_MASK 0x000FFFFFFFFFFFFFull
#define DS_HIGH_CONST 0x3FF0000000000000ull

int kor_dsfmt_size(void) { return (int)sizeof(kor_dsfmt); }

static void ds_rec(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t* lung)
{
	const uint64_t t0 = a[0], t1 = a[1], L0 = lung[0], L1 = lung[1];
	lung[0] = (t0 << DS_SL1) ^ (L1 >> 32) ^ (L1 << 32) ^ b[0];
	lung[1] = (t1 << DS_SL1) ^ (L0 >> 32) ^ (L0 << 32) ^ b[1];
	r[0] = (lung[0] >> DS_SR) ^ (lung[0] & DS_MSK1) ^ t0;
	r[1] = (lung[1] >> DS_SR) ^ (lung[1] & DS_MSK2) ^ t1;
}

static void ds_gen_all(kor_dsfmt* d)
{
	uint64_t* s = d->st.u64;
	uint64_t lung[2] = {s[2 * DS_N], s[2 * DS_N + 1]};
	int i;
	ds_rec(&s[0], &s[0], &s[2 * DS_POS1], lung);
	for(i = 1; i < DS_N - DS_POS1; ++i)
		ds_rec(&s[2 * i], &s[2 * i], &s[2 * (i + DS_POS1)], lung);
	for(; i < DS_N; ++i)
		ds_rec(&s[2 * i], &s[2 * i], &s[2 * (i + DS_POS1 - DS_N)], lung);
	s[2 * DS_N] = lung[0];
	s[2 * DS_N + 1] = lung[1];
}

void kor_dsfmt_init(kor_dsfmt* d, uint32_t seed)
{
	uint32_t* p = d->st.u32;
	p[0] = seed;
	for(int i = 1; i < (DS_N + 1) * 4; ++i)
		p[i] = 1812433253u * (p[i - 1] ^ (p[i - 1] >> 30)) + (uint32_t)i;
	for(int i = 0; i < DS_N * 2; ++i)
		d->st.u64[i] = (d->st.u64[i] & DS_LOW_MASK) | DS_HIGH_CONST;
	/* period certification */
	uint64_t t0 = d->st.u64[2 * DS_N] ^ DS_FIX1;
	uint64_t t1 = d->st.u64[2 * DS_N + 1] ^ DS_FIX2;
	uint64_t inner = (t0 & DS_PCV1) ^ (t1 & DS_PCV2);
	for(int i = 32; i > 0; i >>= 1)
		inner ^= inner >> i;
	if(!(inner & 1))
		d->st.u64[2 * DS_N + 1] ^= 1; /* PCV2 has bit 0 set */
	d->idx = DS_N64;
}

double kor_dsfmt_close_open(kor_dsfmt* d)
{
	if(d->idx >= DS_N64)
	{
		ds_gen_all(d);
		d->idx = 0;
	}
	double r;
	memcpy(&r, &d->st.u64[d->idx++], sizeof r);
	return r - 1.0;
}

/* kpz/schedulerService.cpp:460-470 */
void kor_reseed(kor_dsfmt* d, uint32_t seed)
{
	kor_dsfmt_init(d, seed);
	for(int a = 0; a < 100000; ++a)
		kor_dsfmt_close_open(d);
}

uint32_t kor_seed32(uint64_t seed) { return (uint32_t)(seed ^ (seed >> 32)); }

/* Kmc::Shuffle<4>::shuffle(dsfmt_t*), inc/shuffle.h:76-88: N=4, L=2, one
 * 32-bit draw covers all four RC4-key-schedule swaps. */
void kor_shuffle4(kor_dsfmt* d, int seq[4])
{
	for(int a = 0; a < 4; ++a)
		seq[a] = a;
	unsigned int rnd = (unsigned int)(kor_dsfmt_close_open(d) * 4294967296.0);
	for(int a = 0; a < 4; ++a)
	{
		const int i = rnd & 3;
		const int v = seq[i];
		seq[i] = seq[a];
		seq[a] = v;
		rnd >>= 2;
	}
}

/* ======================================================================
 * SLCG64, inc/kmcRandom.h:59-60 (A, C), :74-117 (skip form), kpz/kpzRandom.cu:
 * 62-68 (output fold, randomFloat), kpz/GPURandom.cu:42-54 (stream seeding).
 * ====================================================================== */

/* skip form of the reference: A_S = A^S, C_S = C * sum_{k=1..S} A^k */
void kor_slcg_skip(uint64_t S, uint64_t* a, uint64_t* c)
{
	uint64_t ak = 1, cs = 0;
	for(uint64_t k = 0; k < S; ++k)
	{
		ak *= KOR_SLCG_A;
		cs += ak * KOR_SLCG_C;
	}
	*a = ak;
	*c = cs;
}

/* k-fold composition of x -> a*x + c, by squaring */
void kor_slcg_jump(uint64_t a, uint64_t c, uint64_t k, uint64_t* ak, uint64_t* ck)
{
	uint64_t ra = 1, rc = 0; /* identity */
	uint64_t ba = a, bc = c;
	while(k)
	{
		if(k & 1)
		{
			rc = ba * rc + bc;
			ra = ba * ra;
		}
		bc = ba * bc + bc;
		ba = ba * ba;
		k >>= 1;
	}
	*ak = ra;
	*ck = rc;
}

/* (contract) skip length: the reference constant 30721 (kpz/kpzConst.h:85)
 * whenever T < 30721 (kpz/GPURandom.cu:44 asserts exactly that), otherwise the
 * smallest odd S > T, so the multiplier stays = 5 mod 8 and C_S stays odd. */
uint64_t kor_default_skip(uint64_t T)
{
	if(T < 30721)
		return 30721;
	return T + 1 + (T & 1);
}

/* SLCG64::randomize, kpz/GPURandom.cu:42-54; (contract) the first master
 * draw is the low half. */
void kor_streams_randomize(kor_dsfmt* master, uint64_t* states, int64_t T)
{
	const double m1 = kor_dsfmt_close_open(master);
	const double m2 = kor_dsfmt_close_open(master);
	uint64_t s = (uint64_t)(m1 * 4294967295.0) ^ ((uint64_t)(m2 * 4294967295.0) << 32);
	for(int64_t t = 0; t < T; ++t)
	{
		states[t] = s;
		s = s * KOR_SLCG_A + KOR_SLCG_C;
	}
}

uint32_t kor_slcg_draw(uint64_t* s, uint64_t a, uint64_t c)
{
	*s = a * *s + c;
	return (uint32_t)(*s ^ (*s >> 32));
}

/* Statistics of n folded SLCG64 draws, as the reference's profiling program
 * inc/slcgProfile.cpp:41-60 prints them: the base generator (A, C) is
 * stepped n times from `state`; of every draw, bits[b] counts bit b of the
 * fold random() = lo ^ hi, j[] is the 12-bucket histogram of
 * (random() >> 24) / 256 * 12 and j2[] the one of randomFloat() * 12, with
 * randomFloat() = (random() & 0x7FFFFF) / 2^23 (kpz/kpzRandom.cu:62-68; float
 * arithmetic throughout, as there).  random() and randomFloat() are this
 * file's own stream cursor functions (gen_random / gen_rf) on two copies of
 * the stream, so the counts check exactly what the schedulers draw.  The n
 * draws are split over nthreads chunks, each started by kor_slcg_jump. */
static uint32_t gen_random_slcg(uint64_t* s, uint64_t a, uint64_t c);
static uint32_t gen_rf_slcg(uint64_t* s, uint64_t a, uint64_t c);

typedef struct {
	uint64_t state, n;
	uint64_t byte[4][256];
	uint64_t j[12], j2[12];
} prof_work;

static void* prof_run(void* arg)
{
	prof_work* W = (prof_work*)arg;
	uint64_t sa = W->state, sb = W->state;
	for(uint64_t i = 0; i < W->n; ++i)
	{
		const uint32_t t = gen_random_slcg(&sa, KOR_SLCG_A, KOR_SLCG_C);
		const uint32_t r = gen_rf_slcg(&sb, KOR_SLCG_A, KOR_SLCG_C);
		++W->byte[0][t & 0xFF];
		++W->byte[1][(t >> 8) & 0xFF];
		++W->byte[2][(t >> 16) & 0xFF];
		++W->byte[3][t >> 24];
		const float f = (float)(t >> 24) / 256.0f * 12.0f;
		++W->j[(int)f];
		const float f2 = (float)r / 8388608.0f * 12.0f;
		++W->j2[(int)f2];
	}
	return NULL;
}

#include <pthread.h>

void kor_slcg_profile(uint64_t state, uint64_t n, int nthreads, uint64_t bits[32], uint64_t j[12], uint64_t j2[12])
{
	if(nthreads < 1)
		nthreads = 1;
	prof_work* wk = (prof_work*)calloc((size_t)nthreads, sizeof(prof_work));
	pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
	uint64_t start = 0;
	for(int w = 0; w < nthreads; ++w)
	{
		const uint64_t cnt = n / (uint64_t)nthreads + ((uint64_t)w < n % (uint64_t)nthreads ? 1 : 0);
		uint64_t ja, jc;
		kor_slcg_jump(KOR_SLCG_A, KOR_SLCG_C, start, &ja, &jc);
		wk[w].state = ja * state + jc;
		wk[w].n = cnt;
		start += cnt;
		pthread_create(&th[w], NULL, prof_run, &wk[w]);
	}
	memset(bits, 0, 32 * sizeof(uint64_t));
	memset(j, 0, 12 * sizeof(uint64_t));
	memset(j2, 0, 12 * sizeof(uint64_t));
	for(int w = 0; w < nthreads; ++w)
	{
		pthread_join(th[w], NULL);
		for(int k = 0; k < 4; ++k)
			for(int v = 0; v < 256; ++v)
				for(int b = 0; b < 8; ++b)
					if(v >> b & 1)
						bits[8 * k + b] += wk[w].byte[k][v];
		for(int k = 0; k < 12; ++k)
		{
			j[k] += wk[w].j[k];
			j2[k] += wk[w].j2[k];
		}
	}
	free(th);
	free(wk);
}

/* ======================================================================
 * TinyMT32 (Saito & Matsumoto 2011, the 127-bit F2-linear generator vendored
 * at inc/tinyMT/tinymt32.h:70-107 (state transition, tempering) and
 * inc/tinyMT/tinymt32.c:59-75 (seeding)); RngIface::TinyMT32,
 * kpz/CPURngIface.h:71-109: random() = the tempered word, randomFloat() =
 * (tempered >> 8) * 2^-24, generator t takes parameter set t and the seed
 * (unsigned)(close_open() * UINT_MAX) from the master, in stream order.
 * ====================================================================== */
static void tmt_next(kor_tmt* g)
{
	uint32_t* st = g->status;
	uint32_t y = st[3];
	uint32_t x = (st[0] & 0x7FFFFFFFu) ^ st[1] ^ st[2];
	x ^= x << 1;
	y ^= (y >> 1) ^ x;
	const uint32_t m = (uint32_t)0 - (y & 1u);
	st[0] = st[1];
	st[1] = st[2] ^ (m & g->mat1);
	st[2] = x ^ (y << 10) ^ (m & g->mat2);
	st[3] = y;
}

static uint32_t tmt_temper(const kor_tmt* g)
{
	const uint32_t t1 = g->status[0] + (g->status[2] >> 8);
	return (g->status[3] ^ t1) ^ (((uint32_t)0 - (t1 & 1u)) & g->tmat);
}

void kor_tmt_init(kor_tmt* g, uint32_t seed)
{
	uint32_t* st = g->status;
	st[0] = seed;
	st[1] = g->mat1;
	st[2] = g->mat2;
	st[3] = g->tmat;
	for(int i = 1; i < 8; ++i)
	{
		const uint32_t p = st[(i - 1) & 3];
		st[i & 3] ^= (uint32_t)i + 1812433253u * (p ^ (p >> 30));
	}
	/* period certification: the all-zero state (bit 31 of word 0 aside) */
	if((st[0] & 0x7FFFFFFFu) == 0 && st[1] == 0 && st[2] == 0 && st[3] == 0)
	{
		st[0] = 'T';
		st[1] = 'I';
		st[2] = 'N';
		st[3] = 'Y';
	}
	for(int i = 0; i < 8; ++i)
		tmt_next(g);
}

uint32_t kor_tmt_u32(kor_tmt* g)
{
	tmt_next(g);
	return tmt_temper(g);
}

float kor_tmt_float(kor_tmt* g) { return (float)(kor_tmt_u32(g) >> 8) * (1.0f / 16777216.0f); }

void kor_tmt_randomize(kor_dsfmt* master, kor_tmt* g, int64_t T)
{
	for(int64_t t = 0; t < T; ++t)
		kor_tmt_init(&g[t], (uint32_t)(kor_dsfmt_close_open(master) * 4294967295.0));
}

/* ----------------------------------------------------------------------
 * A scheduler's T worker streams, of either generator, and one stream's
 * cursor.  random() is the 32-bit output; randomFloat() is rf * 2^-bits with
 * rf = the SLCG fold's low 23 bits (kpz/kpzRandom.cu:67-68) or the TinyMT
 * word's top 24 bits. */
typedef struct {
	int tmt;
	uint64_t* slcg;
	uint64_t a, c;
	kor_tmt* g;
} streams_t;

typedef struct {
	int tmt;
	uint64_t s, a, c;
	kor_tmt g;
} gen_t;

static inline gen_t gen_get(const streams_t* S, int64_t t)
{
	gen_t g;
	g.tmt = S->tmt;
	g.a = S->a;
	g.c = S->c;
	if(S->tmt)
	{
		g.g = S->g[t];
		g.s = 0;
	}
	else
	{
		memset(&g.g, 0, sizeof g.g);
		g.s = S->slcg[t];
	}
	return g;
}

static inline void gen_put(const streams_t* S, int64_t t, const gen_t* g)
{
	if(S->tmt)
		S->g[t] = g->g;
	else
		S->slcg[t] = g->s;
}

static inline uint32_t gen_random(gen_t* g) { return g->tmt ? kor_tmt_u32(&g->g) : kor_slcg_draw(&g->s, g->a, g->c); }
static inline int gen_bits(int tmt) { return tmt ? 24 : 23; }
static inline uint32_t gen_rf(gen_t* g) { return g->tmt ? kor_tmt_u32(&g->g) >> 8 : kor_slcg_draw(&g->s, g->a, g->c) & 0x7FFFFFu; }

/* the SLCG64 cursor functions above on a bare state (kor_slcg_profile) */
static uint32_t gen_random_slcg(uint64_t* s, uint64_t a, uint64_t c)
{
	gen_t g;
	g.tmt = 0;
	g.s = *s;
	g.a = a;
	g.c = c;
	const uint32_t v = gen_random(&g);
	*s = g.s;
	return v;
}

static uint32_t gen_rf_slcg(uint64_t* s, uint64_t a, uint64_t c)
{
	gen_t g;
	g.tmt = 0;
	g.s = *s;
	g.a = a;
	g.c = c;
	const uint32_t v = gen_rf(&g);
	*s = g.s;
	return v;
}

static streams_t streams_slcg(uint64_t* states, uint64_t a, uint64_t c)
{
	streams_t S = {0, states, a, c, NULL};
	return S;
}

static streams_t streams_tmt(kor_tmt* g)
{
	streams_t S = {1, NULL, 0, 0, g};
	return S;
}

/* ======================================================================
 * lattice encodings, kpz/systemSize.h:94-165, kpz/lattice.h:29-96
 * ====================================================================== */
static inline uint32_t loc_index(int lx, int x, int y) { return ((uint32_t)(y >> 2) << (lx - 2)) + (uint32_t)(x >> 2); }
static inline int loc_shift(int x, int y) { return ((x & 3) << 1) | ((y & 3) << 3); }

static inline uint64_t size_w(int lx, int ly) { return 1ull << (lx + ly - 4); }

static inline uint64_t cb_index(int lx, int ly, int x, int y, int dir)
{
	const uint64_t sw = size_w(lx, ly);
	const int sub = (x ^ y) & 1;
	uint64_t w = ((uint64_t)y << (lx - 6)) + (uint64_t)(x >> 6);
	if(sub)
		w += sw >> 1;
	if(dir)
		w += sw >> 2;
	return w;
}

int kor_get_slope(const uint32_t* lat, int enc, int lx, int ly, int x, int y, int dir)
{
	if(enc == 0)
		return (lat[loc_index(lx, x, y)] >> (loc_shift(x, y) | dir)) & 1;
	return (lat[cb_index(lx, ly, x, y, dir)] >> ((x >> 1) & 31)) & 1;
}

static inline void flip_slope(uint32_t* lat, int enc, int lx, int ly, int x, int y, int dir)
{
	if(enc == 0)
		lat[loc_index(lx, x, y)] ^= 1u << (loc_shift(x, y) | dir);
	else
		lat[cb_index(lx, ly, x, y, dir)] ^= 1u << ((x >> 1) & 31);
}

/* SchedulerService::initHomogeneous, kpz/schedulerService.cpp:156-175 */
void kor_init_flat_local(uint32_t* lat, int lx, int ly)
{
	const uint64_t sw = size_w(lx, ly);
	for(uint64_t i = 0; i < sw; ++i)
		lat[i] = 0x33CC33CCu;
}

void kor_init_flat_cbbit(uint32_t* lat, int lx, int ly)
{
	const uint64_t sw = size_w(lx, ly);
	memset(lat, 0, (size_t)(sw << 1));
	memset(lat + (sw >> 1), 0xFF, (size_t)(sw << 1));
}

void kor_local_to_cbbit(const uint32_t* in, uint32_t* out, int lx, int ly)
{
	memset(out, 0, (size_t)(size_w(lx, ly) * 4));
	for(int y = 0; y < (1 << ly); ++y)
		for(int x = 0; x < (1 << lx); ++x)
			for(int d = 0; d < 2; ++d)
				if(kor_get_slope(in, 0, lx, ly, x, y, d))
					flip_slope(out, 1, lx, ly, x, y, d);
}

void kor_cbbit_to_local(const uint32_t* in, uint32_t* out, int lx, int ly)
{
	memset(out, 0, (size_t)(size_w(lx, ly) * 4));
	for(int y = 0; y < (1 << ly); ++y)
		for(int x = 0; x < (1 << lx); ++x)
			for(int d = 0; d < 2; ++d)
				if(kor_get_slope(in, 1, lx, ly, x, y, d))
					flip_slope(out, 0, lx, ly, x, y, d);
}

/* ======================================================================
 * heights and reductions, kpz/correlator.cpp:37-106 (Correlator::set height
 * convention), :127-164 (correlate: slope overlap), :450-455 (ch, cs),
 * kpz/roughness.cpp:38-43 (w2)
 * ====================================================================== */
void kor_heights(const uint32_t* lat, int enc, int lx, int ly, int32_t* h)
{
	const int X = 1 << lx, Y = 1 << ly;
	int hh = 0;
	for(int y = 0; y < Y; ++y)
	{
		hh += kor_get_slope(lat, enc, lx, ly, 0, y, 1) ? 1 : -1;
		for(int x = 1; x < X; ++x)
		{
			hh += kor_get_slope(lat, enc, lx, ly, x, y, 0) ? 1 : -1;
			h[(size_t)y * X + x] = hh;
		}
		hh += kor_get_slope(lat, enc, lx, ly, 0, y, 0) ? 1 : -1;
		h[(size_t)y * X] = hh;
	}
}

void kor_roughness(const uint32_t* lat, int enc, int lx, int ly, int64_t out[3])
{
	const size_t N = (size_t)1 << (lx + ly);
	int32_t* h = (int32_t*)malloc(N * sizeof(int32_t));
	kor_heights(lat, enc, lx, ly, h);
	int64_t s1 = 0;
	__int128 s2 = 0;
	for(size_t i = 0; i < N; ++i)
	{
		s1 += h[i];
		s2 += (int64_t)h[i] * h[i];
	}
	free(h);
	out[0] = s1;
	out[1] = (int64_t)(uint64_t)s2;
	out[2] = (int64_t)(s2 >> 64);
}

void kor_correlate(const uint32_t* sys, const uint32_t* snap, int enc, int lx, int ly, int64_t out[7])
{
	const size_t N = (size_t)1 << (lx + ly);
	int32_t* h = (int32_t*)malloc(N * sizeof(int32_t));
	int32_t* g = (int32_t*)malloc(N * sizeof(int32_t));
	kor_heights(sys, enc, lx, ly, h);
	kor_heights(snap, enc, lx, ly, g);
	int64_t s1 = 0, t1 = 0, tcs = 0;
	__int128 s2 = 0, s11 = 0;
	for(size_t i = 0; i < N; ++i)
	{
		s1 += h[i];
		t1 += g[i];
		s2 += (int64_t)h[i] * h[i];
		s11 += (int64_t)h[i] * g[i];
	}
	for(int y = 0; y < (1 << ly); ++y)
		for(int x = 0; x < (1 << lx); ++x)
			for(int d = 0; d < 2; ++d)
				tcs += (kor_get_slope(sys, enc, lx, ly, x, y, d) == kor_get_slope(snap, enc, lx, ly, x, y, d)) ? 1 : -1;
	free(h);
	free(g);
	out[0] = s1;
	out[1] = t1;
	out[2] = (int64_t)(uint64_t)s2;
	out[3] = (int64_t)(s2 >> 64);
	out[4] = (int64_t)(uint64_t)s11;
	out[5] = (int64_t)(s11 >> 64);
	out[6] = tcs;
}

/* The same sums as kor_roughness / kor_correlate, row by row on nthreads
 * threads, with no lattice-sized height array (a 65536^2 lattice would need
 * 16 GiB per height field).  Row y is walked exactly as kor_heights walks it,
 * starting from 0 instead of the height carried in from row y-1: h_rel(x,y) =
 * sY(0,y) + sum_{x'=1..x} sX(x',y) for x >= 1 and h_rel(0,y) = that plus
 * sX(0,y) =: D(y), the row's carry.  With base(y) = sum_{y'<y} D(y') the true
 * height is base(y) + h_rel(x,y), so per row only sum h_rel, sum h_rel^2
 * (and against the snapshot sum g_rel, sum h_rel g_rel, the slope overlap)
 * are kept; the host combines rows in 128-bit arithmetic. */
typedef struct {
	const uint32_t *sys, *snap;
	int enc, lx, ly, y0, y1;
	int64_t *D, *H1, *H2, *DG, *G1, *HG, *CS;
} rows_work;

static void* rows_run(void* arg)
{
	rows_work* W = (rows_work*)arg;
	const int X = 1 << W->lx;
	for(int y = W->y0; y < W->y1; ++y)
	{
		int64_t h = kor_get_slope(W->sys, W->enc, W->lx, W->ly, 0, y, 1) ? 1 : -1;
		int64_t g = 0, h1 = 0, h2 = 0, g1 = 0, hg = 0, cs = 0;
		if(W->snap)
			g = kor_get_slope(W->snap, W->enc, W->lx, W->ly, 0, y, 1) ? 1 : -1;
		for(int x = 1; x <= X; ++x)
		{
			const int xs = x & (X - 1); /* x = X adds the slope at x = 0 last */
			h += kor_get_slope(W->sys, W->enc, W->lx, W->ly, xs, y, 0) ? 1 : -1;
			h1 += h;
			h2 += h * h;
			if(W->snap)
			{
				g += kor_get_slope(W->snap, W->enc, W->lx, W->ly, xs, y, 0) ? 1 : -1;
				g1 += g;
				hg += h * g;
			}
		}
		W->D[y] = h;
		W->H1[y] = h1;
		W->H2[y] = h2;
		if(W->snap)
		{
			for(int x = 0; x < X; ++x)
				for(int d = 0; d < 2; ++d)
					cs += (kor_get_slope(W->sys, W->enc, W->lx, W->ly, x, y, d)
						== kor_get_slope(W->snap, W->enc, W->lx, W->ly, x, y, d)) ? 1 : -1;
			W->DG[y] = g;
			W->G1[y] = g1;
			W->HG[y] = hg;
			W->CS[y] = cs;
		}
	}
	return NULL;
}

static void rows_sums(const uint32_t* sys, const uint32_t* snap, int enc, int lx, int ly, int nthreads, int64_t out[7])
{
	const int Y = 1 << ly;
	const int64_t X = 1ll << lx;
	if(nthreads < 1)
		nthreads = 1;
	if(nthreads > Y)
		nthreads = Y;
	int64_t* buf = (int64_t*)calloc((size_t)Y * 7, sizeof(int64_t));
	rows_work* wk = (rows_work*)malloc(sizeof(rows_work) * (size_t)nthreads);
	pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
	for(int w = 0; w < nthreads; ++w)
	{
		rows_work* W = &wk[w];
		W->sys = sys;
		W->snap = snap;
		W->enc = enc;
		W->lx = lx;
		W->ly = ly;
		W->y0 = (int)((int64_t)Y * w / nthreads);
		W->y1 = (int)((int64_t)Y * (w + 1) / nthreads);
		W->D = buf;
		W->H1 = buf + Y;
		W->H2 = buf + 2 * (size_t)Y;
		W->DG = buf + 3 * (size_t)Y;
		W->G1 = buf + 4 * (size_t)Y;
		W->HG = buf + 5 * (size_t)Y;
		W->CS = buf + 6 * (size_t)Y;
		pthread_create(&th[w], NULL, rows_run, W);
	}
	for(int w = 0; w < nthreads; ++w)
		pthread_join(th[w], NULL);
	const rows_work* R = &wk[0];
	int64_t bh = 0, bg = 0, s1 = 0, t1 = 0, tcs = 0;
	__int128 s2 = 0, s11 = 0;
	for(int y = 0; y < Y; ++y)
	{
		s1 += X * bh + R->H1[y];
		s2 += (__int128)X * bh * bh + (__int128)2 * bh * R->H1[y] + R->H2[y];
		if(snap)
		{
			t1 += X * bg + R->G1[y];
			s11 += (__int128)X * bh * bg + (__int128)bh * R->G1[y] + (__int128)bg * R->H1[y] + R->HG[y];
			tcs += R->CS[y];
			bg += R->DG[y];
		}
		bh += R->D[y];
	}
	free(th);
	free(wk);
	free(buf);
	out[0] = s1;
	out[1] = t1;
	out[2] = (int64_t)(uint64_t)s2;
	out[3] = (int64_t)(s2 >> 64);
	out[4] = (int64_t)(uint64_t)s11;
	out[5] = (int64_t)(s11 >> 64);
	out[6] = tcs;
}

void kor_roughness_rows(const uint32_t* lat, int enc, int lx, int ly, int nthreads, int64_t out[3])
{
	int64_t o[7];
	rows_sums(lat, NULL, enc, lx, ly, nthreads, o);
	out[0] = o[0];
	out[1] = o[2];
	out[2] = o[3];
}

void kor_correlate_rows(const uint32_t* sys, const uint32_t* snap, int enc, int lx, int ly, int nthreads, int64_t out[7])
{
	rows_sums(sys, snap, enc, lx, ly, nthreads, out);
}

/* SchedulerService::consitencyCheck, kpz/schedulerService.cpp:721-854:
 * row sums of x slopes, column sums of y slopes, plaquette closure */
int kor_consistent(const uint32_t* lat, int enc, int lx, int ly)
{
	const int X = 1 << lx, Y = 1 << ly;
	for(int y = 0; y < Y; ++y)
	{
		int s = 0;
		for(int x = 0; x < X; ++x)
			s += kor_get_slope(lat, enc, lx, ly, x, y, 0) ? 1 : -1;
		if(s)
			return 0;
	}
	for(int x = 0; x < X; ++x)
	{
		int s = 0;
		for(int y = 0; y < Y; ++y)
			s += kor_get_slope(lat, enc, lx, ly, x, y, 1) ? 1 : -1;
		if(s)
			return 0;
	}
	for(int y = 0; y < Y; ++y)
		for(int x = 0; x < X; ++x)
		{
			const int x1 = (x + 1) & (X - 1), y1 = (y + 1) & (Y - 1);
			/* h(x+1,y+1)-h(x,y) along right-then-up and up-then-right */
			const int a = (kor_get_slope(lat, enc, lx, ly, x1, y, 0) ? 1 : -1) + (kor_get_slope(lat, enc, lx, ly, x1, y1, 1) ? 1 : -1);
			const int b = (kor_get_slope(lat, enc, lx, ly, x, y1, 1) ? 1 : -1) + (kor_get_slope(lat, enc, lx, ly, x1, y1, 0) ? 1 : -1);
			if(a != b)
				return 0;
		}
	return 1;
}

/* ======================================================================
 * update rules
 * ====================================================================== */

/* randomFloat() < p  <=>  rf < ceil(p * 2^bits), exact because randomFloat
 * is rf * 2^-bits exactly (bits = 23 for SLCG64, 24 for TinyMT32) and the
 * comparison is done in double (kpz/schedulerCPUBaseThreadedRng.inl:34-44). */
static uint32_t thresh_bits(double p, int bits)
{
	if(!(p > 0.0))
		return 0;
	if(p >= 1.0)
		return 1u << bits;
	return (uint32_t)ceil(p * (double)(1u << bits));
}

/* thresholds per disorder class and the disorder map */
typedef struct {
	uint32_t tp[4], tq[4];
	const uint32_t* dis;
} thr_t;

static thr_t thr_pq_bits(double p, double q, int bits)
{
	thr_t t;
	memset(&t, 0, sizeof t);
	t.tp[0] = thresh_bits(p, bits);
	t.tq[0] = thresh_bits(q, bits);
	return t;
}

static thr_t thr_pq(double p, double q) { return thr_pq_bits(p, q, 23); }

static thr_t thr_rates_bits(const kor_rates* r, int bits)
{
	thr_t t;
	for(int k = 0; k < 4; ++k)
	{
		t.tp[k] = thresh_bits(r->p[k], bits);
		t.tq[k] = thresh_bits(r->q[k], bits);
	}
	t.dis = r->disorder;
	return t;
}

static thr_t thr_rates(const kor_rates* r) { return thr_rates_bits(r, 23); }

/* SchedulerService::generateDisorder2, kpz/schedulerService.cpp:54-69 */
void kor_generate_disorder2(kor_dsfmt* master, uint32_t* dis, uint64_t size_w, double fill)
{
	for(uint64_t a = 0; a < size_w; ++a)
	{
		uint32_t t = 0;
		for(int s = 0; s < 32; s += 2)
			if(kor_dsfmt_close_open(master) < fill)
				t |= 1u << s;
		dis[a] = t;
	}
}

/* SchedulerCPU::pqOneUpdate, kpz/schedulerCPU.inl:21-33 (upd = KPZ),
 * Func_MCSBase::pqSyncRngUpdate, kpz/schedulerCPUBaseThreadedRng.inl:25-48
 * (upd = PQ; rf is the randomFloat numerator, on the thresholds' scale),
 * disorder2SyncRngUpdate :50-75 (upd = DIS: rates of the class c(m_disorder),
 * the 2 bits of the site in the disorder map) and SyncRngEW,
 * kpz/kpzSimulationPrimitivesTemplates.cu:199-234 (upd = EW) */
static inline void upd_local(uint32_t* lat, int lx, int ly, int x, int y, int upd, uint32_t rf, const thr_t* th)
{
	const uint32_t ic = loc_index(lx, x, y);
	const int sc = loc_shift(x, y);
	const int xr = (x + 1) & ((1 << lx) - 1);
	const int yu = (y + 1) & ((1 << ly) - 1);
	const uint32_t ir = loc_index(lx, xr, y);
	const int sr = loc_shift(xr, y);
	const uint32_t il = loc_index(lx, x, yu);
	const int sl = loc_shift(x, yu) | 1;
	const uint32_t c = (lat[ic] >> sc) & 3;
	const uint32_t r = (lat[ir] >> sr) & 1;
	const uint32_t l = (lat[il] >> sl) & 1;
	if(upd == KOR_UPD_KPZ)
	{
		if(!(c == 0 && r && l))
			return;
	}
	else
	{
		const uint32_t config = c | (r << 2) | (l << 3);
		const int k = (upd == KOR_UPD_DIS && th->dis) ? (int)((th->dis[ic] >> sc) & 3) : 0;
		if(config == 0xC)
		{
			if(upd != KOR_UPD_EW && !(rf < th->tp[k]))
				return;
		}
		else if(config == 0x3)
		{
			if(upd != KOR_UPD_EW && !(rf < th->tq[k]))
				return;
		}
		else
			return;
	}
	lat[ic] ^= 3u << sc;
	lat[ir] ^= 1u << sr;
	lat[il] ^= 1u << sl;
}

void kor_update_local(uint32_t* lat, int lx, int ly, int x, int y, int upd, float rnd, double p, double q)
{
	const thr_t th = thr_pq(p, q);
	upd_local(lat, lx, ly, x, y, upd, (uint32_t)(rnd * 8388608.0f), &th);
}

void kor_update_local_rates(uint32_t* lat, int lx, int ly, int x, int y, int upd, float rnd, const kor_rates* r)
{
	const thr_t th = thr_rates(r);
	upd_local(lat, lx, ly, x, y, upd, (uint32_t)(rnd * 8388608.0f), &th);
}

/* Func_MCSLocalBase::pqSyncRngUpdate on ENC_CBBIT,
 * kpz/schedulerCPUCBBit_nested.inl:262-295, and disorder2SyncRngUpdate
 * :298-331 (dis: class = cx(m_disorder), the x-slope bit of the site in the
 * disorder map read with the ENC_CBBIT layout) */
static inline void upd_cbbit(uint32_t* lat, int lx, int ly, int x, int y, uint32_t rf, const thr_t* th, int dis)
{
	const int xr = (x + 1) & ((1 << lx) - 1);
	const int yu = (y + 1) & ((1 << ly) - 1);
	const uint32_t config = (uint32_t)kor_get_slope(lat, 1, lx, ly, x, y, 0)
		| ((uint32_t)kor_get_slope(lat, 1, lx, ly, x, y, 1) << 1)
		| ((uint32_t)kor_get_slope(lat, 1, lx, ly, xr, y, 0) << 2)
		| ((uint32_t)kor_get_slope(lat, 1, lx, ly, x, yu, 1) << 3);
	const int k = (dis && th->dis) ? kor_get_slope(th->dis, 1, lx, ly, x, y, 0) : 0;
	if(config == 0xC)
	{
		if(!(rf < th->tp[k]))
			return;
	}
	else if(config == 0x3)
	{
		if(!(rf < th->tq[k]))
			return;
	}
	else
		return;
	flip_slope(lat, 1, lx, ly, x, y, 0);
	flip_slope(lat, 1, lx, ly, x, y, 1);
	flip_slope(lat, 1, lx, ly, xr, y, 0);
	flip_slope(lat, 1, lx, ly, x, yu, 1);
}

void kor_update_cbbit(uint32_t* lat, int lx, int ly, int x, int y, float rnd, double p, double q)
{
	const thr_t th = thr_pq(p, q);
	upd_cbbit(lat, lx, ly, x, y, (uint32_t)(rnd * 8388608.0f), &th, 0);
}

void kor_update_cbbit_rates(uint32_t* lat, int lx, int ly, int x, int y, float rnd, const kor_rates* r)
{
	const thr_t th = thr_rates(r);
	upd_cbbit(lat, lx, ly, x, y, (uint32_t)(rnd * 8388608.0f), &th, 1);
}

/* ======================================================================
 * double tiling: Kmc::DoubleTilingDD<2>::setBlock, inc/doubleTilingDD.h:71-78
 * with the BlockDD layout of inc/blockDD.h:137-160
 * ====================================================================== */
void kor_dt_block_offset(int lx, int ly, int lbx, int lby, int domain_set,
		const unsigned shift[2], int i, int off[2])
{
	const int lBlocks[2] = {lx - lbx, ly - lby};
	const int lBlockDim[2] = {lbx, lby};
	for(int a = 0; a < 2; ++a)
	{
		const int mBlocks = (1 << lBlocks[a]) - 1;
		i <<= 1;
		off[a] = (int)(shift[a] & ((1u << lBlockDim[a]) - 1))
			+ (((i & mBlocks) + ((domain_set >> a) & 1)) << lBlockDim[a]);
		i >>= lBlocks[a];
	}
}

/* SchedulerCPUDT<Rng>::mcs, kpz/schedulerCPUDT.cpp:65-90 with
 * Func_MCSBase::exec, kpz/schedulerCPUDT_nested.h:46-66 (random_origin = 0),
 * and SchedulerCPUDTr<Rng>::doMcs / Func_MCS::exec,
 * kpz/schedulerCPUDTr.cpp:66-98, :147-175 (random_origin = 1).
 * Rng = SLCG64 stream t with close_open() = draw * 2^-32 (contract), so
 * (int)(close_open()*2^b) = draw >> (32-b); randomFloat()*2^b = rf >> (bits-b)
 * (SLCG64 and, for SchedulerCPUDTr<TinyMT32> (kpz/kpzFrontend.cpp:308-310),
 * TinyMT32; the static DT draws close_open(), which RngIface::TinyMT32 lacks).
 * Blocks of one domain set are disjoint including their +1 halo, so the
 * sequential walk t = 0..T-1 equals any thread interleaving. */
static int dt_mcs(uint32_t* lat, int lx, int ly, int lbx, int lby,
		const streams_t* S, int64_t T,
		kor_dsfmt* master, int n, int random_origin, int upd, const thr_t* th)
{
	if(lbx < 2 || lby < 2 || lx - lbx < 1 || ly - lby < 1 || T < 1 || lbx > 23 || lby > 23)
		return -1;
	if(S->tmt && !random_origin)
		return -1;
	const int bits = gen_bits(S->tmt);
	const int64_t nb = 1ll << (lx - lbx + ly - lby - 2);
	const int64_t bs = 1ll << (lbx + lby);
	const int mx = (1 << lx) - 1, my = (1 << ly) - 1;
	for(int m = 0; m < n; ++m)
	{
		int seq[4];
		kor_shuffle4(master, seq);
		unsigned shift[2] = {0, 0};
		if(random_origin)
		{
			shift[0] = (unsigned)(kor_dsfmt_close_open(master) * (double)(1 << lbx));
			shift[1] = (unsigned)(kor_dsfmt_close_open(master) * (double)(1 << lby));
		}
		for(int d = 0; d < 4; ++d)
		{
			for(int64_t t = 0; t < T; ++t)
			{
				gen_t g = gen_get(S, t);
				for(int64_t i = t; i < nb; i += T)
				{
					int off[2];
					kor_dt_block_offset(lx, ly, lbx, lby, seq[d], shift, (int)i, off);
					for(int64_t k = 0; k < bs; ++k)
					{
						int x, y;
						if(random_origin)
						{
							x = ((int)(gen_rf(&g) >> (bits - lbx)) + off[0]) & mx;
							y = ((int)(gen_rf(&g) >> (bits - lby)) + off[1]) & my;
						}
						else
						{
							x = (int)(gen_random(&g) >> (32 - lbx)) + off[0];
							y = (int)(gen_random(&g) >> (32 - lby)) + off[1];
						}
						uint32_t rf = 0;
						if(upd == KOR_UPD_PQ || upd == KOR_UPD_DIS)
							rf = gen_rf(&g);
						upd_local(lat, lx, ly, x, y, upd, rf, th);
					}
				}
				gen_put(S, t, &g);
			}
		}
	}
	return 0;
}

int kor_dt_mcs(uint32_t* lat, int lx, int ly, int lbx, int lby,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int random_origin, int upd, double p, double q)
{
	const thr_t th = thr_pq(p, q);
	const streams_t S = streams_slcg(states, a, c);
	return dt_mcs(lat, lx, ly, lbx, lby, &S, T, master, n, random_origin, upd, &th);
}

/* SchedulerCPUDTr / SchedulerCPUDT with per-class rates
 * (mcsDisorder2SyncRng, kpz/schedulerCPUDTr.cpp:115-119) */
int kor_dt_mcs_rates(uint32_t* lat, int lx, int ly, int lbx, int lby,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int random_origin, int upd, const kor_rates* r)
{
	const thr_t th = thr_rates(r);
	const streams_t S = streams_slcg(states, a, c);
	return dt_mcs(lat, lx, ly, lbx, lby, &S, T, master, n, random_origin, upd, &th);
}

/* SchedulerCPUDTr<TinyMT32> (kpz/kpzFrontend.cpp:308-310): upd KPZ / PQ / EW
 * with (p, q), DIS with the rates r (r may be NULL otherwise) */
int kor_dtr_mcs_tmt(uint32_t* lat, int lx, int ly, int lbx, int lby, kor_tmt* g, int64_t T,
		kor_dsfmt* master, int n, int upd, double p, double q, const kor_rates* r)
{
	if(upd == KOR_UPD_DIS && !r)
		return -1;
	const thr_t th = (upd == KOR_UPD_DIS) ? thr_rates_bits(r, 24) : thr_pq_bits(p, q, 24);
	const streams_t S = streams_tmt(g);
	return dt_mcs(lat, lx, ly, lbx, lby, &S, T, master, n, 1, upd, &th);
}

/* Threaded form of the above for timing: worker w runs streams w, w+nthreads,
 * ...; one pthread per worker per domain set, spawned and joined per phase
 * as SchedulerCPUpThread::run / joinMCS (kpz/schedulerCPUpThread.h:44-63,
 * kpz/schedulerCPUDT.cpp:75-86).  Same result as kor_dt_mcs. */
#include <pthread.h>

typedef struct {
	uint32_t* lat;
	int lx, ly, lbx, lby, ds, random_origin, upd;
	unsigned shift[2];
	uint64_t* states;
	int64_t T, w, nw, nb, bs;
	uint64_t a, c;
	thr_t th;
} dt_work;

static void* dt_worker(void* arg)
{
	dt_work* W = (dt_work*)arg;
	const int mx = (1 << W->lx) - 1, my = (1 << W->ly) - 1;
	for(int64_t t = W->w; t < W->T; t += W->nw)
	{
		uint64_t s = W->states[t];
		for(int64_t i = t; i < W->nb; i += W->T)
		{
			int off[2];
			kor_dt_block_offset(W->lx, W->ly, W->lbx, W->lby, W->ds, W->shift, (int)i, off);
			for(int64_t k = 0; k < W->bs; ++k)
			{
				int x, y;
				if(W->random_origin)
				{
					x = ((int)((kor_slcg_draw(&s, W->a, W->c) & 0x7FFFFF) >> (23 - W->lbx)) + off[0]) & mx;
					y = ((int)((kor_slcg_draw(&s, W->a, W->c) & 0x7FFFFF) >> (23 - W->lby)) + off[1]) & my;
				}
				else
				{
					x = (int)(kor_slcg_draw(&s, W->a, W->c) >> (32 - W->lbx)) + off[0];
					y = (int)(kor_slcg_draw(&s, W->a, W->c) >> (32 - W->lby)) + off[1];
				}
				uint32_t r23 = 0;
				if(W->upd == KOR_UPD_PQ || W->upd == KOR_UPD_DIS)
					r23 = kor_slcg_draw(&s, W->a, W->c) & 0x7FFFFF;
				upd_local(W->lat, W->lx, W->ly, x, y, W->upd, r23, &W->th);
			}
		}
		W->states[t] = s;
	}
	return NULL;
}

int kor_dt_mcs_threaded(uint32_t* lat, int lx, int ly, int lbx, int lby,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int random_origin, int upd, double p, double q, int nthreads)
{
	if(lbx < 2 || lby < 2 || lx - lbx < 1 || ly - lby < 1 || T < 1 || lbx > 23 || lby > 23 || nthreads < 1)
		return -1;
	pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
	dt_work* wk = (dt_work*)malloc(sizeof(dt_work) * (size_t)nthreads);
	for(int m = 0; m < n; ++m)
	{
		int seq[4];
		kor_shuffle4(master, seq);
		unsigned shift[2] = {0, 0};
		if(random_origin)
		{
			shift[0] = (unsigned)(kor_dsfmt_close_open(master) * (double)(1 << lbx));
			shift[1] = (unsigned)(kor_dsfmt_close_open(master) * (double)(1 << lby));
		}
		for(int d = 0; d < 4; ++d)
		{
			for(int w = 0; w < nthreads; ++w)
			{
				dt_work* W = &wk[w];
				W->lat = lat;
				W->lx = lx;
				W->ly = ly;
				W->lbx = lbx;
				W->lby = lby;
				W->ds = seq[d];
				W->random_origin = random_origin;
				W->upd = upd;
				W->shift[0] = shift[0];
				W->shift[1] = shift[1];
				W->states = states;
				W->T = T;
				W->w = w;
				W->nw = nthreads;
				W->nb = 1ll << (lx - lbx + ly - lby - 2);
				W->bs = 1ll << (lbx + lby);
				W->a = a;
				W->c = c;
				W->th = thr_pq(p, q);
				pthread_create(&th[w], NULL, dt_worker, W);
			}
			for(int w = 0; w < nthreads; ++w)
				pthread_join(th[w], NULL);
		}
	}
	free(th);
	free(wk);
	return 0;
}

/* ======================================================================
 * SchedulerCPUDB<Rng>::mcs, kpz/schedulerCPUDB.cpp:65-88, with the
 * dead-border worker Func_MCSBase::exec, kpz/schedulerCPUDB_nested.h:56-80,
 * the dynamic origin Kmc::DynamicOriginDD<2>, inc/dynamicOriginDD.h:51-75,
 * and the block layout of Kmc::BlockDD::setLayout, inc/blockDD.h:137-160.
 * Per MCS the origin is drawn from the master (x first, then y:
 * o = (int)(close_open * 2^(lbd - lORest)), origin = o << lORest); every
 * block i (x block = low bits) then gets (2^lbx - 1)(2^lby - 1) updates at
 * ((int)(close_open*(2^lb - 1)) + offset) & (L - 1), x drawn before y,
 * pqOneUpdate (DEAD_BORDER_DIM = 1, kpz/kpzConst.h:59).  Worker t walks
 * blocks t, t+T, ... on its own stream (DDIterator::next(T),
 * inc/ddIterators.h:47-69).  The layout is explicit here (contract): the
 * reference derives it from the host's L1 size (layoutFromL1,
 * inc/cacheLayout.h) unless setLayout is called.  Blocks are disjoint
 * including the one-site border every update reaches, so the order of the
 * workers does not matter (threads toggle bits atomically: blocks off the
 * word grid share words).  close_open() of an SLCG stream is draw * 2^-32
 * (contract), so (int)(close_open()*m) = (draw * m) >> 32 exactly.
 * ====================================================================== */
/* pqOneUpdate with word-atomic loads and bit toggles: the blocks own
 * disjoint sites, but with lorest > 0 their origins are not on the 4-site
 * word grid, so two threads' blocks can share a lattice word */
static inline void upd_local_kpz_atomic(uint32_t* lat, int lx, int ly, int x, int y)
{
	const uint32_t ic = loc_index(lx, x, y);
	const int sc = loc_shift(x, y);
	const int xr = (x + 1) & ((1 << lx) - 1);
	const int yu = (y + 1) & ((1 << ly) - 1);
	const uint32_t ir = loc_index(lx, xr, y);
	const int sr = loc_shift(xr, y);
	const uint32_t il = loc_index(lx, x, yu);
	const int sl = loc_shift(x, yu) | 1;
	const uint32_t c = (__atomic_load_n(&lat[ic], __ATOMIC_RELAXED) >> sc) & 3;
	const uint32_t r = (__atomic_load_n(&lat[ir], __ATOMIC_RELAXED) >> sr) & 1;
	const uint32_t l = (__atomic_load_n(&lat[il], __ATOMIC_RELAXED) >> sl) & 1;
	if(!(c == 0 && r && l))
		return;
	__atomic_fetch_xor(&lat[ic], 3u << sc, __ATOMIC_RELAXED);
	__atomic_fetch_xor(&lat[ir], 1u << sr, __ATOMIC_RELAXED);
	__atomic_fetch_xor(&lat[il], 1u << sl, __ATOMIC_RELAXED);
}

typedef struct {
	uint32_t* lat;
	int lx, ly, lbx, lby;
	int org[2];
	uint64_t* states;
	int64_t T, w, nw, nb;
	uint64_t a, c;
	thr_t th;
} db_work;

static void* db_worker(void* arg)
{
	db_work* W = (db_work*)arg;
	const int mx = (1 << W->lx) - 1, my = (1 << W->ly) - 1;
	const uint64_t ax = (1u << W->lbx) - 1, ay = (1u << W->lby) - 1;
	const int mbx = (1 << (W->lx - W->lbx)) - 1;
	for(int64_t t = W->w; t < W->T; t += W->nw)
	{
		uint64_t s = W->states[t];
		for(int64_t i = t; i < W->nb; i += W->T)
		{
			const int offx = (int)((i & mbx) << W->lbx) + W->org[0];
			const int offy = (int)((i >> (W->lx - W->lbx)) << W->lby) + W->org[1];
			for(uint64_t k = 0; k < ax * ay; ++k)
			{
				const int x = ((int)(((uint64_t)kor_slcg_draw(&s, W->a, W->c) * ax) >> 32) + offx) & mx;
				const int y = ((int)(((uint64_t)kor_slcg_draw(&s, W->a, W->c) * ay) >> 32) + offy) & my;
				if(W->nw > 1)
					upd_local_kpz_atomic(W->lat, W->lx, W->ly, x, y);
				else
					upd_local(W->lat, W->lx, W->ly, x, y, KOR_UPD_KPZ, 0, &W->th);
			}
		}
		W->states[t] = s;
	}
	return NULL;
}

int kor_db_mcs(uint32_t* lat, int lx, int ly, int lbx, int lby, int lorest,
		uint64_t* states, int64_t T, uint64_t a, uint64_t c,
		kor_dsfmt* master, int n, int nthreads)
{
	if(lbx < 2 || lby < 2 || lbx > lx || lby > ly || lbx > 16 || lby > 16 || lorest < 0 || lorest > lbx ||
	   lorest > lby || T < 1 || nthreads < 1)
		return -1;
	const int64_t nb = 1ll << (lx - lbx + ly - lby);
	if(T > nb)
		return -1; /* a worker without a block would run on an unset iterator */
	pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
	db_work* wk = (db_work*)malloc(sizeof(db_work) * (size_t)nthreads);
	const int lbd[2] = {lbx, lby};
	for(int m = 0; m < n; ++m)
	{
		int org[2];
		for(int d = 0; d < 2; ++d)
		{
			const int eff = 1 << (lbd[d] - lorest);
			org[d] = (int)(kor_dsfmt_close_open(master) * (double)eff) << lorest;
		}
		for(int w = 0; w < nthreads; ++w)
		{
			db_work* W = &wk[w];
			W->lat = lat;
			W->lx = lx;
			W->ly = ly;
			W->lbx = lbx;
			W->lby = lby;
			W->org[0] = org[0];
			W->org[1] = org[1];
			W->states = states;
			W->T = T;
			W->w = w;
			W->nw = nthreads;
			W->nb = nb;
			W->a = a;
			W->c = c;
			W->th = thr_pq(1.0, 0.0);
			if(nthreads > 1)
				pthread_create(&th[w], NULL, db_worker, W);
			else
				db_worker(W);
		}
		if(nthreads > 1)
			for(int w = 0; w < nthreads; ++w)
				pthread_join(th[w], NULL);
	}
	free(th);
	free(wk);
	return 0;
}

/* ======================================================================
 * checkerboard SCA on ENC_CBBIT: SchedulerCPUCBBit<Rng>::doMcs,
 * kpz/schedulerCPUCBBit.cpp:43-79, Func_MCS::exec
 * kpz/schedulerCPUCBBit_nested.h:80-95 (rows [yMin, yCenter) then
 * [yCenter, ySup), i.e. ascending), pSyncRngUpdateVL3
 * kpz/schedulerCPUCBBit_nested.inl:200-259, pqSyncRngUpdate :262-295.
 * ====================================================================== */
static inline uint64_t ld64(const uint32_t* p, uint64_t i)
{
	uint64_t v;
	memcpy(&v, p + 2 * i, 8);
	return v;
}
static inline void st64(uint32_t* p, uint64_t i, uint64_t v) { memcpy(p + 2 * i, &v, 8); }

/* one row of the bit-parallel half-sweep.  wbits = 64 (VL3: two draws per
 * word, first draw is the high half -- contract for the unsequenced
 * (r()<<32)^r() of :240) or 32 (L = 64 only: one draw per u32 word --
 * contract; the reference's 64-bit loop does nothing when L < 128). */
static void cbbit_row_words(uint32_t* lat, int lx, int ly, int s, int y, int wbits,
		gen_t* g, const uint64_t* masks)
{
	const int lw = (wbits == 64) ? 7 : 6; /* log2 sites per word pair */
	const uint64_t Q = size_w(lx, ly) >> (wbits == 64 ? 3 : 2);
	const uint64_t W = (uint64_t)1 << (lx - lw);
	const uint64_t MSB = (wbits == 64) ? (1ull << 63) : (1ull << 31);
	const uint64_t FULL = (wbits == 64) ? ~0ull : 0xFFFFFFFFull;
	const int parity = (s == (y & 1));
	const uint64_t thisBase = s ? (Q << 1) : 0;
	const uint64_t othBase = s ? 0 : (Q << 1);
#define LD(i) ((wbits == 64) ? ld64(lat, (i)) : (uint64_t)lat[(i)])
#define ST(i, v) do { if(wbits == 64) st64(lat, (i), (v)); else lat[(i)] = (uint32_t)(v); } while(0)
	for(uint64_t xw = 0; xw < W; ++xw)
	{
		const uint64_t off = xw + (uint64_t)y * W;
		const uint64_t tXi = thisBase + off, tYi = thisBase + Q + off;
		const uint64_t tX = LD(tXi), tY = LD(tYi);
		const uint64_t yIdx = othBase + Q + xw + (uint64_t)((y + 1) & ((1 << ly) - 1)) * W;
		const uint64_t yY = LD(yIdx);
		const uint64_t xIdx = othBase + off;
		uint64_t xX = LD(xIdx);
		uint64_t nIdx = xIdx;
		if(!parity)
		{
			nIdx = (xw + 1 == W) ? xIdx - xw : xIdx + 1;
			xX = ((xX >> 1) & ~MSB) | ((LD(nIdx) << (wbits - 1)) & MSB);
		}
		uint64_t m = (~(tX | tY)) & xX & yY & FULL;
		uint64_t rnd;
		if(masks)
			rnd = masks[off];
		else if(wbits == 64)
		{
			const uint64_t hi = gen_random(g);
			const uint64_t lo = gen_random(g);
			rnd = (hi << 32) ^ lo;
		}
		else
			rnd = gen_random(g);
		m &= rnd;
		uint64_t xm = m;
		if(!parity)
		{
			ST(nIdx, LD(nIdx) ^ ((m >> (wbits - 1)) & 1));
			xm = (m << 1) & FULL;
		}
		ST(xIdx, LD(xIdx) ^ xm);
		ST(tXi, tX ^ m);
		ST(tYi, tY ^ m);
		ST(yIdx, yY ^ m);
	}
#undef LD
#undef ST
}

void kor_cbbit_halfsweep_masked(uint32_t* lat, int lx, int ly, int signum, const uint64_t* masks)
{
	const int wbits = (lx >= 7) ? 64 : 32;
	for(int y = 0; y < (1 << ly); ++y)
		cbbit_row_words(lat, lx, ly, signum, y, wbits, NULL, masks);
}

static void cbbit_rows(int L, int64_t T, int64_t t, int* yMin, int* ySup)
{
	const int64_t per = L / T, rem = L % T;
	const int64_t lo = t * per + (t < rem ? t : rem);
	*yMin = (int)lo;
	*ySup = (int)(lo + per + (t < rem ? 1 : 0));
}

/* worker t's rows of half-sweep s (kpz/schedulerCPUCBBit_nested.inl:200-295) */
static void cbbit_worker(uint32_t* lat, int lx, int ly, const streams_t* S, int64_t T, int64_t t,
		int s, int mode, const thr_t* th)
{
	const int X = 1 << lx, Y = 1 << ly;
	const int wbits = (lx >= 7) ? 64 : 32;
	int y0, y1;
	cbbit_rows(Y, T, t, &y0, &y1);
	gen_t g = gen_get(S, t);
	for(int y = y0; y < y1; ++y)
	{
		if(mode == 0)
			cbbit_row_words(lat, lx, ly, s, y, wbits, &g, NULL);
		else
			for(int x = ((y & 1) ^ s); x < X; x += 2)
				upd_cbbit(lat, lx, ly, x, y, gen_rf(&g), th, mode == 2);
	}
	gen_put(S, t, &g);
}

static int cbbit_mcs(uint32_t* lat, int lx, int ly, const streams_t* S, int64_t T,
		int n, int mode, const thr_t* th)
{
	if(lx < 6 || ly < 1 || T < 1)
		return -1;
	for(int m = 0; m < n; ++m)
		for(int s = 0; s < 2; ++s)
			for(int64_t t = 0; t < T; ++t)
				cbbit_worker(lat, lx, ly, S, T, t, s, mode, th);
	return 0;
}

int kor_cbbit_mcs(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int mode, double p, double q)
{
	const thr_t th = thr_pq(p, q);
	const streams_t S = streams_slcg(states, a, c);
	return cbbit_mcs(lat, lx, ly, &S, T, n, mode, &th);
}

int kor_cbbit_mcs_rates(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int mode, const kor_rates* r)
{
	if(mode != 1 && mode != 2)
		return -1;
	const thr_t th = thr_rates(r);
	const streams_t S = streams_slcg(states, a, c);
	return cbbit_mcs(lat, lx, ly, &S, T, n, mode, &th);
}

/* SchedulerCPUCBBit<TinyMT32> (kpz/kpzFrontend.cpp:328-331): mode 0 = VL3,
 * 1 = per-site pq with (p, q), 2 = disorder2SyncRngUpdate with r */
int kor_cbbit_mcs_tmt(uint32_t* lat, int lx, int ly, kor_tmt* g, int64_t T,
		int n, int mode, double p, double q, const kor_rates* r)
{
	if(mode < 0 || mode > 2 || (mode == 2 && !r))
		return -1;
	const thr_t th = (mode == 2) ? thr_rates_bits(r, 24) : thr_pq_bits(p, q, 24);
	const streams_t S = streams_tmt(g);
	return cbbit_mcs(lat, lx, ly, &S, T, n, mode, &th);
}

/* ======================================================================
 * checkerboard SCA on ENC_LOCAL: SchedulerCPUCB<Rng>::doMcs,
 * kpz/schedulerCPUCB.cpp:43-80, Func_MCS::updateLoop :128-142 and exec
 * :153-168 (strip split at xCenter, first half entirely before second).
 * ====================================================================== */
static void cb_loop(uint32_t* lat, int lx, int ly, int s, int xMin, int xSup,
		gen_t* g, int upd, const thr_t* th)
{
	for(int yW = 0; yW < (1 << ly); yW += 4)
		for(int xW = xMin; xW < xSup; xW += 4)
			for(int y = yW; y < yW + 4; ++y)
			{
				const int x = (xW + (y & 1)) ^ s;
				upd_local(lat, lx, ly, x, y, upd, gen_rf(g), th);
				upd_local(lat, lx, ly, x + 2, y, upd, gen_rf(g), th);
			}
}

/* worker t's strip of half-sweep s: columns [xMin, xSup), a whole number of
 * 4-site cells, split at xCenter (kpz/schedulerCPUCB.cpp:153-168) */
static void cb_worker(uint32_t* lat, int lx, int ly, const streams_t* S, int64_t T, int64_t t,
		int s, int upd, const thr_t* th)
{
	const int64_t dimXW = (int64_t)1 << (lx - 2);
	const int64_t per = dimXW / T, rem = dimXW % T;
	const int xMin = (int)(4 * (t * per + (t < rem ? t : rem)));
	const int xSup = xMin + (int)(4 * (per + (t < rem ? 1 : 0)));
	const int xCenter = xMin + (((xSup - xMin) / 2) & ~3);
	gen_t g = gen_get(S, t);
	cb_loop(lat, lx, ly, s, xMin, xCenter, &g, upd, th);
	cb_loop(lat, lx, ly, s, xCenter, xSup, &g, upd, th);
	gen_put(S, t, &g);
}

static int cb_mcs(uint32_t* lat, int lx, int ly, const streams_t* S, int64_t T,
		int n, int upd, const thr_t* th)
{
	if(lx < 2 || ly < 2 || T < 1)
		return -1;
	for(int m = 0; m < n; ++m)
		for(int s = 0; s < 2; ++s)
			for(int64_t t = 0; t < T; ++t)
				cb_worker(lat, lx, ly, S, T, t, s, upd, th);
	return 0;
}

int kor_cb_mcs(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, double p, double q)
{
	const thr_t th = thr_pq(p, q);
	const streams_t S = streams_slcg(states, a, c);
	return cb_mcs(lat, lx, ly, &S, T, n, KOR_UPD_PQ, &th);
}

/* SchedulerCPUCB::mcsDisorder2SyncRng, kpz/schedulerCPUCB.cpp:103-108 (upd = DIS) */
int kor_cb_mcs_rates(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int upd, const kor_rates* r)
{
	if(upd != KOR_UPD_PQ && upd != KOR_UPD_DIS)
		return -1;
	const thr_t th = thr_rates(r);
	const streams_t S = streams_slcg(states, a, c);
	return cb_mcs(lat, lx, ly, &S, T, n, upd, &th);
}

/* SchedulerCPUCB<TinyMT32> (kpz/kpzFrontend.cpp:324-327): upd PQ with (p, q)
 * or DIS with r */
int kor_cb_mcs_tmt(uint32_t* lat, int lx, int ly, kor_tmt* g, int64_t T,
		int n, int upd, double p, double q, const kor_rates* r)
{
	if((upd != KOR_UPD_PQ && upd != KOR_UPD_DIS) || (upd == KOR_UPD_DIS && !r))
		return -1;
	const thr_t th = (upd == KOR_UPD_DIS) ? thr_rates_bits(r, 24) : thr_pq_bits(p, q, 24);
	const streams_t S = streams_tmt(g);
	return cb_mcs(lat, lx, ly, &S, T, n, upd, &th);
}

/* ======================================================================
 * Threaded forms of the two checkerboard loops, for checks at large sizes.
 * Within a half-sweep no update reads a bit that another update writes
 * (sites of one sublattice own disjoint bits), and each worker draws from its
 * own stream in its own order, so any schedule of the workers gives the
 * sequential result.  What threads must not share is a word written by two
 * of them at once:
 *  - CBBit: worker t writes rows y0..y1-1 of quadrants s and the other
 *    sublattice's x slopes, and row y1 of the other sublattice's y slopes,
 *    which no other worker writes in this half-sweep: any split is safe;
 *  - CB: a strip writes its own cells and the x slopes of the first column
 *    right of it, a cell of the next strip.  Each thread takes a contiguous
 *    run of at least two workers and runs its first half of them, then (after
 *    all threads) its second half, so that concurrent writes are always at
 *    least one whole worker strip apart.
 * ====================================================================== */
typedef struct {
	uint32_t* lat;
	int lx, ly, s, mode, cb;
	const streams_t* S;
	int64_t T, t0, t1;
	const thr_t* th;
} sweep_work;

static void* sweep_run(void* arg)
{
	const sweep_work* W = (const sweep_work*)arg;
	for(int64_t t = W->t0; t < W->t1; ++t)
	{
		if(W->cb)
			cb_worker(W->lat, W->lx, W->ly, W->S, W->T, t, W->s, W->mode, W->th);
		else
			cbbit_worker(W->lat, W->lx, W->ly, W->S, W->T, t, W->s, W->mode, W->th);
	}
	return NULL;
}

static int sweep_threaded(uint32_t* lat, int lx, int ly, const streams_t* S, int64_t T, int n,
		int cb, int mode, const thr_t* th, int nthreads)
{
	if(nthreads < 1)
		return -1;
	int64_t nt = nthreads;
	if(cb && nt > T / 2)
		nt = T / 2;
	if(nt > T)
		nt = T;
	if(nt < 1)
		nt = 1;
	pthread_t* th_ = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
	sweep_work* wk = (sweep_work*)malloc(sizeof(sweep_work) * (size_t)nt);
	const int phases = cb ? 2 : 1;
	for(int m = 0; m < n; ++m)
		for(int s = 0; s < 2; ++s)
			for(int ph = 0; ph < phases; ++ph)
			{
				for(int64_t k = 0; k < nt; ++k)
				{
					const int64_t a = T * k / nt, b = T * (k + 1) / nt, mid = (a + b) / 2;
					sweep_work* W = &wk[k];
					W->lat = lat;
					W->lx = lx;
					W->ly = ly;
					W->s = s;
					W->mode = mode;
					W->cb = cb;
					W->S = S;
					W->T = T;
					W->t0 = (phases == 1) ? a : (ph == 0 ? a : mid);
					W->t1 = (phases == 1) ? b : (ph == 0 ? mid : b);
					W->th = th;
					pthread_create(&th_[k], NULL, sweep_run, W);
				}
				for(int64_t k = 0; k < nt; ++k)
					pthread_join(th_[k], NULL);
			}
	free(wk);
	free(th_);
	return 0;
}

/* SchedulerCPUCBBit (mode 0 VL3, 1 per-site pq) on nthreads threads */
int kor_cbbit_mcs_threaded(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, int mode, double p, double q, int nthreads)
{
	if(lx < 6 || ly < 1 || T < 1 || mode < 0 || mode > 1)
		return -1;
	const thr_t th = thr_pq(p, q);
	const streams_t S = streams_slcg(states, a, c);
	return sweep_threaded(lat, lx, ly, &S, T, n, 0, mode, &th, nthreads);
}

/* SchedulerCPUCB (pqSyncRngUpdate) on nthreads threads */
int kor_cb_mcs_threaded(uint32_t* lat, int lx, int ly, uint64_t* states, int64_t T,
		uint64_t a, uint64_t c, int n, double p, double q, int nthreads)
{
	if(lx < 2 || ly < 2 || T < 1)
		return -1;
	const thr_t th = thr_pq(p, q);
	const streams_t S = streams_slcg(states, a, c);
	return sweep_threaded(lat, lx, ly, &S, T, n, 1, KOR_UPD_PQ, &th, nthreads);
}

/* ======================================================================
 * plain random-sequential SchedulerCPU: SchedulerCPU::mcs(n, pred),
 * kpz/schedulerCPU.inl:74-88 with pqOneUpdate :21-33, pqOneOneUpdate :35-48,
 * pqUpdate :50-71; mcsPQ / mcsDisorder2 / mcsSyncRng / mcsPQSyncRng /
 * mcsDisorder2SyncRng, kpz/schedulerCPU.cpp:35-179 (RngPackage: three
 * master draws, kpz/schedulerService.cpp:519-523).
 * ====================================================================== */
int kor_rs_mcs(uint32_t* lat, int lx, int ly, kor_dsfmt* master, int n, int mode, const kor_rates* r)
{
	if(lx < 2 || ly < 2 || mode < KOR_RS_MCS || mode > KOR_RS_DIS_SYNC)
		return -1;
	if((mode == KOR_RS_DIS || mode == KOR_RS_DIS_SYNC) && (!r || !r->disorder))
		return -1;
	const int X = 1 << lx, Y = 1 << ly;
	const uint64_t N = (uint64_t)X * (uint64_t)Y;
	const int sync = (mode == KOR_RS_SYNC || mode == KOR_RS_PQ_SYNC || mode == KOR_RS_DIS_SYNC);
	const int ew = (mode == KOR_RS_PQ && r && r->p[0] == 1.0 && r->q[0] == 1.0);
	for(int m = 0; m < n; ++m)
		for(uint64_t a = 0; a < N; ++a)
		{
			const double rx = kor_dsfmt_close_open(master);
			const double ry = kor_dsfmt_close_open(master);
			const double rz = sync ? kor_dsfmt_close_open(master) : 0.0;
			const int x = (int)(rx * X), y = (int)(ry * Y);
			const uint32_t ic = loc_index(lx, x, y);
			const int sc = loc_shift(x, y);
			const int xr = (x + 1) & (X - 1), yu = (y + 1) & (Y - 1);
			const uint32_t ir = loc_index(lx, xr, y), il = loc_index(lx, x, yu);
			const int sr = loc_shift(xr, y), sl = loc_shift(x, yu) | 1;
			const uint32_t c = (lat[ic] >> sc) & 3, rb = (lat[ir] >> sr) & 1, lb = (lat[il] >> sl) & 1;
			const uint32_t config = c | (rb << 2) | (lb << 3);
			if(mode == KOR_RS_MCS || mode == KOR_RS_SYNC)
			{
				if(config != 0xC)
					continue;
			}
			else if(ew)
			{
				if(config != 0xC && config != 0x3)
					continue;
			}
			else
			{
				const int k = (mode == KOR_RS_DIS || mode == KOR_RS_DIS_SYNC) ? (int)((r->disorder[ic] >> sc) & 3) : 0;
				if(config == 0xC)
				{
					const double v = sync ? rz : kor_dsfmt_close_open(master);
					if(v >= r->p[k])
						continue;
				}
				else if(config == 0x3)
				{
					const double v = sync ? rz : kor_dsfmt_close_open(master);
					if(v >= r->q[k])
						continue;
				}
				else
					continue;
			}
			lat[ic] ^= 3u << sc;
			lat[ir] ^= 1u << sr;
			lat[il] ^= 1u << sl;
		}
	return 0;
}
