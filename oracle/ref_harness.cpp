/*
 * ORACLE (reference harness) -- test infrastructure only.
 *
 * Exposes, through a C ABI, the pieces of the reference that compile from
 * their own sources without any missing dependency: the vendored dSFMT
 * (inc/dSFMT/dSFMT.cpp), Kmc::Shuffle<4> (inc/shuffle.h), the double-tiling
 * block geometry Kmc::DoubleTilingDD<2> / Kmc::DDIterator
 * (inc/doubleTilingDD.h, inc/blockDD.h, inc/ddIterators.h) and the SLCG
 * skip constants (inc/kmcRandom.h, inc/helper/slcgInst.h), the vendored
 * TinyMT32 (inc/tinyMT/tinymt32.c) and the CPU schedulers' TinyMT32 stream
 * interface (kpz/CPURngIface.h, header part only: its parameter table
 * tinymt32_param.h is absent, so parameters are set per generator here).
 * The reference's scheduler classes need libSplash (absent), so they are not
 * built here.
 *
 * Built by oracle/Makefile straight from the files under the reference tree
 * into oracle/_ref/; used only by tests/golden/make_golden.py to write the
 * committed golden fixtures.
 */
#define KMC_NO_RNG_SHUFFLING
#include <dSFMT/dSFMT.h>
#include <shuffle.h>
#include <doubleTilingDD.h>
#include <ddIterators.h>
#include <kmcRandom.h>
#include <tinyMT/tinymt32.h>
#include <kpz/CPURngIface.h>

#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

/* dsfmt_init_gen_rand(seed), `discard` draws thrown away, then n draws */
void ref_dsfmt_draws(uint32_t seed, int discard, int n, double* out)
{
	dsfmt_t d;
	dsfmt_init_gen_rand(&d, seed);
	for(int a = 0; a < discard; ++a)
		dsfmt_genrand_close_open(&d);
	for(int a = 0; a < n; ++a)
		out[a] = dsfmt_genrand_close_open(&d);
}

/* n consecutive Kmc::Shuffle<4> permutations from a master seeded like
 * SchedulerService::reSeed (init + `discard` draws) */
void ref_shuffle4(uint32_t seed, int discard, int n, int* out)
{
	dsfmt_t d;
	dsfmt_init_gen_rand(&d, seed);
	for(int a = 0; a < discard; ++a)
		dsfmt_genrand_close_open(&d);
	for(int k = 0; k < n; ++k)
	{
		Kmc::Shuffle<4> seq(&d);
		for(int a = 0; a < 4; ++a)
			out[4 * k + a] = seq[a];
	}
}

/* offsets of every DT block (index order) of one domain set, after
 * setLayout({lbx,lby}, {lx,ly}) and setShift({sx,sy}); returns nBlocks */
int ref_dt_offsets(int lx, int ly, int lbx, int lby, int ds, unsigned sx, unsigned sy, int* out)
{
	dsfmt_gv_init_gen_rand(1); /* setLayout shuffles with the global dSFMT */
	Kmc::DoubleTilingDD<2> dd;
	const int layout[2] = {lbx, lby};
	const int size[2] = {lx, ly};
	dd.setLayout(layout, size);
	const unsigned shift[2] = {sx, sy};
	dd.setShift(shift);
	dd.setDomainSet(ds);
	const int nb = dd.nBlocks();
	for(int i = 0; i < nb; ++i)
	{
		Kmc::DDIterator<Kmc::DoubleTilingDD<2> > it(dd, i);
		out[2 * i] = it.blockOffset(0);
		out[2 * i + 1] = it.blockOffset(1);
	}
	return nb;
}

/* block indices visited by worker `id` of T (Func_MCSBase::exec loop) */
int ref_dt_walk(int lx, int ly, int lbx, int lby, int id, int T, int* out)
{
	dsfmt_gv_init_gen_rand(1);
	Kmc::DoubleTilingDD<2> dd;
	const int layout[2] = {lbx, lby};
	const int size[2] = {lx, ly};
	dd.setLayout(layout, size);
	Kmc::DDIterator<Kmc::DoubleTilingDD<2> > it(dd, 0);
	int n = 0;
	if(!it.setIndex(id))
		return 0;
	do {
		out[n++] = it.index();
	} while(it.next(T));
	return n;
}

/* the reference's run-time skip functions SLCGskipA/SLCGskipC */
void ref_slcg_skip_fn(int n, uint64_t* a, uint64_t* c)
{
	*a = SLCGskipA(n);
	*c = SLCGskipC(n);
}

/* hard-coded instantiations of inc/helper/slcgInst.h, in the order
 * 3840, 7680, 15360, 30720, 61440, 30721: a then c */
void ref_slcg_skip_table(uint64_t* out)
{
	out[0] = SLCGskipAt<3840>::a; out[1] = SLCGskipCt<3840>::c;
	out[2] = SLCGskipAt<7680>::a; out[3] = SLCGskipCt<7680>::c;
	out[4] = SLCGskipAt<15360>::a; out[5] = SLCGskipCt<15360>::c;
	out[6] = SLCGskipAt<30720>::a; out[7] = SLCGskipCt<30720>::c;
	out[8] = SLCGskipAt<61440>::a; out[9] = SLCGskipCt<61440>::c;
	out[10] = SLCGskipAt<30721>::a; out[11] = SLCGskipCt<30721>::c;
}

/* generic template metaprogram for small S: 1, 2, 3, 5, 8, 17, 64, 101 */
void ref_slcg_skip_tmpl_small(uint64_t* out)
{
	out[0] = SLCGskipAt<1>::a; out[1] = SLCGskipCt<1>::c;
	out[2] = SLCGskipAt<2>::a; out[3] = SLCGskipCt<2>::c;
	out[4] = SLCGskipAt<3>::a; out[5] = SLCGskipCt<3>::c;
	out[6] = SLCGskipAt<5>::a; out[7] = SLCGskipCt<5>::c;
	out[8] = SLCGskipAt<8>::a; out[9] = SLCGskipCt<8>::c;
	out[10] = SLCGskipAt<17>::a; out[11] = SLCGskipCt<17>::c;
	out[12] = SLCGskipAt<64>::a; out[13] = SLCGskipCt<64>::c;
	out[14] = SLCGskipAt<101>::a; out[15] = SLCGskipCt<101>::c;
}

uint64_t ref_slcg_gen(uint64_t s) { return SLCGen(s); }

/* tinymt32_init(seed) with the given parameters: the state after seeding,
 * then n tinymt32_generate_uint32 outputs; a second generator seeded alike
 * gives n tinymt32_generate_float outputs */
void ref_tmt(uint32_t mat1, uint32_t mat2, uint32_t tmat, uint32_t seed, int n,
		uint32_t* init_state, uint32_t* u32, float* f)
{
	tinymt32_t g;
	g.mat1 = mat1;
	g.mat2 = mat2;
	g.tmat = tmat;
	tinymt32_init(&g, seed);
	std::memcpy(init_state, g.status, 16);
	for(int a = 0; a < n; ++a)
		u32[a] = tinymt32_generate_uint32(&g);
	tinymt32_t h;
	h.mat1 = mat1;
	h.mat2 = mat2;
	h.tmat = tmat;
	tinymt32_init(&h, seed);
	for(int a = 0; a < n; ++a)
		f[a] = tinymt32_generate_float(&h);
}

/* Kpz::RngIface::TinyMT32 streams as a CPU scheduler holds them: T
 * generators with parameters (mat1[t], mat2[t], tmat[t]), randomize()d from
 * a master dSFMT seeded with init_gen_rand(seed) + `discard` draws; then per
 * generator n random() and n randomFloat() (alternating: random, randomFloat) */
void ref_tmt_iface(uint32_t seed, int discard, int T, const uint32_t* mat1, const uint32_t* mat2,
		const uint32_t* tmat, int n, uint32_t* u32, float* f)
{
	dsfmt_t d;
	dsfmt_init_gen_rand(&d, seed);
	for(int a = 0; a < discard; ++a)
		dsfmt_genrand_close_open(&d);
	std::vector<Kpz::RngIface::TinyMT32> rngs(T);
	for(int t = 0; t < T; ++t)
	{
		tinymt32_t* g = reinterpret_cast<tinymt32_t*>(&rngs[t]);
		g->mat1 = mat1[t];
		g->mat2 = mat2[t];
		g->tmat = tmat[t];
	}
	Kpz::RngIface::TinyMT32::randomize(rngs, d);
	for(int t = 0; t < T; ++t)
		for(int a = 0; a < n; ++a)
		{
			u32[t * n + a] = rngs[t].random();
			f[t * n + a] = rngs[t].randomFloat();
		}
}

}
