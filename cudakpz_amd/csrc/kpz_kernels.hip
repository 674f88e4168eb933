// gfx950 kernels of the KPZ octahedron-model engine.
//
// Design notes (MI355X / CDNA4, wave64):
//  * DT: one lane owns one double-tiling block for a whole phase.  The block
//    plus its +1 halo (right column of x slopes, top row of y slopes) is
//    staged in LDS as bit-plane rows {X row, Y row} (uint2 per site row),
//    lane-interleaved ([row][lane]) so every lane's random row access hits
//    its own bank pair: no LDS bank conflicts whatever the random site.
//    The SLCG state lives in registers; one update is 2 (3 for pq) SLCG
//    steps, two LDS reads, a 4-bit configuration test and two LDS writes.
//  * SCABit: a group of G <= 64 lanes owns one site row of one half-sweep;
//    lane g handles u64 words g, g+G, ...  HBM traffic is fully coalesced
//    8-byte-per-lane streams; the +-1 bit carry between neighbouring words
//    crosses lanes with wave shuffles.  Per-site random numbers come from
//    each worker's stream via precomputed affine jump-ahead, so any lane
//    order reproduces the CPU counterpart's sequential draw order.
//  * Reductions: one wave per site row (or word row) computes exact int64
//    sums relative to the row's start height from LDS walk tables (8 sites
//    per lookup); two small passes combine the rows in int128 on the device.
#include "kpz_kernels.hpp"

#include <mutex>

namespace kpz {

namespace {

// ---------------------------------------------------------------- bit utils
// bits 0,2,..,30 -> 0..15
__device__ __forceinline__ uint32_t even16(uint32_t v)
{
	v &= 0x55555555u;
	v = (v | (v >> 1)) & 0x33333333u;
	v = (v | (v >> 2)) & 0x0F0F0F0Fu;
	v = (v | (v >> 4)) & 0x00FF00FFu;
	v = (v | (v >> 8)) & 0x0000FFFFu;
	return v;
}

// 0..15 -> bits 0,2,..,30
__device__ __forceinline__ uint32_t spread16(uint32_t v)
{
	v &= 0xFFFFu;
	v = (v | (v << 8)) & 0x00FF00FFu;
	v = (v | (v << 4)) & 0x0F0F0F0Fu;
	v = (v | (v << 2)) & 0x33333333u;
	v = (v | (v << 1)) & 0x55555555u;
	return v;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src, int width)
{
	const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, width);
	const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, width);
	return ((uint64_t)hi << 32) | lo;
}

// The high state word of an SLCG64 step mod 2^24: h aLo + lo aHi + carry, as
// two v_mad_u32_u24 (a[23:0] * b[23:0] + c) in one asm block.  Left to itself
// the compiler widens products used mod 2^23 to v_mul_lo_u32 + v_add3_u32;
// one block instead of two halves the wait states it pads asm blocks with.
__device__ __forceinline__ uint32_t hi24_step(uint32_t h, uint32_t aLo, uint32_t lo, uint32_t aHi, uint32_t carry)
{
	uint32_t r;
	asm("v_mad_u32_u24 %0, %3, %4, %5\n\tv_mad_u32_u24 %0, %1, %2, %0"
	    : "=&v"(r) : "v"(h), "v"(aLo), "v"(lo), "v"(aHi), "v"(carry));
	return r;
}

__device__ __forceinline__ uint32_t shflw(uint32_t v, int src, int width) { return (uint32_t)__shfl((int)v, src, width); }

// Acceptance bits of 32 consecutive draws, draw k -> bit k.  The difference
// d = r - t of a draw's r < 2^23 (or 2^24) and a threshold t <= 2^23 (2^24)
// has bits 24..31 all equal to its sign, which is the acceptance r < t.  One
// v_bitop3 (a | (d & m)) copies that sign into any single bit m of those:
// draws 0-7, 8-15 and 16-23 go to bits 24-31 of three accumulators, shifted
// down once at the end, draws 24-31 straight to their own bits of a fourth.
// Three fast-class instructions per draw and threshold fewer slow ones than
// funnel-shifting the sign in (v_alignbit) and reversing the word.
struct SignBits32 {
	uint32_t a = 0, b = 0, c = 0, d = 0;
	template<int K>
	__device__ __forceinline__ void put(uint32_t diff)
	{
		static_assert(K >= 0 && K < 32, "draw index");
		constexpr uint32_t m = 1u << (24 + (K & 7));
		if constexpr(K < 8)
			a = __builtin_amdgcn_bitop3_b32(a, diff, m, 0xF8);
		else if constexpr(K < 16)
			b = __builtin_amdgcn_bitop3_b32(b, diff, m, 0xF8);
		else if constexpr(K < 24)
			c = __builtin_amdgcn_bitop3_b32(c, diff, m, 0xF8);
		else
			d = __builtin_amdgcn_bitop3_b32(d, diff, m, 0xF8);
	}
	__device__ __forceinline__ uint32_t bits() const { return (a >> 24) | (b >> 16) | (c >> 8) | d; }
};

// compile-time loop: f(integral_constant<int, K>) for K = 0..N-1
template<int N, int K = 0, class F>
__device__ __forceinline__ void static_for(F&& f)
{
	if constexpr(K < N)
	{
		f(std::integral_constant<int, K>{});
		static_for<N, K + 1>(f);
	}
}

// ======================================================================
// DT phase kernel
// ======================================================================
constexpr int DT_WG = 64;
constexpr int DT_LOG2_ROW_BYTES = 8; // a tile row: DT_WG lanes x 4 bytes
static_assert((1 << DT_LOG2_ROW_BYTES) == DT_WG * 4, "row stride");

// A lane's column of the DT LDS tile: the bit-plane rows in two planes,
// X rows 0..BY-1 and Y rows 0..BY (BY = the halo row above the block), each
// [row][lane] with a 4-byte entry, so that every access of the update loop
// (a row's X, its Y, the Y of the row above) is a conflict-free 4-byte access.
struct DtTile {
	uint32_t* x;
	uint32_t* y;
	uint32_t* base;   // start of the tile array (X plane row 0, lane 0)
	uint32_t lane;    // this lane's byte offset in a row (lane * 4 < DT_WG * 4)
	__device__ __forceinline__ uint32_t X(int r) const { return x[r * DT_WG]; }
	__device__ __forceinline__ uint32_t Y(int r) const { return y[r * DT_WG]; }
	__device__ __forceinline__ void setX(int r, uint32_t v) const { x[r * DT_WG] = v; }
	__device__ __forceinline__ void setY(int r, uint32_t v) const { y[r * DT_WG] = v; }
	// by byte offset from the tile start (row * DT_WG * 4 + lane; the Y plane
	// follows BY rows of X)
	__device__ __forceinline__ static uint32_t* at(uint32_t* p, uint32_t b)
	{
		return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(p) + b);
	}
	template<int BY> __device__ __forceinline__ uint32_t Xb(uint32_t b) const { return *at(base, b); }
	template<int BY> __device__ __forceinline__ uint32_t Yb(uint32_t b) const { return *at(base, BY * DT_WG * 4 + b); }
	template<int BY> __device__ __forceinline__ void setXb(uint32_t b, uint32_t v) const { *at(base, b) = v; }
	template<int BY> __device__ __forceinline__ void setYb(uint32_t b, uint32_t v) const
	{
		*at(base, BY * DT_WG * 4 + b) = v;
	}
};

// First three delta swaps of the outer unshuffle: a 32-bit row of 16
// interleaved sites (bit 2c = X_c, 2c+1 = Y_c) becomes the bytes
// {X 0-7, Y 0-7, X 8-15, Y 8-15}; the last (byte) swap is left to the v_perm
// that splits the planes.  zip_bytes is the inverse.
__device__ __forceinline__ uint32_t unzip_bytes(uint32_t x)
{
	uint32_t t;
	t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
	t = (x ^ (x >> 2)) & 0x0C0C0C0Cu; x ^= t ^ (t << 2);
	t = (x ^ (x >> 4)) & 0x00F000F0u; x ^= t ^ (t << 4);
	return x;
}
__device__ __forceinline__ uint32_t zip_bytes(uint32_t x)
{
	uint32_t t;
	t = (x ^ (x >> 4)) & 0x00F000F0u; x ^= t ^ (t << 4);
	t = (x ^ (x >> 2)) & 0x0C0C0C0Cu; x ^= t ^ (t << 2);
	t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
	return x;
}

// v_perm_b32 selector: byte i of the result = byte sel_i of {hi:lo}
// (0-3 lo, 4-7 hi, 12 = zero)
constexpr uint32_t perm_sel(int b0, int b1, int b2, int b3)
{
	return (uint32_t)b0 | ((uint32_t)b1 << 8) | ((uint32_t)b2 << 16) | ((uint32_t)b3 << 24);
}

// Static DT with 16-site-wide blocks (origin on the 4x4 word grid): every
// ENC_LOCAL byte is one site row of 4 sites, bits x0 y0 x1 y1 x2 y2 x3 y3.
// Stage: per word row one 16-byte load of the block's 4 words (+ the halo
// word); a site row is byte rr of the 4 words gathered by v_perm, then
// unzipped into its X (even) and Y (odd) bit planes.
template<int NWY>
__device__ __forceinline__ void dt_stage_x16(const uint32_t* __restrict__ lat, int XW, int mXW, int mYW, int c0,
                                             int wr0, uint32_t (&raw)[NWY][5], const DtTile& my)
{
#pragma unroll
	for(int wk = 0; wk < NWY; ++wk)
	{
		const uint32_t* rp = lat + (size_t)((wr0 + wk) & mYW) * XW;
		const uint4 q = *reinterpret_cast<const uint4*>(rp + c0);
		raw[wk][0] = q.x;
		raw[wk][1] = q.y;
		raw[wk][2] = q.z;
		raw[wk][3] = q.w;
		raw[wk][4] = (wk < NWY - 1) ? rp[(c0 + 4) & mXW] : 0u; // top halo row: corner unused
	}
#pragma unroll
	for(int wk = 0; wk < NWY; ++wk)
	{
#pragma unroll
		for(int rr = 0; rr < 4; ++rr)
		{
			if(wk == NWY - 1 && rr > 0)
				break;
			const uint32_t b = __builtin_amdgcn_perm(raw[wk][1], raw[wk][0], perm_sel(rr, 4 + rr, 12, 12))
			                 | __builtin_amdgcn_perm(raw[wk][3], raw[wk][2], perm_sel(12, 12, rr, 4 + rr));
			const uint32_t u = unzip_bytes(b);
			// X: 16 columns + the halo column (x = 0 of the next word: bit 0 of
			// its byte rr, moved to bit 16 with that byte); the top halo row
			// has Y only
			if(wk < NWY - 1)
				my.setX(4 * wk + rr, __builtin_amdgcn_perm(raw[wk][4], u, perm_sel(0, 2, 4 + rr, 12)) & 0x1FFFFu);
			my.setY(4 * wk + rr, __builtin_amdgcn_perm(u, u, perm_sel(1, 3, 12, 12)));
		}
	}
}

// Inverse of dt_stage_x16: zip each row back to its interleaved byte form,
// transpose 4 rows x 4 words of bytes with v_perm, one 16-byte store per
// word row; halo column (X of x = 16) and top halo row (Y of row BY) are
// merged into the words read at staging (no other lane of this phase
// writes them).
template<int NWY>
__device__ __forceinline__ void dt_writeback_x16(uint32_t* __restrict__ lat, int XW, int mXW, int mYW, int c0,
                                                 int wr0, const uint32_t (&raw)[NWY][5],
                                                 const DtTile& my)
{
#pragma unroll
	for(int wk = 0; wk < NWY - 1; ++wk)
	{
		uint32_t* rp = lat + (size_t)((wr0 + wk) & mYW) * XW;
		uint32_t B[4], hb = 0;
#pragma unroll
		for(int rr = 0; rr < 4; ++rr)
		{
			const uint32_t tx = my.X(4 * wk + rr), ty = my.Y(4 * wk + rr);
			B[rr] = zip_bytes(__builtin_amdgcn_perm(ty, tx, perm_sel(0, 4, 1, 5)));
			hb |= ((tx >> 16) & 1u) << (8 * rr);
		}
		const uint32_t T0 = __builtin_amdgcn_perm(B[1], B[0], perm_sel(0, 4, 2, 6));
		const uint32_t T1 = __builtin_amdgcn_perm(B[1], B[0], perm_sel(1, 5, 3, 7));
		const uint32_t T2 = __builtin_amdgcn_perm(B[3], B[2], perm_sel(0, 4, 2, 6));
		const uint32_t T3 = __builtin_amdgcn_perm(B[3], B[2], perm_sel(1, 5, 3, 7));
		uint4 q;
		q.x = __builtin_amdgcn_perm(T2, T0, perm_sel(0, 1, 4, 5));
		q.y = __builtin_amdgcn_perm(T3, T1, perm_sel(0, 1, 4, 5));
		q.z = __builtin_amdgcn_perm(T2, T0, perm_sel(2, 3, 6, 7));
		q.w = __builtin_amdgcn_perm(T3, T1, perm_sel(2, 3, 6, 7));
		*reinterpret_cast<uint4*>(rp + c0) = q;
		rp[(c0 + 4) & mXW] = (raw[wk][4] & ~0x01010101u) | hb;
	}
	// top halo row: only the Y bits of its first site row (byte 0, odd bits)
	constexpr int wk = NWY - 1;
	uint32_t* rp = lat + (size_t)((wr0 + wk) & mYW) * XW;
	const uint32_t b = zip_bytes(__builtin_amdgcn_perm(my.Y((NWY - 1) * 4), 0u, perm_sel(12, 4, 12, 5)));
	uint4 q;
	q.x = (raw[wk][0] & ~0xAAu) | (b & 0xAAu);
	q.y = (raw[wk][1] & ~0xAAu) | ((b >> 8) & 0xAAu);
	q.z = (raw[wk][2] & ~0xAAu) | ((b >> 16) & 0xAAu);
	q.w = (raw[wk][3] & ~0xAAu) | ((b >> 24) & 0xAAu);
	*reinterpret_cast<uint4*>(rp + c0) = q;
}

// The first two delta swaps of unzip_bytes: every byte x0 y0 x1 y1 x2 y2 x3 y3
// (one site row of a word's 4 sites) becomes x0 x1 x2 x3 y0 y1 y2 y3;
// join_nibbles is its inverse.
__device__ __forceinline__ uint32_t split_nibbles(uint32_t x)
{
	uint32_t t;
	t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
	t = (x ^ (x >> 2)) & 0x0C0C0C0Cu; x ^= t ^ (t << 2);
	return x;
}
__device__ __forceinline__ uint32_t join_nibbles(uint32_t x)
{
	uint32_t t;
	t = (x ^ (x >> 2)) & 0x0C0C0C0Cu; x ^= t ^ (t << 2);
	t = (x ^ (x >> 1)) & 0x22222222u; x ^= t ^ (t << 1);
	return x;
}

// The bits of the 5 x 5 words of a random-origin 16 x 16 block that the
// block owns, for each origin offset (shx, shy) inside its first word (entry
// 4 shy + shx): X of rows 0..15 at columns 0..16 (16 = the halo column), Y of
// rows 0..16 (16 = the halo row) at columns 0..15; bit 8 rr + 2 j (X) and
// + 1 (Y) of word k of word row wk is site (4 k + j - shx, 4 wk + rr - shy).
struct DtShiftMasks {
	uint32_t m[16][25];
};
constexpr DtShiftMasks make_dt_shift_masks()
{
	DtShiftMasks t{};
	for(int sh = 0; sh < 16; ++sh)
		for(int wk = 0; wk < 5; ++wk)
			for(int k = 0; k < 5; ++k)
			{
				uint32_t v = 0;
				for(int rr = 0; rr < 4; ++rr)
					for(int j = 0; j < 4; ++j)
					{
						const int r = 4 * wk + rr - (sh >> 2), x = 4 * k + j - (sh & 3);
						if(r >= 0 && r < 16 && x >= 0 && x <= 16)
							v |= 1u << (8 * rr + 2 * j);
						if(r >= 0 && r <= 16 && x >= 0 && x <= 15)
							v |= 1u << (8 * rr + 2 * j + 1);
					}
				t.m[sh][5 * wk + k] = v;
			}
	return t;
}
__constant__ DtShiftMasks kDtShiftMasks = make_dt_shift_masks();

// Random-origin DT (DTr) with 16 x 16-site blocks: the block starts shx sites
// into its first word and shy site rows into its first word row (uniform per
// launch), so it spans 5 x 5 words (20 x 20 sites).  Stage: per site row the
// bytes of words 0-3 are gathered and unzipped as in dt_stage_x16, word 4's
// byte (sites 16-19, split into nibbles once per word row) is appended, and
// the 20-site X and Y rows are shifted down by shx.
template<class Tile>
__device__ __forceinline__ void dt_stage_x16s(const uint32_t* __restrict__ lat, int XW, int mXW, int mYW, int c0,
                                              int wr0, int shx, int shy, uint32_t (&raw)[5][5], const Tile& my)
{
#pragma unroll
	for(int wk = 0; wk < 5; ++wk)
	{
		const uint32_t* rp = lat + (size_t)((wr0 + wk) & mYW) * XW;
#pragma unroll
		for(int k = 0; k < 5; ++k)
			raw[wk][k] = rp[(c0 + k) & mXW];
	}
#pragma unroll
	for(int wk = 0; wk < 5; ++wk)
	{
		const uint32_t u4 = split_nibbles(raw[wk][4]);
		const uint32_t u4y = u4 >> 4; // byte rr: the y nibble of u4's byte rr in bits 0-3
#pragma unroll
		for(int rr = 0; rr < 4; ++rr)
		{
			const int r = 4 * wk + rr - shy; // launch-uniform
			if(r < 0 || r > 16)
				continue;
			const uint32_t b = __builtin_amdgcn_perm(raw[wk][1], raw[wk][0], perm_sel(rr, 4 + rr, 12, 12))
			                 | __builtin_amdgcn_perm(raw[wk][3], raw[wk][2], perm_sel(12, 12, rr, 4 + rr));
			const uint32_t u = unzip_bytes(b);
			const uint32_t x20 = __builtin_amdgcn_perm(u4, u, perm_sel(0, 2, 4 + rr, 12)) & 0xFFFFFu;
			const uint32_t y20 = __builtin_amdgcn_perm(u4y, u, perm_sel(1, 3, 4 + rr, 12)) & 0xFFFFFu;
			if(r < 16)
				my.setX(r, (x20 >> shx) & 0x1FFFFu);
			my.setY(r, (y20 >> shx) & 0xFFFFu);
		}
	}
}

// Inverse of dt_stage_x16s: every site row is shifted up by shx and zipped
// back (sites 0-15 as in dt_writeback_x16, sites 16-19 as nibble pairs that
// are joined per word row), and each of the 25 words is merged into the word
// read at staging under the bits of that word the block owns
// (kDtShiftMasks).  No other lane of the phase writes these words.
template<class Tile>
__device__ __forceinline__ void dt_writeback_x16s(uint32_t* __restrict__ lat, int XW, int mXW, int mYW, int c0,
                                                  int wr0, int shx, int shy, const uint32_t (&raw)[5][5],
                                                  const Tile& my)
{
	const uint32_t* m = kDtShiftMasks.m[4 * shy + shx];
#pragma unroll
	for(int wk = 0; wk < 5; ++wk)
	{
		uint32_t B[4], Q = 0;
#pragma unroll
		for(int rr = 0; rr < 4; ++rr)
		{
			const int r = 4 * wk + rr - shy; // launch-uniform
			const uint32_t tx = (r >= 0 && r < 16) ? my.X(r) : 0u;
			const uint32_t ty = (r >= 0 && r <= 16) ? my.Y(r) : 0u;
			const uint32_t x20 = tx << shx, y20 = ty << shx;
			B[rr] = zip_bytes(__builtin_amdgcn_perm(y20, x20, perm_sel(0, 4, 1, 5)));
			Q |= (((x20 >> 16) & 0xFu) | ((y20 >> 12) & 0xF0u)) << (8 * rr);
		}
		const uint32_t T0 = __builtin_amdgcn_perm(B[1], B[0], perm_sel(0, 4, 2, 6));
		const uint32_t T1 = __builtin_amdgcn_perm(B[1], B[0], perm_sel(1, 5, 3, 7));
		const uint32_t T2 = __builtin_amdgcn_perm(B[3], B[2], perm_sel(0, 4, 2, 6));
		const uint32_t T3 = __builtin_amdgcn_perm(B[3], B[2], perm_sel(1, 5, 3, 7));
		const uint32_t v[5] = {__builtin_amdgcn_perm(T2, T0, perm_sel(0, 1, 4, 5)),
		                       __builtin_amdgcn_perm(T3, T1, perm_sel(0, 1, 4, 5)),
		                       __builtin_amdgcn_perm(T2, T0, perm_sel(2, 3, 6, 7)),
		                       __builtin_amdgcn_perm(T3, T1, perm_sel(2, 3, 6, 7)), join_nibbles(Q)};
		uint32_t* rp = lat + (size_t)((wr0 + wk) & mYW) * XW;
#pragma unroll
		for(int k = 0; k < 5; ++k)
			rp[(c0 + k) & mXW] = __builtin_amdgcn_bitop3_b32(raw[wk][k], v[k], m[5 * wk + k], 0xD8); // (raw & ~m) | (v & m)
	}
}

// ======================================================================
// TinyMT32 worker streams: the 127-bit generator of
// inc/tinyMT/tinymt32.h:70-107 (transition, tempering) and
// inc/tinyMT/tinymt32.c:59-75 (seeding).  On the device a stream set is 7
// planes of T words: status[0..3], mat1, mat2, tmat (the layout of
// TinyMT32::d_Random, kpz/GPUTinyMT.cu:38-61, and of its `.tinyMT` dataset).
// ======================================================================
// v + v as v_add_u32: left shifts issue at half the rate of adds on gfx950,
// and the compiler turns v + v written in C into a shift
__device__ __forceinline__ uint32_t twice(uint32_t v)
{
	uint32_t r;
	asm("v_add_u32 %0, %1, %1" : "=v"(r) : "v"(v));
	return r;
}

struct TinyMT {
	uint32_t s0, s1, s2, s3, m1, m2, tm;

	// the transition with three-input logic ops: 0x6A is (a & b) ^ c, 0x96 a ^ b ^ c
	__device__ __forceinline__ void next()
	{
		uint32_t x = __builtin_amdgcn_bitop3_b32(s0, 0x7FFFFFFFu, s1, 0x6A) ^ s2;
		x ^= twice(x);
		const uint32_t y = __builtin_amdgcn_bitop3_b32(s3, s3 >> 1, x, 0x96);
		const uint32_t m = 0u - (y & 1u);
		s0 = s1;
		s1 = __builtin_amdgcn_bitop3_b32(m, m1, s2, 0x6A);
		s2 = __builtin_amdgcn_bitop3_b32(m, m2, x, 0x6A) ^ (y << 10);
		s3 = y;
	}
	__device__ __forceinline__ uint32_t temper() const
	{
		const uint32_t t1 = s0 + (s2 >> 8);
		return __builtin_amdgcn_bitop3_b32(0u - (t1 & 1u), tm, s3, 0x6A) ^ t1;
	}
	__device__ __forceinline__ uint32_t u32()
	{
		next();
		return temper();
	}
	__device__ __forceinline__ static TinyMT load(const uint32_t* __restrict__ g, int64_t T, int64_t t)
	{
		TinyMT r;
		r.s0 = g[t];
		r.s1 = g[T + t];
		r.s2 = g[2 * T + t];
		r.s3 = g[3 * T + t];
		r.m1 = g[4 * T + t];
		r.m2 = g[5 * T + t];
		r.tm = g[6 * T + t];
		return r;
	}
	__device__ __forceinline__ void store(uint32_t* __restrict__ g, int64_t T, int64_t t) const
	{
		g[t] = s0;
		g[T + t] = s1;
		g[2 * T + t] = s2;
		g[3 * T + t] = s3;
	}
};

// tinymt32_init with the seed in plane 0 (TinyMT32::randomize + seedGPU,
// kpz/GPUTinyMT.cu:63-98; the CPU streams are seeded alike,
// kpz/CPURngIface.h:99-103)
__global__ __launch_bounds__(256) void tmt_seed_kernel(uint32_t* __restrict__ g, int64_t T)
{
	const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if(t >= T)
		return;
	TinyMT r = TinyMT::load(g, T, t);
	uint32_t st[4] = {r.s0, r.m1, r.m2, r.tm};
#pragma unroll
	for(int i = 1; i < 8; ++i)
	{
		const uint32_t p = st[(i - 1) & 3];
		st[i & 3] ^= (uint32_t)i + 1812433253u * (p ^ (p >> 30));
	}
	if((st[0] & 0x7FFFFFFFu) == 0 && st[1] == 0 && st[2] == 0 && st[3] == 0)
	{
		st[0] = 'T';
		st[1] = 'I';
		st[2] = 'N';
		st[3] = 'Y';
	}
	r.s0 = st[0];
	r.s1 = st[1];
	r.s2 = st[2];
	r.s3 = st[3];
#pragma unroll
	for(int i = 0; i < 8; ++i)
		r.next();
	r.store(g, T, t);
}

// Stream cursors of a DT block.  TOP: the coordinates are the draw's top
// bits; rf(): the randomFloat() numerator compared with the thresholds.
//
// SLCG64: block j of stream t starts j blocks of draws in (jump table).
// DTr uses every draw as randomFloat() = (lo ^ hi) & 0x7FFFFF: the low state
// word evolves on its own and the high word is needed mod 2^24 only (one
// v_mad_u64_u32 and two v_mad_u32_u24 per draw); the full state after the
// block is one jump from its start.  The static DT's coordinates take the top
// bits of lo ^ hi and need the whole state.
template<bool RO>
struct DtSlcg {
	static constexpr bool TOP = !RO;
	__device__ __forceinline__ static uint32_t rf(uint32_t u) { return u & 0x7FFFFFu; }
	uint64_t s, sBlock;
	uint32_t lo, h24, aLo, aHi;
	Affine step;
	int64_t t;
	bool last;

	__device__ __forceinline__ void begin(const DtPhaseArgs& a, int64_t i, int64_t nb)
	{
		const bool own = a.T >= nb;
		t = own ? i : i % a.T;
		const int64_t j = own ? 0 : i / a.T;
		last = own || i + a.T >= nb;
		s = a.st_in[t];
		if(!own)
			s = a.jumps[j](s);
		step = a.step;
		sBlock = s;
		lo = (uint32_t)s;
		h24 = (uint32_t)(s >> 32);
		aLo = (uint32_t)step.a;
		aHi = (uint32_t)(step.a >> 32);
	}
	__device__ __forceinline__ uint32_t draw()
	{
		if constexpr(RO)
		{
			const uint64_t pr = (uint64_t)lo * aLo + step.c;
			h24 = hi24_step(h24, aLo, lo, aHi, (uint32_t)(pr >> 32));
			lo = (uint32_t)pr;
			return lo ^ h24; // bits 0..22 are those of the full fold
		}
		else
		{
			s = step(s);
			return slcg_fold(s);
		}
	}
	__device__ __forceinline__ void end(const DtPhaseArgs& a)
	{
		if(last)
			a.st_out[t] = RO ? a.blockJ(sBlock) : s;
	}
};

// TinyMT32 (SchedulerCPUDTr<TinyMT32>, kpz/kpzFrontend.cpp:308-310):
// randomFloat() = (u >> 8) * 2^-24, so (int)(randomFloat() * 2^b) = u >> (32-b);
// the lane holds its stream across its blocks
struct DtTmt {
	static constexpr bool TOP = true;
	__device__ __forceinline__ static uint32_t rf(uint32_t u) { return u >> 8; }
	TinyMT g;
	__device__ __forceinline__ void begin(const DtPhaseArgs&, int64_t, int64_t) {}
	__device__ __forceinline__ uint32_t draw() { return g.u32(); }
	__device__ __forceinline__ void end(const DtPhaseArgs&) {}
};

// One DT block i of one phase, staged in this lane's LDS column `my`, with
// the stream cursor `cur`.  nb = blocks per domain set.  UPD: 0 = KPZ
// (pqOneUpdate), 1 = pqSyncRngUpdate, 2 = disorder2SyncRngUpdate (class rows
// of the disorder map in myD, 2 bits per site), 3 = EW.
template<int LBX, int LBY, int UPD, bool RO, class Cur>
__device__ __forceinline__ void dt_block(const DtPhaseArgs& a, int64_t i, int64_t nb, const DtTile& my,
                                         uint32_t* __restrict__ myD, Cur& cur)
{
	constexpr int BX = 1 << LBX, BY = 1 << LBY;
	constexpr int NWX = BX / 4 + 1; // words covering BX+1 sites (+ shift)
	constexpr int NWY = BY / 4 + 1; // word rows covering BY+1 site rows
	constexpr uint32_t MX = (2u << BX) - 1u; // X row: columns 0..BX (BX = halo)
	constexpr uint32_t MY = (1u << BX) - 1u; // Y row: columns 0..BX-1

	// block geometry, Kmc::DoubleTilingDD<2>::setBlock (inc/doubleTilingDD.h:71-78)
	const int lhx = a.lx - LBX - 1;
	const int64_t bxi = i & ((1ll << lhx) - 1), byi = i >> lhx;
	const int ox = (int)(((2 * bxi + (a.ds & 1)) << LBX) + a.sx);
	const int oy = (int)(((2 * byi + (a.ds >> 1)) << LBY) + a.sy);
	const int XW = 1 << (a.lx - 2);
	const int mXW = XW - 1, mYW = (1 << (a.ly - 2)) - 1;
	const int shx = RO ? (ox & 3) : 0;
	const int shy = RO ? (oy & 3) : 0;
	const int c0 = ox >> 2, wr0 = oy >> 2;
	uint32_t* __restrict__ lat = a.lat;

	// ---- stage ENC_LOCAL words -> bit-plane rows.  The raw words stay in
	// registers: the write-back merges into them instead of re-reading HBM
	// (no other lane of this phase writes them).
	// Static 16-wide blocks take the byte-gather path; shifted origins (DTr)
	// and narrower blocks the general nibble path below.
	constexpr bool X16 = !RO && LBX == 4;
	constexpr bool X16S = RO && LBX == 4 && LBY == 4; // DTr, 16 x 16 blocks
	uint32_t raw[NWY][NWX];
	if constexpr(X16)
		dt_stage_x16<NWY>(lat, XW, mXW, mYW, c0, wr0, raw, my);
	else if constexpr(X16S)
		dt_stage_x16s(lat, XW, mXW, mYW, c0, wr0, shx, shy, raw, my);
	else
	{
#pragma unroll
	for(int wk = 0; wk < NWY; ++wk)
	{
		const size_t rowBase = (size_t)((wr0 + wk) & mYW) * XW;
#pragma unroll
		for(int k = 0; k < NWX; ++k)
		{
			const bool needed = RO || wk < NWY - 1 || k < NWX - 1; // static DT: corner unused
			raw[wk][k] = needed ? lat[rowBase + ((c0 + k) & mXW)] : 0u;
		}
	}
#pragma unroll
	for(int wk = 0; wk < NWY; ++wk)
	{
		uint32_t cx[NWX], cy[NWX];
#pragma unroll
		for(int k = 0; k < NWX; ++k)
		{
			cx[k] = even16(raw[wk][k]);
			cy[k] = even16(raw[wk][k] >> 1);
		}
#pragma unroll
		for(int rr = 0; rr < 4; ++rr)
		{
			const int r = 4 * wk + rr - shy;
			if(r < 0 || r > BY)
				continue;
			uint32_t sxv = 0, syv = 0;
#pragma unroll
			for(int k = 0; k < NWX; ++k)
			{
				sxv |= ((cx[k] >> (4 * rr)) & 0xFu) << (4 * k);
				syv |= ((cy[k] >> (4 * rr)) & 0xFu) << (4 * k);
			}
			if(r < BY)
				my.setX(r, (sxv >> shx) & MX);
			my.setY(r, (syv >> shx) & MY);
		}
	}
	}
	if constexpr(UPD == 2)
	{
		// class rows of the disorder map (c(m_disorder): the site's 2 bits,
		// kpz/schedulerCPUBaseThreadedRng.inl:50-75): byte rr of a word is
		// one site row, 2 bits per site, so a row is 2*BX bits
		constexpr uint32_t DM = (uint32_t)((1ull << (2 * BX)) - 1u);
#pragma unroll
		for(int wk = 0; wk < NWY; ++wk)
		{
			if(!RO && wk == NWY - 1)
				continue; // static DT: the top halo row needs no classes
			const size_t rowBase = (size_t)((wr0 + wk) & mYW) * XW;
			uint32_t dw[NWX];
#pragma unroll
			for(int k = 0; k < NWX; ++k)
				dw[k] = (RO || k < NWX - 1) ? a.dis[rowBase + ((c0 + k) & mXW)] : 0u;
#pragma unroll
			for(int rr = 0; rr < 4; ++rr)
			{
				const int r = 4 * wk + rr - shy;
				if(r < 0 || r >= BY)
					continue;
				uint64_t dv = 0;
#pragma unroll
				for(int k = 0; k < NWX; ++k)
					dv |= (uint64_t)((dw[k] >> (8 * rr)) & 0xFFu) << (8 * k);
				myD[r * DT_WG] = (uint32_t)(dv >> (2 * shx)) & DM;
			}
		}
	}

	// ---- the block's stream
	cur.begin(a, i, nb);
	const uint32_t tp = a.tp, tq = a.tq;

	// ---- BX*BY random-sequential updates (Func_MCSBase::exec,
	// kpz/schedulerCPUDT_nested.h:46-66; pqOneUpdate kpz/schedulerCPU.inl:21-33)
#pragma unroll 2
	for(int k = 0; k < BX * BY; ++k)
	{
		// x comes as xr = (BX - 1) - x, the same bits of ~r1 (one v_xnor_b32 with
		// the fold), so that the column mask 1 << x is a right shift: right
		// shifts issue at twice the rate of left shifts
		const uint32_t n1 = ~cur.draw();
		const uint32_t r2 = cur.draw();
		// the row's LDS byte address comes straight from the draw's bits: a right
		// shift, then one v_bitop3 for (bits & row mask) | lane offset
		constexpr int LW = DT_LOG2_ROW_BYTES; // log2 of the row stride in bytes
		uint32_t xr, ys;
		if constexpr(!Cur::TOP)
		{
			uint32_t t = n1 >> (23 - LBX);
			asm("" : "+v"(t)); // a shift and an and, not one v_bfe_u32
			xr = t & (BX - 1);
			ys = r2 >> (23 - LBY - LW);
		}
		else
		{
			xr = n1 >> (32 - LBX);
			ys = r2 >> (32 - LBY - LW);
		}
		const uint32_t yo = __builtin_amdgcn_bitop3_b32(ys, (uint32_t)((BY - 1) << LW), my.lane, 0xEA);
		const uint32_t cm = (1u << (BX - 1)) >> xr; // 1 << x
		uint2 row = make_uint2(my.template Xb<BY>(yo), my.template Yb<BY>(yo));
		const uint32_t up = my.template Yb<BY>(yo + DT_WG * 4);
		// whole-row bit algebra, then select column x: a local minimum has
		// own X = 0, right X = 1, own Y = 0, up Y = 1; a maximum the opposite
		const uint32_t rx = row.x >> 1;
		if(UPD == 0)
		{
			// at a minimum own X = 0 and right X = 1, own Y = 0 and up Y = 1:
			// the flip is X - m (bit x set, bit x+1 cleared), Y | m, up - m
			const uint32_t fm = rx & ~row.x & ~row.y & up & cm;
			my.template setXb<BY>(yo, row.x - fm);
			my.template setYb<BY>(yo, row.y | fm);
			my.template setYb<BY>(yo + DT_WG * 4, up - fm);
		}
		else
		{
			// all-ones when accepted: sign of r23 - t (both at most 2^24); the EW
			// rule (SyncRngEW, kpz/kpzSimulationPrimitivesTemplates.cu:199-234:
			// p = q = 1) accepts every extremum without a draw
			uint32_t aMin = ~0u, aMax = ~0u;
			if(UPD != 3)
			{
				const uint32_t r23 = Cur::rf(cur.draw());
				uint32_t tpv = tp, tqv = tq;
				if(UPD == 2)
				{
					const uint32_t x = (BX - 1) - xr;
					const uint32_t cls = (*DtTile::at(myD, yo & ((BY - 1) << LW)) >> (2 * x)) & 3u;
					tpv = (cls & 2) ? ((cls & 1) ? a.tpc[3] : a.tpc[2]) : ((cls & 1) ? a.tpc[1] : a.tpc[0]);
					tqv = (cls & 2) ? ((cls & 1) ? a.tqc[3] : a.tqc[2]) : ((cls & 1) ? a.tqc[1] : a.tqc[0]);
				}
				aMin = (uint32_t)((int32_t)(r23 - tpv) >> 31);
				aMax = (uint32_t)((int32_t)(r23 - tqv) >> 31);
			}
			const uint32_t mn = rx & ~row.x & ~row.y & up;
			const uint32_t mx = ~rx & row.x & row.y & ~up;
			// a minimum flips as X - m (bit x set, bit x+1 cleared), a maximum
			// as X + m (the opposite); Y and the Y above toggle either way
			const uint32_t fn = mn & aMin & cm, fx = mx & aMax & cm;
			my.template setXb<BY>(yo, row.x - fn + fx);
			my.template setYb<BY>(yo, row.y ^ fn ^ fx);
			my.template setYb<BY>(yo + DT_WG * 4, up ^ fn ^ fx);
		}
	}

	cur.end(a);

	// ---- bit-plane rows -> ENC_LOCAL words (masked read-modify-write at the
	// block edges; no other lane of this phase touches these words)
	if constexpr(X16)
	{
		dt_writeback_x16<NWY>(lat, XW, mXW, mYW, c0, wr0, raw, my);
		return;
	}
	if constexpr(X16S)
	{
		dt_writeback_x16s(lat, XW, mXW, mYW, c0, wr0, shx, shy, raw, my);
		return;
	}
	constexpr uint32_t SEGX = MX, SEGY = MY;
#pragma unroll
	for(int wk = 0; wk < NWY; ++wk)
	{
		const size_t rowBase = (size_t)((wr0 + wk) & mYW) * XW;
#pragma unroll
		for(int k = 0; k < NWX; ++k)
		{
			const uint32_t nibMX = ((SEGX << shx) >> (4 * k)) & 0xFu;
			const uint32_t nibMY = ((SEGY << shx) >> (4 * k)) & 0xFu;
			uint32_t mx16 = 0, my16 = 0, vx16 = 0, vy16 = 0;
#pragma unroll
			for(int rr = 0; rr < 4; ++rr)
			{
				const int r = 4 * wk + rr - shy;
				if(r < 0 || r > BY)
					continue;
				const uint32_t tx = (r < BY) ? my.X(r) : 0u, ty = my.Y(r);
				const uint32_t mX = (r < BY) ? nibMX : 0u;
				vx16 |= (((tx << shx) >> (4 * k)) & mX) << (4 * rr);
				mx16 |= mX << (4 * rr);
				vy16 |= (((ty << shx) >> (4 * k)) & nibMY) << (4 * rr);
				my16 |= nibMY << (4 * rr);
			}
			if((mx16 | my16) == 0u)
				continue;
			const uint32_t m32 = spread16(mx16) | (spread16(my16) << 1);
			const uint32_t v32 = spread16(vx16) | (spread16(vy16) << 1);
			lat[rowBase + ((c0 + k) & mXW)] = (raw[wk][k] & ~m32) | v32;
		}
	}
}

template<int LBX, int LBY, int UPD, bool RO>
__global__ __launch_bounds__(DT_WG) void dt_phase_kernel(DtPhaseArgs a, int64_t nb)
{
	// X plane then Y plane in one array: both planes share one address
	// register, the plane offset goes into the LDS instructions' offset
	__shared__ uint32_t tile[(2 * (1 << LBY) + 1) * DT_WG];
	__shared__ uint32_t tileD[(UPD == 2) ? (1 << LBY) * DT_WG : 1];
	const int64_t i = (int64_t)blockIdx.x * DT_WG + threadIdx.x;
	if(i >= nb)
		return; // no barrier in this kernel: idle lanes may leave
	DtSlcg<RO> cur;
	dt_block<LBX, LBY, UPD, RO>(a, i, nb,
	                            DtTile{tile + threadIdx.x, tile + (1 << LBY) * DT_WG + threadIdx.x, tile, threadIdx.x * 4},
	                            (UPD == 2) ? tileD + threadIdx.x : nullptr, cur);
}

// SchedulerCPUDTr<TinyMT32> (SchedulerCPUDTr::Func_MCS::exec,
// kpz/schedulerCPUDTr.cpp:147-175: worker t updates blocks t, t+T, ... of the
// domain set in order with its own stream).  With T >= blocks every worker
// has one block at most and the lane of block i works in place on stream i.
// Otherwise lane c runs a chain of C = a.tmtChain consecutive blocks of worker
// t = c % T, blocks t + T (k C + m) with k = c / T, m < C, in order with one
// stream jumped to the chain's start (a.tmtJ[c] = x^(k C draws per block)
// mod the orbit polynomial): a jump costs 128 recurrence steps, so long
// walks share it over C blocks; the lane of a worker's last block writes the
// worker's state to a.tmtOut.
template<int LBX, int LBY, int UPD>
__global__ __launch_bounds__(DT_WG) void dt_tmt_kernel(DtPhaseArgs a, int64_t nb, uint32_t* __restrict__ g)
{
	__shared__ uint32_t tile[(2 * (1 << LBY) + 1) * DT_WG];
	__shared__ uint32_t tileD[(UPD == 2) ? (1 << LBY) * DT_WG : 1];
	const int64_t c = (int64_t)blockIdx.x * DT_WG + threadIdx.x;
	const DtTile my{tile + threadIdx.x, tile + (1 << LBY) * DT_WG + threadIdx.x, tile, threadIdx.x * 4};
	uint32_t* const myD = (UPD == 2) ? tileD + threadIdx.x : nullptr;
	DtTmt cur;
	if(a.T >= nb)
	{
		if(c >= nb)
			return; // no barrier in this kernel; workers without a block keep their stream
		cur.g = TinyMT::load(g, a.T, c);
		dt_block<LBX, LBY, UPD, true>(a, c, nb, my, myD, cur);
		cur.g.store(g, a.T, c);
		return;
	}
	const int C = a.tmtChain;
	const int64_t t = c % a.T, k = c / a.T;
	int64_t i = t + a.T * (k * C);
	if(i >= nb)
		return; // the chain starts past the worker's last block
	cur.g = TinyMT::load(g, a.T, t);
	if(k > 0)
	{
		uint32_t st[4] = {cur.g.s0, cur.g.s1, cur.g.s2, cur.g.s3};
		tmt_jump4(st, a.tmtJ[c], a.tmtOrb[t], cur.g.m1, cur.g.m2);
		cur.g.s0 = st[0];
		cur.g.s1 = st[1];
		cur.g.s2 = st[2];
		cur.g.s3 = st[3];
	}
	for(int m = 0; m < C && i < nb; ++m, i += a.T)
	{
		dt_block<LBX, LBY, UPD, true>(a, i, nb, my, myD, cur);
		if(i + a.T >= nb)
			cur.g.store(a.tmtOut, a.T, t);
	}
}

// tmtJ for dt_tmt_kernel: lane c gets x^((c / T) * drawsPerEntry) mod the
// orbit polynomial of worker c % T (drawsPerEntry: the draws of a chain)
__global__ __launch_bounds__(256) void tmt_block_polys_kernel(const TmtOrbit* __restrict__ orb, int64_t T, int64_t n,
		uint64_t drawsPerEntry, Poly128* __restrict__ out)
{
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if(i >= n)
		return;
	out[i] = poly_xpow((uint64_t)(i / T) * drawsPerEntry, orb[i % T]);
}

// ======================================================================
// SCABit half-sweep (ENC_CBBIT)
// ======================================================================
// Site row y of one half-sweep, lane g of the row's group of G = 1 << a.lG
// lanes (groups are aligned to G within the wave; every lane of the wave
// must call this, the shuffles are wave-wide).
// Per-site acceptance bits of one SCABit word: the WB draws of the word from
// stream state sW, site k -> bit k of pb (minimum accepted: r23 < tp) and qb
// (maximum accepted: r23 < tq); NT = thresholds per draw: 1 (q = 0), 2 (p and
// q), 4 (two disorder classes: pb1, qb1 from tp1, tq1).  Only the low 23 bits
// of lo ^ hi are used: the low state word evolves on its own (lo' = A_lo lo +
// C mod 2^32) and hi' mod 2^24 needs only hi mod 2^24 and lo, so inside the
// word the high half is carried mod 2^24 with two 24-bit multiply-adds (the
// caller rebuilds the full state from the word's start by a jump).  The sign
// of r23 - t of every draw is copied into its own bit (SignBits32).
template<typename W, int NT>
__device__ __forceinline__ void cbbit_word_draws(uint64_t sW, const Affine& step, uint32_t tp, uint32_t tq,
                                                 uint32_t tp1, uint32_t tq1, W& pb, W& qb, W& pb1, W& qb1)
{
	const uint32_t aLo = (uint32_t)step.a, aHi = (uint32_t)(step.a >> 32);
	const uint64_t cc = step.c;
	uint32_t lo = (uint32_t)sW, h24 = (uint32_t)(sW >> 32);
	static_for<(int)sizeof(W) / 4>([&](auto hc) {
		constexpr int h = decltype(hc)::value;
		SignBits32 P, Q, P1, Q1;
		static_for<32>([&](auto kc) {
			constexpr int k = decltype(kc)::value;
			const uint64_t pr = (uint64_t)lo * aLo + cc;
			h24 = hi24_step(h24, aLo, lo, aHi, (uint32_t)(pr >> 32));
			lo = (uint32_t)pr;
			const uint32_t r23 = (lo ^ h24) & 0x7FFFFFu;
			P.template put<k>(r23 - tp);
			if constexpr(NT >= 2)
				Q.template put<k>(r23 - tq);
			if constexpr(NT == 4)
			{
				P1.template put<k>(r23 - tp1);
				Q1.template put<k>(r23 - tq1);
			}
		});
		pb |= (W)P.bits() << (32 * h);
		qb |= (W)Q.bits() << (32 * h);
		if constexpr(NT == 4)
		{
			pb1 |= (W)P1.bits() << (32 * h);
			qb1 |= (W)Q1.bits() << (32 * h);
		}
	});
}

template<typename W, int MODE>
__device__ __forceinline__ void cbbit_row(const CbHalfArgs& a, int y, int g)
{
	constexpr int WB = sizeof(W) * 8;
	constexpr int LW = (WB == 64) ? 7 : 6; // log2 sites per word (two sublattices)
	const int lG = a.lG, G = 1 << lG;
	const int Y = 1 << a.ly;
	W* __restrict__ lat = (W*)a.lat;
	const int Wn = 1 << (a.lx - LW); // words per row and quadrant
	const int M = Wn >> lG;
	const int64_t Q = (int64_t)1 << (a.lx + a.ly - LW);

	// worker t owns rows [yMin_t, ySup_t) (kpz/schedulerCPUCBBit.cpp:46-62)
	const int big = a.rem * (a.per + 1);
	int t, r;
	if(y < big)
	{
		t = y / (a.per + 1);
		r = y - t * (a.per + 1);
	}
	else
	{
		const int yy = y - big;
		t = a.rem + yy / a.per;
		r = yy - (t - a.rem) * a.per;
	}
	uint64_t s = 0;
	if constexpr(MODE < 3)
	{
		const uint64_t s0 = a.st_in[t];
		s = a.laneJ[g](a.rowJ[r](s0));
		if(r == 0 && g == 0)
			a.st_out[t] = (t < a.rem ? a.fullBig : a.fullSmall)(s0);
	}

	const bool par = (a.signum == (y & 1));
	const int64_t thisB = a.signum ? 2 * Q : 0, othB = a.signum ? 0 : 2 * Q;
	const int64_t row = (int64_t)y * Wn, rowUp = (int64_t)((y + 1) & (Y - 1)) * Wn;
	const Affine step = a.step, stepJ = a.stepJ;
	const uint32_t tp = a.tp, tq = a.tq;
	const int gUp = (g + 1) & (G - 1), gDn = (g - 1) & (G - 1);

	W pending = 0, x0old = 0, f0 = 0;
	for(int m = 0; m < M; ++m)
	{
		const uint64_t sW = s; // stream state at this word's first draw
		const int xw = g + (m << lG);
		const W tX = lat[thisB + row + xw];
		const W tY = lat[thisB + Q + row + xw];
		const W xX = lat[othB + row + xw];
		const W yY = lat[othB + Q + rowUp + xw];
		// the last lane of a group takes its LSB source from memory (below);
		// the other lanes load their own xX word again (same cache lines, no
		// extra HBM traffic) so that the load needs no divergent branch and
		// stays in flight behind the draws.  Only bit 0 of the source is used,
		// so only its low 32 bits are loaded.
		uint32_t nbrMem = 0;
		if(!par) // uniform: the row's parity
			nbrMem = reinterpret_cast<const uint32_t*>(lat + othB + row)[(size_t)((g == G - 1) ? ((xw + 1) & (Wn - 1)) : xw) *
			                                                             (sizeof(W) / 4)];
		W D = 0; // disorder classes of the word's sites (mode 2)
		if constexpr(MODE == 2)
			D = ((const W*)a.dis)[thisB + row + xw];
		// per-site p/q mode: the word's acceptance bits depend on the stream
		// alone, so they are drawn before the loaded words are first used and
		// the loads complete behind the draws
		W pb = 0, qb = 0, pb1 = 0, qb1 = 0;
		if constexpr(MODE == 1)
		{
			if(tq)
				cbbit_word_draws<W, 2>(sW, step, tp, tq, 0, 0, pb, qb, pb1, qb1);
			else
				cbbit_word_draws<W, 1>(sW, step, tp, tq, 0, 0, pb, qb, pb1, qb1);
		}
		// LSB source for the shifted neighbour slope (pSyncRngUpdateVL3,
		// kpz/schedulerCPUCBBit_nested.inl:226-233): lane g+1, or for the last
		// lane the first word of the next stride (wrapping to word 0)
		uint32_t nbr = shflw((uint32_t)xX, gUp, G);
		if(g == G - 1)
			nbr = nbrMem;
		const W xXs = par ? xX : (W)((xX >> 1) | ((W)nbr << (WB - 1)));
		const W minm = ~(tX | tY) & xXs & yY;
		W flip;
		if constexpr(MODE == 3)
			flip = minm & ((const W*)a.rnd)[row + xw]; // VL3 random words of this half-sweep
		else if constexpr(MODE == 4)
		{
			// per-site acceptance masks of this half-sweep (minima, maxima)
			const W maxm = tX & tY & ~xXs & ~yY;
			flip = (minm & ((const W*)a.rnd)[row + xw]) | (maxm & ((const W*)a.rnd)[Q + row + xw]);
		}
		else if(MODE == 0)
		{
			W rnd;
			uint64_t u = sW;
			if(WB == 64)
			{
				u = step(u);
				const uint64_t hi = slcg_fold(u);
				u = step(u);
				const uint64_t lo = slcg_fold(u);
				rnd = (W)((hi << 32) ^ lo);
			}
			else
			{
				u = step(u);
				rnd = (W)slcg_fold(u);
			}
			flip = minm & rnd;
		}
		else
		{
			const W maxm = tX & tY & ~xXs & ~yY;
			if constexpr(MODE == 2)
			{
				// the disorder mode draws here, after the lattice logic: its four
				// accumulators then do not share the live range of the loads
				cbbit_word_draws<W, 4>(sW, step, tp, tq, a.tp1, a.tq1, pb, qb, pb1, qb1);
				pb = (pb & ~D) | (pb1 & D);
				qb = (qb & ~D) | (qb1 & D);
			}
			flip = (minm & pb) | (maxm & qb);
		}
		if constexpr(MODE < 3)
			s = stepJ(sW);
		lat[thisB + row + xw] = tX ^ flip;
		lat[thisB + Q + row + xw] = tY ^ flip;
		lat[othB + Q + rowUp + xw] = yY ^ flip;
		// carry of the shifted flip into the next word's LSB
		const W recv = shflw((uint32_t)(flip >> (WB - 1)), gDn, G);
		W cin;
		if(g == 0)
		{
			cin = pending;
			pending = recv;
		}
		else
			cin = recv;
		if(par)
			lat[othB + row + xw] = xX ^ flip;
		else if(g == 0 && m == 0)
		{
			x0old = xX;
			f0 = flip;
		}
		else
			lat[othB + row + xw] = xX ^ (W)(flip << 1) ^ cin;
	}
	if(!par && g == 0)
		lat[othB + row] = x0old ^ (W)(f0 << 1) ^ pending;
}

template<typename W, int MODE>
__global__ __launch_bounds__(256) void cbbit_half_kernel(CbHalfArgs a)
{
	const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	cbbit_row<W, MODE>(a, (int)(gid >> a.lG), (int)(gid & ((1 << a.lG) - 1)));
}

// ======================================================================
// SCA on ENC_LOCAL (SchedulerCPUCB, kpz/schedulerCPUCB.cpp:43-170)
// ======================================================================
// In a checkerboard half-sweep every bit an active site reads (own slopes,
// x slope of x+1, y slope of y+1) is written only by that site, so the
// result does not depend on the order of the sites, and a site's flips of
// its right and up neighbours' bits can be applied to their words at any
// time before the half-sweep ends.
//
// One lane per 4x4-site cell column of a band of 2^lNB cell rows; a group of
// G = 2^lG lanes (G <= 64, aligned in the wave) covers G adjacent columns.
// Per cell row the lane loads its own word (kept from the previous row's
// "up" load), the right word and the up word once, tests all 8 active sites
// bit-parallel and stores its word once: own flips, the left lane's flips of
// column 0 (a shuffle) and the flips of row 0 from the cell below (kept in a
// register) are merged into one plain store.  Flips that leave a group
// (right edge) into a group of the same workgroup are handed over in LDS
// before that group stores the row (one barrier per row); the others, and
// those that leave a band (top row), go to small edge buffers, which
// cb_local_fix_kernel applies afterwards with one thread per target word.
__device__ __forceinline__ int cb_class(const CbLocalArgs& a, int v)
{
	int k = 0;
	for(int i = 1; i < a.nV; ++i)
		if(a.V[i] == v)
			k = i;
	return k;
}

// active sites of a cell in draw order (yy = 0..3, x = x0, x0 + 2 with
// x0 = (yy & 1) ^ signum), X-bit positions 2x + 8yy:
//   signum 0: 0 4 10 14 16 20 26 30      signum 1: 2 6 8 12 18 22 24 28
// The acceptance bit of draw i (the sign of r23 - t) lands on bit P[i]
// (SignBits32::put<P[i]>).
template<int S>
struct CbCell {
	static constexpr uint32_t MS = S ? 0x11441144u : 0x44114411u; // active X positions
	static constexpr int P[8] = {S ? 2 : 0, S ? 6 : 4, S ? 8 : 10, S ? 12 : 14,
	                             S ? 18 : 16, S ? 22 : 20, S ? 24 : 26, S ? 28 : 30};
};

template<int S>
__device__ __forceinline__ void cb_flips(uint32_t ap, uint32_t aq, uint32_t own, uint32_t right, uint32_t up,
                                         uint32_t& fOwn, uint32_t& fRight, uint32_t& fUp);

// one cell: 8 draws (stream state at the cell's first draw in sc), flips of
// the own word, of column 0 of the right word and of row 0 of the up word.
// DIS: disorder2SyncRngUpdate (kpz/schedulerCPUBaseThreadedRng.inl:50-75),
// the thresholds of each site's class, the site's 2 bits of the map word dw.
// The per-draw step is (aLo, aHi, cc); only (lo ^ hi) & 0x7FFFFF is used, so
// the high state half is carried mod 2^24.
template<int S, bool DIS>
__device__ __forceinline__ void cb_cell(uint64_t sc, uint32_t aLo, uint32_t aHi, uint64_t cc, uint32_t tp, uint32_t tq,
                                        uint32_t own, uint32_t right, uint32_t up, uint32_t& fOwn, uint32_t& fRight,
                                        uint32_t& fUp, uint32_t dw, const uint32_t* tpc, const uint32_t* tqc)
{
	uint32_t lo = (uint32_t)sc, h24 = (uint32_t)(sc >> 32);
	SignBits32 ap, aq;
	// NT = thresholds per draw: 3 (the site's class picks them), 2 (p, q), 1 (q = 0)
	auto draws = [&](auto ntc) {
		constexpr int NT = decltype(ntc)::value;
		static_for<8>([&](auto ic) {
			constexpr int i = decltype(ic)::value;
			constexpr int P = CbCell<S>::P[i];
			const uint64_t pr = (uint64_t)lo * aLo + cc;
			h24 = hi24_step(h24, aLo, lo, aHi, (uint32_t)(pr >> 32));
			lo = (uint32_t)pr;
			const uint32_t r23 = (lo ^ h24) & 0x7FFFFFu;
			if constexpr(NT == 3)
			{
				const uint32_t cls = (dw >> P) & 3u;
				const uint32_t tpv = (cls & 2) ? ((cls & 1) ? tpc[3] : tpc[2]) : ((cls & 1) ? tpc[1] : tpc[0]);
				const uint32_t tqv = (cls & 2) ? ((cls & 1) ? tqc[3] : tqc[2]) : ((cls & 1) ? tqc[1] : tqc[0]);
				ap.template put<P>(r23 - tpv);
				aq.template put<P>(r23 - tqv);
			}
			else
			{
				ap.template put<P>(r23 - tp);
				if constexpr(NT == 2)
					aq.template put<P>(r23 - tq);
			}
		});
	};
	if(DIS)
		draws(std::integral_constant<int, 3>{});
	else if(tq)
		draws(std::integral_constant<int, 2>{});
	else
		draws(std::integral_constant<int, 1>{});
	cb_flips<S>(ap.bits(), aq.bits(), own, right, up, fOwn, fRight, fUp);
}

// the flips of one cell from its acceptance bits (bit 2x + 8y of ap / aq:
// the draw of active site (x, y) accepts a minimum / a maximum)
template<int S>
__device__ __forceinline__ void cb_flips(uint32_t ap, uint32_t aq, uint32_t own, uint32_t right, uint32_t up,
                                         uint32_t& fOwn, uint32_t& fRight, uint32_t& fUp)
{
	// slopes at the X positions 2x + 8y: own X, own Y, X of x+1, Y of y+1
	const uint32_t X = own & 0x55555555u;
	const uint32_t Yb = (own >> 1) & 0x55555555u;
	const uint32_t rX = ((X >> 2) & 0x15151515u) | ((right & 0x01010101u) << 6);
	const uint32_t uY = ((Yb >> 8) & 0x00555555u) | (((up >> 1) & 0x55u) << 24);
	// pqSyncRngUpdate configurations 0b1100 (minimum) and 0b0011 (maximum)
	const uint32_t mn = ~X & ~Yb & rX & uY;
	const uint32_t mx = X & Yb & ~rX & ~uY;
	const uint32_t f = ((mn & ap) | (mx & aq)) & CbCell<S>::MS;
	fOwn = f | (f << 1) | ((f & 0x15151515u) << 2) | ((f & 0x00555555u) << 9);
	fRight = (f >> 6) & 0x01010101u;
	fUp = ((f >> 24) & 0x55u) << 1;
}

// With 64-lane groups and at least four groups per row, a 256-lane
// workgroup holds four consecutive groups of one band (grp = band * ngc + cg)
// and every workgroup is full: a group whose right neighbour group is one of
// them hands its right-edge flips over in LDS, row by row.
__device__ __forceinline__ bool cb_mail(int lG, int ngc) { return lG == 6 && ngc >= 4; }
__device__ __forceinline__ bool cb_right_local(int lG, int64_t grp, int cg, int ngc)
{
	return cb_mail(lG, ngc) && cg + 1 < ngc && (grp & 3) != 3;
}

// DIS: disorder2SyncRngUpdate; ACC: the acceptance bits of every active site
// come from a.acc (a word per cell: bit 2x + 8y accepts a minimum, the bit
// above it a maximum; the TinyMT32 workers), no stream is drawn
template<int S, bool DIS, bool ACC = false>
__global__ __launch_bounds__(256) void cb_local_band_kernel(CbLocalArgs a)
{
	__shared__ uint32_t mail[4][2]; // [group in the workgroup][row parity]
	const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int lXC = a.lx - 2;
	const int XC = 1 << lXC;
	const int G = 1 << a.lG, NB = 1 << a.lNB;
	const int ngc = XC >> a.lG; // groups per band
	// whole groups are live or not (the shuffles below stay inside live
	// groups); with the LDS hand-over every workgroup is full
	if(gid >= ((int64_t)XC << (a.ly - 2 - a.lNB)))
		return;
	const int g = (int)(gid & (G - 1));
	const int64_t grp = gid >> a.lG;
	const int cg = (int)(grp % ngc), band = (int)(grp / ngc);
	const int c = cg * G + g;
	const int yb0 = band * NB;
	uint32_t* __restrict__ lat = a.lat;
	// worker of this cell column and the cell's place in its strip
	const int big = a.rem * (a.pc + 1);
	int t, c0;
	if(c < big)
	{
		t = c / (a.pc + 1);
		c0 = t * (a.pc + 1);
	}
	else
	{
		t = a.rem + (c - big) / a.pc;
		c0 = big + (t - a.rem) * a.pc;
	}
	const int w = a.pc + (t < a.rem ? 1 : 0);
	const int cells0 = w / 2; // xCenter = xMin + ((w*4/2) & ~3)
	const int cc = c - c0;
	const bool h = cc >= cells0;
	const int ci = h ? cc - cells0 : cc;
	const int cellsH = h ? w - cells0 : cells0;
	uint64_t s = 0;
	Affine rowStep{0, 0};
	if constexpr(!ACC)
	{
		const int clsH = cb_class(a, cellsH);
		const uint64_t s0 = a.st_in[t];
		s = s0;
		if(h)
			s = a.baseJ[cb_class(a, cells0)](s);
		s = a.rowJ[clsH * a.nyb + yb0](s);
		s = a.cellJ[ci](s);
		if(yb0 == 0 && cc == 0)
			a.st_out[t] = (t < a.rem ? a.fullBig : a.fullSmall)(s0);
		// the same cell of the next row: 8 draws per cell of the half strip later
		rowStep = a.rowJ[clsH * a.nyb + 1];
	}
	const uint32_t rLo = (uint32_t)rowStep.a, rHi = (uint32_t)(rowStep.a >> 32);

	const uint32_t tp = a.tp, tq = a.tq;
	// step constants stay in VGPRs for the whole loop (a VOP3 multiply-add
	// takes one scalar operand at most)
	uint32_t aLo = (uint32_t)a.step.a, aHi = (uint32_t)(a.step.a >> 32);
	uint64_t stepC = a.step.c, rowC = rowStep.c;
	uint32_t rLoV = rLo, rHiV = rHi;
	asm volatile("" : "+v"(aLo), "+v"(aHi), "+v"(rLoV), "+v"(rHiV), "+v"(stepC), "+v"(rowC));
	const bool wrapG = (XC == G); // the group spans the row: the right edge wraps by shuffle
	const bool useMail = cb_mail(a.lG, ngc); // launch-uniform
	const bool mailR = cb_right_local(a.lG, grp, cg, ngc);
	const bool mailL = cg > 0 && cb_right_local(a.lG, grp - 1, cg - 1, ngc);
	// 32-bit byte offsets into the lattice and the map (below 4 GiB): the
	// loads and stores take the scalar base plus one VGPR offset
	const char* const base = (const char*)lat;
	const uint32_t rowBytes = (uint32_t)XC * 4u;
	const uint32_t latMask = (uint32_t)(((uint64_t)a.nyb << lXC) * 4u - 1u);
	const uint32_t c4 = (uint32_t)c * 4u, cr4 = (uint32_t)((c + 1) & (XC - 1)) * 4u;
	uint32_t rowOff = ((uint32_t)yb0 << lXC) * 4u;
	uint32_t own = *(const uint32_t*)(base + (rowOff + c4));
	uint32_t fromBelow = 0;
	for(int it = 0; it < NB; ++it)
	{
		const uint32_t upOff = (rowOff + rowBytes) & latMask;
		const uint32_t right = *(const uint32_t*)(base + (rowOff + cr4));
		const uint32_t up = *(const uint32_t*)(base + (upOff + c4));
		uint32_t fOwn, fRight, fUp;
		if constexpr(ACC)
		{
			const uint32_t acc = *(const uint32_t*)((const char*)a.acc + (rowOff + c4));
			cb_flips<S>(acc, acc >> 1, own, right, up, fOwn, fRight, fUp);
		}
		else
		{
			const uint32_t dw = DIS ? *(const uint32_t*)((const char*)a.dis + (rowOff + c4)) : 0u;
			cb_cell<S, DIS>(s, aLo, aHi, stepC, tp, tq, own, right, up, fOwn, fRight, fUp, dw, a.tpc, a.tqc);
		}
		uint32_t fromLeft = (uint32_t)__shfl((int)fRight, (g - 1) & (G - 1), G);
		if(g == 0 && !wrapG)
			fromLeft = 0;
		if(useMail)
		{
			// the left group's right-edge flips of this row, through LDS (two
			// slots: a slot is rewritten two barriers after it was read)
			if(g == G - 1 && mailR)
				mail[grp & 3][it & 1] = fRight;
			__syncthreads();
			if(g == 0 && mailL)
				fromLeft = mail[(grp - 1) & 3][it & 1];
		}
		*(uint32_t*)((char*)lat + (rowOff + c4)) = own ^ fOwn ^ fromLeft ^ fromBelow;
		if(g == G - 1 && !wrapG && !mailR)
			a.edgeR[(size_t)(yb0 + it) * ngc + cg] = fRight;
		fromBelow = fUp;
		own = up; // no other lane writes the bits this lane reads from it
		rowOff = upOff;
		if constexpr(!ACC)
		{
			// the jump to the next row's cell, kept exact in the low state word
			// and mod 2^24 in the high one: all cb_cell uses (the stream's final
			// state comes from s0 above)
			const uint64_t pr = (uint64_t)(uint32_t)s * rLoV + rowC;
			const uint32_t h24 = hi24_step((uint32_t)(s >> 32), rLoV, (uint32_t)s, rHiV, (uint32_t)(pr >> 32));
			s = ((uint64_t)h24 << 32) | (uint32_t)pr;
		}
	}
	a.edgeU[(size_t)band * XC + c] = fromBelow;
}

// flips that crossed a group's right edge or a band's top row
__global__ __launch_bounds__(256) void cb_local_fix_kernel(CbLocalArgs a)
{
	const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int lXC = a.lx - 2;
	const int XC = 1 << lXC;
	const int G = 1 << a.lG;
	const int ngc = XC >> a.lG;
	const bool edges = XC > G; // groups do not span the row
	const int64_t nU = (int64_t)(a.nyb >> a.lNB) * XC;
	const int64_t nR = edges ? ((int64_t)a.nyb * ngc) : 0;
	// every target word has exactly one thread here: plain read-modify-write
	if(gid < nU)
	{
		// first row of band b + 1, column c: the top-row flips of band b and,
		// at a group start, the right-edge flips of the group to its left
		// when they did not go through LDS
		const int b = (int)(gid >> lXC), c = (int)(gid & (XC - 1));
		const int r = ((b + 1) << a.lNB) & (a.nyb - 1);
		uint32_t v = a.edgeU[gid];
		if(edges && (c & (G - 1)) == 0)
		{
			const int cgL = ((c >> a.lG) + ngc - 1) & (ngc - 1);
			if(!cb_right_local(a.lG, (int64_t)(r >> a.lNB) * ngc + cgL, cgL, ngc))
				v ^= a.edgeR[(size_t)r * ngc + cgL];
		}
		if(v)
			a.lat[((size_t)r << lXC) + c] ^= v;
	}
	else if(gid < nU + nR)
	{
		// row r (not a band's first row), the start of the group right of cg
		const int64_t k = gid - nU;
		const int r = (int)(k >> (lXC - a.lG)), cg = (int)(k & (ngc - 1));
		if((r & ((1 << a.lNB) - 1)) == 0 || cb_right_local(a.lG, (int64_t)(r >> a.lNB) * ngc + cg, cg, ngc))
			return;
		const uint32_t v = a.edgeR[k];
		if(v)
			a.lat[((size_t)r << lXC) + (((cg + 1) * G) & (XC - 1))] ^= v;
	}
}

// ======================================================================
// TinyMT32 workers of the checkerboard schedulers (kpz/kpzFrontend.cpp:
// 324-331).  A TinyMT32 stream has no cheap jump ahead, so lane t is worker t
// and replays the worker's own walk with its own stream.
// ======================================================================
// worker t of row y and the row's index r in the worker's rows
// (kpz/schedulerCPUCBBit.cpp:46-62: the first rem workers have per + 1 rows)
__device__ __forceinline__ void cbbit_worker_row(int64_t y, int per, int rem, int64_t& t, int64_t& r)
{
	const int64_t big = (int64_t)rem * (per + 1);
	if(y < big)
	{
		t = y / (per + 1);
		r = y - t * (per + 1);
	}
	else
	{
		t = rem + (y - big) / per;
		r = (y - big) - (t - rem) * per;
	}
}

// SchedulerCPUCBBit<TinyMT32>, the random part of one half-sweep: lane
// c * Y + y is chunk c of the 2^lC chunks of site row y; it takes its
// worker's stream to the chunk's first draw (a jump, rowJ[(y << lC) + c] =
// x^(r * draws per row + c * draws per chunk) mod the stream's orbit
// polynomial) and draws the chunk's words in order
// (kpz/schedulerCPUCBBit.cpp:43-79): MODE 0 = pSyncRngUpdateVL3's random
// words (two draws per u64 word, the first the high half; one per u32 word
// when L = 64), 1 = per-site pq, 2 = disorder2SyncRngUpdate (one
// randomFloat() per active site, x ascending, accepted as (u >> 8) < t; the
// class is the site's x-slope bit of the map).  Output: the words (MODE 0) or
// the minimum and maximum acceptance masks (Q words each), which
// cbbit_half_kernel<W, 3 / 4> applies.  A workgroup is 256 rows of one
// chunk: blocks of B words per row are staged in LDS and stored row segment
// by row segment (a lane per row would store 256 scattered words).  The lane
// of a worker's last row writes the worker's state after the half-sweep.
template<class W, int MODE>
__global__ __launch_bounds__(256) void cbbit_tmt_gen_kernel(CbHalfArgs a, int lC, const uint32_t* __restrict__ gin,
		uint32_t* __restrict__ gout, const Poly128* __restrict__ rowJ, const TmtOrbit* __restrict__ orb,
		W* __restrict__ out)
{
	constexpr int WB = (int)sizeof(W) * 8;
	constexpr int NO = (MODE == 0) ? 1 : 2;        // output planes
	constexpr int B = (MODE == 0) ? 16 : 8;        // staged words per row
	__shared__ W stage[NO][256][B + 1];
	const int64_t Y = 1ll << a.ly;
	const int64_t lane = (int64_t)blockIdx.x * 256 + threadIdx.x;
	const uint64_t Wn = 1ull << (a.lx - (WB == 64 ? 7 : 6)); // words per row and quadrant
	const uint64_t wc = Wn >> lC;                             // words per chunk
	const bool staged = Y >= 256 && wc >= (uint64_t)B;        // launch-uniform
	if(!staged && lane >= (Y << lC))
		return; // no barrier on this path
	const int c = (int)(lane >> a.ly);
	const int64_t y = lane & (Y - 1);
	int64_t t, r;
	cbbit_worker_row(y, a.per, a.rem, t, r);
	TinyMT g = TinyMT::load(gin, a.T, t);
	if(r > 0 || c > 0)
	{
		uint32_t st[4] = {g.s0, g.s1, g.s2, g.s3};
		tmt_jump4(st, rowJ[(y << lC) + c], orb[t], g.m1, g.m2);
		g.s0 = st[0];
		g.s1 = st[1];
		g.s2 = st[2];
		g.s3 = st[3];
	}
	const uint64_t Q = Wn << a.ly;
	const uint64_t row = (uint64_t)y * Wn;
	const W* __restrict__ dis = (MODE == 2) ? reinterpret_cast<const W*>(a.dis) + (a.signum ? 2 * Q : 0) : nullptr;
	auto word = [&](uint64_t xw, W& v0, W& v1) {
		if constexpr(MODE == 0)
		{
			if constexpr(WB == 64)
			{
				const uint64_t hi = g.u32();
				const uint64_t lo = g.u32();
				v0 = (W)((hi << 32) ^ lo);
			}
			else
				v0 = (W)g.u32();
			v1 = 0;
		}
		else
		{
			// acceptance bits: the sign of rf - t (both at most 2^24) of draw k on
			// bit k of the word (SignBits32); the disorder mode keeps both
			// classes' bits and selects per site
			W pb = 0, qb = 0, pb1 = 0, qb1 = 0;
			static_for<WB / 32>([&](auto hc) {
				constexpr int h = decltype(hc)::value;
				SignBits32 pa, qa, pa1, qa1;
				static_for<32>([&](auto kc) {
					constexpr int k = decltype(kc)::value;
					const uint32_t rf = g.u32() >> 8;
					pa.template put<k>(rf - a.tp);
					qa.template put<k>(rf - a.tq);
					if constexpr(MODE == 2)
					{
						pa1.template put<k>(rf - a.tp1);
						qa1.template put<k>(rf - a.tq1);
					}
				});
				pb |= (W)pa.bits() << (32 * h);
				qb |= (W)qa.bits() << (32 * h);
				if constexpr(MODE == 2)
				{
					pb1 |= (W)pa1.bits() << (32 * h);
					qb1 |= (W)qa1.bits() << (32 * h);
				}
			});
			if(MODE == 2)
			{
				const W cl = dis[row + xw]; // class: the site's x-slope bit of the map
				pb = (pb & ~cl) | (pb1 & cl);
				qb = (qb & ~cl) | (qb1 & cl);
			}
			v0 = pb;
			v1 = qb;
		}
	};
	const uint64_t x0 = (uint64_t)c * wc, x1 = x0 + wc;
	if(!staged)
	{
		for(uint64_t xw = x0; xw < x1; ++xw)
		{
			W v0, v1;
			word(xw, v0, v1);
			out[row + xw] = v0;
			if(NO == 2)
				out[Q + row + xw] = v1;
		}
	}
	else
	{
		const int tid = (int)threadIdx.x;
		const uint64_t y0 = (uint64_t)(y - tid); // first row of the workgroup
		for(uint64_t xb = x0; xb < x1; xb += B)
		{
			for(int w = 0; w < B; ++w)
			{
				W v0, v1;
				word(xb + w, v0, v1);
				stage[0][tid][w] = v0;
				if(NO == 2)
					stage[NO - 1][tid][w] = v1;
			}
			__syncthreads();
			for(int s = 0; s < B; ++s)
			{
				const int e = s * 256 + tid, rr = e / B, w = e % B;
				const uint64_t o = (y0 + (uint64_t)rr) * Wn + xb + (uint64_t)w;
				out[o] = stage[0][rr][w];
				if(NO == 2)
					out[Q + o] = stage[NO - 1][rr][w];
			}
			__syncthreads();
		}
	}
	const int64_t rows = a.per + (t < a.rem ? 1 : 0);
	if(r == rows - 1 && c == (1 << lC) - 1)
		g.store(gout, a.T, t);
}

// out[(y << lC) + c] = x^(r * drawsPerRow + c * drawsPerRow / 2^lC) mod the
// orbit polynomial of row y's worker
__global__ __launch_bounds__(256) void tmt_row_polys_kernel(const TmtOrbit* __restrict__ orb, int64_t nrows, int lC,
		int per, int rem, uint64_t drawsPerRow, Poly128* __restrict__ out)
{
	const int64_t lane = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if(lane >= (nrows << lC))
		return;
	const int64_t y = lane >> lC;
	const uint64_t c = (uint64_t)(lane & ((1 << lC) - 1));
	int64_t t, r;
	cbbit_worker_row(y, per, rem, t, r);
	out[lane] = poly_xpow((uint64_t)r * drawsPerRow + c * (drawsPerRow >> lC), orb[t]);
}

// SchedulerCPUCB<TinyMT32>, the random part of one half-sweep: lane (t, k)
// draws chunk k (CB_CHUNK cells, 8 draws each) of worker t's walk -- the
// first half of its strip of cell columns, then the second, cell row by cell
// row, cell by cell, two active sites per site row, one randomFloat() each
// (kpz/schedulerCPUCB.cpp:128-168) -- with the worker's stream jumped to the
// chunk's first draw (chunkJ[lane] = x^(8 CB_CHUNK k) mod its orbit
// polynomial).  Output: one word per cell, bit 2x + 8y set when the draw of
// active site (x, y) accepts a minimum ((u >> 8) < tp of the site's class),
// the bit above it when it accepts a maximum; cb_local_band_kernel<S, *, ACC>
// applies them.  The lane of a worker's last chunk writes the worker's state.
constexpr int CB_CHUNK = 256;
static_assert(CB_TMT_CHUNK_DRAWS == (uint64_t)CB_CHUNK * 8, "chunk of the SCA TinyMT32 walk");

template<int S, bool DIS>
__global__ __launch_bounds__(256) void cb_local_tmt_gen_kernel(CbLocalArgs a, int chunks, const uint32_t* __restrict__ gin,
		uint32_t* __restrict__ gout, const Poly128* __restrict__ chunkJ, const TmtOrbit* __restrict__ orb,
		uint32_t* __restrict__ acc)
{
	// lane k * T + t: neighbouring lanes are neighbouring workers at the same
	// place of their walks, so their stores of a step land side by side
	const int64_t lane = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if(lane >= a.T * chunks)
		return;
	const int k = (int)(lane / a.T);
	const int64_t t = lane - (int64_t)k * a.T;
	const int64_t XC = 1ll << (a.lx - 2), YC = 1ll << (a.ly - 2);
	const int64_t w = a.pc + (t < a.rem ? 1 : 0);
	const int64_t c0 = t * a.pc + (t < a.rem ? t : a.rem);
	const int64_t wl = w / 2; // xCenter = xMin + ((w*4/2) & ~3)
	const int64_t L = w * YC;  // cells of the worker's walk
	const int64_t q0 = (int64_t)k * CB_CHUNK;
	if(q0 >= L)
		return;
	const int64_t q1 = (q0 + CB_CHUNK < L) ? q0 + CB_CHUNK : L;
	TinyMT g = TinyMT::load(gin, a.T, t);
	if(k > 0)
	{
		uint32_t st[4] = {g.s0, g.s1, g.s2, g.s3};
		tmt_jump4(st, chunkJ[t * chunks + k], orb[t], g.m1, g.m2);
		g.s0 = st[0];
		g.s1 = st[1];
		g.s2 = st[2];
		g.s3 = st[3];
	}
	// cell q0 of the walk: half, cell row, cell in the half row
	int half;
	int64_t r, ci;
	if(q0 < wl * YC)
	{
		half = 0;
		r = q0 / wl;
		ci = q0 - r * wl;
	}
	else
	{
		half = 1;
		const int64_t q = q0 - wl * YC;
		r = q / (w - wl);
		ci = q - r * (w - wl);
	}
	// the acceptance word of cell io: its 8 draws in site order, the sign of
	// rf - t of draw i on bit CbCell<S>::P[i] = 2x + 8y (minimum), the bit
	// above it for a maximum
	auto cell = [&](size_t io) {
		const uint32_t dw = DIS ? a.dis[io] : 0u;
		SignBits32 ap, aq;
		static_for<8>([&](auto ic) {
			constexpr int P = CbCell<S>::P[decltype(ic)::value];
			const uint32_t rf = g.u32() >> 8;
			uint32_t tp = a.tp, tq = a.tq;
			if(DIS)
			{
				const uint32_t cls = (dw >> P) & 3u;
				tp = a.tpc[cls];
				tq = a.tqc[cls];
			}
			ap.template put<P>(rf - tp);
			aq.template put<P>(rf - tq);
		});
		return (ap.bits() & CbCell<S>::MS) | ((aq.bits() & CbCell<S>::MS) << 1);
	};
	// Sixteen cells of one half row at a time where they fill an aligned
	// 64-byte segment (four 16-byte stores): neighbouring lanes' walks are a
	// strip apart, and single 4-byte stores would reach HBM as partial writes
	constexpr int V = 16;
	int64_t q = q0;
	while(q < q1)
	{
		const int64_t hw = half ? w - wl : wl;
		const size_t io = (size_t)(r * XC + c0 + (half ? wl : 0) + ci);
		int64_t n = 1;
		if((io & (V - 1)) == 0 && ci + V <= hw && q + V <= q1)
		{
			uint32_t v[V];
#pragma unroll
			for(int u = 0; u < V; ++u)
				v[u] = cell(io + u);
			uint4* d = reinterpret_cast<uint4*>(acc + io);
#pragma unroll
			for(int u = 0; u < V / 4; ++u)
				d[u] = make_uint4(v[4 * u], v[4 * u + 1], v[4 * u + 2], v[4 * u + 3]);
			n = V;
		}
		else
			acc[io] = cell(io);
		q += n;
		ci += n;
		if(ci == hw)
		{
			ci = 0;
			if(++r == YC)
			{
				r = 0;
				half = 1;
			}
		}
	}
	if(q1 == L)
		g.store(gout, a.T, t);
}

// chunkJ for cb_local_tmt_gen_kernel: lane t * chunks + k gets
// x^(k * drawsPerChunk) mod the orbit polynomial of worker t
__global__ __launch_bounds__(256) void tmt_chunk_polys_kernel(const TmtOrbit* __restrict__ orb, int64_t T, int chunks,
		uint64_t drawsPerChunk, Poly128* __restrict__ out)
{
	const int64_t lane = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if(lane >= T * chunks)
		return;
	const int64_t t = lane / chunks;
	const uint64_t k = (uint64_t)(lane - t * chunks);
	out[lane] = poly_xpow(k * drawsPerChunk, orb[t]);
}

// ======================================================================
// row statistics (roughness / correlator)
// ======================================================================
// x-slope bits of sites x = 64j .. 64j+63 of row y (bit i <-> x = 64j+i)
template<int ENC>
__device__ __forceinline__ uint64_t xbits(const uint32_t* lat, int lx, int ly, int j, int y)
{
	if(ENC == ENC_LOCAL)
	{
		const int nw = (lx >= 6) ? 16 : (1 << (lx - 2));
		const size_t base = ((size_t)(y >> 2) << (lx - 2)) + (size_t)j * 16;
		const int sh = (y & 3) << 3;
		uint64_t b = 0;
		if(nw == 16)
		{
			// 16 words at a multiple of 16 words from the lattice start: four 16-byte loads
			const uint4* p = reinterpret_cast<const uint4*>(lat + base);
#pragma unroll
			for(int k4 = 0; k4 < 4; ++k4)
			{
				const uint4 v = p[k4];
				const uint32_t ws[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
				for(int kk = 0; kk < 4; ++kk)
					b |= (uint64_t)(even16((ws[kk] >> sh) & 0xFFu) & 0xFu) << (16 * k4 + 4 * kk);
			}
			return b;
		}
		for(int k = 0; k < nw; ++k)
		{
			const uint32_t byte = (lat[base + k] >> sh) & 0xFFu;
			b |= (uint64_t)(even16(byte) & 0xFu) << (4 * k);
		}
		return b;
	}
	else
	{
		const size_t sw = (size_t)1 << (lx + ly - 4);
		const size_t w = ((size_t)y << (lx - 6)) + j;
		const uint32_t e = lat[w + ((y & 1) ? (sw >> 1) : 0)];       // even x: sub = y&1
		const uint32_t o = lat[w + ((y & 1) ? 0 : (sw >> 1))];       // odd x
		const uint64_t se = (uint64_t)spread16(e & 0xFFFFu) | ((uint64_t)spread16(e >> 16) << 32);
		const uint64_t so = (uint64_t)spread16(o & 0xFFFFu) | ((uint64_t)spread16(o >> 16) << 32);
		return se | (so << 1);
	}
}

// matching slope bits (x and y) of the chunk's sites: count of equal bits
template<int ENC>
__device__ __forceinline__ int eqbits(const uint32_t* a, const uint32_t* b, int lx, int ly, int j, int y)
{
	if(ENC == ENC_LOCAL)
	{
		const int nw = (lx >= 6) ? 16 : (1 << (lx - 2));
		const size_t base = ((size_t)(y >> 2) << (lx - 2)) + (size_t)j * 16;
		const uint32_t m = 0xFFu << ((y & 3) << 3);
		int e = 0;
		for(int k = 0; k < nw; ++k)
			e += __popc(~(a[base + k] ^ b[base + k]) & m);
		return e;
	}
	else
	{
		const size_t q = (size_t)1 << (lx + ly - 6);
		const size_t w = ((size_t)y << (lx - 6)) + j;
		int e = 0;
		for(int k = 0; k < 4; ++k)
			e += __popc(~(a[w + k * q] ^ b[w + k * q]));
		return e;
	}
}

template<int ENC>
__device__ __forceinline__ int ybit0(const uint32_t* lat, int lx, int ly, int y)
{
	if(ENC == ENC_LOCAL)
		return (lat[(size_t)(y >> 2) << (lx - 2)] >> (((y & 3) << 3) | 1)) & 1;
	const size_t sw = (size_t)1 << (lx + ly - 4);
	return lat[(sw >> 2) + ((y & 1) ? (sw >> 1) : 0) + ((size_t)y << (lx - 6))] & 1;
}

// x- and y-slope bits of sites x = 64j .. 64j+63 of row y (bit i <-> x = 64j+i)
template<int ENC>
__device__ __forceinline__ void xybits(const uint32_t* lat, int lx, int ly, int j, int y, uint64_t& X, uint64_t& Yb)
{
	X = 0;
	Yb = 0;
	if(ENC == ENC_LOCAL)
	{
		const size_t base = ((size_t)(y >> 2) << (lx - 2)) + (size_t)j * 16;
		const int sh = (y & 3) << 3;
		if(lx >= 6)
		{
			const uint4* v = reinterpret_cast<const uint4*>(lat + base);
#pragma unroll
			for(int k = 0; k < 4; ++k)
			{
				const uint4 w = v[k];
				const uint32_t b = ((w.x >> sh) & 0xFFu) | (((w.y >> sh) & 0xFFu) << 8) | (((w.z >> sh) & 0xFFu) << 16)
				                 | (((w.w >> sh) & 0xFFu) << 24);
				X |= (uint64_t)even16(b) << (16 * k);
				Yb |= (uint64_t)even16(b >> 1) << (16 * k);
			}
		}
		else
		{
			for(int k = 0; k < (1 << (lx - 2)); ++k)
			{
				const uint32_t b = (lat[base + k] >> sh) & 0xFFu;
				X |= (uint64_t)(even16(b) & 0xFu) << (4 * k);
				Yb |= (uint64_t)(even16(b >> 1) & 0xFu) << (4 * k);
			}
		}
	}
	else
	{
		X = xbits<ENC>(lat, lx, ly, j, y);
		const size_t sw = (size_t)1 << (lx + ly - 4);
		const size_t w = (sw >> 2) + ((size_t)y << (lx - 6)) + j;
		const uint32_t e = lat[w + ((y & 1) ? (sw >> 1) : 0)];
		const uint32_t o = lat[w + ((y & 1) ? 0 : (sw >> 1))];
		const uint64_t se = (uint64_t)spread16(e & 0xFFFFu) | ((uint64_t)spread16(e >> 16) << 32);
		const uint64_t so = (uint64_t)spread16(o & 0xFFFFu) | ((uint64_t)spread16(o >> 16) << 32);
		Yb = se | (so << 1);
	}
}

// the y slope of site (x, y)
template<int ENC>
__device__ __forceinline__ uint32_t yslope(const uint32_t* lat, int lx, int ly, int x, int y)
{
	if(ENC == ENC_LOCAL)
		return (lat[((size_t)(y >> 2) << (lx - 2)) + (size_t)(x >> 2)] >> ((((x & 3) << 1) | ((y & 3) << 3)) + 1)) & 1u;
	const size_t sw = (size_t)1 << (lx + ly - 4);
	const size_t w = (sw >> 2) + (((x ^ y) & 1) ? (sw >> 1) : 0) + ((size_t)y << (lx - 6)) + (size_t)(x >> 6);
	return (lat[w] >> ((x >> 1) & 31)) & 1u;
}

// SchedulerService::consitencyCheck (kpz/schedulerService.cpp:721-854) on the
// device: every row's x slopes and every column's y slopes sum to zero as
// +-1, and every loop closes, sX(x,y) + sY(x-1,y) = sY(x,y) + sX(x,y-1).
// Lane (band, j) takes the CK-site chunk j of the rows of one band of CB rows
// (chunks fastest, NC chunks per row): rowOnes[y] += ones of the row's x
// slopes (summed over the row's lanes by shuffles), part[band][x] = ones of
// column x's y slopes in the band (bit-sliced counters, one 16-byte-store
// run per lane; column_sum_kernel adds the bands), badLoops += loops that do
// not close.  Every lane runs every shuffle; lanes past the end carry zeros.
constexpr int CK_BAND = 256;
template<int ENC>
__global__ __launch_bounds__(256) void consistency_kernel(const uint32_t* __restrict__ lat, int lx, int ly,
		int* __restrict__ rowOnes, uint16_t* __restrict__ part, unsigned long long* __restrict__ badLoops)
{
	const int L = 1 << lx, Y = 1 << ly;
	const int CK = L < 64 ? L : 64;               // sites per chunk
	const int NC = L / CK;                        // chunks per row (a power of two)
	const int CB = Y < CK_BAND ? Y : CK_BAND;     // rows per band
	const int64_t lane = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	const int64_t lanes = (int64_t)NC * (Y / CB);
	const bool live = lane < lanes;
	const int j = (int)(lane % NC), band = (int)(lane / NC);
	const uint64_t wm = (CK == 64) ? ~0ull : ((1ull << CK) - 1);
	const int y0 = band * CB;
	uint64_t c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; // bit-sliced column counts (CB <= 256)
	unsigned long long bad = 0;
	uint64_t xPrev = 0, yDummy;
	if(live)
		xybits<ENC>(lat, lx, ly, j, (y0 - 1) & (Y - 1), xPrev, yDummy);
	const int xm1 = (j * CK - 1) & (L - 1); // the site left of the chunk
	for(int r = 0; r < CB; ++r)
	{
		const int y = y0 + r;
		uint64_t X = 0, Yb = 0;
		if(live)
		{
			xybits<ENC>(lat, lx, ly, j, y, X, Yb);
			const uint64_t q = ((Yb << 1) | yslope<ENC>(lat, lx, ly, xm1, y)) & wm; // sY(x-1, y)
			const uint64_t e = ((X ^ q) ^ (Yb ^ xPrev)) | ((X & q) ^ (Yb & xPrev));
			bad += (unsigned long long)__popcll(e & wm);
			xPrev = X;
			uint64_t carry = Yb;
#pragma unroll
			for(int k = 0; k < 9; ++k)
			{
				const uint64_t t = c[k] & carry;
				c[k] ^= carry;
				carry = t;
			}
		}
		// the row's ones: the NC lanes of one row are consecutive, NC-aligned
		int ones = __popcll(X);
		for(int o = 1; o < NC && o < 64; o <<= 1)
			ones += __shfl_xor(ones, o, 64);
		if(live && (j & 63) == 0)
			atomicAdd(rowOnes + y, ones);
	}
	if(live)
	{
		uint16_t* dst = part + ((size_t)band << lx) + (size_t)j * CK;
		for(int i0 = 0; i0 < CK; i0 += 8)
		{
			uint32_t n[8];
#pragma unroll
			for(int u = 0; u < 8; ++u)
			{
				n[u] = 0;
#pragma unroll
				for(int k = 0; k < 9; ++k)
					n[u] |= (uint32_t)((c[k] >> (i0 + u)) & 1) << k;
			}
			if(CK >= 8)
				*reinterpret_cast<uint4*>(dst + i0) =
					make_uint4(n[0] | (n[1] << 16), n[2] | (n[3] << 16), n[4] | (n[5] << 16), n[6] | (n[7] << 16));
			else
				for(int u = 0; u < CK; ++u)
					dst[u] = (uint16_t)n[u];
		}
	}
	// one atomic per wave for the loops
	for(int o = 32; o > 0; o >>= 1)
		bad += __shfl_xor(bad, o, 64);
	if((threadIdx.x & 63) == 0 && bad)
		atomicAdd(badLoops, bad);
}

// colOnes[x] = sum over the bands of part[band][x]
__global__ __launch_bounds__(256) void column_sum_kernel(const uint16_t* __restrict__ part, int nbands, int lx,
		int* __restrict__ colOnes)
{
	const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if(x >= (1ll << lx))
		return;
	int n = 0;
	for(int b = 0; b < nbands; ++b)
		n += part[((size_t)b << lx) + (size_t)x];
	colOnes[x] = n;
}

// local prefix statistics of one chunk: q = sum d, sq = sum Q_i, sq2 = sum Q_i^2
// over the n counted sites (site x = 0 is excluded from the row prefix)
__device__ __forceinline__ void chunk_prefix(uint64_t bits, int nbits, bool skip0, int& q, int64_t& sq, int64_t& sq2)
{
	int qq = 0, a1 = 0, a2 = 0;
	for(int k = skip0 ? 1 : 0; k < nbits; ++k)
	{
		qq += ((bits >> k) & 1) ? 1 : -1;
		a1 += qq;
		a2 += qq * qq;
	}
	q = qq;
	sq = a1;
	sq2 = a2;
}

__device__ __forceinline__ void chunk_prefix2(uint64_t b1, uint64_t b2, int nbits, bool skip0,
		int& q1, int64_t& s1, int64_t& s11, int& q2, int64_t& s2, int64_t& s22, int64_t& s12)
{
	int a = 0, b = 0, sa = 0, saa = 0, sb = 0, sbb = 0, sab = 0;
	for(int k = skip0 ? 1 : 0; k < nbits; ++k)
	{
		a += ((b1 >> k) & 1) ? 1 : -1;
		b += ((b2 >> k) & 1) ? 1 : -1;
		sa += a;
		saa += a * a;
		sb += b;
		sbb += b * b;
		sab += a * b;
	}
	q1 = a;
	s1 = sa;
	s11 = saa;
	q2 = b;
	s2 = sb;
	s22 = sbb;
	s12 = sab;
}

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v, int lane)
{
#pragma unroll
	for(int o = 1; o < 64; o <<= 1)
	{
		const int64_t u = (int64_t)shfl64((uint64_t)v, lane >= o ? lane - o : lane, 64);
		if(lane >= o)
			v += u;
	}
	return v;
}

__device__ __forceinline__ int64_t wave_sum(int64_t v)
{
#pragma unroll
	for(int o = 32; o > 0; o >>= 1)
		v += (int64_t)shfl64((uint64_t)v, (threadIdx.x & 63) ^ o, 64);
	return v;
}

// One wave computes the statistics of site row y; lane 0 writes them to o[]
// (ROW_STATS values, or the first 4 when !CORR and `short4` is set).
template<int ENC, bool CORR>
__device__ __forceinline__ void row_stats_row(const uint32_t* __restrict__ lat, const uint32_t* __restrict__ snap,
		int lx, int ly, int y, int lane, int64_t* __restrict__ o, bool short4)
{
	const int nbits = (lx >= 6) ? 64 : (1 << lx);
	const int nch = (lx >= 6) ? (1 << (lx - 6)) : 1;
	int64_t S1 = 0, S2 = 0, T1 = 0, T2 = 0, S12 = 0, tcs = 0;
	int64_t run = 0, runS = 0;
	int d0 = 0, d0s = 0;
	for(int base = 0; base < nch; base += 64)
	{
		const int j = base + lane;
		const bool valid = j < nch;
		int q = 0, qs = 0;
		int64_t sq = 0, sq2 = 0, ss = 0, ss2 = 0, s12 = 0;
		int n = 0;
		uint64_t b = 0, bs = 0;
		if(valid)
		{
			b = xbits<ENC>(lat, lx, ly, j, y);
			n = (j == 0) ? nbits - 1 : nbits;
			if(CORR)
			{
				bs = xbits<ENC>(snap, lx, ly, j, y);
				chunk_prefix2(b, bs, nbits, j == 0, q, sq, sq2, qs, ss, ss2, s12);
				tcs += 2 * eqbits<ENC>(lat, snap, lx, ly, j, y) - 2 * nbits;
			}
			else
				chunk_prefix(b, nbits, j == 0, q, sq, sq2);
			if(j == 0)
			{
				d0 = (b & 1) ? 1 : -1;
				d0s = (bs & 1) ? 1 : -1;
			}
		}
		const int64_t incl = wave_incl_scan(q, lane);
		const int64_t c = run + incl - q;
		S1 += n * c + sq;
		S2 += n * c * c + 2 * c * sq + sq2;
		run += (int64_t)shfl64((uint64_t)incl, 63, 64);
		if(CORR)
		{
			const int64_t incs = wave_incl_scan(qs, lane);
			const int64_t cs = runS + incs - qs;
			T1 += n * cs + ss;
			T2 += n * cs * cs + 2 * cs * ss + ss2;
			S12 += n * c * cs + c * ss + cs * sq + s12;
			runS += (int64_t)shfl64((uint64_t)incs, 63, 64);
		}
	}
	S1 = wave_sum(S1);
	S2 = wave_sum(S2);
	if(CORR)
	{
		T1 = wave_sum(T1);
		T2 = wave_sum(T2);
		S12 = wave_sum(S12);
		tcs = wave_sum(tcs);
	}
	d0 = __shfl(d0, 0, 64);
	d0s = __shfl(d0s, 0, 64);
	if(lane == 0)
	{
		// site x = 0 sits at the end of the row walk: P_0 = D (all x slopes)
		const int64_t D = run + d0;
		o[0] = S1 + D;
		o[1] = S2 + D * D;
		o[2] = D;
		o[3] = ybit0<ENC>(lat, lx, ly, y);
		if(short4)
			return;
		if(CORR)
		{
			const int64_t Ds = runS + d0s;
			o[4] = T1 + Ds;
			o[5] = Ds;
			o[6] = ybit0<ENC>(snap, lx, ly, y);
			o[7] = S12 + D * Ds;
			o[8] = tcs;
			o[9] = T2 + Ds * Ds;
		}
		else
		{
			o[4] = o[5] = o[6] = o[7] = o[8] = o[9] = 0;
		}
	}
}

// Row statistics of ENC_LOCAL lattices without a snapshot, table-driven:
// one wave per word row (4 site rows, one byte of every word each).  Two
// consecutive words give 8 consecutive sites of a row; their x-slope bits
// (even bits of the row's byte) interleaved form an 8-bit index into an LDS
// table of the walk over those 8 sites from 0: its end q, sum A and sum of
// squares B of the inclusive prefixes.  A lane walks K words per round with
// a running offset c relative to its segment start, a wave scan of the
// segment ends gives each segment's offset C, and the row sums in int64 are
// sum P = sum_lanes (n C + s1), sum P^2 = sum_lanes (n C^2 + 2 C s1 + s2),
// with P the inclusive prefix from site 0.  The heights of Correlator::set
// start the walk at site 1 and put site 0 last (h_0 = D, the row's slope
// sum), so with s0 = site 0's x slope: S1 = sum P - L s0 + D and
// S2 = sum P^2 - 2 s0 sum P + L + D^2 (the values row_stats_row writes).
// Needs 2^(lx-2) >= 2 words per row.
constexpr int RS_K = 16; // words per lane and round

__global__ __launch_bounds__(256) void row_stats_local_kernel(const uint32_t* __restrict__ lat, int lx, int ly,
		int64_t* __restrict__ out)
{
	__shared__ uint32_t tab[256];
	{
		// entry of index i: bit 2k = site k, bit 2k+1 = site 4+k (k < 4);
		// (q + 8) | (A + 36) << 5 | B << 12
		const int i = threadIdx.x;
		int p = 0, A = 0, B = 0;
#pragma unroll
		for(int j = 0; j < 8; ++j)
		{
			p += ((i >> (j < 4 ? 2 * j : 2 * (j - 4) + 1)) & 1) ? 1 : -1;
			A += p;
			B += p * p;
		}
		tab[i] = (uint32_t)(p + 8) | ((uint32_t)(A + 36) << 5) | ((uint32_t)B << 12);
	}
	__syncthreads();
	const int yw = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if(yw >= (1 << (ly - 2)))
		return; // uniform per wave; no barrier below
	const int XW = 1 << (lx - 2);
	const int K = XW < RS_K ? XW : RS_K;     // even: XW >= 2
	const int span = 64 * K;                 // words per round
	const uint32_t* __restrict__ row = lat + ((size_t)yw << (lx - 2));
	int64_t S1[4] = {0, 0, 0, 0}, S2[4] = {0, 0, 0, 0};
	int base[4] = {0, 0, 0, 0};
	for(int w0 = 0; w0 < XW; w0 += span)
	{
		const int wl = w0 + lane * K;
		const bool live = wl < XW;
		int c[4] = {0, 0, 0, 0}, s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
		// 8 sites of each of the 4 rows from the word pair (a, b)
		auto pair = [&](uint32_t a, uint32_t b) {
#pragma unroll
			for(int r = 0; r < 4; ++r)
			{
				const uint32_t t0 = a >> (8 * r), t1 = (b << 1) >> (8 * r);
				const uint32_t e = tab[(t0 & 0x55u) | (t1 & 0xAAu)];
				const int q = (int)(e & 31u) - 8, A = (int)((e >> 5) & 127u) - 36, B = (int)(e >> 12);
				s1[r] += 8 * c[r] + A;
				s2[r] += 8 * c[r] * c[r] + 2 * c[r] * A + B;
				c[r] += q;
			}
		};
		if(live && K == RS_K)
		{
			// 16-byte loads of the lane's 16 consecutive words
#pragma unroll
			for(int k = 0; k < RS_K; k += 4)
			{
				const uint4 v = *reinterpret_cast<const uint4*>(row + wl + k);
				pair(v.x, v.y);
				pair(v.z, v.w);
			}
		}
		else if(live)
		{
			for(int k = 0; k < K; k += 2)
				pair(row[wl + k], row[wl + k + 1]);
		}
		const int n = live ? 4 * K : 0;
#pragma unroll
		for(int r = 0; r < 4; ++r)
		{
			int incl = c[r];
#pragma unroll
			for(int o = 1; o < 64; o <<= 1)
			{
				const int u = __shfl_up(incl, o, 64);
				if(lane >= o)
					incl += u;
			}
			const int64_t C = (int64_t)base[r] + (incl - c[r]);
			S1[r] += (int64_t)n * C + s1[r];
			S2[r] += (int64_t)n * C * C + 2 * C * s1[r] + s2[r];
			base[r] += __shfl(incl, 63, 64);
		}
	}
	const int64_t L = (int64_t)1 << lx;
#pragma unroll
	for(int r = 0; r < 4; ++r)
	{
		const int64_t sp = wave_sum(S1[r]), spp = wave_sum(S2[r]);
		if(lane == 0)
		{
			const int y = 4 * yw + r;
			const int64_t D = base[r];
			const int64_t s0 = ((row[0] >> (8 * r)) & 1u) ? 1 : -1;
			int64_t* o = out + (size_t)y * ROW_STATS;
			o[0] = sp - L * s0 + D;
			o[1] = spp - 2 * s0 * sp + L + D * D;
			o[2] = D;
			o[3] = ybit0<ENC_LOCAL>(lat, lx, ly, y);
			o[4] = o[5] = o[6] = o[7] = o[8] = o[9] = 0;
		}
	}
}

// Row statistics of an ENC_CBBIT lattice (roughness), table-driven as above.
// Site row y's x slopes are bit i of the row's word j in two quadrants
// (kpz/systemSize.h:143-165): the even sites x = 64j + 2i in quadrant
// 2 (y & 1), the odd sites x = 64j + 2i + 1 in quadrant 2 (1 - (y & 1)).  A
// byte holding a nibble of each (4 even sites in bits 0-3, the 4 odd sites
// between them in bits 4-7) indexes the 8-site walk table; the bytes of a
// word pair (e, o) are the bytes of (e & 0x0F0F0F0F) | (o << 4 & 0xF0F0F0F0)
// (nibbles 0, 2, 4, 6) and of (e >> 4 & 0x0F0F0F0F) | (o & 0xF0F0F0F0)
// (nibbles 1, 3, 5, 7).  One wave per site row, CB_RS_K word pairs per lane
// and round (sums per lane and round stay below 2^31).
constexpr int CB_RS_K = 4;

__global__ __launch_bounds__(256) void row_stats_cbbit_kernel(const uint32_t* __restrict__ lat, int lx, int ly,
		int64_t* __restrict__ out)
{
	__shared__ uint32_t tab[256];
	{
		// entry of index i: bit j = site 2j, bit 4 + j = site 2j + 1 (j < 4);
		// (q + 8) | (A + 36) << 5 | B << 12
		const int i = threadIdx.x;
		int p = 0, A = 0, B = 0;
#pragma unroll
		for(int k = 0; k < 8; ++k)
		{
			p += ((i >> ((k & 1) ? 4 + (k >> 1) : (k >> 1))) & 1) ? 1 : -1;
			A += p;
			B += p * p;
		}
		tab[i] = (uint32_t)(p + 8) | ((uint32_t)(A + 36) << 5) | ((uint32_t)B << 12);
	}
	__syncthreads();
	const int y = blockIdx.x * 4 + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if(y >= (1 << ly))
		return; // uniform per wave; no barrier below
	const int WR = 1 << (lx - 6); // words per row and quadrant
	const size_t sw = (size_t)1 << (lx + ly - 4);
	const size_t rowW = (size_t)y << (lx - 6);
	const uint32_t* __restrict__ ev = lat + rowW + ((y & 1) ? (sw >> 1) : 0);
	const uint32_t* __restrict__ od = lat + rowW + ((y & 1) ? 0 : (sw >> 1));
	const int K = WR < CB_RS_K ? WR : CB_RS_K;
	const int span = 64 * K;
	int64_t S1 = 0, S2 = 0;
	int base = 0;
	for(int w0 = 0; w0 < WR; w0 += span)
	{
		const int wl = w0 + lane * K;
		const bool live = wl < WR;
		int c = 0, s1 = 0, s2 = 0;
		auto chunk = [&](uint32_t idx) {
			const uint32_t e = tab[idx];
			const int q = (int)(e & 31u) - 8, A = (int)((e >> 5) & 127u) - 36, B = (int)(e >> 12);
			s1 += 8 * c + A;
			s2 += 8 * c * c + 2 * c * A + B;
			c += q;
		};
		auto pair = [&](uint32_t e, uint32_t o) {
			const uint32_t lo = __builtin_amdgcn_bitop3_b32(e, o << 4, 0x0F0F0F0Fu, 0xE4); // nibbles 0, 2, 4, 6
			const uint32_t hi = __builtin_amdgcn_bitop3_b32(e >> 4, o, 0x0F0F0F0Fu, 0xE4); // nibbles 1, 3, 5, 7
#pragma unroll
			for(int k = 0; k < 4; ++k)
			{
				chunk((lo >> (8 * k)) & 0xFFu);
				chunk((hi >> (8 * k)) & 0xFFu);
			}
		};
		if(live && K == CB_RS_K)
		{
			const uint4 a = *reinterpret_cast<const uint4*>(ev + wl);
			const uint4 b = *reinterpret_cast<const uint4*>(od + wl);
			pair(a.x, b.x);
			pair(a.y, b.y);
			pair(a.z, b.z);
			pair(a.w, b.w);
		}
		else if(live)
		{
			for(int k = 0; k < K; ++k)
				pair(ev[wl + k], od[wl + k]);
		}
		const int n = live ? 64 * K : 0;
		int incl = c;
#pragma unroll
		for(int o = 1; o < 64; o <<= 1)
		{
			const int u = __shfl_up(incl, o, 64);
			if(lane >= o)
				incl += u;
		}
		const int64_t C = (int64_t)base + (incl - c);
		S1 += (int64_t)n * C + s1;
		S2 += (int64_t)n * C * C + 2 * C * s1 + s2;
		base += __shfl(incl, 63, 64);
	}
	const int64_t L = (int64_t)1 << lx;
	const int64_t sp = wave_sum(S1), spp = wave_sum(S2);
	if(lane == 0)
	{
		const int64_t D = base;
		const int64_t s0 = (ev[0] & 1u) ? 1 : -1;
		int64_t* o = out + (size_t)y * ROW_STATS;
		o[0] = sp - L * s0 + D;
		o[1] = spp - 2 * s0 * sp + L + D * D;
		o[2] = D;
		o[3] = ybit0<ENC_CBBIT>(lat, lx, ly, y);
		o[4] = o[5] = o[6] = o[7] = o[8] = o[9] = 0;
	}
}

// The correlator against a snapshot, 8 sites per lookup.  A byte of each
// lattice (8 consecutive x slopes of a row) indexes the 8-site walk tables of
// that lattice, as in the roughness kernels, and the two bytes together index
// the cross table (the sum of P P' over the 8 sites of the two walks from 0).
// Byte layouts: ENC_LOCAL bit 2k = site k, bit 2k+1 = site 4+k (k < 4);
// ENC_CBBIT bit j = site 2j, bit 4+j = site 2j+1.  The walk table packs a
// chunk's end q, sum A and square sum B in one u32 ((q + 8) | (A + 36) << 5 |
// B << 12); the cross table is 65536 int16.  The 129 KiB of tables fill most
// of a CU's LDS, so a workgroup of 1024 lanes builds them
// once and its 16 waves then take units (ENC_LOCAL: word rows of 4 site rows;
// ENC_CBBIT: site rows) in a grid-stride loop.  Sums per lane and round stay
// below 2^31 (at most 256 sites per lane and round).
constexpr int CORR_WG = 1024;

template<int ENC>
__device__ __forceinline__ int corr_site_bit(uint32_t idx, int i)
{
	if(ENC == ENC_LOCAL)
		return (int)(idx >> (i < 4 ? 2 * i : 2 * (i - 4) + 1)) & 1;
	return (int)(idx >> ((i & 1) ? 4 + (i >> 1) : (i >> 1))) & 1;
}

struct CorrTables {
	uint32_t* own; // the 8-site walk of a byte, packed
	int16_t* X;    // [v | w << 8]: sum of P P' of the two walks
};

// running sums of one site row over a lane's chunks, relative to the lane's
// first site: c, cs the walks' ends so far
struct CorrAcc {
	int c = 0, cs = 0, s1 = 0, s2 = 0, t1 = 0, t2 = 0, x11 = 0;
	// chunk of bytes v (this lattice), w (snapshot); xa = the cross entry's byte
	// address (2 (v | w << 8))
	__device__ __forceinline__ void chunk(const CorrTables& T, uint32_t v, uint32_t w, uint32_t xa)
	{
		const uint32_t e = T.own[v], f = T.own[w];
		const int X = *reinterpret_cast<const int16_t*>(reinterpret_cast<const char*>(T.X) + xa);
		const int q = (int)(e & 31u) - 8, A = (int)((e >> 5) & 127u) - 36, B = (int)(e >> 12);
		const int qs = (int)(f & 31u) - 8, As = (int)((f >> 5) & 127u) - 36, Bs = (int)(f >> 12);
		s1 += 8 * c + A;
		s2 += c * (8 * c + 2 * A) + B;
		t1 += 8 * cs + As;
		t2 += cs * (8 * cs + 2 * As) + Bs;
		x11 += c * (8 * cs + As) + cs * A + X;
		c += q;
		cs += qs;
		// one running value per sum: without this the compiler regroups the
		// additions of all the round's chunks into trees whose terms wait in
		// registers
		asm("" : "+v"(c), "+v"(cs), "+v"(s1), "+v"(s2), "+v"(t1), "+v"(t2), "+v"(x11));
	}
};

// the lane's round sums into the row totals: a wave scan of the chunk ends
// gives every lane's offsets C, Cs; n sites per lane this round
struct CorrRow {
	int64_t S1 = 0, S2 = 0, T1 = 0, T2 = 0, S11 = 0;
	int base = 0, baseS = 0;
	__device__ __forceinline__ void add(const CorrAcc& a, int n, int lane)
	{
		int incl = a.c, incs = a.cs;
#pragma unroll
		for(int o = 1; o < 64; o <<= 1)
		{
			const int u = __shfl_up(incl, o, 64), us = __shfl_up(incs, o, 64);
			if(lane >= o)
			{
				incl += u;
				incs += us;
			}
		}
		const int64_t C = (int64_t)base + (incl - a.c), Cs = (int64_t)baseS + (incs - a.cs);
		S1 += (int64_t)n * C + a.s1;
		S2 += (int64_t)n * C * C + 2 * C * a.s1 + a.s2;
		T1 += (int64_t)n * Cs + a.t1;
		T2 += (int64_t)n * Cs * Cs + 2 * Cs * a.t1 + a.t2;
		S11 += (int64_t)n * C * Cs + C * a.t1 + Cs * a.s1 + a.x11;
		base += __shfl(incl, 63, 64);
		baseS += __shfl(incs, 63, 64);
	}
	// lane 0: the row's ROW_STATS values (heights of Correlator::set: the walk
	// starts at site 1 and site 0 comes last, s0 / s0s its x slopes)
	__device__ __forceinline__ void write(int64_t* o, int lx, int64_t s0, int64_t s0s, int64_t y0, int64_t y0s,
	                                      int64_t tcs)
	{
		const int64_t L = (int64_t)1 << lx;
		const int64_t sp = wave_sum(S1), spp = wave_sum(S2), tp = wave_sum(T1), tpp = wave_sum(T2),
		              x = wave_sum(S11);
		if((threadIdx.x & 63) == 0)
		{
			const int64_t D = base, Ds = baseS;
			o[0] = sp - L * s0 + D;
			o[1] = spp - 2 * s0 * sp + L + D * D;
			o[2] = D;
			o[3] = y0;
			o[4] = tp - L * s0s + Ds;
			o[5] = Ds;
			o[6] = y0s;
			o[7] = x - s0s * sp - s0 * tp + L * s0 * s0s + D * Ds;
			o[8] = tcs;
			o[9] = tpp - 2 * s0s * tp + L + Ds * Ds;
		}
	}
};

// four chunks from a word V of this lattice's bytes and a word W of the
// snapshot's (byte k of each: chunk k); acc(k) is the accumulator of chunk k
template<class Acc>
__device__ __forceinline__ void corr_chunks4(const CorrTables& T, uint32_t V, uint32_t W, Acc&& acc)
{
	// the cross entries' byte addresses, two per v_perm: [v_k, w_k, v_k+1, w_k+1]
	// doubled
	const uint32_t p01 = __builtin_amdgcn_perm(W, V, 0x05010400u), p23 = __builtin_amdgcn_perm(W, V, 0x07030602u);
	const uint32_t x0 = twice(p01) & 0x1FFFEu, x1 = (p01 >> 15) & 0x1FFFEu;
	const uint32_t x2 = twice(p23) & 0x1FFFEu, x3 = (p23 >> 15) & 0x1FFFEu;
	acc(0).chunk(T, V & 0xFFu, W & 0xFFu, x0);
	acc(1).chunk(T, (V >> 8) & 0xFFu, (W >> 8) & 0xFFu, x1);
	acc(2).chunk(T, (V >> 16) & 0xFFu, (W >> 16) & 0xFFu, x2);
	acc(3).chunk(T, V >> 24, W >> 24, x3);
}

constexpr int CORR_LOCAL_K = 16; // words per lane and round (ENC_LOCAL): 64 sites of each row
constexpr int CORR_CB_K = 4;     // word positions per lane and round (ENC_CBBIT): 256 sites

// ENC_LOCAL word row yw (site rows 4 yw .. 4 yw + 3); needs 2^(lx-2) >= 2.
// A round loads all of a lane's K words of both lattices first, so that a
// wave's loads use whole cache lines while they are resident, keeps per word
// pair only the row bytes V (this lattice) and W (snapshot), and then walks
// the four site rows one after the other (one row's running sums live at a
// time).
__device__ __forceinline__ void corr_unit_local(const CorrTables& T, const uint32_t* __restrict__ lat,
		const uint32_t* __restrict__ snap, int lx, int ly, int yw, int lane, int64_t* __restrict__ out)
{
	constexpr int NP = CORR_LOCAL_K / 2; // word pairs per lane and round
	const int XW = 1 << (lx - 2);
	const int K = XW < CORR_LOCAL_K ? XW : CORR_LOCAL_K; // even
	const int span = 64 * K;
	const size_t rowOff = (size_t)yw << (lx - 2);
	const uint32_t* __restrict__ ra = lat + rowOff;
	const uint32_t* __restrict__ rb = snap + rowOff;
	CorrRow R[4];
	int tcs = 0;
	for(int w0 = 0; w0 < XW; w0 += span)
	{
		const int wl = w0 + lane * K;
		const bool live = wl < XW;
		uint32_t V[NP], W[NP];
		// 8 sites of each row r from the word pair (a, b): byte r of a holds
		// sites 0..3 on its even bits, byte r of b sites 4..7 (moved to the odd
		// bits), so byte r of V = (a & 0x55..) | (2b & 0xAA..) is row r's chunk
		auto pair = [&](int p, uint32_t a, uint32_t b, uint32_t as, uint32_t bs) {
			tcs += 2 * (__popc(~(a ^ as)) + __popc(~(b ^ bs))) - 64; // +1 per equal bit, -1 per different one
			V[p] = __builtin_amdgcn_bitop3_b32(a, twice(b), 0x55555555u, 0xE4);
			W[p] = __builtin_amdgcn_bitop3_b32(as, twice(bs), 0x55555555u, 0xE4);
		};
		if(live && K == CORR_LOCAL_K)
		{
			uint4 a[NP / 2], b[NP / 2];
#pragma unroll
			for(int k = 0; k < NP / 2; ++k)
			{
				a[k] = *reinterpret_cast<const uint4*>(ra + wl + 4 * k);
				b[k] = *reinterpret_cast<const uint4*>(rb + wl + 4 * k);
			}
#pragma unroll
			for(int k = 0; k < NP / 2; ++k)
			{
				pair(2 * k, a[k].x, a[k].y, b[k].x, b[k].y);
				pair(2 * k + 1, a[k].z, a[k].w, b[k].z, b[k].w);
			}
		}
		else
		{
#pragma unroll
			for(int p = 0; p < NP; ++p)
			{
				V[p] = W[p] = 0;
				if(live && 2 * p < K)
					pair(p, ra[wl + 2 * p], ra[wl + 2 * p + 1], rb[wl + 2 * p], rb[wl + 2 * p + 1]);
			}
		}
#pragma unroll
		for(int r = 0; r < 4; ++r)
		{
			CorrAcc A;
#pragma unroll
			for(int p = 0; p < NP; ++p)
			{
				if(2 * p >= K) // uniform: K < CORR_LOCAL_K only on rows under 64 sites
					break;
				// cross entry byte address 2 (v | w << 8): bytes r of V and W side by side
				const uint32_t xa = twice(__builtin_amdgcn_perm(W[p], V[p], perm_sel(r, 4 + r, 12, 12)));
				A.chunk(T, (V[p] >> (8 * r)) & 0xFFu, (W[p] >> (8 * r)) & 0xFFu, xa);
			}
			if(!live)
				A = CorrAcc{};
			R[r].add(A, live ? 4 * K : 0, lane);
		}
	}
	// tcs counts both slope bits of all 4 site rows: the word row's first row
	// carries it (the host adds the rows)
	const int64_t tcsRow = wave_sum(tcs);
#pragma unroll
	for(int r = 0; r < 4; ++r)
	{
		const int y = 4 * yw + r;
		const int64_t s0 = ((ra[0] >> (8 * r)) & 1u) ? 1 : -1, s0s = ((rb[0] >> (8 * r)) & 1u) ? 1 : -1;
		R[r].write(out + (size_t)y * ROW_STATS, lx, s0, s0s, ybit0<ENC_LOCAL>(lat, lx, ly, y),
		           ybit0<ENC_LOCAL>(snap, lx, ly, y), r == 0 ? tcsRow : 0);
	}
}

// ENC_CBBIT site row y: x slopes of the even sites in quadrant 2 (y & 1), of
// the odd sites in quadrant 2 (1 - (y & 1)), y slopes in quadrants 1 and 3
// (kpz/systemSize.h:143-165); tcs counts all four quadrants' bits of the row
__device__ __forceinline__ void corr_unit_cbbit(const CorrTables& T, const uint32_t* __restrict__ lat,
		const uint32_t* __restrict__ snap, int lx, int ly, int y, int lane, int64_t* __restrict__ out)
{
	const int WR = 1 << (lx - 6); // words per row and quadrant
	const size_t q = (size_t)1 << (lx + ly - 6);
	const size_t rowW = (size_t)y << (lx - 6);
	const size_t oe = rowW + ((y & 1) ? 2 * q : 0), oo = rowW + ((y & 1) ? 0 : 2 * q);
	const int K = WR < CORR_CB_K ? WR : CORR_CB_K;
	const int span = 64 * K;
	CorrRow R;
	int tcs = 0;
	for(int w0 = 0; w0 < WR; w0 += span)
	{
		const int wl = w0 + lane * K;
		const bool live = wl < WR;
		CorrAcc A;
		// word position j: 64 sites; (e, o) the even / odd sites' x slopes;
		// byte k of lo holds sites 16k .. 16k+7, byte k of hi the next 8
		auto word = [&](uint32_t e, uint32_t o, uint32_t es, uint32_t os, uint32_t y1, uint32_t y3, uint32_t y1s,
		                uint32_t y3s) {
			tcs += 2 * (__popc(~(e ^ es)) + __popc(~(o ^ os)) + __popc(~(y1 ^ y1s)) + __popc(~(y3 ^ y3s))) - 128;
			const uint32_t lo = __builtin_amdgcn_bitop3_b32(e, o << 4, 0x0F0F0F0Fu, 0xE4);   // nibbles 0, 2, 4, 6
			const uint32_t hi = __builtin_amdgcn_bitop3_b32(e >> 4, o, 0x0F0F0F0Fu, 0xE4);   // nibbles 1, 3, 5, 7
			const uint32_t los = __builtin_amdgcn_bitop3_b32(es, os << 4, 0x0F0F0F0Fu, 0xE4);
			const uint32_t his = __builtin_amdgcn_bitop3_b32(es >> 4, os, 0x0F0F0F0Fu, 0xE4);
			// chunk order along the row: lo byte 0, hi byte 0, lo byte 1, ...
			const uint32_t V0 = __builtin_amdgcn_perm(hi, lo, 0x05010400u), V1 = __builtin_amdgcn_perm(hi, lo, 0x07030602u);
			const uint32_t W0 = __builtin_amdgcn_perm(his, los, 0x05010400u), W1 = __builtin_amdgcn_perm(his, los, 0x07030602u);
			corr_chunks4(T, V0, W0, [&](int) -> CorrAcc& { return A; });
			corr_chunks4(T, V1, W1, [&](int) -> CorrAcc& { return A; });
		};
		if(live && K == CORR_CB_K)
		{
			const uint4 e = *reinterpret_cast<const uint4*>(lat + oe + wl);
			const uint4 o = *reinterpret_cast<const uint4*>(lat + oo + wl);
			const uint4 es = *reinterpret_cast<const uint4*>(snap + oe + wl);
			const uint4 os = *reinterpret_cast<const uint4*>(snap + oo + wl);
			const uint4 y1 = *reinterpret_cast<const uint4*>(lat + q + rowW + wl);
			const uint4 y3 = *reinterpret_cast<const uint4*>(lat + 3 * q + rowW + wl);
			const uint4 y1s = *reinterpret_cast<const uint4*>(snap + q + rowW + wl);
			const uint4 y3s = *reinterpret_cast<const uint4*>(snap + 3 * q + rowW + wl);
			word(e.x, o.x, es.x, os.x, y1.x, y3.x, y1s.x, y3s.x);
			word(e.y, o.y, es.y, os.y, y1.y, y3.y, y1s.y, y3s.y);
			word(e.z, o.z, es.z, os.z, y1.z, y3.z, y1s.z, y3s.z);
			word(e.w, o.w, es.w, os.w, y1.w, y3.w, y1s.w, y3s.w);
		}
		else if(live)
		{
			for(int k = 0; k < K; ++k)
			{
				const size_t j = (size_t)(wl + k);
				word(lat[oe + j], lat[oo + j], snap[oe + j], snap[oo + j], lat[q + rowW + j], lat[3 * q + rowW + j],
				     snap[q + rowW + j], snap[3 * q + rowW + j]);
			}
		}
		R.add(A, live ? 64 * K : 0, lane);
	}
	const int64_t tcsRow = wave_sum(tcs);
	const int64_t s0 = (lat[oe] & 1u) ? 1 : -1, s0s = (snap[oe] & 1u) ? 1 : -1;
	R.write(out + (size_t)y * ROW_STATS, lx, s0, s0s, ybit0<ENC_CBBIT>(lat, lx, ly, y),
	        ybit0<ENC_CBBIT>(snap, lx, ly, y), tcsRow);
}

template<int ENC>
__global__ __launch_bounds__(CORR_WG) void row_corr_table_kernel(const uint32_t* __restrict__ lat,
		const uint32_t* __restrict__ snap, int lx, int ly, int64_t* __restrict__ out)
{
	__shared__ uint32_t tOwn[256];
	__shared__ int16_t tX[1 << 16];
	{
		// the walk of this lane's own byte v, then the cross sums against the
		// bytes w = w0, w0 + 4, ... (64 of them)
		const uint32_t v = threadIdx.x & 255u;
		int a[8];
		int p = 0, A = 0, B = 0;
#pragma unroll
		for(int i = 0; i < 8; ++i)
		{
			p += corr_site_bit<ENC>(v, i) ? 1 : -1;
			a[i] = p;
			A += p;
			B += p * p;
		}
		if(threadIdx.x < 256)
			tOwn[v] = (uint32_t)(p + 8) | ((uint32_t)(A + 36) << 5) | ((uint32_t)B << 12);
		for(uint32_t w = threadIdx.x >> 8; w < 256; w += CORR_WG / 256)
		{
			int b = 0, X = 0;
#pragma unroll
			for(int i = 0; i < 8; ++i)
			{
				b += corr_site_bit<ENC>(w, i) ? 1 : -1;
				X += a[i] * b;
			}
			tX[v | (w << 8)] = (int16_t)X;
		}
	}
	__syncthreads();
	const CorrTables T{tOwn, tX};
	const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
	const int units = (ENC == ENC_LOCAL) ? (1 << (ly - 2)) : (1 << ly);
	// no barrier below: waves leave the loop independently
	constexpr int WAVES = CORR_WG / 64;
	for(int u = (int)blockIdx.x * WAVES + wave; u < units; u += (int)gridDim.x * WAVES)
	{
		if(ENC == ENC_LOCAL)
			corr_unit_local(T, lat, snap, lx, ly, u, lane, out);
		else
			corr_unit_cbbit(T, lat, snap, lx, ly, u, lane, out);
	}
}

template<int ENC, bool CORR>
__global__ __launch_bounds__(256) void row_stats_kernel(const uint32_t* __restrict__ lat, const uint32_t* __restrict__ snap,
		int lx, int ly, int64_t* __restrict__ out)
{
	const int y = blockIdx.x * 4 + (threadIdx.x >> 6);
	if(y >= (1 << ly))
		return; // uniform per wave
	row_stats_row<ENC, CORR>(lat, snap, lx, ly, y, threadIdx.x & 63, out + (size_t)y * ROW_STATS, false);
}

// Exact lattice totals from the row statistics, on the device, in two passes.
// Row y starts at height hs_y = hs_{y-1} + D_{y-1} +- 1 (its y slope at x = 0,
// Correlator::set, kpz/correlator.cpp:50-64).  Pass 1: one workgroup per
// RC_ROWS rows, one lane per row; a workgroup scan of the rows' height changes
// gives each row's start height h relative to the workgroup's first row, and
// the workgroup's sums relative to that start are written as one record
// (with H the true start: sum h = A1 + H L n, sum h^2 = A2 + 2 H A1 + H^2 L n,
// and against a snapshot starting at G: sum h g = B12 + H T1 + G A1 + H G L n).
// Pass 2: one workgroup scans the records' height changes for every H and G
// and adds them up in int128.  res = {sum h, sum h^2 lo, hi, sum h_s,
// sum h h_s lo, hi, tcs}; hs (may be null) receives every row's start height
// (pass 1 writes it relative, pass 3 adds H).
constexpr int RC_ROWS = 256;
enum { RC_DA, RC_DB, RC_A1, RC_A2LO, RC_A2HI, RC_T1, RC_B12LO, RC_B12HI, RC_TCS, RC_N, RC_H, RC_G, RC_REC };

__device__ __forceinline__ __int128 wave_sum128(__int128 v)
{
	uint64_t lo = (uint64_t)v;
	int64_t hi = (int64_t)(v >> 64);
#pragma unroll
	for(int o = 32; o > 0; o >>= 1)
	{
		const int src = (threadIdx.x & 63) ^ o;
		const uint64_t lo2 = shfl64(lo, src, 64);
		const int64_t hi2 = (int64_t)shfl64((uint64_t)hi, src, 64);
		const uint64_t s = lo + lo2;
		hi = hi + hi2 + (s < lo ? 1 : 0);
		lo = s;
	}
	return ((__int128)hi << 64) | lo;
}

__global__ __launch_bounds__(RC_ROWS) void row_part_kernel(const int64_t* __restrict__ rows, int lx, int ly, int corr,
		int64_t* __restrict__ hsRel, int64_t* __restrict__ part)
{
	constexpr int NW = RC_ROWS / 64;
	__shared__ int64_t wa[NW], wb[NW], w1[NW], wt[NW], wc[NW];
	__shared__ __int128 w2[NW], w12[NW];
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int64_t ny = (int64_t)1 << ly;
	const int64_t y = (int64_t)blockIdx.x * RC_ROWS + threadIdx.x;
	const bool live = y < ny;
	int64_t o0 = 0, o1 = 0, o2 = 0, o3 = 0, o4 = 0, o5 = 0, o6 = 0, o7 = 0, o8 = 0;
	if(live)
	{
		const int64_t* o = rows + (size_t)y * ROW_STATS;
		o0 = o[0];
		o1 = o[1];
		o2 = o[2];
		o3 = o[3];
		if(corr)
		{
			o4 = o[4];
			o5 = o[5];
			o6 = o[6];
			o7 = o[7];
			o8 = o[8];
		}
	}
	const int64_t sy = o3 ? 1 : -1, syS = o6 ? 1 : -1;
	const int64_t da = live ? sy + o2 : 0, db = (live && corr) ? syS + o5 : 0;
	// workgroup exclusive scan of the row changes
	const int64_t ia = wave_incl_scan(da, lane), ib = corr ? wave_incl_scan(db, lane) : 0;
	if(lane == 63)
	{
		wa[wv] = ia;
		wb[wv] = ib;
	}
	__syncthreads();
	int64_t offA = 0, offB = 0, totA = 0, totB = 0;
#pragma unroll
	for(int k = 0; k < NW; ++k)
	{
		if(k < wv)
		{
			offA += wa[k];
			offB += wb[k];
		}
		totA += wa[k];
		totB += wb[k];
	}
	const int64_t L = (int64_t)1 << lx;
	const int64_t h = offA + ia - da + sy; // this row's start height, relative
	int64_t c1 = 0, t1 = 0, tcs = 0;
	__int128 c2 = 0, c12 = 0;
	if(live)
	{
		if(hsRel)
			hsRel[y] = h;
		c1 = L * h + o0;
		c2 = (__int128)L * h * h + (__int128)2 * h * o0 + o1;
		if(corr)
		{
			const int64_t g = offB + ib - db + syS;
			t1 = L * g + o4;
			c12 = (__int128)L * h * g + (__int128)h * o4 + (__int128)g * o0 + o7;
			tcs = o8;
		}
	}
	c1 = wave_sum(c1);
	c2 = wave_sum128(c2);
	if(corr)
	{
		t1 = wave_sum(t1);
		c12 = wave_sum128(c12);
		tcs = wave_sum(tcs);
	}
	if(lane == 0)
	{
		w1[wv] = c1;
		w2[wv] = c2;
		wt[wv] = t1;
		w12[wv] = c12;
		wc[wv] = tcs;
	}
	__syncthreads();
	if(threadIdx.x == 0)
	{
		int64_t A1 = 0, T1 = 0, TC = 0;
		__int128 A2 = 0, B12 = 0;
#pragma unroll
		for(int k = 0; k < NW; ++k)
		{
			A1 += w1[k];
			A2 += w2[k];
			T1 += wt[k];
			B12 += w12[k];
			TC += wc[k];
		}
		const int64_t n = (ny - (int64_t)blockIdx.x * RC_ROWS) < RC_ROWS ? ny - (int64_t)blockIdx.x * RC_ROWS : RC_ROWS;
		int64_t* r = part + (size_t)blockIdx.x * RC_REC;
		r[RC_DA] = totA;
		r[RC_DB] = totB;
		r[RC_A1] = A1;
		r[RC_A2LO] = (int64_t)(uint64_t)A2;
		r[RC_A2HI] = (int64_t)(A2 >> 64);
		r[RC_T1] = T1;
		r[RC_B12LO] = (int64_t)(uint64_t)B12;
		r[RC_B12HI] = (int64_t)(B12 >> 64);
		r[RC_TCS] = TC;
		r[RC_N] = n;
	}
}

// pass 2: one 1024-lane workgroup over the nb records; lane i takes records
// [i R, (i+1) R); writes every record's start heights H, G (for pass 3)
__global__ __launch_bounds__(1024) void row_final_kernel(int64_t* __restrict__ part, int64_t nb, int lx, int corr,
		int64_t* __restrict__ res)
{
	__shared__ int64_t sa[16], sb[16];
	__shared__ int64_t q1[16], qt[16], qc[16];
	__shared__ __int128 q2[16], q12[16];
	const int i = threadIdx.x, lane = i & 63, wv = i >> 6;
	const int64_t R = (nb + 1023) / 1024;
	const int64_t b0 = i * R, b1 = (b0 + R < nb) ? b0 + R : nb;
	int64_t da = 0, db = 0;
	for(int64_t b = b0; b < b1; ++b)
	{
		da += part[b * RC_REC + RC_DA];
		db += part[b * RC_REC + RC_DB];
	}
	const int64_t ia = wave_incl_scan(da, lane), ib = wave_incl_scan(db, lane);
	if(lane == 63)
	{
		sa[wv] = ia;
		sb[wv] = ib;
	}
	__syncthreads();
	int64_t H = ia - da, G = ib - db;
	for(int k = 0; k < wv; ++k)
	{
		H += sa[k];
		G += sb[k];
	}
	const int64_t L = (int64_t)1 << lx;
	int64_t s1 = 0, t1 = 0, tcs = 0;
	__int128 s2 = 0, s12 = 0;
	for(int64_t b = b0; b < b1; ++b)
	{
		int64_t* r = part + b * RC_REC;
		const int64_t n = r[RC_N], A1 = r[RC_A1];
		const __int128 A2 = ((__int128)r[RC_A2HI] << 64) | (uint64_t)r[RC_A2LO];
		r[RC_H] = H;
		r[RC_G] = G;
		s1 += A1 + H * L * n;
		s2 += A2 + (__int128)2 * H * A1 + (__int128)H * H * L * n;
		if(corr)
		{
			const int64_t T1 = r[RC_T1];
			const __int128 B12 = ((__int128)r[RC_B12HI] << 64) | (uint64_t)r[RC_B12LO];
			t1 += T1 + G * L * n;
			s12 += B12 + (__int128)H * T1 + (__int128)G * A1 + (__int128)H * G * L * n;
			tcs += r[RC_TCS];
		}
		H += r[RC_DA];
		G += r[RC_DB];
	}
	s1 = wave_sum(s1);
	s2 = wave_sum128(s2);
	t1 = wave_sum(t1);
	s12 = wave_sum128(s12);
	tcs = wave_sum(tcs);
	if(lane == 0)
	{
		q1[wv] = s1;
		q2[wv] = s2;
		qt[wv] = t1;
		q12[wv] = s12;
		qc[wv] = tcs;
	}
	__syncthreads();
	if(i == 0)
	{
		int64_t S1 = 0, T1 = 0, TC = 0;
		__int128 S2 = 0, S12 = 0;
		for(int k = 0; k < 16; ++k)
		{
			S1 += q1[k];
			S2 += q2[k];
			T1 += qt[k];
			S12 += q12[k];
			TC += qc[k];
		}
		res[0] = S1;
		res[1] = (int64_t)(uint64_t)S2;
		res[2] = (int64_t)(S2 >> 64);
		res[3] = corr ? T1 : 0;
		res[4] = corr ? (int64_t)(uint64_t)S12 : 0;
		res[5] = corr ? (int64_t)(S12 >> 64) : 0;
		res[6] = corr ? TC : 0;
	}
}

// pass 3 (hs requested): every row's start height = its workgroup's H plus
// the relative start pass 1 wrote
__global__ __launch_bounds__(256) void row_heights_kernel(const int64_t* __restrict__ part, int ly,
		int64_t* __restrict__ hs)
{
	const int64_t y = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if(y < ((int64_t)1 << ly))
		hs[y] += part[(y / RC_ROWS) * RC_REC + RC_H];
}

// Height moments of site row y: out[y*7 + k-1] = sum_x h(x, y)^k, k = 1..7, in
// double (Roughness::addH, kpz/schedulerService.h:56-61), with the row's start
// height hs[y] from the exact row statistics.  Heights as in row_stats_row:
// sites x >= 1 at hs + (x slopes 1..x), site 0 at hs + D.
template<int ENC>
__global__ __launch_bounds__(256) void row_moments_kernel(const uint32_t* __restrict__ lat, int lx, int ly,
		const int64_t* __restrict__ hs, double* __restrict__ out)
{
	const int y = blockIdx.x * 4 + (threadIdx.x >> 6);
	if(y >= (1 << ly))
		return; // uniform per wave
	const int lane = threadIdx.x & 63;
	const int nbits = (lx >= 6) ? 64 : (1 << lx);
	const int nch = (lx >= 6) ? (1 << (lx - 6)) : 1;
	const int64_t h0 = hs[y];
	double m[7] = {0, 0, 0, 0, 0, 0, 0};
	int64_t run = 0;
	int d0 = 0;
	for(int base = 0; base < nch; base += 64)
	{
		const int j = base + lane;
		const bool valid = j < nch;
		uint64_t b = 0;
		int q = 0;
		if(valid)
		{
			b = xbits<ENC>(lat, lx, ly, j, y);
			const uint64_t cnt = (j == 0) ? (b & ~1ull) : b;
			const int n = (j == 0) ? nbits - 1 : nbits;
			const uint64_t live = (nbits == 64) ? ~0ull : ((1ull << nbits) - 1);
			q = 2 * __popcll(cnt & live) - n;
			if(j == 0)
				d0 = (b & 1) ? 1 : -1;
		}
		const int64_t incl = wave_incl_scan(q, lane);
		int64_t h = h0 + run + incl - q;
		if(valid)
		{
			for(int k = (j == 0) ? 1 : 0; k < nbits; ++k)
			{
				h += ((b >> k) & 1) ? 1 : -1;
				const double v = (double)h;
				double p = v;
				m[0] += p;
#pragma unroll
				for(int e = 1; e < 7; ++e)
				{
					p *= v;
					m[e] += p;
				}
			}
		}
		run += (int64_t)shfl64((uint64_t)incl, 63, 64);
	}
	d0 = __shfl(d0, 0, 64);
	if(lane == 0)
	{
		// site x = 0 sits at the end of the row walk: hs + D
		const double v = (double)(h0 + run + d0);
		double p = v;
		m[0] += p;
#pragma unroll
		for(int e = 1; e < 7; ++e)
		{
			p *= v;
			m[e] += p;
		}
	}
#pragma unroll
	for(int e = 0; e < 7; ++e)
	{
#pragma unroll
		for(int o = 32; o > 0; o >>= 1)
			m[e] += __shfl_xor(m[e], o, 64);
	}
	if(lane == 0)
	{
#pragma unroll
		for(int e = 0; e < 7; ++e)
			out[(size_t)y * 7 + e] = m[e];
	}
}

// ---------------------------------------------------------------- fill
__global__ void fill_kernel(uint32_t* lat, uint64_t n, uint64_t split, uint32_t lo, uint32_t hi)
{
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for(uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
		lat[k] = k < split ? lo : hi;
}

// Integer VALU issue-rate probe: 8 independent chains per lane, 8
// instructions per iteration in one asm block (no hazard waits between
// them), so every SIMD issues back to back from its resident waves.
// FAST = 0: v_xad_u32, the rate of shifts, multiplies, add3, bfe, perm and
// alignbit on gfx950; FAST = 1: v_xor_b32, the rate of plain xor / add /
// bitop3, about twice as high.
template<int FAST>
__global__ __launch_bounds__(256) void valu_probe_kernel(uint32_t* __restrict__ out, int iters, uint32_t k0, uint32_t k1)
{
	uint32_t a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
	const uint32_t kv = k1 ^ threadIdx.x;
	for(int it = 0; it < iters; ++it)
	{
		if(FAST)
			asm volatile(
				"v_xor_b32 %0, %8, %0\n\t"
				"v_xor_b32 %1, %8, %1\n\t"
				"v_xor_b32 %2, %8, %2\n\t"
				"v_xor_b32 %3, %8, %3\n\t"
				"v_xor_b32 %4, %8, %4\n\t"
				"v_xor_b32 %5, %8, %5\n\t"
				"v_xor_b32 %6, %8, %6\n\t"
				"v_xor_b32 %7, %8, %7"
				: "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
				: "v"(kv));
		else
			asm volatile(
				"v_xad_u32 %0, %0, %8, %9\n\t"
				"v_xad_u32 %1, %1, %8, %9\n\t"
				"v_xad_u32 %2, %2, %8, %9\n\t"
				"v_xad_u32 %3, %3, %8, %9\n\t"
				"v_xad_u32 %4, %4, %8, %9\n\t"
				"v_xad_u32 %5, %5, %8, %9\n\t"
				"v_xad_u32 %6, %6, %8, %9\n\t"
				"v_xad_u32 %7, %7, %8, %9"
				: "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
				: "s"(k0), "v"(kv));
	}
	out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

// Issue rate of an instruction mix: blocks of 8 independent VALU instructions,
// NS of them v_mul_lo_u32 (slow class) spread evenly among v_xor_b32 (fast
// class).  On gfx950 runs of fast-class instructions issue at about twice the
// rate of slow-class ones, but a fast-class instruction next to a slow-class
// one issues at about the slow rate, so a kernel's bound is the rate of its
// own mix, not the fast-class peak.
#define KPZ_VS(i) "v_mul_lo_u32 %" #i ", %" #i ", %8\n\t"
#define KPZ_VF(i) "v_xor_b32 %" #i ", %8, %" #i "\n\t"
#define KPZ_MIX_ASM(body) \
	asm volatile(body : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(kv))
template<int NS>
__global__ __launch_bounds__(256) void valu_mix_kernel(uint32_t* __restrict__ out, int iters, uint32_t k1)
{
	uint32_t a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
	const uint32_t kv = (k1 ^ threadIdx.x) | 1u;
	for(int it = 0; it < iters; ++it)
	{
		if constexpr(NS == 2)
			KPZ_MIX_ASM(KPZ_VS(0) KPZ_VF(1) KPZ_VF(2) KPZ_VF(3) KPZ_VS(4) KPZ_VF(5) KPZ_VF(6) KPZ_VF(7));
		else if constexpr(NS == 3)
			KPZ_MIX_ASM(KPZ_VS(0) KPZ_VF(1) KPZ_VF(2) KPZ_VS(3) KPZ_VF(4) KPZ_VF(5) KPZ_VS(6) KPZ_VF(7));
		else if constexpr(NS == 4)
			KPZ_MIX_ASM(KPZ_VS(0) KPZ_VF(1) KPZ_VS(2) KPZ_VF(3) KPZ_VS(4) KPZ_VF(5) KPZ_VS(6) KPZ_VF(7));
		else if constexpr(NS == 5)
			KPZ_MIX_ASM(KPZ_VS(0) KPZ_VS(1) KPZ_VF(2) KPZ_VS(3) KPZ_VS(4) KPZ_VF(5) KPZ_VS(6) KPZ_VF(7));
		else if constexpr(NS == 8)
			KPZ_MIX_ASM(KPZ_VS(0) KPZ_VS(1) KPZ_VS(2) KPZ_VS(3) KPZ_VS(4) KPZ_VS(5) KPZ_VS(6) KPZ_VS(7));
		else
			KPZ_MIX_ASM(KPZ_VF(0) KPZ_VF(1) KPZ_VF(2) KPZ_VF(3) KPZ_VF(4) KPZ_VF(5) KPZ_VF(6) KPZ_VF(7));
	}
	out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}
#undef KPZ_MIX_ASM
#undef KPZ_VF
#undef KPZ_VS

__global__ __launch_bounds__(256) void copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n)
{
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for(uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
		dst[k] = src[k];
}

// ======================================================================
// batched replicas
// ======================================================================
// One DT phase of every DT/DTr replica: lane -> {replica, block}.  The host
// keeps each wave within one (kind, mode) class, so the dispatch below is
// wave-uniform.  Block layout 16 x 16 sites.
// RO: the DTr lanes [nstatic, nlanes), else the static DT lanes [0, nstatic):
// one kernel per kind, so that the static DT lanes keep the register budget
// (and occupancy) of the static update loop rather than the DTr staging's
template<bool RO>
__global__ __launch_bounds__(DT_WG) void batch_dt_kernel(BatchDtArgs a)
{
	__shared__ uint32_t tile[33 * DT_WG]; // X plane (16 rows), Y plane (17 rows)
	const int64_t lane = (int64_t)blockIdx.x * DT_WG + threadIdx.x + (RO ? a.nstatic : 0);
	if(lane >= (RO ? a.nlanes : a.nstatic))
		return;
	const uint2 w = a.lanes[lane];
	if(w.x == 0xFFFFFFFFu)
		return;
	if(a.limit && a.m >= a.limit[w.x])
		return;
	const BatchRep& rep = a.reps[w.x];
	DtPhaseArgs p;
	p.lat = rep.lat;
	p.lx = rep.lx;
	p.ly = rep.ly;
	p.lbx = p.lby = 4;
	p.st_in = p.st_out = rep.st[0];
	p.jumps = nullptr;
	p.step = rep.step;
	p.blockJ = rep.full;
	p.tp = rep.tp;
	p.tq = rep.tq;
	const size_t mi = (size_t)a.m * a.nreps + w.x;
	p.ds = (a.order[mi] >> (2 * a.d)) & 3;
	const int64_t nb = 1ll << (rep.lx + rep.ly - 10);
	p.T = nb;
	const DtTile my{tile + threadIdx.x, tile + 16 * DT_WG + threadIdx.x, tile, threadIdx.x * 4};
	if constexpr(RO)
	{
		const uint16_t sh = a.shift[mi];
		p.sx = sh & 0xFF;
		p.sy = sh >> 8;
		DtSlcg<true> cur;
		if(rep.mode)
			dt_block<4, 4, 1, true>(p, w.y, nb, my, nullptr, cur);
		else
			dt_block<4, 4, 0, true>(p, w.y, nb, my, nullptr, cur);
	}
	else
	{
		p.sx = p.sy = 0;
		DtSlcg<false> cur;
		if(rep.mode)
			dt_block<4, 4, 1, false>(p, w.y, nb, my, nullptr, cur);
		else
			dt_block<4, 4, 0, false>(p, w.y, nb, my, nullptr, cur);
	}
}

// One half-sweep of every SCABit replica: wave -> {replica, first row}; the
// wave covers 64 >> lG rows of G = 1 << lG lanes each.
__global__ __launch_bounds__(256) void batch_cbbit_kernel(BatchCbArgs a)
{
	const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	if(gw >= a.nwaves)
		return; // whole wave
	const uint2 w = a.waves[gw];
	if(a.limit && a.m >= a.limit[w.x])
		return; // whole wave: one replica per wave
	const BatchRep& rep = a.reps[w.x];
	const int lane = threadIdx.x & 63;
	CbHalfArgs p;
	p.lat = rep.lat;
	p.lx = rep.lx;
	p.ly = rep.ly;
	p.signum = a.signum;
	p.mode = rep.mode;
	p.st_in = rep.st[a.buf];
	p.st_out = rep.st[a.buf ^ 1];
	p.T = 1ll << rep.ly;
	p.per = 1;
	p.rem = 0;
	p.rowJ = rep.rowJ;
	p.laneJ = rep.laneJ;
	p.stepJ = rep.stepJ;
	p.step = rep.step;
	p.fullBig = p.fullSmall = rep.full;
	p.tp = rep.tp;
	p.tq = rep.tq;
	p.lG = rep.lG;
	const int y = (int)w.y + (lane >> rep.lG);
	const int g = lane & ((1 << rep.lG) - 1);
	if(rep.wide)
	{
		if(rep.mode)
			cbbit_row<uint64_t, 1>(p, y, g);
		else
			cbbit_row<uint64_t, 0>(p, y, g);
	}
	else
	{
		if(rep.mode)
			cbbit_row<uint32_t, 1>(p, y, g);
		else
			cbbit_row<uint32_t, 0>(p, y, g);
	}
}

// ---- per-row roughness statistics of every replica ----
// Walk of one byte of x slopes (bit i set: step +1, else -1; bit 0 first):
// end height q, sum A and square sum B of the heights after each of the 8
// steps, packed (q + 8) | (A + 36) << 5 | B << 12 (q in -8..8, A in -36..36,
// B <= 204)
struct ByteWalk {
	uint32_t e[256];
};
constexpr ByteWalk make_byte_walk()
{
	ByteWalk t{};
	for(int v = 0; v < 256; ++v)
	{
		int q = 0, a = 0, b = 0;
		for(int i = 0; i < 8; ++i)
		{
			q += ((v >> i) & 1) ? 1 : -1;
			a += q;
			b += q * q;
		}
		t.e[v] = (uint32_t)(q + 8) | ((uint32_t)(a + 36) << 5) | ((uint32_t)b << 12);
	}
	return t;
}
__constant__ ByteWalk kByteWalk = make_byte_walk();

// chunk_prefix from the byte walks (tab = kByteWalk in LDS): nbits a multiple
// of 8.  Skipping bit 0 (step s0) lowers every later height by s0, so with
// the full walk's end Q, sum S1 and square sum S2 over N steps the walk from
// bit 1 has end Q - s0, sum S1 - N s0 and square sum S2 - 2 s0 S1 + N.
__device__ __forceinline__ void chunk_prefix_tab(const uint32_t* tab, uint64_t bits, int nbits, bool skip0, int& q,
		int64_t& sq, int64_t& sq2)
{
	int c = 0, s1 = 0, s2 = 0;
	for(int t = 0; t < nbits; t += 8)
	{
		const uint32_t e = tab[(uint32_t)(bits >> t) & 0xFFu];
		const int qe = (int)(e & 31u) - 8, a = (int)((e >> 5) & 127u) - 36, b = (int)(e >> 12);
		s1 += 8 * c + a;
		s2 += 8 * c * c + 2 * c * a + b;
		c += qe;
	}
	if(skip0)
	{
		const int s0 = (bits & 1) ? 1 : -1;
		s2 += nbits - 2 * s0 * s1;
		s1 -= nbits * s0;
		c -= s0;
	}
	q = c;
	sq = s1;
	sq2 = s2;
}

// The first four row statistics of site row y (as row_stats_row with short4)
// from an aligned segment of G lanes (G a power of two dividing the row's
// 64-site chunk count, or 1): segment lane sl takes chunks sl, sl + G, ...;
// the scans and sums stay inside the segment, so the 64 / G segments of a
// wave handle 64 / G rows at once.  tab: kByteWalk in LDS.
template<int ENC>
__device__ __forceinline__ void row_stats_seg(const uint32_t* __restrict__ lat, int lx, int ly, int y, int lane, int G,
		const uint32_t* tab, int64_t* __restrict__ o)
{
	const int sl = lane & (G - 1), s0 = lane - sl;
	const int nbits = (lx >= 6) ? 64 : (1 << lx);
	const int nch = (lx >= 6) ? (1 << (lx - 6)) : 1;
	int64_t S1 = 0, S2 = 0, run = 0;
	int d0 = 0;
	for(int base = 0; base < nch; base += G)
	{
		const int j = base + sl;
		int q = 0;
		int64_t sq = 0, sq2 = 0;
		const uint64_t b = xbits<ENC>(lat, lx, ly, j, y);
		const int n = (j == 0) ? nbits - 1 : nbits;
		chunk_prefix_tab(tab, b, nbits, j == 0, q, sq, sq2);
		if(j == 0)
			d0 = (b & 1) ? 1 : -1;
		int64_t incl = q;
		for(int k = 1; k < G; k <<= 1)
		{
			const int64_t u = (int64_t)shfl64((uint64_t)incl, sl >= k ? lane - k : lane, 64);
			if(sl >= k)
				incl += u;
		}
		const int64_t c = run + incl - q;
		S1 += n * c + sq;
		S2 += n * c * c + 2 * c * sq + sq2;
		run += (int64_t)shfl64((uint64_t)incl, s0 + G - 1, 64);
	}
	for(int k = G >> 1; k > 0; k >>= 1)
	{
		S1 += (int64_t)shfl64((uint64_t)S1, lane ^ k, 64);
		S2 += (int64_t)shfl64((uint64_t)S2, lane ^ k, 64);
	}
	d0 = __shfl(d0, s0, 64);
	if(sl == 0)
	{
		// site x = 0 sits at the end of the row walk: P_0 = D (all x slopes)
		const int64_t D = run + d0;
		o[0] = S1 + D;
		o[1] = S2 + D * D;
		o[2] = D;
		o[3] = ybit0<ENC>(lat, lx, ly, y);
	}
}

// waves[] = {replica, first row}: the wave's 64 / G segments of G lanes take
// consecutive rows (segments past the replica's last row leave; the rest only
// exchange values inside their own segment)
__global__ __launch_bounds__(256) void batch_rowstats_kernel(const BatchRep* __restrict__ reps, const uint2* __restrict__ waves,
		int64_t nwaves, int64_t* __restrict__ scratch)
{
	__shared__ uint32_t tab[256];
	tab[threadIdx.x] = kByteWalk.e[threadIdx.x]; // blockDim.x == 256
	__syncthreads();
	const int64_t gw = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	if(gw >= nwaves)
		return;
	const uint2 w = waves[gw];
	const BatchRep& rep = reps[w.x];
	const int lane = threadIdx.x & 63;
	const int G = rep.lx >= 12 ? 64 : (rep.lx >= 6 ? 1 << (rep.lx - 6) : 1);
	const int y = (int)w.y + lane / G;
	if(y >= (1 << rep.ly))
		return;
	int64_t* o = scratch + 4 * (rep.rowBase + y);
	if(rep.kind == BK_SCABIT)
		row_stats_seg<ENC_CBBIT>(rep.lat, rep.lx, rep.ly, y, lane, G, tab, o);
	else
		row_stats_seg<ENC_LOCAL>(rep.lat, rep.lx, rep.ly, y, lane, G, tab, o);
}

// exact combination per replica: one workgroup of BC_WG lanes per replica.
// Lane k takes a contiguous run of rows and sums them relative to the run's
// start (row y starts at the previous row's start plus its change, as in
// row_part_kernel); a workgroup scan of the runs' height changes gives every
// run's true start H, and sum h = A1 + H L n, sum h^2 = A2 + 2 H A1 + H^2 L n
// (int128) are added over the workgroup.
constexpr int BC_WG = 256;
__global__ __launch_bounds__(BC_WG) void batch_combine_kernel(const BatchRep* __restrict__ reps, int nreps,
		const int64_t* __restrict__ scratch, int64_t* __restrict__ out)
{
	constexpr int NW = BC_WG / 64;
	__shared__ int64_t wd[NW];
	__shared__ __int128 w1[NW], w2[NW];
	const int r = blockIdx.x;
	if(r >= nreps)
		return; // uniform per workgroup
	const BatchRep& rep = reps[r];
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	const int64_t L = 1ll << rep.lx, ny = 1ll << rep.ly;
	const int64_t per = (ny + BC_WG - 1) / BC_WG;
	const int64_t y0 = ((int64_t)threadIdx.x * per < ny) ? (int64_t)threadIdx.x * per : ny;
	const int64_t y1 = (y0 + per < ny) ? y0 + per : ny;
	const int64_t* s = scratch + 4 * (rep.rowBase + y0);
	int64_t rel = 0; // start height of the next row, relative to the run's start
	__int128 a1 = 0, a2 = 0;
	for(int64_t y = y0; y < y1; ++y, s += 4)
	{
		const int64_t hs = rel + (s[3] ? 1 : -1);
		a1 += (__int128)L * hs + s[0];
		a2 += (__int128)(L * hs) * hs + (__int128)(2 * hs) * s[0] + s[1];
		rel = hs + s[2];
	}
	const int64_t n = y1 - y0;
	const int64_t inc = wave_incl_scan(rel, lane);
	if(lane == 63)
		wd[wv] = inc;
	__syncthreads();
	int64_t H = inc - rel;
#pragma unroll
	for(int k = 0; k < NW; ++k)
		if(k < wv)
			H += wd[k];
	const __int128 s1 = a1 + (__int128)H * L * n;
	const __int128 s2 = a2 + (__int128)2 * H * a1 + (__int128)H * H * L * n;
	const __int128 t1 = wave_sum128(s1), t2 = wave_sum128(s2);
	if(lane == 0)
	{
		w1[wv] = t1;
		w2[wv] = t2;
	}
	__syncthreads();
	if(threadIdx.x == 0)
	{
		__int128 S1 = 0, S2 = 0;
#pragma unroll
		for(int k = 0; k < NW; ++k)
		{
			S1 += w1[k];
			S2 += w2[k];
		}
		out[3 * r] = (int64_t)S1;
		out[3 * r + 1] = (int64_t)(uint64_t)S2;
		out[3 * r + 2] = (int64_t)(S2 >> 64);
	}
}

template<int LBX, int LBY, int UPD>
void dt_launch(const DtPhaseArgs& a, int64_t nb, dim3 grid, hipStream_t stream)
{
	if(a.tmt)
		hipLaunchKernelGGL((dt_tmt_kernel<LBX, LBY, UPD>), grid, dim3(DT_WG), 0, stream, a, nb, a.tmt);
	else if(a.random_origin)
		hipLaunchKernelGGL((dt_phase_kernel<LBX, LBY, UPD, true>), grid, dim3(DT_WG), 0, stream, a, nb);
	else
		hipLaunchKernelGGL((dt_phase_kernel<LBX, LBY, UPD, false>), grid, dim3(DT_WG), 0, stream, a, nb);
}

template<int LBX, int LBY>
hipError_t dt_dispatch_upd(const DtPhaseArgs& a, int64_t nb, dim3 grid, hipStream_t stream)
{
	switch(a.upd)
	{
		case 0: dt_launch<LBX, LBY, 0>(a, nb, grid, stream); break;
		case 1: dt_launch<LBX, LBY, 1>(a, nb, grid, stream); break;
		case 2:
			if(!a.dis)
				return hipErrorInvalidValue;
			dt_launch<LBX, LBY, 2>(a, nb, grid, stream);
			break;
		case 3: dt_launch<LBX, LBY, 3>(a, nb, grid, stream); break;
		default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

template<int LBX>
hipError_t dt_dispatch_y(const DtPhaseArgs& a, int64_t nb, dim3 grid, hipStream_t stream)
{
	switch(a.lby)
	{
		case 2: return dt_dispatch_upd<LBX, 2>(a, nb, grid, stream);
		case 3: return dt_dispatch_upd<LBX, 3>(a, nb, grid, stream);
		case 4: return dt_dispatch_upd<LBX, 4>(a, nb, grid, stream);
		default: return hipErrorInvalidValue;
	}
}

} // namespace

bool dt_layout_supported(int lbx, int lby) { return lbx >= 2 && lbx <= 4 && lby >= 2 && lby <= 4; }

hipError_t launch_tmt_row_polys(const TmtOrbit* orb, int64_t nrows, int lC, int per, int rem, uint64_t drawsPerRow,
		Poly128* out, hipStream_t stream)
{
	if(!orb || !out || nrows < 1 || lC < 0 || lC > 8 || per < 0 || rem < 0 || (drawsPerRow & ((1ull << lC) - 1)))
		return hipErrorInvalidValue;
	const int64_t lanes = nrows << lC;
	hipLaunchKernelGGL(tmt_row_polys_kernel, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, stream, orb, nrows, lC,
	                   per, rem, drawsPerRow, out);
	return hipGetLastError();
}

hipError_t launch_cbbit_tmt_gen(const CbHalfArgs& a, int mode, int lC, const uint32_t* gin, uint32_t* gout,
		const Poly128* rowJ, const TmtOrbit* orb, void* out, hipStream_t stream)
{
	const int lWn = a.lx - (a.lx >= 7 ? 7 : 6);
	if(a.lx < 6 || a.ly < 1 || a.T < 1 || a.T > (1ll << a.ly) || mode < 0 || mode > 2 || (mode == 2 && !a.dis) ||
	   lC < 0 || lC > lWn || !gin || !gout || !out ||
	   ((lC > 0 || a.per > 1 || (a.per == 1 && a.rem > 0)) && (!rowJ || !orb)))
		return hipErrorInvalidValue;
	const dim3 grid((unsigned)(((1ll << (a.ly + lC)) + 255) / 256));
	const bool wide = a.lx >= 7;
	switch(mode * 2 + (wide ? 1 : 0))
	{
		case 0:
			hipLaunchKernelGGL((cbbit_tmt_gen_kernel<uint32_t, 0>), grid, dim3(256), 0, stream, a, lC, gin, gout, rowJ, orb,
			                   (uint32_t*)out);
			break;
		case 1:
			hipLaunchKernelGGL((cbbit_tmt_gen_kernel<uint64_t, 0>), grid, dim3(256), 0, stream, a, lC, gin, gout, rowJ, orb,
			                   (uint64_t*)out);
			break;
		case 2:
			hipLaunchKernelGGL((cbbit_tmt_gen_kernel<uint32_t, 1>), grid, dim3(256), 0, stream, a, lC, gin, gout, rowJ, orb,
			                   (uint32_t*)out);
			break;
		case 3:
			hipLaunchKernelGGL((cbbit_tmt_gen_kernel<uint64_t, 1>), grid, dim3(256), 0, stream, a, lC, gin, gout, rowJ, orb,
			                   (uint64_t*)out);
			break;
		case 4:
			hipLaunchKernelGGL((cbbit_tmt_gen_kernel<uint32_t, 2>), grid, dim3(256), 0, stream, a, lC, gin, gout, rowJ, orb,
			                   (uint32_t*)out);
			break;
		default:
			hipLaunchKernelGGL((cbbit_tmt_gen_kernel<uint64_t, 2>), grid, dim3(256), 0, stream, a, lC, gin, gout, rowJ, orb,
			                   (uint64_t*)out);
			break;
	}
	return hipGetLastError();
}

int cb_tmt_chunks(int lx, int ly, int64_t T)
{
	const int64_t cells = (((1ll << (lx - 2)) + T - 1) / T) << (ly - 2); // the longest walk
	return (int)((cells + CB_CHUNK - 1) / CB_CHUNK);
}

hipError_t launch_tmt_chunk_polys(const TmtOrbit* orb, int64_t T, int chunks, uint64_t drawsPerChunk, Poly128* out,
		hipStream_t stream)
{
	if(!orb || !out || T < 1 || chunks < 1)
		return hipErrorInvalidValue;
	const int64_t lanes = T * chunks;
	hipLaunchKernelGGL(tmt_chunk_polys_kernel, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, stream, orb, T, chunks,
	                   drawsPerChunk, out);
	return hipGetLastError();
}

hipError_t launch_cb_local_tmt_gen(const CbLocalArgs& a, const uint32_t* gin, uint32_t* gout, const Poly128* chunkJ,
		const TmtOrbit* orb, uint32_t* acc, hipStream_t stream)
{
	const int chunks = (a.T >= 1 && a.lx >= 3 && a.ly >= 3) ? cb_tmt_chunks(a.lx, a.ly, a.T) : 0;
	if(a.lx < 3 || a.ly < 3 || a.T < 1 || a.T > (1ll << (a.lx - 2)) || a.pc != (int)((1ll << (a.lx - 2)) / a.T) ||
	   a.rem != (int)((1ll << (a.lx - 2)) % a.T) || !gin || !gout || !acc || (chunks > 1 && (!chunkJ || !orb)))
		return hipErrorInvalidValue;
	const int64_t lanes = a.T * chunks;
	const dim3 grid((unsigned)((lanes + 255) / 256));
	switch((a.dis ? 2 : 0) + (a.signum ? 1 : 0))
	{
		case 0:
			hipLaunchKernelGGL((cb_local_tmt_gen_kernel<0, false>), grid, dim3(256), 0, stream, a, chunks, gin, gout, chunkJ,
			                   orb, acc);
			break;
		case 1:
			hipLaunchKernelGGL((cb_local_tmt_gen_kernel<1, false>), grid, dim3(256), 0, stream, a, chunks, gin, gout, chunkJ,
			                   orb, acc);
			break;
		case 2:
			hipLaunchKernelGGL((cb_local_tmt_gen_kernel<0, true>), grid, dim3(256), 0, stream, a, chunks, gin, gout, chunkJ,
			                   orb, acc);
			break;
		default:
			hipLaunchKernelGGL((cb_local_tmt_gen_kernel<1, true>), grid, dim3(256), 0, stream, a, chunks, gin, gout, chunkJ,
			                   orb, acc);
			break;
	}
	return hipGetLastError();
}

hipError_t launch_tmt_block_polys(const TmtOrbit* orb, int64_t T, int64_t n, uint64_t drawsPerEntry, Poly128* out,
		hipStream_t stream)
{
	if(!orb || !out || T < 1 || n < 1)
		return hipErrorInvalidValue;
	hipLaunchKernelGGL(tmt_block_polys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, orb, T, n,
	                   drawsPerEntry, out);
	return hipGetLastError();
}

hipError_t launch_tmt_seed(uint32_t* g, int64_t T, hipStream_t stream)
{
	if(!g || T < 1)
		return hipErrorInvalidValue;
	hipLaunchKernelGGL(tmt_seed_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, stream, g, T);
	return hipGetLastError();
}

hipError_t launch_dt_phase(const DtPhaseArgs& a, hipStream_t stream)
{
	if(!dt_layout_supported(a.lbx, a.lby) || a.T < 1 || (a.tmt && !a.random_origin))
		return hipErrorInvalidValue;
	const int64_t nb = 1ll << (a.lx - a.lbx + a.ly - a.lby - 2);
	const bool chains = a.tmt && a.T < nb;
	if(chains && (!a.tmtOut || !a.tmtJ || !a.tmtOrb || a.tmtChain < 1))
		return hipErrorInvalidValue;
	const int64_t lanes = chains ? dt_tmt_lanes(nb, a.T, a.tmtChain) : nb;
	const dim3 grid((unsigned)((lanes + DT_WG - 1) / DT_WG));
	switch(a.lbx)
	{
		case 2: return dt_dispatch_y<2>(a, nb, grid, stream);
		case 3: return dt_dispatch_y<3>(a, nb, grid, stream);
		case 4: return dt_dispatch_y<4>(a, nb, grid, stream);
		default: return hipErrorInvalidValue;
	}
}

hipError_t launch_cbbit_half(const CbHalfArgs& a, hipStream_t stream)
{
	if(a.mode == 3 || a.mode == 4)
	{
		// random words / acceptance masks from a buffer (the TinyMT32 workers)
		if(!a.rnd)
			return hipErrorInvalidValue;
		const int64_t threads = (int64_t)1 << (a.ly + a.lG);
		const int block = (int)(threads < 256 ? threads : 256);
		const dim3 grid((unsigned)(threads / block));
		const bool wide = a.lx >= 7;
		if(a.mode == 3)
		{
			if(wide)
				hipLaunchKernelGGL((cbbit_half_kernel<uint64_t, 3>), grid, dim3(block), 0, stream, a);
			else
				hipLaunchKernelGGL((cbbit_half_kernel<uint32_t, 3>), grid, dim3(block), 0, stream, a);
		}
		else
		{
			if(wide)
				hipLaunchKernelGGL((cbbit_half_kernel<uint64_t, 4>), grid, dim3(block), 0, stream, a);
			else
				hipLaunchKernelGGL((cbbit_half_kernel<uint32_t, 4>), grid, dim3(block), 0, stream, a);
		}
		return hipGetLastError();
	}
	const int64_t threads = (int64_t)1 << (a.ly + a.lG);
	const int block = (int)(threads < 256 ? threads : 256);
	const dim3 grid((unsigned)(threads / block));
	const bool wide = a.lx >= 7;
	if(a.mode == 0)
	{
		if(wide)
			hipLaunchKernelGGL((cbbit_half_kernel<uint64_t, 0>), grid, dim3(block), 0, stream, a);
		else
			hipLaunchKernelGGL((cbbit_half_kernel<uint32_t, 0>), grid, dim3(block), 0, stream, a);
	}
	else if(a.mode == 1)
	{
		if(wide)
			hipLaunchKernelGGL((cbbit_half_kernel<uint64_t, 1>), grid, dim3(block), 0, stream, a);
		else
			hipLaunchKernelGGL((cbbit_half_kernel<uint32_t, 1>), grid, dim3(block), 0, stream, a);
	}
	else if(a.mode == 2 && a.dis)
	{
		if(wide)
			hipLaunchKernelGGL((cbbit_half_kernel<uint64_t, 2>), grid, dim3(block), 0, stream, a);
		else
			hipLaunchKernelGGL((cbbit_half_kernel<uint32_t, 2>), grid, dim3(block), 0, stream, a);
	}
	else
		return hipErrorInvalidValue;
	return hipGetLastError();
}

hipError_t launch_row_moments(const uint32_t* lat, int enc, int lx, int ly, const int64_t* hs, double* out,
		hipStream_t stream)
{
	const dim3 grid((unsigned)(((1 << ly) + 3) / 4));
	if(enc == ENC_LOCAL)
		hipLaunchKernelGGL((row_moments_kernel<ENC_LOCAL>), grid, dim3(256), 0, stream, lat, lx, ly, hs, out);
	else
		hipLaunchKernelGGL((row_moments_kernel<ENC_CBBIT>), grid, dim3(256), 0, stream, lat, lx, ly, hs, out);
	return hipGetLastError();
}

int64_t consistency_part_words(int lx, int ly)
{
	const int64_t Y = 1ll << ly;
	return (Y < CK_BAND ? 1 : Y / CK_BAND) << lx;
}

hipError_t launch_consistency(const uint32_t* lat, int enc, int lx, int ly, int* rowOnes, int* colOnes,
		uint16_t* part, unsigned long long* badLoops, hipStream_t stream)
{
	if((enc == ENC_LOCAL ? (lx < 2 || ly < 2) : (lx < 6 || ly < 1)) || lx > 30 || ly > 30 || !rowOnes || !colOnes
	   || !part || !badLoops)
		return hipErrorInvalidValue;
	const int L = 1 << lx, Y = 1 << ly;
	const int CK = L < 64 ? L : 64;
	const int CB = Y < CK_BAND ? Y : CK_BAND;
	const int64_t lanes = (int64_t)(L / CK) * (Y / CB);
	const dim3 grid((unsigned)((lanes + 255) / 256));
	if(enc == ENC_LOCAL)
		hipLaunchKernelGGL((consistency_kernel<ENC_LOCAL>), grid, dim3(256), 0, stream, lat, lx, ly, rowOnes, part,
		                   badLoops);
	else
		hipLaunchKernelGGL((consistency_kernel<ENC_CBBIT>), grid, dim3(256), 0, stream, lat, lx, ly, rowOnes, part,
		                   badLoops);
	hipLaunchKernelGGL(column_sum_kernel, dim3((unsigned)((L + 255) / 256)), dim3(256), 0, stream, part, Y / CB, lx,
	                   colOnes);
	return hipGetLastError();
}

namespace {
// workgroups of row_corr_table_kernel: one per CU (its tables fill the CU's
// LDS), fewer when the lattice has fewer units than 16 per CU
unsigned corr_grid(int64_t units)
{
	int dev = 0, cus = 256;
	if(hipGetDevice(&dev) != hipSuccess ||
	   hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
		cus = 256;
	const int64_t need = (units + CORR_WG / 64 - 1) / (CORR_WG / 64);
	return (unsigned)(need < cus ? need : cus);
}
} // namespace

hipError_t launch_row_stats(const uint32_t* lat, const uint32_t* snap, int enc, int lx, int ly,
		int64_t* out, hipStream_t stream)
{
	const dim3 grid((unsigned)(((1 << ly) + 3) / 4));
	if(enc == ENC_LOCAL)
	{
		if(snap && lx >= 3 && ly >= 2)
			hipLaunchKernelGGL(row_corr_table_kernel<ENC_LOCAL>, dim3(corr_grid(1ll << (ly - 2))), dim3(CORR_WG), 0,
			                   stream, lat, snap, lx, ly, out);
		else if(snap)
			hipLaunchKernelGGL((row_stats_kernel<ENC_LOCAL, true>), grid, dim3(256), 0, stream, lat, snap, lx, ly, out);
		else if(lx >= 3 && ly >= 2)
			// one wave per word row, four word rows per block
			hipLaunchKernelGGL(row_stats_local_kernel, dim3((unsigned)(((1 << (ly - 2)) + 3) / 4)), dim3(256), 0,
			                   stream, lat, lx, ly, out);
		else
			hipLaunchKernelGGL((row_stats_kernel<ENC_LOCAL, false>), grid, dim3(256), 0, stream, lat, snap, lx, ly, out);
	}
	else
	{
		if(snap && lx >= 6)
			hipLaunchKernelGGL(row_corr_table_kernel<ENC_CBBIT>, dim3(corr_grid(1ll << ly)), dim3(CORR_WG), 0,
			                   stream,
			                   lat, snap, lx, ly, out);
		else if(snap)
			hipLaunchKernelGGL((row_stats_kernel<ENC_CBBIT, true>), grid, dim3(256), 0, stream, lat, snap, lx, ly, out);
		else if(lx >= 6)
			// one wave per site row, the 8-site walk table
			hipLaunchKernelGGL(row_stats_cbbit_kernel, grid, dim3(256), 0, stream, lat, lx, ly, out);
		else
			hipLaunchKernelGGL((row_stats_kernel<ENC_CBBIT, false>), grid, dim3(256), 0, stream, lat, snap, lx, ly, out);
	}
	return hipGetLastError();
}

int64_t row_combine_part_words(int ly)
{
	return ((((int64_t)1 << ly) + RC_ROWS - 1) / RC_ROWS) * RC_REC;
}

hipError_t launch_row_combine(const int64_t* rows, int lx, int ly, bool corr, int64_t* hs, int64_t* part,
		int64_t* res, hipStream_t stream)
{
	if(!rows || !part || !res || lx < 0 || ly < 0 || lx > 30 || ly > 30)
		return hipErrorInvalidValue;
	const int64_t nb = (((int64_t)1 << ly) + RC_ROWS - 1) / RC_ROWS;
	hipLaunchKernelGGL(row_part_kernel, dim3((unsigned)nb), dim3(RC_ROWS), 0, stream, rows, lx, ly, corr ? 1 : 0, hs,
	                   part);
	hipLaunchKernelGGL(row_final_kernel, dim3(1), dim3(1024), 0, stream, part, nb, lx, corr ? 1 : 0, res);
	if(hs)
		hipLaunchKernelGGL(row_heights_kernel, dim3((unsigned)((((int64_t)1 << ly) + 255) / 256)), dim3(256), 0,
		                   stream, part, ly, hs);
	return hipGetLastError();
}

// ---- allocation-free roughness -------------------------------------------
// The row statistics, the combine's per-workgroup records and the totals of
// launch_roughness_noalloc live in this file-scope device array instead of
// buffers the library allocates.  In host code `scratch` is only a stand-in
// for the symbol: its device address (one copy per GPU) comes from
// hipGetSymbolAddress with that GPU current, resolved on the first call per
// device and kept in g_scratchDev; kernels and copies get that address only.
namespace {
constexpr int64_t NOALLOC_ROW_WORDS = ((int64_t)1 << NOALLOC_MAX_LY) * ROW_STATS;
constexpr int64_t NOALLOC_PART_WORDS = (((int64_t)1 << NOALLOC_MAX_LY) + RC_ROWS - 1) / RC_ROWS * RC_REC;
constexpr int64_t NOALLOC_WORDS = NOALLOC_ROW_WORDS + NOALLOC_PART_WORDS + 8;
static __device__ __align__(256) int64_t scratch[NOALLOC_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratchMu;
int64_t* g_scratchDev[NOALLOC_MAX_DEVICES] = {};

hipError_t scratch_on_current_device(int64_t** out)
{
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if(e != hipSuccess)
		return e;
	if(dev < 0 || dev >= NOALLOC_MAX_DEVICES)
		return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> lock(g_scratchMu);
	if(!g_scratchDev[dev])
	{
		void* p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess)
			return e;
		g_scratchDev[dev] = static_cast<int64_t*>(p);
	}
	*out = g_scratchDev[dev];
	return hipSuccess;
}
} // namespace

int64_t noalloc_workspace_bytes() { return NOALLOC_WORDS * (int64_t)sizeof(int64_t); }

hipError_t launch_roughness_noalloc(const uint32_t* lat, int enc, int lx, int ly, int64_t* out, hipStream_t stream)
{
	if(!lat || !out || ly > NOALLOC_MAX_LY)
		return hipErrorInvalidValue;
	int64_t* ws = nullptr;
	hipError_t e = scratch_on_current_device(&ws);
	if(e != hipSuccess)
		return e;
	int64_t* rows = ws;
	int64_t* part = ws + NOALLOC_ROW_WORDS;
	int64_t* res = part + NOALLOC_PART_WORDS;
	// every word of rows, part and res that the combine reads is written first
	// by this call's own kernels: nothing in the workspace needs clearing
	e = launch_row_stats(lat, nullptr, enc, lx, ly, rows, stream);
	if(e != hipSuccess)
		return e;
	e = launch_row_combine(rows, lx, ly, false, nullptr, part, res, stream);
	if(e != hipSuccess)
		return e;
	return hipMemcpyAsync(out, res, 3 * sizeof(int64_t), hipMemcpyDeviceToDevice, stream);
}

hipError_t launch_fill_flat(uint32_t* lat, int enc, int lx, int ly, hipStream_t stream)
{
	const uint64_t n = 1ull << (lx + ly - 4);
	const unsigned blocks = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
	if(enc == ENC_LOCAL)
		hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, stream, lat, n, n, 0x33CC33CCu, 0x33CC33CCu);
	else
		hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, stream, lat, n, n >> 1, 0u, 0xFFFFFFFFu);
	return hipGetLastError();
}

hipError_t launch_cb_local_half(const CbLocalArgs& a, hipStream_t stream)
{
	// the band kernel's shape: groups of 2^lG <= 64 columns inside a row of
	// 2^(lx-2) cells, bands of 2^lNB cell rows, at least 2 cell rows
	// (and a lattice below 4 GiB: the band kernel addresses it with 32-bit byte offsets)
	if(a.lx < 3 || a.ly < 3 || a.lx + a.ly > 33 || a.lG < 0 || a.lG > 6 || a.lG > a.lx - 2 || a.lNB < 0 ||
	   a.lNB > a.ly - 2 || !a.edgeU || (a.lG < a.lx - 2 && !a.edgeR))
		return hipErrorInvalidValue;
	const int64_t lanes = 1ll << (a.lx - 2 + a.ly - 2 - a.lNB);
	const unsigned grid = (unsigned)((lanes + 255) / 256);
	if(a.acc)
	{
		if(a.signum)
			hipLaunchKernelGGL((cb_local_band_kernel<1, false, true>), dim3(grid), dim3(256), 0, stream, a);
		else
			hipLaunchKernelGGL((cb_local_band_kernel<0, false, true>), dim3(grid), dim3(256), 0, stream, a);
	}
	else if(a.dis)
	{
		if(a.signum)
			hipLaunchKernelGGL((cb_local_band_kernel<1, true>), dim3(grid), dim3(256), 0, stream, a);
		else
			hipLaunchKernelGGL((cb_local_band_kernel<0, true>), dim3(grid), dim3(256), 0, stream, a);
	}
	else
	{
		if(a.signum)
			hipLaunchKernelGGL((cb_local_band_kernel<1, false>), dim3(grid), dim3(256), 0, stream, a);
		else
			hipLaunchKernelGGL((cb_local_band_kernel<0, false>), dim3(grid), dim3(256), 0, stream, a);
	}
	const int XC = 1 << (a.lx - 2);
	const int64_t nR = (a.lG < a.lx - 2) ? ((int64_t)a.nyb * (XC >> a.lG)) : 0;
	const int64_t nU = (int64_t)(a.nyb >> a.lNB) * XC;
	hipLaunchKernelGGL(cb_local_fix_kernel, dim3((unsigned)((nR + nU + 255) / 256)), dim3(256), 0, stream, a);
	return hipGetLastError();
}

hipError_t launch_batch_dt(const BatchDtArgs& a, hipStream_t stream)
{
	if(a.nstatic < 0 || a.nstatic > a.nlanes)
		return hipErrorInvalidValue;
	if(a.nstatic > 0)
		hipLaunchKernelGGL(batch_dt_kernel<false>, dim3((unsigned)((a.nstatic + DT_WG - 1) / DT_WG)), dim3(DT_WG), 0,
		                   stream, a);
	if(a.nlanes > a.nstatic)
		hipLaunchKernelGGL(batch_dt_kernel<true>, dim3((unsigned)((a.nlanes - a.nstatic + DT_WG - 1) / DT_WG)),
		                   dim3(DT_WG), 0, stream, a);
	return hipGetLastError();
}

hipError_t launch_batch_cbbit(const BatchCbArgs& a, hipStream_t stream)
{
	if(a.nwaves <= 0)
		return hipSuccess;
	hipLaunchKernelGGL(batch_cbbit_kernel, dim3((unsigned)((a.nwaves + 3) / 4)), dim3(256), 0, stream, a);
	return hipGetLastError();
}

hipError_t launch_batch_roughness(const BatchRep* reps, int nreps, const uint2* waves, int64_t nwaves,
		int64_t* scratch, int64_t* out, hipStream_t stream)
{
	if(nwaves > 0)
		hipLaunchKernelGGL(batch_rowstats_kernel, dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, stream, reps, waves, nwaves,
		                   scratch);
	if(nreps > 0)
		hipLaunchKernelGGL(batch_combine_kernel, dim3((unsigned)nreps), dim3(BC_WG), 0, stream, reps, nreps, scratch, out);
	return hipGetLastError();
}

hipError_t launch_valu_probe(uint32_t* out, unsigned blocks, int iters, bool fast, hipStream_t stream)
{
	if(fast)
		hipLaunchKernelGGL(valu_probe_kernel<1>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x9E3779B9u, 0x7F4A7C15u);
	else
		hipLaunchKernelGGL(valu_probe_kernel<0>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x9E3779B9u, 0x7F4A7C15u);
	return hipGetLastError();
}

hipError_t launch_valu_mix(uint32_t* out, unsigned blocks, int iters, int slow, hipStream_t stream)
{
	switch(slow)
	{
		case 0: hipLaunchKernelGGL(valu_mix_kernel<0>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x7F4A7C15u); break;
		case 2: hipLaunchKernelGGL(valu_mix_kernel<2>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x7F4A7C15u); break;
		case 3: hipLaunchKernelGGL(valu_mix_kernel<3>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x7F4A7C15u); break;
		case 4: hipLaunchKernelGGL(valu_mix_kernel<4>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x7F4A7C15u); break;
		case 5: hipLaunchKernelGGL(valu_mix_kernel<5>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x7F4A7C15u); break;
		case 8: hipLaunchKernelGGL(valu_mix_kernel<8>, dim3(blocks), dim3(256), 0, stream, out, iters, 0x7F4A7C15u); break;
		default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, uint64_t bytes, hipStream_t stream)
{
	const uint64_t n = bytes / 16;
	const unsigned blocks = (unsigned)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
	hipLaunchKernelGGL(copy_kernel, dim3(blocks ? blocks : 1), dim3(256), 0, stream, (uint4*)dst, (const uint4*)src, n);
	return hipGetLastError();
}

} // namespace kpz
