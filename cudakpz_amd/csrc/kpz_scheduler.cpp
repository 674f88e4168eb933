#include "kpz_scheduler.hpp"

#include <algorithm>
#include <cstring>
#include <thread>

namespace kpz {

void check(hipError_t e, const char* what)
{
	if(e != hipSuccess)
		throw Error(-2, std::string(what) + ": " + hipGetErrorString(e));
}

DevBuf::~DevBuf()
{
	if(p_)
		(void)hipFree(p_);
}

void DevBuf::alloc(size_t bytes)
{
	if(p_)
	{
		(void)hipFree(p_);
		p_ = nullptr;
		n_ = 0;
	}
	if(bytes == 0)
		return;
	const hipError_t e = hipMalloc(&p_, bytes);
	if(e != hipSuccess)
	{
		p_ = nullptr;
		throw Error(-3, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
	}
	n_ = bytes;
}

// ======================================================================
SchedulerService::SchedulerService(int lx, int ly, Encoding enc, const Params& p)
	: params_(p), size_{lx, ly}, enc_(enc), T_(p.threads), S_(p.skip), device_(p.device)
{
	if(lx < 2 || ly < 2 || lx > 30 || ly > 30)
		throw Error(-1, "system size out of range (log2 sizes 2..30)");
	if(p.rng != RNG_SLCG64 && p.rng != RNG_TINYMT32)
		throw Error(-1, "unknown stream generator (0 = SLCG64, 1 = TinyMT32)");
	if(device_ >= 0)
		check(hipSetDevice(device_), "hipSetDevice");
	else
		check(hipGetDevice(&device_), "hipGetDevice");
	check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
	lat_.alloc(size_.sizeW() * sizeof(uint32_t));
	// row statistics, then the combine's per-workgroup partial sums
	rows_.alloc(((size_t)size_.dimY() * ROW_STATS + (size_t)row_combine_part_words(size_.ly)) * sizeof(int64_t));
	totals_.alloc(8 * sizeof(int64_t));
	master_.reseed(0);
}

SchedulerService::~SchedulerService()
{
	if(stream_)
	{
		(void)hipStreamSynchronize(stream_);
		(void)hipStreamDestroy(stream_);
	}
}

void SchedulerService::initHomogeneous()
{
	check(launch_fill_flat(deviceLattice(), enc_, size_.lx, size_.ly, stream_), "fill_flat");
}

void SchedulerService::supplySeed(uint64_t seed)
{
	master_.reseed(seed32(seed));
	randomize();
}

void SchedulerService::allocStreams(bool twoBuffers)
{
	if(tinyMT())
	{
		tmt_[0].alloc((size_t)T_ * 7 * 4);
		tmt_[1].alloc((size_t)T_ * 7 * 4);
		return;
	}
	st_[0].alloc((size_t)T_ * 8);
	if(twoBuffers)
		st_[1].alloc((size_t)T_ * 8);
}

// SLCG64::randomize, kpz/GPURandom.cu:42-54 (first master draw = low half);
// TinyMT32: RngIface::TinyMT32::randomize (kpz/CPURngIface.h:99-103), one
// master draw per stream as its tinymt32_init seed
void SchedulerService::randomize()
{
	if(tinyMT())
	{
		if(!tmtParams_)
			throw Error(-4, "TinyMT32 streams need their parameter sets first (setTinyMTParams)");
		std::vector<uint32_t> seed((size_t)T_);
		for(auto& v : seed)
			v = (uint32_t)(master_.close_open() * 4294967295.0);
		check(hipMemcpyAsync(tmtIn(), seed.data(), seed.size() * 4, hipMemcpyHostToDevice, stream_), "seeds H2D");
		check(launch_tmt_seed(tmtIn(), T_, stream_), "tmt_seed");
		sync();
		tmtStreamsChanged();
		return;
	}
	const double m1 = master_.close_open();
	const double m2 = master_.close_open();
	uint64_t s = (uint64_t)(m1 * 4294967295.0) ^ ((uint64_t)(m2 * 4294967295.0) << 32);
	std::vector<uint64_t> h((size_t)T_);
	for(int64_t t = 0; t < T_; ++t)
	{
		h[(size_t)t] = s;
		s = s * SLCG_A + SLCG_C;
	}
	setStreams(h.data());
}

bool SchedulerService::setPQ(double p, double q)
{
	if(!(p >= 0.0 && p <= 1.0 && q >= 0.0 && q <= 1.0))
		return false;
	p_ = p;
	q_ = q;
	return true;
}

// SchedulerService::setPQ(p, q, n) (kpz/schedulerService.h:185-191)
bool SchedulerService::setPQ(double p, double q, unsigned n)
{
	if(n >= 4)
		return false;
	if(n == 0)
		return setPQ(p, q);
	if(!(p >= 0.0 && p <= 1.0 && q >= 0.0 && q <= 1.0))
		return false;
	cp_[n] = p;
	cq_[n] = q;
	cset_[n] = true;
	return true;
}

void SchedulerService::classThresholds(uint32_t tp[4], uint32_t tq[4]) const
{
	for(unsigned k = 0; k < 4; ++k)
	{
		tp[k] = thresh(rateP(k));
		tq[k] = thresh(rateQ(k));
	}
}

// SchedulerService::generateDisorder2 (kpz/schedulerService.cpp:54-69): the
// map has the lattice's size and layout; every word takes 16 master draws
void SchedulerService::generateDisorder2(double fill)
{
	std::vector<uint32_t> h((size_t)size_.sizeW());
	for(auto& w : h)
	{
		uint32_t t = 0;
		for(int s = 0; s < 32; s += 2)
			if(master_.close_open() < fill)
				t |= 1u << s;
		w = t;
	}
	setDisorder(h.data());
}

void SchedulerService::setDisorder(const uint32_t* host)
{
	if(!dis_.get())
		dis_.alloc(size_.sizeW() * 4);
	check(hipMemcpyAsync(dis_.get(), host, size_.sizeW() * 4, hipMemcpyHostToDevice, stream_), "disorder H2D");
	sync();
}

void SchedulerService::getDisorder(uint32_t* host)
{
	if(!dis_.get())
		throw Error(-4, "no disorder map (generateDisorder2 / setDisorder first)");
	check(hipMemcpyAsync(host, dis_.get(), size_.sizeW() * 4, hipMemcpyDeviceToHost, stream_), "disorder D2H");
	sync();
}

void SchedulerService::mcsDisorder2SyncRng(int)
{
	throw Error(-1, "mcsDisorder2SyncRng is not available for this scheduler");
}

void SchedulerService::mcsEW(int)
{
	throw Error(-1, "the EW fast path (mcsEW) exists for the DT schedulers only");
}

void SchedulerService::pushSystem(const uint32_t* host)
{
	check(hipMemcpyAsync(deviceLattice(), host, size_.sizeW() * 4, hipMemcpyHostToDevice, stream_), "pushSystem");
	sync();
}

void SchedulerService::popSystem(uint32_t* host)
{
	check(hipMemcpyAsync(host, deviceLattice(), size_.sizeW() * 4, hipMemcpyDeviceToHost, stream_), "popSystem");
	sync();
}

void SchedulerService::copySystem(const SchedulerService& other)
{
	if(other.size_.lx != size_.lx || other.size_.ly != size_.ly || other.enc_ != enc_)
		throw Error(-1, "copySystem: size or encoding differs");
	check(hipStreamSynchronize(other.stream_), "hipStreamSynchronize");
	const uint64_t bytes = size_.sizeW() * 4;
	if(bytes % 16)
		check(hipMemcpyAsync(deviceLattice(), other.deviceLattice(), bytes, hipMemcpyDeviceToDevice, stream_), "copySystem");
	else
		check(launch_copy(deviceLattice(), other.deviceLattice(), bytes, stream_), "copy");
}

void SchedulerService::sync() { check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

void SchedulerService::getStreams(uint64_t* out)
{
	if(tinyMT())
		throw Error(-1, "TinyMT32 streams: use getTinyMT");
	check(hipMemcpyAsync(out, stIn(), (size_t)T_ * 8, hipMemcpyDeviceToHost, stream_), "getStreams");
	sync();
}

void SchedulerService::setStreams(const uint64_t* in)
{
	if(tinyMT())
		throw Error(-1, "TinyMT32 streams: use setTinyMT");
	check(hipMemcpyAsync(stIn(), in, (size_t)T_ * 8, hipMemcpyHostToDevice, stream_), "setStreams");
	sync();
}

void SchedulerService::setTinyMTParams(const uint32_t* mat1, const uint32_t* mat2, const uint32_t* tmat)
{
	if(!tinyMT())
		throw Error(-1, "setTinyMTParams: this scheduler uses SLCG64 streams");
	const size_t T = (size_t)T_;
	const uint32_t* src[3] = {mat1, mat2, tmat};
	for(int b = 0; b < 2; ++b)
		for(int k = 0; k < 3; ++k)
			check(hipMemcpyAsync(tmt_[b].as<uint32_t>() + (4 + k) * T, src[k], T * 4, hipMemcpyHostToDevice, stream_),
					"TinyMT parameters H2D");
	sync();
	tmtParams_ = true;
	tmtStreamsChanged();
}

void SchedulerService::getTinyMT(uint32_t* out)
{
	if(!tinyMT())
		throw Error(-1, "getTinyMT: this scheduler uses SLCG64 streams");
	check(hipMemcpyAsync(out, tmtIn(), (size_t)T_ * 28, hipMemcpyDeviceToHost, stream_), "getTinyMT");
	sync();
}

void SchedulerService::setTinyMT(const uint32_t* in)
{
	if(!tinyMT())
		throw Error(-1, "setTinyMT: this scheduler uses SLCG64 streams");
	for(int b = 0; b < 2; ++b)
		check(hipMemcpyAsync(tmt_[b].get(), in, (size_t)T_ * 28, hipMemcpyHostToDevice, stream_), "setTinyMT");
	sync();
	tmtParams_ = true;
	tmtStreamsChanged();
}

// orbit polynomial of every stream's current state (host threads; the
// elimination takes ~40 us per stream)
const TmtOrbit* SchedulerService::tmtOrbits()
{
	if(orbitValid_)
		return orbit_.as<TmtOrbit>();
	const size_t T = (size_t)T_;
	std::vector<uint32_t> g(T * 7);
	getTinyMT(g.data());
	std::vector<TmtOrbit> o(T);
	const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
	std::vector<std::thread> pool;
	for(unsigned w = 0; w < nth; ++w)
		pool.emplace_back([&, w] {
			for(size_t t = w; t < T; t += nth)
			{
				const uint32_t st[4] = {g[t], g[T + t], g[2 * T + t], g[3 * T + t]};
				o[t] = tmt_orbit(st, g[4 * T + t], g[5 * T + t]);
			}
		});
	for(auto& th : pool)
		th.join();
	orbit_.alloc(T * sizeof(TmtOrbit));
	check(hipMemcpyAsync(orbit_.get(), o.data(), T * sizeof(TmtOrbit), hipMemcpyHostToDevice, stream_), "orbits H2D");
	sync();
	orbitValid_ = true;
	return orbit_.as<TmtOrbit>();
}

const Affine* SchedulerService::upload(std::vector<Affine>& host, DevBuf& buf)
{
	buf.alloc(host.size() * sizeof(Affine));
	check(hipMemcpyAsync(buf.get(), host.data(), host.size() * sizeof(Affine), hipMemcpyHostToDevice, stream_), "upload");
	sync();
	return buf.as<Affine>();
}

// exact totals: per-row partial sums relative to each row's start height
// (row_stats), combined on the device in int128 (row_combine); heights follow
// Correlator::set (kpz/correlator.cpp:37-106).  hs (optional) receives every
// row's start height.
Sums SchedulerService::totals(const uint32_t* snap, int64_t* hs)
{
	check(launch_row_stats(deviceLattice(), snap, enc_, size_.lx, size_.ly, rows_.as<int64_t>(), stream_), "row_stats");
	check(launch_row_combine(rows_.as<int64_t>(), size_.lx, size_.ly, snap != nullptr, hs,
	                         rows_.as<int64_t>() + (size_t)size_.dimY() * ROW_STATS, totals_.as<int64_t>(), stream_),
	      "row_combine");
	int64_t r[7];
	check(hipMemcpyAsync(r, totals_.get(), sizeof r, hipMemcpyDeviceToHost, stream_), "totals D2H");
	sync();
	Sums s;
	s.sumH = r[0];
	s.sumH2 = ((__int128)r[2] << 64) | (uint64_t)r[1];
	if(snap)
	{
		s.sumHs = r[3];
		s.sumHHs = ((__int128)r[5] << 64) | (uint64_t)r[4];
		s.tcs = r[6];
	}
	return s;
}

Sums SchedulerService::roughness() { return totals(nullptr, nullptr); }

// counts of one lattice (device or snapshot), added to out[0..2]
void SchedulerService::consistencyOf(const uint32_t* lat, int64_t out[3])
{
	const size_t nx = (size_t)1 << size_.lx, ny = (size_t)1 << size_.ly;
	const size_t bytes = (ny + nx) * sizeof(int) + sizeof(unsigned long long);
	const size_t partOff = (bytes + 255) & ~(size_t)255; // 16-byte stores into the per-band counts
	const size_t partBytes = (size_t)consistency_part_words(size_.lx, size_.ly) * sizeof(uint16_t);
	if(check_.bytes() < partOff + partBytes)
		check_.alloc(partOff + partBytes);
	int* rowOnes = check_.as<int>();
	int* colOnes = rowOnes + ny;
	unsigned long long* bad = (unsigned long long*)(colOnes + nx);
	uint16_t* part = (uint16_t*)(check_.as<char>() + partOff);
	check(hipMemsetAsync(check_.get(), 0, bytes, stream_), "consistency clear");
	check(launch_consistency(lat, enc_, size_.lx, size_.ly, rowOnes, colOnes, part, bad, stream_), "consistency");
	std::vector<int> h(ny + nx + 2);
	check(hipMemcpyAsync(h.data(), check_.get(), bytes, hipMemcpyDeviceToHost, stream_), "consistency D2H");
	sync();
	unsigned long long nbad;
	std::memcpy(&nbad, h.data() + ny + nx, sizeof nbad);
	for(size_t y = 0; y < ny; ++y)
		out[0] += (size_t)h[y] * 2 != nx;
	for(size_t x = 0; x < nx; ++x)
		out[1] += (size_t)h[ny + x] * 2 != ny;
	out[2] += (int64_t)nbad;
}

bool SchedulerService::consistencyCheck(bool full, int64_t out[3])
{
	out[0] = out[1] = out[2] = 0;
	consistencyOf(deviceLattice(), out);
	if(full)
		for(const Tag& t : tags_)
			consistencyOf(t.snap->as<uint32_t>(), out);
	return !out[0] && !out[1] && !out[2];
}

void SchedulerService::moments(long double out[7])
{
	const size_t ny = (size_t)1 << size_.ly;
	if(!rowStart_.get())
	{
		rowStart_.alloc(ny * 8);
		rowMoments_.alloc(ny * 7 * 8);
	}
	totals(nullptr, rowStart_.as<int64_t>()); // every row's exact start height, on the device
	check(launch_row_moments(deviceLattice(), enc_, size_.lx, size_.ly, rowStart_.as<int64_t>(),
			rowMoments_.as<double>(), stream_), "row_moments");
	std::vector<double> m(ny * 7);
	check(hipMemcpyAsync(m.data(), rowMoments_.get(), m.size() * 8, hipMemcpyDeviceToHost, stream_), "moments D2H");
	sync();
	long double acc[7] = {0, 0, 0, 0, 0, 0, 0};
	for(size_t y = 0; y < ny; ++y)
		for(int k = 0; k < 7; ++k)
			acc[k] += m[y * 7 + (size_t)k];
	const long double N = (long double)size_.sites();
	for(int k = 0; k < 7; ++k)
		out[k] = acc[k] / N;
}

void SchedulerService::correlateTag(unsigned int stop)
{
	Tag t;
	t.snap = std::make_unique<DevBuf>(size_.sizeW() * 4);
	t.mcs = mcs_;
	t.stop = stop;
	const uint64_t bytes = size_.sizeW() * 4;
	if(bytes % 16)
		check(hipMemcpyAsync(t.snap->get(), deviceLattice(), bytes, hipMemcpyDeviceToDevice, stream_), "snapshot");
	else
		check(launch_copy(t.snap->get(), deviceLattice(), bytes, stream_), "snapshot");
	tags_.push_back(std::move(t));
}

Sums SchedulerService::correlate()
{
	if(tags_.empty())
		throw Error(-4, "correlate() needs a correlateTag() snapshot");
	return correlateWith(*tags_.back().snap);
}

std::vector<SchedulerService::TagResult> SchedulerService::correlateAll()
{
	std::vector<TagResult> out;
	for(auto it = tags_.begin(); it != tags_.end();)
	{
		if(mcs_ > it->stop)
		{
			it = tags_.erase(it);
			continue;
		}
		out.push_back(TagResult{it->mcs, correlateWith(*it->snap)});
		++it;
	}
	return out;
}

Sums SchedulerService::correlateWith(const DevBuf& snap) { return totals(snap.as<uint32_t>(), nullptr); }

// ======================================================================
SchedulerDT::SchedulerDT(int lx, int ly, const Params& p, bool randomOrigin)
	: SchedulerService(lx, ly, ENC_LOCAL, p), lbx_(p.layout_x), lby_(p.layout_y), ro_(randomOrigin)
{
	if(!dt_layout_supported(lbx_, lby_))
		throw Error(-1, "DT block layout must be 2..4 x 2..4 (log2 sites)");
	if(lx - lbx_ < 1 || ly - lby_ < 1)
		throw Error(-1, "DT needs an even number (>= 2) of blocks per dimension");
	nb_ = 1ll << (lx - lbx_ + ly - lby_ - 2);
	if(T_ <= 0)
		T_ = nb_;
	if(tinyMT() && !ro_)
		throw Error(-1, "TinyMT32 workers exist for the random-origin DT (SchedulerCPUDTr); "
		                "SchedulerCPUDT draws close_open(), which RngIface::TinyMT32 lacks");
	if(S_ == 0)
		S_ = default_skip((uint64_t)T_);
	step_ = slcg_skip(S_);
	allocStreams(T_ < nb_);
	p_ = 1.0;
	q_ = 0.0;
}

const Affine* SchedulerDT::jumps(int du)
{
	if(T_ >= nb_)
		return nullptr;
	auto it = jumpBuf_.find(du);
	if(it != jumpBuf_.end())
		return it->second->as<Affine>();
	const int64_t nj = (nb_ + T_ - 1) / T_;
	const Affine blk = affine_pow(step_, (uint64_t)du << (lbx_ + lby_));
	std::vector<Affine> h((size_t)nj);
	Affine acc{1, 0};
	for(int64_t j = 0; j < nj; ++j)
	{
		h[(size_t)j] = acc;
		acc = compose(blk, acc);
	}
	auto buf = std::make_unique<DevBuf>();
	const Affine* d = upload(h, *buf);
	jumpBuf_[du] = std::move(buf);
	return d;
}

void SchedulerDT::run(int n, int upd)
{
	if(upd == 2 && !hasDisorder())
		throw Error(-4, "mcsDisorder2SyncRng needs a disorder map (generateDisorder2 / setDisorder)");
	// draws per update: x, y, and the acceptance draw of the pq and disorder rules
	const int du = (upd == 1 || upd == 2) ? 3 : 2;
	DtPhaseArgs a{};
	a.dis = dis_.as<uint32_t>();
	classThresholds(a.tpc, a.tqc);
	a.lat = deviceLattice();
	a.lx = size_.lx;
	a.ly = size_.ly;
	a.lbx = lbx_;
	a.lby = lby_;
	a.T = T_;
	a.tmt = tinyMT() ? tmtIn() : nullptr;
	a.jumps = tinyMT() ? nullptr : jumps(du);
	if(tinyMT() && T_ < nb_)
	{
		// blocks start mid-stream: a lane runs a chain of C consecutive blocks of
		// one worker from a jump to the chain's start (draws per block = du <<
		// (lbx + lby)); chains of 4 where they still leave >= 2^19 lanes (8
		// waves per SIMD), single blocks on smaller lattices
		a.tmtOrb = tmtOrbits();
		const int C = (nb_ >= (1ll << 21) && nb_ >= 16 * T_) ? 4 : 1;
		a.tmtChain = C;
		const uint64_t dpb = (uint64_t)du << (lbx_ + lby_);
		const int64_t lanes = dt_tmt_lanes(nb_, T_, C);
		auto& tab = tmtTab_[(dpb << 4) | (uint64_t)C];
		if(!tab)
		{
			tab = std::make_unique<DevBuf>((size_t)lanes * sizeof(Poly128));
			check(launch_tmt_block_polys(a.tmtOrb, T_, lanes, dpb * (uint64_t)C, tab->as<Poly128>(), stream_),
			      "tmt_block_polys");
		}
		a.tmtJ = tab->as<Poly128>();
	}
	a.step = step_;
	a.blockJ = affine_pow(step_, (uint64_t)du << (lbx_ + lby_));
	a.upd = upd;
	a.random_origin = ro_;
	a.tp = thresh(upd ? p_ : 1.0);
	a.tq = thresh(upd ? q_ : 0.0);
	for(int m = 0; m < n; ++m)
	{
		// SchedulerCPUDT::mcs (kpz/schedulerCPUDT.cpp:65-90) / SchedulerCPUDTr::doMcs
		// (kpz/schedulerCPUDTr.cpp:66-98): domain order, then the origin shift
		int seq[4];
		master_.shuffle4(seq);
		a.sx = a.sy = 0;
		if(ro_)
		{
			a.sx = (int)(unsigned)(master_.close_open() * (double)(1 << lbx_));
			a.sy = (int)(unsigned)(master_.close_open() * (double)(1 << lby_));
		}
		for(int d = 0; d < 4; ++d)
		{
			a.ds = seq[d];
			if(tinyMT())
			{
				a.tmt = tmtIn();
				a.tmtOut = tmtOut();
				check(launch_dt_phase(a, stream_), "dt_phase");
				if(T_ < nb_)
					tmtCur_ ^= 1;
				continue;
			}
			a.st_in = stIn();
			a.st_out = (T_ < nb_) ? stOut() : stIn();
			check(launch_dt_phase(a, stream_), "dt_phase");
			if(T_ < nb_)
				swapStreams();
		}
		++mcs_;
	}
}

void SchedulerDT::mcs(int n) { run(n, 0); }
void SchedulerDT::mcsPQSyncRng(int n) { run(n, 1); }
// SchedulerCPUDTr::mcsDisorder2SyncRng (kpz/schedulerCPUDTr.cpp:115-119)
void SchedulerDT::mcsDisorder2SyncRng(int n) { run(n, 2); }
void SchedulerDT::mcsEW(int n) { run(n, 3); }

// ======================================================================
SchedulerSCABit::SchedulerSCABit(int lx, int ly, const Params& p)
	: SchedulerService(lx, ly, ENC_CBBIT, p)
{
	if(lx < 6 || ly < 1)
		throw Error(-1, "ENC_CBBIT needs log2 width >= 6");
	const int Y = 1 << ly;
	if(T_ <= 0)
		T_ = Y;
	if(T_ > Y)
		throw Error(-1, "SCABit: more workers than rows");
	if(S_ == 0)
		S_ = default_skip((uint64_t)T_);
	step_ = slcg_skip(S_);
	per_ = (int)(Y / T_);
	rem_ = (int)(Y % T_);
	const int lw = (lx >= 7) ? 7 : 6;
	const int lWn = lx - lw;
	lG_ = lWn < 6 ? lWn : 6;
	allocStreams(true);
	p_ = 0.5;
	q_ = 0.0;
}

SchedulerSCABit::Tables& SchedulerSCABit::tables(int mode)
{
	auto it = tab_.find(mode);
	if(it != tab_.end())
		return *it->second;
	const bool wide = size_.lx >= 7;
	const int wb = wide ? 64 : 32;
	const uint64_t du = mode == 0 ? (wide ? 2 : 1) : (uint64_t)wb; // draws per word
	const uint64_t Wn = 1ull << (size_.lx - (wide ? 7 : 6));
	const int G = 1 << lG_;
	auto t = std::make_unique<Tables>();
	std::vector<Affine> row((size_t)per_ + 1), lane((size_t)G);
	const Affine rowStep = affine_pow(step_, du * Wn);
	Affine acc{1, 0};
	for(size_t r = 0; r < row.size(); ++r)
	{
		row[r] = acc;
		acc = compose(rowStep, acc);
	}
	const Affine wordStep = affine_pow(step_, du);
	acc = Affine{1, 0};
	for(int g = 0; g < G; ++g)
	{
		lane[(size_t)g] = acc;
		acc = compose(wordStep, acc);
	}
	upload(row, t->row);
	upload(lane, t->lane);
	t->stepJ = affine_pow(step_, du * (uint64_t)G);
	t->fullBig = affine_pow(rowStep, (uint64_t)per_ + 1);
	t->fullSmall = affine_pow(rowStep, (uint64_t)per_);
	Tables& ref = *t;
	tab_[mode] = std::move(t);
	return ref;
}

void SchedulerSCABit::run(int n, int mode)
{
	if(mode == 2 && !hasDisorder())
		throw Error(-4, "mcsDisorder2SyncRng needs a disorder map (generateDisorder2 / setDisorder)");
	if(tinyMT())
	{
		// TinyMT32 workers: per half-sweep the random part first (one lane per
		// row, a stream jumped to the row's first draw), then the bit-parallel
		// kernel applies it
		const bool wide = size_.lx >= 7;
		const uint64_t WB = wide ? 64 : 32;
		const uint64_t Wn = 1ull << (size_.lx - (wide ? 7 : 6));
		const uint64_t Q = Wn << size_.ly;
		const uint64_t drawsPerRow = Wn * (mode == 0 ? (wide ? 2 : 1) : WB);
		// rows of many draws are split into chunks of >= 4096 draws (more lanes,
		// each jumped to its chunk); a chunk is at least one word
		int lC = 0;
		while(lC < 4 && (drawsPerRow >> (lC + 1)) >= 4096 && (Wn >> (lC + 1)) >= 1)
			++lC;
		const bool split = lC > 0 || per_ > 1 || (per_ == 1 && rem_ > 0); // lanes starting mid-stream
		const TmtOrbit* orb = split ? tmtOrbits() : nullptr;
		const Poly128* rowJ = nullptr;
		if(split)
		{
			auto& tab = tmtTab_[(drawsPerRow << 4) | (uint64_t)lC];
			if(!tab)
			{
				tab = std::make_unique<DevBuf>(((size_t)size_.dimY() << lC) * sizeof(Poly128));
				check(launch_tmt_row_polys(orb, size_.dimY(), lC, per_, rem_, drawsPerRow, tab->as<Poly128>(), stream_),
						"tmt_row_polys");
			}
			rowJ = tab->as<Poly128>();
		}
		const size_t need = (size_t)(mode == 0 ? Q : 2 * Q) * (WB / 8);
		if(rndBuf_.bytes() < need)
			rndBuf_.alloc(need);
		CbHalfArgs a{};
		a.lat = deviceLattice();
		a.lx = size_.lx;
		a.ly = size_.ly;
		a.mode = mode == 0 ? 3 : 4;
		a.T = T_;
		a.per = per_;
		a.rem = rem_;
		a.tp = thresh(p_);
		a.tq = thresh(q_);
		a.dis = dis_.as<uint32_t>();
		a.tp1 = thresh(rateP(1));
		a.tq1 = thresh(rateQ(1));
		a.lG = lG_;
		a.rnd = rndBuf_.get();
		for(int m = 0; m < n; ++m)
		{
			for(int s = 0; s < 2; ++s)
			{
				a.signum = s;
				check(launch_cbbit_tmt_gen(a, mode, lC, tmtIn(), tmtOut(), rowJ, orb, rndBuf_.get(), stream_),
						"cbbit_tmt_gen");
				check(launch_cbbit_half(a, stream_), "cbbit_half");
				tmtCur_ ^= 1;
			}
			++mcs_;
		}
		return;
	}
	Tables& tb = tables(mode == 2 ? 1 : mode); // one draw per site in both per-site modes
	CbHalfArgs a{};
	a.dis = dis_.as<uint32_t>();
	a.tp1 = thresh23(rateP(1));
	a.tq1 = thresh23(rateQ(1));
	a.lat = deviceLattice();
	a.lx = size_.lx;
	a.ly = size_.ly;
	a.mode = mode;
	a.T = T_;
	a.per = per_;
	a.rem = rem_;
	a.rowJ = tb.row.as<Affine>();
	a.laneJ = tb.lane.as<Affine>();
	a.stepJ = tb.stepJ;
	a.step = step_;
	a.fullBig = tb.fullBig;
	a.fullSmall = tb.fullSmall;
	a.tp = thresh23(p_);
	a.tq = thresh23(q_);
	a.lG = lG_;
	for(int m = 0; m < n; ++m)
	{
		// SchedulerCPUCBBit::doMcs (kpz/schedulerCPUCBBit.cpp:64-78)
		for(int s = 0; s < 2; ++s)
		{
			a.signum = s;
			a.st_in = stIn();
			a.st_out = stOut();
			check(launch_cbbit_half(a, stream_), "cbbit_half");
			swapStreams();
		}
		++mcs_;
	}
}

// SchedulerSCABit::mcs -> mcsSyncRng sets p=.5, q=0 (kpz/schedulerSCABit.h:61-67)
void SchedulerSCABit::mcs(int n)
{
	setPQ(0.5, 0.0);
	run(n, 0);
}

// SchedulerCPUCBBit::mcsPQSyncRng (kpz/schedulerCPUCBBit.cpp:95-102)
void SchedulerSCABit::mcsPQSyncRng(int n) { run(n, (q_ == 0.0 && p_ == 0.5) ? 0 : 1); }
// SchedulerCPUCBBit::mcsDisorder2SyncRng (kpz/schedulerCPUCBBit.cpp:105-109): the
// class is the site's x-slope bit of the map, so classes 0 and 1 exist
void SchedulerSCABit::mcsDisorder2SyncRng(int n) { run(n, 2); }

// ======================================================================
SchedulerSCA::SchedulerSCA(int lx, int ly, const Params& p) : SchedulerService(lx, ly, ENC_LOCAL, p)
{
	if(lx < 3 || ly < 3)
		throw Error(-1, "SCA: lattice sides must be at least 8 sites");
	const int dimXW = 1 << (lx - 2);
	const int nyb = 1 << (ly - 2);
	if(T_ <= 0)
		T_ = dimXW;
	if(T_ > dimXW)
		throw Error(-1, "SCA: more workers than 4-site columns");
	if(S_ == 0)
		S_ = default_skip((uint64_t)T_);
	step_ = slcg_skip(S_);
	allocStreams(true);
	// worker strips (kpz/schedulerCPUCB.cpp:46-62) in cells of 4 sites
	const int pc = (int)(dimXW / T_), rem = (int)(dimXW % T_);
	args_.lx = lx;
	args_.ly = ly;
	args_.pc = pc;
	args_.rem = rem;
	args_.nyb = nyb;
	args_.T = T_;
	p_ = 0.95;
	q_ = 0.0;
	std::vector<int> V;
	auto addV = [&](int v) { // v = 0: an empty first half (one-cell strips)
		if(std::find(V.begin(), V.end(), v) == V.end())
			V.push_back(v);
	};
	for(int w : {pc, pc + 1})
	{
		if((w == pc && rem == T_) || (w == pc + 1 && rem == 0))
			continue;
		addV(w / 2);
		addV(w - w / 2);
	}
	if(V.empty())
		V.push_back(1);
	const Affine cell = affine_pow(step_, 8);
	std::vector<Affine> rowJ(V.size() * (size_t)nyb), baseJ(V.size());
	int maxV = 0;
	for(size_t k = 0; k < V.size(); ++k)
	{
		const Affine rowStep = affine_pow(cell, (uint64_t)V[k]);
		Affine acc{1, 0};
		for(int yb = 0; yb < nyb; ++yb)
		{
			rowJ[k * (size_t)nyb + (size_t)yb] = acc;
			acc = compose(rowStep, acc);
		}
		baseJ[k] = acc; // J(nyb * V * 8)
		maxV = std::max(maxV, V[k]);
	}
	std::vector<Affine> cellJ((size_t)maxV + 1);
	Affine acc{1, 0};
	for(auto& c : cellJ)
	{
		c = acc;
		acc = compose(cell, acc);
	}
	args_.nV = (int)V.size();
	for(size_t k = 0; k < V.size(); ++k)
		args_.V[k] = V[k];
	args_.rowJ = upload(rowJ, rowJ_);
	args_.baseJ = upload(baseJ, baseJ_);
	args_.cellJ = upload(cellJ, cellJ_);
	args_.fullBig = affine_pow(cell, (uint64_t)nyb * (uint64_t)(pc + 1));
	args_.fullSmall = affine_pow(cell, (uint64_t)nyb * (uint64_t)pc);
	args_.step = step_;
	// band kernel shape: up to 64 cell columns per lane group, 64 cell rows per band
	args_.lG = std::min(lx - 2, 6);
	args_.lNB = std::min(ly - 2, 6);
	const size_t groups = (size_t)dimXW >> args_.lG;
	edgeR_.alloc(std::max<size_t>(1, (size_t)nyb * groups) * 4);
	edgeU_.alloc(((size_t)nyb >> args_.lNB) * (size_t)dimXW * 4);
	args_.edgeR = edgeR_.as<uint32_t>();
	args_.edgeU = edgeU_.as<uint32_t>();
}

void SchedulerSCA::mcs(int n)
{
	setPQ(0.95, 0.0);
	mcsPQSyncRng(n);
}

void SchedulerSCA::mcsPQSyncRng(int n) { run(n, false); }

// SchedulerCPUCB::mcsDisorder2SyncRng (kpz/schedulerCPUCB.cpp:103-108)
void SchedulerSCA::mcsDisorder2SyncRng(int n)
{
	if(!hasDisorder())
		throw Error(-4, "mcsDisorder2SyncRng needs a disorder map (generateDisorder2 / setDisorder)");
	run(n, true);
}

void SchedulerSCA::run(int n, bool disorder)
{
	CbLocalArgs a = args_;
	a.lat = deviceLattice();
	a.tp = thresh(p_);
	a.tq = thresh(q_);
	a.dis = disorder ? dis_.as<uint32_t>() : nullptr;
	classThresholds(a.tpc, a.tqc);
	if(tinyMT())
	{
		// TinyMT32 workers: per half-sweep the acceptance words first (one lane
		// per chunk of a worker's walk, its stream jumped to the chunk), then
		// the band kernel applies them
		const int chunks = cb_tmt_chunks(size_.lx, size_.ly, T_);
		const TmtOrbit* orb = chunks > 1 ? tmtOrbits() : nullptr;
		const Poly128* chunkJ = nullptr;
		if(chunks > 1)
		{
			auto& tab = tmtTab_[CB_TMT_CHUNK_DRAWS];
			if(!tab)
			{
				tab = std::make_unique<DevBuf>((size_t)T_ * chunks * sizeof(Poly128));
				check(launch_tmt_chunk_polys(orb, T_, chunks, CB_TMT_CHUNK_DRAWS, tab->as<Poly128>(), stream_),
						"tmt_chunk_polys");
			}
			chunkJ = tab->as<Poly128>();
		}
		if(rndBuf_.bytes() < size_.sizeW() * 4)
			rndBuf_.alloc(size_.sizeW() * 4);
		CbLocalArgs b = a;
		b.acc = rndBuf_.as<uint32_t>();
		for(int m = 0; m < n; ++m)
		{
			for(int s = 0; s < 2; ++s)
			{
				a.signum = b.signum = s;
				check(launch_cb_local_tmt_gen(a, tmtIn(), tmtOut(), chunkJ, orb, rndBuf_.as<uint32_t>(), stream_),
						"cb_local_tmt_gen");
				check(launch_cb_local_half(b, stream_), "cb_local_half");
				tmtCur_ ^= 1;
			}
			++mcs_;
		}
		return;
	}
	for(int m = 0; m < n; ++m)
	{
		for(int s = 0; s < 2; ++s)
		{
			a.signum = s;
			a.st_in = stIn();
			a.st_out = stOut();
			check(launch_cb_local_half(a, stream_), "cb_local_half");
			swapStreams();
		}
		++mcs_;
	}
}

std::unique_ptr<SchedulerService> makeScheduler(int kind, int lx, int ly, const Params& p)
{
	std::unique_ptr<SchedulerService> s;
	switch(kind)
	{
		case KIND_DT: s = std::make_unique<SchedulerDT>(lx, ly, p, false); break;
		case KIND_DTR: s = std::make_unique<SchedulerDT>(lx, ly, p, true); break;
		case KIND_SCABIT: s = std::make_unique<SchedulerSCABit>(lx, ly, p); break;
		case KIND_SCA: s = std::make_unique<SchedulerSCA>(lx, ly, p); break;
		default: throw Error(-1, "unknown scheduler kind");
	}
	s->kind_ = kind;
	return s;
}

} // namespace kpz
