#include "kpz_server.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace kpz {

namespace {
// log2 of the most lanes that share one SCABit row of a job
constexpr int SERVE_MAX_LANES_LOG2 = 3;

int ceilLog2(size_t v)
{
	int l = 0;
	while((1ull << l) < v)
		++l;
	return l;
}
int floorLog2(size_t v)
{
	int l = 0;
	while((1ull << (l + 1)) <= v)
		++l;
	return l;
}
} // namespace

// ---------------------------------------------------------------- buddy arena
void BuddyArena::init(size_t bytes, size_t minBlock)
{
	minOrder_ = ceilLog2(minBlock < 256 ? 256 : minBlock);
	maxOrder_ = floorLog2(bytes);
	if(maxOrder_ < minOrder_)
		throw Error(-1, "HBM budget smaller than one block");
	cap_ = 1ull << maxOrder_;
	buf_.alloc(cap_);
	free_.assign((size_t)maxOrder_ + 1, {});
	free_[(size_t)maxOrder_].push_back(0);
	live_.clear();
	used_ = 0;
}

bool BuddyArena::alloc(size_t bytes, uint64_t& off)
{
	const int o = std::max(minOrder_, ceilLog2(bytes));
	if(o > maxOrder_)
		return false;
	int k = o;
	while(k <= maxOrder_ && free_[(size_t)k].empty())
		++k;
	if(k > maxOrder_)
		return false;
	const uint64_t b = free_[(size_t)k].back();
	free_[(size_t)k].pop_back();
	while(k > o)
	{
		--k;
		free_[(size_t)k].push_back(b + (1ull << k));
	}
	live_[b] = o;
	used_ += 1ull << o;
	off = b;
	return true;
}

void BuddyArena::release(uint64_t off)
{
	auto it = live_.find(off);
	if(it == live_.end())
		return;
	int o = it->second;
	live_.erase(it);
	used_ -= 1ull << o;
	uint64_t b = off;
	while(o < maxOrder_)
	{
		const uint64_t buddy = b ^ (1ull << o);
		auto& fl = free_[(size_t)o];
		auto f = std::find(fl.begin(), fl.end(), buddy);
		if(f == fl.end())
			break;
		fl.erase(f);
		b = std::min(b, buddy);
		++o;
	}
	free_[(size_t)o].push_back(b);
}

// ---------------------------------------------------------------- pinned staging
PinnedStage::~PinnedStage()
{
	for(auto& b : blocks_)
		(void)hipHostFree(b.p);
}

void PinnedStage::reset()
{
	for(auto& b : blocks_)
		b.used = 0;
	cur_ = 0;
}

void* PinnedStage::take(size_t bytes)
{
	bytes = (bytes + 255) & ~(size_t)255;
	for(; cur_ < blocks_.size(); ++cur_)
	{
		Block& b = blocks_[cur_];
		if(b.used + bytes <= b.size)
		{
			void* p = b.p + b.used;
			b.used += bytes;
			return p;
		}
	}
	const size_t sz = std::max<size_t>(bytes, blocks_.empty() ? ((size_t)4 << 20) : 2 * blocks_.back().size);
	Block b{nullptr, sz, bytes};
	check(hipHostMalloc((void**)&b.p, sz, hipHostMallocDefault), "hipHostMalloc");
	blocks_.push_back(b);
	cur_ = blocks_.size() - 1;
	return b.p;
}

// ---------------------------------------------------------------- server
Server::Server(size_t hbmBudget, int maxSlots, int maxTickMcs, int device)
	: maxSlots_(maxSlots), maxTick_(maxTickMcs < 1 ? 1 : maxTickMcs)
{
	if(maxSlots_ < 1)
		throw Error(-1, "need at least one slot");
	if(device >= 0)
		check(hipSetDevice(device), "hipSetDevice");
	check(hipGetDevice(&device_), "hipGetDevice");
	check(hipStreamCreateWithFlags(&s0_, hipStreamNonBlocking), "hipStreamCreate");
	check(hipStreamCreateWithFlags(&s1_, hipStreamNonBlocking), "hipStreamCreate");
	check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
	check(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "hipEventCreate");
	arena_.init(hbmBudget, 4096);
	slots_.assign((size_t)maxSlots_, nullptr);
	hostReps_.assign((size_t)maxSlots_, BatchRep{});
	repTable_.alloc(sizeof(BatchRep) * (size_t)maxSlots_);
	identity_.alloc(sizeof(Affine));
	const Affine id{1, 0};
	check(hipMemcpy(identity_.get(), &id, sizeof id, hipMemcpyHostToDevice), "identity");
	orderBuf_.alloc((size_t)maxTick_ * (size_t)maxSlots_);
	shiftBuf_.alloc((size_t)maxTick_ * (size_t)maxSlots_ * 2);
	limitBuf_.alloc((size_t)maxSlots_ * sizeof(int));
	for(Tick& t : ticks_)
	{
		check(hipEventCreate(&t.begin), "hipEventCreate");
		check(hipEventCreate(&t.end), "hipEventCreate");
	}
	check(hipEventCreate(&first_), "hipEventCreate");
	t0_ = std::chrono::steady_clock::now();
}

Server::~Server()
{
	if(s0_)
		(void)hipStreamSynchronize(s0_);
	if(s1_)
		(void)hipStreamSynchronize(s1_);
	for(Job* j : queue_)
		delete j;
	for(Job* j : slots_)
		delete j;
	for(Tick& t : ticks_)
	{
		if(t.begin)
			(void)hipEventDestroy(t.begin);
		if(t.end)
			(void)hipEventDestroy(t.end);
	}
	if(first_)
		(void)hipEventDestroy(first_);
	if(fork_)
		(void)hipEventDestroy(fork_);
	if(join_)
		(void)hipEventDestroy(join_);
	if(s1_)
		(void)hipStreamDestroy(s1_);
	if(s0_)
		(void)hipStreamDestroy(s0_);
}

double Server::now() const
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

void Server::busy(double& gpuSeconds, double& spanSeconds) const
{
	gpuSeconds = gpuMs_ * 1e-3;
	spanSeconds = spanMs_ * 1e-3;
}

// device buffers grow geometrically; a tick may still read the old one
void Server::ensure(DevBuf& b, size_t bytes)
{
	if(b.bytes() < bytes)
	{
		check(hipStreamSynchronize(s0_), "grow sync"); // s1_'s work is joined into s0_
		b.alloc(std::max(bytes, 2 * b.bytes()));
	}
}

size_t Server::footprint(const JobSpec& s, size_t* parts)
{
	parts[0] = 1ull << (s.lx + s.ly - 2);
	if(s.kind == KIND_SCABIT)
	{
		parts[1] = 2 * 8 * (1ull << s.ly);
		parts[2] = 64 * sizeof(Affine);
	}
	else
	{
		parts[1] = 8 * (1ull << (s.lx + s.ly - 10));
		parts[2] = 0;
	}
	return parts[0] + parts[1] + parts[2];
}

uint64_t Server::submit(const JobSpec& j)
{
	const bool dt = j.kind == KIND_DT || j.kind == KIND_DTR;
	if(!(dt || j.kind == KIND_SCABIT))
		throw Error(-1, "unknown job kind");
	if(dt && (j.lx < 5 || j.ly < 5 || j.lx > 24 || j.ly > 24))
		throw Error(-1, "DT jobs need 32..2^24 sites per side");
	if(!dt && (j.lx < 6 || j.ly < 6 || j.lx > 24 || j.ly > 24))
		throw Error(-1, "SCABit jobs need 64..2^24 sites per side");
	if(j.mcs < 1 || !(j.p >= 0 && j.p <= 1 && j.q >= 0 && j.q <= 1))
		throw Error(-1, "bad job parameters");
	size_t parts[3];
	footprint(j, parts);
	if((1ull << ceilLog2(parts[0])) > arena_.capacity())
		throw Error(-1, "job larger than the HBM budget");
	Job* job = new Job;
	job->id = nextId_++;
	job->spec = j;
	job->res.id = job->id;
	job->res.spec = j;
	job->res.tSubmit = now();
	queue_.push_back(job);
	return job->id;
}

// reserve the job's memory, seed it and queue its set-up on the server
// stream (host data staged in pinned memory: nothing here waits for the GPU)
bool Server::admit(Job& j, PinnedStage& stage)
{
	size_t parts[3];
	footprint(j.spec, parts);
	int k = 0;
	for(; k < 3; ++k)
		if(parts[k] && !arena_.alloc(parts[k], j.off[k]))
			break;
	if(k < 3)
	{
		for(int u = 0; u < k; ++u)
			if(parts[u])
				arena_.release(j.off[u]);
		return false;
	}
	const JobSpec& s = j.spec;
	BatchRep& d = j.dev;
	d = BatchRep{};
	d.lx = s.lx;
	d.ly = s.ly;
	char* base = arena_.base();
	d.lat = (uint32_t*)(base + j.off[0]);
	if(s.kind == KIND_SCABIT)
	{
		d.kind = BK_SCABIT;
		d.mode = (s.p == 0.5 && s.q == 0.0) ? 0 : 1;
		d.wide = s.lx >= 7 ? 1 : 0;
		// at most 8 lanes per row: a lane then walks >= 8 words of a wide row,
		// so its stream set-up (dependent loads of the job table, the stream
		// and its jumps) is spread over >= 512 sites (the choice the batch
		// autotuner makes on the 4096-replica mix)
		d.lG = std::min(s.lx - (d.wide ? 7 : 6), SERVE_MAX_LANES_LOG2);
		j.T = 1ll << s.ly;
	}
	else
	{
		d.kind = s.kind == KIND_DT ? BK_DT : BK_DTR;
		d.mode = (s.p == 1.0 && s.q == 0.0) ? 0 : 1;
		j.T = 1ll << (s.lx + s.ly - 10);
	}
	d.tp = thresh23((d.kind != BK_SCABIT && d.mode == 0) ? 1.0 : s.p);
	d.tq = thresh23((d.kind != BK_SCABIT && d.mode == 0) ? 0.0 : s.q);
	d.step = slcg_skip(default_skip((uint64_t)j.T));
	d.st[0] = (uint64_t*)(base + j.off[1]);
	d.st[1] = (d.kind == BK_SCABIT) ? d.st[0] + j.T : d.st[0];
	d.rowJ = identity_.as<Affine>();
	d.laneJ = identity_.as<Affine>();
	// seeding: reSeed + SLCG64::randomize, exactly as a stand-alone scheduler
	j.master.reseed(seed32(s.seed));
	const double m1 = j.master.close_open();
	const double m2 = j.master.close_open();
	uint64_t x = (uint64_t)(m1 * 4294967295.0) ^ ((uint64_t)(m2 * 4294967295.0) << 32);
	uint64_t* st = (uint64_t*)stage.take((size_t)j.T * 8);
	for(int64_t t = 0; t < j.T; ++t)
	{
		st[t] = x;
		x = x * SLCG_A + SLCG_C;
	}
	check(hipMemcpyAsync(d.st[0], st, (size_t)j.T * 8, hipMemcpyHostToDevice, s0_), "job streams");
	if(d.kind == BK_SCABIT)
	{
		const uint64_t du = d.mode == 0 ? (d.wide ? 2 : 1) : (d.wide ? 64 : 32);
		const uint64_t Wn = 1ull << (s.lx - (d.wide ? 7 : 6));
		const int G = 1 << d.lG;
		Affine* lane = (Affine*)stage.take((size_t)G * sizeof(Affine));
		const Affine word = affine_pow(d.step, du);
		Affine acc{1, 0};
		for(int g = 0; g < G; ++g)
		{
			lane[g] = acc;
			acc = compose(word, acc);
		}
		d.laneJ = (const Affine*)(base + j.off[2]);
		check(hipMemcpyAsync((void*)d.laneJ, lane, (size_t)G * sizeof(Affine), hipMemcpyHostToDevice, s0_), "lane table");
		d.stepJ = affine_pow(d.step, du * (uint64_t)G);
		d.full = affine_pow(d.step, du * Wn);
	}
	else
		d.full = affine_pow(d.step, (d.mode ? 3u : 2u) * 256u); // one 16 x 16 block
	check(launch_fill_flat(d.lat, d.kind == BK_SCABIT ? ENC_CBBIT : ENC_LOCAL, s.lx, s.ly, s0_), "fill");
	// sample times (kpz/sampling.h:84-100) and the final MCS
	j.samples.clear();
	for(double t = 0;;)
	{
		t = std::ceil(std::exp(std::log(t + 10.0) + 0.1));
		if(t >= (double)s.mcs)
			break;
		j.samples.push_back((uint32_t)t);
	}
	j.samples.push_back((uint32_t)s.mcs);
	j.nextPlan = 0;
	j.planned = 0;
	j.done = 0;
	j.res.tStart = now();
	return true;
}

// device tables of the jobs that still have MCS to run, per launch chain: DT
// lanes (one per block) grouped by (kind, mode), SCABit waves (rows).  Each
// running job goes to the chain with fewer sites so far, in slot order.
void Server::rebuild(PinnedStage& stage)
{
	std::vector<uint2> lanes, waves;
	auto runs = [](const Job* j) { return j && j->planned < (uint32_t)j->spec.mcs; };
	uint64_t load[2] = {0, 0};
	for(Job* j : slots_)
		if(runs(j))
		{
			j->group = load[1] < load[0] ? 1 : 0;
			load[j->group] += 1ull << (j->spec.lx + j->spec.ly);
		}
	for(int g = 0; g < 2; ++g)
	{
		laneOff_[g] = (int64_t)lanes.size();
		for(int kind : {BK_DT, BK_DTR})
		{
			if(kind == BK_DTR)
				nStatic_[g] = (int64_t)lanes.size() - laneOff_[g];
			for(int mode : {0, 1})
			{
				for(int sl = 0; sl < maxSlots_; ++sl)
				{
					Job* j = slots_[(size_t)sl];
					if(!runs(j) || j->group != g || j->dev.kind != kind || j->dev.mode != mode)
						continue;
					for(int64_t b = 0; b < j->T; ++b)
						lanes.push_back(make_uint2((unsigned)sl, (unsigned)b));
				}
				while(lanes.size() % 64)
					lanes.push_back(make_uint2(0xFFFFFFFFu, 0));
			}
		}
		nLanes_[g] = (int64_t)lanes.size() - laneOff_[g];
		waveOff_[g] = (int64_t)waves.size();
		for(int sl = 0; sl < maxSlots_; ++sl)
		{
			Job* j = slots_[(size_t)sl];
			if(!runs(j) || j->group != g || j->dev.kind != BK_SCABIT)
				continue;
			const int rpw = 64 >> j->dev.lG;
			for(int y = 0; y < (1 << j->dev.ly); y += rpw)
				waves.push_back(make_uint2((unsigned)sl, (unsigned)y));
		}
		nWaves_[g] = (int64_t)waves.size() - waveOff_[g];
	}
	for(int sl = 0; sl < maxSlots_; ++sl)
	{
		Job* j = slots_[(size_t)sl];
		hostReps_[(size_t)sl] = runs(j) ? j->dev : BatchRep{};
		if(j)
			j->inTable = runs(j);
	}
	if(!lanes.empty())
		ensure(laneTable_, lanes.size() * sizeof(uint2));
	if(!waves.empty())
		ensure(waveTable_, waves.size() * sizeof(uint2));
	auto put = [&](DevBuf& dst, const void* src, size_t bytes, const char* what) {
		void* p = stage.take(bytes);
		std::memcpy(p, src, bytes);
		check(hipMemcpyAsync(dst.get(), p, bytes, hipMemcpyHostToDevice, s0_), what);
	};
	put(repTable_, hostReps_.data(), sizeof(BatchRep) * hostReps_.size(), "reps");
	if(!lanes.empty())
		put(laneTable_, lanes.data(), lanes.size() * sizeof(uint2), "lanes");
	if(!waves.empty())
		put(waveTable_, waves.data(), waves.size() * sizeof(uint2), "waves");
	dirty_ = false;
}

// plan the next tick and queue all of its device work; nothing here waits
// for the tick that may still be running (all work is ordered on s0_, into
// which each tick joins its s1_ chain)
bool Server::enqueue(Tick& t)
{
	t.stage.reset();
	t.plan.clear();
	t.ran.clear();
	t.live = false;
	if(nActive_ == 0 && queue_.empty())
		return false;
	if(!started_)
	{
		check(hipEventRecord(first_, s0_), "first event");
		started_ = true;
	}
	check(hipEventRecord(t.begin, s0_), "tick begin");
	t.live = true;

	// admission: FIFO with bounded backfill by HBM footprint, into free slots
	int looked = 0;
	for(auto it = queue_.begin(); it != queue_.end() && looked < 64; ++looked)
	{
		int sl = -1;
		for(int k = 0; k < maxSlots_; ++k)
			if(!slots_[(size_t)k])
			{
				sl = k;
				break;
			}
		if(sl < 0)
			break;
		Job* j = *it;
		if(admit(*j, t.stage))
		{
			j->slot = sl;
			slots_[(size_t)sl] = j;
			++nActive_;
			dirty_ = true;
			it = queue_.erase(it);
		}
		else
			++it;
	}

	// fixed-length tick: job j runs lim_j = min(maxTick, MCS left) MCS; the
	// kernels skip a job once m >= lim_j
	const size_t R = (size_t)maxSlots_;
	int* limit = (int*)t.stage.take(R * sizeof(int));
	std::fill(limit, limit + R, 0);
	int k = 0;
	for(Job* j : slots_)
		if(j && j->planned < (uint32_t)j->spec.mcs)
		{
			const int lim = std::min<int>(maxTick_, j->spec.mcs - (int)j->planned);
			limit[j->slot] = lim;
			k = std::max(k, lim);
			if(!j->inTable)
				dirty_ = true;
		}
	if(k == 0)
	{
		check(hipEventRecord(t.end, s0_), "tick end");
		return true;
	}
	if(dirty_)
		rebuild(t.stage);

	// master draws of the DT jobs, per MCS of the tick (domain order, DTr shift)
	uint8_t* order = (uint8_t*)t.stage.take((size_t)k * R);
	uint16_t* shift = (uint16_t*)t.stage.take((size_t)k * R * 2);
	std::memset(order, 0, (size_t)k * R);
	std::memset(shift, 0, (size_t)k * R * 2);
	for(int m = 0; m < k; ++m)
		for(Job* j : slots_)
		{
			if(!j || j->dev.kind == BK_SCABIT || m >= limit[j->slot])
				continue;
			int seq[4];
			j->master.shuffle4(seq);
			order[(size_t)m * R + (size_t)j->slot] = (uint8_t)(seq[0] | (seq[1] << 2) | (seq[2] << 4) | (seq[3] << 6));
			uint16_t sh = 0;
			if(j->dev.kind == BK_DTR)
			{
				const unsigned sx = (unsigned)(j->master.close_open() * 16.0);
				const unsigned sy = (unsigned)(j->master.close_open() * 16.0);
				sh = (uint16_t)(sx | (sy << 8));
			}
			shift[(size_t)m * R + (size_t)j->slot] = sh;
		}
	check(hipMemcpyAsync(orderBuf_.get(), order, (size_t)k * R, hipMemcpyHostToDevice, s0_), "order");
	check(hipMemcpyAsync(shiftBuf_.get(), shift, (size_t)k * R * 2, hipMemcpyHostToDevice, s0_), "shift");
	check(hipMemcpyAsync(limitBuf_.get(), limit, R * sizeof(int), hipMemcpyHostToDevice, s0_), "limit");

	// sample plan: (m, job) for every sample time in (planned, planned + lim]
	for(Job* j : slots_)
		if(j && limit[j->slot] > 0)
			for(size_t s = j->nextPlan; s < j->samples.size() && j->samples[s] <= j->planned + (uint32_t)limit[j->slot]; ++s)
				t.plan.push_back({(int)(j->samples[s] - j->planned) - 1, j});
	std::stable_sort(t.plan.begin(), t.plan.end(), [](const Sample& a, const Sample& b) {
		return a.m != b.m ? a.m < b.m : a.job->group < b.job->group;
	});
	const size_t np = t.plan.size();
	// row-statistics waves (64 / batch_row_lanes(lx) rows each) of every sample
	std::vector<int64_t> rowStart(np + 1);
	std::vector<size_t> groupStart; // plan index where each (m, chain) group starts
	int64_t nrows = 0;
	for(size_t i = 0; i < np; ++i)
	{
		rowStart[i] = nrows;
		const int rpw = 64 / batch_row_lanes(t.plan[i].job->spec.lx);
		nrows += ((1 << t.plan[i].job->spec.ly) + rpw - 1) / rpw;
	}
	rowStart[np] = nrows;
	BatchRep* sreps = (BatchRep*)t.stage.take(std::max<size_t>(np, 1) * sizeof(BatchRep));
	uint2* srows = (uint2*)t.stage.take((size_t)std::max<int64_t>(nrows, 1) * sizeof(uint2));
	int64_t rb = 0, maxGroupRows = 0;
	for(size_t i = 0; i < np; ++i)
	{
		if(i == 0 || t.plan[i].m != t.plan[i - 1].m || t.plan[i].job->group != t.plan[i - 1].job->group)
		{
			groupStart.push_back(i);
			rb = 0; // scratch rows restart per group (a chain's groups run in its stream order)
		}
		sreps[i] = t.plan[i].job->dev;
		sreps[i].rowBase = rb;
		const int ny = 1 << sreps[i].ly, rpw = 64 / batch_row_lanes(sreps[i].lx);
		for(int y = 0, k = 0; y < ny; y += rpw, ++k)
			srows[rowStart[i] + k] = make_uint2((unsigned)(i - groupStart.back()), (unsigned)y);
		rb += ny; // scratch: 4 int64 per site row
		maxGroupRows = std::max(maxGroupRows, rb);
	}
	groupStart.push_back(np);
	t.out = (int64_t*)t.stage.take(std::max<size_t>(np, 1) * 3 * sizeof(int64_t));
	if(np)
	{
		ensure(sampleReps_, np * sizeof(BatchRep));
		ensure(rowTable_, (size_t)nrows * sizeof(uint2));
		ensure(rowScratch_, 2 * (size_t)maxGroupRows * 4 * sizeof(int64_t)); // one region per chain
		ensure(outBuf_, np * 3 * sizeof(int64_t));
		check(hipMemcpyAsync(sampleReps_.get(), sreps, np * sizeof(BatchRep), hipMemcpyHostToDevice, s0_), "sample reps");
		check(hipMemcpyAsync(rowTable_.get(), srows, (size_t)nrows * sizeof(uint2), hipMemcpyHostToDevice, s0_), "rows");
	}

	BatchDtArgs da[2];
	BatchCbArgs ca[2];
	for(int c = 0; c < 2; ++c)
	{
		da[c] = BatchDtArgs{};
		da[c].reps = repTable_.as<BatchRep>();
		da[c].lanes = laneTable_.as<uint2>() + laneOff_[c];
		da[c].order = orderBuf_.as<uint8_t>();
		da[c].shift = shiftBuf_.as<uint16_t>();
		da[c].nreps = maxSlots_;
		da[c].nlanes = nLanes_[c];
		da[c].nstatic = nStatic_[c];
		da[c].limit = limitBuf_.as<int>();
		ca[c] = BatchCbArgs{};
		ca[c].reps = repTable_.as<BatchRep>();
		ca[c].waves = waveTable_.as<uint2>() + waveOff_[c];
		ca[c].nwaves = nWaves_[c];
		ca[c].limit = limitBuf_.as<int>();
	}
	const hipStream_t chain[2] = {s0_, s1_};
	const bool two = nLanes_[1] > 0 || nWaves_[1] > 0;
	if(two)
	{
		check(hipEventRecord(fork_, s0_), "fork");
		check(hipStreamWaitEvent(s1_, fork_, 0), "fork wait");
	}
	size_t g = 0;
	for(int m = 0; m < k; ++m)
		for(int c = 0; c < 2; ++c)
		{
			da[c].m = m;
			ca[c].m = m;
			for(int d = 0; d < 4 && da[c].nlanes > 0; ++d)
			{
				da[c].d = d;
				check(launch_batch_dt(da[c], chain[c]), "batch_dt");
			}
			for(int s = 0; s < 2 && ca[c].nwaves > 0; ++s)
			{
				ca[c].signum = s;
				ca[c].buf = s;
				check(launch_batch_cbbit(ca[c], chain[c]), "batch_cbbit");
			}
			// exact roughness of the chain's jobs whose sample time is this MCS
			if(g + 1 < groupStart.size() && t.plan[groupStart[g]].m == m && t.plan[groupStart[g]].job->group == c)
			{
				const size_t i0 = groupStart[g], i1 = groupStart[g + 1];
				check(launch_batch_roughness(sampleReps_.as<BatchRep>() + i0, (int)(i1 - i0),
				                             rowTable_.as<uint2>() + rowStart[i0], rowStart[i1] - rowStart[i0],
				                             rowScratch_.as<int64_t>() + (size_t)c * maxGroupRows * 4,
				                             outBuf_.as<int64_t>() + 3 * i0, chain[c]),
				      "batch_roughness");
				++g;
			}
		}
	if(two)
	{
		check(hipEventRecord(join_, s1_), "join");
		check(hipStreamWaitEvent(s0_, join_, 0), "join wait");
	}
	if(np)
		check(hipMemcpyAsync(t.out, outBuf_.get(), np * 3 * sizeof(int64_t), hipMemcpyDeviceToHost, s0_), "roughness D2H");
	check(hipEventRecord(t.end, s0_), "tick end");

	// bookkeeping of what this tick runs; a job that has now planned all its
	// MCS leaves the tables of the next tick
	for(Job* j : slots_)
		if(j && limit[j->slot] > 0)
		{
			const int lim = limit[j->slot];
			t.ran.push_back({j, lim});
			while(j->nextPlan < j->samples.size() && j->samples[j->nextPlan] <= j->planned + (uint32_t)lim)
				++j->nextPlan;
			j->planned += (uint32_t)lim;
			if(j->planned == (uint32_t)j->spec.mcs)
				dirty_ = true;
		}
	return true;
}

// wait for a tick, record its samples, retire the jobs it finished
void Server::finish(Tick& t)
{
	if(!t.live)
		return;
	check(hipEventSynchronize(t.end), "tick sync");
	float ms = 0, span = 0;
	check(hipEventElapsedTime(&ms, t.begin, t.end), "tick time");
	check(hipEventElapsedTime(&span, first_, t.end), "span time");
	gpuMs_ += ms;
	spanMs_ = span;
	for(size_t i = 0; i < t.plan.size(); ++i)
	{
		Job* j = t.plan[i].job;
		JobSample smp;
		smp.mcs = j->done + (uint32_t)t.plan[i].m + 1;
		smp.sumH = t.out[3 * i];
		smp.sumH2 = ((__int128)t.out[3 * i + 2] << 64) | (__int128)(uint64_t)t.out[3 * i + 1];
		j->res.series.push_back(smp);
	}
	for(auto& r : t.ran)
	{
		Job* j = r.first;
		j->done += (uint32_t)r.second;
		siteUpdates_ += (uint64_t)r.second << (j->spec.lx + j->spec.ly);
		if(j->done != (uint32_t)j->spec.mcs)
			continue;
		// retire: results out, memory back to the arena, slot free (no tick
		// still queued touches this job: it left the tables after its last MCS)
		if(j->spec.keepLattice)
		{
			j->res.lattice.resize((size_t)1 << (j->spec.lx + j->spec.ly - 4));
			check(hipMemcpy(j->res.lattice.data(), j->dev.lat, j->res.lattice.size() * 4, hipMemcpyDeviceToHost),
			      "final lattice");
		}
		size_t parts[3];
		footprint(j->spec, parts);
		for(int u = 0; u < 3; ++u)
			if(parts[u])
				arena_.release(j->off[u]);
		j->res.tEnd = now();
		done_.push_back(std::move(j->res));
		slots_[(size_t)j->slot] = nullptr;
		delete j;
		--nActive_;
	}
	t.ran.clear();
	t.plan.clear();
	t.live = false;
}

size_t Server::step()
{
	enqueue(ticks_[cur_]);
	finish(ticks_[cur_ ^ 1]);
	cur_ ^= 1;
	return queue_.size() + nActive_;
}

bool Server::pop(JobResult& out)
{
	if(done_.empty())
		return false;
	out = std::move(done_.front());
	done_.pop_front();
	return true;
}

} // namespace kpz
