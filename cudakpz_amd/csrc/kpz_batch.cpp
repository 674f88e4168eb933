#include "kpz_batch.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace kpz {

namespace {
uint64_t align256(uint64_t b) { return (b + 255) & ~255ull; }
} // namespace

Batch::Batch(const std::vector<ReplicaSpec>& specs, const BatchTuning& tune, int device, bool quickSeed) : tune_(tune)
{
	if(specs.empty())
		throw Error(-1, "empty batch");
	if(tune_.chunk < 1)
		tune_.chunk = 1;
	if(device >= 0)
		check(hipSetDevice(device), "hipSetDevice");
	check(hipGetDevice(&device_), "hipGetDevice");
	check(hipStreamCreateWithFlags(&s0_, hipStreamNonBlocking), "hipStreamCreate");
	check(hipStreamCreateWithFlags(&s1_, hipStreamNonBlocking), "hipStreamCreate");
	check(hipStreamCreateWithFlags(&s2_, hipStreamNonBlocking), "hipStreamCreate");
	check(hipStreamCreateWithFlags(&s3_, hipStreamNonBlocking), "hipStreamCreate");
	check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
	for(auto& e : join_)
		check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
	for(auto& e : copied_)
		check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");

	reps_.resize(specs.size());
	uint64_t latBytes = 0, stWords = 0, tabEntries = 1; // entry 0: identity
	for(size_t r = 0; r < specs.size(); ++r)
	{
		Rep& R = reps_[r];
		R.spec = specs[r];
		const ReplicaSpec& s = R.spec;
		if(!(s.p >= 0 && s.p <= 1 && s.q >= 0 && s.q <= 1))
			throw Error(-1, "replica " + std::to_string(r) + ": p, q must lie in [0, 1]");
		BatchRep& d = R.dev;
		d.lx = s.lx;
		d.ly = s.ly;
		if(s.kind == KIND_DT || s.kind == KIND_DTR)
		{
			if(s.lx < 5 || s.ly < 5 || s.lx > 24 || s.ly > 24)
				throw Error(-1, "batch DT replicas need 32..2^24 sites per side (16x16 blocks)");
			d.kind = s.kind == KIND_DT ? BK_DT : BK_DTR;
			d.mode = (s.p == 1.0 && s.q == 0.0) ? 0 : 1;
			R.T = 1ll << (s.lx + s.ly - 10);
			d.tp = thresh23(d.mode ? s.p : 1.0);
			d.tq = thresh23(d.mode ? s.q : 0.0);
			R.stOff = stWords;
			stWords += (uint64_t)R.T;
			hasDt_ = true;
		}
		else if(s.kind == KIND_SCABIT)
		{
			if(s.lx < 6 || s.ly < 1 || s.lx > 24 || s.ly > 24)
				throw Error(-1, "batch SCABit replicas need log2 width 6..24");
			d.kind = BK_SCABIT;
			d.mode = (s.p == 0.5 && s.q == 0.0) ? 0 : 1;
			d.wide = (s.lx >= 7 && (d.mode == 0 || tune_.pqWords64)) ? 1 : 0;
			const int lWn = s.lx - (d.wide ? 7 : 6);
			d.lG = std::min(lWn, std::max(0, std::min(6, tune_.maxLanesLog2)));
			if(s.ly + d.lG < 6)
				throw Error(-1, "batch SCABit replica too small for one wave (rows x lanes < 64)");
			R.T = 1ll << s.ly;
			d.tp = thresh23(s.p);
			d.tq = thresh23(s.q);
			R.stOff = stWords;
			stWords += 2 * (uint64_t)R.T;
			tabEntries += 1ull << d.lG;
			hasCb_ = true;
		}
		else
			throw Error(-1, "unknown replica kind");
		R.latOff = latBytes;
		latBytes += align256((1ull << (s.lx + s.ly - 4)) * 4);
		sites_ += 1ull << (s.lx + s.ly);
		d.step = slcg_skip(default_skip((uint64_t)R.T));
		if(quickSeed)
			R.master.init(seed32(s.seed));
		else
			R.master.reseed(seed32(s.seed));
	}
	latArena_.alloc(latBytes);
	stArena_.alloc(stWords * 8);
	tabArena_.alloc(tabEntries * sizeof(Affine));

	// stream states (SLCG64::randomize) and SCABit jump tables
	std::vector<uint64_t> st(stWords);
	std::vector<Affine> tab(tabEntries);
	tab[0] = Affine{1, 0};
	uint64_t tabPos = 1;
	uint32_t* latBase = latArena_.as<uint32_t>();
	uint64_t* stBase = stArena_.as<uint64_t>();
	Affine* tabBase = tabArena_.as<Affine>();
	for(Rep& R : reps_)
	{
		const double m1 = R.master.close_open();
		const double m2 = R.master.close_open();
		uint64_t s = (uint64_t)(m1 * 4294967295.0) ^ ((uint64_t)(m2 * 4294967295.0) << 32);
		for(int64_t t = 0; t < R.T; ++t)
		{
			st[R.stOff + (uint64_t)t] = s;
			s = s * SLCG_A + SLCG_C;
		}
		BatchRep& d = R.dev;
		d.lat = latBase + R.latOff / 4;
		d.st[0] = stBase + R.stOff;
		d.st[1] = (d.kind == BK_SCABIT) ? stBase + R.stOff + (uint64_t)R.T : d.st[0];
		d.rowJ = tabBase;
		if(d.kind == BK_SCABIT)
		{
			const uint64_t du = d.mode == 0 ? (d.wide ? 2 : 1) : (d.wide ? 64 : 32);
			const uint64_t Wn = 1ull << (d.lx - (d.wide ? 7 : 6));
			const int G = 1 << d.lG;
			const Affine word = affine_pow(d.step, du);
			Affine acc{1, 0};
			d.laneJ = tabBase + tabPos;
			for(int g = 0; g < G; ++g)
			{
				tab[tabPos++] = acc;
				acc = compose(word, acc);
			}
			d.stepJ = affine_pow(d.step, du * (uint64_t)G);
			d.full = affine_pow(d.step, du * Wn);
		}
		else
		{
			d.laneJ = tabBase;
			d.full = affine_pow(d.step, (d.mode ? 3u : 2u) * 256u); // one 16 x 16 block
		}
	}
	check(hipMemcpy(stArena_.get(), st.data(), st.size() * 8, hipMemcpyHostToDevice), "upload streams");
	check(hipMemcpy(tabArena_.get(), tab.data(), tab.size() * sizeof(Affine), hipMemcpyHostToDevice), "upload tables");
	buildTables();
}

void Batch::buildTables()
{
	const int R = size();
	std::vector<int> order(R);
	std::iota(order.begin(), order.end(), 0);
	if(tune_.sortBySize)
		std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
			return reps_[a].spec.lx + reps_[a].spec.ly > reps_[b].spec.lx + reps_[b].spec.ly;
		});
	// replica groups: with concurrent == 2 each replica goes whole to the group
	// of its engine (DT or SCABit) with fewer sites so far, in launch order;
	// else every replica to group 0
	std::vector<int> group(R, 0);
	if(tune_.concurrent >= 2)
	{
		uint64_t load[2][2] = {{0, 0}, {0, 0}};
		for(int r : order)
		{
			uint64_t* l = load[reps_[r].dev.kind == BK_SCABIT ? 1 : 0];
			group[r] = l[1] < l[0] ? 1 : 0;
			l[group[r]] += 1ull << (reps_[r].spec.lx + reps_[r].spec.ly);
		}
	}
	// DT lanes of each group, one class per (kind, mode) so every wave is uniform
	std::vector<uint2> lanes;
	for(int g = 0; g < 2; ++g)
	{
		dtOff_[g] = (int64_t)lanes.size();
		for(int kind : {BK_DT, BK_DTR})
		{
			if(kind == BK_DTR)
				dtStatic_[g] = (int64_t)lanes.size() - dtOff_[g];
			for(int mode : {0, 1})
			{
				for(int r : order)
				{
					const BatchRep& d = reps_[r].dev;
					if(d.kind != kind || d.mode != mode || group[r] != g)
						continue;
					for(int64_t b = 0; b < reps_[r].T; ++b)
						lanes.push_back(make_uint2((unsigned)r, (unsigned)b));
				}
				while(lanes.size() % 64)
					lanes.push_back(make_uint2(0xFFFFFFFFu, 0));
			}
		}
		dtLanes_[g] = (int64_t)lanes.size() - dtOff_[g];
	}
	// SCABit waves of each group: 64 >> lG rows of one replica each
	std::vector<uint2> waves;
	for(int g = 0; g < 2; ++g)
	{
		cbOff_[g] = (int64_t)waves.size();
		for(int r : order)
		{
			const BatchRep& d = reps_[r].dev;
			if(d.kind != BK_SCABIT || group[r] != g)
				continue;
			const int rpw = 64 >> d.lG;
			for(int y = 0; y < (1 << d.ly); y += rpw)
				waves.push_back(make_uint2((unsigned)r, (unsigned)y));
		}
		cbWaves_[g] = (int64_t)waves.size() - cbOff_[g];
	}
	// all rows, for the reductions: waves of 64 / batch_row_lanes(lx) rows
	std::vector<uint2> rows;
	int64_t rowBase = 0;
	for(int r = 0; r < R; ++r)
	{
		reps_[r].dev.rowBase = rowBase;
		const int rpw = 64 / batch_row_lanes(reps_[r].dev.lx);
		for(int y = 0; y < (1 << reps_[r].dev.ly); y += rpw)
			rows.push_back(make_uint2((unsigned)r, (unsigned)y));
		rowBase += 1 << reps_[r].dev.ly;
	}
	nRows_ = rowBase;
	nRowWaves_ = (int64_t)rows.size();

	std::vector<BatchRep> table(R);
	for(int r = 0; r < R; ++r)
		table[r] = reps_[r].dev;
	repTable_.alloc(sizeof(BatchRep) * R);
	check(hipMemcpy(repTable_.get(), table.data(), sizeof(BatchRep) * R, hipMemcpyHostToDevice), "rep table");
	if(!lanes.empty())
	{
		laneTable_.alloc(lanes.size() * sizeof(uint2));
		check(hipMemcpy(laneTable_.get(), lanes.data(), lanes.size() * sizeof(uint2), hipMemcpyHostToDevice), "lanes");
	}
	if(!waves.empty())
	{
		waveTable_.alloc(waves.size() * sizeof(uint2));
		check(hipMemcpy(waveTable_.get(), waves.data(), waves.size() * sizeof(uint2), hipMemcpyHostToDevice), "waves");
	}
	rowTable_.alloc(rows.size() * sizeof(uint2));
	check(hipMemcpy(rowTable_.get(), rows.data(), rows.size() * sizeof(uint2), hipMemcpyHostToDevice), "rows");
	rowScratch_.alloc((size_t)nRows_ * 4 * sizeof(int64_t));
	outBuf_.alloc((size_t)R * 3 * sizeof(int64_t));
	const size_t cells = (size_t)tune_.chunk * R;
	for(int b = 0; b < 2; ++b)
	{
		orderBuf_[b].alloc(cells);
		shiftBuf_[b].alloc(cells * 2);
		check(hipHostMalloc((void**)&orderHost_[b], cells, hipHostMallocDefault), "hipHostMalloc");
		check(hipHostMalloc((void**)&shiftHost_[b], cells * 2, hipHostMallocDefault), "hipHostMalloc");
		check(hipEventRecord(copied_[b], s0_), "hipEventRecord");
	}
}

Batch::~Batch()
{
	if(s0_)
		(void)hipStreamSynchronize(s0_);
	if(s1_)
		(void)hipStreamSynchronize(s1_);
	if(s2_)
		(void)hipStreamSynchronize(s2_);
	if(s3_)
		(void)hipStreamSynchronize(s3_);
	for(auto& kv : graphs_)
	{
		if(kv.second.exec)
			(void)hipGraphExecDestroy(kv.second.exec);
		if(kv.second.graph)
			(void)hipGraphDestroy(kv.second.graph);
	}
	for(int b = 0; b < 2; ++b)
	{
		if(orderHost_[b])
			(void)hipHostFree(orderHost_[b]);
		if(shiftHost_[b])
			(void)hipHostFree(shiftHost_[b]);
		if(copied_[b])
			(void)hipEventDestroy(copied_[b]);
	}
	if(fork_)
		(void)hipEventDestroy(fork_);
	for(auto& e : join_)
		if(e)
			(void)hipEventDestroy(e);
	if(s3_)
		(void)hipStreamDestroy(s3_);
	if(s2_)
		(void)hipStreamDestroy(s2_);
	if(s1_)
		(void)hipStreamDestroy(s1_);
	if(s0_)
		(void)hipStreamDestroy(s0_);
}

void Batch::sync() { check(hipStreamSynchronize(s0_), "hipStreamSynchronize"); }

void Batch::initHomogeneous()
{
	for(Rep& R : reps_)
		check(launch_fill_flat(R.dev.lat, R.dev.kind == BK_SCABIT ? ENC_CBBIT : ENC_LOCAL, R.spec.lx, R.spec.ly, s0_), "fill");
}

// per-MCS master draws of every DT replica (kpz/schedulerCPUDT.cpp:75,
// kpz/schedulerCPUDTr.cpp:76-82)
void Batch::drawChunk(int b, int len)
{
	const int R = size();
	for(int m = 0; m < len; ++m)
		for(int r = 0; r < R; ++r)
		{
			Rep& P = reps_[r];
			if(P.dev.kind == BK_SCABIT)
				continue;
			int seq[4];
			P.master.shuffle4(seq);
			orderHost_[b][(size_t)m * R + r] = (uint8_t)(seq[0] | (seq[1] << 2) | (seq[2] << 4) | (seq[3] << 6));
			uint16_t sh = 0;
			if(P.dev.kind == BK_DTR)
			{
				const unsigned sx = (unsigned)(P.master.close_open() * 16.0);
				const unsigned sy = (unsigned)(P.master.close_open() * 16.0);
				sh = (uint16_t)(sx | (sy << 8));
			}
			shiftHost_[b][(size_t)m * R + r] = sh;
		}
}

Batch::Chunk& Batch::graphFor(int b, int len)
{
	auto key = std::make_pair(b, len);
	auto it = graphs_.find(key);
	if(it != graphs_.end())
		return it->second;
	Chunk c;
	const hipStream_t dtStream[2] = {s0_, s2_};
	const hipStream_t cbStream[2] = {tune_.concurrent ? s1_ : s0_, s3_};
	// side streams in use, forked from s0_ at the start and joined at the end
	const hipStream_t side[3] = {s1_, s2_, s3_};
	const bool used[3] = {tune_.concurrent && cbWaves_[0] > 0, dtLanes_[1] > 0, cbWaves_[1] > 0};
	check(hipStreamBeginCapture(s0_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
	check(hipEventRecord(fork_, s0_), "fork");
	for(int i = 0; i < 3; ++i)
		if(used[i])
			check(hipStreamWaitEvent(side[i], fork_, 0), "fork wait");
	BatchDtArgs dg[2];
	for(int g = 0; g < 2; ++g)
	{
		BatchDtArgs& da = dg[g];
		da = BatchDtArgs{};
		da.reps = repTable_.as<BatchRep>();
		da.lanes = laneTable_.as<uint2>() + dtOff_[g];
		da.order = orderBuf_[b].as<uint8_t>();
		da.shift = shiftBuf_[b].as<uint16_t>();
		da.nreps = size();
		da.nlanes = dtLanes_[g];
		da.nstatic = dtStatic_[g];
	}
	BatchCbArgs cg[2];
	for(int g = 0; g < 2; ++g)
	{
		cg[g] = BatchCbArgs{};
		cg[g].reps = repTable_.as<BatchRep>();
		cg[g].waves = waveTable_.as<uint2>() + cbOff_[g];
		cg[g].nwaves = cbWaves_[g];
	}
	for(int m = 0; m < len; ++m)
	{
		if(hasDt_)
			for(int g = 0; g < 2; ++g)
				for(int d = 0; d < 4; ++d)
				{
					dg[g].m = m;
					dg[g].d = d;
					if(dg[g].nlanes > 0)
						check(launch_batch_dt(dg[g], dtStream[g]), "batch_dt");
				}
		if(hasCb_)
			for(int g = 0; g < 2; ++g)
				for(int s = 0; s < 2; ++s)
				{
					cg[g].signum = s;
					cg[g].buf = s;
					if(cg[g].nwaves > 0)
						check(launch_batch_cbbit(cg[g], cbStream[g]), "batch_cbbit");
				}
	}
	for(int i = 0; i < 3; ++i)
		if(used[i])
		{
			check(hipEventRecord(join_[i], side[i]), "join");
			check(hipStreamWaitEvent(s0_, join_[i], 0), "join wait");
		}
	check(hipStreamEndCapture(s0_, &c.graph), "hipStreamEndCapture");
	check(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
	return graphs_[key] = c;
}

void Batch::mcs(int n)
{
	const int R = size();
	int done = 0;
	while(done < n)
	{
		const int len = std::min(tune_.chunk, n - done);
		const int b = cur_;
		check(hipEventSynchronize(copied_[b]), "hipEventSynchronize"); // host buffer b is free again
		drawChunk(b, len);
		check(hipMemcpyAsync(orderBuf_[b].get(), orderHost_[b], (size_t)len * R, hipMemcpyHostToDevice, s0_), "order H2D");
		check(hipMemcpyAsync(shiftBuf_[b].get(), shiftHost_[b], (size_t)len * R * 2, hipMemcpyHostToDevice, s0_), "shift H2D");
		check(hipEventRecord(copied_[b], s0_), "hipEventRecord");
		Chunk& c = graphFor(b, len);
		check(hipGraphLaunch(c.exec, s0_), "hipGraphLaunch");
		cur_ ^= 1;
		done += len;
	}
	mcs_ += (unsigned)n;
}

void Batch::roughness(int64_t* out)
{
	check(launch_batch_roughness(repTable_.as<BatchRep>(), size(), rowTable_.as<uint2>(), nRowWaves_,
			rowScratch_.as<int64_t>(), outBuf_.as<int64_t>(), s0_), "batch_roughness");
	check(hipMemcpyAsync(out, outBuf_.get(), (size_t)size() * 3 * 8, hipMemcpyDeviceToHost, s0_), "roughness D2H");
	sync();
}

void Batch::download(int r, uint32_t* host)
{
	const Rep& P = reps_.at(r);
	check(hipMemcpyAsync(host, P.dev.lat, sizeWords(r) * 4, hipMemcpyDeviceToHost, s0_), "download");
	sync();
}

void Batch::upload(int r, const uint32_t* host)
{
	const Rep& P = reps_.at(r);
	check(hipMemcpyAsync(P.dev.lat, host, sizeWords(r) * 4, hipMemcpyHostToDevice, s0_), "upload");
	sync();
}

void Batch::getStreams(int r, uint64_t* out)
{
	const Rep& P = reps_.at(r);
	check(hipMemcpyAsync(out, P.dev.st[0], (size_t)P.T * 8, hipMemcpyDeviceToHost, s0_), "streams D2H");
	sync();
}

static double timeTrial(const std::vector<ReplicaSpec>& specs, int mcs, int device, const BatchTuning& t)
{
	Batch b(specs, t, device, true);
	b.initHomogeneous();
	// capture + warm-up: the third call alternates the host buffers as the
	// first did, so the timed call replays graphs captured already
	b.mcs(mcs);
	b.mcs(mcs);
	b.sync();
	hipEvent_t e0, e1;
	check(hipEventCreate(&e0), "hipEventCreate");
	check(hipEventCreate(&e1), "hipEventCreate");
	check(hipEventRecord(e0, b.stream()), "hipEventRecord");
	b.mcs(mcs);
	check(hipEventRecord(e1, b.stream()), "hipEventRecord");
	check(hipEventSynchronize(e1), "hipEventSynchronize");
	float ms = 0;
	check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return ms / mcs;
}

double autotuneBatch(const std::vector<ReplicaSpec>& specs, int mcs, int device, BatchTuning& best)
{
	if(mcs < 1)
		mcs = 1;
	double bestMs = timeTrial(specs, mcs, device, best);
	auto tryOne = [&](BatchTuning t) {
		const double ms = timeTrial(specs, mcs, device, t);
		if(ms < bestMs)
		{
			bestMs = ms;
			best = t;
		}
	};
	for(int c : {10, 25, 100})
	{
		BatchTuning t = best;
		t.chunk = c;
		tryOne(t);
	}
	for(int c = 0; c <= 2; ++c)
	{
		BatchTuning t = best;
		if(t.concurrent == c)
			continue;
		t.concurrent = c;
		tryOne(t);
	}
	for(int l = 2; l <= 6; ++l)
	{
		BatchTuning t = best;
		if(t.maxLanesLog2 == l)
			continue;
		t.maxLanesLog2 = l;
		tryOne(t);
	}
	{
		BatchTuning t = best;
		t.pqWords64 ^= 1;
		tryOne(t);
	}
	{
		BatchTuning t = best;
		t.sortBySize ^= 1;
		tryOne(t);
	}
	return bestMs;
}

} // namespace kpz
