// extern "C" boundary over the C++ schedulers (include/kpz_amd.h).
#include "../../include/kpz_amd.h"
#include "kpz_batch.hpp"
#include "kpz_capi_internal.hpp"
#include "kpz_tinymt.hpp"
#include "kpz_scheduler.hpp"
#include "kpz_server.hpp"

#include <algorithm>
#include <cstring>
#include <string>

struct kpz_batch {
	std::unique_ptr<kpz::Batch> b;
};

struct kpz_server {
	std::unique_ptr<kpz::Server> s;
};

namespace {
thread_local std::string g_err;
}

int kpz::capiFail(int code, const std::string& msg)
{
	g_err = msg;
	return code;
}

namespace {
int fail(int code, const std::string& msg) { return kpz::capiFail(code, msg); }

template<class F>
int guard(F&& f)
{
	return kpz::capiGuard(std::forward<F>(f));
}

#define HANDLE(h)                                     \
	if(!(h) || !(h)->s)                               \
		return fail(KPZ_EINVAL, "null scheduler handle");
} // namespace

extern "C" {

const char* kpz_last_error(void) { return g_err.c_str(); }

int kpz_create(int kind, int lx, int ly, const kpz_params* params, kpz_scheduler** out)
{
	if(!out)
		return fail(KPZ_EINVAL, "null output pointer");
	*out = nullptr;
	kpz::Params p;
	if(params)
	{
		if(params->threads < 0)
			return fail(KPZ_EINVAL, "threads must be >= 0");
		p.threads = params->threads;
		p.layout_x = params->layout_x ? params->layout_x : 4;
		p.layout_y = params->layout_y ? params->layout_y : 4;
		p.skip = params->skip;
		p.device = params->device;
		p.rng = params->rng;
		if(p.skip && !(p.skip & 1))
			return fail(KPZ_EINVAL, "SLCG skip must be odd");
	}
	return guard([&] {
		auto h = new kpz_scheduler;
		try
		{
			h->s = kpz::makeScheduler(kind, lx, ly, p);
		}
		catch(...)
		{
			delete h;
			throw;
		}
		*out = h;
	});
}

int kpz_destroy(kpz_scheduler* h)
{
	if(!h)
		return KPZ_OK;
	return guard([&] { delete h; });
}

int kpz_save_state(kpz_scheduler* h, const char* path, uint32_t aux)
{
	HANDLE(h);
	if(!path)
		return fail(KPZ_EINVAL, "null path");
	return guard([&] { h->s->saveState(path, aux); });
}

int kpz_load_state(const char* path, int device, kpz_scheduler** out, uint32_t* aux)
{
	if(!out || !path)
		return fail(KPZ_EINVAL, "null argument");
	*out = nullptr;
	return guard([&] {
		auto h = new kpz_scheduler;
		try
		{
			h->s = kpz::SchedulerService::loadState(path, device, aux);
		}
		catch(...)
		{
			delete h;
			throw;
		}
		*out = h;
	});
}

int kpz_kind(const kpz_scheduler* h)
{
	HANDLE(h);
	return h->s->kind();
}

int kpz_size(const kpz_scheduler* h, int* lx, int* ly)
{
	HANDLE(h);
	if(!lx || !ly)
		return fail(KPZ_EINVAL, "null output");
	*lx = h->s->size().lx;
	*ly = h->s->size().ly;
	return KPZ_OK;
}

int kpz_set_seed(kpz_scheduler* h, uint64_t seed)
{
	HANDLE(h);
	return guard([&] { h->s->supplySeed(seed); });
}

int kpz_randomize(kpz_scheduler* h)
{
	HANDLE(h);
	return guard([&] { h->s->randomize(); });
}

int kpz_set_pq(kpz_scheduler* h, double p, double q)
{
	HANDLE(h);
	if(!h->s->setPQ(p, q))
		return fail(KPZ_EINVAL, "p and q must lie in [0, 1]");
	return KPZ_OK;
}

int kpz_init_flat(kpz_scheduler* h)
{
	HANDLE(h);
	return guard([&] { h->s->initHomogeneous(); });
}

int kpz_upload(kpz_scheduler* h, const uint32_t* lattice)
{
	HANDLE(h);
	if(!lattice)
		return fail(KPZ_EINVAL, "null lattice");
	return guard([&] { h->s->pushSystem(lattice); });
}

int kpz_download(kpz_scheduler* h, uint32_t* lattice)
{
	HANDLE(h);
	if(!lattice)
		return fail(KPZ_EINVAL, "null lattice");
	return guard([&] { h->s->popSystem(lattice); });
}

int kpz_mcs(kpz_scheduler* h, int n)
{
	HANDLE(h);
	if(n < 0)
		return fail(KPZ_EINVAL, "negative MCS count");
	return guard([&] { h->s->mcs(n); });
}

int kpz_mcs_pq(kpz_scheduler* h, int n)
{
	HANDLE(h);
	if(n < 0)
		return fail(KPZ_EINVAL, "negative MCS count");
	return guard([&] { h->s->mcsPQSyncRng(n); });
}

int kpz_sync(kpz_scheduler* h)
{
	HANDLE(h);
	return guard([&] { h->s->sync(); });
}

int kpz_copy_system(kpz_scheduler* dst, const kpz_scheduler* src)
{
	HANDLE(dst);
	HANDLE(src);
	return guard([&] { dst->s->copySystem(*src->s); });
}

int kpz_roughness(kpz_scheduler* h, int64_t out[3])
{
	HANDLE(h);
	if(!out)
		return fail(KPZ_EINVAL, "null output");
	return guard([&] {
		const kpz::Sums s = h->s->roughness();
		out[0] = s.sumH;
		out[1] = (int64_t)(uint64_t)s.sumH2;
		out[2] = (int64_t)(s.sumH2 >> 64);
	});
}

int kpz_roughness_noalloc(const uint32_t* lattice, int lx, int ly, int encoding, int64_t* out, void* stream)
{
	if(!lattice || !out)
		return fail(KPZ_EINVAL, "null lattice or output pointer");
	if(encoding != KPZ_ENC_LOCAL && encoding != KPZ_ENC_CBBIT)
		return fail(KPZ_EINVAL, "encoding must be KPZ_ENC_LOCAL or KPZ_ENC_CBBIT");
	const bool cb = encoding == KPZ_ENC_CBBIT;
	if(lx < (cb ? 6 : 2) || lx > 30 || ly < (cb ? 1 : 2))
		return fail(KPZ_EINVAL, "lattice size out of range: lx " + std::to_string(lx) + ", ly " + std::to_string(ly));
	if(ly > kpz::NOALLOC_MAX_LY)
		return fail(KPZ_EINVAL, "ly " + std::to_string(ly) + " exceeds the fixed workspace (" +
		                            std::to_string(kpz::noalloc_workspace_bytes()) + " bytes): at most 2^" +
		                            std::to_string(kpz::NOALLOC_MAX_LY) + " site rows");
	const hipError_t e = kpz::launch_roughness_noalloc(lattice, encoding, lx, ly, out, static_cast<hipStream_t>(stream));
	if(e != hipSuccess)
		return fail(KPZ_EDEVICE, std::string("kpz_roughness_noalloc: ") + hipGetErrorString(e));
	return KPZ_OK;
}

int kpz_set_pq_class(kpz_scheduler* h, double p, double q, unsigned n)
{
	HANDLE(h);
	if(n >= 4)
		return fail(KPZ_EINVAL, "disorder class must be 0..3");
	return guard([&] {
		if(!h->s->setPQ(p, q, n))
			throw kpz::Error(KPZ_EINVAL, "rates must lie in [0, 1]");
	});
}

int kpz_get_pq_class(const kpz_scheduler* h, unsigned n, double* p, double* q)
{
	HANDLE(h);
	if(n >= 4)
		return fail(KPZ_EINVAL, "disorder class must be 0..3");
	if(p)
		*p = h->s->rateP(n);
	if(q)
		*q = h->s->rateQ(n);
	return KPZ_OK;
}

int kpz_generate_disorder2(kpz_scheduler* h, double fill)
{
	HANDLE(h);
	return guard([&] { h->s->generateDisorder2(fill); });
}

int kpz_set_disorder(kpz_scheduler* h, const uint32_t* map)
{
	HANDLE(h);
	if(!map)
		return fail(KPZ_EINVAL, "null map");
	return guard([&] { h->s->setDisorder(map); });
}

int kpz_get_disorder(kpz_scheduler* h, uint32_t* map)
{
	HANDLE(h);
	if(!map)
		return fail(KPZ_EINVAL, "null map");
	return guard([&] { h->s->getDisorder(map); });
}

int kpz_mcs_disorder2(kpz_scheduler* h, int n)
{
	HANDLE(h);
	if(n < 0)
		return fail(KPZ_EINVAL, "negative MCS count");
	return guard([&] { h->s->mcsDisorder2SyncRng(n); });
}

int kpz_mcs_ew(kpz_scheduler* h, int n)
{
	HANDLE(h);
	if(n < 0)
		return fail(KPZ_EINVAL, "negative MCS count");
	return guard([&] { h->s->mcsEW(n); });
}

int kpz_moments(kpz_scheduler* h, double out[7])
{
	HANDLE(h);
	if(!out)
		return fail(KPZ_EINVAL, "null output");
	return guard([&] {
		long double m[7];
		h->s->moments(m);
		for(int k = 0; k < 7; ++k)
			out[k] = (double)m[k];
	});
}

int kpz_consistency_check(kpz_scheduler* h, int full, int64_t out[3])
{
	HANDLE(h);
	if(!out)
		return fail(KPZ_EINVAL, "null output");
	bool ok = false;
	const int r = guard([&] { ok = h->s->consistencyCheck(full != 0, out); });
	return r < 0 ? r : (ok ? 1 : 0);
}

int kpz_corr_tag(kpz_scheduler* h)
{
	HANDLE(h);
	return guard([&] { h->s->correlateTag(); });
}

int kpz_corr_tag_stop(kpz_scheduler* h, uint32_t stop)
{
	HANDLE(h);
	return guard([&] { h->s->correlateTag(stop); });
}

int kpz_correlate_tags(kpz_scheduler* h, int64_t* out, int max_tags)
{
	HANDLE(h);
	if(!out || max_tags < 0)
		return fail(KPZ_EINVAL, "null output");
	int n = 0;
	const int rc = guard([&] {
		const auto res = h->s->correlateAll();
		for(const auto& r : res)
		{
			if(n >= max_tags)
				break;
			int64_t* o = out + 8 * n;
			o[0] = r.tagMcs;
			o[1] = r.sums.sumH;
			o[2] = r.sums.sumHs;
			o[3] = (int64_t)(uint64_t)r.sums.sumH2;
			o[4] = (int64_t)(r.sums.sumH2 >> 64);
			o[5] = (int64_t)(uint64_t)r.sums.sumHHs;
			o[6] = (int64_t)(r.sums.sumHHs >> 64);
			o[7] = r.sums.tcs;
			++n;
		}
	});
	return rc ? rc : n;
}

int kpz_correlate(kpz_scheduler* h, int64_t out[7])
{
	HANDLE(h);
	if(!out)
		return fail(KPZ_EINVAL, "null output");
	if(!h->s->hasTag())
		return fail(KPZ_ESTATE, "correlate() needs a correlateTag() snapshot");
	return guard([&] {
		const kpz::Sums s = h->s->correlate();
		out[0] = s.sumH;
		out[1] = s.sumHs;
		out[2] = (int64_t)(uint64_t)s.sumH2;
		out[3] = (int64_t)(s.sumH2 >> 64);
		out[4] = (int64_t)(uint64_t)s.sumHHs;
		out[5] = (int64_t)(s.sumHHs >> 64);
		out[6] = s.tcs;
	});
}

int64_t kpz_threads(const kpz_scheduler* h) { return (h && h->s) ? h->s->threads() : 0; }

int kpz_get_streams(kpz_scheduler* h, uint64_t* states)
{
	HANDLE(h);
	if(!states)
		return fail(KPZ_EINVAL, "null output");
	return guard([&] { h->s->getStreams(states); });
}

int kpz_set_streams(kpz_scheduler* h, const uint64_t* states)
{
	HANDLE(h);
	if(!states)
		return fail(KPZ_EINVAL, "null input");
	return guard([&] { h->s->setStreams(states); });
}

int kpz_rng(const kpz_scheduler* h) { return (h && h->s) ? h->s->rng() : KPZ_EINVAL; }

int kpz_set_tinymt_params(kpz_scheduler* h, const uint32_t* mat1, const uint32_t* mat2, const uint32_t* tmat)
{
	HANDLE(h);
	if(!mat1 || !mat2 || !tmat)
		return fail(KPZ_EINVAL, "null input");
	return guard([&] { h->s->setTinyMTParams(mat1, mat2, tmat); });
}

int kpz_tinymt_jump(const uint32_t gen[7], uint64_t n, uint32_t status[4])
{
	if(!gen || !status)
		return fail(KPZ_EINVAL, "null argument");
	const kpz::TmtOrbit m = kpz::tmt_orbit(gen, gen[4], gen[5]);
	uint32_t s[4] = {gen[0], gen[1], gen[2], gen[3]};
	kpz::tmt_jump4(s, kpz::poly_xpow(n, m), m, gen[4], gen[5]);
	for(int k = 0; k < 4; ++k)
		status[k] = s[k];
	return (int)m.deg;
}

int kpz_get_tinymt(kpz_scheduler* h, uint32_t* planes)
{
	HANDLE(h);
	if(!planes)
		return fail(KPZ_EINVAL, "null output");
	return guard([&] { h->s->getTinyMT(planes); });
}

int kpz_set_tinymt(kpz_scheduler* h, const uint32_t* planes)
{
	HANDLE(h);
	if(!planes)
		return fail(KPZ_EINVAL, "null input");
	return guard([&] { h->s->setTinyMT(planes); });
}

int kpz_master_bytes(void) { return (int)kpz::MasterRng::STATE_BYTES; }

int kpz_get_master(kpz_scheduler* h, uint8_t* state)
{
	HANDLE(h);
	if(!state)
		return fail(KPZ_EINVAL, "null output");
	h->s->master().save(state);
	return KPZ_OK;
}

int kpz_set_master(kpz_scheduler* h, const uint8_t* state)
{
	HANDLE(h);
	if(!state)
		return fail(KPZ_EINVAL, "null input");
	h->s->master().load(state);
	return KPZ_OK;
}

double kpz_master_draw(kpz_scheduler* h) { return (h && h->s) ? h->s->master().close_open() : -1.0; }

uint32_t kpz_mcs_count(const kpz_scheduler* h) { return (h && h->s) ? h->s->mcsCount() : 0; }
double kpz_mcs_scale(const kpz_scheduler* h) { return (h && h->s) ? h->s->mcsScale() : 0.0; }
int kpz_encoding(const kpz_scheduler* h) { return (h && h->s) ? (int)h->s->encoding() : -1; }
uint64_t kpz_size_words(const kpz_scheduler* h) { return (h && h->s) ? h->s->size().sizeW() : 0; }
uint64_t kpz_skip(const kpz_scheduler* h) { return (h && h->s) ? h->s->skip() : 0; }
void* kpz_device_lattice(const kpz_scheduler* h) { return (h && h->s) ? (void*)h->s->deviceLattice() : nullptr; }
void* kpz_stream(const kpz_scheduler* h) { return (h && h->s) ? (void*)h->s->stream() : nullptr; }

int kpz_valu_peak(int device, int reps, double* slow_class, double* fast_class)
{
	if(!slow_class || reps < 1)
		return fail(KPZ_EINVAL, "null output or reps < 1");
	return guard([&] {
		if(device >= 0)
			kpz::check(hipSetDevice(device), "hipSetDevice");
		int dev = 0, cus = 0;
		kpz::check(hipGetDevice(&dev), "hipGetDevice");
		kpz::check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
		// 32 blocks of 256 lanes per CU: enough resident waves to keep every
		// SIMD issuing (fewer leave the measured rate below the peak)
		const unsigned blocks = (unsigned)(cus > 0 ? cus : 1) * 32u;
		const int iters = 1024;
		kpz::DevBuf out((size_t)blocks * 256 * 4);
		hipEvent_t e0, e1;
		kpz::check(hipEventCreate(&e0), "event");
		kpz::check(hipEventCreate(&e1), "event");
		for(int fast = 0; fast < (fast_class ? 2 : 1); ++fast)
		{
			double best = 0;
			for(int r = 0; r <= reps; ++r) // the first launch warms up
			{
				kpz::check(hipEventRecord(e0, nullptr), "event");
				kpz::check(kpz::launch_valu_probe(out.as<uint32_t>(), blocks, iters, fast != 0, nullptr), "valu probe");
				kpz::check(hipEventRecord(e1, nullptr), "event");
				kpz::check(hipEventSynchronize(e1), "event sync");
				float ms = 0;
				kpz::check(hipEventElapsedTime(&ms, e0, e1), "event time");
				const double ops = (double)blocks * 256.0 * iters * 8.0;
				if(r > 0 && ms > 0)
					best = std::max(best, ops / (ms * 1e-3));
			}
			*(fast ? fast_class : slow_class) = best;
		}
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
	});
}

int kpz_valu_mix_rate(int device, int reps, int slow_of_8, double* lane_ops_per_s)
{
	if(!lane_ops_per_s || reps < 1)
		return fail(KPZ_EINVAL, "null output or reps < 1");
	if(slow_of_8 != 0 && (slow_of_8 < 2 || slow_of_8 > 5) && slow_of_8 != 8)
		return fail(KPZ_EINVAL, "slow_of_8 must be 0, 2, 3, 4, 5 or 8");
	return guard([&] {
		if(device >= 0)
			kpz::check(hipSetDevice(device), "hipSetDevice");
		int dev = 0, cus = 0;
		kpz::check(hipGetDevice(&dev), "hipGetDevice");
		kpz::check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
		const unsigned blocks = (unsigned)(cus > 0 ? cus : 1) * 32u; // 8 waves per SIMD
		const int iters = 1024;
		kpz::DevBuf out((size_t)blocks * 256 * 4);
		hipEvent_t e0, e1;
		kpz::check(hipEventCreate(&e0), "event");
		kpz::check(hipEventCreate(&e1), "event");
		double best = 0;
		for(int r = 0; r <= reps; ++r) // the first launch warms up
		{
			kpz::check(hipEventRecord(e0, nullptr), "event");
			kpz::check(kpz::launch_valu_mix(out.as<uint32_t>(), blocks, iters, slow_of_8, nullptr), "valu mix");
			kpz::check(hipEventRecord(e1, nullptr), "event");
			kpz::check(hipEventSynchronize(e1), "event sync");
			float ms = 0;
			kpz::check(hipEventElapsedTime(&ms, e0, e1), "event time");
			const double ops = (double)blocks * 256.0 * iters * 8.0;
			if(r > 0 && ms > 0)
				best = std::max(best, ops / (ms * 1e-3));
		}
		*lane_ops_per_s = best;
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
	});
}

// ---------------------------------------------------------------- batches
static bool to_specs(const kpz_replica* reps, int n, std::vector<kpz::ReplicaSpec>& out)
{
	if(!reps || n <= 0)
		return false;
	out.resize((size_t)n);
	for(int i = 0; i < n; ++i)
		out[(size_t)i] = kpz::ReplicaSpec{reps[i].kind, reps[i].lx, reps[i].ly, reps[i].p, reps[i].q, reps[i].seed};
	return true;
}

static kpz::BatchTuning to_tuning(const kpz_batch_tuning* t)
{
	kpz::BatchTuning k;
	if(t)
	{
		k.maxLanesLog2 = t->max_lanes_log2;
		k.pqWords64 = t->pq_words64;
		k.chunk = t->chunk;
		k.concurrent = t->concurrent;
		k.sortBySize = t->sort_by_size;
	}
	return k;
}

#define BHANDLE(h)                                    \
	if(!(h) || !(h)->b)                               \
		return fail(KPZ_EINVAL, "null batch handle");

int kpz_batch_create(const kpz_replica* reps, int n, const kpz_batch_tuning* tuning, int device, kpz_batch** out)
{
	if(!out)
		return fail(KPZ_EINVAL, "null output pointer");
	*out = nullptr;
	std::vector<kpz::ReplicaSpec> specs;
	if(!to_specs(reps, n, specs))
		return fail(KPZ_EINVAL, "no replicas");
	return guard([&] {
		auto h = new kpz_batch;
		try
		{
			h->b = std::make_unique<kpz::Batch>(specs, to_tuning(tuning), device);
		}
		catch(...)
		{
			delete h;
			throw;
		}
		*out = h;
	});
}

int kpz_batch_destroy(kpz_batch* b)
{
	if(!b)
		return KPZ_OK;
	return guard([&] { delete b; });
}

int kpz_batch_init_flat(kpz_batch* b)
{
	BHANDLE(b);
	return guard([&] { b->b->initHomogeneous(); });
}

int kpz_batch_mcs(kpz_batch* b, int n)
{
	BHANDLE(b);
	if(n < 0)
		return fail(KPZ_EINVAL, "negative MCS count");
	return guard([&] { b->b->mcs(n); });
}

int kpz_batch_roughness(kpz_batch* b, int64_t* out)
{
	BHANDLE(b);
	if(!out)
		return fail(KPZ_EINVAL, "null output");
	return guard([&] { b->b->roughness(out); });
}

int kpz_batch_download(kpz_batch* b, int replica, uint32_t* lattice)
{
	BHANDLE(b);
	if(!lattice || replica < 0 || replica >= b->b->size())
		return fail(KPZ_EINVAL, "bad replica index or null lattice");
	return guard([&] { b->b->download(replica, lattice); });
}

int kpz_batch_get_streams(kpz_batch* b, int replica, uint64_t* states)
{
	BHANDLE(b);
	if(!states || replica < 0 || replica >= b->b->size())
		return fail(KPZ_EINVAL, "bad replica index or null output");
	return guard([&] { b->b->getStreams(replica, states); });
}

int64_t kpz_batch_threads(const kpz_batch* b, int replica)
{
	if(!b || !b->b || replica < 0 || replica >= b->b->size())
		return 0;
	return b->b->threads(replica);
}

uint64_t kpz_batch_sites(const kpz_batch* b) { return (b && b->b) ? b->b->sites() : 0; }

int kpz_batch_sync(kpz_batch* b)
{
	BHANDLE(b);
	return guard([&] { b->b->sync(); });
}

void* kpz_batch_stream(const kpz_batch* b) { return (b && b->b) ? (void*)b->b->stream() : nullptr; }

// ---------------------------------------------------------------- server
#define SHANDLE(h)                                    \
	if(!(h) || !(h)->s)                               \
		return fail(KPZ_EINVAL, "null server handle");

int kpz_server_create(uint64_t hbm_budget_bytes, int max_slots, int max_tick_mcs, int device, kpz_server** out)
{
	if(!out)
		return fail(KPZ_EINVAL, "null output pointer");
	*out = nullptr;
	return guard([&] {
		auto h = new kpz_server;
		try
		{
			h->s = std::make_unique<kpz::Server>((size_t)hbm_budget_bytes, max_slots, max_tick_mcs, device);
		}
		catch(...)
		{
			delete h;
			throw;
		}
		*out = h;
	});
}

int kpz_server_destroy(kpz_server* s)
{
	if(!s)
		return KPZ_OK;
	return guard([&] { delete s; });
}

int kpz_server_submit(kpz_server* s, const kpz_job* job, uint64_t* id)
{
	SHANDLE(s);
	if(!job)
		return fail(KPZ_EINVAL, "null job");
	return guard([&] {
		const uint64_t i = s->s->submit(kpz::JobSpec{job->kind, job->lx, job->ly, job->mcs, job->p, job->q, job->seed,
				job->keep_lattice});
		if(id)
			*id = i;
	});
}

int64_t kpz_server_step(kpz_server* s)
{
	SHANDLE(s);
	int64_t left = 0;
	const int rc = guard([&] { left = (int64_t)s->s->step(); });
	return rc ? rc : left;
}

namespace {
thread_local kpz::JobResult g_pending;
thread_local const kpz_server* g_pendingOwner = nullptr;
}

int kpz_server_peek(kpz_server* s, int* nsamples, uint64_t* lattice_words)
{
	SHANDLE(s);
	if(g_pendingOwner != s)
	{
		if(!s->s->pop(g_pending))
			return 0;
		g_pendingOwner = s;
	}
	if(nsamples)
		*nsamples = (int)g_pending.series.size();
	if(lattice_words)
		*lattice_words = g_pending.lattice.size();
	return 1;
}

int kpz_server_pop(kpz_server* s, uint64_t* id, double times[3], uint32_t* mcs, int64_t* sums, uint32_t* lattice)
{
	SHANDLE(s);
	if(g_pendingOwner != s && !kpz_server_peek(s, nullptr, nullptr))
		return fail(KPZ_ESTATE, "no finished job");
	const kpz::JobResult& r = g_pending;
	if(id)
		*id = r.id;
	if(times)
	{
		times[0] = r.tSubmit;
		times[1] = r.tStart;
		times[2] = r.tEnd;
	}
	for(size_t i = 0; i < r.series.size(); ++i)
	{
		if(mcs)
			mcs[i] = r.series[i].mcs;
		if(sums)
		{
			sums[3 * i] = r.series[i].sumH;
			sums[3 * i + 1] = (int64_t)(uint64_t)r.series[i].sumH2;
			sums[3 * i + 2] = (int64_t)(r.series[i].sumH2 >> 64);
		}
	}
	if(lattice && !r.lattice.empty())
		std::memcpy(lattice, r.lattice.data(), r.lattice.size() * 4);
	g_pendingOwner = nullptr;
	g_pending = kpz::JobResult{};
	return KPZ_OK;
}

int kpz_server_stats(kpz_server* s, uint64_t out[5])
{
	SHANDLE(s);
	if(!out)
		return fail(KPZ_EINVAL, "null output");
	out[0] = s->s->queued();
	out[1] = s->s->active();
	out[2] = s->s->hbmUsed();
	out[3] = s->s->hbmCapacity();
	out[4] = s->s->siteUpdates();
	return KPZ_OK;
}

double kpz_server_now(const kpz_server* s) { return (s && s->s) ? s->s->now() : -1.0; }

int kpz_server_busy(const kpz_server* s, double* gpu_seconds, double* span_seconds)
{
	SHANDLE(s);
	double g = 0, w = 0;
	s->s->busy(g, w);
	if(gpu_seconds)
		*gpu_seconds = g;
	if(span_seconds)
		*span_seconds = w;
	return KPZ_OK;
}

int kpz_batch_autotune(const kpz_replica* reps, int n, int mcs, int device, kpz_batch_tuning* tuning, double* ms_per_mcs)
{
	std::vector<kpz::ReplicaSpec> specs;
	if(!to_specs(reps, n, specs) || !tuning)
		return fail(KPZ_EINVAL, "no replicas or null tuning");
	return guard([&] {
		kpz::BatchTuning t = to_tuning(tuning);
		const double ms = kpz::autotuneBatch(specs, mcs, device, t);
		tuning->max_lanes_log2 = t.maxLanesLog2;
		tuning->pq_words64 = t.pqWords64;
		tuning->chunk = t.chunk;
		tuning->concurrent = t.concurrent;
		tuning->sort_by_size = t.sortBySize;
		if(ms_per_mcs)
			*ms_per_mcs = ms;
	});
}

} // extern "C"
