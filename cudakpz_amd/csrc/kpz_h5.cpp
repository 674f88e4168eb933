// HDF5 checkpoints with the reference's dataset names (include/kpz_amd_h5.h):
// SchedulerService::collectH5 / retrieveH5 (kpz/schedulerService.cpp:583-676),
// Correlator::save / fromH5 (kpz/correlator.cpp:709-751) and
// KmcSplash::writeDSFMT / readDSFMT (inc/KMCsplash.cpp:30-54), written with
// the HDF5 C API in the group layout of libSplash's SerialDataCollector
// (/data/<id>/<name>, global attributes in /custom, 2-d datasets stored
// [y][x] with x fastest).  Built as its own library so the engine does not
// depend on HDF5.
#include "../../include/kpz_amd_h5.h"
#include "kpz_capi_internal.hpp"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

using kpz::Error;

// owner of one HDF5 identifier
class Hid {
public:
	using Closer = herr_t (*)(hid_t);
	Hid() = default;
	Hid(hid_t id, Closer c, const std::string& what) : id_(id), close_(c)
	{
		if(id_ < 0)
			throw Error(KPZ_EINVAL, "HDF5: " + what);
	}
	Hid(const Hid&) = delete;
	Hid& operator=(const Hid&) = delete;
	Hid(Hid&& o) noexcept : id_(o.id_), close_(o.close_) { o.id_ = -1; }
	~Hid()
	{
		if(id_ >= 0 && close_)
			close_(id_);
	}
	operator hid_t() const { return id_; }

private:
	hid_t id_ = -1;
	Closer close_ = nullptr;
};

void ok(herr_t e, const std::string& what)
{
	if(e < 0)
		throw Error(KPZ_EINVAL, "HDF5: " + what);
}

// HDF5 prints its error stack by default; the C ABI reports errors itself
struct QuietErrors {
	H5E_auto2_t fn = nullptr;
	void* data = nullptr;
	QuietErrors()
	{
		H5Eget_auto2(H5E_DEFAULT, &fn, &data);
		H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
	}
	~QuietErrors() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

std::string fullName(const char* file)
{
	std::string f(file);
	if(f.size() >= 3 && f.compare(f.size() - 3, 3, ".h5") == 0)
		return f;
	return f + "_0_0_0.h5";
}

bool linkExists(hid_t loc, const std::string& name) { return H5Lexists(loc, name.c_str(), H5P_DEFAULT) > 0; }
bool attrExists(hid_t obj, const std::string& name) { return H5Aexists(obj, name.c_str()) > 0; }

Hid openOrCreateGroup(hid_t loc, const std::string& name)
{
	if(linkExists(loc, name))
		return Hid(H5Gopen2(loc, name.c_str(), H5P_DEFAULT), H5Gclose, "open group " + name);
	return Hid(H5Gcreate2(loc, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
			"create group " + name);
}

void writeDataset(hid_t grp, const std::string& name, hid_t type, int rank, const hsize_t* dims, const void* data)
{
	Hid space(H5Screate_simple(rank, dims, nullptr), H5Sclose, "dataspace " + name);
	Hid ds(H5Dcreate2(grp, name.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
			"create dataset " + name);
	ok(H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write " + name);
}

void writeAttr(hid_t obj, const std::string& name, hid_t type, const void* data)
{
	if(attrExists(obj, name))
		ok(H5Adelete(obj, name.c_str()), "replace attribute " + name);
	Hid space(H5Screate(H5S_SCALAR), H5Sclose, "scalar dataspace");
	Hid a(H5Acreate2(obj, name.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose,
			"create attribute " + name);
	ok(H5Awrite(a, type, data), "write attribute " + name);
}

// a dataset's extents in HDF5 order (slowest first)
std::vector<hsize_t> extents(hid_t grp, const std::string& name)
{
	Hid ds(H5Dopen2(grp, name.c_str(), H5P_DEFAULT), H5Dclose, "open dataset " + name);
	Hid space(H5Dget_space(ds), H5Sclose, "dataspace of " + name);
	const int rank = H5Sget_simple_extent_ndims(space);
	if(rank < 0)
		throw Error(KPZ_EINVAL, "HDF5: rank of " + name);
	std::vector<hsize_t> d((size_t)rank);
	if(rank)
		ok(H5Sget_simple_extent_dims(space, d.data(), nullptr), "extents of " + name);
	return d;
}

template<class T>
std::vector<T> readDataset(hid_t grp, const std::string& name, hid_t memType, size_t expect)
{
	Hid ds(H5Dopen2(grp, name.c_str(), H5P_DEFAULT), H5Dclose, "open dataset " + name);
	Hid space(H5Dget_space(ds), H5Sclose, "dataspace of " + name);
	const hssize_t n = H5Sget_simple_extent_npoints(space);
	if(n < 0 || (size_t)n != expect)
		throw Error(KPZ_EINVAL, "HDF5: " + name + " has " + std::to_string((long long)n) + " elements, expected " +
				std::to_string(expect));
	std::vector<T> v(expect);
	ok(H5Dread(ds, memType, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "read " + name);
	return v;
}

template<class T>
bool readAttr(hid_t obj, const std::string& name, hid_t memType, T* out)
{
	if(!attrExists(obj, name))
		return false;
	Hid a(H5Aopen(obj, name.c_str(), H5P_DEFAULT), H5Aclose, "open attribute " + name);
	ok(H5Aread(a, memType, out), "read attribute " + name);
	return true;
}

int log2exact(hsize_t v)
{
	if(v == 0 || (v & (v - 1)))
		return -1;
	int l = 0;
	while((hsize_t(1) << l) < v)
		++l;
	return l;
}

// dsfmt_t of the reference build: 192 x 128-bit words + int idx, padded to
// 8 bytes (its Makefile defines no HAVE_SSE2, so w128_t is 8-byte aligned)
constexpr size_t DSFMT_WORDS = 770;

herr_t collectId(hid_t, const char* name, const H5L_info_t*, void* op)
{
	auto* ids = static_cast<std::vector<long long>*>(op);
	char* end = nullptr;
	const long long v = std::strtoll(name, &end, 10);
	if(end != name && *end == '\0' && v >= 0 && v <= 0xFFFFFFFFll)
		ids->push_back(v);
	return 0;
}

void writeH5(kpz::SchedulerService& s, const char* file, const std::string& pre, uint32_t pmo)
{
	QuietErrors quiet;
	const std::string fname = fullName(file);
	const auto& sz = s.size();
	const hsize_t dims[2] = {hsize_t(1) << (sz.ly - 2), hsize_t(1) << (sz.lx - 2)};
	const uint64_t words = sz.sizeW();
	const unsigned int mcs = s.mcsCount();

	std::vector<uint32_t> lat((size_t)words);
	s.popSystem(lat.data());
	std::vector<uint32_t> dsfmt(DSFMT_WORDS, 0u);
	s.master().save(reinterpret_cast<uint8_t*>(dsfmt.data()));
	std::vector<uint64_t> streams;
	std::vector<uint32_t> tmt;
	if(s.tinyMT())
	{
		tmt.resize((size_t)s.threads() * 7);
		s.getTinyMT(tmt.data());
	}
	else
	{
		streams.resize((size_t)s.threads());
		s.getStreams(streams.data());
	}
	std::vector<kpz::SchedulerService::TagImage> tags = s.exportTags();

	Hid f(H5Fcreate(fname.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "cannot create " + fname);
	Hid data = openOrCreateGroup(f, "data");
	Hid custom = openOrCreateGroup(f, "custom");
	{
		Hid g = openOrCreateGroup(data, std::to_string(mcs));
		writeDataset(g, pre + ".system", H5T_NATIVE_UINT32, 2, dims, lat.data());
		if(s.hasDisorder())
		{
			s.getDisorder(lat.data());
			writeDataset(g, pre + ".disorder", H5T_NATIVE_UINT32, 2, dims, lat.data());
		}
		const hsize_t nd = DSFMT_WORDS;
		writeDataset(g, pre + ".dsfmt", H5T_NATIVE_UINT32, 1, &nd, dsfmt.data());
		if(s.tinyMT())
		{
			// TinyMT32::writeH5 (kpz/GPUTinyMT.cu:113-120): the 7 planes of T
			// words, libSplash Dimensions(T, 7) = HDF5 extents {7, T}
			const hsize_t td[2] = {7, (hsize_t)s.threads()};
			writeDataset(g, pre + ".tinyMT", H5T_NATIVE_UINT32, 2, td, tmt.data());
		}
		else
		{
			const hsize_t nt = streams.size();
			writeDataset(g, pre + ".slcg64", H5T_NATIVE_UINT64, 1, &nt, streams.data());
			Hid ds(H5Dopen2(g, (pre + ".slcg64").c_str(), H5P_DEFAULT), H5Dclose, "open .slcg64");
			const uint64_t skip = s.skip();
			const int64_t threads = s.threads();
			writeAttr(ds, "skip", H5T_NATIVE_UINT64, &skip);
			writeAttr(ds, "threads", H5T_NATIVE_INT64, &threads);
		}
	}
	const double nSites = std::ldexp(1.0, sz.lx + sz.ly);
	for(const auto& t : tags)
	{
		Hid g = openOrCreateGroup(data, std::to_string(t.mcs));
		const std::string sys = pre + ".system";
		// a tag taken at the current MCS shares the group and its .system
		// (the snapshot equals the lattice then)
		if(!linkExists(g, sys))
			writeDataset(g, sys, H5T_NATIVE_UINT32, 2, dims, t.snap.data());
		writeAttr(g, pre + ".corrTag", H5T_NATIVE_UINT32, &t.stop);
		Hid ds(H5Dopen2(g, sys.c_str(), H5P_DEFAULT), H5Dclose, "open tag snapshot");
		const double avgH = (double)t.sumH / nSites;
		writeAttr(ds, "avgH", H5T_NATIVE_DOUBLE, &avgH);
	}
	{
		const hsize_t four = 4;
		Hid d4(H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, &four), H5Tclose, "array type");
		double p[4], q[4];
		for(unsigned n = 0; n < 4; ++n)
		{
			p[n] = s.rateP(n);
			q[n] = s.rateQ(n);
		}
		writeAttr(custom, pre + ".disorderP", d4, p);
		writeAttr(custom, pre + ".disorderQ", d4, q);
		writeAttr(custom, "prevMeasureOverride", H5T_NATIVE_UINT32, &pmo);
	}
	ok(H5Fflush(f, H5F_SCOPE_GLOBAL), "flush " + fname);
}

std::unique_ptr<kpz::SchedulerService> readH5(const char* file, const std::string& pre, int kind,
		const kpz::Params& prm, uint32_t* pmo, uint32_t* restored)
{
	QuietErrors quiet;
	const std::string fname = fullName(file);
	uint32_t got = 0;
	Hid f(H5Fopen(fname.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "cannot open " + fname);
	if(!linkExists(f, "data"))
		throw Error(KPZ_EINVAL, fname + ": no /data group");
	Hid data(H5Gopen2(f, "data", H5P_DEFAULT), H5Gclose, "open /data");
	std::vector<long long> ids;
	hsize_t idx = 0;
	ok(H5Literate(data, H5_INDEX_NAME, H5_ITER_NATIVE, &idx, collectId, &ids), "list /data");
	if(ids.empty())
		throw Error(KPZ_EINVAL, fname + ": no ids under /data");
	std::sort(ids.begin(), ids.end());
	const long long mcs = ids.back();
	const std::string sys = pre + ".system";
	Hid g(H5Gopen2(data, std::to_string(mcs).c_str(), H5P_DEFAULT), H5Gclose, "open /data/" + std::to_string(mcs));
	if(!linkExists(g, sys))
		throw Error(KPZ_EINVAL, fname + ": no " + sys + " at id " + std::to_string(mcs));
	const std::vector<hsize_t> d = extents(g, sys);
	// libSplash keeps a 2-d selection as (x, y, 1); in HDF5 order [y][x]
	const int lwy = d.size() == 2 ? log2exact(d[0]) : -1, lwx = d.size() == 2 ? log2exact(d[1]) : -1;
	if(lwx < 0 || lwy < 0)
		throw Error(KPZ_EINVAL, fname + ": " + sys + " is not a 2-d power-of-two lattice");
	const int lx = lwx + 2, ly = lwy + 2;
	const size_t words = size_t(1) << (lwx + lwy);

	// read everything before any device work
	std::vector<uint32_t> lat = readDataset<uint32_t>(g, sys, H5T_NATIVE_UINT32, words);
	std::vector<uint32_t> dis;
	if(linkExists(g, pre + ".disorder") && extents(g, pre + ".disorder") == d)
		dis = readDataset<uint32_t>(g, pre + ".disorder", H5T_NATIVE_UINT32, words);
	std::vector<uint8_t> master;
	if(linkExists(g, pre + ".dsfmt"))
	{
		const std::vector<hsize_t> dd = extents(g, pre + ".dsfmt");
		hsize_t n = 1;
		for(hsize_t v : dd)
			n *= v;
		// 770 words (8-byte aligned build) or 772 (16-byte aligned build)
		if(n * 4 >= kpz::MasterRng::STATE_BYTES && n * 4 <= kpz::MasterRng::STATE_BYTES + 12)
		{
			std::vector<uint32_t> w = readDataset<uint32_t>(g, pre + ".dsfmt", H5T_NATIVE_UINT32, (size_t)n);
			master.assign(reinterpret_cast<uint8_t*>(w.data()),
					reinterpret_cast<uint8_t*>(w.data()) + kpz::MasterRng::STATE_BYTES);
		}
	}
	std::vector<uint64_t> streams;
	uint64_t fileSkip = 0;
	int64_t fileThreads = -1;
	if(linkExists(g, pre + ".slcg64"))
	{
		const std::vector<hsize_t> sd = extents(g, pre + ".slcg64");
		Hid ds(H5Dopen2(g, (pre + ".slcg64").c_str(), H5P_DEFAULT), H5Dclose, "open .slcg64");
		if(sd.size() == 1 && readAttr(ds, "skip", H5T_NATIVE_UINT64, &fileSkip) &&
		   readAttr(ds, "threads", H5T_NATIVE_INT64, &fileThreads) && (hsize_t)fileThreads == sd[0])
			streams = readDataset<uint64_t>(g, pre + ".slcg64", H5T_NATIVE_UINT64, (size_t)sd[0]);
	}
	// TinyMT32::readH5 (kpz/GPUTinyMT.cu:122-144): {7, T} planes
	std::vector<uint32_t> tmt;
	hsize_t tmtStreams = 0;
	if(linkExists(g, pre + ".tinyMT"))
	{
		const std::vector<hsize_t> td = extents(g, pre + ".tinyMT");
		if(td.size() == 2 && td[0] == 7)
		{
			tmtStreams = td[1];
			tmt = readDataset<uint32_t>(g, pre + ".tinyMT", H5T_NATIVE_UINT32, (size_t)(7 * td[1]));
		}
	}
	struct TagIn {
		unsigned int mcs, stop;
		std::vector<uint32_t> snap;
	};
	std::vector<TagIn> tags;
	for(long long id : ids)
	{
		Hid tg(H5Gopen2(data, std::to_string(id).c_str(), H5P_DEFAULT), H5Gclose, "open /data/" + std::to_string(id));
		uint32_t stop;
		if(!readAttr(tg, pre + ".corrTag", H5T_NATIVE_UINT32, &stop) &&
		   !readAttr(tg, pre + "corrTag", H5T_NATIVE_UINT32, &stop))
			continue;
		if(!linkExists(tg, sys) || extents(tg, sys) != d)
			throw Error(KPZ_EINVAL, fname + ": tag at " + std::to_string(id) + " has no matching " + sys);
		tags.push_back(TagIn{(unsigned int)id, stop, readDataset<uint32_t>(tg, sys, H5T_NATIVE_UINT32, words)});
	}
	double rp[4], rq[4];
	bool rates = false;
	uint32_t pmoVal = 0;
	bool hasPmo = false;
	// global attributes: libSplash's serial collector keeps them in /custom
	for(const char* where : {"custom", "/"})
	{
		if(std::strcmp(where, "/") && !linkExists(f, where))
			continue;
		Hid grp(H5Gopen2(f, where, H5P_DEFAULT), H5Gclose, std::string("open ") + where);
		const hsize_t four = 4;
		Hid d4(H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, &four), H5Tclose, "array type");
		if(!rates && readAttr(grp, pre + ".disorderP", d4, rp) && readAttr(grp, pre + ".disorderQ", d4, rq))
			rates = true;
		if(!hasPmo)
			hasPmo = readAttr(grp, "prevMeasureOverride", H5T_NATIVE_UINT32, &pmoVal);
	}

	std::unique_ptr<kpz::SchedulerService> s = kpz::makeScheduler(kind, lx, ly, prm);
	s->pushSystem(lat.data());
	s->setMcs((unsigned int)mcs);
	if(rates)
	{
		s->setPQ(rp[0], rq[0]);
		for(unsigned n = 1; n < 4; ++n)
			s->setPQ(rp[n], rq[n], n);
		got |= KPZ_H5_RATES;
	}
	if(!dis.empty())
	{
		s->setDisorder(dis.data());
		got |= KPZ_H5_DISORDER;
	}
	for(const TagIn& t : tags)
		s->importTag(t.mcs, t.stop, t.snap.data());
	if(!master.empty())
	{
		s->master().load(master.data());
		got |= KPZ_H5_MASTER;
	}
	if(s->tinyMT())
	{
		// the parameter sets travel with the state; without them the streams
		// cannot be re-seeded here
		if(tmt.empty() || tmtStreams != (hsize_t)s->threads())
			throw Error(KPZ_EINVAL, fname + ": no " + pre + ".tinyMT state for " +
			                        std::to_string(s->threads()) + " TinyMT32 streams");
		s->setTinyMT(tmt.data());
		got |= KPZ_H5_STREAMS;
	}
	else if(!streams.empty() && fileThreads == s->threads() && fileSkip == s->skip())
	{
		s->setStreams(streams.data());
		got |= KPZ_H5_STREAMS;
	}
	else
		s->randomize(); // retrieveH5 of the GPU schedulers, kpz/schedulerDT.cu:161-171
	s->sync();
	if(hasPmo)
		got |= KPZ_H5_PMO;
	if(pmo)
		*pmo = hasPmo ? pmoVal : 0;
	if(restored)
		*restored = got;
	return s;
}

} // namespace

extern "C" {

int kpz_write_h5(kpz_scheduler* h, const char* file, const char* prefix, uint32_t prevMeasureOverride)
{
	if(!h || !h->s)
		return kpz::capiFail(KPZ_EINVAL, "null scheduler handle");
	if(!file)
		return kpz::capiFail(KPZ_EINVAL, "null file name");
	return kpz::capiGuard([&] { writeH5(*h->s, file, prefix ? prefix : "", prevMeasureOverride); });
}

int kpz_read_h5(const char* file, const char* prefix, int kind, const kpz_params* params,
		kpz_scheduler** out, uint32_t* prevMeasureOverride, uint32_t* restored)
{
	if(!out || !file)
		return kpz::capiFail(KPZ_EINVAL, "null argument");
	*out = nullptr;
	kpz::Params p;
	if(params)
	{
		if(params->threads < 0)
			return kpz::capiFail(KPZ_EINVAL, "threads must be >= 0");
		p.threads = params->threads;
		p.layout_x = params->layout_x ? params->layout_x : 4;
		p.layout_y = params->layout_y ? params->layout_y : 4;
		p.skip = params->skip;
		p.device = params->device;
		p.rng = params->rng;
		if(p.skip && !(p.skip & 1))
			return kpz::capiFail(KPZ_EINVAL, "SLCG skip must be odd");
	}
	return kpz::capiGuard([&] {
		auto h = new kpz_scheduler;
		try
		{
			h->s = readH5(file, prefix ? prefix : "", kind, p, prevMeasureOverride, restored);
		}
		catch(...)
		{
			delete h;
			throw;
		}
		*out = h;
	});
}

} // extern "C"
