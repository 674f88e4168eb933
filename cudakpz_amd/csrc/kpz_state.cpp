// Checkpoint / resume of one scheduler (the role of collectH5 / retrieveH5,
// kpz/schedulerService.cpp:583-676, kpz/kpzScaling.cpp:118-139, 217-222).
//
// One little-endian file holds everything the scheduler needs to continue
// bit for bit: kind, sizes and construction parameters, MCS count, (p, q),
// the master dSFMT state (dsfmt_t bytes), every stream state (SLCG64, which the
// reference never saved, or TinyMT32 as its `.tinyMT` dataset,
// kpz/GPUTinyMT.cu:113-144), the lattice in reference layout, the rates of the
// disorder classes and the disorder map (collectH5 writes .disorderP/Q and
// .disorder), and every live correlator tag (MCS, stop time, snapshot).  An
// FNV-1a 64 checksum over the file's bytes ends it.
//
//   "KPZSTAT1"  u32 version (3; 2 is read as well: SLCG64 streams)
//   i32 kind, lx, ly | i64 threads | i32 layout_x, layout_y | u64 skip
//   u32 mcs, aux | f64 p, q | f64 p1..p3, q1..q3 | u32 classes set (bits 1..3)
//   u32 master bytes | u64 streams | u64 words | u32 tags | u32 has disorder map
//   [version 3: u32 generator, 0 = SLCG64, 1 = TinyMT32]
//   master bytes | streams x u64 (SLCG64) or 7 x streams x u32 (TinyMT32:
//   status[0..3], mat1, mat2, tmat planes) | words x u32 | [map: words x u32]
//   tags x (u32 mcs, u32 stop, words x u32) | u64 checksum
#include "kpz_scheduler.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

namespace kpz {

namespace {

constexpr char MAGIC[8] = {'K', 'P', 'Z', 'S', 'T', 'A', 'T', '1'};
constexpr uint32_t VERSION = 3;

struct Fnv {
	uint64_t h = 1469598103934665603ull;
	void add(const void* p, size_t n)
	{
		const unsigned char* b = (const unsigned char*)p;
		for(size_t i = 0; i < n; ++i)
		{
			h ^= b[i];
			h *= 1099511628211ull;
		}
	}
};

struct Writer {
	FILE* f;
	Fnv sum;
	void put(const void* p, size_t n)
	{
		if(n && std::fwrite(p, 1, n, f) != n)
			throw Error(-1, "state file: write failed");
		sum.add(p, n);
	}
	template<class T> void val(T v) { put(&v, sizeof v); }
};

struct Reader {
	const std::vector<unsigned char>& b;
	size_t pos = 0;
	void get(void* p, size_t n)
	{
		if(pos + n > b.size())
			throw Error(-1, "state file: truncated");
		std::memcpy(p, b.data() + pos, n);
		pos += n;
	}
	template<class T> T val()
	{
		T v;
		get(&v, sizeof v);
		return v;
	}
};

struct Header {
	int32_t kind, lx, ly;
	int64_t threads;
	int32_t layoutX, layoutY;
	uint64_t skip;
	uint32_t mcs, aux;
	double p, q, cp[3], cq[3];
	uint32_t classesSet;
	uint32_t masterBytes;
	uint64_t streams, words;
	uint32_t tags, hasMap;
	uint32_t rng;
};

} // namespace

std::vector<SchedulerService::TagImage> SchedulerService::exportTags()
{
	const uint64_t words = size_.sizeW();
	std::vector<TagImage> out;
	for(const Tag& t : tags_)
	{
		TagImage im;
		im.mcs = t.mcs;
		im.stop = t.stop;
		im.snap.resize((size_t)words);
		check(hipMemcpyAsync(im.snap.data(), t.snap->get(), words * 4, hipMemcpyDeviceToHost, stream_), "tag D2H");
		sync();
		im.sumH = correlateWith(*t.snap).sumHs;
		out.push_back(std::move(im));
	}
	return out;
}

void SchedulerService::importTag(unsigned int mcs, unsigned int stop, const uint32_t* snap)
{
	const uint64_t words = size_.sizeW();
	Tag t;
	t.mcs = mcs;
	t.stop = stop;
	t.snap = std::make_unique<DevBuf>(words * 4);
	check(hipMemcpyAsync(t.snap->get(), snap, words * 4, hipMemcpyHostToDevice, stream_), "tag H2D");
	sync();
	tags_.push_back(std::move(t));
}

void SchedulerService::saveState(const char* path, uint32_t aux)
{
	if(kind_ < 0)
		throw Error(-1, "saveState: scheduler kind unknown");
	const uint64_t words = size_.sizeW();
	std::vector<uint8_t> master(MasterRng::STATE_BYTES);
	master_.save(master.data());
	std::vector<uint64_t> st;
	std::vector<uint32_t> tm;
	if(tinyMT())
	{
		tm.resize((size_t)T_ * 7);
		getTinyMT(tm.data());
	}
	else
	{
		st.resize((size_t)T_);
		getStreams(st.data());
	}
	std::vector<uint32_t> lat((size_t)words);
	popSystem(lat.data());

	std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "wb"), &std::fclose);
	if(!f)
		throw Error(-1, std::string("saveState: cannot open ") + path);
	Writer w{f.get()};
	w.put(MAGIC, sizeof MAGIC);
	w.val(VERSION);
	w.val((int32_t)kind_);
	w.val((int32_t)size_.lx);
	w.val((int32_t)size_.ly);
	w.val((int64_t)T_);
	w.val((int32_t)params_.layout_x);
	w.val((int32_t)params_.layout_y);
	w.val((uint64_t)S_);
	w.val((uint32_t)mcs_);
	w.val(aux);
	w.val(p_);
	w.val(q_);
	uint32_t set = 0;
	for(int k = 1; k < 4; ++k)
	{
		w.val(cp_[k]);
		set |= cset_[k] ? (1u << k) : 0u;
	}
	for(int k = 1; k < 4; ++k)
		w.val(cq_[k]);
	w.val(set);
	w.val((uint32_t)master.size());
	w.val((uint64_t)T_);
	w.val(words);
	w.val((uint32_t)tags_.size());
	w.val((uint32_t)(hasDisorder() ? 1 : 0));
	w.val((uint32_t)params_.rng);
	w.put(master.data(), master.size());
	if(tinyMT())
		w.put(tm.data(), tm.size() * 4);
	else
		w.put(st.data(), st.size() * 8);
	w.put(lat.data(), lat.size() * 4);
	if(hasDisorder())
	{
		getDisorder(lat.data());
		w.put(lat.data(), lat.size() * 4);
	}
	for(const Tag& t : tags_)
	{
		w.val((uint32_t)t.mcs);
		w.val((uint32_t)t.stop);
		check(hipMemcpyAsync(lat.data(), t.snap->get(), words * 4, hipMemcpyDeviceToHost, stream_), "tag D2H");
		sync();
		w.put(lat.data(), lat.size() * 4);
	}
	const uint64_t sum = w.sum.h;
	if(std::fwrite(&sum, 8, 1, f.get()) != 1)
		throw Error(-1, "state file: write failed");
	if(std::fflush(f.get()) != 0)
		throw Error(-1, "state file: write failed");
}

std::unique_ptr<SchedulerService> SchedulerService::loadState(const char* path, int device, uint32_t* aux)
{
	// read and validate the whole file before any device work
	std::vector<unsigned char> buf;
	{
		std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), &std::fclose);
		if(!f)
			throw Error(-1, std::string("loadState: cannot open ") + path);
		unsigned char chunk[1 << 16];
		size_t n;
		while((n = std::fread(chunk, 1, sizeof chunk, f.get())) > 0)
			buf.insert(buf.end(), chunk, chunk + n);
	}
	if(buf.size() < sizeof MAGIC + 8 || std::memcmp(buf.data(), MAGIC, sizeof MAGIC))
		throw Error(-1, "loadState: not a state file");
	Fnv sum;
	sum.add(buf.data(), buf.size() - 8);
	uint64_t stored;
	std::memcpy(&stored, buf.data() + buf.size() - 8, 8);
	if(stored != sum.h)
		throw Error(-1, "loadState: checksum mismatch (file damaged)");
	Reader r{buf};
	r.pos = sizeof MAGIC;
	const uint32_t version = r.val<uint32_t>();
	if(version != 2 && version != VERSION)
		throw Error(-1, "loadState: unsupported version");
	Header h;
	h.kind = r.val<int32_t>();
	h.lx = r.val<int32_t>();
	h.ly = r.val<int32_t>();
	h.threads = r.val<int64_t>();
	h.layoutX = r.val<int32_t>();
	h.layoutY = r.val<int32_t>();
	h.skip = r.val<uint64_t>();
	h.mcs = r.val<uint32_t>();
	h.aux = r.val<uint32_t>();
	h.p = r.val<double>();
	h.q = r.val<double>();
	for(int k = 0; k < 3; ++k)
		h.cp[k] = r.val<double>();
	for(int k = 0; k < 3; ++k)
		h.cq[k] = r.val<double>();
	h.classesSet = r.val<uint32_t>();
	h.masterBytes = r.val<uint32_t>();
	h.streams = r.val<uint64_t>();
	h.words = r.val<uint64_t>();
	h.tags = r.val<uint32_t>();
	h.hasMap = r.val<uint32_t>();
	h.rng = (version >= 3) ? r.val<uint32_t>() : (uint32_t)RNG_SLCG64;
	if(h.kind < KIND_DT || h.kind > KIND_SCA || h.lx < 2 || h.ly < 2 || h.lx > 30 || h.ly > 30 ||
	   h.words != (1ull << (h.lx + h.ly - 4)) || h.masterBytes != (uint32_t)MasterRng::STATE_BYTES ||
	   h.streams == 0 || (h.threads > 0 && (uint64_t)h.threads != h.streams) || h.hasMap > 1 ||
	   (h.classesSet & ~0xEu) || h.rng > RNG_TINYMT32)
		throw Error(-1, "loadState: inconsistent header");
	const size_t streamBytes = h.rng == RNG_TINYMT32 ? 28 : 8;
	const size_t need = r.pos + h.masterBytes + h.streams * streamBytes + h.words * 4 * (1 + h.hasMap) +
	                    (size_t)h.tags * (8 + h.words * 4) + 8;
	if(need != buf.size())
		throw Error(-1, "loadState: size does not match the header");

	Params prm;
	prm.threads = (int64_t)h.streams;
	prm.layout_x = h.layoutX;
	prm.layout_y = h.layoutY;
	prm.skip = h.skip;
	prm.device = device;
	prm.rng = (int)h.rng;
	std::unique_ptr<SchedulerService> s = makeScheduler(h.kind, h.lx, h.ly, prm);
	if(s->threads() != (int64_t)h.streams || s->skip() != h.skip)
		throw Error(-1, "loadState: stream layout differs");
	s->setPQ(h.p, h.q);
	for(int k = 1; k < 4; ++k)
	{
		s->cp_[k] = h.cp[k - 1];
		s->cq_[k] = h.cq[k - 1];
		s->cset_[k] = (h.classesSet >> k) & 1;
	}
	s->master_.load(buf.data() + r.pos);
	r.pos += h.masterBytes;
	if(h.rng == RNG_TINYMT32)
		s->setTinyMT(reinterpret_cast<const uint32_t*>(buf.data() + r.pos));
	else
		s->setStreams(reinterpret_cast<const uint64_t*>(buf.data() + r.pos));
	r.pos += h.streams * streamBytes;
	s->pushSystem(reinterpret_cast<const uint32_t*>(buf.data() + r.pos));
	r.pos += h.words * 4;
	if(h.hasMap)
	{
		s->setDisorder(reinterpret_cast<const uint32_t*>(buf.data() + r.pos));
		r.pos += h.words * 4;
	}
	for(uint32_t k = 0; k < h.tags; ++k)
	{
		Tag t;
		t.mcs = r.val<uint32_t>();
		t.stop = r.val<uint32_t>();
		t.snap = std::make_unique<DevBuf>(h.words * 4);
		check(hipMemcpyAsync(t.snap->get(), buf.data() + r.pos, h.words * 4, hipMemcpyHostToDevice, s->stream_),
				"tag H2D");
		s->sync();
		r.pos += h.words * 4;
		s->tags_.push_back(std::move(t));
	}
	s->mcs_ = h.mcs;
	s->sync();
	if(aux)
		*aux = h.aux;
	return s;
}

} // namespace kpz
