// kpz_scaling: the reference's kpzScaling front end (kpz/kpzScaling.cpp:47-237)
// on the MI355X engine, through the C ABI of include/kpz_amd.h.
//
//   kpz_scaling mcs [lx [ly]] [options]
//
// Arguments, sampling and event loop follow kpz/kpzFrontend.cpp:400-859 and
// kpz/sampling.{h,cpp}; the output table is the reference's: one line per
// measurement, "mcs  w2  <h> <h^2> ... <h^7>" (operator<< of Roughness,
// kpz/schedulerService.cpp:73-79), with "  tagMcs  ch  cs" per live
// correlator tag (kpz/schedulerService.cpp:216-251, kpz/correlator.cpp:754),
// and "mcs  w2  <h>...<h^7>  set corr tag" when a tag is taken, printed with
// 12 significant digits.  All analysis runs on the device: w2, ch and the
// slope overlap are exact, the higher moments are double per row summed in
// long double.
#include "kpz_amd.h"
#ifdef KPZ_WITH_H5
#include "kpz_amd_h5.h"
#endif

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <list>
#include <sstream>
#include <string>
#include <vector>

namespace {

const char* VERSION = "kpz_scaling (MI355X engine, gfx950)";

// ---------------------------------------------------------------- arguments
template<class T>
bool getNum(T& v, int a, int argc, char** argv)
{
	if(a >= argc)
		return false;
	std::istringstream s(argv[a]);
	T t;
	s >> t;
	if(s.fail() || !s.eof())
		return false;
	v = t;
	return true;
}

// SamplingParam (kpz/sampling.h:55-96)
struct Sampling {
	enum Kind { EXP, LINEAR, RANDOM } kind = EXP;
	double density = .1;
	unsigned interval = 0;
	double mean = 0;
};

// SamplingEvent kinds (kpz/sampling.h:27-112)
struct Event {
	enum Kind { BARRIER, END, SAMPLING, TAG } kind;
	unsigned start;
	bool countAsMeasurement;
	Sampling sampling;
	unsigned stop;
};

struct Args {
	unsigned mcs = 0;
	int lx = 10, ly = 10;
	std::string scheduler = "default";
	int64_t mcsThreads = 0;
	int dd[2] = {0, 0};
	double p[4] = {NAN, NAN, NAN, NAN}, q[4] = {NAN, NAN, NAN, NAN}; // --pq[n] per disorder class
	double disorderFill = NAN;                                         // --disorder2 [fill]
	int anaThreads = 1;
	bool countTags = false, save = false;
	double maxH = 0;
	std::string file, saveState, loadState, saveH5, loadH5;
	bool haveSeed = false;
	uint64_t seed = 0;
	int device = -1;
	int rng = KPZ_RNG_SLCG64;     // --rng slcg64|tinymt32
	std::string tinymtParams;     // --tinymtParams file (tinymt32dc output)
	Sampling current;
	std::list<Event> events;
	bool finished = false;
};

void help()
{
	std::cout << "kpz_scaling mcs [x [y]] [options]\n"
	             "  x, y: log2 lattice sizes (default 10; y defaults to x)\n"
	             "  --scheduler dt|dtStatic|scaBit|sca|cb|default [--mcsThreads n] [--ddBlocksize bx [by]]\n"
	             "  --pq p [q]                rates (mcsPQSyncRng); --pq1..--pq3: disorder classes 1..3\n"
	             "  --disorder2 [fill]        disorder map (default fill .5), mcsDisorder2SyncRng with --pq\n"
	             "  --sampleRate d [start]    measurements at ceil(exp(log(t+10)+d)) (default d = 0.1)\n"
	             "  --linSample n [start]     every n MCS\n"
	             "  --randomSample m [start]  uniform in [m/2, 3m/2] MCS apart (master dSFMT)\n"
	             "  --corrStart t [stop]      correlator tag at t, kept until stop\n"
	             "  --countTagsAsMeasurement | --noCountTagsAsMeasurement\n"
	             "  --anaThreads n            accepted; analysis runs on the GPU\n"
	             "  --maxH hours              wall-clock limit (prints TIMEOUT mcs)\n"
	             "  --save | --noSave         write KPZheightmap_<mcs>.xyz at the end\n"
	             "  -f, --file name           output table to a file (-- = stdout)\n"
	             "  --saveState file          checkpoint at the end (lattice, RNG states, tags)\n"
	             "  --loadState file          resume from a checkpoint (continues bit for bit)\n"
	             "  --saveH5 prefix           HDF5 checkpoint <prefix>_0_0_0.h5 (reference dataset names\n"
	             "                            + SLCG stream states)\n"
	             "  --loadH5 prefix           resume from an HDF5 checkpoint (size from the file)\n"
	             "  --seed s                  64-bit seed (default: /dev/urandom and time)\n"
	             "  --rng slcg64|tinymt32     worker generator (tinymt32: dt, sca, scaBit)\n"
	             "  --tinymtParams file       TinyMT32 parameter sets, tinymt32dc output lines\n"
	             "                            (characteristic,type,id,mat1,mat2,tmat,...; set t = line t)\n"
	             "  --device n                HIP device\n";
}

// the leading positional words "mcs [x [y]]" (kpz/kpzFrontend.cpp:470-488);
// returns the index of the first option word, -1 without an MCS count
int parsePositional(Args& A, int argc, char** argv)
{
	if(!getNum(A.mcs, 1, argc, argv))
	{
		std::cerr << "[EE][kpz_scaling] Number of mcs was not supplied (kpz_scaling mcs [x [y]] ...)\n";
		return -1;
	}
	int sizes[2];
	int n = 0;
	while(n < 2 && getNum(sizes[n], 2 + n, argc, argv))
		++n;
	A.lx = n > 0 ? sizes[0] : 10;
	A.ly = n > 1 ? sizes[1] : A.lx;
	return 2 + n;
}

void pushSampling(Args& A, unsigned start, const Sampling& s)
{
	if(start == 0)
		A.current = s;
	else
		A.events.push_back(Event{Event::SAMPLING, start, true, s, 0});
}

// ---------------------------------------------------------------- option table
// Options are rows of a table: names and a handler.  A handler reads its
// values through the cursor: need() takes the next word or fails with the
// given message, maybe() takes it only when it parses (else the default
// stays).  Option names and meanings are those of kpz/kpzFrontend.cpp:93-232
// (scheduler options) and :494-733 (front-end options); scheduler options are
// looked up first while the words after --scheduler are read, as there.
struct Cursor {
	int argc;
	char** argv;
	int a;              // index of the option word being handled
	bool sched = false; // inside the scheduler's own options

	const char* word() const { return argv[a]; }
	template<class T>
	bool maybe(T& v)
	{
		if(!getNum(v, a + 1, argc, argv))
			return false;
		++a;
		return true;
	}
	bool maybeStr(std::string& v)
	{
		if(a + 1 >= argc)
			return false;
		v = argv[++a];
		return true;
	}
	template<class T>
	bool need(T& v, const std::string& err)
	{
		if(maybe(v))
			return true;
		std::cerr << err << '\n';
		return false;
	}
	bool needStr(std::string& v, const std::string& err)
	{
		if(maybeStr(v))
			return true;
		std::cerr << err << '\n';
		return false;
	}
};

struct Option {
	std::vector<std::string> names;
	bool (*run)(Cursor&, Args&);
};

bool warn(const std::string& msg)
{
	std::cerr << "[WW][kpz_scaling] " << msg << '\n';
	return true;
}

bool refuse(const std::string& msg)
{
	std::cerr << "[EE][kpz_scaling] " << msg << '\n';
	return false;
}

// a start time after a sampling option: optional, 0 = replace the current rule
bool samplingStart(Cursor& c, Args& A, const Sampling& s)
{
	unsigned start = 0;
	c.maybe(start);
	pushSampling(A, start, s);
	return true;
}

const std::vector<Option>& schedulerOptions()
{
	static const std::vector<Option> t = {
		{{"--yield", "--anaGPU"}, [](Cursor& c, Args&) { return warn(std::string(c.word()) + " does nothing on this engine"); }},
		{{"--byRank"}, [](Cursor& c, Args&) {
			 int rank;
			 c.maybe(rank);
			 return warn("--byRank: one GPU per process here; pick it with --device");
		 }},
		{{"-l", "--localLayout"}, [](Cursor& c, Args&) {
			 std::string id;
			 if(!c.needStr(id, "--localLayout needs a layout preset id"))
				 return false;
			 return warn("--localLayout does nothing on this engine");
		 }},
		{{"--1d"}, [](Cursor&, Args&) { return refuse("--1d: this engine has no 1+1d schedulers"); }},
		{{"--mcsThreads"}, [](Cursor& c, Args& A) {
			 if(!c.need(A.mcsThreads, "--mcsThreads needs a thread count"))
				 return false;
			 if(A.mcsThreads < 1)
			 {
				 A.mcsThreads = 1;
				 warn("--mcsThreads below 1, using 1");
			 }
			 return true;
		 }},
		{{"--ddBlocksize"}, [](Cursor& c, Args& A) {
			 if(!c.need(A.dd[0], "--ddBlocksize needs a block size (log2 x [log2 y])"))
				 return false;
			 if(!c.maybe(A.dd[1]))
				 A.dd[1] = A.dd[0];
			 return true;
		 }},
	};
	return t;
}

const std::vector<Option>& frontendOptions()
{
	static const std::vector<Option> t = {
		{{"--scheduler", "--gpuscheduler", "--cpuscheduler"}, [](Cursor& c, Args& A) {
			 int config = 0;
			 if(c.maybe(config) && config != 0)
				 return refuse("scheduler configuration " + std::to_string(config) + " does not exist (only 0)");
			 if(!c.needStr(A.scheduler, "--scheduler needs a scheduler id"))
				 return false;
			 if(A.scheduler == "cb")
				 A.scheduler = "sca";
			 static const char* const known[] = {"dt", "dtStatic", "scaBit", "sca", "default"};
			 if(std::find_if(std::begin(known), std::end(known), [&](const char* k) { return A.scheduler == k; }) ==
			    std::end(known))
				 return refuse("No such scheduler: " + A.scheduler + " (available: dt, dtStatic, scaBit, sca/cb)");
			 c.sched = true;
			 return true;
		 }},
		{{"--nschedulers", "--mcsFct", "--retryH5", "--prevMeasureOverride"},
		 [](Cursor& c, Args&) { return refuse(std::string(c.word()) + " is not supported by this engine"); }},
		{{"--saveH5", "--loadH5"}, [](Cursor& c, Args& A) {
#ifdef KPZ_WITH_H5
			 const std::string opt = c.word();
			 return c.needStr(opt == "--saveH5" ? A.saveH5 : A.loadH5, opt + " needs an HDF5 file-name prefix");
#else
			 (void)A;
			 return refuse(std::string(c.word()) + ": built without HDF5; --saveState / --loadState do the same");
#endif
		 }},
		{{"--saveState", "--loadState"}, [](Cursor& c, Args& A) {
			 const std::string opt = c.word();
			 return c.needStr(opt == "--saveState" ? A.saveState : A.loadState, opt + " needs a file name");
		 }},
		{{"--disorder2"}, [](Cursor& c, Args& A) {
			 A.disorderFill = 0.5;
			 c.maybe(A.disorderFill);
			 return true;
		 }},
		{{"--pq", "--pq0", "--pq1", "--pq2", "--pq3"}, [](Cursor& c, Args& A) {
			 const std::string opt = c.word();
			 const int n = opt.size() == 5 ? opt[4] - '0' : 0;
			 if(!c.need(A.p[n], opt + " needs p (and optionally q)"))
				 return false;
			 A.q[n] = 0.;
			 c.maybe(A.q[n]);
			 return true;
		 }},
		{{"--anaThreads"}, [](Cursor& c, Args& A) {
			 if(!c.need(A.anaThreads, "--anaThreads needs a thread count"))
				 return false;
			 if(A.anaThreads < 1)
			 {
				 A.anaThreads = 1;
				 warn("--anaThreads below 1, using 1");
			 }
			 return true;
		 }},
		{{"--save", "--noSave"}, [](Cursor& c, Args& A) {
			 A.save = !strcmp(c.word(), "--save");
			 return true;
		 }},
		{{"--countTagsAsMeasurement", "--noCountTagsAsMeasurement"}, [](Cursor& c, Args& A) {
			 A.countTags = !strcmp(c.word(), "--countTagsAsMeasurement");
			 return true;
		 }},
		{{"--maxH"}, [](Cursor& c, Args& A) {
			 float h = 0;
			 if(!c.maybe(h) || h <= 0.f)
				 return refuse("--maxH needs a positive number of hours");
			 A.maxH = h;
			 return true;
		 }},
		{{"--version"}, [](Cursor&, Args&) {
			 std::cout << VERSION << '\n';
			 return true;
		 }},
		{{"--help"}, [](Cursor&, Args&) {
			 help();
			 std::cout << VERSION << '\n';
			 return true;
		 }},
		{{"--sampleRate"}, [](Cursor& c, Args& A) {
			 Sampling s;
			 return c.need(s.density, "--sampleRate needs a rate [start MCS]") && samplingStart(c, A, s);
		 }},
		{{"--linSample"}, [](Cursor& c, Args& A) {
			 Sampling s;
			 s.kind = Sampling::LINEAR;
			 if(!c.maybe(s.interval) || s.interval == 0)
				 return refuse("--linSample needs an interval of at least 1 MCS [start MCS]");
			 return samplingStart(c, A, s);
		 }},
		{{"--randomSample"}, [](Cursor& c, Args& A) {
			 Sampling s;
			 s.kind = Sampling::RANDOM;
			 double mean;
			 if(!c.need(mean, "--randomSample needs a mean interval [start MCS]"))
				 return false;
			 s.mean = mean / 2.; // intervals uniform in [mean/2, 3 mean/2]
			 return samplingStart(c, A, s);
		 }},
		{{"--corrStart"}, [](Cursor& c, Args& A) {
			 unsigned t, stop = 0xFFFFFFFFu;
			 if(!c.need(t, "--corrStart needs the tag time in MCS [stop MCS]"))
				 return false;
			 c.maybe(stop);
			 A.events.push_back(Event{Event::TAG, t, A.countTags, Sampling{}, stop});
			 return true;
		 }},
		{{"--seed"}, [](Cursor& c, Args& A) {
			 A.haveSeed = c.need(A.seed, "--seed needs an unsigned 64-bit integer");
			 return A.haveSeed;
		 }},
		{{"--rng"}, [](Cursor& c, Args& A) {
			 std::string r;
			 if(!c.maybeStr(r) || (r != "slcg64" && r != "tinymt32"))
				 return refuse("--rng takes slcg64 or tinymt32");
			 A.rng = (r == "tinymt32") ? KPZ_RNG_TINYMT32 : KPZ_RNG_SLCG64;
			 return true;
		 }},
		{{"--tinymtParams"}, [](Cursor& c, Args& A) {
			 return c.needStr(A.tinymtParams, "--tinymtParams needs a file name");
		 }},
		{{"--device"}, [](Cursor& c, Args& A) { return c.need(A.device, "--device needs a device ordinal"); }},
		{{"-f", "--file"}, [](Cursor& c, Args& A) {
			 return c.needStr(A.file, "-f / --file needs a file name (-- for stdout)");
		 }},
	};
	return t;
}

const Option* lookup(const std::vector<Option>& table, const char* w)
{
	for(const Option& o : table)
		if(std::find(o.names.begin(), o.names.end(), w) != o.names.end())
			return &o;
	return nullptr;
}

// every word after "mcs [x [y]]"; false on the first bad option
bool parseOptions(Args& A, int argc, char** argv, int first)
{
	Cursor c{argc, argv, first};
	for(; c.a < argc; ++c.a)
	{
		const Option* o = c.sched ? lookup(schedulerOptions(), c.word()) : nullptr;
		if(!o)
		{
			c.sched = false;
			o = lookup(frontendOptions(), c.word());
		}
		if(!o)
		{
			if(!strncmp(c.word(), "--pq", 4))
				return refuse(std::string(c.word()) + ": rate sets are --pq, --pq0 .. --pq3");
			return refuse(std::string("Unknown argument: ") + c.word());
		}
		if(!o->run(c, A))
			return false;
	}
	return true;
}

// TinyMT32 parameter sets from tinymt32dc output, the lines
// inc/tinyMT/genParam.sh turns into tinymt32_param.h: "characteristic,type,
// id,mat1,mat2,tmat,weight,delta" with hex words; '#' lines are comments.
// Set t is the t-th data line, as tinyMTmat1[t] of the generated table.
bool readTinyMTParams(const std::string& file, size_t T, std::vector<uint32_t> (&m)[3])
{
	std::ifstream in(file);
	if(!in)
	{
		std::cerr << "[EE] cannot open TinyMT32 parameter file " << file << '\n';
		return false;
	}
	std::string line;
	size_t lineNo = 0;
	while(m[0].size() < T && std::getline(in, line))
	{
		++lineNo;
		if(line.empty() || line[0] == '#')
			continue;
		std::vector<std::string> f;
		size_t b = 0;
		for(size_t e; (e = line.find(',', b)) != std::string::npos; b = e + 1)
			f.push_back(line.substr(b, e - b));
		f.push_back(line.substr(b));
		if(f.size() < 6)
		{
			std::cerr << "[EE] " << file << ":" << lineNo << ": expected tinymt32dc output (>= 6 fields)\n";
			return false;
		}
		for(int k = 0; k < 3; ++k)
		{
			char* end = nullptr;
			const unsigned long v = std::strtoul(f[(size_t)(3 + k)].c_str(), &end, 16);
			if(end == f[(size_t)(3 + k)].c_str() || v > 0xFFFFFFFFul)
			{
				std::cerr << "[EE] " << file << ":" << lineNo << ": field " << (4 + k) << " is not a hex word\n";
				return false;
			}
			m[k].push_back((uint32_t)v);
		}
	}
	if(m[0].size() < T)
	{
		std::cerr << "[EE] " << file << " holds " << m[0].size() << " TinyMT32 parameter sets, " << T
		          << " workers need one each\n";
		return false;
	}
	return true;
}

// ---------------------------------------------------------------- engine
struct Engine {
	kpz_scheduler* h = nullptr;
	int lx = 0, ly = 0;
	unsigned mcs = 0;
	enum Mode { MCS, PQ, DISORDER } mode = MCS;
	std::ostream* out = &std::cout;

	~Engine()
	{
		if(h)
			kpz_destroy(h);
	}
	static void check(int rc, const char* what)
	{
		if(rc < 0)
		{
			std::cerr << "[EE][kpz_scaling] " << what << ": " << kpz_last_error() << '\n';
			std::exit(1);
		}
	}
	long double N() const { return std::ldexp(1.0L, lx + ly); }

	// roughness line fields: w2 and <h^k>, k = 1..7 (Roughness, operator<<)
	void roughness(long double r[8])
	{
		int64_t s[3];
		double m[7];
		check(kpz_roughness(h, s), "roughness");
		check(kpz_moments(h, m), "moments");
		const __int128 s1 = s[0];
		const __int128 s2 = (__int128)(((unsigned __int128)(uint64_t)s[2] << 64) | (uint64_t)s[1]);
		const long double n = N();
		r[0] = (long double)(((__int128)1 << (lx + ly)) * s2 - s1 * s1) / (n * n);
		r[1] = (long double)s1 / n;
		r[2] = (long double)s2 / n;
		for(int k = 2; k < 7; ++k)
			r[k + 1] = m[k];
	}
	void printRoughness(const long double r[8])
	{
		for(int k = 0; k < 8; ++k)
			*out << '\t' << r[k];
	}

	// SchedulerService::correlate (kpz/schedulerService.cpp:216-251)
	void correlate()
	{
		long double r[8];
		roughness(r);
		std::vector<int64_t> c(8 * 64);
		int n = 0;
		for(;;)
		{
			n = kpz_correlate_tags(h, c.data(), (int)(c.size() / 8));
			check(n, "correlate");
			if((size_t)n < c.size() / 8)
				break;
			c.resize(c.size() * 2);
		}
		std::ostringstream text;
		text << std::setprecision(12);
		std::ostream* keep = out;
		out = &text;
		for(int i = 0; i < n; ++i)
		{
			const int64_t* t = &c[(size_t)i * 8];
			const __int128 sh = t[1], shs = t[2];
			const __int128 shhs = (__int128)(((unsigned __int128)(uint64_t)t[6] << 64) | (uint64_t)t[5]);
			const long double nn = N();
			const long double ch = (long double)(((__int128)1 << (lx + ly)) * shhs - sh * shs) / (nn * nn);
			const long double cs = (long double)t[7] / nn / 2.0L;
			text << mcs;
			printRoughness(r);
			text << '\t' << (unsigned)t[0] << '\t' << ch << '\t' << cs << '\n';
		}
		if(n == 0)
		{
			text << mcs;
			printRoughness(r);
			text << '\n';
		}
		out = keep;
		*out << text.str();
		out->flush();
	}

	// SchedulerService::correlateTag (kpz/schedulerService.cpp:199-211, 279-284)
	void correlateTag(unsigned stop)
	{
		check(kpz_corr_tag_stop(h, stop), "correlateTag");
		long double r[8];
		roughness(r);
		std::ostringstream text;
		text << std::setprecision(12) << mcs;
		std::ostream* keep = out;
		out = &text;
		printRoughness(r);
		out = keep;
		text << "\tset\tcorr\ttag\n";
		*out << text.str();
		out->flush();
	}

	void run(unsigned n)
	{
		if(n)
			check(mode == PQ ? kpz_mcs_pq(h, (int)n) : mode == DISORDER ? kpz_mcs_disorder2(h, (int)n) : kpz_mcs(h, (int)n),
			      "mcs");
		mcs += n;
	}

	// KPZheightmap_<mcs>.xyz (SchedulerService::saveXYZ, kpz/schedulerService.cpp:403-450)
	void saveXYZ(const std::string& name)
	{
		std::vector<uint32_t> lat(kpz_size_words(h));
		check(kpz_download(h, lat.data()), "download");
		const int enc = kpz_encoding(h);
		const int X = 1 << lx, Y = 1 << ly;
		auto bit = [&](int x, int y, int d) -> int {
			if(enc == KPZ_ENC_LOCAL)
			{
				const size_t w = ((size_t)(y >> 2) << (lx - 2)) + (size_t)(x >> 2);
				return (lat[w] >> ((((x & 3) << 1) | ((y & 3) << 3)) + d)) & 1;
			}
			const size_t q = (size_t)1 << (lx + ly - 6);
			const int sub = (x ^ y) & 1;
			const size_t w = ((size_t)y << (lx - 6)) + (size_t)(x >> 6);
			return (lat[(size_t)(2 * sub + d) * q + w] >> ((x >> 1) & 31)) & 1;
		};
		std::ofstream o(name);
		o << (4 + (uint64_t)X * Y) << "\n\n";
		o << "H\t0\t0\t0\t0\n"
		  << "H\t" << X << "\t0\t0\t0\n"
		  << "H\t" << X << '\t' << Y << "\t0\t0\n"
		  << "H\t0\t" << Y << "\t0\t0\n";
		double hh = 0;
		for(int y = 0; y < Y; ++y)
		{
			hh += bit(0, y, 1) ? 1 : -1;
			for(int x = 1; x < X; ++x)
			{
				hh += bit(x, y, 0) ? 1 : -1;
				o << "Fe\t" << x << '\t' << y << '\t' << hh << '\t' << hh << '\n';
			}
			hh += bit(0, y, 0) ? 1 : -1;
			o << "Fe\t0\t" << y << '\t' << hh << '\t' << hh << '\n';
		}
	}
};

// SamplingParam::nextEval (kpz/sampling.h:71-88)
unsigned nextEval(const Sampling& s, unsigned mcs, Engine& e)
{
	if(s.kind == Sampling::LINEAR)
		return mcs + s.interval;
	if(s.kind == Sampling::EXP)
		return (unsigned)std::ceil(std::exp(std::log((double)mcs + 10) + s.density));
	return mcs + (unsigned)(kpz_master_draw(e.h) * s.mean + s.mean);
}

// FrontendArgs::handleNextEvent (kpz/kpzFrontend.cpp:805-830); returns
// whether this point counts as a measurement for the sampling schedule
bool handleNextEvent(Args& A, Engine& e)
{
	if(A.events.empty())
		return true;
	bool count = A.events.front().start > e.mcs;
	while(A.events.front().start <= e.mcs)
	{
		const Event ev = A.events.front();
		A.events.pop_front();
		if(ev.start == e.mcs)
		{
			switch(ev.kind)
			{
			case Event::END:
				A.finished = true;
				break;
			case Event::SAMPLING:
				A.current = ev.sampling;
				break;
			case Event::TAG:
				e.correlateTag(ev.stop);
				break;
			case Event::BARRIER:
				break;
			}
			count |= ev.countAsMeasurement;
		}
		else
			std::cerr << "[WW][FrontendArgs][handleNextEvent] Event skipped: " << ev.start << '<' << e.mcs << '\n';
		if(A.events.empty())
			break;
	}
	return count;
}

} // namespace

int main(int argc, char** argv)
{
	if(argc < 2)
	{
		help();
		return 1;
	}
	Args A;
	int a = parsePositional(A, argc, argv);
	if(a < 0)
		return 1;
	if(!parseOptions(A, argc, argv, a))
		return 1;
	// FrontendArgs::init (kpz/kpzFrontend.cpp:735-745)
	std::cout << "[MM][FrontendArgs] running " << A.anaThreads << " threads for analysis (per process)\n";
	std::cout << std::setprecision(12);
	std::cout << "[MM][FrontendArgs] output precision= " << 12 << '\n';
	A.events.push_back(Event{Event::END, A.mcs, false, Sampling{}, 0});
	A.events.sort([](const Event& x, const Event& y) { return x.start < y.start; });

	// SchedulerConfig::mkScheduler (kpz/kpzFrontend.cpp:234-395)
	int kind = KPZ_DTR;
	if(A.scheduler == "dtStatic")
		kind = KPZ_DT;
	else if(A.scheduler == "scaBit")
		kind = KPZ_SCABIT;
	else if(A.scheduler == "sca")
		kind = KPZ_SCA;
	kpz_params prm{A.mcsThreads, A.dd[0], A.dd[1], 0, A.device, A.rng};
	if(A.rng == KPZ_RNG_TINYMT32 && A.loadState.empty() && A.loadH5.empty() && A.tinymtParams.empty())
	{
		std::cerr << "[EE][FrontendArgs] --rng tinymt32 needs --tinymtParams (one parameter set per worker).\n";
		return 1;
	}
	Engine e;
	uint32_t lastMeasurement = 0;
	if(!A.loadState.empty() && !A.loadH5.empty())
	{
		std::cerr << "[EE][FrontendArgs] --loadState and --loadH5 exclude each other.\n";
		return 1;
	}
	const bool resumed = !A.loadState.empty() || !A.loadH5.empty();
	if(!resumed)
	{
		e.lx = A.lx;
		e.ly = A.ly;
		Engine::check(kpz_create(kind, A.lx, A.ly, &prm, &e.h), "create scheduler");
	}
	else if(!A.loadState.empty())
	{
		// --loadH5 path (kpz/kpzScaling.cpp:118-139): size, kind, RNG and tags
		// come from the file
		Engine::check(kpz_load_state(A.loadState.c_str(), A.device, &e.h, &lastMeasurement), "load state");
		Engine::check(kpz_size(e.h, &e.lx, &e.ly), "size");
		e.mcs = kpz_mcs_count(e.h);
		std::cout << "[readState] Loading state at t = " << e.mcs << " MCS\n";
	}
#ifdef KPZ_WITH_H5
	else
	{
		// SchedulerService::readH5 (kpz/kpzScaling.cpp:121-136): the size comes
		// from the file, the kind and worker layout from the command line
		uint32_t restored = 0;
		Engine::check(kpz_read_h5(A.loadH5.c_str(), "", kind, &prm, &e.h, &lastMeasurement, &restored), "read HDF5");
		Engine::check(kpz_size(e.h, &e.lx, &e.ly), "size");
		e.mcs = kpz_mcs_count(e.h);
		std::cout << "[readH5] Loading Hdf5 file at t = " << e.mcs << " MCS\n";
		if(restored & KPZ_H5_MASTER)
			std::cout << "[readH5] restored dSFMT state from file.\n";
		if(!(restored & KPZ_H5_STREAMS))
			std::cout << "[retrieveH5][WW] Failed to restore GPU RNG state from file, randomizing. Check if dSFMT "
			             "was restored.\n";
	}
#endif
	std::cout << "mcsScale " << kpz_mcs_scale(e.h) << '\n';

	std::ofstream fileOut;
	if(A.file.length() && A.file != "--")
	{
		fileOut.open(A.file);
		if(!fileOut)
		{
			std::cerr << "Could not open file: '" << A.file << "' for writing.\n";
			return 1;
		}
		e.out = &fileOut;
	}

	// seeding (SchedulerService::initURandom / supplySeed,
	// kpz/schedulerService.cpp:460-495), then initHomogeneous + randomize
	if(!resumed)
	{
		if(!A.haveSeed)
		{
			std::ifstream r("/dev/urandom", std::ios_base::binary);
			uint64_t u = 0;
			r.read((char*)&u, sizeof u);
			A.seed = u ^ (uint64_t)std::time(nullptr);
			std::cout << "[SchedulerService][initURandom] seed= " << A.seed << '\n';
		}
		else
			std::cout << "[SchedulerService][supplySeed] seed= " << A.seed << '\n';
		if(A.rng == KPZ_RNG_TINYMT32)
		{
			// RngIface::TinyMT32::initialize: parameter set t for worker t
			std::vector<uint32_t> m[3];
			const size_t T = (size_t)kpz_threads(e.h);
			if(!readTinyMTParams(A.tinymtParams, T, m))
				return 1;
			Engine::check(kpz_set_tinymt_params(e.h, m[0].data(), m[1].data(), m[2].data()), "TinyMT32 parameters");
			std::cout << "TinyMT32: Using parametersets: 0 .. " << T - 1 << '\n';
		}
		Engine::check(kpz_set_seed(e.h, A.seed), "seed");
		Engine::check(kpz_init_flat(e.h), "initHomogeneous");
		if(!std::isnan(A.disorderFill))
			Engine::check(kpz_generate_disorder2(e.h, A.disorderFill), "generateDisorder2");
	}
	else
	{
		// FrontendArgs::fastForwardEvents (kpz/kpzFrontend.cpp:790-800)
		while(!A.events.empty() && A.events.front().start < e.mcs)
		{
			if(A.events.front().kind == Event::SAMPLING)
				A.current = A.events.front().sampling;
			A.events.pop_front();
		}
	}
	// rates and MCS function (kpz/kpzScaling.cpp:146-175, FrontendArgs::setPQ
	// kpz/kpzFrontend.cpp:852-859): rate sets up to the first unset one; the
	// disorder dynamics need --pq
	if(!std::isnan(A.p[0]))
	{
		if(A.p[0] == 1. && A.q[0] == 0.)
			std::cerr << "[WW] Running dynamic pq implementation with p=1 and "
			             "q=0, using the fixed implementation would be wore efficient in this case\n";
		std::cout << "Rates: p = " << A.p[0] << " , q = " << A.q[0] << '\n';
		for(unsigned n = 0; n < 4 && !std::isnan(A.p[n]); ++n)
			Engine::check(kpz_set_pq_class(e.h, A.p[n], A.q[n], n), "setPQ");
		if(std::isnan(A.disorderFill))
			e.mode = Engine::PQ;
		else
		{
			std::cout << "Disorder-rates: p1 = " << A.p[1] << " , q1 = " << A.q[1] << '\n';
			e.mode = Engine::DISORDER;
		}
	}

	const auto t0 = std::chrono::steady_clock::now();
	auto timedOut = [&] {
		return A.maxH > 0 &&
		       std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > A.maxH * 3600.;
	};

	// main loop (kpz/kpzScaling.cpp:176-215)
	e.correlate();
	unsigned prev = e.mcs;
	// a resumed run continues the sampling schedule from the last measurement
	// stored with the state (overridePrevMeasurementTime, kpz/kpzFrontend.cpp:752-788)
	// (random sampling draws from the master and is not replayed)
	if(resumed && A.current.kind != Sampling::RANDOM && lastMeasurement <= e.mcs &&
	   nextEval(A.current, lastMeasurement, e) >= e.mcs)
		prev = lastMeasurement;
	bool timeout = false;
	while(!A.finished)
	{
		unsigned next = nextEval(A.current, prev, e);
		if(!A.events.empty())
			next = std::min(A.events.front().start, next);
		e.run(next - e.mcs);
		if(timedOut())
		{
			timeout = true;
			break;
		}
		e.correlate();
		if(handleNextEvent(A, e))
			prev = e.mcs;
	}
	Engine::check(kpz_sync(e.h), "sync");
	if(timeout)
		std::cout << "TIMEOUT " << e.mcs << '\n';
	if(!A.saveState.empty())
		Engine::check(kpz_save_state(e.h, A.saveState.c_str(), prev), "save state");
#ifdef KPZ_WITH_H5
	// writeH5 + the prevMeasureOverride attribute (kpz/kpzScaling.cpp:225-229)
	if(!A.saveH5.empty())
		Engine::check(kpz_write_h5(e.h, A.saveH5.c_str(), "", prev), "write HDF5");
#endif
	if(A.save)
	{
		std::ostringstream ss;
		ss << "KPZheightmap_" << e.mcs << ".xyz";
		e.saveXYZ(ss.str());
	}
	return 0;
}
