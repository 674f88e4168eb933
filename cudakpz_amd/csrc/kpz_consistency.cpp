// kpz_consistency_check: the reference's kpzConsistencyCheck tool
// (kpz/kpzConsistencyCheck.cpp:38-86) on the engine.  Every file named
// <dir>/<name>_<r>_0_0.h5 is read as a plain 4x4-cell lattice (the reference
// reads it into a bare SchedulerService(4,4); prefix "dis" when r is 1, the
// disordered part of an autoResponse run), its lattice and correlator tags are
// checked on the device (kpz_consistency_check, full), and the tool prints
// "[MM] Good: <file>" or "[MM] Bad: <file>".  --anaThreads n is accepted as
// the reference's option (the check runs on the GPU, not on analysis threads).
#include "kpz_amd.h"
#include "kpz_amd_h5.h"

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <regex>
#include <string>

int main(int argc, char* argv[])
{
	// subexpressions: directory, name, rank (kpz/kpzConsistencyCheck.cpp:43)
	const std::regex rxName("(.*/)(.*)_(.)_0_0\\.h5");
	for(int a = 1; a < argc; ++a)
	{
		if(!std::strcmp(argv[a], "--anaThreads"))
		{
			char* end = nullptr;
			const long n = (a + 1 < argc) ? std::strtol(argv[a + 1], &end, 10) : 0;
			if(a + 1 >= argc || !end || *end || end == argv[a + 1])
			{
				std::cerr << "Expected number of threads for --anaThreads.\n";
				return 1;
			}
			++a;
			std::cout << "[MM] running " << n << " threads for analysis (per process)\n";
			continue;
		}
		const std::string fname(argv[a]);
		std::smatch match;
		if(!std::regex_match(fname, match, rxName))
		{
			std::cerr << "[WW] Skipping file with non-matching name: " << fname << '\n';
			continue;
		}
		const char* prefix = (match[3] == "1") ? "dis" : "";
		kpz_scheduler* h = nullptr;
		if(kpz_read_h5(fname.c_str(), prefix, KPZ_SCA, nullptr, &h, nullptr, nullptr) != KPZ_OK)
		{
			std::cerr << "[EE] Failed to read file: " << fname << " (" << kpz_last_error() << ")\n";
			continue;
		}
		int64_t counts[3];
		const int ok = kpz_consistency_check(h, 1, counts);
		if(ok < 0)
		{
			std::cerr << "[EE] " << kpz_last_error() << '\n';
			kpz_destroy(h);
			return 1;
		}
		if(!ok)
			std::cout << "[MM][consitencyCheck] Found inconsistency(s) in m_system: " << counts[0] << " lines, "
			          << counts[1] << " columns, " << counts[2] << " loops\n";
		std::cout << (ok ? "[MM] Good: " : "[MM] Bad: ") << fname << '\n';
		kpz_destroy(h);
	}
	return 0;
}
