// Pieces of the C ABI shared by the libraries that implement it
// (libkpz_amd.so and the HDF5 checkpoint library libkpz_amd_h5.so).
#pragma once

#include "../../include/kpz_amd.h"
#include "kpz_scheduler.hpp"

#include <memory>
#include <new>
#include <string>

struct kpz_scheduler {
	std::unique_ptr<kpz::SchedulerService> s;
};

namespace kpz {

// records msg as the text kpz_last_error() returns (per thread) and passes
// code through
int capiFail(int code, const std::string& msg);

// runs f, turning every exception into a status code (no exception crosses
// the C ABI)
template<class F>
int capiGuard(F&& f)
{
	try
	{
		f();
		return KPZ_OK;
	}
	catch(const Error& e)
	{
		return capiFail(e.code, e.what());
	}
	catch(const std::bad_alloc&)
	{
		return capiFail(KPZ_ENOMEM, "host allocation failed");
	}
	catch(const std::exception& e)
	{
		return capiFail(KPZ_EDEVICE, e.what());
	}
}

} // namespace kpz
