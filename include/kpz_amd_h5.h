/*
 * kpz_amd_h5.h -- HDF5 checkpoints with the reference's dataset names.
 *
 * Implemented by libkpz_amd_h5.so (links libkpz_amd.so and libhdf5), so that
 * the engine library itself has no HDF5 dependency.  Same conventions as
 * kpz_amd.h: KPZ_OK or a negative KPZ_E* code, kpz_last_error() explains.
 *
 * File layout (SchedulerService::collectH5, kpz/schedulerService.cpp:583-612,
 * through libSplash's SerialDataCollector; ids are MCS counts):
 *   /data/<mcs>/<prefix>.system      u32 [dimYW][dimXW]  the lattice words
 *   /data/<mcs>/<prefix>.disorder    u32 [dimYW][dimXW]  Disorder2 map, if set
 *   /data/<mcs>/<prefix>.dsfmt       u32 [770]           master dsfmt_t bytes
 *                                                        (KmcSplash::writeDSFMT,
 *                                                        inc/KMCsplash.cpp:30-34)
 *   /data/<t>/<prefix>.system        snapshot of the correlator tag taken at t,
 *       attribute avgH (f64) on it, attribute <prefix>.corrTag (u32 stop) on
 *       the group /data/<t> (Correlator::save, kpz/correlator.cpp:709-716)
 *   /custom  attributes <prefix>.disorderP, <prefix>.disorderQ (f64[4]) and
 *            prevMeasureOverride (u32, kpz/kpzScaling.cpp:225-229)
 * and, which the reference never saved (its SLCG64 writeH5 is empty,
 * inc/kpzCudaRng/GPUSLCG64.h:55-56):
 *   /data/<mcs>/<prefix>.slcg64      u64 [T] every worker stream's state,
 *       attributes skip (u64) and threads (i64)
 * The file name is <file>_0_0_0.h5 (the serial collector's rank suffix,
 * KmcSplash::getFullSerialFilename, inc/KMCsplash.cpp:56-62) unless <file>
 * already ends in ".h5".  Datasets are contiguous and uncompressed.
 */
#ifndef KPZ_AMD_H5_H
#define KPZ_AMD_H5_H

#include "kpz_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* what kpz_read_h5 restored from the file (bit mask) */
enum {
	KPZ_H5_MASTER = 1,   /* the master dSFMT state */
	KPZ_H5_STREAMS = 2,  /* the SLCG64 streams (else: re-randomized from the master) */
	KPZ_H5_DISORDER = 4, /* a Disorder2 map */
	KPZ_H5_RATES = 8,    /* the class rates .disorderP / .disorderQ */
	KPZ_H5_PMO = 16      /* the prevMeasureOverride attribute */
};

/* SchedulerService::writeH5 (kpz/schedulerService.cpp:527-553) followed by
 * the front end's prevMeasureOverride attribute (kpz/kpzScaling.cpp:225-229).
 * prefix may be NULL (= ""). */
int kpz_write_h5(kpz_scheduler* h, const char* file, const char* prefix, uint32_t prevMeasureOverride);

/* SchedulerService::readH5 / retrieveH5 (kpz/schedulerService.cpp:555-676)
 * on a new scheduler of the given kind: sizes from <prefix>.system at the
 * largest id, which becomes the MCS count; the lattice, the Disorder2 map,
 * every tag (a group with a <prefix>.corrTag attribute, also the legacy
 * "<prefix>corrTag"; unlike the reference, a tag at the largest id is kept
 * too), the master, the class rates (the reference leaves them to the
 * caller), and the SLCG streams when the file holds them for the same T and
 * skip -- otherwise the streams are re-seeded from the restored master, as
 * the reference's retrieveH5 does.  restored (may be NULL) gets KPZ_H5_*
 * bits; prevMeasureOverride (may be NULL) gets the attribute or 0. */
int kpz_read_h5(const char* file, const char* prefix, int kind, const kpz_params* params,
		kpz_scheduler** out, uint32_t* prevMeasureOverride, uint32_t* restored);

#ifdef __cplusplus
}
#endif
#endif
