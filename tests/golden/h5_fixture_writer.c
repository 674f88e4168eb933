/*
 * Writes HDF5 checkpoint fixtures in the layout the reference's
 * SchedulerService::collectH5 produces through libSplash's serial collector
 * (kpz/schedulerService.cpp:583-612, kpz/correlator.cpp:709-716,
 * inc/KMCsplash.cpp:30-34, kpz/kpzScaling.cpp:225-229), written directly with
 * the HDF5 C API and independently of the engine's own writer.  Test data
 * only; make_h5_fixture.py builds and runs it.
 *
 *   h5_fixture_writer out.h5 variant wx wy mcs lattice.bin tag1.bin tag2.bin dsfmt.bin avgH1 avgH2
 *
 * variant: ref     -- lattice at /data/<mcs>, tags at /data/<mcs/2> (".corrTag",
 *                     stop 1000) and /data/<mcs/4> (legacy "corrTag", stop 30),
 *                     .dsfmt of 770 words, /custom .disorderP/.disorderQ and
 *                     prevMeasureOverride; no SLCG states (the reference has none)
 *          rank3   -- .system stored as a 3-d dataset
 *          npot    -- .system with a width that is not a power of two
 *          tagdims -- a tag snapshot whose extents differ from the lattice
 *          nodata  -- no /data group
 */
#include <hdf5.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void* slurp(const char* path, size_t n)
{
	void* p = calloc(1, n ? n : 1);
	FILE* f = fopen(path, "rb");
	if(!f || fread(p, 1, n, f) != n)
	{
		fprintf(stderr, "cannot read %zu bytes from %s\n", n, path);
		exit(1);
	}
	fclose(f);
	return p;
}

static void dataset(hid_t g, const char* name, hid_t type, int rank, const hsize_t* dims, const void* data)
{
	hid_t sp = H5Screate_simple(rank, dims, NULL);
	hid_t ds = H5Dcreate2(g, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	if(ds < 0 || H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
	{
		fprintf(stderr, "dataset %s failed\n", name);
		exit(1);
	}
	H5Dclose(ds);
	H5Sclose(sp);
}

static void attribute(hid_t obj, const char* name, hid_t type, const void* data)
{
	hid_t sp = H5Screate(H5S_SCALAR);
	hid_t a = H5Acreate2(obj, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
	if(a < 0 || H5Awrite(a, type, data) < 0)
	{
		fprintf(stderr, "attribute %s failed\n", name);
		exit(1);
	}
	H5Aclose(a);
	H5Sclose(sp);
}

static hid_t group(hid_t parent, unsigned id)
{
	char name[32];
	snprintf(name, sizeof name, "%u", id);
	return H5Gcreate2(parent, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

int main(int argc, char** argv)
{
	if(argc != 12)
	{
		fprintf(stderr, "usage: %s out.h5 variant wx wy mcs lattice.bin tag1.bin tag2.bin dsfmt.bin avgH1 avgH2\n",
				argv[0]);
		return 2;
	}
	const char* variant = argv[2];
	const unsigned wx = (unsigned)atoi(argv[3]), wy = (unsigned)atoi(argv[4]), mcs = (unsigned)atoi(argv[5]);
	const size_t words = (size_t)wx * wy;
	uint32_t* lat = slurp(argv[6], words * 4);
	uint32_t* tag1 = slurp(argv[7], words * 4);
	uint32_t* tag2 = slurp(argv[8], words * 4);
	uint32_t* dsfmt = slurp(argv[9], 770 * 4);

	hid_t f = H5Fcreate(argv[1], H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
	if(f < 0)
		return 1;
	hid_t custom = H5Gcreate2(f, "custom", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	const hsize_t four = 4;
	hid_t d4 = H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, &four);
	const double P[4] = {0.9, 0.3, 0.6, 0.1}, Q[4] = {0.05, 0.7, 0.2, 0.5};
	attribute(custom, ".disorderP", d4, P);
	attribute(custom, ".disorderQ", d4, Q);
	const uint32_t pmo = 37;
	attribute(custom, "prevMeasureOverride", H5T_NATIVE_UINT32, &pmo);
	H5Tclose(d4);
	H5Gclose(custom);
	if(!strcmp(variant, "nodata"))
	{
		H5Fclose(f);
		return 0;
	}

	hid_t data = H5Gcreate2(f, "data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
	/* the lattice and the master at the largest id */
	hid_t g = group(data, mcs);
	if(!strcmp(variant, "rank3"))
	{
		const hsize_t d3[3] = {1, wy, wx};
		dataset(g, ".system", H5T_NATIVE_UINT32, 3, d3, lat);
	}
	else if(!strcmp(variant, "npot"))
	{
		const hsize_t d2[2] = {wy, wx - 1};
		dataset(g, ".system", H5T_NATIVE_UINT32, 2, d2, lat);
	}
	else
	{
		const hsize_t d2[2] = {wy, wx};
		dataset(g, ".system", H5T_NATIVE_UINT32, 2, d2, lat);
	}
	const hsize_t nd = 770;
	dataset(g, ".dsfmt", H5T_NATIVE_UINT32, 1, &nd, dsfmt);
	H5Gclose(g);

	/* correlator tags: snapshot .system with avgH on it, stop time on the group */
	const unsigned ids[2] = {mcs / 2, mcs / 4};
	const uint32_t stops[2] = {1000, 30};
	const char* names[2] = {".corrTag", "corrTag"};
	uint32_t* snaps[2] = {tag1, tag2};
	for(int k = 0; k < 2; ++k)
	{
		hid_t tg = group(data, ids[k]);
		hsize_t d2[2] = {wy, wx};
		if(k == 0 && !strcmp(variant, "tagdims"))
			d2[0] = wy / 2;
		dataset(tg, ".system", H5T_NATIVE_UINT32, 2, d2, snaps[k]);
		attribute(tg, names[k], H5T_NATIVE_UINT32, &stops[k]);
		hid_t ds = H5Dopen2(tg, ".system", H5P_DEFAULT);
		const double avgH = atof(argv[10 + k]);
		attribute(ds, "avgH", H5T_NATIVE_DOUBLE, &avgH);
		H5Dclose(ds);
		H5Gclose(tg);
	}
	/* an id group without a tag: ignored by the reader */
	hid_t other = group(data, mcs / 8);
	H5Gclose(other);
	H5Gclose(data);
	H5Fclose(f);
	free(lat);
	free(tag1);
	free(tag2);
	free(dsfmt);
	return 0;
}
