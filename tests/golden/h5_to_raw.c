/*
 * Fixture converter: dumps every dataset in the root group of an HDF5 file
 * as raw little-endian binary plus a manifest line, so that the fixtures the
 * reference's own tests read (test/data/*.hdf5) can be committed as .npz.
 *
 * Build (this container only; links the image's libhdf5):
 *   gcc -O2 h5_to_raw.c -I/opt/conda/include -L/opt/conda/lib -lhdf5 \
 *       -Wl,-rpath,/opt/conda/lib -o h5_to_raw
 * Usage: h5_to_raw <file.hdf5> <outdir>
 *   writes <outdir>/<name>.bin and prints "<name> <f8|i8> <d0> <d1> ..." per dataset.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hdf5.h>

static const char* g_outdir;

static herr_t visit(hid_t group, const char* name, const H5L_info_t* info, void* op_data)
{
	(void)info; (void)op_data;
	hid_t ds = H5Dopen2(group, name, H5P_DEFAULT);
	if (ds < 0) { return 0; }  /* not a dataset */
	hid_t space = H5Dget_space(ds);
	int ndims = H5Sget_simple_extent_ndims(space);
	hsize_t dims[16] = { 0 };
	H5Sget_simple_extent_dims(space, dims, NULL);
	hssize_t npoints = H5Sget_simple_extent_npoints(space);
	hid_t dtype = H5Dget_type(ds);
	H5T_class_t cls = H5Tget_class(dtype);

	const int is_int = (cls == H5T_INTEGER);
	void* buf = malloc((size_t)(npoints > 0 ? npoints : 1) * 8);
	herr_t st = H5Dread(ds, is_int ? H5T_NATIVE_INT64 : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
	if (st < 0) {
		fprintf(stderr, "reading dataset %s failed\n", name);
		exit(1);
	}

	char path[4096];
	snprintf(path, sizeof(path), "%s/%s.bin", g_outdir, name);
	FILE* f = fopen(path, "wb");
	fwrite(buf, 8, (size_t)npoints, f);
	fclose(f);

	printf("%s %s", name, is_int ? "i8" : "f8");
	for (int i = 0; i < ndims; i++) { printf(" %llu", (unsigned long long)dims[i]); }
	printf("\n");

	free(buf);
	H5Tclose(dtype);
	H5Sclose(space);
	H5Dclose(ds);
	return 0;
}

int main(int argc, char** argv)
{
	if (argc != 3) {
		fprintf(stderr, "usage: %s <file.hdf5> <outdir>\n", argv[0]);
		return 2;
	}
	g_outdir = argv[2];
	hid_t file = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
	if (file < 0) {
		fprintf(stderr, "cannot open %s\n", argv[1]);
		return 1;
	}
	H5Literate(file, H5_INDEX_NAME, H5_ITER_INC, NULL, visit, NULL);
	H5Fclose(file);
	return 0;
}
