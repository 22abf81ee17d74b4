/*
 * sp_h5.c -- HDF5 files of the distribution diagnostics for the Fortran host.
 *
 * The reference writes its distributions with the HDF5 Fortran API
 * (diagnostics.f90:1244-1650: save_mu_pbins, save_global_distributions,
 * save_escaped_distributions, save_5d_local_dist_data,
 * save_5d_local_escaped_dist_data): one file per frame, FP64 datasets whose
 * dimensions are declared in Fortran order.  The Fortran API stores a
 * dataset declared (n1, n2, ..., nk) as a C dataspace (nk, ..., n2, n1) over
 * the same bytes, so a C reader (h5py, the reference's notebooks) sees the
 * reversed shape.  These helpers take the Fortran-order dimensions and write
 * exactly that dataspace, so the files match the reference's dataset for
 * dataset.  Only the serial HDF5 C library is needed; the Fortran HDF5
 * modules of the image are in another compiler's format.
 */
#include <hdf5.h>
#include <stdint.h>

/* Creates (create != 0, truncating) or opens read-write an HDF5 file.
 * Returns the file id (> 0) or a negative value. */
int64_t sp_h5_open(const char* fname, int32_t create)
{
    if (!fname) return -1;
    if (create) return (int64_t)H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    return (int64_t)H5Fopen(fname, H5F_ACC_RDWR, H5P_DEFAULT);
}

/* Writes an FP64 dataset `name` of rank 1..8 with Fortran-order dimensions
 * fdims (first index fastest in `data`).  Returns 0 or -1. */
int32_t sp_h5_write(int64_t fid, const char* name, int32_t rank, const int64_t* fdims,
                    const double* data)
{
    hsize_t dims[8];
    if (fid < 0 || !name || !fdims || !data || rank < 1 || rank > 8) return -1;
    for (int i = 0; i < rank; ++i) {
        if (fdims[i] < 0) return -1;
        dims[i] = (hsize_t)fdims[rank - 1 - i];
    }
    const hid_t space = H5Screate_simple(rank, dims, NULL);
    if (space < 0) return -1;
    const hid_t dset = H5Dcreate2((hid_t)fid, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT,
                                  H5P_DEFAULT, H5P_DEFAULT);
    herr_t st = -1;
    if (dset >= 0) {
        st = H5Dwrite(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
        if (H5Dclose(dset) < 0) st = -1;
    }
    if (H5Sclose(space) < 0) st = -1;
    return st < 0 ? -1 : 0;
}

int32_t sp_h5_close(int64_t fid)
{
    if (fid < 0) return -1;
    return H5Fclose((hid_t)fid) < 0 ? -1 : 0;
}
