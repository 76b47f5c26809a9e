/*
 * h5writer.h -- the HDF5 layer of ParaxialSimulation::save, kept in its own
 * library (libpaxpro_h5.so, built only where the HDF5 C library is found) so
 * that the device path does not depend on HDF5.  libpaxpro.so loads it on the
 * first save() and calls pxh5_write.
 *
 * Layout written (src/paraxialSimulation.cpp:137-193,286-317,454-511 and
 * include/hdf5DataspaceCreator.hpp:30-66 of the reference):
 *   group `group` with scalar attributes: doubles, and fixed-length strings
 *   (H5T_C_S1 of the string's length) where text is given;
 *   one dataset `group/name` per entry: H5T_NATIVE_DOUBLE or H5T_NATIVE_UINT8,
 *   rank and dims as given (the caller passes the reference's dataspace: a
 *   vector {n}, a matrix {n_cols, n_rows}, a cube {n_rows, n_cols, n_slices},
 *   each over the column-major buffer as stored), with scalar int or double
 *   attributes.  The file is created anew (H5F_ACC_TRUNC) and closed.
 */
#ifndef PAXPRO_H5WRITER_H
#define PAXPRO_H5WRITER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const char *name;
  int is_int;      /* H5T_NATIVE_INT (value converted) or H5T_NATIVE_DOUBLE */
  double value;
} pxh5_attr;

typedef struct {
  const char *name;
  const char *text;  /* non-null: a string attribute */
  double value;
} pxh5_group_attr;

typedef struct {
  const char *name;
  int rank;                      /* 1..3 */
  unsigned long long dims[3];
  int uint8;                     /* element type: 1 = unsigned char, 0 = double */
  const void *data;
  const pxh5_attr *attrs;
  int n_attrs;
} pxh5_dataset;

/* 0 on success; otherwise a message in err (errlen bytes). */
int pxh5_write(const char *fname, const char *group, const pxh5_group_attr *gattrs, int n_gattrs,
               const pxh5_dataset *dsets, int n_dsets, char *err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
