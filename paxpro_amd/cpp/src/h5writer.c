/* HDF5 writer of ParaxialSimulation::save (see paxpro/h5writer.h), on the
 * HDF5 C library. */
#include <hdf5.h>
#include <stdio.h>
#include <string.h>

#include "paxpro/h5writer.h"

static int fail(char *err, size_t errlen, const char *what, const char *name)
{
  if (err && errlen) snprintf(err, errlen, "HDF5: %s%s%s", what, name ? " " : "", name ? name : "");
  return 1;
}

/* scalar attribute of a group or dataset */
static int put_scalar(hid_t obj, const char *name, hid_t type, const void *value)
{
  hid_t space = H5Screate(H5S_SCALAR);
  if (space < 0) return -1;
  hid_t att = H5Acreate2(obj, name, type, space, H5P_DEFAULT, H5P_DEFAULT);
  int rc = att < 0 ? -1 : (H5Awrite(att, type, value) < 0 ? -1 : 0);
  if (att >= 0) H5Aclose(att);
  H5Sclose(space);
  return rc;
}

static int put_string(hid_t obj, const char *name, const char *text)
{
  const size_t len = strlen(text);
  if (len == 0) return -1;
  hid_t type = H5Tcopy(H5T_C_S1);
  if (type < 0) return -1;
  int rc = H5Tset_size(type, len) < 0 ? -1 : put_scalar(obj, name, type, text);
  H5Tclose(type);
  return rc;
}

int pxh5_write(const char *fname, const char *group, const pxh5_group_attr *gattrs, int n_gattrs,
               const pxh5_dataset *dsets, int n_dsets, char *err, size_t errlen)
{
  if (!fname || !group || (n_gattrs && !gattrs) || (n_dsets && !dsets))
    return fail(err, errlen, "null argument", NULL);
  H5Eset_auto2(H5E_DEFAULT, NULL, NULL);  /* errors come back as status codes */
  hid_t file = H5Fcreate(fname, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return fail(err, errlen, "cannot create file", fname);
  int rc = 0;
  hid_t grp = H5Gcreate2(file, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (grp < 0) rc = fail(err, errlen, "cannot create group", group);
  for (int i = 0; !rc && i < n_gattrs; ++i) {
    const pxh5_group_attr *a = &gattrs[i];
    const int r = a->text ? put_string(grp, a->name, a->text)
                          : put_scalar(grp, a->name, H5T_NATIVE_DOUBLE, &a->value);
    if (r < 0) rc = fail(err, errlen, "cannot write group attribute", a->name);
  }
  for (int i = 0; !rc && i < n_dsets; ++i) {
    const pxh5_dataset *d = &dsets[i];
    unsigned long long count = 1;
    for (int k = 0; k < d->rank && k < 3; ++k) count *= d->dims[k];
    if (d->rank < 1 || d->rank > 3 || !d->name || (count && !d->data)) {
      rc = fail(err, errlen, "bad dataset", d->name);
      break;
    }
    hsize_t dims[3] = {d->dims[0], d->dims[1], d->dims[2]};
    char path[1024];
    snprintf(path, sizeof path, "%s/%s", group, d->name);
    const hid_t type = d->uint8 ? H5T_NATIVE_UINT8 : H5T_NATIVE_DOUBLE;
    hid_t space = H5Screate_simple(d->rank, dims, NULL);
    hid_t ds = space < 0 ? -1 : H5Dcreate2(file, path, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) {
      rc = fail(err, errlen, "cannot create dataset", path);
    } else {
      for (int a = 0; !rc && a < d->n_attrs; ++a) {
        const pxh5_attr *at = &d->attrs[a];
        int r;
        if (at->is_int) {
          const int v = (int)at->value;
          r = put_scalar(ds, at->name, H5T_NATIVE_INT, &v);
        } else {
          r = put_scalar(ds, at->name, H5T_NATIVE_DOUBLE, &at->value);
        }
        if (r < 0) rc = fail(err, errlen, "cannot write attribute", at->name);
      }
      /* an empty result (a module without that overload) leaves an empty dataset */
      if (!rc && count && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, d->data) < 0)
        rc = fail(err, errlen, "cannot write dataset", path);
      H5Dclose(ds);
    }
    if (space >= 0) H5Sclose(space);
  }
  if (grp >= 0) H5Gclose(grp);
  if (H5Fclose(file) < 0 && !rc) rc = fail(err, errlen, "cannot close file", fname);
  return rc;
}
