// ml_chombo_h5.cpp — Chombo HDF5 checkpoint/plot files (the reference's
// on-disk format, srcSubcycle/AMRLevelMushyLayerIO.cpp:131-656) over the HDF5
// C API.  Host code only; the layout is described in include/mlio.h.
//
// The reference writes these files through Chombo's HDF5Handle,
// HDF5HeaderData and write(HDF5Handle&, LevelData&) [ext]; the layout below
// is the one its files hold (test/unit/data/chk000100.2d.hdf5,
// test/unit/data/plt000100.2d.hdf5): header values as scalar attributes
// (ints as 32-bit integers, reals as doubles, strings as fixed-length
// strings of their own length, boxes as compounds lo_i..hi_k), a level's
// boxes as a compound dataset "boxes" with "Processors", and each variable
// as "<name>:datatype=0" (doubles) + "<name>:offsets=0" (int64) + a
// "<name>_attributes" group.
#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mlio.h"

struct ml_h5_s {
  hid_t file = -1;
  int dim = 2;
  bool writable = false;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

const char* axis(int d) { return d == 0 ? "i" : d == 1 ? "j" : "k"; }

// scoped HDF5 identifier
struct Id {
  hid_t v = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id() = default;
  Id(hid_t h, herr_t (*c)(hid_t)) : v(h), close(c) {}
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  Id(Id&& o) noexcept : v(o.v), close(o.close) { o.v = -1; }
  ~Id() {
    if (v >= 0 && close) close(v);
  }
  operator hid_t() const { return v; }
  bool ok() const { return v >= 0; }
};

std::string group_path(const char* group) {
  if (!group || !*group || std::strcmp(group, "/") == 0) return "/";
  return std::string(group);
}

std::string level_group(int level) { return "level_" + std::to_string(level); }

// open a group, creating it (and only it) when writing
Id open_group(ml_h5 f, const std::string& path, bool create) {
  if (path == "/") return Id(H5Gopen2(f->file, "/", H5P_DEFAULT), H5Gclose);
  if (H5Lexists(f->file, path.c_str(), H5P_DEFAULT) > 0) return Id(H5Gopen2(f->file, path.c_str(), H5P_DEFAULT), H5Gclose);
  if (!create) return Id();
  return Id(H5Gcreate2(f->file, path.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
}

Id box_type(int dim) {
  Id t(H5Tcreate(H5T_COMPOUND, sizeof(int) * 2 * dim), H5Tclose);
  for (int d = 0; d < dim; ++d) H5Tinsert(t, ("lo_" + std::string(axis(d))).c_str(), sizeof(int) * d, H5T_NATIVE_INT);
  for (int d = 0; d < dim; ++d)
    H5Tinsert(t, ("hi_" + std::string(axis(d))).c_str(), sizeof(int) * (dim + d), H5T_NATIVE_INT);
  return t;
}

Id intvect_type(int dim) {
  Id t(H5Tcreate(H5T_COMPOUND, sizeof(int) * dim), H5Tclose);
  for (int d = 0; d < dim; ++d) H5Tinsert(t, ("intvect" + std::string(axis(d))).c_str(), sizeof(int) * d, H5T_NATIVE_INT);
  return t;
}

int write_attr(ml_h5 f, const char* group, const char* name, hid_t type, const void* data) {
  if (!f || !f->writable || !name || !*name) return fail(ML_H5_EINVAL, "attribute: bad file or name");
  Id g = open_group(f, group_path(group), true);
  if (!g.ok()) return fail(ML_H5_EIO, "cannot open group " + group_path(group));
  if (H5Aexists(g, name) > 0) H5Adelete(g, name);
  Id sp(H5Screate(H5S_SCALAR), H5Sclose);
  Id a(H5Acreate2(g, name, type, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Awrite(a, type, data) < 0) return fail(ML_H5_EIO, std::string("cannot write attribute ") + name);
  return ML_H5_OK;
}

int read_attr(ml_h5 f, const char* group, const char* name, hid_t type, void* data) {
  if (!f || !name) return fail(ML_H5_EINVAL, "attribute: bad file or name");
  Id g = open_group(f, group_path(group), false);
  if (!g.ok()) return fail(ML_H5_ENOTFOUND, "no group " + group_path(group));
  if (H5Aexists(g, name) <= 0) return fail(ML_H5_ENOTFOUND, std::string("no attribute ") + name);
  Id a(H5Aopen(g, name, H5P_DEFAULT), H5Aclose);
  if (!a.ok() || H5Aread(a, type, data) < 0) return fail(ML_H5_EIO, std::string("cannot read attribute ") + name);
  return ML_H5_OK;
}

int write_dataset(ml_h5 f, const std::string& path, hid_t type, hsize_t n, const void* data) {
  if (H5Lexists(f->file, path.c_str(), H5P_DEFAULT) > 0) H5Ldelete(f->file, path.c_str(), H5P_DEFAULT);
  Id sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
  Id ds(H5Dcreate2(f->file, path.c_str(), type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(ML_H5_EIO, "cannot create dataset " + path);
  if (n > 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    return fail(ML_H5_EIO, "cannot write dataset " + path);
  return ML_H5_OK;
}

// a 1-D dataset read whole into a vector of `T` through memory type `type`
template <class T>
int read_dataset(ml_h5 f, const std::string& path, hid_t type, std::vector<T>& out) {
  if (H5Lexists(f->file, path.c_str(), H5P_DEFAULT) <= 0) return fail(ML_H5_ENOTFOUND, "no dataset " + path);
  Id ds(H5Dopen2(f->file, path.c_str(), H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) return fail(ML_H5_EIO, "cannot open dataset " + path);
  Id sp(H5Dget_space(ds), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  if (n < 0) return fail(ML_H5_EIO, "bad extent of " + path);
  out.resize(size_t(n));
  if (n > 0 && H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
    return fail(ML_H5_EIO, "cannot read dataset " + path);
  return ML_H5_OK;
}

struct BoxDims {
  int lo[3], hi[3];
  long long len(int d, int grow) const { return (long long)(hi[d] - lo[d] + 1 + 2 * grow); }
};

bool get_box(int dim, const int lo[], const int hi[], int b, BoxDims& B) {
  for (int d = 0; d < 3; ++d) {
    B.lo[d] = d < dim ? lo[3 * b + d] : 0;
    B.hi[d] = d < dim ? hi[3 * b + d] : 0;
    if (B.hi[d] < B.lo[d]) return false;
  }
  return true;
}

long long cells(const BoxDims& B, int dim, int grow) {
  long long n = 1;
  for (int d = 0; d < dim; ++d) n *= B.len(d, grow);
  return n;
}

int put_var_attrs(ml_h5 f, const std::string& lev, const std::string& name, int ncomp, int ghost, int out_ghost,
                  const char* object_type) {
  const std::string g = lev + "/" + name + "_attributes";
  int gv[3] = {ghost, ghost, ghost}, ov[3] = {out_ghost, out_ghost, out_ghost};
  Id iv = intvect_type(f->dim);
  int rc;
  if ((rc = ml_h5_put_int(f, g.c_str(), "comps", ncomp))) return rc;
  if ((rc = write_attr(f, g.c_str(), "ghost", iv, gv))) return rc;
  if ((rc = write_attr(f, g.c_str(), "outputGhost", iv, ov))) return rc;
  return ml_h5_put_string(f, g.c_str(), "objectType", object_type);
}

// ---- describe -----------------------------------------------------------

std::string class_of(hid_t t) {
  switch (H5Tget_class(t)) {
    case H5T_INTEGER:
      return "int";
    case H5T_FLOAT:
      return "float";
    case H5T_STRING:
      return "string";
    case H5T_COMPOUND: {
      std::string s = "compound(";
      const int n = H5Tget_nmembers(t);
      for (int m = 0; m < n; ++m) {
        char* nm = H5Tget_member_name(t, unsigned(m));
        s += (m ? "," : "") + std::string(nm ? nm : "?");
        if (nm) H5free_memory(nm);
      }
      return s + ")";
    }
    default:
      return "other";
  }
}

struct DescribeCtx {
  std::vector<std::string> lines;
  std::string path;
};

herr_t attr_cb(hid_t loc, const char* name, const H5A_info_t*, void* op) {
  auto* C = static_cast<DescribeCtx*>(op);
  Id a(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  Id t(H5Aget_type(a), H5Tclose);
  C->lines.push_back(C->path + "@" + name + " attr " + class_of(t));
  return 0;
}

herr_t obj_cb(hid_t root, const char* name, const H5O_info_t* info, void* op) {
  auto* C = static_cast<DescribeCtx*>(op);
  const std::string p = std::strcmp(name, ".") == 0 ? "/" : "/" + std::string(name);
  if (info->type == H5O_TYPE_GROUP) {
    C->lines.push_back(p + " group");
    Id g(H5Gopen2(root, name, H5P_DEFAULT), H5Gclose);
    C->path = p;
    hsize_t idx = 0;
    H5Aiterate2(g, H5_INDEX_NAME, H5_ITER_INC, &idx, attr_cb, C);
  } else if (info->type == H5O_TYPE_DATASET) {
    Id ds(H5Dopen2(root, name, H5P_DEFAULT), H5Dclose);
    Id t(H5Dget_type(ds), H5Tclose);
    Id sp(H5Dget_space(ds), H5Sclose);
    C->lines.push_back(p + " dataset " + class_of(t) + " " + std::to_string((long long)H5Sget_simple_extent_npoints(sp)));
    C->path = p;
    hsize_t idx = 0;
    H5Aiterate2(ds, H5_INDEX_NAME, H5_ITER_INC, &idx, attr_cb, C);
  }
  return 0;
}

}  // namespace

extern "C" {

const char* ml_h5_error(void) { return g_err.c_str(); }

int ml_h5_create(const char* path, int dim, ml_h5* out) {
  if (!path || !out || dim < 2 || dim > 3) return fail(ML_H5_EINVAL, "ml_h5_create: bad arguments");
  quiet();
  hid_t fid = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (fid < 0) return fail(ML_H5_EIO, std::string("cannot create ") + path);
  ml_h5 f = new ml_h5_s;
  f->file = fid;
  f->dim = dim;
  f->writable = true;
  // HDF5Handle::CREATE writes the space dimension and a precision probe
  int rc = ml_h5_put_int(f, "Chombo_global", "SpaceDim", dim);
  if (!rc) rc = ml_h5_put_real(f, "Chombo_global", "testReal", 0.0);
  if (rc) {
    ml_h5_close(f);
    return rc;
  }
  *out = f;
  return ML_H5_OK;
}

int ml_h5_open(const char* path, ml_h5* out) {
  if (!path || !out) return fail(ML_H5_EINVAL, "ml_h5_open: bad arguments");
  quiet();
  hid_t fid = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (fid < 0) return fail(ML_H5_EIO, std::string("cannot open ") + path);
  ml_h5 f = new ml_h5_s;
  f->file = fid;
  int dim = 0;
  if (read_attr(f, "Chombo_global", "SpaceDim", H5T_NATIVE_INT, &dim) || dim < 2 || dim > 3) {
    ml_h5_close(f);
    return fail(ML_H5_EIO, std::string("not a Chombo file (no /Chombo_global/SpaceDim): ") + path);
  }
  f->dim = dim;
  *out = f;
  return ML_H5_OK;
}

int ml_h5_close(ml_h5 f) {
  if (!f) return ML_H5_EINVAL;
  int rc = ML_H5_OK;
  if (f->file >= 0 && H5Fclose(f->file) < 0) rc = fail(ML_H5_EIO, "H5Fclose failed");
  delete f;
  return rc;
}

int ml_h5_dim(ml_h5 f) { return f ? f->dim : 0; }

int ml_h5_put_int(ml_h5 f, const char* group, const char* name, int v) {
  return write_attr(f, group, name, H5T_NATIVE_INT, &v);
}

int ml_h5_put_real(ml_h5 f, const char* group, const char* name, double v) {
  return write_attr(f, group, name, H5T_NATIVE_DOUBLE, &v);
}

int ml_h5_put_string(ml_h5 f, const char* group, const char* name, const char* v) {
  if (!v) return fail(ML_H5_EINVAL, "null string");
  Id t(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(t, std::max<size_t>(1, std::strlen(v)));
  std::string s(v);
  s.resize(std::max<size_t>(1, s.size()), '\0');
  return write_attr(f, group, name, t, s.data());
}

int ml_h5_put_box(ml_h5 f, const char* group, const char* name, const int lo[], const int hi[]) {
  if (!f || !lo || !hi) return fail(ML_H5_EINVAL, "box: bad arguments");
  int v[6];
  for (int d = 0; d < f->dim; ++d) {
    v[d] = lo[d];
    v[f->dim + d] = hi[d];
  }
  Id t = box_type(f->dim);
  return write_attr(f, group, name, t, v);
}

int ml_h5_get_int(ml_h5 f, const char* group, const char* name, int* v) {
  return v ? read_attr(f, group, name, H5T_NATIVE_INT, v) : ML_H5_EINVAL;
}

int ml_h5_get_real(ml_h5 f, const char* group, const char* name, double* v) {
  return v ? read_attr(f, group, name, H5T_NATIVE_DOUBLE, v) : ML_H5_EINVAL;
}

int ml_h5_get_string(ml_h5 f, const char* group, const char* name, char* buf, int cap) {
  if (!f || !name || !buf || cap < 1) return fail(ML_H5_EINVAL, "string: bad arguments");
  Id g = open_group(f, group_path(group), false);
  if (!g.ok()) return fail(ML_H5_ENOTFOUND, "no group " + group_path(group));
  if (H5Aexists(g, name) <= 0) return fail(ML_H5_ENOTFOUND, std::string("no attribute ") + name);
  Id a(H5Aopen(g, name, H5P_DEFAULT), H5Aclose);
  Id ft(H5Aget_type(a), H5Tclose);
  if (H5Tget_class(ft) != H5T_STRING || H5Tis_variable_str(ft) > 0)
    return fail(ML_H5_EIO, std::string("not a fixed-length string: ") + name);
  const size_t n = H5Tget_size(ft);
  Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_size(mt, n + 1);
  std::vector<char> tmp(n + 1, '\0');
  if (H5Aread(a, mt, tmp.data()) < 0) return fail(ML_H5_EIO, std::string("cannot read attribute ") + name);
  const size_t len = std::min(std::strlen(tmp.data()), size_t(cap - 1));
  std::memcpy(buf, tmp.data(), len);
  buf[len] = '\0';
  return ML_H5_OK;
}

int ml_h5_get_box(ml_h5 f, const char* group, const char* name, int lo[], int hi[]) {
  if (!f || !lo || !hi) return fail(ML_H5_EINVAL, "box: bad arguments");
  int v[6] = {0, 0, 0, 0, 0, 0};
  Id t = box_type(f->dim);
  int rc = read_attr(f, group, name, t, v);
  if (rc) return rc;
  for (int d = 0; d < f->dim; ++d) {
    lo[d] = v[d];
    hi[d] = v[f->dim + d];
  }
  return ML_H5_OK;
}

int ml_h5_put_boxes(ml_h5 f, int level, int nbox, const int lo[], const int hi[]) {
  if (!f || !f->writable || nbox < 0 || (nbox > 0 && (!lo || !hi))) return fail(ML_H5_EINVAL, "boxes: bad arguments");
  const std::string lev = level_group(level);
  Id g = open_group(f, lev, true);
  if (!g.ok()) return fail(ML_H5_EIO, "cannot open " + lev);
  std::vector<int> v(size_t(nbox) * 2 * f->dim);
  for (int b = 0; b < nbox; ++b)
    for (int d = 0; d < f->dim; ++d) {
      v[size_t(b) * 2 * f->dim + d] = lo[3 * b + d];
      v[size_t(b) * 2 * f->dim + f->dim + d] = hi[3 * b + d];
    }
  Id t = box_type(f->dim);
  int rc = write_dataset(f, lev + "/boxes", t, hsize_t(nbox), v.data());
  if (rc) return rc;
  std::vector<int> procs(size_t(nbox), 0);
  return write_dataset(f, lev + "/Processors", H5T_NATIVE_INT, hsize_t(nbox), procs.data());
}

int ml_h5_get_boxes(ml_h5 f, int level, int cap, int* nbox, int lo[], int hi[]) {
  if (!f || !nbox || cap < 0) return fail(ML_H5_EINVAL, "boxes: bad arguments");
  Id t = box_type(f->dim);
  std::vector<int> v;
  // read through the compound type: a vector of ints, 2*dim per box
  const std::string path = level_group(level) + "/boxes";
  if (H5Lexists(f->file, level_group(level).c_str(), H5P_DEFAULT) <= 0 ||
      H5Lexists(f->file, path.c_str(), H5P_DEFAULT) <= 0)
    return fail(ML_H5_ENOTFOUND, "no dataset " + path);
  Id ds(H5Dopen2(f->file, path.c_str(), H5P_DEFAULT), H5Dclose);
  Id sp(H5Dget_space(ds), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  if (n < 0) return fail(ML_H5_EIO, "bad extent of " + path);
  v.resize(size_t(n) * 2 * f->dim);
  if (n > 0 && H5Dread(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
    return fail(ML_H5_EIO, "cannot read " + path);
  *nbox = int(n);
  for (int b = 0; b < std::min<int>(int(n), cap); ++b)
    for (int d = 0; d < 3; ++d) {
      lo[3 * b + d] = d < f->dim ? v[size_t(b) * 2 * f->dim + d] : 0;
      hi[3 * b + d] = d < f->dim ? v[size_t(b) * 2 * f->dim + f->dim + d] : 0;
    }
  return ML_H5_OK;
}

int ml_h5_put_fabs(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int ncomp,
                   int ghost, int out_ghost, const double* const fabs[]) {
  if (!f || !f->writable || !name || !*name || nbox < 0 || ncomp < 1 || ghost < 0 || out_ghost < 0 ||
      out_ghost > ghost || (nbox > 0 && (!lo || !hi || !fabs)))
    return fail(ML_H5_EINVAL, "put_fabs: bad arguments");
  const int dim = f->dim;
  const std::string lev = level_group(level);
  Id g = open_group(f, lev, true);
  if (!g.ok()) return fail(ML_H5_EIO, "cannot open " + lev);
  std::vector<long long> off(static_cast<size_t>(nbox) + 1, 0);
  std::vector<BoxDims> boxes(static_cast<size_t>(nbox));
  for (int b = 0; b < nbox; ++b) {
    if (!get_box(dim, lo, hi, b, boxes[b]) || !fabs[b]) return fail(ML_H5_EINVAL, "put_fabs: bad box or FAB");
    off[b + 1] = off[b] + ncomp * cells(boxes[b], dim, out_ghost);
  }
  std::vector<double> data(size_t(off[nbox]));
  for (int b = 0; b < nbox; ++b) {
    const BoxDims& B = boxes[b];
    const long long fx = B.len(0, ghost), fy = B.len(1, ghost), fz = dim == 3 ? B.len(2, ghost) : 1;
    const long long fn = fx * fy * fz;
    const long long ox = B.len(0, out_ghost), oy = B.len(1, out_ghost), oz = dim == 3 ? B.len(2, out_ghost) : 1;
    const int s = ghost - out_ghost;
    double* dst = data.data() + off[b];
    for (int c = 0; c < ncomp; ++c)
      for (long long k = 0; k < oz; ++k)
        for (long long j = 0; j < oy; ++j)
          for (long long i = 0; i < ox; ++i)
            *dst++ = fabs[b][c * fn + (i + s) + fx * ((j + s) + fy * (k + (dim == 3 ? s : 0)))];
  }
  int rc = write_dataset(f, lev + "/" + name + ":datatype=0", H5T_NATIVE_DOUBLE, hsize_t(data.size()), data.data());
  if (rc) return rc;
  rc = write_dataset(f, lev + "/" + name + ":offsets=0", H5T_NATIVE_LLONG, hsize_t(off.size()), off.data());
  if (rc) return rc;
  return put_var_attrs(f, lev, name, ncomp, out_ghost, out_ghost, "FArrayBox");
}

int ml_h5_fab_info(ml_h5 f, int level, const char* name, int* ncomp, int* out_ghost) {
  if (!f || !name || !ncomp || !out_ghost) return fail(ML_H5_EINVAL, "fab_info: bad arguments");
  const std::string g = level_group(level) + "/" + name + "_attributes";
  int rc = ml_h5_get_int(f, g.c_str(), "comps", ncomp);
  if (rc) return rc;
  int ov[3] = {0, 0, 0};
  Id iv = intvect_type(f->dim);
  rc = read_attr(f, g.c_str(), "outputGhost", iv, ov);
  if (rc == ML_H5_ENOTFOUND) {
    *out_ghost = 0;
    return ML_H5_OK;
  }
  if (rc) return rc;
  for (int d = 1; d < f->dim; ++d)
    if (ov[d] != ov[0]) return fail(ML_H5_EIO, std::string("anisotropic outputGhost of ") + name);
  *out_ghost = ov[0];
  return ML_H5_OK;
}

int ml_h5_get_fabs(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int c0,
                   int ncomp, int ghost, double* const fabs[]) {
  if (!f || !name || nbox < 0 || c0 < 0 || ncomp < 1 || ghost < 0 || (nbox > 0 && (!lo || !hi || !fabs)))
    return fail(ML_H5_EINVAL, "get_fabs: bad arguments");
  const int dim = f->dim;
  int fcomp = 0, og = 0;
  int rc = ml_h5_fab_info(f, level, name, &fcomp, &og);
  if (rc) return rc;
  if (c0 + ncomp > fcomp) return fail(ML_H5_EINVAL, std::string("get_fabs: ") + name + " has fewer components");
  const std::string lev = level_group(level);
  std::vector<long long> off;
  std::vector<double> data;
  if ((rc = read_dataset(f, lev + "/" + name + ":offsets=0", H5T_NATIVE_LLONG, off))) return rc;
  if ((rc = read_dataset(f, lev + "/" + name + ":datatype=0", H5T_NATIVE_DOUBLE, data))) return rc;
  if (off.size() != size_t(nbox) + 1) return fail(ML_H5_EINVAL, std::string("get_fabs: box count differs for ") + name);
  for (int b = 0; b < nbox; ++b) {
    BoxDims B;
    if (!get_box(dim, lo, hi, b, B) || !fabs[b]) return fail(ML_H5_EINVAL, "get_fabs: bad box or FAB");
    const long long per = cells(B, dim, og);
    if (off[b + 1] - off[b] != per * fcomp || off[b + 1] > (long long)data.size())
      return fail(ML_H5_EIO, std::string("get_fabs: box sizes differ from the file for ") + name);
    const long long fx = B.len(0, ghost), fy = B.len(1, ghost), fz = dim == 3 ? B.len(2, ghost) : 1;
    const long long fn = fx * fy * fz;
    const long long ox = B.len(0, og), oy = B.len(1, og), oz = dim == 3 ? B.len(2, og) : 1;
    const int g = std::min(og, ghost);   // stored ghost layers that fit in the FAB
    for (int c = 0; c < ncomp; ++c) {
      const double* src = data.data() + off[b] + (c0 + c) * per;
      for (long long k = 0; k < oz; ++k)
        for (long long j = 0; j < oy; ++j)
          for (long long i = 0; i < ox; ++i) {
            // position relative to the valid box
            const long long ri = i - og, rj = j - og, rk = dim == 3 ? k - og : 0;
            if (ri < -g || rj < -g || rk < -(dim == 3 ? g : 0)) continue;
            if (ri >= B.len(0, 0) + g || rj >= B.len(1, 0) + g || (dim == 3 && rk >= B.len(2, 0) + g)) continue;
            fabs[b][c * fn + (ri + ghost) + fx * ((rj + ghost) + fy * (rk + (dim == 3 ? ghost : 0)))] =
                src[i + ox * (j + oy * k)];
          }
    }
  }
  return ML_H5_OK;
}

int ml_h5_put_faces(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int ncomp,
                    const double* const faces[]) {
  if (!f || !f->writable || !name || !*name || nbox < 0 || ncomp < 1 || (nbox > 0 && (!lo || !hi || !faces)))
    return fail(ML_H5_EINVAL, "put_faces: bad arguments");
  const int dim = f->dim;
  const std::string lev = level_group(level);
  Id g = open_group(f, lev, true);
  if (!g.ok()) return fail(ML_H5_EIO, "cannot open " + lev);
  std::vector<long long> off(static_cast<size_t>(nbox) + 1, 0);
  std::vector<BoxDims> boxes(static_cast<size_t>(nbox));
  for (int b = 0; b < nbox; ++b) {
    if (!get_box(dim, lo, hi, b, boxes[b])) return fail(ML_H5_EINVAL, "put_faces: bad box");
    long long n = 0;
    for (int d = 0; d < dim; ++d) {
      if (!faces[b * dim + d]) return fail(ML_H5_EINVAL, "put_faces: null face FAB");
      n += cells(boxes[b], dim, 0) / boxes[b].len(d, 0) * (boxes[b].len(d, 0) + 1);
    }
    off[b + 1] = off[b] + ncomp * n;
  }
  std::vector<double> data(size_t(off[nbox]));
  double* dst = data.data();
  for (int b = 0; b < nbox; ++b)
    for (int d = 0; d < dim; ++d) {
      const long long n = cells(boxes[b], dim, 0) / boxes[b].len(d, 0) * (boxes[b].len(d, 0) + 1);
      std::memcpy(dst, faces[b * dim + d], sizeof(double) * size_t(ncomp * n));
      dst += ncomp * n;
    }
  int rc = write_dataset(f, lev + "/" + name + ":datatype=0", H5T_NATIVE_DOUBLE, hsize_t(data.size()), data.data());
  if (rc) return rc;
  rc = write_dataset(f, lev + "/" + name + ":offsets=0", H5T_NATIVE_LLONG, hsize_t(off.size()), off.data());
  if (rc) return rc;
  return put_var_attrs(f, lev, name, ncomp, 0, 0, "unknown");
}

int ml_h5_get_faces(ml_h5 f, int level, const char* name, int nbox, const int lo[], const int hi[], int ncomp,
                    double* const faces[]) {
  if (!f || !name || nbox < 0 || ncomp < 1 || (nbox > 0 && (!lo || !hi || !faces)))
    return fail(ML_H5_EINVAL, "get_faces: bad arguments");
  const int dim = f->dim;
  int fcomp = 0, og = 0;
  int rc = ml_h5_fab_info(f, level, name, &fcomp, &og);
  if (rc) return rc;
  if (fcomp != ncomp || og != 0) return fail(ML_H5_EINVAL, std::string("get_faces: layout differs for ") + name);
  const std::string lev = level_group(level);
  std::vector<long long> off;
  std::vector<double> data;
  if ((rc = read_dataset(f, lev + "/" + name + ":offsets=0", H5T_NATIVE_LLONG, off))) return rc;
  if ((rc = read_dataset(f, lev + "/" + name + ":datatype=0", H5T_NATIVE_DOUBLE, data))) return rc;
  if (off.size() != size_t(nbox) + 1) return fail(ML_H5_EINVAL, std::string("get_faces: box count differs for ") + name);
  for (int b = 0; b < nbox; ++b) {
    BoxDims B;
    if (!get_box(dim, lo, hi, b, B)) return fail(ML_H5_EINVAL, "get_faces: bad box");
    const double* src = data.data() + off[b];
    long long tot = 0;
    for (int d = 0; d < dim; ++d) {
      const long long n = cells(B, dim, 0) / B.len(d, 0) * (B.len(d, 0) + 1);
      if (!faces[b * dim + d]) return fail(ML_H5_EINVAL, "get_faces: null face FAB");
      if (off[b] + tot + ncomp * n > (long long)data.size()) return fail(ML_H5_EIO, "get_faces: file too short");
      std::memcpy(faces[b * dim + d], src + tot, sizeof(double) * size_t(ncomp * n));
      tot += ncomp * n;
    }
    if (off[b + 1] - off[b] != tot) return fail(ML_H5_EIO, std::string("get_faces: box sizes differ for ") + name);
  }
  return ML_H5_OK;
}

namespace {

int level_common(ml_h5 f, int level, const ml_h5_level* L) {
  if (!f || !L || level < 0) return fail(ML_H5_EINVAL, "level header: bad arguments");
  const std::string g = level_group(level);
  int rc;
  if ((rc = ml_h5_put_int(f, g.c_str(), "ref_ratio", L->ref_ratio))) return rc;
  if ((rc = ml_h5_put_real(f, g.c_str(), "dx", L->dx))) return rc;
  if ((rc = ml_h5_put_real(f, g.c_str(), "dt", L->dt))) return rc;
  if ((rc = ml_h5_put_real(f, g.c_str(), "time", L->time))) return rc;
  return ml_h5_put_box(f, g.c_str(), "prob_domain", L->lo, L->hi);
}

int names_and_count(ml_h5 f, int ncomp, const char* const names[]) {
  if (ncomp < 0 || (ncomp > 0 && !names)) return fail(ML_H5_EINVAL, "component names: bad arguments");
  int rc = ml_h5_put_int(f, "/", "num_components", ncomp);
  for (int c = 0; c < ncomp && !rc; ++c) {
    if (!names[c]) return fail(ML_H5_EINVAL, "null component name");
    rc = ml_h5_put_string(f, "/", ("component_" + std::to_string(c)).c_str(), names[c]);
  }
  return rc;
}

int amr_root(ml_h5 f, int num_levels, int iteration, double time) {
  if (!f || num_levels < 1) return fail(ML_H5_EINVAL, "root header: bad arguments");
  int rc;
  if ((rc = ml_h5_put_int(f, "/", "max_level", num_levels - 1))) return rc;
  if ((rc = ml_h5_put_int(f, "/", "num_levels", num_levels))) return rc;
  if ((rc = ml_h5_put_int(f, "/", "iteration", iteration))) return rc;
  return ml_h5_put_real(f, "/", "time", time);
}

}  // namespace

int ml_h5_write_checkpoint_header(ml_h5 f, int num_levels, int iteration, double time, const int regrid_interval[],
                                  const int steps_since_regrid[], int ncomp, const char* const names[]) {
  int rc = amr_root(f, num_levels, iteration, time);
  if (rc) return rc;
  if (num_levels > 1 && (!regrid_interval || !steps_since_regrid))
    return fail(ML_H5_EINVAL, "checkpoint header: regrid arrays needed for several levels");
  for (int l = 0; l < num_levels - 1 && !rc; ++l) {
    rc = ml_h5_put_int(f, "/", ("regrid_interval_" + std::to_string(l)).c_str(), regrid_interval[l]);
    if (!rc) rc = ml_h5_put_int(f, "/", ("steps_since_regrid_" + std::to_string(l)).c_str(), steps_since_regrid[l]);
  }
  return rc ? rc : names_and_count(f, ncomp, names);
}

int ml_h5_write_checkpoint_level_header(ml_h5 f, int level, const ml_h5_level* L) {
  int rc = level_common(f, level, L);
  if (rc) return rc;
  const std::string g = level_group(level);
  if ((rc = ml_h5_put_int(f, g.c_str(), "tag_buffer_size", L->tag_buffer_size))) return rc;
  for (int d = 0; d < f->dim && !rc; ++d)
    rc = ml_h5_put_int(f, g.c_str(), ("is_periodic_" + std::to_string(d)).c_str(), L->periodic[d] ? 1 : 0);
  return rc;
}

int ml_h5_write_plot_header(ml_h5 f, int num_levels, int iteration, double time, int ncomp, const char* const names[],
                            const ml_h5_plot_params* p) {
  int rc = amr_root(f, num_levels, iteration, time);
  if (!rc) rc = names_and_count(f, ncomp, names);
  if (rc || !p) return rc;
  const struct {
    const char* k;
    double v;
  } reals[] = {{"C", p->compositionRatio}, {"k", p->heatConductivityRatio}, {"cp", p->specificHeatRatio},
               {"pc", p->waterDistributionCoeff}, {"S", p->stefan}, {"L", p->lewis}, {"D", p->darcy}};
  for (const auto& r : reals)
    if ((rc = ml_h5_put_real(f, "/", r.k, r.v))) return rc;
  if ((rc = ml_h5_put_string(f, "/", "timescale", p->timescale ? p->timescale : ""))) return rc;
  return ml_h5_put_string(f, "/", "velocity_scale", p->velocity_scale ? p->velocity_scale : "");
}

int ml_h5_write_plot_level_header(ml_h5 f, int level, const ml_h5_level* L) {
  int rc = level_common(f, level, L);
  if (rc) return rc;
  return ml_h5_put_int(f, level_group(level).c_str(), "timestepFailed", L->timestep_failed);
}

int ml_h5_read_level_header(ml_h5 f, int level, ml_h5_level* L) {
  if (!f || !L || level < 0) return fail(ML_H5_EINVAL, "read_level_header: bad arguments");
  std::memset(L, 0, sizeof(*L));
  const std::string g = level_group(level);
  int rc;
  if ((rc = ml_h5_get_int(f, g.c_str(), "ref_ratio", &L->ref_ratio))) return rc;
  if ((rc = ml_h5_get_real(f, g.c_str(), "dx", &L->dx))) return rc;
  if ((rc = ml_h5_get_real(f, g.c_str(), "dt", &L->dt))) return rc;
  if ((rc = ml_h5_get_real(f, g.c_str(), "time", &L->time))) return rc;
  if ((rc = ml_h5_get_box(f, g.c_str(), "prob_domain", L->lo, L->hi))) return rc;
  for (int d = 0; d < f->dim; ++d) {
    int v = 0;
    rc = ml_h5_get_int(f, g.c_str(), ("is_periodic_" + std::to_string(d)).c_str(), &v);
    if (rc && rc != ML_H5_ENOTFOUND) return rc;
    L->periodic[d] = rc ? 0 : (v == 1);
  }
  rc = ml_h5_get_int(f, g.c_str(), "tag_buffer_size", &L->tag_buffer_size);
  if (rc && rc != ML_H5_ENOTFOUND) return rc;
  rc = ml_h5_get_int(f, g.c_str(), "timestepFailed", &L->timestep_failed);
  if (rc && rc != ML_H5_ENOTFOUND) return rc;
  return ML_H5_OK;
}

int ml_h5_describe(ml_h5 f, char* buf, int cap) {
  if (!f || cap < 0 || (cap > 0 && !buf)) return fail(ML_H5_EINVAL, "describe: bad arguments");
  DescribeCtx C;
  if (H5Ovisit(f->file, H5_INDEX_NAME, H5_ITER_INC, obj_cb, &C) < 0) return fail(ML_H5_EIO, "H5Ovisit failed");
  std::sort(C.lines.begin(), C.lines.end());
  std::string s;
  for (const auto& l : C.lines) s += l + "\n";
  if (cap > 0) {
    const size_t n = std::min(s.size(), size_t(cap - 1));
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return int(s.size());
}

}  // extern "C"
