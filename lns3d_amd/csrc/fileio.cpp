// Reference file formats: Fortran sequential unformatted files of the
// restart field (src/impdrv3d.f90:393-399, read by src/genini3d.f90:26-57),
// the mean field (src/genmean.f90:28-33), the grid (src/gengrid.f90:17-32)
// and the mesh metrics (src/gengrid.f90:76-80).  Host code only.
//
// Record framing (gfortran / flang): each record is one or more subrecords,
// each between a 4-byte leading and a 4-byte trailing length marker (little
// endian).  A leading marker is negative when the record continues in the
// next subrecord; a trailing marker is negative when its subrecord continues
// an earlier one.  Integers are 4 bytes, reals 8 bytes (the reference is
// built with real*8 defaults), complex 16 bytes.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lns3d_amd.h"

namespace lns3d {
int set_error(const char *file, int line, const char *msg);   // capi_penta.cpp
}

namespace {

constexpr int64_t kMaxSub = 2147483639;   // longest subrecord written

struct File {
  FILE *f = nullptr;
  ~File() { if (f) fclose(f); }
};

int fail(int code, const std::string &m) {
  lns3d::set_error("fileio", 0, m.c_str());
  return code;
}

bool rd32(FILE *f, int32_t &v) { return fread(&v, 4, 1, f) == 1; }

// one whole record (its subrecords joined)
int read_record(FILE *f, std::vector<char> &out, const char *what) {
  out.clear();
  bool more = true, first = true;
  while (more) {
    int32_t lead, trail;
    if (!rd32(f, lead)) return fail(LNS3D_EINVAL, std::string(what) + ": record missing (end of file)");
    const int64_t len = lead < 0 ? -(int64_t)lead : lead;
    more = lead < 0;
    const size_t at = out.size();
    out.resize(at + (size_t)len);
    if (len && fread(out.data() + at, 1, (size_t)len, f) != (size_t)len)
      return fail(LNS3D_EINVAL, std::string(what) + ": record shorter than its length marker");
    if (!rd32(f, trail)) return fail(LNS3D_EINVAL, std::string(what) + ": trailing record marker missing");
    const int64_t tl = trail < 0 ? -(int64_t)trail : trail;
    if (tl != len || (trail < 0) != !first)
      return fail(LNS3D_EINVAL, std::string(what) + ": record markers do not match");
    first = false;
  }
  return 0;
}

int write_record(FILE *f, const void *data, int64_t len, const char *what) {
  const char *p = (const char *)data;
  int64_t done = 0;
  do {
    const int64_t n = len - done < kMaxSub ? len - done : kMaxSub;
    const bool more = done + n < len, cont = done > 0;
    const int32_t lead = more ? -(int32_t)n : (int32_t)n, trail = cont ? -(int32_t)n : (int32_t)n;
    if (fwrite(&lead, 4, 1, f) != 1 || (n && fwrite(p + done, 1, (size_t)n, f) != (size_t)n) ||
        fwrite(&trail, 4, 1, f) != 1)
      return fail(LNS3D_EINVAL, std::string(what) + ": write failed");
    done += n;
  } while (done < len);
  return 0;
}

// the header record of restart and mean files: lstep, time, nx, ny, nz,
// ndof, Re, Ma, Pr, gamma, cv (68 bytes)
constexpr size_t kHeader = 4 + 8 + 4 * 4 + 5 * 8;

void pack_header(const lns3d_field_header *h, char *b) {
  size_t o = 0;
  auto put = [&](const void *v, size_t n) { memcpy(b + o, v, n); o += n; };
  const int32_t ls = h->lstep, dims[4] = {h->nx, h->ny, h->nz, h->ndof};
  const double ph[5] = {h->Re, h->Ma, h->Pr, h->gamma, h->cv};
  put(&ls, 4);
  put(&h->time, 8);
  put(dims, 16);
  put(ph, 40);
}

void unpack_header(const char *b, lns3d_field_header *h) {
  int32_t ls, dims[4];
  double ph[5];
  memcpy(&ls, b, 4);
  memcpy(&h->time, b + 4, 8);
  memcpy(dims, b + 12, 16);
  memcpy(ph, b + 28, 40);
  h->lstep = ls;
  h->nx = dims[0]; h->ny = dims[1]; h->nz = dims[2]; h->ndof = dims[3];
  h->Re = ph[0]; h->Ma = ph[1]; h->Pr = ph[2]; h->gamma = ph[3]; h->cv = ph[4];
}

int open_file(File &F, const char *path, const char *mode, const char *what) {
  if (!path) return fail(LNS3D_EINVAL, std::string(what) + ": null path");
  F.f = fopen(path, mode);
  if (!F.f) return fail(LNS3D_EINVAL, std::string(what) + ": cannot open " + path);
  return 0;
}

int read_header_rec(FILE *f, lns3d_field_header *h, const char *what) {
  std::vector<char> rec;
  if (int rc = read_record(f, rec, what)) return rc;
  if (rec.size() != kHeader)
    return fail(LNS3D_EINVAL, std::string(what) + ": first record is not a field header (68 bytes)");
  unpack_header(rec.data(), h);
  if (h->nx < 1 || h->ny < 1 || h->nz < 1 || h->ndof < 1)
    return fail(LNS3D_EINVAL, std::string(what) + ": header sizes out of range");
  return 0;
}

int64_t field_points(const lns3d_field_header *h) { return (int64_t)h->ndof * h->nx * h->ny * h->nz; }

// reads a data record of exactly `bytes` bytes into dst
int read_data(FILE *f, void *dst, int64_t bytes, const char *what) {
  std::vector<char> rec;
  if (int rc = read_record(f, rec, what)) return rc;
  if ((int64_t)rec.size() != bytes)
    return fail(LNS3D_EINVAL, std::string(what) + ": data record of " + std::to_string(rec.size()) +
                                  " bytes, expected " + std::to_string(bytes));
  memcpy(dst, rec.data(), (size_t)bytes);
  return 0;
}

int write_field(const char *path, const lns3d_field_header *h, const void *a, const void *b, int64_t es,
                const char *what) {
  if (!h || !a) return fail(LNS3D_EINVAL, std::string(what) + ": null argument");
  if (h->nx < 1 || h->ny < 1 || h->nz < 1 || h->ndof < 1)
    return fail(LNS3D_EINVAL, std::string(what) + ": header sizes out of range");
  File F;
  if (int rc = open_file(F, path, "wb", what)) return rc;
  char hb[kHeader];
  pack_header(h, hb);
  if (int rc = write_record(F.f, hb, kHeader, what)) return rc;
  const int64_t bytes = field_points(h) * es;
  if (int rc = write_record(F.f, a, bytes, what)) return rc;
  if (b)
    if (int rc = write_record(F.f, b, bytes, what)) return rc;
  if (fflush(F.f) != 0) return fail(LNS3D_EINVAL, std::string(what) + ": write failed");
  return 0;
}

}  // namespace

extern "C" int lns3d_read_header(const char *path, lns3d_field_header *h) {
  if (!h) return fail(LNS3D_EINVAL, "lns3d_read_header: null header");
  File F;
  if (int rc = open_file(F, path, "rb", "lns3d_read_header")) return rc;
  return read_header_rec(F.f, h, "lns3d_read_header");
}

extern "C" int lns3d_read_restart(const char *path, lns3d_field_header *h, void *v, void *vold,
                                  int *got_vold) {
  const char *what = "lns3d_read_restart";
  if (!h || !v) return fail(LNS3D_EINVAL, "lns3d_read_restart: null argument");
  File F;
  if (int rc = open_file(F, path, "rb", what)) return rc;
  if (int rc = read_header_rec(F.f, h, what)) return rc;
  const int64_t bytes = field_points(h) * 16;
  if (int rc = read_data(F.f, v, bytes, what)) return rc;
  int got = 0;
  if (vold) {
    int32_t probe;
    if (rd32(F.f, probe)) {                      // a third record: the previous field
      fseek(F.f, -4, SEEK_CUR);
      if (int rc = read_data(F.f, vold, bytes, what)) return rc;
      got = 1;
    }
  }
  if (got_vold) *got_vold = got;
  return 0;
}

extern "C" int lns3d_write_restart(const char *path, const lns3d_field_header *h, const void *v,
                                   const void *vold) {
  return write_field(path, h, v, vold, 16, "lns3d_write_restart");
}

extern "C" int lns3d_read_mean(const char *path, lns3d_field_header *h, double *vm) {
  const char *what = "lns3d_read_mean";
  if (!h || !vm) return fail(LNS3D_EINVAL, "lns3d_read_mean: null argument");
  File F;
  if (int rc = open_file(F, path, "rb", what)) return rc;
  if (int rc = read_header_rec(F.f, h, what)) return rc;
  return read_data(F.f, vm, field_points(h) * 8, what);
}

extern "C" int lns3d_write_mean(const char *path, const lns3d_field_header *h, const double *vm) {
  return write_field(path, h, vm, nullptr, 8, "lns3d_write_mean");
}

extern "C" int lns3d_read_grid(const char *path, int *nx, int *ny, int *nz, double *x, double *y) {
  const char *what = "lns3d_read_grid";
  if (!nx || !ny || !nz) return fail(LNS3D_EINVAL, "lns3d_read_grid: null size argument");
  File F;
  if (int rc = open_file(F, path, "rb", what)) return rc;
  std::vector<char> rec;
  if (int rc = read_record(F.f, rec, what)) return rc;
  if (rec.size() != 12) return fail(LNS3D_EINVAL, "lns3d_read_grid: first record is not nx, ny, nz");
  int32_t d[3];
  memcpy(d, rec.data(), 12);
  if (d[0] < 1 || d[1] < 1 || d[2] < 1) return fail(LNS3D_EINVAL, "lns3d_read_grid: sizes out of range");
  *nx = d[0]; *ny = d[1]; *nz = d[2];
  if (!x && !y) return 0;                       // sizes only
  const int64_t n = (int64_t)d[0] * d[1] * d[2];
  if (int rc = read_record(F.f, rec, what)) return rc;
  if ((int64_t)rec.size() != 3 * n * 8)
    return fail(LNS3D_EINVAL, "lns3d_read_grid: coordinate record is not x, y, z of nx*ny*nz reals");
  if (x) memcpy(x, rec.data(), (size_t)(n * 8));
  if (y) memcpy(y, rec.data() + n * 8, (size_t)(n * 8));
  return 0;
}

extern "C" int lns3d_write_grid(const char *path, int nx, int ny, int nz, const double *x, const double *y) {
  const char *what = "lns3d_write_grid";
  if (nx < 1 || ny < 1 || nz < 1 || !x || !y) return fail(LNS3D_EINVAL, "lns3d_write_grid: bad argument");
  File F;
  if (int rc = open_file(F, path, "wb", what)) return rc;
  const int32_t d[3] = {nx, ny, nz};
  if (int rc = write_record(F.f, d, 12, what)) return rc;
  const int64_t n = (int64_t)nx * ny * nz;
  std::vector<double> xyz((size_t)(3 * n), 0.0);     // z (third array) is zero, as gengrid sets it
  memcpy(xyz.data(), x, (size_t)(n * 8));
  memcpy(xyz.data() + n, y, (size_t)(n * 8));
  if (int rc = write_record(F.f, xyz.data(), 3 * n * 8, what)) return rc;
  if (fflush(F.f) != 0) return fail(LNS3D_EINVAL, "lns3d_write_grid: write failed");
  return 0;
}

extern "C" int lns3d_read_metric(const char *path, int nx, int ny, double *metric) {
  const char *what = "lns3d_read_metric";
  if (nx < 1 || ny < 1 || !metric) return fail(LNS3D_EINVAL, "lns3d_read_metric: bad argument");
  File F;
  if (int rc = open_file(F, path, "rb", what)) return rc;
  return read_data(F.f, metric, (int64_t)nx * ny * 10 * 8, what);
}

extern "C" int lns3d_write_metric(const char *path, int nx, int ny, const double *metric) {
  const char *what = "lns3d_write_metric";
  if (nx < 1 || ny < 1 || !metric) return fail(LNS3D_EINVAL, "lns3d_write_metric: bad argument");
  File F;
  if (int rc = open_file(F, path, "wb", what)) return rc;
  if (int rc = write_record(F.f, metric, (int64_t)nx * ny * 10 * 8, what)) return rc;
  if (fflush(F.f) != 0) return fail(LNS3D_EINVAL, "lns3d_write_metric: write failed");
  return 0;
}
