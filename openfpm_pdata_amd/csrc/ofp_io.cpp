// ofp_io.cpp -- vector_dist save / load in the reference's HDF5 checkpoint
// layout (see include/ofp_io.h).  Host code over the HDF5 C library.
#include "ofp_io.h"

#include <hdf5.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ofp_nn.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// H5 handles closed on scope exit
struct H5Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Id() {
    if (id >= 0 && close) close(id);
  }
  H5Id(const H5Id&) = delete;
  H5Id& operator=(const H5Id&) = delete;
};

// the library prints its error stack by default; the codes are reported here
struct QuietH5 {
  H5E_auto2_t fn = nullptr;
  void* data = nullptr;
  QuietH5() {
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~QuietH5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

int check_fields(const ofp_h5_field* f, int n, uint32_t obj) {
  if (n < 0 || (n > 0 && !f)) return fail(OFP_E_ARG, "null field list");
  for (int k = 0; k < n; ++k) {
    const uint32_t e = f[k].elem_bytes;
    if (!f[k].ptr) return fail(OFP_E_ARG, "field %d: null pointer", k);
    if (e != 1 && e != 2 && e != 4 && e != 8) return fail(OFP_E_ARG, "field %d: elem_bytes %u", k, e);
    if (uint64_t(f[k].offset) + e > obj)
      return fail(OFP_E_ARG, "field %d: offset %u + %u beyond the object (%u bytes)", k, f[k].offset, e, obj);
  }
  return OFP_OK;
}

size_t pos_bytes(int dtype) { return dtype == OFP_F64 ? 8 : 4; }

// Byte image of a checkpoint: the blocks' sizes and their bytes.
struct Image {
  std::vector<int32_t> sizes;
  std::vector<char> bytes;
};

int read_image(const char* filename, Image* im) {
  QuietH5 quiet;
  H5Id file(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return fail(OFP_E_ARG, "cannot open %s as HDF5", filename);
  for (int which = 0; which < 2; ++which) {
    const char* name = which == 0 ? "metadata" : "vector_dist";
    H5Id ds(H5Dopen2(file.id, name, H5P_DEFAULT), H5Dclose);
    if (ds.id < 0) return fail(OFP_E_ARG, "%s: no dataset \"%s\"", filename, name);
    H5Id sp(H5Dget_space(ds.id), H5Sclose);
    if (sp.id < 0 || H5Sget_simple_extent_ndims(sp.id) != 1)
      return fail(OFP_E_ARG, "%s: dataset \"%s\" is not one-dimensional", filename, name);
    hsize_t dim = 0;
    H5Sget_simple_extent_dims(sp.id, &dim, nullptr);
    herr_t rc;
    if (which == 0) {
      im->sizes.resize(size_t(dim));
      rc = dim ? H5Dread(ds.id, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, im->sizes.data()) : 0;
    } else {
      im->bytes.resize(size_t(dim));
      rc = dim ? H5Dread(ds.id, H5T_NATIVE_SCHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, im->bytes.data()) : 0;
    }
    if (rc < 0) return fail(OFP_E_ARG, "%s: cannot read \"%s\"", filename, name);
  }
  uint64_t total = 0;
  for (int32_t b : im->sizes) {
    if (b < 0) return fail(OFP_E_ARG, "%s: negative block size", filename);
    total += uint64_t(b);
  }
  if (total != im->bytes.size())
    return fail(OFP_E_ARG, "%s: block sizes sum to %llu, data holds %zu bytes", filename,
                (unsigned long long)total, im->bytes.size());
  return OFP_OK;
}

// Walk the blocks: fn(block bytes, n particles) for each; checks every block's
// structure against (position size, object size).
template <class F>
int walk_blocks(const Image& im, const char* filename, size_t psz, uint32_t obj, F&& fn) {
  size_t at = 0;
  for (size_t b = 0; b < im.sizes.size(); ++b) {
    const size_t len = size_t(im.sizes[b]);
    const char* p = im.bytes.data() + at;
    uint64_t n1 = 0, n2 = 0;
    if (len < 16) return fail(OFP_E_ARG, "%s: block %zu shorter than its two counts", filename, b);
    std::memcpy(&n1, p, 8);
    const uint64_t after_pos = 8 + n1 * psz;
    if (after_pos + 8 > len)
      return fail(OFP_E_ARG, "%s: block %zu: %llu positions of %zu bytes do not fit %zu bytes", filename, b,
                  (unsigned long long)n1, psz, len);
    std::memcpy(&n2, p + after_pos, 8);
    if (n2 != n1 || after_pos + 8 + n2 * uint64_t(obj) != len)
      return fail(OFP_E_ARG,
                  "%s: block %zu: %llu positions, %llu property objects, %zu bytes: not a layout of %zu-byte "
                  "positions and %u-byte property objects",
                  filename, b, (unsigned long long)n1, (unsigned long long)n2, len, psz, obj);
    int rc = fn(p + 8, p + after_pos + 8, size_t(n1));
    if (rc != OFP_OK) return rc;
    at += len;
  }
  return OFP_OK;
}

}  // namespace

extern "C" {

const char* ofp_io_last_error(void) { return g_err.c_str(); }

int ofp_h5_save(const char* filename, int dtype, const void* pos, size_t pos_stride, size_t n,
                const ofp_h5_field* fields, int n_fields, uint32_t obj_bytes) {
  if (!filename) return fail(OFP_E_ARG, "ofp_h5_save: null file name");
  if (dtype != OFP_F32 && dtype != OFP_F64) return fail(OFP_E_ARG, "ofp_h5_save: bad dtype %d", dtype);
  if (n > 0 && !pos) return fail(OFP_E_ARG, "ofp_h5_save: null positions");
  if (pos_stride < n) return fail(OFP_E_ARG, "ofp_h5_save: stride %zu < n %zu", pos_stride, n);
  int rc = check_fields(fields, n_fields, obj_bytes);
  if (rc != OFP_OK) return rc;
  const size_t s = pos_bytes(dtype), psz = 3 * s, per = psz + obj_bytes;
  // blocks of at most 2^30 bytes: the size field is an int32
  const size_t max_n = per ? ((size_t(1) << 30) - 16) / per : n;
  std::vector<int32_t> sizes;
  std::vector<char> bytes;
  bytes.reserve(16 * (n / (max_n ? max_n : 1) + 1) + n * per);
  size_t done = 0;
  do {
    const size_t m = n - done < max_n ? n - done : max_n;
    const size_t start = bytes.size();
    const uint64_t m64 = m;
    bytes.resize(start + 16 + m * per);
    char* p = bytes.data() + start;
    std::memcpy(p, &m64, 8);
    for (size_t i = 0; i < m; ++i)
      for (int k = 0; k < 3; ++k)
        std::memcpy(p + 8 + i * psz + k * s, static_cast<const char*>(pos) + (k * pos_stride + done + i) * s, s);
    char* q = p + 8 + m * psz;
    std::memcpy(q, &m64, 8);
    q += 8;
    std::memset(q, 0, m * obj_bytes);  // padding between members is written as zeros
    for (int f = 0; f < n_fields; ++f) {
      const uint32_t e = fields[f].elem_bytes;
      const char* src = static_cast<const char*>(fields[f].ptr) + done * e;
      for (size_t i = 0; i < m; ++i) std::memcpy(q + i * obj_bytes + fields[f].offset, src + i * e, e);
    }
    sizes.push_back(int32_t(16 + m * per));
    done += m;
  } while (done < n);

  QuietH5 quiet;
  H5Id file(H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (file.id < 0) return fail(OFP_E_ARG, "cannot create %s", filename);
  for (int which = 0; which < 2; ++which) {
    const hsize_t dim = which == 0 ? hsize_t(sizes.size()) : hsize_t(bytes.size());
    H5Id sp(H5Screate_simple(1, &dim, nullptr), H5Sclose);
    H5Id ds(H5Dcreate2(file.id, which == 0 ? "metadata" : "vector_dist", which == 0 ? H5T_STD_I32LE : H5T_STD_I8LE,
                       sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
            H5Dclose);
    if (sp.id < 0 || ds.id < 0) return fail(OFP_E_ARG, "%s: cannot create the datasets", filename);
    const herr_t w = dim == 0 ? 0
                     : which == 0
                         ? H5Dwrite(ds.id, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, sizes.data())
                         : H5Dwrite(ds.id, H5T_NATIVE_SCHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, bytes.data());
    if (w < 0) return fail(OFP_E_ARG, "%s: write failed", filename);
  }
  return OFP_OK;
}

int ofp_h5_load_count(const char* filename, int dtype, uint32_t obj_bytes, size_t* n) {
  if (!filename || !n) return fail(OFP_E_ARG, "ofp_h5_load_count: null argument");
  if (dtype != OFP_F32 && dtype != OFP_F64) return fail(OFP_E_ARG, "ofp_h5_load_count: bad dtype %d", dtype);
  Image im;
  int rc = read_image(filename, &im);
  if (rc != OFP_OK) return rc;
  size_t tot = 0;
  rc = walk_blocks(im, filename, 3 * pos_bytes(dtype), obj_bytes, [&](const char*, const char*, size_t m) {
    tot += m;
    return OFP_OK;
  });
  if (rc != OFP_OK) return rc;
  *n = tot;
  return OFP_OK;
}

int ofp_h5_load(const char* filename, int dtype, void* pos, size_t pos_stride, size_t n_cap,
                const ofp_h5_field* fields, int n_fields, uint32_t obj_bytes, size_t* n) {
  if (!filename || !n) return fail(OFP_E_ARG, "ofp_h5_load: null argument");
  if (dtype != OFP_F32 && dtype != OFP_F64) return fail(OFP_E_ARG, "ofp_h5_load: bad dtype %d", dtype);
  if (pos_stride < n_cap) return fail(OFP_E_ARG, "ofp_h5_load: stride %zu < capacity %zu", pos_stride, n_cap);
  int rc = check_fields(fields, n_fields, obj_bytes);
  if (rc != OFP_OK) return rc;
  Image im;
  rc = read_image(filename, &im);
  if (rc != OFP_OK) return rc;
  const size_t s = pos_bytes(dtype), psz = 3 * s;
  size_t tot = 0;
  rc = walk_blocks(im, filename, psz, obj_bytes, [&](const char*, const char*, size_t m) {
    tot += m;
    return OFP_OK;
  });
  if (rc != OFP_OK) return rc;
  if (tot > n_cap) return fail(OFP_E_CAPACITY, "ofp_h5_load: %zu particles, room for %zu", tot, n_cap);
  if (tot > 0 && !pos) return fail(OFP_E_ARG, "ofp_h5_load: null positions");
  size_t at = 0;
  rc = walk_blocks(im, filename, psz, obj_bytes, [&](const char* pp, const char* op, size_t m) {
    for (size_t i = 0; i < m; ++i)
      for (int k = 0; k < 3; ++k)
        std::memcpy(static_cast<char*>(pos) + (k * pos_stride + at + i) * s, pp + i * psz + k * s, s);
    for (int f = 0; f < n_fields; ++f) {
      const uint32_t e = fields[f].elem_bytes;
      char* dst = static_cast<char*>(fields[f].ptr) + at * e;
      for (size_t i = 0; i < m; ++i) std::memcpy(dst + i * e, op + i * obj_bytes + fields[f].offset, e);
    }
    at += m;
    return OFP_OK;
  });
  if (rc != OFP_OK) return rc;
  *n = tot;
  return OFP_OK;
}

}  // extern "C"
