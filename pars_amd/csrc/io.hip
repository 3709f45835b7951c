// Block loader: data_io::load_vector_field (pa/source/stages/data_io.cpp:91-115)
// reading one rank's ghosted hyperslab of the HDF5 dataset vectors[X][Y][Z][3]
// and the file attribute spacing[3], normalised by its largest component, into
// padded device bricks (FP32, BF16 or FP16) without the reference's
// multi_array copy.  libhdf5 (1.10 or later: 64-bit hid_t) is opened at run
// time, so the library itself does not depend on it.
#include <dlfcn.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "pa_internal.hpp"

namespace pa_amd {
namespace {

using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kDefault = 0;    // H5P_DEFAULT
constexpr unsigned kReadOnly = 0u;  // H5F_ACC_RDONLY
constexpr int kSelectSet = 0;    // H5S_SELECT_SET

struct Hdf5 {
  void* so = nullptr;
  herr_t (*open)();
  herr_t (*get_libversion)(unsigned*, unsigned*, unsigned*);
  herr_t (*eset_auto)(hid_t, void*, void*);
  hid_t (*fopen)(const char*, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*dopen)(hid_t, const char*, hid_t);
  herr_t (*dclose)(hid_t);
  hid_t (*dget_space)(hid_t);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t*, hsize_t*);
  long long (*sget_npoints)(hid_t);
  herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t*, const hsize_t*, const hsize_t*, const hsize_t*);
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
  herr_t (*sclose)(hid_t);
  hid_t (*aopen)(hid_t, const char*, hid_t);
  hid_t (*aget_space)(hid_t);
  herr_t (*aread)(hid_t, hid_t, void*);
  herr_t (*aclose)(hid_t);
  hid_t* native_float = nullptr;  // H5T_NATIVE_FLOAT_g, valid after H5open
  std::string error;
};

template <typename F>
bool bind(Hdf5& h, F& f, const char* name) {
  f = reinterpret_cast<F>(dlsym(h.so, name));
  if (!f) h.error = std::string("libhdf5 lacks ") + name;
  return f != nullptr;
}

Hdf5& hdf5() {
  static Hdf5 h;
  static std::once_flag once;
  std::call_once(once, [] {
    const char* names[] = {"libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so"};
    for (const char* n : names)
      if ((h.so = dlopen(n, RTLD_NOW | RTLD_LOCAL)) != nullptr) break;
    if (!h.so) {
      h.error = "libhdf5 not found (tried libhdf5.so, /opt/conda/lib/libhdf5.so)";
      return;
    }
    bool ok = bind(h, h.open, "H5open") && bind(h, h.get_libversion, "H5get_libversion") &&
              bind(h, h.eset_auto, "H5Eset_auto2") && bind(h, h.fopen, "H5Fopen") && bind(h, h.fclose, "H5Fclose") &&
              bind(h, h.dopen, "H5Dopen2") && bind(h, h.dclose, "H5Dclose") && bind(h, h.dget_space, "H5Dget_space") &&
              bind(h, h.dread, "H5Dread") && bind(h, h.sget_ndims, "H5Sget_simple_extent_ndims") &&
              bind(h, h.sget_dims, "H5Sget_simple_extent_dims") &&
              bind(h, h.sget_npoints, "H5Sget_simple_extent_npoints") &&
              bind(h, h.sselect_hyperslab, "H5Sselect_hyperslab") && bind(h, h.screate_simple, "H5Screate_simple") &&
              bind(h, h.sclose, "H5Sclose") && bind(h, h.aopen, "H5Aopen") && bind(h, h.aget_space, "H5Aget_space") &&
              bind(h, h.aread, "H5Aread") && bind(h, h.aclose, "H5Aclose");
    if (ok) {
      h.native_float = reinterpret_cast<hid_t*>(dlsym(h.so, "H5T_NATIVE_FLOAT_g"));
      if (!h.native_float) ok = false, h.error = "libhdf5 lacks H5T_NATIVE_FLOAT_g";
    }
    unsigned maj = 0, min = 0, rel = 0;
    if (ok && (h.open() < 0 || h.get_libversion(&maj, &min, &rel) < 0 || maj != 1 || min < 10))
      ok = false, h.error = "libhdf5 1.10 or later is needed (64-bit identifiers)";
    if (ok) h.eset_auto(0, nullptr, nullptr);  // errors come back as status codes, not printed
    if (!ok) {
      dlclose(h.so);
      h.so = nullptr;
    }
  });
  return h;
}

// Closes HDF5 identifiers in reverse order of opening.
struct Closer {
  Hdf5& h;
  hid_t file = -1, dset = -1, fspace = -1, mspace = -1, attr = -1, aspace = -1;
  ~Closer() {
    if (aspace >= 0) h.sclose(aspace);
    if (attr >= 0) h.aclose(attr);
    if (mspace >= 0) h.sclose(mspace);
    if (fspace >= 0) h.sclose(fspace);
    if (dset >= 0) h.dclose(dset);
    if (file >= 0) h.fclose(file);
  }
};

pa_status read_block(const char* path, const int64_t offset[3], const int32_t ghosted[3], const int32_t block[3],
                     float* xyz, float offset_out[3], float size_out[3], float spacing_out[3], std::string& err) {
  if (!path || !offset || !ghosted || !block || !xyz) {
    err = "read_vector_block: null argument";
    return PA_E_INVALID;
  }
  for (int i = 0; i < 3; ++i)
    if (offset[i] < 0 || ghosted[i] < 2 || block[i] < 1) {
      err = "read_vector_block: offsets >= 0, ghosted block sizes >= 2, block sizes >= 1";
      return PA_E_INVALID;
    }
  Hdf5& h = hdf5();
  if (!h.so) {
    err = h.error;
    return PA_E_UNSUPPORTED;
  }
  static std::mutex reader;  // HDF5 builds are not always thread-safe: one reader at a time
  std::lock_guard<std::mutex> lock(reader);
  Closer c{h};
  if ((c.file = h.fopen(path, kReadOnly, kDefault)) < 0) {
    err = std::string("cannot open ") + path;
    return PA_E_INVALID;
  }
  if ((c.dset = h.dopen(c.file, "vectors", kDefault)) < 0) {
    err = "no dataset 'vectors'";
    return PA_E_INVALID;
  }
  c.fspace = h.dget_space(c.dset);
  hsize_t dims[4] = {0, 0, 0, 0};
  if (c.fspace < 0 || h.sget_ndims(c.fspace) != 4 || h.sget_dims(c.fspace, dims, nullptr) != 4 || dims[3] != 3) {
    err = "'vectors' is not a [X][Y][Z][3] dataset";
    return PA_E_INVALID;
  }
  // the ghosted hyperslab at the rank's offset (data_io.cpp:94-97)
  const hsize_t start[4] = {hsize_t(offset[0]), hsize_t(offset[1]), hsize_t(offset[2]), 0};
  const hsize_t count[4] = {hsize_t(ghosted[0]), hsize_t(ghosted[1]), hsize_t(ghosted[2]), 3};
  for (int i = 0; i < 3; ++i)
    if (start[i] + count[i] > dims[i]) {
      err = "hyperslab outside 'vectors' on axis " + std::to_string(i);
      return PA_E_INVALID;
    }
  const hsize_t one[4] = {1, 1, 1, 1};
  if (h.sselect_hyperslab(c.fspace, kSelectSet, start, one, count, nullptr) < 0 ||
      (c.mspace = h.screate_simple(4, count, nullptr)) < 0 ||
      h.dread(c.dset, *h.native_float, c.mspace, c.fspace, kDefault, xyz) < 0) {
    err = "reading the hyperslab of 'vectors' failed";
    return PA_E_INVALID;
  }
  // the spacing attribute, divided by its largest component (data_io.cpp:108-111)
  float raw[3] = {0, 0, 0};
  if ((c.attr = h.aopen(c.file, "spacing", kDefault)) < 0 || (c.aspace = h.aget_space(c.attr)) < 0 ||
      h.sget_npoints(c.aspace) != 3 || h.aread(c.attr, *h.native_float, raw) < 0) {
    err = "no 3-element attribute 'spacing'";
    return PA_E_INVALID;
  }
  const float mx = std::max(std::max(raw[0], raw[1]), raw[2]);  // Eigen maxCoeff
  if (!(mx > 0.0f)) {
    err = "'spacing' has no positive component";
    return PA_E_INVALID;
  }
  float sp[3];
  for (int i = 0; i < 3; ++i) sp[i] = raw[i] / mx;
  // offset and size in FP32 as data_io.cpp:113-114 computes them
  for (int i = 0; i < 3; ++i) {
    if (spacing_out) spacing_out[i] = sp[i];
    if (offset_out) offset_out[i] = static_cast<float>(static_cast<int32_t>(offset[i])) * sp[i];
    if (size_out) size_out[i] = static_cast<float>(block[i]) * sp[i];
  }
  return PA_OK;
}

pa_status read_dims(const char* path, int64_t dims_out[3], std::string& err) {
  if (!path || !dims_out) {
    err = "read_vector_dims: null argument";
    return PA_E_INVALID;
  }
  Hdf5& h = hdf5();
  if (!h.so) {
    err = h.error;
    return PA_E_UNSUPPORTED;
  }
  static std::mutex reader;
  std::lock_guard<std::mutex> lock(reader);
  Closer c{h};
  if ((c.file = h.fopen(path, kReadOnly, kDefault)) < 0) {
    err = std::string("cannot open ") + path;
    return PA_E_INVALID;
  }
  if ((c.dset = h.dopen(c.file, "vectors", kDefault)) < 0) {
    err = "no dataset 'vectors'";
    return PA_E_INVALID;
  }
  c.fspace = h.dget_space(c.dset);
  hsize_t dims[4] = {0, 0, 0, 0};
  if (c.fspace < 0 || h.sget_ndims(c.fspace) != 4 || h.sget_dims(c.fspace, dims, nullptr) != 4 || dims[3] != 3) {
    err = "'vectors' is not a [X][Y][Z][3] dataset";
    return PA_E_INVALID;
  }
  for (int i = 0; i < 3; ++i) dims_out[i] = int64_t(dims[i]);
  return PA_OK;
}

}  // namespace
}  // namespace pa_amd

using namespace pa_amd;

extern "C" {

pa_status pa_gpu_read_vector_block(const char* path, const int64_t offset[3], const int32_t ghosted[3],
                                  const int32_t block[3], float* h_xyz, float offset_out[3], float size_out[3],
                                  float spacing_out[3], char* error, size_t error_len) {
  std::string err;
  const pa_status s = read_block(path, offset, ghosted, block, h_xyz, offset_out, size_out, spacing_out, err);
  if (error && error_len) {
    std::strncpy(error, err.c_str(), error_len - 1);
    error[error_len - 1] = '\0';
  }
  return s;
}

pa_status pa_gpu_read_vector_dims(const char* path, int64_t dims[3], char* error, size_t error_len) {
  std::string err;
  const pa_status s = read_dims(path, dims, err);
  if (error && error_len) {
    std::strncpy(error, err.c_str(), error_len - 1);
    error[error_len - 1] = '\0';
  }
  return s;
}

pa_status pa_gpu_field_load(pa_gpu_ctx* ctx, const char* path, const int64_t offset[3], const int32_t ghosted[3],
                            const int32_t block[3], pa_storage storage, pa_gpu_field** out) {
  if (!ctx || !out || !ghosted) return PA_E_INVALID;
  *out = nullptr;
  for (int i = 0; i < 3; ++i)
    if (ghosted[i] < 2) return fail(ctx, PA_E_INVALID, "field_load: ghosted block sizes >= 2");
  const size_t n = size_t(ghosted[0]) * size_t(ghosted[1]) * size_t(ghosted[2]);
  // read into pinned memory, so the upload runs at full link speed
  float* staging = nullptr;
  PA_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&staging), n * 3 * sizeof(float), hipHostMallocDefault));
  float off[3], size[3], sp[3];
  std::string err;
  pa_status s = read_block(path, offset, ghosted, block, staging, off, size, sp, err);
  if (s == PA_OK) s = pa_gpu_field_create(ctx, staging, ghosted, off, size, sp, storage, out);
  else fail(ctx, s, "field_load: " + err);
  const hipError_t e = hipHostFree(staging);
  if (s == PA_OK && e != hipSuccess) return check_hip(ctx, e, "hipHostFree");
  return s;
}

}  // extern "C"
