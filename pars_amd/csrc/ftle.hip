// FTLE post-processing in FP64 on the device: flow-map Jacobian, right
// Cauchy-Green tensor C = J^T J, symmetric 3x3 eigenvalues, FTLE or
// fractional anisotropy.  The reference's structure is
// pa/source/stages/ftle_map_generator.cpp:22-60; its gradient
// (vector_field.cpp:49-93) is replaced by textbook central differences.
#include "pa_internal.hpp"

#pragma clang fp contract(off)

namespace pa_amd {
namespace {

constexpr double kPi = 3.14159265358979323846;

// Eigenvalues of a symmetric 3x3 matrix by the trigonometric closed form
// (O. K. Smith, "Eigenvalues of a symmetric 3x3 matrix", CACM 4(4), 1961):
// A = q I + p B with q = tr(A)/3, p = sqrt(tr((A-qI)^2)/6), det(B)/2 =
// cos(3 phi).  Returns ascending e[0] <= e[1] <= e[2].
__device__ __forceinline__ void eig_sym3(double a00, double a01, double a02, double a11, double a12, double a22,
                                         double e[3]) {
  const double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  const double q = (a00 + a11 + a22) / 3.0;
  if (p1 == 0.0) {
    double x = a00, y = a11, z = a22, t;
    if (x > y) { t = x; x = y; y = t; }
    if (y > z) { t = y; y = z; z = t; }
    if (x > y) { t = x; x = y; y = t; }
    e[0] = x; e[1] = y; e[2] = z;
    return;
  }
  const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  const double p = sqrt(p2 / 6.0);
  const double ip = 1.0 / p;
  const double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip;
  const double c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
  const double det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02);
  const double r = det / 2.0;
  double phi;
  if (r <= -1.0) phi = kPi / 3.0;
  else if (r >= 1.0) phi = 0.0;
  else phi = acos(r) / 3.0;
  const double hi = q + 2.0 * p * cos(phi);
  const double lo = q + 2.0 * p * cos(phi + 2.0 * kPi / 3.0);
  e[2] = hi;
  e[0] = lo;
  e[1] = 3.0 * q - hi - lo;
}

__global__ __launch_bounds__(256) void k_ftle(const float* __restrict__ fm, int d0, int d1, int d2, double h0,
                                              double h1, double h2, double abs_t, int mode,
                                              float* __restrict__ out, double* __restrict__ lmax) {
  const int64_t n = int64_t(d0) * d1 * d2;
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t z = v % d2, y = (v / d2) % d1, x = v / (int64_t(d1) * d2);
  const int64_t ijk[3] = {x, y, z};
  const int dims[3] = {d0, d1, d2};
  const int64_t stride[3] = {int64_t(d1) * d2, d2, 1};
  const double hs[3] = {h0, h1, h2};
  double J[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int64_t i = ijk[j];
    int64_t lo = i, hi = i;
    double den = 0.0;
    if (dims[j] > 1) {
      if (i > 0 && i < dims[j] - 1) {
        lo = i - 1; hi = i + 1; den = 2.0 * hs[j];
      } else if (i == 0) {
        hi = i + 1; den = hs[j];
      } else {
        lo = i - 1; den = hs[j];
      }
    }
    const float* fh = fm + 3 * (v + (hi - i) * stride[j]);
    const float* fl = fm + 3 * (v + (lo - i) * stride[j]);
#pragma unroll
    for (int c = 0; c < 3; ++c) J[c][j] = den > 0.0 ? (double(fh[c]) - double(fl[c])) / den : 0.0;
  }
  double C[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) C[a][b] = J[0][a] * J[0][b] + J[1][a] * J[1][b] + J[2][a] * J[2][b];
  double e[3];
  eig_sym3(C[0][0], C[0][1], C[0][2], C[1][1], C[1][2], C[2][2], e);
  double value;
  if (mode == PA_FTLE_REGULAR) {
    value = log(sqrt(e[2])) / abs_t;
  } else {
    const double num = (e[0] - e[1]) * (e[0] - e[1]) + (e[1] - e[2]) * (e[1] - e[2]) + (e[2] - e[0]) * (e[2] - e[0]);
    const double den = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
    value = sqrt(0.5) * sqrt(num) / sqrt(den);
  }
  out[v] = float(value);
  if (lmax) lmax[v] = e[2];
}

__global__ __launch_bounds__(256) void k_eig_sym3(const double* __restrict__ a, int64_t n, double* __restrict__ e) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* m = a + 6 * i;
  double v[3];
  eig_sym3(m[0], m[1], m[2], m[3], m[4], m[5], v);
  e[3 * i] = v[0];
  e[3 * i + 1] = v[1];
  e[3 * i + 2] = v[2];
}

}  // namespace
}  // namespace pa_amd

using namespace pa_amd;

extern "C" pa_status pa_gpu_sym3_eigenvalues(pa_gpu_ctx* ctx, const double* d_sym, size_t n, double* d_eigenvalues) {
  if (!ctx || (n && (!d_sym || !d_eigenvalues))) return PA_E_INVALID;
  if (n == 0) return PA_OK;
  launch_begin(ctx, "eig_sym3");
  hipLaunchKernelGGL(k_eig_sym3, dim3(unsigned((n + 255) / 256)), dim3(256), 0, ctx->stream, d_sym, int64_t(n),
                     d_eigenvalues);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}

extern "C" pa_status pa_gpu_ftle_generate(pa_gpu_ctx* ctx, const float* d_flow_map, const int32_t dims[3],
                                          const float spacing[3], double time, pa_ftle_mode mode, float* d_ftle,
                                          double* d_lambda_max) {
  if (!ctx || !d_flow_map || !dims || !spacing || !d_ftle) return PA_E_INVALID;
  if (mode != PA_FTLE_REGULAR && mode != PA_FTLE_FRACTIONAL_ANISOTROPY)
    return fail(ctx, PA_E_INVALID, "unknown FTLE mode");
  for (int i = 0; i < 3; ++i)
    if (dims[i] < 1 || !(spacing[i] > 0.0f)) return fail(ctx, PA_E_INVALID, "bad flow map geometry");
  if (!(time != 0.0)) return fail(ctx, PA_E_INVALID, "FTLE time must be non-zero");
  const int64_t n = int64_t(dims[0]) * dims[1] * dims[2];
  const int64_t blocks = (n + 255) / 256;
  launch_begin(ctx, "ftle");
  hipLaunchKernelGGL(k_ftle, dim3(unsigned(blocks)), dim3(256), 0, ctx->stream, d_flow_map, dims[0], dims[1], dims[2],
                     double(spacing[0]), double(spacing[1]), double(spacing[2]), fabs(time), int(mode), d_ftle,
                     d_lambda_max);
  launch_end(ctx);
  PA_HIP(ctx, hipGetLastError());
  return PA_OK;
}
