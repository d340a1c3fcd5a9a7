// direct.hip -- FP64 O(N^2) direct sum on the GPU (the reference's
// DirectCalc3D / EvaluateField, bbfmm.c:471-504, 880-913, with start = 0):
//   out_t = q_t * sum_{j : x_j != x_t} q_j / |x_t - x_j|
// for a list of targets.  It is the accuracy reference of the solve.
// Sources are split into chunks over blockIdx.y; per-chunk partial sums are
// written and reduced in a fixed order (deterministic result).
#include "common.h"
#include "kernels.h"

#include <vector>

namespace bb {

constexpr int kDirectChunks = 64;

__global__ void __launch_bounds__(256) k_direct(int N, const double4 *__restrict__ src, int nt,
                                                const int *__restrict__ targets, double *__restrict__ part)
{
  __shared__ double4 tile[256];
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int chunk = blockIdx.y;
  const int per = (N + kDirectChunks - 1) / kDirectChunks;
  const int j0 = chunk * per, j1 = min(N, j0 + per);
  double4 x = make_double4(0, 0, 0, 0);
  if (t < nt) x = src[targets[t]];
  double s = 0;
  for (int base = j0; base < j1; base += 256) {
    const int j = base + threadIdx.x;
    if (j < j1) tile[threadIdx.x] = src[j];
    __syncthreads();
    const int nk = min(256, j1 - base);
    for (int k = 0; k < nk; k++) {
      const double4 y = tile[k];
      if (y.x != x.x || y.y != x.y || y.z != x.z) {
        const double dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
        s += y.w / sqrt(dx * dx + dy * dy + dz * dz);
      }
    }
    __syncthreads();
  }
  if (t < nt) part[(size_t)chunk * nt + t] = s;
}

__global__ void k_direct_reduce(int nt, const int *__restrict__ targets, const double4 *__restrict__ src,
                                const double *__restrict__ part, double *__restrict__ out)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double s = 0;
  for (int c = 0; c < kDirectChunks; c++) s += part[(size_t)c * nt + t];
  out[t] = src[targets[t]].w * s;
}

} // namespace bb

extern "C" int bbfmm_amd_direct(int N, const double *pos, const double *q, int nt, const int *targets,
                                double *out)
{
  using namespace bb;
  clear_error();
  if (N < 0 || nt < 0 || (N > 0 && (!pos || !q)) || (nt > 0 && !out))
    return set_error(BBFMM_E_ARG, "bbfmm_amd_direct: invalid argument");
  if (nt == 0) return BBFMM_OK;
  std::vector<int> tg((size_t)nt);
  for (int t = 0; t < nt; t++) {
    tg[t] = targets ? targets[t] : t;
    if (tg[t] < 0 || tg[t] >= N) return set_error(BBFMM_E_ARG, "bbfmm_amd_direct: target %d out of range", tg[t]);
  }
  std::vector<double4> h((size_t)N);
  for (int i = 0; i < N; i++) h[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], q[i]);
  double4 *d_src = nullptr;
  int *d_tg = nullptr;
  double *d_part = nullptr, *d_out = nullptr;
  int rc = BBFMM_OK;
  hipError_t e = hipSuccess;
  do {
    if ((e = hipMalloc(&d_src, sizeof(double4) * N)) != hipSuccess) break;
    if ((e = hipMalloc(&d_tg, sizeof(int) * nt)) != hipSuccess) break;
    if ((e = hipMalloc(&d_part, sizeof(double) * (size_t)nt * kDirectChunks)) != hipSuccess) break;
    if ((e = hipMalloc(&d_out, sizeof(double) * nt)) != hipSuccess) break;
    if ((e = hipMemcpy(d_src, h.data(), sizeof(double4) * N, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemcpy(d_tg, tg.data(), sizeof(int) * nt, hipMemcpyHostToDevice)) != hipSuccess) break;
    hipLaunchKernelGGL(k_direct, dim3((nt + 255) / 256, kDirectChunks), dim3(256), 0, nullptr, N, d_src, nt, d_tg,
                       d_part);
    if ((e = hipGetLastError()) != hipSuccess) break;
    hipLaunchKernelGGL(k_direct_reduce, dim3((nt + 255) / 256), dim3(256), 0, nullptr, nt, d_tg, d_src, d_part,
                       d_out);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if ((e = hipMemcpy(out, d_out, sizeof(double) * nt, hipMemcpyDeviceToHost)) != hipSuccess) break;
  } while (0);
  if (e != hipSuccess) rc = set_error(BBFMM_E_HIP, "bbfmm_amd_direct: %s", hipGetErrorString(e));
  if (d_src) (void)hipFree(d_src);
  if (d_tg) (void)hipFree(d_tg);
  if (d_part) (void)hipFree(d_part);
  if (d_out) (void)hipFree(d_out);
  return rc;
}

// ---------------------------------------------------------------- potentials and gradients
namespace bb {

// One vector per blockIdx.z; sources over blockIdx.y chunks; per target
// (potential, gradient x, y, z) partial sums written per chunk.
__global__ void __launch_bounds__(256) k_direct_field(int N, const double4 *__restrict__ xyz,
                                                      const double *__restrict__ q, int nt,
                                                      const int *__restrict__ targets, int nchunk,
                                                      double4 *__restrict__ part)
{
  __shared__ double4 tile[256];
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int chunk = blockIdx.y, v = blockIdx.z;
  const int per = (N + nchunk - 1) / nchunk;
  const int j0 = chunk * per, j1 = min(N, j0 + per);
  double4 x = make_double4(0, 0, 0, 0);
  if (t < nt) x = xyz[targets[t]];
  const double *qv = q + (size_t)v * N;
  double s = 0, gx = 0, gy = 0, gz = 0;
  for (int base = j0; base < j1; base += 256) {
    const int j = base + threadIdx.x;
    if (j < j1) {
      double4 y = xyz[j];
      y.w = qv[j];
      tile[threadIdx.x] = y;
    }
    __syncthreads();
    const int nk = min(256, j1 - base);
    for (int k = 0; k < nk; k++) {
      const double4 y = tile[k];
      if (y.x != x.x || y.y != x.y || y.z != x.z) {
        const double dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
        const double ri = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
        const double p = y.w * ri;
        const double g = p * ri * ri;  // q / r^3; d(1/|x - y|)/dx = (y - x) / r^3
        s += p;
        gx += g * dx;
        gy += g * dy;
        gz += g * dz;
      }
    }
    __syncthreads();
  }
  if (t < nt) part[((size_t)v * nchunk + chunk) * nt + t] = make_double4(s, gx, gy, gz);
}

__global__ void k_direct_field_reduce(int nt, int nvec, int nchunk, const double4 *__restrict__ part,
                                      double *__restrict__ pot, double *__restrict__ grad)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (t >= nt || v >= nvec) return;
  double s = 0, gx = 0, gy = 0, gz = 0;
  for (int c = 0; c < nchunk; c++) {
    const double4 p = part[((size_t)v * nchunk + c) * nt + t];
    s += p.x;
    gx += p.y;
    gy += p.z;
    gz += p.w;
  }
  const size_t o = (size_t)v * nt + t;
  if (pot) pot[o] = s;
  if (grad) {
    grad[3 * o] = gx;
    grad[3 * o + 1] = gy;
    grad[3 * o + 2] = gz;
  }
}

} // namespace bb

extern "C" int bbfmm_amd_direct_field(int N, const double *pos, int nvec, const double *q, int nt,
                                      const int *targets, double *pot, double *grad)
{
  using namespace bb;
  clear_error();
  if (N < 0 || nt < 0 || nvec < 1 || nvec > 65535 || (N > 0 && (!pos || !q)) || (nt > 0 && !pot && !grad))
    return set_error(BBFMM_E_ARG, "bbfmm_amd_direct_field: invalid argument");
  if (nt == 0) return BBFMM_OK;
  BB_TRY(check_device());
  std::vector<int> tg((size_t)nt);
  for (int t = 0; t < nt; t++) {
    tg[t] = targets ? targets[t] : t;
    if (tg[t] < 0 || tg[t] >= N)
      return set_error(BBFMM_E_ARG, "bbfmm_amd_direct_field: target %d out of range", tg[t]);
  }
  std::vector<double4> h((size_t)N);
  for (int i = 0; i < N; i++) h[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0);
  // enough blocks to fill the GPU, bounded partial-sum storage
  const int tblocks = (nt + 255) / 256;
  int nchunk = (4096 + tblocks * nvec - 1) / (tblocks * nvec);
  nchunk = nchunk < 1 ? 1 : nchunk > 64 ? 64 : nchunk;
  double4 *d_xyz = nullptr, *d_part = nullptr;
  double *d_q = nullptr, *d_pot = nullptr, *d_grad = nullptr;
  int *d_tg = nullptr;
  int rc = BBFMM_OK;
  hipError_t e = hipSuccess;
  const size_t nout = (size_t)nvec * nt;
  do {
    if ((e = hipMalloc(&d_xyz, sizeof(double4) * N)) != hipSuccess) break;
    if ((e = hipMalloc(&d_q, sizeof(double) * (size_t)nvec * N)) != hipSuccess) break;
    if ((e = hipMalloc(&d_tg, sizeof(int) * nt)) != hipSuccess) break;
    if ((e = hipMalloc(&d_part, sizeof(double4) * nout * nchunk)) != hipSuccess) break;
    if ((e = hipMalloc(&d_pot, sizeof(double) * nout)) != hipSuccess) break;
    if ((e = hipMalloc(&d_grad, sizeof(double) * 3 * nout)) != hipSuccess) break;
    if ((e = hipMemcpy(d_xyz, h.data(), sizeof(double4) * N, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemcpy(d_q, q, sizeof(double) * (size_t)nvec * N, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemcpy(d_tg, tg.data(), sizeof(int) * nt, hipMemcpyHostToDevice)) != hipSuccess) break;
    hipLaunchKernelGGL(k_direct_field, dim3(tblocks, nchunk, nvec), dim3(256), 0, nullptr, N, d_xyz, d_q, nt, d_tg,
                       nchunk, d_part);
    if ((e = hipGetLastError()) != hipSuccess) break;
    hipLaunchKernelGGL(k_direct_field_reduce, dim3(tblocks, nvec), dim3(256), 0, nullptr, nt, nvec, nchunk, d_part,
                       d_pot, d_grad);
    if ((e = hipGetLastError()) != hipSuccess) break;
    if (pot && (e = hipMemcpy(pot, d_pot, sizeof(double) * nout, hipMemcpyDeviceToHost)) != hipSuccess) break;
    if (grad && (e = hipMemcpy(grad, d_grad, sizeof(double) * 3 * nout, hipMemcpyDeviceToHost)) != hipSuccess) break;
  } while (0);
  if (e != hipSuccess) rc = set_error(BBFMM_E_HIP, "bbfmm_amd_direct_field: %s", hipGetErrorString(e));
  for (void *ptr : {(void *)d_xyz, (void *)d_part, (void *)d_q, (void *)d_pot, (void *)d_grad, (void *)d_tg})
    if (ptr) (void)hipFree(ptr);
  return rc;
}
