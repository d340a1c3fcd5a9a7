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
