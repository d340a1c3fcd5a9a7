// ops_gpu.hip -- FP64 operator precompute on the GPU (ComputeKernelSVD,
// bbfmm.c:1209-1421), for n = 8 where the host version takes tens of
// seconds: K0 is sampled by a kernel, K0 K0^T and the 316 products
// E_k = U^T K0_k U are plain FP64 GEMMs (rocBLAS), and the n^3 x n^3
// symmetric eigenproblem is solved on the host by the same Jacobi routine
// as the host precompute.  Nodes and weights are computed on the host so
// that both paths sample from identical values.
#include "common.h"
#include "kernels.h"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace bb {

// K0[(op * n3 + s) * n3 + f] = 1 / |c_op + x_s / 2 - x_f / 2| / (w_s w_f)
__global__ void k_sample_K0(int n, const double *__restrict__ x, const double *__restrict__ w,
                            const int4 *__restrict__ offs, double *__restrict__ K0)
{
  const int n3 = n * n * n;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)kNumOps * n3 * n3) return;
  const int f = (int)(t % n3);
  const long long col = t / n3;
  const int s = (int)(col % n3);
  const int4 o = offs[col / n3];
  const double sx = o.x + 0.5 * x[s / (n * n)], sy = o.y + 0.5 * x[(s / n) % n], sz = o.z + 0.5 * x[s % n];
  const double dx = sx - 0.5 * x[f / (n * n)], dy = sy - 0.5 * x[(f / n) % n], dz = sz - 0.5 * x[f % n];
  K0[t] = 1.0 / sqrt(dx * dx + dy * dy + dz * dz) / (w[s] * w[f]);
}

namespace {
struct DevBufs {
  double *x = nullptr, *w = nullptr, *K0 = nullptr, *G = nullptr, *U = nullptr, *KU = nullptr, *E = nullptr;
  int4 *offs = nullptr;
  rocblas_handle h = nullptr;
  ~DevBufs()
  {
    for (void *p : {(void *)x, (void *)w, (void *)K0, (void *)G, (void *)U, (void *)KU, (void *)E, (void *)offs})
      if (p) (void)hipFree(p);
    if (h) (void)rocblas_destroy_handle(h);
  }
};
}  // namespace

#define RB_TRY(call)                                                                                 \
  do {                                                                                               \
    const rocblas_status st_ = (call);                                                               \
    if (st_ != rocblas_status_success)                                                               \
      return set_error(BBFMM_E_HIP, "%s failed: %s", #call, rocblas_status_to_string(st_));         \
  } while (0)

int operators_fp64_device(int n, int r, double *U, double *K, double *sigma)
{
  const int n3 = n * n * n;
  const long long ncol = (long long)kNumOps * n3;
  std::vector<double> x(n), w(n3);
  for (int m = 0; m < n; m++) x[m] = std::cos(M_PI * (2.0 * m + 1) / (2.0 * n));
  for (int a = 0; a < n3; a++) {
    const int i = a / (n * n), j = (a / n) % n, k = a % n;
    w[a] = 1.0 / (std::sqrt(1 - x[i] * x[i]) * std::sqrt(1 - x[j] * x[j]) * std::sqrt(1 - x[k] * x[k]));
  }
  std::vector<int4> offs;
  for (int ox = -3; ox <= 3; ox++)
    for (int oy = -3; oy <= 3; oy++)
      for (int oz = -3; oz <= 3; oz++)
        if (std::abs(ox) > 1 || std::abs(oy) > 1 || std::abs(oz) > 1) offs.push_back(int4{ox, oy, oz, 0});

  DevBufs d;
  const size_t nK0 = (size_t)ncol * n3;
  BB_HIP(hipMalloc(&d.x, sizeof(double) * n));
  BB_HIP(hipMalloc(&d.w, sizeof(double) * n3));
  BB_HIP(hipMalloc(&d.offs, sizeof(int4) * kNumOps));
  if (hipMalloc(&d.K0, sizeof(double) * nK0) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(BBFMM_E_NOMEM, "hipMalloc of %zu bytes for K0 failed", sizeof(double) * nK0);
  }
  BB_HIP(hipMalloc(&d.G, sizeof(double) * n3 * n3));
  BB_HIP(hipMemcpy(d.x, x.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  BB_HIP(hipMemcpy(d.w, w.data(), sizeof(double) * n3, hipMemcpyHostToDevice));
  BB_HIP(hipMemcpy(d.offs, offs.data(), sizeof(int4) * kNumOps, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sample_K0, dim3((unsigned)((nK0 + 255) / 256)), dim3(256), 0, 0, n, d.x, d.w, d.offs, d.K0);
  BB_HIP(hipGetLastError());

  RB_TRY(rocblas_create_handle(&d.h));
  const double one = 1.0, zero = 0.0;
  // G = K0 K0^T (K0 is n3 x ncol, column-major)
  RB_TRY(rocblas_dgemm(d.h, rocblas_operation_none, rocblas_operation_transpose, n3, n3, (rocblas_int)ncol, &one,
                       d.K0, n3, d.K0, n3, &zero, d.G, n3));
  std::vector<double> G((size_t)n3 * n3);
  BB_HIP(hipMemcpy(G.data(), d.G, sizeof(double) * G.size(), hipMemcpyDeviceToHost));
  // the GEMM fills the whole matrix; make it exactly symmetric from the
  // upper triangle (rows i <= column j), as the host path does
  for (int j = 0; j < n3; j++)
    for (int i = 0; i < j; i++) G[(size_t)i * n3 + j] = G[(size_t)j * n3 + i];
  std::vector<double> ev, V;
  jacobi_eig(n3, G, ev, V);
  for (int i = 0; i < n3; i++) sigma[i] = std::sqrt(std::max(ev[i], 0.0));
  for (int c = 0; c < r; c++)
    for (int i = 0; i < n3; i++) U[(size_t)c * n3 + i] = V[(size_t)c * n3 + i];

  BB_HIP(hipMalloc(&d.U, sizeof(double) * n3 * r));
  BB_HIP(hipMalloc(&d.KU, sizeof(double) * (size_t)kNumOps * n3 * r));
  BB_HIP(hipMalloc(&d.E, sizeof(double) * (size_t)kNumOps * r * r));
  BB_HIP(hipMemcpy(d.U, U, sizeof(double) * n3 * r, hipMemcpyHostToDevice));
  // KU_k = K0_k U, E_k = U^T KU_k for the 316 blocks
  RB_TRY(rocblas_dgemm_strided_batched(d.h, rocblas_operation_none, rocblas_operation_none, n3, r, n3, &one, d.K0, n3,
                                       (rocblas_stride)n3 * n3, d.U, n3, 0, &zero, d.KU, n3, (rocblas_stride)n3 * r,
                                       kNumOps));
  RB_TRY(rocblas_dgemm_strided_batched(d.h, rocblas_operation_transpose, rocblas_operation_none, r, r, n3, &one, d.U,
                                       n3, 0, d.KU, n3, (rocblas_stride)n3 * r, &zero, d.E, r, (rocblas_stride)r * r,
                                       kNumOps));
  BB_HIP(hipMemcpy(K, d.E, sizeof(double) * (size_t)kNumOps * r * r, hipMemcpyDeviceToHost));
  return BBFMM_OK;
}

}  // namespace bb
