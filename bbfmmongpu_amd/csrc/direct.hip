// direct.hip -- FP64 O(N^2) direct sum on the GPU (the reference's
// DirectCalc3D / EvaluateField, bbfmm.c:471-504, 880-913, with start = 0):
//   out_t = q_t * sum_{j : x_j != x_t} q_j / |x_t - x_j|
// for a list of targets.  It is the accuracy reference of the solve.
// Sources are split into chunks over blockIdx.y; per-chunk partial sums are
// written and reduced in a fixed order (deterministic result).
#include "common.h"
#include "kernels.h"

#include <mutex>
#include <vector>

namespace bb {

constexpr int kDirectChunks = 64;

// Partial sums of the direct sum: out_t needs (chunk, t) for chunk <
// kDirectChunks; pos is N x 3, q is N (device), targets NULL: t itself.
__global__ void __launch_bounds__(256) k_direct(int N, const double *__restrict__ pos, const double *__restrict__ q,
                                                int nt, const int *__restrict__ targets, double *__restrict__ part)
{
  __shared__ double4 tile[256];
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int chunk = blockIdx.y;
  const int per = (N + kDirectChunks - 1) / kDirectChunks;
  const int j0 = chunk * per, j1 = min(N, j0 + per);
  double3 x = make_double3(0, 0, 0);
  if (t < nt) {
    const int i = targets ? targets[t] : t;
    x = make_double3(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]);
  }
  double s = 0;
  for (int base = j0; base < j1; base += 256) {
    const int j = base + threadIdx.x;
    if (j < j1) tile[threadIdx.x] = make_double4(pos[3 * (size_t)j], pos[3 * (size_t)j + 1], pos[3 * (size_t)j + 2], q[j]);
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

// out_t = q_t * (sum of the chunk partials in chunk order)
__global__ void k_direct_reduce(int nt, const int *__restrict__ targets, const double *__restrict__ q,
                                const double *__restrict__ part, double *__restrict__ out)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  double s = 0;
  for (int c = 0; c < kDirectChunks; c++) s += part[(size_t)c * nt + t];
  out[t] = q[targets ? targets[t] : t] * s;
}

static int launch_direct(int N, const double *pos, const double *q, int nt, const int *targets, double *part,
                         double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_direct, dim3((nt + 255) / 256, kDirectChunks), dim3(256), 0, s, N, pos, q, nt, targets, part);
  BB_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_direct_reduce, dim3((nt + 255) / 256), dim3(256), 0, s, nt, targets, q, part, out);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
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
  BB_TRY(check_device());
  double *d_pos = nullptr, *d_q = nullptr, *d_part = nullptr, *d_out = nullptr;
  int *d_tg = nullptr;
  int rc = BBFMM_OK;
  hipError_t e = hipSuccess;
  do {
    if ((e = hipMalloc(&d_pos, sizeof(double) * 3 * (size_t)N)) != hipSuccess) break;
    if ((e = hipMalloc(&d_q, sizeof(double) * (size_t)N)) != hipSuccess) break;
    if ((e = hipMalloc(&d_tg, sizeof(int) * nt)) != hipSuccess) break;
    if ((e = hipMalloc(&d_part, sizeof(double) * (size_t)nt * kDirectChunks)) != hipSuccess) break;
    if ((e = hipMalloc(&d_out, sizeof(double) * nt)) != hipSuccess) break;
    if ((e = hipMemcpy(d_pos, pos, sizeof(double) * 3 * (size_t)N, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemcpy(d_q, q, sizeof(double) * (size_t)N, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((e = hipMemcpy(d_tg, tg.data(), sizeof(int) * nt, hipMemcpyHostToDevice)) != hipSuccess) break;
    if ((rc = launch_direct(N, d_pos, d_q, nt, d_tg, d_part, d_out, nullptr)) != BBFMM_OK) break;
    if ((e = hipMemcpy(out, d_out, sizeof(double) * nt, hipMemcpyDeviceToHost)) != hipSuccess) break;
  } while (0);
  if (e != hipSuccess) rc = set_error(BBFMM_E_HIP, "bbfmm_amd_direct: %s", hipGetErrorString(e));
  for (void *ptr : {(void *)d_pos, (void *)d_q, (void *)d_tg, (void *)d_part, (void *)d_out})
    if (ptr) (void)hipFree(ptr);
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

// ---------------------------------------------------------------- without allocation
// bbfmm_amd_direct for callers that cannot allocate device memory: the
// chunk partial sums go to a file-scope device array (each GPU the library
// runs on holds its own copy until the process ends).  Host code never
// passes the array's name to a kernel: its device address is taken once per
// device by hipGetSymbolAddress and kept in g_scratch_ptr.
namespace bb {
static __device__ __align__(256) double scratch[kDirectChunks * BBFMM_DIRECT_NOALLOC_MAX_TARGETS];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (resolved on the first
// call per device)
static int scratch_on_current_device(double **out)
{
  int dev = -1;
  BB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(BBFMM_E_ARG, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(BBFMM_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return BBFMM_OK;
}
} // namespace bb

extern "C" int bbfmm_amd_direct_noalloc(int N, const double *pos, const double *q, int nt, const int *targets,
                                        double *out, void *stream)
{
  using namespace bb;
  clear_error();
  if (N < 0 || nt < 0 || (N > 0 && (!pos || !q)) || (nt > 0 && !out))
    return set_error(BBFMM_E_ARG, "bbfmm_amd_direct_noalloc: invalid argument");
  if (nt > BBFMM_DIRECT_NOALLOC_MAX_TARGETS)
    return set_error(BBFMM_E_ARG, "bbfmm_amd_direct_noalloc: %d targets exceed the maximum of %d", nt,
                     BBFMM_DIRECT_NOALLOC_MAX_TARGETS);
  if (!targets && nt > N) return set_error(BBFMM_E_ARG, "bbfmm_amd_direct_noalloc: %d targets > N = %d", nt, N);
  if (nt == 0) return BBFMM_OK;
  BB_TRY(check_device());
  double *part = nullptr;
  BB_TRY(scratch_on_current_device(&part));
  // every partial sum the reduction reads is written by the same call, so
  // the array needs no clearing
  return launch_direct(N, pos, q, nt, targets, part, out, (hipStream_t)stream);
}
