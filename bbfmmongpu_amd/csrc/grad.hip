// grad.hip -- adjoint of the solve (no counterpart in the reference).
//
// For the potentials phi = A(x) q of m charge vectors and a loss with
// g = dLoss/dphi, the FMM operator A is symmetric (the same interpolants
// in P2M and L2P, M2M the transpose of L2L, E_{-v} = E_v^T over the
// interaction lists), so
//     dLoss/dq_v   = A g_v                       (a solve with charges g),
//     dLoss/dx_k   = sum_v g_vk grad Phi_{q_v}(x_k) + q_vk grad Phi_{g_v}(x_k),
// where Phi_c is the FMM potential of charges c.  One solve of the 2m
// stacked vectors [q; g] recomputes all expansions; the far part of the
// position gradient is the gradient of the L2P interpolant (Chebyshev
// derivative) of the leaf's local values F, the near part the gradient of
// 1/r over the same leaf neighbourhoods as P2P.
#include "common.h"
#include "kernels.h"

namespace bb {

__device__ __forceinline__ real rsqrt_g(real x)
{
#ifdef SINGLE
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0 / sqrt(x);
#endif
}

// S_n(x_m, xi) and dS_n/dxi for m = 0..NN-1, xi in [-1, 1]:
// S = 1/2 + sum_{k>=1} T_k(x_m) T_k(xi),  dS/dxi = sum_{k>=1} T_k(x_m) k U_{k-1}(xi).
template <int NN>
__device__ __forceinline__ void cheb_val_der(const real *__restrict__ Tkz, real xi, real *s, real *ds)
{
  real T[NN], Ud[NN];
  T[0] = 1;
  Ud[0] = 1;  // U_0
  if (NN > 1) {
    T[1] = xi;
    Ud[1] = 2 * xi;  // U_1
  }
#pragma unroll
  for (int k = 2; k < NN; k++) {
    T[k] = 2 * xi * T[k - 1] - T[k - 2];
    Ud[k] = 2 * xi * Ud[k - 1] - Ud[k - 2];
  }
#pragma unroll
  for (int m = 0; m < NN; m++) {
    real v = (real)0.5 * Tkz[m * NN], d = 0;
#pragma unroll
    for (int k = 1; k < NN; k++) {
      const real c = Tkz[m * NN + k];
      v += c * T[k];
      d += c * (real)k * Ud[k - 1];
    }
    s[m] = v;
    ds[m] = d;
  }
}

// Far part: dx[c][i] = prefac3 / (L_leaf / 2) * d/dxi_c sum_l S_l(xi) Fc_l,
// Fc = sum_v g_vi F_{q_v} + q_vi F_{g_v} on the particle's leaf.  One block
// per leaf, a thread per particle.
template <int NN, int MMAX>
__global__ void __launch_bounds__(64) k_l2p_grad(GradArgs a)
{
  constexpr int N3 = NN * NN * NN;
  const int leaf = blockIdx.x;
  const int A = a.leafcell[leaf];
  const real ih = a.leafihalf[leaf];
  const int sta = a.leafsta[leaf], cnt = a.leafcnt[leaf];
  const real4 c = a.center[A];
  for (int j = threadIdx.x; j < cnt; j += 64) {
    const int i = sta + j;
    const real4 x = a.pos[i];
    real sx[NN], dsx[NN], sy[NN], dsy[NN], sz[NN], dsz[NN];
    cheb_val_der<NN>(a.Tkz, ih * (x.x - c.x), sx, dsx);
    cheb_val_der<NN>(a.Tkz, ih * (x.y - c.y), sy, dsy);
    cheb_val_der<NN>(a.Tkz, ih * (x.z - c.z), sz, dsz);
    if (NN % 4 == 3 && a.l2p_ref) {  // the reference's L2P z weights (anotherDownwardPassX.c:304)
      sz[NN - 1] = sz[NN - 2];
      dsz[NN - 1] = dsz[NN - 2];
    }
    real qk[MMAX], gk[MMAX];
#pragma unroll
    for (int v = 0; v < MMAX; v++) {
      qk[v] = v < a.mb ? a.qg[(size_t)(a.v0 + v) * a.N + i] : (real)0;
      gk[v] = v < a.mb ? a.qg[(size_t)(a.m + a.v0 + v) * a.N + i] : (real)0;
    }
    real gx = 0, gy = 0, gz = 0;
#pragma unroll 1
    for (int l1 = 0; l1 < NN; l1++)
#pragma unroll 1
      for (int l2 = 0; l2 < NN; l2++) {
        real in_s = 0, in_d = 0;
#pragma unroll
        for (int l3 = 0; l3 < NN; l3++) {
          const size_t l = (size_t)(l1 * NN + l2) * NN + l3;
          real f = 0;
#pragma unroll
          for (int v = 0; v < MMAX; v++)
            if (v < a.mb)
              f += gk[v] * a.F[((size_t)(a.v0 + v) * a.ncell + A) * N3 + l] +
                   qk[v] * a.F[((size_t)(a.m + a.v0 + v) * a.ncell + A) * N3 + l];
          in_s += sz[l3] * f;
          in_d += dsz[l3] * f;
        }
        gx += dsx[l1] * sy[l2] * in_s;
        gy += sx[l1] * dsy[l2] * in_s;
        gz += sx[l1] * sy[l2] * in_d;
      }
    const real sc = a.prefac3 * ih;
    if (a.v0 > 0) {  // a later block of vectors: add to the earlier blocks' sum
      a.dx[i] += sc * gx;
      a.dx[(size_t)a.N + i] += sc * gy;
      a.dx[(size_t)2 * a.N + i] += sc * gz;
    } else {
      a.dx[i] = sc * gx;
      a.dx[(size_t)a.N + i] = sc * gy;
      a.dx[(size_t)2 * a.N + i] = sc * gz;
    }
  }
}

// Near part, added to dx: for target k of the leaf,
//   sum_{j in neighbour leaves, r_kj != 0} (sum_v g_vk q_vj + q_vk g_vj) * (x_j - x_k) / r^3.
// One block per leaf; sources staged through LDS 64 at a time with their
// 2m charges.
template <int MMAX>
__global__ void __launch_bounds__(64) k_p2p_grad(GradArgs a)
{
  constexpr int CH = 64;
  __shared__ real4 sp[CH];
  __shared__ real sqg[CH][2 * MMAX];
  const int leaf = blockIdx.x;
  const int sta = a.leafsta[leaf], cnt = a.leafcnt[leaf];
  const int nn = a.nnbr[leaf];
  const int *nb = a.nbr + (size_t)leaf * 27;
  for (int tb = 0; tb < cnt; tb += CH) {
    const int k = tb + (int)threadIdx.x;
    const bool valid = k < cnt;
    const int i = sta + (valid ? k : cnt - 1);
    const real4 xk = a.pos[i];
    real qk[MMAX], gk[MMAX];
#pragma unroll
    for (int v = 0; v < MMAX; v++) {
      qk[v] = v < a.mb ? a.qg[(size_t)(a.v0 + v) * a.N + i] : (real)0;
      gk[v] = v < a.mb ? a.qg[(size_t)(a.m + a.v0 + v) * a.N + i] : (real)0;
    }
    real ax = 0, ay = 0, az = 0;
    // chunks of up to 64 sources, neighbour by neighbour; the next chunk is
    // loaded into registers while the current one is summed from LDS
    int b = 0, jb = 0;
    real4 rp = real4{0, 0, 0, 0};
    real rq[2 * MMAX];
    auto fetch = [&](int bb, int jj0) {  // lane's source of chunk (bb, jj0), if any
      if (bb >= nn) return;
      const int B = nb[bb];
      const int bs = a.leafsta[B], bc = a.leafcnt[B];
      if ((int)threadIdx.x < min(CH, bc - jj0)) {
        const int j = bs + jj0 + threadIdx.x;
        rp = a.pos[j];
#pragma unroll
        for (int v = 0; v < MMAX; v++) {
          rq[v] = v < a.mb ? a.qg[(size_t)(a.v0 + v) * a.N + j] : (real)0;
          rq[MMAX + v] = v < a.mb ? a.qg[(size_t)(a.m + a.v0 + v) * a.N + j] : (real)0;
        }
      }
    };
    // skip neighbours without particles (none in practice: lists hold non-empty leaves)
    while (b < nn && a.leafcnt[nb[b]] == 0) b++;
    fetch(b, 0);
    while (b < nn) {
      const int bc = a.leafcnt[nb[b]];
      const int nj = min(CH, bc - jb);
      __syncthreads();  // the previous chunk has been summed
      if ((int)threadIdx.x < nj) {
        sp[threadIdx.x] = rp;
#pragma unroll
        for (int v = 0; v < 2 * MMAX; v++) sqg[threadIdx.x][v] = rq[v];
      }
      __syncthreads();
      // the next chunk's indices, and its loads in flight during the sums
      if (jb + CH < bc) {
        jb += CH;
      } else {
        jb = 0;
        b++;
        while (b < nn && a.leafcnt[nb[b]] == 0) b++;
      }
      fetch(b, jb);
      {
        // four sources per iteration, each weight one chain of FMAs (the
        // LDS reads of the next overlap the arithmetic of the previous)
        auto pair = [&](int jj) {
          const real dx = sp[jj].x - xk.x, dy = sp[jj].y - xk.y, dz = sp[jj].z - xk.z;
          const real rr = dx * dx + dy * dy + dz * dz;
          real w = 0;
#pragma unroll
          for (int v = 0; v < MMAX; v++) {
            w += gk[v] * sqg[jj][v];
            w += qk[v] * sqg[jj][MMAX + v];
          }
          const real ri = rr != (real)0 ? rsqrt_g(rr) : (real)0;
          const real s = w * (ri * ri * ri);
          ax += s * dx;
          ay += s * dy;
          az += s * dz;
        };
        int jj = 0;
        for (; jj + 3 < nj; jj += 4) {
          pair(jj);
          pair(jj + 1);
          pair(jj + 2);
          pair(jj + 3);
        }
        for (; jj < nj; jj++) pair(jj);
      }
    }
    if (valid) {
      a.dx[i] += ax;
      a.dx[(size_t)a.N + i] += ay;
      a.dx[(size_t)2 * a.N + i] += az;
    }
  }
}

// dx (leaf order, [3][N]) -> out (original order, [N][3])
__global__ void k_scatter_vec3(int N, const int *__restrict__ iperm, const real *__restrict__ src,
                               real *__restrict__ out)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const int i = iperm[j];
  out[(size_t)3 * j] = src[i];
  out[(size_t)3 * j + 1] = src[(size_t)N + i];
  out[(size_t)3 * j + 2] = src[(size_t)2 * N + i];
}

template <int NN, int MMAX>
static void grad_t(const GradArgs &a, hipStream_t s)
{
  hipLaunchKernelGGL((k_l2p_grad<NN, MMAX>), dim3(a.nleaf), dim3(64), 0, s, a);
  hipLaunchKernelGGL((k_p2p_grad<MMAX>), dim3(a.nleaf), dim3(64), 0, s, a);
}

// one block of a.mb <= 16 vectors (a.v0 the first)
template <int NN>
static void grad_n(const GradArgs &a, hipStream_t s)
{
  if (a.mb <= 1)
    grad_t<NN, 1>(a, s);
  else if (a.mb <= 4)
    grad_t<NN, 4>(a, s);
  else if (a.mb <= 8)
    grad_t<NN, 8>(a, s);
  else
    grad_t<NN, 16>(a, s);
}

int launch_gradient(const GradArgs &a0, const int *iperm, real *dx_out, hipStream_t s)
{
  if (a0.nleaf <= 0 || a0.N <= 0) return BBFMM_OK;
  // blocks of at most 16 vectors, summed into dx in order
  for (int v0 = 0; v0 < a0.m; v0 += 16) {
    GradArgs a = a0;
    a.v0 = v0;
    a.mb = a0.m - v0 < 16 ? a0.m - v0 : 16;
    BB_WITH_ORDER(a.n, grad_n<NN>(a, s));
  }
  BB_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scatter_vec3, dim3((a0.N + 255) / 256), dim3(256), 0, s, a0.N, iperm, a0.dx, dx_out);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

} // namespace bb
