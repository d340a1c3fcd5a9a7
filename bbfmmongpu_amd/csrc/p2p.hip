// p2p.hip -- near-field P2P on the leaf level: for every particle i of leaf A,
//   phi_i += q_i * sum_{B in neighbours(A)} sum_{j in B, r_ij != 0} q_j / r_ij
// (the semantics of the reference GPU kernel direct.cu:69-209 and the CPU
// anotherNearField.c:137-206; the r != 0 test is applied in the self leaf).
//
// One 64-lane wavefront per leaf; the neighbour particles are staged in LDS
// 64 at a time and every lane (target) sweeps them.  NR right-hand sides are
// processed together so that the geometry (dx, dy, dz, 1/r) is computed once
// per pair and shared by NR charge vectors.
// 1/r: fp32 uses the hardware v_rsq_f32 (AMD's own reciprocal square root,
// about 1 ulp); fp64 uses the device math library's rsqrt.
#include "common.h"
#include "kernels.h"

namespace bb {

__device__ __forceinline__ real rsqrt_r(real x)
{
#ifdef SINGLE
  return __builtin_amdgcn_rsqf(x);
#else
  return rsqrt(x);
#endif
}

template <int NR>
__global__ void __launch_bounds__(64) k_p2p(P2PArgs a, int rhs0)
{
  constexpr int CH = 64;
  __shared__ real4 sp[CH];
  __shared__ real sq[NR][CH];
  __shared__ int nsta[27], ncnt[27], nself[27];
  const int leaf = blockIdx.x;
  const int t = threadIdx.x;
  const int sta = a.leafsta[leaf], cnt = a.leafcnt[leaf];
  const int nn = a.nnbr[leaf];
  if (t < nn) {
    const int B = a.nbr[(size_t)leaf * 27 + t];
    nsta[t] = a.leafsta[B];
    ncnt[t] = a.leafcnt[B];
    nself[t] = (B == leaf);
  }
  __syncthreads();
  const size_t N = a.N;
  for (int tb = 0; tb < cnt; tb += CH) {
    const int i = tb + t;
    const bool act = i < cnt;
    real4 xf;
    if (act) {
      xf = a.pos[sta + i];
    } else {
      xf.x = xf.y = xf.z = xf.w = 0;
    }
    real acc[NR], acc2[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r] = acc2[r] = 0;
    for (int b = 0; b < nn; b++) {
      const int jsta = nsta[b], nj = ncnt[b];
      const bool self = nself[b] != 0;
      for (int jb = 0; jb < nj; jb += CH) {
        const int nk = min(CH, nj - jb);
        if (t < nk) {
          sp[t] = a.pos[jsta + jb + t];
#pragma unroll
          for (int r = 0; r < NR; r++) sq[r][t] = a.q[(size_t)(rhs0 + r) * N + jsta + jb + t];
        }
        __syncthreads();
        if (!self) {
          int k = 0;
          for (; k + 1 < nk; k += 2) {
            const real4 s0 = sp[k], s1 = sp[k + 1];
            const real dx0 = s0.x - xf.x, dy0 = s0.y - xf.y, dz0 = s0.z - xf.z;
            const real dx1 = s1.x - xf.x, dy1 = s1.y - xf.y, dz1 = s1.z - xf.z;
            const real w0 = rsqrt_r(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
            const real w1 = rsqrt_r(dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
#pragma unroll
            for (int r = 0; r < NR; r++) {
              acc[r] += w0 * sq[r][k];
              acc2[r] += w1 * sq[r][k + 1];
            }
          }
          if (k < nk) {
            const real4 s0 = sp[k];
            const real dx0 = s0.x - xf.x, dy0 = s0.y - xf.y, dz0 = s0.z - xf.z;
            const real w0 = rsqrt_r(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
#pragma unroll
            for (int r = 0; r < NR; r++) acc[r] += w0 * sq[r][k];
          }
        } else {
          for (int k = 0; k < nk; k++) {
            const real4 s0 = sp[k];
            const real dx0 = s0.x - xf.x, dy0 = s0.y - xf.y, dz0 = s0.z - xf.z;
            const real rr = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
            const real w0 = rr != (real)0 ? rsqrt_r(rr) : (real)0;
#pragma unroll
            for (int r = 0; r < NR; r++) acc[r] += w0 * sq[r][k];
          }
        }
        __syncthreads();
      }
    }
    if (act) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const size_t o = (size_t)(rhs0 + r) * N + sta + i;
        a.phi[o] += (a.potential ? (real)1 : a.q[o]) * (acc[r] + acc2[r]);
      }
    }
  }
}

// Sparse forests, where leaves hold a few particles: a lane per target
// particle, sweeping its own leaf's neighbour list (the neighbour leaves'
// particles straight from memory; lanes of one leaf read the same source at
// the same time).  Same sums as above; the r != 0 test is applied to every
// neighbour leaf, which changes nothing since coincident particles share a
// leaf.
template <int NR>
__global__ void __launch_bounds__(256) k_p2p_lane(P2PArgs a, int rhs0)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.N) return;
  const int lf = a.leafof[i];
  const real4 x = a.pos[i];
  const int nn = a.nnbr[lf];
  const int *nb = a.nbr + (size_t)lf * 27;
  const size_t N = a.N;
  real acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = 0;
  // the next neighbour's range is fetched while this one's sources are summed
  int Bn = nb[0];  // nn >= 1: the leaf itself
  int j0n = a.leafsta[Bn], cn = a.leafcnt[Bn];
  for (int b = 0; b < nn; b++) {
    const int j0 = j0n, j1 = j0 + cn;
    if (b + 1 < nn) {
      Bn = nb[b + 1];
      j0n = a.leafsta[Bn];
      cn = a.leafcnt[Bn];
    }
    for (int j = j0; j < j1; j++) {
      const real4 y = a.pos[j];
      const real dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
      const real rr = dx * dx + dy * dy + dz * dz;
      const real w = rr != (real)0 ? rsqrt_r(rr) : (real)0;
#pragma unroll
      for (int r = 0; r < NR; r++) acc[r] += w * a.q[(size_t)(rhs0 + r) * N + j];
    }
  }
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const size_t o = (size_t)(rhs0 + r) * N + i;
    a.phi[o] += (a.potential ? (real)1 : a.q[o]) * acc[r];
  }
}

// The same for one vector, the sources read as (x, y, z, q) in one load.
__global__ void __launch_bounds__(256) k_p2p_lane_pq(P2PArgs a, int rhs)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.N) return;
  const int lf = a.leafof[i];
  const real4 x = a.pq[i];
  const int nn = a.nnbr[lf];
  const int *nb = a.nbr + (size_t)lf * 27;
  real acc = 0;
  // the next neighbour's range is fetched while this one's sources are summed
  int Bn = nb[0];  // nn >= 1: the leaf itself
  int j0n = a.leafsta[Bn], cn = a.leafcnt[Bn];
  for (int b = 0; b < nn; b++) {
    const int j0 = j0n, j1 = j0 + cn;
    if (b + 1 < nn) {
      Bn = nb[b + 1];
      j0n = a.leafsta[Bn];
      cn = a.leafcnt[Bn];
    }
    for (int j = j0; j < j1; j++) {
      const real4 y = a.pq[j];
      const real dx = y.x - x.x, dy = y.y - x.y, dz = y.z - x.z;
      const real rr = dx * dx + dy * dy + dz * dz;
      const real w = rr != (real)0 ? rsqrt_r(rr) : (real)0;
      acc += w * y.w;
    }
  }
  const size_t o = (size_t)rhs * a.N + i;
  a.phi[o] += (a.potential ? (real)1 : x.w) * acc;
}

// pq[i] = (pos[i].xyz, q[rhs][i])
__global__ void k_pack_pq(int N, const real4 *__restrict__ pos, const real *__restrict__ q, real4 *__restrict__ pq)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  real4 v = pos[i];
  v.w = q[i];
  pq[i] = v;
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2):
// give XCD x the x-th contiguous eighth of the leaves (Morton order), so that
// neighbouring leaves, which read the same sources, share an L2.
__device__ __forceinline__ int xcd_leaf(int b, int nleaf)
{
  if (nleaf & 7) return b;
  return (b & 7) * (nleaf >> 3) + (b >> 3);
}

#ifdef SINGLE
typedef float f32x4 __attribute__((ext_vector_type(4)));

#endif

// ---------------------------------------------------------------- padded layout
// The kernel below reads a private copy of the particles in which every
// leaf's range starts at a multiple of 4 and is padded to a multiple of 4
// with dummy sources far outside the root cell and of zero charge, so that
// its source loops need no tail masks.  seg[leaf][0..nseg) lists the
// (padded start, padded count) of the leaf's non-empty neighbours, the leaf
// itself first.  Targets are read from the same copy; dummy targets are
// never written.

#ifdef SINGLE
template <int NT>
__device__ __forceinline__ void p2p16_step(const float4 &z, float bq, const float4 *x, f32x4 *acc)
{
  float g[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const float dx = z.x - x[t].x, dy = z.y - x[t].y, dz = z.z - x[t].z;
    g[t] = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
  }
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(g[t], bq, acc[t], 0, 0, 0);
}

// One pass over all sources for NT tiles of 16 targets starting at target
// tb of the leaf; results added into phi (q * potential, or the potential).
// BUF: the source positions and charge rows are read with buffer loads whose
// group offset is a scalar register (the padded arrays then stay below 2 GB),
// so no vector instructions go into the per-group addresses.
template <int NT, bool BUF>
__device__ __forceinline__ void p2p16_pass(const P2PArgs &a, int rhs0, int2 sg, int nseg, int T, int sta,
                                           int cnt, int psta, int pself, int tb)
{
  const int l = threadIdx.x;
  const int ti = l & 15, kq = l >> 4;
  const float4 *P = reinterpret_cast<const float4 *>(a.ppos);
  const float *Qt = a.qt;
  const int qoff = kq * 16 + ti;
  const __amdgpu_buffer_rsrc_t rP = __builtin_amdgcn_make_buffer_rsrc((void *)a.ppos, 0, a.npad * 16, 0x00020000);
  const __amdgpu_buffer_rsrc_t rQ = __builtin_amdgcn_make_buffer_rsrc((void *)a.qt, 0, a.npad * 64, 0x00020000);
  // source group at padded index p: lane l reads the position of source p + l / 16 and its charge of vector l % 16
  auto src_pos = [&](int p) -> float4 {
    if (BUF) return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rP, kq * 16, p * 16, 0));
    return P[p + kq];
  };
  auto src_q = [&](int p) -> float {
    if (BUF) return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rQ, qoff * 4, p * 64, 0));
    return Qt[(size_t)p * 16 + qoff];
  };
  float4 x[NT];
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    x[t] = P[psta + min(tb + 16 * t + ti, cnt - 1)];
    acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  // the self leaf (r != 0 exclusion), loaded one group ahead
  auto self_step = [&](const float4 &z, float bq) {
    float g[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const float dx = z.x - x[t].x, dy = z.y - x[t].y, dz = z.z - x[t].z;
      const float rr = dx * dx + dy * dy + dz * dz;
      g[t] = rr != 0.f ? __builtin_amdgcn_rsqf(rr) : 0.f;
    }
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(g[t], bq, acc[t], 0, 0, 0);
  };
  if (pself > 0) {
    float4 zA = src_pos(psta);
    float qA = src_q(psta);
    for (int k = 0; k < pself; k += 8) {
      const int k1 = min(k + 4, pself - 4);
      const float4 zB = src_pos(psta + k1);
      const float qB = src_q(psta + k1);
      self_step(zA, qA);
      if (k + 4 >= pself) break;
      const int k2 = min(k + 8, pself - 4);
      zA = src_pos(psta + k2);
      qA = src_q(psta + k2);
      self_step(zB, qB);
    }
  }
  if (T > 0) {
    int b = 1;
    int pos = __builtin_amdgcn_readlane(sg.x, 1);
    int end = pos + __builtin_amdgcn_readlane(sg.y, 1);
    // next group; stays on the last group at the end (loaded, never summed)
    auto advance = [&]() {
      if (pos + 4 < end) {
        pos += 4;
      } else if (b + 1 < nseg) {
        b++;
        pos = __builtin_amdgcn_readlane(sg.x, b);
        end = pos + __builtin_amdgcn_readlane(sg.y, b);
      }
    };
    // four groups in flight: a ring of four register sets, loaded three
    // groups ahead (the loop is unrolled so that no set is copied)
    float4 y0, y1, y2, y3;
    float q0, q1, q2, q3;
    auto load = [&](float4 &y, float &q) {
      y = src_pos(pos);
      q = src_q(pos);
      advance();
    };
    load(y0, q0);
    load(y1, q1);
    load(y2, q2);
    for (int t = 0; t < T; t += 4) {
      load(y3, q3);
      p2p16_step<NT>(y0, q0, x, acc);
      if (t + 1 >= T) break;
      load(y0, q0);
      p2p16_step<NT>(y1, q1, x, acc);
      if (t + 2 >= T) break;
      load(y1, q1);
      p2p16_step<NT>(y2, q2, x, acc);
      if (t + 3 >= T) break;
      load(y2, q2);
      p2p16_step<NT>(y3, q3, x, acc);
    }
  }
  if (a.phin) {
    // particle rows of their own: the near-field potential
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = tb + 16 * t + 4 * kq + v;
        if (i < cnt) a.phin[(size_t)(psta + i) * 16 + ti] = acc[t][v];
      }
    return;
  }
  if (a.phit) {
    // particle rows: far field + near field, then the output's q factor
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = tb + 16 * t + 4 * kq + v;
        const size_t o = (size_t)(psta + i) * 16 + ti;
        if (i < cnt) a.phit[o] = (a.potential ? 1.0f : Qt[o]) * (a.phit[o] + acc[t][v]);
      }
    return;
  }
  const size_t row = (size_t)(rhs0 + ti) * a.N + sta;
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = tb + 16 * t + 4 * kq + v;
      if (i < cnt) a.phi[row + i] += (a.potential ? 1.0f : a.q[row + i]) * acc[t][v];
    }
}

// Sixteen vectors on MFMA; qt is [padded N][16].  Lane l computes
// G[target l%16 of each 16-target tile][source 4j + l/16] and feeds the
// charges of that source for vector l%16 as the B operand.  The self leaf is
// summed first with the r != 0 exclusion; the other neighbours form one
// flattened sequence of 4-source groups, loaded three groups ahead into a
// ring of four register sets.  Up to four target tiles share a pass.
template <bool BUF>
__global__ void __launch_bounds__(64) k_p2p_seg16(P2PArgs a, int rhs0)
{
  const int leaf = xcd_leaf(blockIdx.x, a.nleaf);
  const int l = threadIdx.x;
  const int sta = a.leafsta[leaf], cnt = a.leafcnt[leaf];
  if (cnt == 0) return;
  const int nseg = a.nseg[leaf];
  const int2 sg = l < nseg ? a.seg[(size_t)leaf * 27 + l] : make_int2(0, 0);
  const int psta = __builtin_amdgcn_readfirstlane(sg.x);
  const int pself = __builtin_amdgcn_readfirstlane(sg.y);
  int T = 0;
  for (int b = 1; b < nseg; b++) T += __builtin_amdgcn_readlane(sg.y, b) >> 2;
  for (int tb = 0; tb < cnt;) {
    const int tiles = (cnt - tb + 15) >> 4;
    if (tiles >= 4) {
      p2p16_pass<4, BUF>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 64;
    } else if (tiles == 3) {
      p2p16_pass<3, BUF>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 48;
    } else if (tiles == 2) {
      p2p16_pass<2, BUF>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 32;
    } else {
      p2p16_pass<1, BUF>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 16;
    }
  }
}

// charges of vectors rhs0..rhs0+15, leaf order [16][N] -> padded [Npad][16]
__global__ void k_pad_charges16(int N, const int *__restrict__ pad_of, const float *__restrict__ q,
                                float *__restrict__ qt)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float4 *o = reinterpret_cast<float4 *>(qt + (size_t)pad_of[i] * 16);
#pragma unroll
  for (int r = 0; r < 16; r += 4)
    o[r / 4] = make_float4(q[(size_t)r * N + i], q[(size_t)(r + 1) * N + i], q[(size_t)(r + 2) * N + i],
                           q[(size_t)(r + 3) * N + i]);
}
#endif

int launch_p2p(const P2PArgs &a, hipStream_t s)
{
  if (a.nleaf <= 0 || a.N <= 0) return BBFMM_OK;
  int rhs = 0;
  if (a.leafof && !a.phit && !a.phin) {
    const int nb = (a.N + 255) / 256;
    while (rhs < a.nrhs) {
      const int left = a.nrhs - rhs;
      if (left >= 4) {
        hipLaunchKernelGGL(k_p2p_lane<4>, dim3(nb), dim3(256), 0, s, a, rhs);
        rhs += 4;
      } else if (a.pq) {
        hipLaunchKernelGGL(k_pack_pq, dim3(nb), dim3(256), 0, s, a.N, a.pos, a.q + (size_t)rhs * a.N, a.pq);
        BB_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_p2p_lane_pq, dim3(nb), dim3(256), 0, s, a, rhs);
        rhs += 1;
      } else {
        hipLaunchKernelGGL(k_p2p_lane<1>, dim3(nb), dim3(256), 0, s, a, rhs);
        rhs += 1;
      }
      BB_HIP(hipGetLastError());
    }
    return BBFMM_OK;
  }
  while (rhs < a.nrhs) {
    const int left = a.nrhs - rhs;
    if (left >= 16) {
#ifdef SINGLE
      if (a.use_mfma && a.qt && a.ppos && a.seg) {
        if (!a.phit && !a.phin) {  // particle rows arrive filled
          hipLaunchKernelGGL(k_pad_charges16, dim3((a.N + 255) / 256), dim3(256), 0, s, a.N, a.pad_of,
                             a.q + (size_t)rhs * a.N, a.qt);
          BB_HIP(hipGetLastError());
        }
        if (a.npad > 0 && a.npad < (1 << 25))  // charge rows below 2 GB: scalar-offset buffer loads
          hipLaunchKernelGGL(k_p2p_seg16<true>, dim3(a.nleaf), dim3(64), 0, s, a, rhs);
        else
          hipLaunchKernelGGL(k_p2p_seg16<false>, dim3(a.nleaf), dim3(64), 0, s, a, rhs);
      } else
#endif
        hipLaunchKernelGGL(k_p2p<16>, dim3(a.nleaf), dim3(64), 0, s, a, rhs);
      rhs += 16;
    } else if (left >= 4) {
      hipLaunchKernelGGL(k_p2p<4>, dim3(a.nleaf), dim3(64), 0, s, a, rhs);
      rhs += 4;
    } else {
      hipLaunchKernelGGL(k_p2p<1>, dim3(a.nleaf), dim3(64), 0, s, a, rhs);
      rhs += 1;
    }
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

} // namespace bb
