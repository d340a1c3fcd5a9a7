// p2p.hip -- near-field P2P on the leaf level: for every particle i of leaf A,
//   phi_i += q_i * sum_{B in neighbours(A)} sum_{j in B, r_ij != 0} q_j / r_ij
// (the semantics of the reference GPU kernel direct.cu:69-209 and the CPU
// anotherNearField.c:137-206; the r != 0 test is applied in the self leaf).
//
// One 64-lane wavefront per leaf; the neighbour particles are staged in LDS
// 64 at a time and every lane (target) sweeps them (or a 1/2, 1/4 share of
// them on leaves of at most 32, 16 particles).  NR right-hand sides are
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

// A leaf of at most 16 (32) targets takes 16 (32) lanes per target and
// splits every source chunk 4 (2) ways over the lane groups (source k to
// group k mod S), the groups' sums added across lanes at the end, so that
// the wavefront's 64 lanes stay busy on leaves of a few dozen particles.
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
  const int S = cnt <= 16 ? 4 : cnt <= 32 ? 2 : 1;  // source groups (uniform in the block)
  const int T = CH / S;                              // targets per pass
  const int tl = t % T, h = t / T;
  for (int tb = 0; tb < cnt; tb += T) {
    const int i = tb + tl;
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
          int k = h;
          for (; k + S < nk; k += 2 * S) {
            const real4 s0 = sp[k], s1 = sp[k + S];
            const real dx0 = s0.x - xf.x, dy0 = s0.y - xf.y, dz0 = s0.z - xf.z;
            const real dx1 = s1.x - xf.x, dy1 = s1.y - xf.y, dz1 = s1.z - xf.z;
            const real w0 = rsqrt_r(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
            const real w1 = rsqrt_r(dx1 * dx1 + dy1 * dy1 + dz1 * dz1);
#pragma unroll
            for (int r = 0; r < NR; r++) {
              acc[r] += w0 * sq[r][k];
              acc2[r] += w1 * sq[r][k + S];
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
          for (int k = h; k < nk; k += S) {
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
    real tot[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
      tot[r] = acc[r] + acc2[r];
      if (S == 4) tot[r] += __shfl_xor(tot[r], 16);
      if (S >= 2) tot[r] += __shfl_xor(tot[r], 32);
    }
    if (act && h == 0) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const size_t o = (size_t)(rhs0 + r) * N + sta + i;
        a.phi[o] += (a.potential ? (real)1 : a.q[o]) * tot[r];
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
  for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bq, g[t], acc[t], 0, 0, 0);
}

// results of NT target tiles of the transposed product (lane l holds the
// vectors 4 (l / 16) .. 4 (l / 16) + 3 of target tb + 16 t + l % 16) added
// into phi (q * potential, or the potential), the particle rows (one
// 16-byte piece of a row per lane), or the near rows
template <int NT>
__device__ __forceinline__ void p2p16_store_t(const P2PArgs &a, int rhs0, int sta, int cnt, int psta, int tb,
                                              const f32x4 *acc)
{
  const int l = threadIdx.x;
  const int ti = l & 15, kq = l >> 4;
  if (a.phin) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int i = tb + 16 * t + ti;
      if (i < cnt) *reinterpret_cast<f32x4 *>(a.phin + (size_t)(psta + i) * 16 + 4 * kq) = acc[t];
    }
    return;
  }
  if (a.phit) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int i = tb + 16 * t + ti;
      if (i < cnt) {
        const size_t o = (size_t)(psta + i) * 16 + 4 * kq;
        const f32x4 far = *reinterpret_cast<const f32x4 *>(a.phit + o);
        const f32x4 q = a.potential ? f32x4{1.f, 1.f, 1.f, 1.f} : *reinterpret_cast<const f32x4 *>(a.qt + o);
        *reinterpret_cast<f32x4 *>(a.phit + o) = q * (far + acc[t]);
      }
    }
    return;
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int i = tb + 16 * t + ti;
    if (i < cnt)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const size_t o = (size_t)(rhs0 + 4 * kq + e) * a.N + sta + i;
        a.phi[o] += (a.potential ? 1.0f : a.q[o]) * acc[t][e];
      }
  }
}

// results of NT target tiles (accumulator layout of the 16x16 MFMA: lane l
// holds targets tb + 16 t + 4 (l / 16) + v, vector l % 16) added into phi
// (q * potential, or the potential), the particle rows, or the near rows
template <int NT>
__device__ __forceinline__ void p2p16_store(const P2PArgs &a, int rhs0, int sta, int cnt, int psta, int tb,
                                            const f32x4 *acc)
{
  const int l = threadIdx.x;
  const int ti = l & 15, kq = l >> 4;
  const float *Qt = a.qt;
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
    for (int t = 0; t < NT; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(bq, g[t], acc[t], 0, 0, 0);
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
  p2p16_store_t<NT>(a, rhs0, sta, cnt, psta, tb, acc);
}

// Sixteen vectors on MFMA; qt is [padded N][16].  Lane l computes
// G[target l%16 of each 16-target tile][source 4j + l/16] and feeds it as the
// B operand, the charges of that source for vector l%16 as the A operand
// (the transposed product: a lane ends with four consecutive vectors of one
// target, stored as one 16-byte piece of its row).  The self leaf is
// summed first with the r != 0 exclusion; the other neighbours form one
// flattened sequence of 4-source groups, loaded three groups ahead into a
// ring of four register sets.  Up to four target tiles share a pass.
template <bool BUF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(7))) k_p2p_seg16(P2PArgs a, int rhs0)
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

// ------------------------------------------------ split-BF16 accumulation
// On gfx950 the FP32 MFMA and the FP32 vector ALU of a SIMD do not run at the
// same time, so in k_p2p_seg16 the 1/r work and the MFMA accumulation (32
// cycles per 64 pairs) add up.  Here G = 1/r (fp32) and the charges are each
// split into bf16 hi + lo parts and the sum over sources takes three
// products (hi*hi + hi*lo + lo*hi) of v_mfma_f32_16x16x16_bf16: 16 targets x
// 16 sources x 16 vectors per product, fp32 accumulation.  The relative
// error of one term is below 3 * 2^-16 (the dropped lo*lo and the rounding of
// the lo parts).
#ifdef SINGLE
typedef __bf16 bf16x4p __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2p __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x3 __attribute__((ext_vector_type(3)));

// two fp32 values -> their bf16 hi parts (round to nearest even) and the
// bf16 of the remainders, each pair in one conversion
__device__ __forceinline__ void split2(float g0, float g1, unsigned &h, unsigned &lo)
{
  h = __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{g0, g1}, bf16x2p));
  const float h0 = __builtin_bit_cast(float, h << 16), h1 = __builtin_bit_cast(float, h & 0xffff0000u);
  lo = __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{g0 - h0, g1 - h1}, bf16x2p));
}

// qt [Npad][16] -> hi and lo parts at [padded source / 4][vector][4 sources]
__global__ void k_split_q16(int ngroups, const float *__restrict__ qt, bf16x4p *__restrict__ qh,
                            bf16x4p *__restrict__ ql)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (group of 4 sources, vector)
  if (i >= ngroups * 16) return;
  const int g = i >> 4, v = i & 15;
  bf16x4p h, lo;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const float x = qt[(size_t)(4 * g + e) * 16 + v];
    h[e] = (__bf16)x;
    lo[e] = (__bf16)(x - (float)h[e]);
  }
  qh[i] = h;
  ql[i] = lo;
}

typedef short i16x4 __attribute__((ext_vector_type(4)));

// One pass over all sources for NT tiles of 16 targets starting at target tb.
// A step takes 4 groups of 4 sources: lane l (target l % 16 of each tile)
// computes G for group l / 16 of the step, whose 4 sources are also its
// k-slots in the A and B operands, so no data moves between lanes.  The self
// leaf's groups are contiguous; the neighbours' groups follow the segment
// table in the order of k_p2p_seg16, walked with scalar registers, two steps
// in flight.  Groups past the end take the last group's position with zero
// charges.
template <int NT>
__device__ __forceinline__ void p2p16x_pass(const P2PArgs &a, int rhs0, int2 sg, int nseg, int T, int sta, int cnt,
                                            int psta, int pself, int tb)
{
  const int l = threadIdx.x;
  const int ti = l & 15, kq = l >> 4;
  const float4 *P = reinterpret_cast<const float4 *>(a.ppos);
  const __amdgpu_buffer_rsrc_t rP = __builtin_amdgcn_make_buffer_rsrc((void *)a.ppos, 0, a.npad * 16, 0x00020000);
  const __amdgpu_buffer_rsrc_t rH = __builtin_amdgcn_make_buffer_rsrc(a.qsh, 0, a.npad * 32, 0x00020000);
  const __amdgpu_buffer_rsrc_t rL = __builtin_amdgcn_make_buffer_rsrc(a.qsl, 0, a.npad * 32, 0x00020000);
  float4 x[NT];
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    x[t] = P[psta + min(tb + 16 * t + ti, cnt - 1)];
    acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  // the lane's group at padded position pa (x, y, z of its 4 sources and
  // their hi / lo charges of vector ti); va: it counts
  struct Src {
    f32x3 y[4];
    i16x4 h, lo;
  };
  auto load = [&](Src &u, int pa, bool va) {
#pragma unroll
    for (int e = 0; e < 4; e++)
      u.y[e] = __builtin_bit_cast(f32x3, __builtin_amdgcn_raw_buffer_load_b96(rP, (pa + e) * 16, 0, 0));
    const int oa = ((pa >> 2) * 16 + ti) * 8;
    u.h = __builtin_bit_cast(i16x4, __builtin_amdgcn_raw_buffer_load_b64(rH, oa, 0, 0));
    u.lo = __builtin_bit_cast(i16x4, __builtin_amdgcn_raw_buffer_load_b64(rL, oa, 0, 0));
    const i16x4 z = {};
    if (!va) u.h = u.lo = z;
  };
  auto step = [&](const Src &u, bool self) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
      float g[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const float dx = u.y[e].x - x[t].x, dy = u.y[e].y - x[t].y, dz = u.y[e].z - x[t].z;
        const float rr = dx * dx + dy * dy + dz * dz;
        g[e] = __builtin_amdgcn_rsqf(rr);
        if (self) g[e] = rr != 0.f ? g[e] : 0.f;
      }
      unsigned hw[2], lw[2];
      split2(g[0], g[1], hw[0], lw[0]);
      split2(g[2], g[3], hw[1], lw[1]);
      typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
      const i16x4 gh = __builtin_bit_cast(i16x4, u32x2{hw[0], hw[1]});
      const i16x4 gl = __builtin_bit_cast(i16x4, u32x2{lw[0], lw[1]});
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gh, u.h, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gh, u.lo, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gl, u.h, acc[t], 0, 0, 0);
    }
  };
  // self leaf: groups 0 .. pself/4 - 1 from psta
  const int gs = pself >> 2;
  for (int g0 = 0; g0 < gs; g0 += 4) {
    const int ga = g0 + kq;
    Src u;
    load(u, psta + 4 * min(ga, gs - 1), ga < gs);
    step(u, true);
  }
  if (T > 0) {
    int b = 1;
    int pos = __builtin_amdgcn_readlane(sg.x, 1);
    int end = pos + __builtin_amdgcn_readlane(sg.y, 1);
    // next group; stays on the last group at the end
    auto advance = [&]() {
      if (pos + 4 < end) {
        pos += 4;
      } else if (b + 1 < nseg) {
        b++;
        pos = __builtin_amdgcn_readlane(sg.x, b);
        end = pos + __builtin_amdgcn_readlane(sg.y, b);
      }
    };
    // the 4 group positions of the next step, the lane's picked out
    auto next = [&](Src &u, int g0) {
      int p[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        p[j] = pos;
        advance();
      }
      const int pa = kq == 0 ? p[0] : kq == 1 ? p[1] : kq == 2 ? p[2] : p[3];
      load(u, pa, g0 + kq < T);
    };
    Src u0, u1;
    next(u0, 0);
    for (int g0 = 0; g0 < T; g0 += 8) {
      if (g0 + 4 < T) next(u1, g0 + 4);
      step(u0, false);
      if (g0 + 4 >= T) break;
      if (g0 + 8 < T) next(u0, g0 + 8);
      step(u1, false);
    }
  }
  p2p16_store<NT>(a, rhs0, sta, cnt, psta, tb, acc);
}

// k_p2p_seg16 with the split-BF16 accumulation; qsh / qsl filled from qt.
// At most NTMAX target tiles per pass: the registers of a kernel are those of
// its largest pass (NTMAX = 3: 80 VGPRs, 6 waves per SIMD; NTMAX = 4: 100).
template <int NTMAX>
__global__ void __launch_bounds__(64) k_p2p_seg16x(P2PArgs a, int rhs0)
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
    const int tiles = min((cnt - tb + 15) >> 4, NTMAX);
    if (NTMAX >= 4 && tiles >= 4) {
      p2p16x_pass<NTMAX >= 4 ? 4 : 1>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 64;
    } else if (NTMAX >= 3 && tiles == 3) {
      p2p16x_pass<NTMAX >= 3 ? 3 : 1>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 48;
    } else if (tiles == 2) {
      p2p16x_pass<2>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 32;
    } else {
      p2p16x_pass<1>(a, rhs0, sg, nseg, T, sta, cnt, psta, pself, tb);
      tb += 16;
    }
  }
}
#endif

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
        if (a.qsh && a.qsl && a.npad > 0 && a.npad < (1 << 25)) {  // split-BF16 accumulation
          const int ng = a.npad >> 2;
          hipLaunchKernelGGL(k_split_q16, dim3((ng * 16 + 255) / 256), dim3(256), 0, s, ng, a.qt,
                             (bf16x4p *)a.qsh, (bf16x4p *)a.qsl);
          BB_HIP(hipGetLastError());
          if (a.N <= 40LL * a.nleaf)  // up to about 40 particles per leaf: passes of 1 to 3 tiles
            hipLaunchKernelGGL(k_p2p_seg16x<3>, dim3(a.nleaf), dim3(64), 0, s, a, rhs);
          else
            hipLaunchKernelGGL(k_p2p_seg16x<4>, dim3(a.nleaf), dim3(64), 0, s, a, rhs);
        } else if (!a.flat && a.npad > 0 && a.npad < (1 << 25))  // charge rows below 2 GB: scalar-offset buffer loads
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
