// passes.hip -- GPU kernels of the upward and downward passes around M2L:
// charge gather, P2M, M2M, post-M2L, L2L, L2P and the result scatter.
//
// Layouts (device, per right-hand side `rhs`, rhs-major):
//   pos   [N]          real4 (x, y, z, 0) in leaf order (particles of a leaf contiguous)
//   q,phi [rhs][N]     leaf order
//   S, F  [rhs][ncell][n^3]   Chebyshev-node source / field values
//   M, L  [rhs][ncell][r]     multipole / local coefficients
// Cells are numbered as the reference numbers them (cell.h:27-28): the
// children of A are 8A+1..8A+8, so the children of a cell are contiguous.
//
// The Chebyshev order is a template parameter, instantiated for n = 3..8
// (the reference takes n at run time, main.c:58); the rank r is a run-time
// argument of every kernel here except the n = 4, r = 32 MFMA kernels.
#include "common.h"
#include "kernels.h"

#include <algorithm>

namespace bb {

// S_n(x_m, x) for m = 0..NN-1 (bbfmm.c:1641-1679, Clenshaw recurrence).
template <int NN>
__device__ __forceinline__ void cheb_row(const real *__restrict__ Tkz, real x, real *out)
{
#pragma unroll
  for (int m = 0; m < NN; m++) {
    real b1 = 0, b2 = 0;
#pragma unroll
    for (int j = NN - 1; j > 0; j--) {
      const real b0 = 2 * x * b1 - b2 + Tkz[m * NN + j];
      b2 = b1;
      b1 = b0;
    }
    out[m] = x * b1 - b2 + (real)0.5 * Tkz[m * NN];
  }
}

// The reference's L2P for n % 4 == 3 weights the last z node's value with the
// second-to-last node's interpolation weight (anotherDownwardPassX.c:304, a
// row reused in its remainder loop); with `on` the z weights are changed the
// same way, otherwise the interpolant is evaluated as it is.
template <int NN>
__device__ __forceinline__ void l2p_reference_z(real *sz, int on)
{
  if (NN % 4 == 3 && on) sz[NN - 1] = sz[NN - 2];
}

// LDS writes of a wave visible to its other lanes
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_gather(int N, const int *__restrict__ perm, const real *__restrict__ src,
                         real *__restrict__ dst)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t off = (size_t)blockIdx.y * N;
  if (i < N) dst[off + i] = src[off + perm[i]];
}

// dst (user order) from src (leaf order) through the inverse permutation:
// reads are scattered, writes coalesced.
__global__ void k_scatter(int N, const int *__restrict__ iperm, const real *__restrict__ src,
                          real *__restrict__ dst)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t off = (size_t)blockIdx.y * N;
  if (j < N) dst[off + j] = src[off + iperm[j]];
}

// Particle-row form of 16 vectors.  Four lanes per user-order particle j:
// lane part k4 reads vectors 4 k4 .. 4 k4 + 3 of its charges (each load of a
// wave covers 16 consecutive particles of 4 vectors) and writes them as one
// 16-byte piece of the particle's 64-byte row at its padded leaf-order slot,
// so every store of a wave writes 16 whole rows; the scatter reads the rows
// of potentials back the same way and writes coalesced values.  Random
// accesses are whole rows, not single words.
__global__ void k_gather_rows16(int N, const int *__restrict__ ipad, const real *__restrict__ q, int nq,
                                real *__restrict__ qt)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = (int)(t >> 2), k0 = 4 * (int)(t & 3);
  if (j >= N) return;
  real v[4];
#pragma unroll
  for (int c = 0; c < 4; c++) v[c] = k0 + c < nq ? q[(size_t)(k0 + c) * N + j] : (real)0;
  reinterpret_cast<real4 *>(qt + (size_t)ipad[j] * 16)[k0 >> 2] = real4{v[0], v[1], v[2], v[3]};
}

// phin (not null): the near field arrived in rows of its own (P2P ran
// beside the far-field passes); the scatter adds it and applies the output's
// q factor (qt, unless potential).
__global__ void k_scatter_rows16(int N, const int *__restrict__ ipad, const real *__restrict__ phit,
                                 const real *__restrict__ phin, const real *__restrict__ qt, int potential,
                                 real *__restrict__ phi, int nq)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int j = (int)(t >> 2), k0 = 4 * (int)(t & 3);
  if (j >= N || k0 >= nq) return;
  const size_t r0 = (size_t)ipad[j] * 16 + k0;
  const real4 a = *reinterpret_cast<const real4 *>(phit + r0);
  real v[4] = {a.x, a.y, a.z, a.w};
  if (phin) {
    const real4 b = *reinterpret_cast<const real4 *>(phin + r0);
    const real4 w = *reinterpret_cast<const real4 *>(qt + r0);
    const real nb[4] = {b.x, b.y, b.z, b.w}, qw[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const real s = v[c] + nb[c];
      v[c] = potential ? s : qw[c] * s;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++)
    if (k0 + c < nq) phi[(size_t)(k0 + c) * N + j] = v[c];
}

// M = U^T (Kw .* S) for one cell; sw holds Kw .* S in LDS.
template <int NN, int BLOCK>
__device__ __forceinline__ void project_u(const real *__restrict__ sw, const real *__restrict__ Urm,
                                          int r, real *__restrict__ Mout)
{
  constexpr int N3 = NN * NN * NN;
  for (int i = threadIdx.x; i < r; i += BLOCK) {
    real m = 0;
#pragma unroll 8
    for (int l = 0; l < N3; l++) m += Urm[(size_t)l * r + i] * sw[l];
    Mout[i] = m;
  }
}

// Vectors per block: NR right-hand sides share the particle geometry and
// the interpolation weights; each thread keeps NR accumulators. Rows of the
// [item][NR] staging arrays in LDS are padded to QS(NR) entries.
template <int NR>
__host__ __device__ constexpr int row_stride()
{
  return NR == 1 ? 1 : NR + 4;
}

// P2M at the leaf level (anotherUpwardPass.c:61-330): one block per leaf and
// group of NR vectors, one thread per Chebyshev node (n=8: two nodes per
// thread); S = sum_j q_j Sx Sy Sz, then M = U^T (Kw .* S).
// ROWS: the 16 charges of a particle are one row of a.qt (padded leaf order).
template <int NN, int BLOCK, int NR, bool ROWS = false>
__global__ void __launch_bounds__(BLOCK) k_p2m(PassArgs a)
{
  static_assert(!ROWS || NR == 16, "particle rows hold 16 vectors");
  constexpr int N3 = NN * NN * NN;
  constexpr int NPT = (N3 + BLOCK - 1) / BLOCK;  // nodes per thread (the last may be past N3)
  constexpr int CH = BLOCK;
  constexpr int QS = row_stride<NR>();
  __shared__ real sx[NN][CH], sy[NN][CH], sz[NN][CH];
  __shared__ __align__(16) real sq[CH * QS];
  __shared__ __align__(16) real sw[N3 * NR];
  const int leaf = blockIdx.x;
  const int rhs0 = blockIdx.y * NR;
  const int A = a.leafcell[leaf];
  const real ih = a.leafihalf[leaf];
  const int sta = a.leafsta[leaf], cnt = a.leafcnt[leaf];
  const real4 c = a.center[A];
  real acc[NPT][NR];
#pragma unroll
  for (int p = 0; p < NPT; p++)
#pragma unroll
    for (int k = 0; k < NR; k++) acc[p][k] = 0;
  for (int base = 0; base < cnt; base += CH) {
    const int nk = min(CH, cnt - base);
    if ((int)threadIdx.x < nk) {
      const real4 x = a.pos[sta + base + threadIdx.x];
      real t[NN];
      cheb_row<NN>(a.Tkz, ih * (x.x - c.x), t);
#pragma unroll
      for (int m = 0; m < NN; m++) sx[m][threadIdx.x] = t[m];
      cheb_row<NN>(a.Tkz, ih * (x.y - c.y), t);
#pragma unroll
      for (int m = 0; m < NN; m++) sy[m][threadIdx.x] = t[m];
      cheb_row<NN>(a.Tkz, ih * (x.z - c.z), t);
#pragma unroll
      for (int m = 0; m < NN; m++) sz[m][threadIdx.x] = t[m];
    }
    if (ROWS) {
      const real *qr = a.qt + (size_t)a.pad_of[sta] * 16;  // the leaf's rows are contiguous
      for (int o = threadIdx.x; o < 16 * CH; o += BLOCK) {
        const int j = o >> 4, k = o & 15;
        if (j < nk) sq[j * QS + k] = qr[(size_t)(base + j) * 16 + k];
      }
    } else {
      for (int o = threadIdx.x; o < NR * CH; o += BLOCK) {
        const int k = o / CH, j = o % CH;
        if (j < nk) sq[j * QS + k] = a.q[(size_t)(rhs0 + k) * a.N + sta + base + j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NPT; p++) {
      const int l = threadIdx.x + p * BLOCK;
      if (N3 % BLOCK != 0 && l >= N3) break;
      const int l1 = l / (NN * NN), l2 = (l / NN) % NN, l3 = l % NN;
      for (int j = 0; j < nk; j++) {
        const real w = sx[l1][j] * sy[l2][j] * sz[l3][j];
#pragma unroll
        for (int k = 0; k < NR; k++) acc[p][k] += sq[j * QS + k] * w;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < NPT; p++) {
    const int l = threadIdx.x + p * BLOCK;
    if (N3 % BLOCK != 0 && l >= N3) break;
    const real kw = a.Kw[l];
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const real v = a.prefac3 * acc[p][k];
      a.S[(size_t)(rhs0 + k) * a.ncell * N3 + (size_t)A * N3 + l] = v;
      sw[l * NR + k] = v * kw;
    }
  }
  __syncthreads();
  // M[k][i] = sum_l U[l][i] sw[l][k]; a thread per (i, group of RS vectors)
  constexpr int RS = NR >= 8 ? NR / 2 : NR, NG = NR / RS;
  for (int o = threadIdx.x; o < a.r * NG; o += BLOCK) {
    const int i = o % a.r, g = o / a.r;
    real m[RS];
#pragma unroll
    for (int k = 0; k < RS; k++) m[k] = 0;
#pragma unroll 4
    for (int l = 0; l < N3; l++) {
      const real u = a.Urm[(size_t)l * a.r + i];
#pragma unroll
      for (int k = 0; k < RS; k++) m[k] += u * sw[l * NR + g * RS + k];
    }
#pragma unroll
    for (int k = 0; k < RS; k++)
      a.M[(size_t)(rhs0 + g * RS + k) * a.ncell * a.r + (size_t)A * a.r + i] = m[k];
  }
}

#ifdef SINGLE
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// v[k] for a lane-dependent k in 0..3 without a dynamically indexed array
__device__ __forceinline__ float pick4(const float *v, int k)
{
  const float a = (k & 1) ? v[1] : v[0];
  const float b = (k & 1) ? v[3] : v[2];
  return (k & 2) ? b : a;
}

// P2M of the particle-row form, n = 4, on MFMA (v_mfma_f32_16x16x4_f32):
// one wave per leaf, four leaves per block.  For 4 particles at a time,
// lane l supplies the A value W[node 16t + l%16][particle l/16] (the
// Chebyshev weight Sx Sy Sz of that node, row tile t = x index) and the B
// value q[particle l/16][vector l%16] from the charge rows, so four MFMAs
// accumulate S (64 nodes x 16 vectors).  M = U^T (Kw .* S) follows on the
// same unit: S goes to LDS, A = U^T (lane: coefficient l%16 of a row tile,
// node l/16 of a k-step), B = Kw .* S.  Dummy rows of the padded leaf have
// zero charge and are masked anyway.
__global__ void __launch_bounds__(256) k_p2m_rows_mfma4(PassArgs a)
{
  constexpr int N3 = 64, LP = 17;  // padded LDS row of 16 vectors
  constexpr int CP = 13;           // LDS row of a particle's 12 Chebyshev values
  constexpr int CH = 64;           // particles staged per chunk
  // per wave: a chunk's Chebyshev values and charge rows, later Kw .* S
  __shared__ float wst[4][CH * (CP + LP)];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int leaf = blockIdx.x * 4 + wave;
  const bool active = leaf < a.nleaf;
  const int l16 = lane & 15, kq = lane >> 4;
  int A = 0, cnt = 0, sta = 0, psta = 0;
  if (active) {
    A = a.leafcell[leaf];
    cnt = a.leafcnt[leaf];
    sta = a.leafsta[leaf];
    psta = cnt > 0 ? a.pad_of[sta] : 0;
  }
  float *sc = wst[wave], *sq = wst[wave] + CH * CP;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (cnt > 0) {
    const real ih = a.leafihalf[leaf];
    const real4 c = a.center[A];
    for (int base = 0; base < cnt; base += CH) {
      const int nk = min(CH, cnt - base);
      // stage: lane j computes particle j's Chebyshev values; the chunk's
      // charge rows are copied with coalesced loads
      if (lane < nk) {
        const real4 x = a.pos[sta + base + lane];
        float t[4];
        cheb_row<4>(a.Tkz, ih * (x.x - c.x), t);
#pragma unroll
        for (int m = 0; m < 4; m++) sc[lane * CP + m] = t[m];
        cheb_row<4>(a.Tkz, ih * (x.y - c.y), t);
#pragma unroll
        for (int m = 0; m < 4; m++) sc[lane * CP + 4 + m] = t[m];
        cheb_row<4>(a.Tkz, ih * (x.z - c.z), t);
#pragma unroll
        for (int m = 0; m < 4; m++) sc[lane * CP + 8 + m] = t[m];
      }
      const float *qr = a.qt + (size_t)(psta + base) * 16;
      for (int o = lane; o < nk * 16; o += 64) sq[(o >> 4) * LP + (o & 15)] = qr[o];
      wave_sync();
      for (int k0 = 0; k0 < nk; k0 += 4) {
        const int p = k0 + kq;
        const bool ok = p < nk;
        // rows past the chunk hold stale LDS contents: both operands masked
        const float *cp = sc + (ok ? p : 0) * CP;
        const float syz = ok ? cp[4 + (l16 >> 2)] * cp[8 + (l16 & 3)] : 0.f;
        const float bq = ok ? sq[p * LP + l16] : 0.f;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const float wt = ok ? cp[t] * syz : 0.f;
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wt, bq, acc[t], 0, 0, 0);
        }
      }
      wave_sync();  // the chunk's staging is read before the next one (or S) overwrites it
    }
  }
  // acc[t][v] = S[node 16t + 4 kq + v][vector l16]
  float *sw = wst[wave];
  if (active) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
      f32x4 v = acc[t] * a.prefac3;
      *reinterpret_cast<f32x4 *>(a.S + ((size_t)l16 * a.ncell + A) * N3 + 16 * t + 4 * kq) = v;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int node = 16 * t + 4 * kq + e;
        sw[node * LP + l16] = v[e] * a.Kw[node];
      }
    }
  }
  __syncthreads();
  if (!active) return;
  // M[coef 16t + i][v] = sum_node U[node][coef] sw[node][v]
  f32x4 m[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll 4
  for (int kk = 0; kk < 16; kk++) {
    const int node = 4 * kk + kq;
    const float b = sw[node * LP + l16];
#pragma unroll
    for (int t = 0; t < 2; t++)
      m[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.Urm[(size_t)node * 32 + 16 * t + l16], b, m[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 2; t++)
    *reinterpret_cast<f32x4 *>(a.M + ((size_t)l16 * a.ncell + A) * 32 + 16 * t + 4 * kq) = m[t];
}

// L2P of the particle-row form, n = 4, on MFMA: one wave per leaf, four
// leaves per block.  The leaf's F (64 nodes x 16 vectors) is staged in LDS;
// per tile of 16 particles and k-step of 4 nodes, lane l supplies
// A = F^T[vector l%16][node 4kk + l/16] and B = W^T[node 4kk + l/16][particle
// l%16], so that a lane holds four consecutive vectors of one particle.  The
// far-field potentials go to the particle rows (no q factor) and vector 0's
// far field, with the output's factor, to farfield.
__global__ void __launch_bounds__(256) k_l2p_rows_mfma4(PassArgs a)
{
  constexpr int N3 = 64, LP = 17;
  __shared__ float sF[4][N3 * LP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int leaf = blockIdx.x * 4 + wave;
  if (leaf >= a.nleaf) return;  // wave-private LDS: no block barrier below
  const int l16 = lane & 15, kq = lane >> 4;
  const int cnt = a.leafcnt[leaf];
  if (cnt == 0) return;
  const int A = a.leafcell[leaf], sta = a.leafsta[leaf];
  const int psta = a.pad_of[sta];
  float *sf = sF[wave];
  // 16 vectors x 64 nodes: lane reads 16 consecutive nodes of vector v
  for (int o = lane; o < 16 * 16; o += 64) {
    const int v = o >> 4, q4 = o & 15;
    const f32x4 f = *reinterpret_cast<const f32x4 *>(a.F + ((size_t)v * a.ncell + A) * N3 + 4 * q4);
#pragma unroll
    for (int e = 0; e < 4; e++) sf[(4 * q4 + e) * LP + v] = f[e];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const real ih = a.leafihalf[leaf];
  const real4 c = a.center[A];
  for (int tb = 0; tb < cnt; tb += 16) {
    const int p = tb + l16;
    const bool ok = p < cnt;
    const real4 x = ok ? a.pos[sta + p] : c;
    float sx[4], sy[4], sz[4];
    cheb_row<4>(a.Tkz, ih * (x.x - c.x), sx);
    cheb_row<4>(a.Tkz, ih * (x.y - c.y), sy);
    cheb_row<4>(a.Tkz, ih * (x.z - c.z), sz);
    const float zq = pick4(sz, kq);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      const float w = sx[kk >> 2] * sy[kk & 3] * zq;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sf[(4 * kk + kq) * LP + l16], w, acc, 0, 0, 0);
    }
    // acc[e] = phi[particle tb + l16][vector 4 kq + e]: one 16-byte piece of
    // the particle's row per lane
    if (ok) {
      const size_t o = (size_t)(psta + p) * 16 + 4 * kq;
      const f32x4 ph = a.prefac3 * acc;
      *reinterpret_cast<f32x4 *>(a.phit + o) = ph;
      if (kq == 0) a.farfield[sta + p] = (a.potential ? 1.0f : a.qt[o]) * ph[0];
    }
  }
}

// Last level of the downward pass, fused (particle rows, n = 4, complete
// tree): one block per parent P on level maxlev - 1 and a wave per child
// leaf.  A leaf's F = scale Kw .* (U L) (post-M2L, on MFMA: A = U rows, B =
// the leaf's L) plus the L2L interpolation of P's F (separable x and y passes
// shared by the block, the z pass per wave in the accumulator layout); then
// the L2P of the leaf's particles as in k_l2p_rows_mfma4.  The leaves' F
// never goes to memory.
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6))) k_down_leaf_v16(PassArgs a, int p0, int leaf0)
{
  constexpr int NN = 4, N3 = 64, N2 = 16, R = 32, LP = 17;
  // per wave: the leaf's L in rows 0..31 ([coefficient][vector]), later its F
  // in all 64 rows; until then rows 32..63 of the waves hold P's F (waves 0,
  // 1) and the x pass (waves 2..5).  3 blocks per CU.
  __shared__ float wbuf[8][N3 * LP];
  __shared__ float fy[16][4 * N3 + 1];   // padded rows: the z pass reads across vectors
  __shared__ float cw[2 * N2];
  constexpr int UPPER = 32 * LP;
  auto fP = [&](int v) { return &wbuf[v >> 3][UPPER + (v & 7) * N3]; };
  auto fx = [&](int v) { return &wbuf[2 + (v >> 2)][UPPER + (v & 3) * 2 * N3]; };
  const int P = p0 + blockIdx.x;  // complete tree: the parents of a level are consecutive
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int ch = 8 * P + 1 + w;  // child w of P (complete tree)
  const int leaf = ch - leaf0;
  const int cnt = a.leafcnt[leaf], sta = a.leafsta[leaf];
  const real4 c = a.center[ch];
  const real4 x0 = l16 < cnt ? a.pos[sta + l16] : c;
  const real4 x1 = 16 + l16 < cnt ? a.pos[sta + 16 + l16] : c;
  if (tid < 2 * N2) cw[tid] = a.Cw[tid];
  for (int o = tid; o < 16 * N3; o += 512) fP(o / N3)[o % N3] = a.F[((size_t)(o / N3) * a.ncell + P) * N3 + o % N3];
  float *wb = wbuf[w];
  for (int o = lane; o < 16 * R; o += 64) {
    const int v = o / R, i = o % R;
    wb[i * LP + v] = a.L[((size_t)v * a.ncell + ch) * R + i];
  }
  __syncthreads();
  // U fragments of the post-M2L product (A = U rows: lane (i, k) holds
  // U[node 16 t + i][coefficient 4 kk + k]), in flight during the x and y passes
  float uc[4][R / 4];
#pragma unroll
  for (int kk = 0; kk < R / 4; kk++)
#pragma unroll
    for (int t = 0; t < 4; t++) uc[t][kk] = a.Ucm[(16 * t + l16) + (size_t)N3 * (4 * kk + kq)];
  for (int o = tid; o < 16 * 2 * N3; o += 512) {  // x
    const int v = o / (2 * N3), r = o % (2 * N3);
    const int xb = r / N3, rem = r % N3, cx = rem / N2, yz = rem % N2;
    const float *f = fP(v);
    float t = 0;
#pragma unroll
    for (int k = 0; k < NN; k++) t += f[k * N2 + yz] * cw[xb * N2 + k * NN + cx];
    fx(v)[r] = t;
  }
  __syncthreads();
  for (int o = tid; o < 16 * 4 * N3; o += 512) {  // y
    const int v = o / (4 * N3), r = o % (4 * N3);
    const int jj = r / N3, rem = r % N3, cx = rem / N2, cy = (rem / NN) % NN, pz = rem % NN;
    const float *f = fx(v);
    float t = 0;
#pragma unroll
    for (int k = 0; k < NN; k++) t += f[(jj >> 1) * N3 + cx * N2 + k * NN + pz] * cw[(jj & 1) * N2 + k * NN + cy];
    fy[v][r] = t;
  }
  __syncthreads();
  // post-M2L: acc[t][e] = (U L)[node 16t + 4kq + e][vector l16]
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < R / 4; kk++) {
    const float b = wb[(4 * kk + kq) * LP + l16];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(uc[t][kk], b, acc[t], 0, 0, 0);
  }
  // + z pass of the L2L for child w: node = 16t + 4kq + e = (l1 = 4t + kq, cz = e)
  const float sc = a.scale[a.maxlev];
  float fv[4][4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int node = 16 * t + 4 * kq + e;
      float inc = 0;
#pragma unroll
      for (int k = 0; k < NN; k++) inc += fy[l16][(w >> 1) * N3 + (4 * t + kq) * NN + k] * cw[(w & 1) * N2 + k * NN + e];
      fv[t][e] = sc * a.Kw[node] * acc[t][e] + a.prefac3 * inc;
    }
  wave_sync();  // every lane has read the leaf's L
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int e = 0; e < 4; e++) wb[(16 * t + 4 * kq + e) * LP + l16] = fv[t][e];
  wave_sync();
  // L2P of leaf ch (positions of its first two tiles loaded at the start)
  if (cnt == 0) return;
  const int psta = a.pad_of[sta];
  const real ih = a.leafihalf[leaf];
  for (int tb = 0; tb < cnt; tb += 16) {
    const int p = tb + l16;
    const bool ok = p < cnt;
    const real4 x = tb == 0 ? x0 : tb == 16 ? x1 : (ok ? a.pos[sta + p] : c);
    float sx[4], sy[4], sz[4];
    cheb_row<4>(a.Tkz, ih * (x.x - c.x), sx);
    cheb_row<4>(a.Tkz, ih * (x.y - c.y), sy);
    cheb_row<4>(a.Tkz, ih * (x.z - c.z), sz);
    const float zq = pick4(sz, kq);
    // vector 0's charge of the far-field output, loaded before the products
    const size_t o = (size_t)(psta + p) * 16 + 4 * kq;
    const float qf = ok && kq == 0 && !a.potential ? a.qt[o] : 1.0f;
    // transposed product (A = F^T, B = W^T): ph[e] = phi[particle tb + l16]
    // [vector 4 kq + e], one 16-byte piece of the particle's row per lane
    f32x4 ph = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      const float wt = sx[kk >> 2] * sy[kk & 3] * zq;
      ph = __builtin_amdgcn_mfma_f32_16x16x4f32(wb[(4 * kk + kq) * LP + l16], wt, ph, 0, 0, 0);
    }
    if (ok) {
      const f32x4 v = a.prefac3 * ph;
      *reinterpret_cast<f32x4 *>(a.phit + o) = v;
      if (kq == 0) a.farfield[sta + p] = qf * v[0];
    }
  }
}

typedef __bf16 bf16x4g __attribute__((ext_vector_type(4)));

// First upward level fused (particle rows, n = 4, complete tree): one block
// per parent P on level maxlev - 1 and a wave per child leaf.  Each wave runs
// the leaf's P2M as k_p2m_rows_mfma4 (32 particles staged at a time) and
// writes the leaf's M; the leaves' S stay in LDS, where the block forms P's S
// by the M2M interpolation (z, y, x passes over the 16 vectors) and P's M.
// The leaves' S never go to memory.
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6))) k_up_leaf_v16(PassArgs a, int p0, int leaf0)
{
  constexpr int NN = 4, N3 = 64, N2 = 16, R = 32, LP = 17, CP = 13, CH = 32;
  constexpr int WB = N3 * LP;  // per wave: staging (CH * (CP + LP) floats), then the leaf's S
  static_assert(CH * (CP + LP) <= WB, "staging fits the wave's buffer");
  __shared__ float wbuf[8][WB];
  __shared__ float szz[16][4 * N3];
  __shared__ float cw[2 * N2];
  const int P = p0 + blockIdx.x;  // complete tree: the parents of a level are consecutive
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int ch = 8 * P + 1 + w;  // child w of P (complete tree)
  const int leaf = ch - leaf0;
  const int cnt = a.leafcnt[leaf], sta = a.leafsta[leaf];
  if (tid < 2 * N2) cw[tid] = a.Cw[tid];
  float *wb = wbuf[w];
  float *sc = wb, *sq = wb + CH * CP;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (cnt > 0) {
    const int psta = a.pad_of[sta];
    const real ih = a.leafihalf[leaf];
    const real4 c = a.center[ch];
    for (int base = 0; base < cnt; base += CH) {
      const int nk = min(CH, cnt - base);
      // the chunk's positions and charge rows: every load issued before any use
      const real4 x = a.pos[sta + base + (lane < nk ? lane : 0)];
      const float *qr = a.qt + (size_t)(psta + base) * 16;
#pragma unroll
      for (int j = 0; j < CH * 16 / 64; j++) {
        const int o = lane + 64 * j;
        if (o < nk * 16) sq[(o >> 4) * LP + (o & 15)] = qr[o];
      }
      if (lane < nk) {
        float t[4];
        cheb_row<4>(a.Tkz, ih * (x.x - c.x), t);
#pragma unroll
        for (int m = 0; m < 4; m++) sc[lane * CP + m] = t[m];
        cheb_row<4>(a.Tkz, ih * (x.y - c.y), t);
#pragma unroll
        for (int m = 0; m < 4; m++) sc[lane * CP + 4 + m] = t[m];
        cheb_row<4>(a.Tkz, ih * (x.z - c.z), t);
#pragma unroll
        for (int m = 0; m < 4; m++) sc[lane * CP + 8 + m] = t[m];
      }
      wave_sync();
      for (int k0 = 0; k0 < nk; k0 += 4) {
        const int p = k0 + kq;
        const bool ok = p < nk;
        // rows past the chunk hold stale LDS contents: both operands masked
        const float *cp = sc + (ok ? p : 0) * CP;
        const float syz = ok ? cp[4 + (l16 >> 2)] * cp[8 + (l16 & 3)] : 0.f;
        const float bq = ok ? sq[p * LP + l16] : 0.f;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const float wt = ok ? cp[t] * syz : 0.f;
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wt, bq, acc[t], 0, 0, 0);
        }
      }
      wave_sync();
    }
  }
  // the leaf's S (64 nodes x 16 vectors) into the wave's buffer
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int e = 0; e < 4; e++) wb[(16 * t + 4 * kq + e) * LP + l16] = a.prefac3 * acc[t][e];
  // U fragments of the M products (A = U^T rows: lane (i, k) holds
  // U[node 4 kk + k][coefficient 16 t + i]), all loads issued at once
  float uf[2][16];
#pragma unroll
  for (int kk = 0; kk < 16; kk++)
#pragma unroll
    for (int t = 0; t < 2; t++) uf[t][kk] = a.Urm[(size_t)(4 * kk + kq) * R + 16 * t + l16];
  wave_sync();
  {  // the leaf's M = U^T (Kw .* S)
    f32x4 m[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      const int node = 4 * kk + kq;
      const float b = wb[node * LP + l16] * a.Kw[node];
#pragma unroll
      for (int t = 0; t < 2; t++) m[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(uf[t][kk], b, m[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 2; t++)
      *reinterpret_cast<f32x4 *>(a.M + ((size_t)l16 * a.ncell + ch) * R + 16 * t + 4 * kq) = m[t];
    if (a.grid_hi) {
      // and into the split-BF16 grid of the deepest level: the leaf is child
      // (sibling class) w of P; slot of class w at P's grid coordinates
      // (ghost layer of width 1), then [vector][coefficient]; hi = the
      // nearest bf16, lo = bf16 of the remainder (as the grid scatter)
      const int lp = a.maxlev - 1;  // P's level
      const int mp = P - ((1 << (3 * lp)) - 1) / 7;
      int gx = 0, gy = 0, gz = 0;
      for (int bt = 0; bt < lp; bt++) {
        const int sb = (mp >> (3 * bt)) & 7;
        gx |= ((sb >> 2) & 1) << bt;
        gy |= ((sb >> 1) & 1) << bt;
        gz |= (sb & 1) << bt;
      }
      const int PP = (1 << lp) + 2;
      const size_t slot = (size_t)a.grid_slot0 + ((w * PP + gx + 1) * PP + gy + 1) * PP + gz + 1;
#pragma unroll
      for (int t = 0; t < 2; t++) {
        bf16x4g hi, lo;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const __bf16 h = (__bf16)m[t][e];
          hi[e] = h;
          lo[e] = (__bf16)(m[t][e] - (float)h);
        }
        const size_t o = (slot * 16 + l16) * R + 16 * t + 4 * kq;
        *reinterpret_cast<bf16x4g *>((__bf16 *)a.grid_hi + o) = hi;
        *reinterpret_cast<bf16x4g *>((__bf16 *)a.grid_lo + o) = lo;
      }
    }
  }
  __syncthreads();
  // M2M of P (as k_m2m): z pass over the 8 leaves' S
  for (int o = tid; o < 16 * 4 * N3; o += 512) {
    const int v = o / (4 * N3), r = o % (4 * N3);
    const int pair = r / N3, rem = r % N3, l1 = rem / NN, pz = rem % NN;
    float t = 0;
#pragma unroll
    for (int zb = 0; zb < 2; zb++)
#pragma unroll
      for (int k = 0; k < NN; k++) t += wbuf[2 * pair + zb][(l1 * NN + k) * LP + v] * cw[zb * N2 + pz * NN + k];
    szz[v][r] = t;
  }
  __syncthreads();
  float *syy = &wbuf[0][0];       // [16][2 * N3], the leaves' S are consumed
  float *sw = &wbuf[2][0];        // [16][N3]
  for (int o = tid; o < 16 * 2 * N3; o += 512) {  // y
    const int v = o / (2 * N3), r = o % (2 * N3);
    const int xb = r / N3, rem = r % N3, cx = rem / N2, py = (rem / NN) % NN, pz = rem % NN;
    float t = 0;
#pragma unroll
    for (int yb = 0; yb < 2; yb++)
#pragma unroll
      for (int k = 0; k < NN; k++) t += szz[v][(2 * xb + yb) * N3 + cx * N2 + k * NN + pz] * cw[yb * N2 + py * NN + k];
    syy[v * 2 * N3 + r] = t;
  }
  __syncthreads();
  // U fragments of P's M product (waves 0 and 1, one 16-coefficient tile
  // each), loaded while the x pass runs
  float up[16];
  if (w < 2) {
#pragma unroll
    for (int kk = 0; kk < 16; kk++) up[kk] = a.Urm[(size_t)(4 * kk + kq) * R + 16 * w + l16];
  }
  for (int o = tid; o < 16 * N3; o += 512) {  // x
    const int v = o / N3, l = o % N3;
    const int px = l / N2, rem = l % N2;
    float t = 0;
#pragma unroll
    for (int xb = 0; xb < 2; xb++)
#pragma unroll
      for (int k = 0; k < NN; k++) t += syy[v * 2 * N3 + xb * N3 + k * N2 + rem] * cw[xb * N2 + px * NN + k];
    t *= a.prefac3;
    a.S[((size_t)v * a.ncell + P) * N3 + l] = t;
    sw[v * N3 + l] = t * a.Kw[l];
  }
  __syncthreads();
  if (w < 2) {  // P's M = U^T sw on MFMA, coefficients 16 w .. 16 w + 15
    f32x4 m = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) m = __builtin_amdgcn_mfma_f32_16x16x4f32(up[kk], sw[l16 * N3 + 4 * kk + kq], m, 0, 0, 0);
    *reinterpret_cast<f32x4 *>(a.M + ((size_t)l16 * a.ncell + P) * R + 16 * w + 4 * kq) = m;
  }
}

int launch_up_leaf(const PassArgs &a, const int *h_ploff, hipStream_t s)
{
  const int lev = a.maxlev - 1;
  const int off = h_ploff[lev], count = h_ploff[lev + 1] - off;
  if (count <= 0) return BBFMM_OK;
  hipLaunchKernelGGL(k_up_leaf_v16, dim3(count), dim3(512), 0, s, a, ((1 << (3 * lev)) - 1) / 7,
                     ((1 << (3 * a.maxlev)) - 1) / 7);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

int launch_down_leaf(const PassArgs &a, const int *h_ploff, hipStream_t s)
{
  const int lev = a.maxlev - 1;
  const int off = h_ploff[lev], count = h_ploff[lev + 1] - off;
  if (count <= 0) return BBFMM_OK;
  hipLaunchKernelGGL(k_down_leaf_v16, dim3(count), dim3(512), 0, s, a, ((1 << (3 * lev)) - 1) / 7,
                     ((1 << (3 * a.maxlev)) - 1) / 7);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}
#endif

// ---------------------------------------------------------------- 16 vectors per block, n = 4
// The per-cell passes below give each of the 16 vectors a wave of its own
// inside one block per cell: the operator tables (U, Cw) are staged in LDS
// once per block and every wave works in its private LDS region, so a cell
// costs one round trip to memory for all vectors instead of one per vector.

// M2M (as k_m2m) for one level, 16 vectors: block = parent, wave = vector.
__global__ void __launch_bounds__(1024) k_m2m_v16(PassArgs a, int off)
{
  constexpr int NN = 4, N3 = 64, N2 = 16, R = 32;
  __shared__ real sU[N3 * R];  // [node][coefficient]
  __shared__ real cw[2 * N2];
  __shared__ real wsc[16][8 * N3];
  __shared__ real wzz[16][4 * N3];
  __shared__ real wyy[16][2 * N3];
  __shared__ real wsw[16][N3];
  const int A = a.plist[off + blockIdx.x];
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  const real *S = a.S + (size_t)v * a.ncell * N3;
  real *sc = wsc[v], *szz = wzz[v], *syy = wyy[v], *sw = wsw[v];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int ch = a.child[(size_t)8 * A + j];
    sc[j * N3 + lane] = ch >= 0 ? S[(size_t)ch * N3 + lane] : (real)0;
  }
  for (int o = threadIdx.x; o < N3 * R; o += 1024) sU[o] = a.Urm[o];
  if (threadIdx.x < 2 * N2) cw[threadIdx.x] = a.Cw[threadIdx.x];
  __syncthreads();
#pragma unroll
  for (int o = lane; o < 4 * N3; o += 64) {  // z
    const int pair = o / N3, rem = o % N3, l1 = rem / NN, pz = rem % NN;
    real t = 0;
#pragma unroll
    for (int zb = 0; zb < 2; zb++)
#pragma unroll
      for (int k = 0; k < NN; k++) t += sc[(2 * pair + zb) * N3 + l1 * NN + k] * cw[zb * N2 + pz * NN + k];
    szz[o] = t;
  }
  wave_sync();
#pragma unroll
  for (int o = lane; o < 2 * N3; o += 64) {  // y
    const int xb = o / N3, rem = o % N3, cx = rem / N2, py = (rem / NN) % NN, pz = rem % NN;
    real t = 0;
#pragma unroll
    for (int yb = 0; yb < 2; yb++)
#pragma unroll
      for (int k = 0; k < NN; k++) t += szz[(2 * xb + yb) * N3 + cx * N2 + k * NN + pz] * cw[yb * N2 + py * NN + k];
    syy[o] = t;
  }
  wave_sync();
  {  // x
    const int px = lane / N2, rem = lane % N2;
    real t = 0;
#pragma unroll
    for (int xb = 0; xb < 2; xb++)
#pragma unroll
      for (int k = 0; k < NN; k++) t += syy[xb * N3 + k * N2 + rem] * cw[xb * N2 + px * NN + k];
    t *= a.prefac3;
    a.S[((size_t)v * a.ncell + A) * N3 + lane] = t;
    sw[lane] = t * a.Kw[lane];
  }
  wave_sync();
  // M = U^T sw: lanes i and i + 32 take half of the nodes each
  const int i = lane & 31, h = lane >> 5;
  real m = 0;
#pragma unroll 8
  for (int l = 32 * h; l < 32 * h + 32; l++) m += sU[l * R + i] * sw[l];
  m += __shfl_down(m, 32);
  if (h == 0) a.M[((size_t)v * a.ncell + A) * R + i] = m;
}

// post-M2L (as k_postm2l), 16 vectors: four cells per block, a wave per
// cell, a lane per node with 16 accumulators; U staged once per block.
__global__ void __launch_bounds__(256) k_postm2l_v16(PassArgs a)
{
  constexpr int N3 = 64, R = 32;
  __shared__ real sU[N3 * R];   // column-major as Ucm: [coefficient][node]
  __shared__ __align__(32) real wl[4][R * 16];  // [coefficient][vector]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int A = blockIdx.x * 4 + w;
  for (int o = threadIdx.x; o < N3 * R; o += 256) sU[o] = a.Ucm[o];
  const bool act = A < a.ncell && A < a.cell_end && a.celllev[A] >= kMinLev;
  real *sl = wl[w];
  if (act) {
#pragma unroll
    for (int o = lane; o < R * 16; o += 64) {
      const int v = o / R, i = o % R;
      sl[i * 16 + v] = a.L[((size_t)v * a.ncell + A) * R + i];
    }
  }
  __syncthreads();
  if (!act) return;
  real acc[16];
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0;
#pragma unroll 4
  for (int i = 0; i < R; i++) {
    const real u = sU[i * N3 + lane];
    const real4 *l4 = reinterpret_cast<const real4 *>(sl + i * 16);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const real4 lv = l4[q];
      acc[4 * q] += u * lv.x;
      acc[4 * q + 1] += u * lv.y;
      acc[4 * q + 2] += u * lv.z;
      acc[4 * q + 3] += u * lv.w;
    }
  }
  const real f = a.scale[a.celllev[A]] * a.Kw[lane];
#pragma unroll
  for (int v = 0; v < 16; v++) a.F[((size_t)v * a.ncell + A) * N3 + lane] = f * acc[v];
}

// L2L (as k_l2l) for one level, 16 vectors: block = parent, wave = vector.
__global__ void __launch_bounds__(1024) k_l2l_v16(PassArgs a, int off)
{
  constexpr int NN = 4, N3 = 64, N2 = 16;
  __shared__ real cw[2 * N2];
  __shared__ real wfa[16][N3];
  __shared__ real wfx[16][2 * N3];
  __shared__ real wfy[16][4 * N3];
  const int A = a.plist[off + blockIdx.x];
  const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
  real *F = a.F + (size_t)v * a.ncell * N3;
  real *fa = wfa[v], *fx = wfx[v], *fy = wfy[v];
  fa[lane] = F[(size_t)A * N3 + lane];
  if (threadIdx.x < 2 * N2) cw[threadIdx.x] = a.Cw[threadIdx.x];
  __syncthreads();
#pragma unroll
  for (int o = lane; o < 2 * N3; o += 64) {  // x
    const int xb = o / N3, rem = o % N3, cx = rem / N2, yz = rem % N2;
    real t = 0;
#pragma unroll
    for (int k = 0; k < NN; k++) t += fa[k * N2 + yz] * cw[xb * N2 + k * NN + cx];
    fx[o] = t;
  }
  wave_sync();
#pragma unroll
  for (int o = lane; o < 4 * N3; o += 64) {  // y
    const int jj = o / N3, rem = o % N3, cx = rem / N2, cy = (rem / NN) % NN, pz = rem % NN;
    real t = 0;
#pragma unroll
    for (int k = 0; k < NN; k++) t += fx[(jj >> 1) * N3 + cx * N2 + k * NN + pz] * cw[(jj & 1) * N2 + k * NN + cy];
    fy[o] = t;
  }
  wave_sync();
#pragma unroll
  for (int j = 0; j < 8; j++) {  // z, added into the children
    const int l1 = lane / NN, cz = lane % NN;
    real t = 0;
#pragma unroll
    for (int k = 0; k < NN; k++) t += fy[(j >> 1) * N3 + l1 * NN + k] * cw[(j & 1) * N2 + k * NN + cz];
    const int ch = a.child[(size_t)8 * A + j];
    if (ch >= 0) F[(size_t)ch * N3 + lane] += a.prefac3 * t;
  }
}

// M2M for one level (anotherUpwardPass.c:345-560): one block per parent,
// separable interpolation of the eight children's node values (z, y, x).
template <int NN, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_m2m(PassArgs a, int off)
{
  constexpr int N3 = NN * NN * NN, N2 = NN * NN;
  __shared__ real sc[8 * N3];
  __shared__ real szz[4 * N3];
  __shared__ real syy[2 * N3];
  __shared__ real sw[N3];
  __shared__ real cw[2 * N2];
  const int A = a.plist[off + blockIdx.x];
  const int rhs = blockIdx.y;
  real *S = a.S + (size_t)rhs * a.ncell * N3;
  for (int o = threadIdx.x; o < 8 * N3; o += BLOCK) {
    const int ch = a.child[(size_t)8 * A + o / N3];
    sc[o] = ch >= 0 ? S[(size_t)ch * N3 + o % N3] : (real)0;
  }
  for (int o = threadIdx.x; o < 2 * N2; o += BLOCK) cw[o] = a.Cw[o];
  __syncthreads();
  for (int o = threadIdx.x; o < 4 * N3; o += BLOCK) { // z
    const int pair = o / N3, rem = o % N3, l1 = rem / NN, pz = rem % NN;
    real v = 0;
#pragma unroll
    for (int zb = 0; zb < 2; zb++) {
      const real *ch = &sc[(2 * pair + zb) * N3 + l1 * NN];
      const real *w = &cw[zb * N2 + pz * NN];
#pragma unroll
      for (int k = 0; k < NN; k++) v += ch[k] * w[k];
    }
    szz[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < 2 * N3; o += BLOCK) { // y
    const int xb = o / N3, rem = o % N3, cx = rem / N2, py = (rem / NN) % NN, pz = rem % NN;
    real v = 0;
#pragma unroll
    for (int yb = 0; yb < 2; yb++) {
      const real *s = &szz[(2 * xb + yb) * N3 + cx * N2 + pz];
      const real *w = &cw[yb * N2 + py * NN];
#pragma unroll
      for (int k = 0; k < NN; k++) v += s[k * NN] * w[k];
    }
    syy[o] = v;
  }
  __syncthreads();
  for (int l = threadIdx.x; l < N3; l += BLOCK) { // x
    const int px = l / N2, rem = l % N2;
    real v = 0;
#pragma unroll
    for (int xb = 0; xb < 2; xb++) {
      const real *s = &syy[xb * N3 + rem];
      const real *w = &cw[xb * N2 + px * NN];
#pragma unroll
      for (int k = 0; k < NN; k++) v += s[k * N2] * w[k];
    }
    v *= a.prefac3;
    S[(size_t)A * N3 + l] = v;
    sw[l] = v * a.Kw[l];
  }
  __syncthreads();
  project_u<NN, BLOCK>(sw, a.Urm, a.r, a.M + (size_t)rhs * a.ncell * a.r + (size_t)A * a.r);
}

// post-M2L (auxAnotherFMMInteraction.c:398-416, scale per level as
// anotherFMMInteraction.cu:89-113): F = (1/L_l)^homogen Kw .* (U L), one
// block per cell and group of NR vectors, a thread per node.
template <int NN, int BLOCK, int NR>
__global__ void __launch_bounds__(BLOCK) k_postm2l(PassArgs a, int Afirst)
{
  constexpr int N3 = NN * NN * NN;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  real *sl = reinterpret_cast<real *>(smem_raw);  // [r][NR]
  const int A = Afirst + blockIdx.x;
  const int rhs0 = blockIdx.y * NR;
  const int level = a.celllev[A];
  if (level < kMinLev) return;  // whole block, before any barrier
  const real scale = a.scale[level];
  for (int o = threadIdx.x; o < a.r * NR; o += BLOCK) {
    const int k = o / a.r, i = o % a.r;
    sl[i * NR + k] = a.L[(size_t)(rhs0 + k) * a.ncell * a.r + (size_t)A * a.r + i];
  }
  __syncthreads();
  for (int l = threadIdx.x; l < N3; l += BLOCK) {
    real s[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) s[k] = 0;
#pragma unroll 4
    for (int i = 0; i < a.r; i++) {
      const real u = a.Ucm[l + (size_t)N3 * i];
#pragma unroll
      for (int k = 0; k < NR; k++) s[k] += u * sl[i * NR + k];
    }
    const real f = scale * a.Kw[l];
#pragma unroll
    for (int k = 0; k < NR; k++) a.F[(size_t)(rhs0 + k) * a.ncell * N3 + (size_t)A * N3 + l] = f * s[k];
  }
}

// L2L for one level (anotherDownwardPassX.c:40-211): one block per parent,
// children F += (2/n)^3 * interpolation of the parent's F (x, y, z).
template <int NN, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_l2l(PassArgs a, int off)
{
  constexpr int N3 = NN * NN * NN, N2 = NN * NN;
  __shared__ real fa[N3];
  __shared__ real fx[2 * N3];
  __shared__ real fy[4 * N3];
  __shared__ real cw[2 * N2];
  const int A = a.plist[off + blockIdx.x];
  const int rhs = blockIdx.y;
  real *F = a.F + (size_t)rhs * a.ncell * N3;
  for (int o = threadIdx.x; o < N3; o += BLOCK) fa[o] = F[(size_t)A * N3 + o];
  for (int o = threadIdx.x; o < 2 * N2; o += BLOCK) cw[o] = a.Cw[o];
  __syncthreads();
  for (int o = threadIdx.x; o < 2 * N3; o += BLOCK) { // x
    const int xb = o / N3, rem = o % N3, cx = rem / N2, yz = rem % N2;
    real v = 0;
#pragma unroll
    for (int k = 0; k < NN; k++) v += fa[k * N2 + yz] * cw[xb * N2 + k * NN + cx];
    fx[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < 4 * N3; o += BLOCK) { // y
    const int jj = o / N3, rem = o % N3, cx = rem / N2, cy = (rem / NN) % NN, pz = rem % NN;
    real v = 0;
#pragma unroll
    for (int k = 0; k < NN; k++)
      v += fx[(jj >> 1) * N3 + cx * N2 + k * NN + pz] * cw[(jj & 1) * N2 + k * NN + cy];
    fy[o] = v;
  }
  __syncthreads();
  for (int o = threadIdx.x; o < 8 * N3; o += BLOCK) { // z, add into the children
    const int j = o / N3, rem = o % N3, l1 = rem / NN, cz = rem % NN;
    real v = 0;
#pragma unroll
    for (int k = 0; k < NN; k++) v += fy[(j >> 1) * N3 + l1 * NN + k] * cw[(j & 1) * N2 + k * NN + cz];
    const int ch = a.child[(size_t)8 * A + j];
    if (ch >= 0) F[(size_t)ch * N3 + rem] += a.prefac3 * v;
  }
}

// L2P (anotherDownwardPassX.c:226-424): one thread per particle,
// phi = q (2/n)^3 sum_{l} S(x) S(y) S(z) F_l.
template <int NN>
__global__ void __launch_bounds__(256) k_l2p(PassArgs a)
{
  constexpr int N3 = NN * NN * NN;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int rhs = blockIdx.y;
  if (i >= a.N) return;
  const int lf = a.leafof[i];
  const int A = a.leafcell[lf];
  const real ih = a.leafihalf[lf];
  const real4 x = a.pos[i];
  const real4 c = a.center[A];
  real sxv[NN], syv[NN], szv[NN];
  cheb_row<NN>(a.Tkz, ih * (x.x - c.x), sxv);
  cheb_row<NN>(a.Tkz, ih * (x.y - c.y), syv);
  cheb_row<NN>(a.Tkz, ih * (x.z - c.z), szv);
  l2p_reference_z<NN>(szv, a.l2p_ref);
  const real *FA = a.F + (size_t)rhs * a.ncell * N3 + (size_t)A * N3;
  real sum = 0;
#pragma unroll
  for (int l1 = 0; l1 < NN; l1++)
#pragma unroll
    for (int l2 = 0; l2 < NN; l2++) {
      real inner = 0;
#pragma unroll
      for (int l3 = 0; l3 < NN; l3++) inner += szv[l3] * FA[(l1 * NN + l2) * NN + l3];
      sum += sxv[l1] * syv[l2] * inner;
    }
  const size_t o = (size_t)rhs * a.N + i;
  a.phi[o] = (a.potential ? (real)1 : a.q[o]) * a.prefac3 * sum;
}

// L2P for NR vectors: one block per leaf, the leaf's F values in LDS
// ([node][NR]); lanes split into NG groups of RS vectors, a lane per
// particle within a group.
// ROWS: the potentials go to rows of a.phit (padded leaf order) without the
// output's q factor, which P2P applies; vector 0's far field with that factor
// goes to a.farfield.
template <int NN, int NR, bool ROWS = false>
__global__ void __launch_bounds__(64) k_l2p_leaf(PassArgs a)
{
  static_assert(!ROWS || NR == 16, "particle rows hold 16 vectors");
  constexpr int N3 = NN * NN * NN;
  constexpr int RS = NR >= 8 ? NR / 2 : NR, NG = NR / RS, PT = 64 / NG;
  __shared__ __align__(16) real sf[N3 * NR];
  const int leaf = blockIdx.x;
  const int rhs0 = blockIdx.y * NR;
  const int A = a.leafcell[leaf];
  const real ih = a.leafihalf[leaf];
  const int sta = a.leafsta[leaf], cnt = a.leafcnt[leaf];
  for (int o = threadIdx.x; o < N3 * NR; o += 64) {
    const int k = o / N3, l = o % N3;
    sf[l * NR + k] = a.F[(size_t)(rhs0 + k) * a.ncell * N3 + (size_t)A * N3 + l];
  }
  __syncthreads();
  const real4 c = a.center[A];
  const int g = threadIdx.x / PT, jl = threadIdx.x % PT;
  for (int j = jl; j < cnt; j += PT) {
    const real4 x = a.pos[sta + j];
    real sxv[NN], syv[NN], szv[NN];
    cheb_row<NN>(a.Tkz, ih * (x.x - c.x), sxv);
    cheb_row<NN>(a.Tkz, ih * (x.y - c.y), syv);
    cheb_row<NN>(a.Tkz, ih * (x.z - c.z), szv);
    l2p_reference_z<NN>(szv, a.l2p_ref);
    real sum[RS];
#pragma unroll
    for (int k = 0; k < RS; k++) sum[k] = 0;
    // kept rolled: unrolled, the LDS reads are hoisted out of the particle
    // loop and the kernel runs out of registers
#pragma unroll 1
    for (int l1 = 0; l1 < NN; l1++)
#pragma unroll(NN > 4 ? 2 : NN)
      for (int l2 = 0; l2 < NN; l2++) {
        real inner[RS];
#pragma unroll
        for (int k = 0; k < RS; k++) inner[k] = 0;
#pragma unroll
        for (int l3 = 0; l3 < NN; l3++) {
          const real *f = &sf[((l1 * NN + l2) * NN + l3) * NR + g * RS];
#pragma unroll
          for (int k = 0; k < RS; k++) inner[k] += szv[l3] * f[k];
        }
        const real w = sxv[l1] * syv[l2];
#pragma unroll
        for (int k = 0; k < RS; k++) sum[k] += w * inner[k];
      }
    if (ROWS) {
      const size_t row = (size_t)a.pad_of[sta + j] * 16 + g * RS;
#pragma unroll
      for (int k = 0; k < RS; k++) a.phit[row + k] = a.prefac3 * sum[k];
      if (g == 0) a.farfield[sta + j] = (a.potential ? (real)1 : a.qt[row]) * a.prefac3 * sum[0];
    } else {
#pragma unroll
      for (int k = 0; k < RS; k++) {
        const size_t o = (size_t)(rhs0 + g * RS + k) * a.N + sta + j;
        a.phi[o] = (a.potential ? (real)1 : a.q[o]) * a.prefac3 * sum[k];
      }
    }
  }
}

// ---------------------------------------------------------------- launchers
int launch_gather(const int *perm, const real *src, real *dst, int N, int nrhs, hipStream_t s)
{
  if (N <= 0) return BBFMM_OK;
  hipLaunchKernelGGL(k_gather, dim3((N + 255) / 256, nrhs), dim3(256), 0, s, N, perm, src, dst);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

int launch_scatter(const int *iperm, const real *src, real *dst, int N, int nrhs, hipStream_t s)
{
  if (N <= 0) return BBFMM_OK;
  hipLaunchKernelGGL(k_scatter, dim3((N + 255) / 256, nrhs), dim3(256), 0, s, N, iperm, src, dst);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

int launch_gather_rows16(const int *ipad, const real *q, int nq, real *qt, int N, hipStream_t s)
{
  if (N <= 0) return BBFMM_OK;
  if (nq < 1 || nq > 16) return set_error(BBFMM_E_ARG, "particle rows: %d vectors (1..16)", nq);
  hipLaunchKernelGGL(k_gather_rows16, dim3((unsigned)((4LL * N + 255) / 256)), dim3(256), 0, s, N, ipad, q, nq, qt);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

int launch_scatter_rows16(const int *ipad, const real *phit, const real *phin, const real *qt, int potential,
                          real *phi, int nq, int N, hipStream_t s)
{
  if (N <= 0) return BBFMM_OK;
  if (nq < 1 || nq > 16) return set_error(BBFMM_E_ARG, "particle rows: %d vectors (1..16)", nq);
  hipLaunchKernelGGL(k_scatter_rows16, dim3((unsigned)((4LL * N + 255) / 256)), dim3(256), 0, s, N, ipad, phit, phin,
                     qt, potential, phi, nq);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

// largest group of vectors per block that divides nrhs
static int vec_group(int nrhs)
{
  return nrhs % 16 == 0 ? 16 : nrhs % 4 == 0 ? 4 : 1;
}

#ifdef SINGLE
// P2M of 32 leaves per wave (n = 4, r = 32; sparse forests, where leaves hold
// a few particles): lane l takes leaf j = l % 32 of the wave and the nodes
// 32h + m (h = l / 32, m = 0..31) and sums S over its leaf's particles
// (anotherUpwardPass.c:61-330, as k_p2m); then M = U^T (Kw .* S) on the matrix
// unit: 32 k-steps of v_mfma_f32_32x32x2_f32 with A = U^T (coefficient l % 32,
// node 32h + s) and B = (Kw .* S)[node 32h + s][leaf j], the lane's entry m = s.
__global__ void __launch_bounds__(256) k_p2m_cols(PassArgs a)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  const int leaf = (blockIdx.x * 4 + wave) * 32 + j;
  const int rhs = blockIdx.y;
  const bool valid = leaf < a.nleaf;
  float acc[32];
#pragma unroll
  for (int m = 0; m < 32; m++) acc[m] = 0.0f;
  int A = 0;
  if (valid) {
    A = a.leafcell[leaf];
    const float ih = a.leafihalf[leaf];
    const int sta = a.leafsta[leaf], cnt = a.leafcnt[leaf];
    const float4 c = a.center[A];
    const float *q = a.q + (size_t)rhs * a.N + sta;
    for (int p = 0; p < cnt; p++) {
      const float4 x = a.pos[sta + p];
      const float qv = q[p];
      float tx[4], ty[4], tz[4];
      cheb_row<4>(a.Tkz, ih * (x.x - c.x), tx);
      cheb_row<4>(a.Tkz, ih * (x.y - c.y), ty);
      cheb_row<4>(a.Tkz, ih * (x.z - c.z), tz);
      // node 32h + m: x index 2h + m / 16, y index (m / 4) % 4, z index m % 4
      const float x0 = h ? tx[2] : tx[0], x1 = h ? tx[3] : tx[1];
      float P[8];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        P[v] = qv * x0 * ty[v];
        P[4 + v] = qv * x1 * ty[v];
      }
#pragma unroll
      for (int m = 0; m < 32; m++) acc[m] += P[4 * (m / 16) + (m / 4) % 4] * tz[m % 4];
    }
    f32x4 *Sp = reinterpret_cast<f32x4 *>(a.S + ((size_t)rhs * a.ncell + A) * 64 + 32 * h);
#pragma unroll
    for (int q = 0; q < 8; q++)
      Sp[q] = f32x4{a.prefac3 * acc[4 * q], a.prefac3 * acc[4 * q + 1], a.prefac3 * acc[4 * q + 2],
                    a.prefac3 * acc[4 * q + 3]};
  }
  f32x16 mc;
#pragma unroll
  for (int e = 0; e < 16; e++) mc[e] = 0.0f;
#pragma unroll
  for (int m = 0; m < 32; m++) {
    const float u = a.Urm[(32 * h + m) * 32 + j];
    const float b = a.prefac3 * acc[m] * a.Kw[32 * h + m];
    mc = __builtin_amdgcn_mfma_f32_32x32x2f32(u, b, mc, 0, 0, 0);
  }
  if (!valid) return;
  // lane: rows 8 (e / 4) + 4 h + e % 4 of its leaf's column
  float *Mp = a.M + ((size_t)rhs * a.ncell + A) * 32 + 4 * h;
#pragma unroll
  for (int q = 0; q < 4; q++)
    *reinterpret_cast<f32x4 *>(Mp + 8 * q) = f32x4{mc[4 * q], mc[4 * q + 1], mc[4 * q + 2], mc[4 * q + 3]};
}

// post-M2L of 32 cells per wave (n = 4, r = 32; sparse forests):
// F = scale_l Kw .* (U L) (auxAnotherFMMInteraction.c:398-416, as k_postm2l)
// is U (64 nodes x 32) times the 32 cells' L columns: two row tiles of 32
// nodes, 16 k-steps of v_mfma_f32_32x32x2_f32.  Lane l: cell j = l % 32 of the
// wave, h = l / 32; A = U[32 t + j][16 h + s], B = L[cell j][16 h + s].
__global__ void __launch_bounds__(256) k_postm2l_cols(PassArgs a)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  const int A = (blockIdx.x * 4 + wave) * 32 + j;
  const int rhs = blockIdx.y;
  const int level = A < a.ncell && A < a.cell_end ? a.celllev[A] : -1;
  const bool act = level >= kMinLev;
  // k-step k pairs coefficient k (h = 0) with 16 + k (h = 1): the lane's
  // coefficients are contiguous (cells without F read cell 0, unused)
  const f32x4 *Lr = reinterpret_cast<const f32x4 *>(a.L + ((size_t)rhs * a.ncell + (act ? A : 0)) * 32 + 16 * h);
  float lv[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const f32x4 v = Lr[q];
#pragma unroll
    for (int e = 0; e < 4; e++) lv[4 * q + e] = v[e];
  }
  f32x16 d0, d1;
#pragma unroll
  for (int e = 0; e < 16; e++) d0[e] = d1[e] = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const float b = lv[k];
    const float *Uc = a.Ucm + 64 * (16 * h + k);
    d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(Uc[j], b, d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(Uc[32 + j], b, d1, 0, 0, 0);
  }
  if (!act) return;
  const float sc = a.scale[level];
  float *Fp = a.F + ((size_t)rhs * a.ncell + A) * 64 + 4 * h;
  // lane: rows 8 (e / 4) + 4 h + e % 4 of its cell's column, per row tile
  auto store = [&](const f32x16 d, int t) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int l = 32 * t + 8 * q + 4 * h;
      f32x4 o;
#pragma unroll
      for (int e = 0; e < 4; e++) o[e] = sc * a.Kw[l + e] * d[4 * q + e];
      *reinterpret_cast<f32x4 *>(Fp + 32 * t + 8 * q) = o;
    }
  };
  store(d0, 0);
  store(d1, 1);
}

// Downward pass of sparse forests, one level at a time (n = 4, r = 32): for
// the 32 cells of one child class c in a wave,
//   F = scale_l Kw .* (U L) + (2/n)^3 T_c F_parent
// (post-M2L, auxAnotherFMMInteraction.c:398-416, and L2L,
// anotherDownwardPassX.c:40-211; T_c = Cx (x) Cy (x) Cz of the class) as one
// product [Kw .* U | T_c] (64 x 96) times the columns [scale_l L ; F_parent]:
// two row tiles of 32 nodes, 16 + 32 k-steps of v_mfma_f32_32x32x2_f32 (no
// parent term at level 2).  Lane l: cell j = l % 32, h = l / 32; the parents'
// F are final since the levels run in order.
__global__ void __launch_bounds__(256) k_down_cols(PassArgs a, int w0, int w1, int level)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  const int w = w0 + blockIdx.x * 4 + wave;
  if (w >= w1) return;  // whole wave; no block-wide barrier below
  const int2 wv = a.dwv[w];
  const int c = wv.y >> 8, cnt = wv.y & 255;
  const bool valid = j < cnt;
  const int2 cp = a.dcp[wv.x + (valid ? j : 0)];
  const int rhs = blockIdx.y;
  const float sc = a.scale[level];
  // k-step k pairs coefficient k (h = 0) with 16 + k (h = 1), and parent node
  // k with 32 + k, so that a lane's operands are contiguous in memory
  const f32x4 *Lr = reinterpret_cast<const f32x4 *>(a.L + ((size_t)rhs * a.ncell + cp.x) * 32 + 16 * h);
  float lv[16];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const f32x4 v = Lr[q];
#pragma unroll
    for (int e = 0; e < 4; e++) lv[4 * q + e] = sc * v[e];
  }
  f32x16 d0, d1;
#pragma unroll
  for (int e = 0; e < 16; e++) d0[e] = d1[e] = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const float *Uc = a.KU + 64 * (16 * h + k);
    d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(Uc[j], lv[k], d0, 0, 0, 0);
    d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(Uc[32 + j], lv[k], d1, 0, 0, 0);
  }
  if (level > kMinLev) {
    const f32x4 *Fr = reinterpret_cast<const f32x4 *>(a.F + ((size_t)rhs * a.ncell + cp.y) * 64 + 32 * h);
    const float *T = a.T8 + (size_t)c * 64 * 64;
    float fv[32];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const f32x4 v = Fr[q];
#pragma unroll
      for (int e = 0; e < 4; e++) fv[4 * q + e] = v[e];
    }
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const float *Tc = T + 64 * (32 * h + k);
      d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(Tc[j], fv[k], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(Tc[32 + j], fv[k], d1, 0, 0, 0);
    }
  }
  if (!valid) return;
  float *Fp = a.F + ((size_t)rhs * a.ncell + cp.x) * 64 + 4 * h;
  // lane: rows 8 (e / 4) + 4 h + e % 4 of its cell's column, per row tile
#pragma unroll
  for (int q = 0; q < 4; q++) {
    *reinterpret_cast<f32x4 *>(Fp + 8 * q) = f32x4{d0[4 * q], d0[4 * q + 1], d0[4 * q + 2], d0[4 * q + 3]};
    *reinterpret_cast<f32x4 *>(Fp + 32 + 8 * q) = f32x4{d1[4 * q], d1[4 * q + 1], d1[4 * q + 2], d1[4 * q + 3]};
  }
}

// M2M of sparse forests, one level (n = 4, r = 32): 32 parents per wave, in
// the order of mplist (parents with the same set of children together);
// S_parent = sum_c T_c^T S_child(c) (anotherUpwardPass.c, as k_m2m) as
// products over the child classes some parent of the wave has
// (v_mfma_f32_32x32x2_f32, A = T_c^T, B = the children's S, zero for absent
// children; k-step k pairs child nodes k and 32 + k), then M = U^T (Kw .* S_parent) on the matrix unit with the
// contraction ordered as the accumulators hold the nodes: lane (j, h) owns
// nodes 32 t + 8 q + 4 h + e of parent j, so k-step (t, q, e) pairs node
// 32 t + 8 q + e (h = 0) with node 32 t + 8 q + 4 + e (h = 1).
__global__ void __launch_bounds__(256) k_m2m_cols(PassArgs a, int off, int count)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane & 31, h = lane >> 5;
  const int base = (blockIdx.x * 4 + wave) * 32;
  if (base >= count) return;  // whole wave; no block-wide barrier below
  const bool valid = base + j < count;
  const int P = a.mplist[off + (valid ? base + j : base)];
  const size_t rs = (size_t)blockIdx.y * a.ncell;
  f32x16 d0, d1;
#pragma unroll
  for (int e = 0; e < 16; e++) d0[e] = d1[e] = 0.0f;
  for (int c = 0; c < 8; c++) {
    const int ch = valid ? a.child[(size_t)8 * P + c] : -1;
    if (__ballot(ch >= 0) == 0) continue;  // no parent of the wave has this child
    // k-step k pairs child node k (h = 0) with 32 + k (h = 1): contiguous rows
    const f32x4 *Sr = reinterpret_cast<const f32x4 *>(a.S + (rs + (ch >= 0 ? ch : P)) * 64 + 32 * h);
    float sv[32];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const f32x4 v = Sr[q];
#pragma unroll
      for (int e = 0; e < 4; e++) sv[4 * q + e] = ch >= 0 ? v[e] : 0.0f;
    }
    const float *Tm = a.T8m + (size_t)c * 64 * 64;
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const float *Tc = Tm + 64 * (32 * h + k);
      d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(Tc[j], sv[k], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(Tc[32 + j], sv[k], d1, 0, 0, 0);
    }
  }
  if (valid) {
    float *Sp = a.S + (rs + P) * 64 + 4 * h;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      *reinterpret_cast<f32x4 *>(Sp + 8 * q) = f32x4{d0[4 * q], d0[4 * q + 1], d0[4 * q + 2], d0[4 * q + 3]};
      *reinterpret_cast<f32x4 *>(Sp + 32 + 8 * q) = f32x4{d1[4 * q], d1[4 * q + 1], d1[4 * q + 2], d1[4 * q + 3]};
    }
  }
  f32x16 mc;
#pragma unroll
  for (int e = 0; e < 16; e++) mc[e] = 0.0f;
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int node = 8 * q + 4 * h + e;
      mc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.Urm[node * 32 + j], a.Kw[node] * d0[4 * q + e], mc, 0, 0, 0);
      mc = __builtin_amdgcn_mfma_f32_32x32x2f32(a.Urm[(32 + node) * 32 + j], a.Kw[32 + node] * d1[4 * q + e], mc,
                                                0, 0, 0);
    }
  if (!valid) return;
  float *Mp = a.M + (rs + P) * 32 + 4 * h;
#pragma unroll
  for (int q = 0; q < 4; q++)
    *reinterpret_cast<f32x4 *>(Mp + 8 * q) = f32x4{mc[4 * q], mc[4 * q + 1], mc[4 * q + 2], mc[4 * q + 3]};
}
#endif

int launch_down_cols(const PassArgs &a, const int *h_dwv_off, hipStream_t s, int nrhs)
{
#ifdef SINGLE
  if (!a.dcp || a.n != 4 || a.r != 32) return set_error(BBFMM_E_ARG, "internal: no sparse downward tables");
  for (int level = kMinLev; level <= a.maxlev; level++) {
    const int w0 = h_dwv_off[level], w1 = h_dwv_off[level + 1];
    if (w1 <= w0) continue;
    hipLaunchKernelGGL(k_down_cols, dim3((w1 - w0 + 3) / 4, nrhs), dim3(256), 0, s, a, w0, w1, level);
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
#else
  (void)a; (void)h_dwv_off; (void)s; (void)nrhs;
  return set_error(BBFMM_E_ARG, "the sparse downward pass on MFMA needs the fp32 build");
#endif
}

template <int NN, int BLOCK>
static void p2m_t(const PassArgs &a, hipStream_t s, int nrhs)
{
  const int nr = vec_group(nrhs);
#ifdef SINGLE
  if (a.rows && NN == 4 && a.r == 32) {
    hipLaunchKernelGGL(k_p2m_rows_mfma4, dim3((a.nleaf + 3) / 4), dim3(256), 0, s, a);
    return;
  }
  if (a.cols && NN == 4 && a.r == 32) {
    hipLaunchKernelGGL(k_p2m_cols, dim3((a.nleaf + 127) / 128, nrhs), dim3(256), 0, s, a);
    return;
  }
#endif
  if (a.rows)
    hipLaunchKernelGGL((k_p2m<NN, BLOCK, 16, true>), dim3(a.nleaf, 1), dim3(BLOCK), 0, s, a);
  else if (nr == 16)
    hipLaunchKernelGGL((k_p2m<NN, BLOCK, 16>), dim3(a.nleaf, nrhs / 16), dim3(BLOCK), 0, s, a);
  else if (nr == 4)
    hipLaunchKernelGGL((k_p2m<NN, BLOCK, 4>), dim3(a.nleaf, nrhs / 4), dim3(BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL((k_p2m<NN, BLOCK, 1>), dim3(a.nleaf, nrhs), dim3(BLOCK), 0, s, a);
}

// threads per block of the per-cell pass kernels: one per Chebyshev node up
// to 256 (n = 7, 8: two nodes per thread in P2M)
__host__ __device__ constexpr int node_block(int nn)
{
  return nn <= 4 ? 64 : nn == 5 ? 128 : 256;
}

int launch_p2m(const PassArgs &a, hipStream_t s, int nrhs)
{
  if (a.nleaf <= 0) return BBFMM_OK;
  BB_WITH_ORDER(a.n, p2m_t<NN, node_block(NN)>(a, s, nrhs));
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

int launch_m2m(const PassArgs &a, const int *h_ploff, hipStream_t s, int nrhs)
{
  if (a.n < kMinOrder || a.n > kMaxOrder)
    return set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", a.n, kMinOrder, kMaxOrder);
  for (int level = std::min(a.maxlev - 1, a.m2m_start); level >= kMinLev; level--) {
    const int off = h_ploff[level], count = h_ploff[level + 1] - off;
    if (count <= 0) continue;
    if (a.n == 4 && a.r == 32 && nrhs == 16)
      hipLaunchKernelGGL(k_m2m_v16, dim3(count), dim3(1024), 0, s, a, off);
#ifdef SINGLE
    else if (a.mplist && a.T8m && a.n == 4 && a.r == 32)
      hipLaunchKernelGGL(k_m2m_cols, dim3((count + 127) / 128, nrhs), dim3(256), 0, s, a, off, count);
#endif
    else
      BB_WITH_ORDER(a.n, hipLaunchKernelGGL((k_m2m<NN, node_block(NN)>), dim3(count, nrhs), dim3(node_block(NN)),
                                            0, s, a, off));
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

// every cell; blocks of cells above level 2 return at once
template <int NN, int BLOCK>
static int postm2l_t(const PassArgs &a, hipStream_t s, int nrhs)
{
  const int first = 0, count = a.ncell;
  if (count <= 0) return BBFMM_OK;
  const int nr = vec_group(nrhs);
  const size_t sh = sizeof(real) * a.r * nr;
  if (nr == 16)
    hipLaunchKernelGGL((k_postm2l<NN, BLOCK, 16>), dim3(count, nrhs / 16), dim3(BLOCK), sh, s, a, first);
  else if (nr == 4)
    hipLaunchKernelGGL((k_postm2l<NN, BLOCK, 4>), dim3(count, nrhs / 4), dim3(BLOCK), sh, s, a, first);
  else
    hipLaunchKernelGGL((k_postm2l<NN, BLOCK, 1>), dim3(count, nrhs), dim3(BLOCK), sh, s, a, first);
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

int launch_postm2l(const PassArgs &a, hipStream_t s, int nrhs)
{
#ifdef SINGLE
  if (a.cols && a.n == 4 && a.r == 32 && nrhs != 16) {
    if (a.ncell <= 0) return BBFMM_OK;
    hipLaunchKernelGGL(k_postm2l_cols, dim3((a.ncell + 127) / 128, nrhs), dim3(256), 0, s, a);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
  }
#endif
  if (a.n == 4 && nrhs == 16 && a.r == 32) {
    if (a.ncell <= 0) return BBFMM_OK;
    hipLaunchKernelGGL(k_postm2l_v16, dim3((a.ncell + 3) / 4), dim3(256), 0, s, a);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
  }
  BB_WITH_ORDER(a.n, return postm2l_t<NN, node_block(NN)>(a, s, nrhs));
  return BBFMM_OK;
}

int launch_l2l(const PassArgs &a, const int *h_ploff, hipStream_t s, int nrhs)
{
  // L2L only (parents at levels minlev..maxlev-1)
  if (a.n < kMinOrder || a.n > kMaxOrder)
    return set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", a.n, kMinOrder, kMaxOrder);
  for (int level = kMinLev; level < std::min(a.maxlev, a.l2l_end); level++) {
    const int off = h_ploff[level], count = h_ploff[level + 1] - off;
    if (count <= 0) continue;
    if (a.n == 4 && nrhs == 16)
      hipLaunchKernelGGL(k_l2l_v16, dim3(count), dim3(1024), 0, s, a, off);
    else
      BB_WITH_ORDER(a.n, hipLaunchKernelGGL((k_l2l<NN, node_block(NN)>), dim3(count, nrhs), dim3(node_block(NN)),
                                            0, s, a, off));
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

template <int NN>
static void l2p_t(const PassArgs &a, hipStream_t s, int nrhs)
{
  const int nr = vec_group(nrhs);
#ifdef SINGLE
  if (a.rows && NN == 4) {
    hipLaunchKernelGGL(k_l2p_rows_mfma4, dim3((a.nleaf + 3) / 4), dim3(256), 0, s, a);
    return;
  }
#endif
  if (a.rows)
    hipLaunchKernelGGL((k_l2p_leaf<NN, 16, true>), dim3(a.nleaf, 1), dim3(64), 0, s, a);
  else if (nr == 16)
    hipLaunchKernelGGL((k_l2p_leaf<NN, 16>), dim3(a.nleaf, nrhs / 16), dim3(64), 0, s, a);
  else if (nr == 4)
    hipLaunchKernelGGL((k_l2p_leaf<NN, 4>), dim3(a.nleaf, nrhs / 4), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((k_l2p<NN>), dim3((a.N + 255) / 256, nrhs), dim3(256), 0, s, a);
}

int launch_l2p(const PassArgs &a, hipStream_t s, int nrhs)
{
  if (a.N <= 0) return BBFMM_OK;
  BB_WITH_ORDER(a.n, l2p_t<NN>(a, s, nrhs));
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
}

} // namespace bb
