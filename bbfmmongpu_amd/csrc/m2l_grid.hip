// m2l_grid.hip -- dense-grid M2L (schemes CL and IJ).
//
// At level l (2^l cells per axis) a field cell f = 2p + s has parent p and
// sibling class s in {0,1}^3.  Its interaction list is every cell c = f + v
// whose parent is a neighbour of p (|p(c) - p|_inf <= 1) and which is not a
// neighbour of f (|v|_inf >= 2) -- 189 offsets per class, with
//   source class  s' = (s + v) mod 2,   parent shift  d = floor((s + v) / 2),
// so the sources of all class-s cells for one offset v form the class-s'
// parent grid shifted by d.  M is scattered into eight per-class parent grids
// with one ghost layer of zeros on every side; then for every class s and
// offset v
//     L_s[:, p] += K_v M_{s'}[:, p + d]      (an r x r by r x ncols product)
// which is a GEMM per (s, v) accumulated in registers -- the natural MFMA
// form.  Cells at the domain boundary read ghost zeros, so the result equals
// the list-driven M2L (anotherInteractionList.c:11-111).
//
// Columns are (parent position, right-hand side) pairs, so many charge
// vectors widen every product.  One launch covers all levels.
//
// Forms:
//   PLAIN   real-precision FMAs on the vector ALUs (drop-in parity);
//   BF16X2  K and M split into bf16 hi + lo, three MFMA products
//           (hi*hi + hi*lo + lo*hi) accumulated in fp32 (fp32 build);
//   FP32    v_mfma_f32_32x32x2_f32 (fp32 build);
//   FP16X3  K and M each split into three fp16 terms t1 + t2 + t3 (about 33
//           bits), six MFMA products (t1t1, t1t2, t2t1, t1t3, t2t2, t3t1)
//           accumulated in fp32 (fp32 build).  fp16 has a narrow exponent
//           range, so K is scaled by a power of two chosen from its largest
//           entry and the M of every level by one chosen from that level's
//           largest entry (measured on the device each solve); all sources
//           of a field cell lie on its level, so the product is unscaled
//           exactly in the epilogue.
// MFMA forms use r = 32 (one 32 x 32 output tile per 32 columns); FP32 also
// r = 256.  PLAIN takes any rank 1..512 (r = 32 and 256 with the rank fixed
// at compile time).
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace bb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kOffsets = 189;
constexpr int kMaxLev = 10;

struct GridArgs {
  int r, nrhs, ncell, minlev, maxlev;
  int blk0[kMaxLev + 2];      // first block of each level (flattened launch)
  int lvl_slot0[kMaxLev + 1];  // first slot of each level's 8 class grids
  int levsta[kMaxLev + 2];
  const int4 *tab;  // [8][189]: (k, source class, packed shift, 0)
  const void *Ka;   // operators in the form's layout
  const void *Mg0;  // grids (hi / real)
  const void *Mg1;  // grids (lo) for BF16X2, (t2) for FP16X3
  const void *Mg2;  // grids (t3) for FP16X3
  unsigned *lvl_max;  // FP16X3: [kMaxLev + 1] bits of each level's largest |M|
  float kscale_inv;   // FP16X3: 1 / the power of two K is scaled by
  real *L;
  int wave_cols;    // columns per wave
  const int *steps;  // [class group][208][kStepWords] step records (step_table)
  int nsteps[8];     // steps of each class group
  int ss;            // source-stationary kernel: field classes per wave (8, 4 or 2); 0: not used
  int ss_tpw;        // source-stationary kernel: 32-column tiles per wave (1 or 2)
  int ss_pair;       // 1: both class groups of a column group in one block sharing source tiles
                     // (k_m2l_mfma_ss_pair: 4 classes, 2 tiles per wave, the 8-class step table)
  int gbytes;        // bytes of one grid array when below 2 GB (scalar-offset buffer loads), else 0
  int kbytes;        // bytes of the operator array
  // source-stationary kernel: each class group's step list in nsplit parts
  // (one block each); part 0 stores into L, part k >= 1 into Lx
  // [k - 1][rhs][up_ncell][32] (cells up_cell0 ...), added to L afterwards
  int nsplit;
  int up_cell0, up_ncell;
  float *Lx;
};

constexpr int kSteps = 208;  // 8 source classes x 26 non-zero parent shifts
constexpr int kStepWords = 8;  // words per step record (see step_table)
constexpr int kZeroOp = kNumOps;  // all-zero operator slot of the split-BF16 layout
// s_waitcnt immediate (gfx9 encoding) for vmcnt(0), expcnt and lgkmcnt not waited on
constexpr int kWaitVm0 = 0x0F70;

struct GridM2L {
  int r = 0, maxlev = 0, nrhs = 1, form = 0, ncell_cap = 0, minlev = kMinLev, variant = 0;
  int flat = 0;  // 1: pointer-form loads even on grids below 2 GB
  // split-BF16 source-stationary kernel with a class-group size of its own
  // for the levels above the deepest: part[0] those levels, part[1] the
  // deepest (one launch each); nparts = 1: one launch, args
  int nparts = 1;
  GridArgs part[2];
  int part_blocks[2] = {0, 0};
  float *d_Lx = nullptr;  // partial sums of the upper part's steps beyond the first part
  GridArgs args{};
  int nblocks = 0;
  size_t nslots = 0;
  int4 *d_tab = nullptr;
  int *d_steps = nullptr;
  void *d_K = nullptr;
  void *d_Mg0 = nullptr;
  void *d_Mg1 = nullptr;
  void *d_Mg2 = nullptr;
  unsigned *d_lvl_max = nullptr;
};

__device__ __forceinline__ int morton3(int x, int y, int z, int bits)
{
  int m = 0;
  for (int b = 0; b < bits; b++)
    m |= ((((x >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((z >> b) & 1)) << (3 * b);
  return m;
}

__device__ __forceinline__ void demorton3(int m, int bits, int &x, int &y, int &z)
{
  x = y = z = 0;
  for (int b = 0; b < bits; b++) {
    const int s = (m >> (3 * b)) & 7;
    x |= ((s >> 2) & 1) << b;
    y |= ((s >> 1) & 1) << b;
    z |= (s & 1) << b;
  }
}

// slot of (level, class, parent p) in the ghost-padded grids
__device__ __forceinline__ int grid_slot(const GridArgs &g, int l, int c, int px, int py, int pz)
{
  const int PP = (1 << (l - 1)) + 2;
  return g.lvl_slot0[l] + ((c * PP + px + 1) * PP + py + 1) * PP + pz + 1;
}

// ---------------------------------------------------------------- FP16X3 scaling
// The power of two that maps a largest magnitude m (bits of a non-negative
// float) into [2^14, 2^15): 2^(15 - e) with m = f 2^e, f in [0.5, 1).
__device__ __forceinline__ int fp16_scale_exp(unsigned mbits)
{
  int e = 0;
  (void)frexpf(__uint_as_float(mbits), &e);
  return 15 - e;
}

__device__ __forceinline__ void split3_f16(float v, _Float16 &t1, _Float16 &t2, _Float16 &t3)
{
  t1 = (_Float16)v;
  const float r1 = v - (float)t1;
  t2 = (_Float16)r1;
  t3 = (_Float16)(r1 - (float)t2);
}

__global__ void k_zero_u32(unsigned *p, int n)
{
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// Largest |M| of every level minlev..maxlev: blockIdx.y = level - minlev; a
// block reduces a grid-strided slice of the level's [rhs][cells][r] values
// and posts one atomicMax on the bits (order-preserving for non-negative
// floats).
__global__ void __launch_bounds__(256) k_level_absmax(GridArgs g, const real *__restrict__ M)
{
  __shared__ unsigned red[4];
  const int l = g.minlev + blockIdx.y;
  const int c0 = g.levsta[l], nc = g.levsta[l + 1] - c0;
  const long long per = (long long)nc * g.r / 4;  // float4 per vector
  const long long total = per * g.nrhs;
  float m = 0.0f;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x) {
    const int rhs = (int)(t / per);
    const long long k = t - (long long)rhs * per;
    const float4 v = *reinterpret_cast<const float4 *>(M + ((size_t)rhs * g.ncell + c0) * g.r + 4 * k);
    m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = __float_as_uint(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned b = red[0];
    for (int w = 1; w < 4; w++) b = max(b, red[w]);
    atomicMax(g.lvl_max + l, b);
  }
}

// ---------------------------------------------------------------- scatter
// M [rhs][ncell][r] -> per-class grids [slot][rhs][r]; one thread per
// (cell, rhs, 8 consecutive coefficients).
template <int FORM, int CW = 8>
__global__ void k_grid_scatter(GridArgs g, const real *__restrict__ M)
{
  static_assert(CW == 8 || FORM == BBFMM_M2L_PLAIN, "the MFMA grids take 8 coefficients per thread");
  const int chunks = g.r / CW;
  const long long total = (long long)(g.levsta[g.maxlev + 1] - g.levsta[g.minlev]) * g.nrhs * chunks;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int ch = (int)(t % chunks);
  const long long cr = t / chunks;
  const int rhs = (int)(cr % g.nrhs);
  const int A = g.levsta[g.minlev] + (int)(cr / g.nrhs);
  int l = g.minlev;
  while (A >= g.levsta[l + 1]) l++;
  int x, y, z;
  demorton3(A - g.levsta[l], l, x, y, z);
  const int c = ((x & 1) << 2) | ((y & 1) << 1) | (z & 1);
  const size_t slot = grid_slot(g, l, c, x >> 1, y >> 1, z >> 1);
  const real *src = M + ((size_t)rhs * g.ncell + A) * g.r + ch * CW;
  const size_t dst = (slot * g.nrhs + rhs) * g.r + ch * CW;
  if (FORM == BBFMM_M2L_BF16X2) {
    bf16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const float v = (float)src[e];
      const __bf16 h = (__bf16)v;
      hi[e] = h;
      lo[e] = (__bf16)(v - (float)h);
    }
    *reinterpret_cast<bf16x8 *>((__bf16 *)g.Mg0 + dst) = hi;
    *reinterpret_cast<bf16x8 *>((__bf16 *)g.Mg1 + dst) = lo;
  } else if (FORM == BBFMM_M2L_FP16X3) {
    const float sc = ldexpf(1.0f, fp16_scale_exp(g.lvl_max[l]));
    f16x8 t1, t2, t3;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      _Float16 a1, a2, a3;
      split3_f16((float)src[e] * sc, a1, a2, a3);
      t1[e] = a1;
      t2[e] = a2;
      t3[e] = a3;
    }
    *reinterpret_cast<f16x8 *>((_Float16 *)g.Mg0 + dst) = t1;
    *reinterpret_cast<f16x8 *>((_Float16 *)g.Mg1 + dst) = t2;
    *reinterpret_cast<f16x8 *>((_Float16 *)g.Mg2 + dst) = t3;
  } else {
    real *d = (real *)g.Mg0 + dst;
#pragma unroll
    for (int e = 0; e < CW; e++) d[e] = src[e];
  }
}

// Blocks are dealt to the 8 XCDs round-robin (block b on XCD b % 8): the
// logical block of launch block b gives each XCD a contiguous run of logical
// blocks -- the 8 classes of a column group and neighbouring column groups,
// which read the same shifted source tiles -- so they share its L2.  The
// launch is rounded up to a multiple of 8; blocks past the end get -1.
__device__ __forceinline__ int xcd_block(const GridArgs &g, int b)
{
  const int nb = g.blk0[g.maxlev + 1];
  const int lb = (b & 7) * ((nb + 7) >> 3) + (b >> 3);
  return lb < nb ? lb : -1;
}

// level / class / column-group of a block of the flattened launch
__device__ __forceinline__ void decode_block(const GridArgs &g, int b, int &l, int &s, int &grp)
{
  l = g.minlev;
  while (l < g.maxlev && b >= g.blk0[l + 1]) l++;
  const int rel = b - g.blk0[l];
  s = rel & 7;
  grp = rel >> 3;
}

// ---------------------------------------------------------------- MFMA kernel
// 256 threads = 4 waves, all of one class s at one level; each wave owns
// TPW tiles of 32 columns.  Per offset v: A = K_v (bf16 hi/lo or fp32, read
// in lane order), B = the shifted sources of the wave's columns.
template <int FORM, int TPW>
__global__ void __launch_bounds__(256) k_m2l_mfma(GridArgs g)
{
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lb = xcd_block(g, blockIdx.x);
  if (lb < 0) return;
  int l, s, grp;
  decode_block(g, lb, l, s, grp);
  const int P = 1 << (l - 1);
  const int PP = P + 2;
  const int ncols = P * P * P * g.nrhs;
  const int half = lane >> 5;  // which k-half of the operand this lane holds
  int base[TPW], rhs[TPW], pcol[TPW];
  bool valid[TPW];
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    const int col = ((grp * 4 + wave) * TPW + t) * 32 + (lane & 31);
    valid[t] = col < ncols;
    const int c = valid[t] ? col : 0;
    const int pl = c / g.nrhs;
    rhs[t] = c - pl * g.nrhs;
    pcol[t] = pl;
    const int pz = pl % P, py = (pl / P) % P, px = pl / (P * P);
    base[t] = ((px + 1) * PP + py + 1) * PP + pz + 1;
  }
  f32x16 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; t++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[t][e] = 0.0f;

  const int4 *tab = g.tab + s * kOffsets;
  const int slot0 = g.lvl_slot0[l];
  for (int v = 0; v < kOffsets; v++) {
    const int4 e = tab[v];
    const int k = e.x;
    const int sc = e.y;
    const int dx = ((e.z >> 4) & 3) - 1, dy = ((e.z >> 2) & 3) - 1, dz = (e.z & 3) - 1;
    const int off = slot0 + sc * PP * PP * PP + (dx * PP + dy) * PP + dz;
    if (FORM == BBFMM_M2L_BF16X2) {
      // A: [k][kstep][hi/lo][lane] bf16x8
      const bf16x8 *Ka = (const bf16x8 *)g.Ka + (size_t)k * 4 * 64 + lane;
      const bf16x8 a0h = Ka[0 * 64], a0l = Ka[1 * 64], a1h = Ka[2 * 64], a1l = Ka[3 * 64];
      const __bf16 *Mh = (const __bf16 *)g.Mg0;
      const __bf16 *Ml = (const __bf16 *)g.Mg1;
#pragma unroll
      for (int t = 0; t < TPW; t++) {
        const size_t o = ((size_t)(off + base[t]) * g.nrhs + rhs[t]) * 32 + half * 8;
        bf16x8 b0h = *(const bf16x8 *)(Mh + o), b0l = *(const bf16x8 *)(Ml + o);
        bf16x8 b1h = *(const bf16x8 *)(Mh + o + 16), b1l = *(const bf16x8 *)(Ml + o + 16);
        if (!valid[t]) {
#pragma unroll
          for (int q = 0; q < 8; q++) b0h[q] = b0l[q] = b1h[q] = b1l[q] = (__bf16)0.0f;
        }
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, b0h, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, b0l, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0l, b0h, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, b1h, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, b1l, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1l, b1h, acc[t], 0, 0, 0);
      }
    } else if (FORM == BBFMM_M2L_FP16X3) {
      // A: [k][kstep][term][lane] f16x8, as the BF16X2 layout with three terms
      const f16x8 *Ka = (const f16x8 *)g.Ka + (size_t)k * 6 * 64 + lane;
      f16x8 a[2][3];
#pragma unroll
      for (int ks = 0; ks < 2; ks++)
#pragma unroll
        for (int q = 0; q < 3; q++) a[ks][q] = Ka[(ks * 3 + q) * 64];
      const _Float16 *Mt[3] = {(const _Float16 *)g.Mg0, (const _Float16 *)g.Mg1, (const _Float16 *)g.Mg2};
#pragma unroll
      for (int t = 0; t < TPW; t++) {
        const size_t o = ((size_t)(off + base[t]) * g.nrhs + rhs[t]) * 32 + half * 8;
        f16x8 b[2][3];
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
          for (int q = 0; q < 3; q++) {
            b[ks][q] = *(const f16x8 *)(Mt[q] + o + 16 * ks);
            if (!valid[t]) b[ks][q] = f16x8{0, 0, 0, 0, 0, 0, 0, 0};
          }
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ks][0], b[ks][0], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ks][0], b[ks][1], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ks][1], b[ks][0], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ks][0], b[ks][2], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ks][1], b[ks][1], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[ks][2], b[ks][0], acc[t], 0, 0, 0);
        }
      }
    } else {
      // FP32: lane holds k = 16*half + ks for ks = 0..15; A: [k][lane][16]
      const f32x4 *Ka = (const f32x4 *)g.Ka + ((size_t)k * 64 + lane) * 4;
      f32x4 a[4];
#pragma unroll
      for (int q = 0; q < 4; q++) a[q] = Ka[q];
      const float *Mf = (const float *)g.Mg0;
#pragma unroll
      for (int t = 0; t < TPW; t++) {
        const size_t o = ((size_t)(off + base[t]) * g.nrhs + rhs[t]) * 32 + half * 16;
        const f32x4 *bp = (const f32x4 *)(Mf + o);
        f32x4 b[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          b[q] = bp[q];
          if (!valid[t]) b[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int w = 0; w < 4; w++)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q][w], b[q][w], acc[t], 0, 0, 0);
      }
    }
  }
  // epilogue: lane holds rows 8*(e/4) + 4*half + e%4 of its column
  float unscale = 1.0f;
  if (FORM == BBFMM_M2L_FP16X3) unscale = ldexpf(g.kscale_inv, -fp16_scale_exp(g.lvl_max[l]));
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    if (!valid[t]) continue;
    if (FORM == BBFMM_M2L_FP16X3)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[t][e] *= unscale;
    const int pl = pcol[t];
    const int pz = pl % P, py = (pl / P) % P, px = pl / (P * P);
    const int fx = 2 * px + ((s >> 2) & 1), fy = 2 * py + ((s >> 1) & 1), fz = 2 * pz + (s & 1);
    const int A = g.levsta[l] + morton3(fx, fy, fz, l);
    float *Lp = (float *)g.L + ((size_t)rhs[t] * g.ncell + A) * 32 + 4 * half;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      f32x4 o = {acc[t][4 * q], acc[t][4 * q + 1], acc[t][4 * q + 2], acc[t][4 * q + 3]};
      *reinterpret_cast<f32x4 *>(Lp + 8 * q) = o;
    }
  }
}

// ---------------------------------------------------------------- ghost steps of a wave
// The parents of a wave's span columns (columns are (parent, vector) pairs,
// vector fastest, then pz, py, px) lie in the box [lo, hi] per axis (the
// whole axis where the columns span more than one row).  A step of shift d
// whose shifted box lies outside [0, P) in some axis reads only the ghost
// layer (zeros) for every column of the wave: its products add nothing, and
// the wave skips them.  A wave without valid columns skips every step.
struct WaveBox {
  int lo[3], hi[3];
};

__device__ __forceinline__ WaveBox wave_box(int col0, int ncols, int nrhs, int P, int span)
{
  WaveBox b;
  if (col0 >= ncols) {
    for (int a = 0; a < 3; a++) {
      b.lo[a] = 2 * P;
      b.hi[a] = -2 * P;
    }
    return b;
  }
  const int f = col0 / nrhs, t = min(col0 + span - 1, ncols - 1) / nrhs;
  const int fx = f / (P * P), fy = (f / P) % P, fz = f % P;
  const int tx = t / (P * P), ty = (t / P) % P, tz = t % P;
  b.lo[0] = fx;
  b.hi[0] = tx;
  const bool row = fx == tx, line = row && fy == ty;
  b.lo[1] = row ? fy : 0;
  b.hi[1] = row ? ty : P - 1;
  b.lo[2] = line ? fz : 0;
  b.hi[2] = line ? tz : P - 1;
  return b;
}

// the step record's first word: source class | packed shift << 3
__device__ __forceinline__ bool ghost_step(const WaveBox &b, int e0, int P)
{
  const int pk = e0 >> 3;
  const int d[3] = {((pk >> 4) & 3) - 1, ((pk >> 2) & 3) - 1, (pk & 3) - 1};
  bool g = false;
#pragma unroll
  for (int a = 0; a < 3; a++) g |= (b.hi[a] + d[a] < 0) | (b.lo[a] + d[a] >= P);
  return g;
}

// ---------------------------------------------------------------- source-stationary MFMA kernel
// Same products, loop order inverted: a field cell of class s at parent p and
// a source of class s' at parent p + d interact through the offset
// v = 2d + s' - s, so for one (s', d) the shifted source tile M_{s'}[p + d]
// serves every field class.  A wave owns TPW tiles of 32 columns (parent,
// vector) of CLS field classes (TPW x CLS accumulators); per step (s', d) it
// loads its source tiles once and multiplies them by the CLS operators K_v,
// which the block stages in LDS (double-buffered) for all its waves.  CLS = 8
// takes all classes of a column group in one block; CLS = 4 or 2 splits a
// column group over 2 or 4 blocks (class groups), which evens out the work
// per CU when the column groups do not fill whole rounds of the chip, and
// drops the steps in which every class of the group has a near-field offset.
// With TPW = 2 every operator fragment read from LDS serves two tiles.
// Split-BF16 (fp32 build).
template <int CLS, int TPW, bool BUF>
__global__ void __launch_bounds__(256) k_m2l_mfma_ss(GridArgs g)
{
  constexpr int WAVES = 4;  // one k-part (k-step x hi/lo) of the operators per wave
  constexpr int NCG = 8 / CLS;
  __shared__ bf16x8 sK[2][CLS][4][64];  // [buffer][class of the group][kstep*2 + hi/lo][lane]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // Blocks are dealt to the 8 XCDs round-robin (block b on XCD b % 8); the
  // logical order gives each XCD a contiguous run of column groups, i.e.
  // neighbouring parents, whose shifted source tiles then share its L2; the
  // class groups of one column group are adjacent.
  const int nb = g.blk0[g.maxlev + 1];
  const int lb = (blockIdx.x & 7) * ((nb + 7) >> 3) + (blockIdx.x >> 3);
  if (lb >= nb) return;
  int l = g.minlev;
  while (l < g.maxlev && lb >= g.blk0[l + 1]) l++;
  const int nsp = g.nsplit > 1 ? g.nsplit : 1;
  const int rel = lb - g.blk0[l];
  const int part = rel % nsp, rq = rel / nsp;
  const int grp = rq / NCG, cg = rq - grp * NCG;
  const int P = 1 << (l - 1);
  const int PP = P + 2;
  const int ncols = P * P * P * g.nrhs;
  const int half = lane >> 5;
  const int sh0 = PP * PP + PP + 1;
  const int col0 = (grp * WAVES + wave) * 32 * TPW;
  // per tile: columns past the end compute a copy of column 0 and are not stored
  bool valid[TPW];
  int rh[TPW], px[TPW], py[TPW], pz[TPW], base[TPW], voff[TPW];
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    const int col = col0 + 32 * t + (lane & 31);
    valid[t] = col < ncols;
    const int cc = valid[t] ? col : 0;
    const int pl = cc / g.nrhs;
    rh[t] = cc - pl * g.nrhs;
    pz[t] = pl % P;
    py[t] = (pl / P) % P;
    px[t] = pl / (P * P);
    base[t] = ((px[t] + 1) * PP + py[t] + 1) * PP + pz[t] + 1;
    // the scalar offset of a buffer load is unsigned: the shift's most
    // negative value, -(PP^2 + PP + 1), moves to the lane part (base >= PP^2 + PP + 1)
    voff[t] = (((base[t] - sh0) * g.nrhs + rh[t]) * 32 + half * 8) * 2;
  }
  const int slot0 = g.lvl_slot0[l];
  const int cls = PP * PP * PP;
  const bf16x8 *Kl = (const bf16x8 *)g.Ka + wave * 64 + lane;
  const __bf16 *Mh = (const __bf16 *)g.Mg0;
  const __bf16 *Ml = (const __bf16 *)g.Mg1;
  const WaveBox box =
      wave_box((grp * WAVES + __builtin_amdgcn_readfirstlane(wave)) * 32 * TPW, ncols, g.nrhs, P, 32 * TPW);

  f32x16 acc[TPW][CLS];
#pragma unroll
  for (int t = 0; t < TPW; t++)
#pragma unroll
    for (int c = 0; c < CLS; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[t][c][e] = 0.0f;

  // Step record of the class group: e[0] = source class | packed shift << 3,
  // e[1 + i/2] holds the operator of the group's classes i (low 16 bits) and
  // i + 1 (high); near-field pairs point at the zero operator, so every step
  // runs all CLS classes.  Each wave moves its k-part of the step's operators
  // (1 KB each) from global memory straight into the other LDS buffer
  // (global_load_lds, 16 bytes per lane at lane-contiguous addresses) while
  // the MFMAs of the current step run; the source tiles of the next step are
  // loaded into registers at the same time.  Records are read two steps ahead.
  typedef __attribute__((address_space(3))) void lds_void;
  // BUF: buffer loads whose per-step part is a scalar offset (the operator's
  // index, the step's shift) and whose per-lane part is fixed, so that no
  // vector instructions go into the step's addresses (grids below 2 GB)
  const __amdgpu_buffer_rsrc_t rK = __builtin_amdgcn_make_buffer_rsrc((void *)g.Ka, 0, g.kbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rMh = __builtin_amdgcn_make_buffer_rsrc((void *)g.Mg0, 0, g.gbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rMl = __builtin_amdgcn_make_buffer_rsrc((void *)g.Mg1, 0, g.gbytes, 0x00020000);
  auto load_k = [&](const int *e, int buf) {
#pragma unroll
    for (int i = 0; i < CLS; i++) {
      const int k = (e[1 + (i >> 1)] >> (16 * (i & 1))) & 0xffff;
      if (BUF)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rK, (lds_void *)&sK[buf][i][wave][0], 16, (wave * 64 + lane) * 16,
                                                 k * 256 * 16, 0, 0);
      else
        __builtin_amdgcn_global_load_lds((const void *)(Kl + (size_t)k * 256), (lds_void *)&sK[buf][i][wave][0], 16,
                                         0, 0);
    }
  };
  // source tiles of step e: hi and lo of k-steps 0 and 1
  struct Src {
    bf16x8 h0, l0, h1, l1;
  };
  auto load_src = [&](const int *e, Src *d) {
    const int sc = e[0] & 7, pk = e[0] >> 3;
    const int dx = ((pk >> 4) & 3) - 1, dy = ((pk >> 2) & 3) - 1, dz = (pk & 3) - 1;
    const int sft = slot0 + sc * cls + (dx * PP + dy) * PP + dz;
#pragma unroll
    for (int t = 0; t < TPW; t++) {
      if (BUF) {
        const int so = (sft + sh0) * g.nrhs * 64;
        d[t].h0 = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rMh, voff[t], so, 0));
        d[t].l0 = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rMl, voff[t], so, 0));
        d[t].h1 = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rMh, voff[t] + 32, so, 0));
        d[t].l1 = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rMl, voff[t] + 32, so, 0));
      } else {
        const size_t o = ((size_t)(sft + base[t]) * g.nrhs + rh[t]) * 32 + half * 8;
        d[t].h0 = *(const bf16x8 *)(Mh + o);
        d[t].l0 = *(const bf16x8 *)(Ml + o);
        d[t].h1 = *(const bf16x8 *)(Mh + o + 16);
        d[t].l1 = *(const bf16x8 *)(Ml + o + 16);
      }
    }
  };
  // this block's part of the class group's steps (all of them when nsplit = 1)
  const int s_lo = part * g.nsteps[cg] / nsp;
  const int nst = (part + 1) * g.nsteps[cg] / nsp - s_lo;
  const int *S = g.steps + ((size_t)cg * kSteps + s_lo) * kStepWords;
  int e1[5];
  int ec = S[0];  // first word of the current step's record
  Src cur[TPW];
  {
    int e0[5];
#pragma unroll
    for (int i = 0; i < 5; i++) e0[i] = S[i];
    const int s1 = nst > 1 ? 1 : 0;
#pragma unroll
    for (int i = 0; i < 5; i++) e1[i] = S[s1 * kStepWords + i];
    load_k(e0, 0);
    load_src(e0, cur);
  }
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  __syncthreads();
  for (int st = 0; st < nst; st++) {
    const int b = st & 1;
    // the last steps reload their own data (harmless, never used)
    const int n2 = st + 2 < nst ? st + 2 : nst - 1;
    // vector loads (operators arriving by DMA keep the compiler from scalar
    // loads here); moved to scalar registers after the end-of-step wait
    const int4 r2a = *reinterpret_cast<const int4 *>(S + n2 * kStepWords);
    const int r2b = CLS > 6 ? S[n2 * kStepWords + 4] : 0;
    const bool skip = ghost_step(box, ec, P);
    // k-step 0 of all classes, then the k-step-1 fragments; only then the
    // loads into the other buffer (the compiler orders LDS reads after any
    // earlier global_load_lds), then k-step 1
    bf16x8 ah[CLS], al[CLS];
    if (!skip) {
#pragma unroll
      for (int c = 0; c < CLS; c++) {
        ah[c] = sK[b][c][0][lane];
        al[c] = sK[b][c][1][lane];
      }
      // all reads issued before the first product (otherwise the scheduler
      // puts each read just before its MFMA, which then waits on it)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int c = 0; c < CLS; c++)
#pragma unroll
        for (int t = 0; t < TPW; t++) {
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], cur[t].h0, acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], cur[t].l0, acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[c], cur[t].h0, acc[t][c], 0, 0, 0);
        }
#pragma unroll
      for (int c = 0; c < CLS; c++) {
        ah[c] = sK[b][c][2][lane];
        al[c] = sK[b][c][3][lane];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    load_k(e1, b ^ 1);
    Src nxt[TPW];
    load_src(e1, nxt);
    __builtin_amdgcn_sched_barrier(0);
    if (!skip) {
#pragma unroll
      for (int c = 0; c < CLS; c++)
#pragma unroll
        for (int t = 0; t < TPW; t++) {
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], cur[t].h1, acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], cur[t].l1, acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[c], cur[t].h1, acc[t][c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < TPW; t++) cur[t] = nxt[t];
    // the next step's operators are in LDS for every wave after this (the
    // scheduling barrier keeps the MFMAs above the wait)
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    ec = e1[0];
    e1[0] = __builtin_amdgcn_readfirstlane(r2a.x);
    e1[1] = __builtin_amdgcn_readfirstlane(r2a.y);
    e1[2] = __builtin_amdgcn_readfirstlane(r2a.z);
    e1[3] = __builtin_amdgcn_readfirstlane(r2a.w);
    e1[4] = __builtin_amdgcn_readfirstlane(r2b);
  }
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    if (!valid[t]) continue;
#pragma unroll
    for (int i = 0; i < CLS; i++) {
      const int c = cg * CLS + i;
      const int fx = 2 * px[t] + ((c >> 2) & 1), fy = 2 * py[t] + ((c >> 1) & 1), fz = 2 * pz[t] + (c & 1);
      const int A = g.levsta[l] + morton3(fx, fy, fz, l);
      float *Lp = part == 0 ? (float *)g.L + ((size_t)rh[t] * g.ncell + A) * 32 + 4 * half
                            : g.Lx + (((size_t)(part - 1) * g.nrhs + rh[t]) * g.up_ncell + (A - g.up_cell0)) * 32 +
                                  4 * half;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        f32x4 v = {acc[t][i][4 * q], acc[t][i][4 * q + 1], acc[t][i][4 * q + 2], acc[t][i][4 * q + 3]};
        *reinterpret_cast<f32x4 *>(Lp + 8 * q) = v;
      }
    }
  }
}

// ---------------------------------------------------------------- class-group pairs sharing source tiles
// k_m2l_mfma_ss with 4 classes and two 32-column tiles per wave, and both
// class groups of a column group in one block of 8 waves (waves 0-3 classes
// 0-3, waves 4-7 classes 4-7, wave w on the columns of tile pair w % 4).
// The block walks the union of the two groups' step lists (the 8-class
// table: 208 steps; a group's classes that are near in a step take the zero
// operator), so the shifted source tiles of a step are the same for both
// groups: the block moves them into LDS once (wave w moves tile w / 4 of its
// own columns, global_load_lds, double-buffered like the operators) and
// every wave reads its B fragments from there.  Per 48 MFMAs a wave moves
// 4 KB of operators and 4 KB of sources from memory (8 + 4 KB when each
// group loads its own tiles).  LDS: 64 KB of operators + 64 KB of source
// tiles, one block per CU.
template <bool BUF>
__global__ void __launch_bounds__(512) k_m2l_mfma_ss_pair(GridArgs g)
{
  constexpr int CLS = 4, TPW = 2;
  __shared__ bf16x8 sK[2][8][4][64];         // [buffer][class][kstep*2 + hi/lo][lane]
  __shared__ bf16x8 sS[2][4][TPW][4][64];    // [buffer][tile pair][tile][h0, l0, h1, l1][lane]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tp = wave & 3, cg = wave >> 2;
  const int nb = g.blk0[g.maxlev + 1];
  const int lb = (blockIdx.x & 7) * ((nb + 7) >> 3) + (blockIdx.x >> 3);
  if (lb >= nb) return;
  int l = g.minlev;
  while (l < g.maxlev && lb >= g.blk0[l + 1]) l++;
  const int grp = lb - g.blk0[l];
  const int P = 1 << (l - 1);
  const int PP = P + 2;
  const int ncols = P * P * P * g.nrhs;
  const int half = lane >> 5;
  const int sh0 = PP * PP + PP + 1;
  const int col0 = (grp * 4 + tp) * 32 * TPW;
  bool valid[TPW];
  int rh[TPW], px[TPW], py[TPW], pz[TPW], base[TPW], voff[TPW];
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    const int col = col0 + 32 * t + (lane & 31);
    valid[t] = col < ncols;
    const int cc = valid[t] ? col : 0;
    const int pl = cc / g.nrhs;
    rh[t] = cc - pl * g.nrhs;
    pz[t] = pl % P;
    py[t] = (pl / P) % P;
    px[t] = pl / (P * P);
    base[t] = ((px[t] + 1) * PP + py[t] + 1) * PP + pz[t] + 1;
    voff[t] = (((base[t] - sh0) * g.nrhs + rh[t]) * 32 + half * 8) * 2;
  }
  const int slot0 = g.lvl_slot0[l];
  const int cls = PP * PP * PP;
  const WaveBox box = wave_box((grp * 4 + __builtin_amdgcn_readfirstlane(tp)) * 32 * TPW, ncols, g.nrhs, P, 32 * TPW);
  f32x16 acc[TPW][CLS];
#pragma unroll
  for (int t = 0; t < TPW; t++)
#pragma unroll
    for (int c = 0; c < CLS; c++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[t][c][e] = 0.0f;

  typedef __attribute__((address_space(3))) void lds_void;
  const __amdgpu_buffer_rsrc_t rK = __builtin_amdgcn_make_buffer_rsrc((void *)g.Ka, 0, g.kbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rMh = __builtin_amdgcn_make_buffer_rsrc((void *)g.Mg0, 0, g.gbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rMl = __builtin_amdgcn_make_buffer_rsrc((void *)g.Mg1, 0, g.gbytes, 0x00020000);
  const bf16x8 *Kl = (const bf16x8 *)g.Ka + tp * 64 + lane;
  const __bf16 *Mh = (const __bf16 *)g.Mg0;
  const __bf16 *Ml = (const __bf16 *)g.Mg1;
  const int my_t = cg;  // the tile of its own columns this wave moves into LDS
  // operators: wave w moves k-part w % 4 of its group's 4 classes
  auto load_k = [&](const int *e, int buf) {
#pragma unroll
    for (int i = 0; i < CLS; i++) {
      const int k = (e[1 + 2 * cg + (i >> 1)] >> (16 * (i & 1))) & 0xffff;
      if (BUF)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rK, (lds_void *)&sK[buf][4 * cg + i][tp][0], 16,
                                                 (tp * 64 + lane) * 16, k * 256 * 16, 0, 0);
      else
        __builtin_amdgcn_global_load_lds((const void *)(Kl + (size_t)k * 256), (lds_void *)&sK[buf][4 * cg + i][tp][0],
                                         16, 0, 0);
    }
  };
  // source tile my_t of this wave's columns for step e: h0, l0, h1, l1
  auto load_s = [&](const int *e, int buf) {
    const int sc = e[0] & 7, pk = e[0] >> 3;
    const int dx = ((pk >> 4) & 3) - 1, dy = ((pk >> 2) & 3) - 1, dz = (pk & 3) - 1;
    const int sft = slot0 + sc * cls + (dx * PP + dy) * PP + dz;
    lds_void *d = (lds_void *)&sS[buf][tp][my_t][0][0];
    if (BUF) {
      const int so = (sft + sh0) * g.nrhs * 64;
      const int vo = my_t ? voff[1] : voff[0];
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rMh, d, 16, vo, so, 0, 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rMl, (lds_void *)&sS[buf][tp][my_t][1][0], 16, vo, so, 0, 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rMh, (lds_void *)&sS[buf][tp][my_t][2][0], 16, vo + 32, so, 0, 0);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rMl, (lds_void *)&sS[buf][tp][my_t][3][0], 16, vo + 32, so, 0, 0);
    } else {
      const int bs = my_t ? base[1] : base[0];
      const int rr = my_t ? rh[1] : rh[0];
      const size_t o = ((size_t)(sft + bs) * g.nrhs + rr) * 32 + half * 8;
      __builtin_amdgcn_global_load_lds((const void *)(Mh + o), d, 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void *)(Ml + o), (lds_void *)&sS[buf][tp][my_t][1][0], 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void *)(Mh + o + 16), (lds_void *)&sS[buf][tp][my_t][2][0], 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void *)(Ml + o + 16), (lds_void *)&sS[buf][tp][my_t][3][0], 16, 0, 0);
    }
  };
  const int nst = g.nsteps[0];
  const int *S = g.steps;
  int e1[5];
  int ec = S[0];
  {
    int e0[5];
#pragma unroll
    for (int i = 0; i < 5; i++) e0[i] = S[i];
    const int s1 = nst > 1 ? 1 : 0;
#pragma unroll
    for (int i = 0; i < 5; i++) e1[i] = S[s1 * kStepWords + i];
    load_k(e0, 0);
    load_s(e0, 0);
  }
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  __syncthreads();
  for (int st = 0; st < nst; st++) {
    const int b = st & 1;
    const int n2 = st + 2 < nst ? st + 2 : nst - 1;
    const int4 r2a = *reinterpret_cast<const int4 *>(S + n2 * kStepWords);
    const int r2b = S[n2 * kStepWords + 4];
    const bool skip = ghost_step(box, ec, P);
    // every LDS read of the step comes before the DMA into the other buffer
    bf16x8 ah[CLS], al[CLS], bh[TPW], bl[TPW];
    if (!skip) {
#pragma unroll
      for (int c = 0; c < CLS; c++) {
        ah[c] = sK[b][4 * cg + c][0][lane];
        al[c] = sK[b][4 * cg + c][1][lane];
      }
#pragma unroll
      for (int t = 0; t < TPW; t++) {
        bh[t] = sS[b][tp][t][0][lane];
        bl[t] = sS[b][tp][t][1][lane];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int c = 0; c < CLS; c++)
#pragma unroll
        for (int t = 0; t < TPW; t++) {
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bh[t], acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bl[t], acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[c], bh[t], acc[t][c], 0, 0, 0);
        }
#pragma unroll
      for (int c = 0; c < CLS; c++) {
        ah[c] = sK[b][4 * cg + c][2][lane];
        al[c] = sK[b][4 * cg + c][3][lane];
      }
#pragma unroll
      for (int t = 0; t < TPW; t++) {
        bh[t] = sS[b][tp][t][2][lane];
        bl[t] = sS[b][tp][t][3][lane];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    load_k(e1, b ^ 1);
    load_s(e1, b ^ 1);
    __builtin_amdgcn_sched_barrier(0);
    if (!skip) {
#pragma unroll
      for (int c = 0; c < CLS; c++)
#pragma unroll
        for (int t = 0; t < TPW; t++) {
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bh[t], acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[c], bl[t], acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[c], bh[t], acc[t][c], 0, 0, 0);
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    ec = e1[0];
    e1[0] = __builtin_amdgcn_readfirstlane(r2a.x);
    e1[1] = __builtin_amdgcn_readfirstlane(r2a.y);
    e1[2] = __builtin_amdgcn_readfirstlane(r2a.z);
    e1[3] = __builtin_amdgcn_readfirstlane(r2a.w);
    e1[4] = __builtin_amdgcn_readfirstlane(r2b);
  }
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    if (!valid[t]) continue;
#pragma unroll
    for (int i = 0; i < CLS; i++) {
      const int c = 4 * cg + i;
      const int fx = 2 * px[t] + ((c >> 2) & 1), fy = 2 * py[t] + ((c >> 1) & 1), fz = 2 * pz[t] + (c & 1);
      const int A = g.levsta[l] + morton3(fx, fy, fz, l);
      float *Lp = (float *)g.L + ((size_t)rh[t] * g.ncell + A) * 32 + 4 * half;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        f32x4 v = {acc[t][i][4 * q], acc[t][i][4 * q + 1], acc[t][i][4 * q + 2], acc[t][i][4 * q + 3]};
        *reinterpret_cast<f32x4 *>(Lp + 8 * q) = v;
      }
    }
  }
}

// L of the cells up_cell0 .. up_cell0 + up_ncell - 1 += the partial sums of
// parts 1 .. nsplit - 1 (in part order), one float4 per thread
__global__ void __launch_bounds__(256) k_m2l_add_parts(GridArgs g)
{
  const size_t n4 = (size_t)g.nrhs * g.up_ncell * 8;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n4) return;
  const size_t rc = i >> 3;  // rhs * up_ncell + cell
  const int q = (int)(i & 7);
  const size_t rh = rc / g.up_ncell, c = rc - rh * g.up_ncell;
  f32x4 *Lp = reinterpret_cast<f32x4 *>(g.L + (rh * g.ncell + g.up_cell0 + c) * 32) + q;
  f32x4 v = *Lp;
  const f32x4 *X = reinterpret_cast<const f32x4 *>(g.Lx);
  for (int k = 0; k + 1 < g.nsplit; k++) v += X[((size_t)k * g.nrhs * g.up_ncell + rc) * 8 + q];
  *Lp = v;
}

// ---------------------------------------------------------------- r = 256 FP32 MFMA kernel
// 256 threads = 4 waves, all of one class s at one level; each wave owns 32
// columns and all 256 rows (8 row tiles of v_mfma_f32_32x32x2f32).  For each
// of the class's 189 offsets v, K_v streams through LDS in 8 chunks of 32
// k-values (32 KB, layout [row tile][lane][16]: lane l holds row l%32 of the
// tile and k = 16*(l/32) + 0..15 of the chunk), double-buffered and loaded
// by LDS-DMA one chunk ahead; the lane's B values are the 16 consecutive
// coefficients k = 16*(l/32) + 0..15 of its source column.
constexpr int kR256Chunks = 8;  // 256 / 32

__global__ void __launch_bounds__(256) k_m2l_fp32_r256(GridArgs g)
{
  __shared__ f32x4 sK[2][8][64][4];  // [buffer][row tile][lane][16 floats]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // XCD-aware order: each XCD takes a contiguous run of (column group, class)
  const int nb = g.blk0[g.maxlev + 1];
  const int lb = (blockIdx.x & 7) * ((nb + 7) >> 3) + (blockIdx.x >> 3);
  if (lb >= nb) return;
  int l, s, grp;
  decode_block(g, lb, l, s, grp);
  const int P = 1 << (l - 1);
  const int PP = P + 2;
  const int ncols = P * P * P * g.nrhs;
  const int half = lane >> 5;
  // columns past the end compute a copy of column 0 and are not stored
  const int col = (grp * 4 + wave) * 32 + (lane & 31);
  const bool valid = col < ncols;
  const int cc = valid ? col : 0;
  const int pl = cc / g.nrhs;
  const int rh = cc - pl * g.nrhs;
  const int pz = pl % P, py = (pl / P) % P, px = pl / (P * P);
  const int base = ((px + 1) * PP + py + 1) * PP + pz + 1;
  const int slot0 = g.lvl_slot0[l];
  const int cls = PP * PP * PP;
  const float *Mf = (const float *)g.Mg0;
  const int4 *tab = g.tab + s * kOffsets;
  // chunk (v, kc) of the operator stream: 32 KB at Ka + (k * 8 + kc) * 32 KB
  const f32x4 *Kq = (const f32x4 *)g.Ka;
  typedef __attribute__((address_space(3))) void lds_void;
  auto load_chunk = [&](int k, int kc, int buf) {
    const f32x4 *src = Kq + ((size_t)k * kR256Chunks + kc) * 2048 + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 8; i++)
      __builtin_amdgcn_global_load_lds((const void *)(src + i * 256),
                                       (lds_void *)((f32x4 *)&sK[buf][0][0][0] + i * 256 + wave * 64), 16, 0, 0);
  };
  auto src_col = [&](int4 e) {
    const int dx = ((e.z >> 4) & 3) - 1, dy = ((e.z >> 2) & 3) - 1, dz = (e.z & 3) - 1;
    const int off = slot0 + e.y * cls + (dx * PP + dy) * PP + dz;
    return ((size_t)(off + base) * g.nrhs + rh) * 256 + half * 16;
  };
  auto load_b = [&](size_t o, int kc, f32x4 *bq) {
    const f32x4 *bp = (const f32x4 *)(Mf + o + kc * 32);
#pragma unroll
    for (int q = 0; q < 4; q++) bq[q] = bp[q];
  };

  f32x16 acc[8];
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[t][e] = 0.0f;

  int4 e = tab[0];
  size_t o = src_col(e);
  f32x4 bq[4];
  load_chunk(e.x, 0, 0);
  load_b(o, 0, bq);
  __builtin_amdgcn_s_waitcnt(kWaitVm0);
  __syncthreads();
  const int nsteps = kOffsets * kR256Chunks;
  for (int st = 0; st < nsteps; st++) {
    const int b = st & 1;
    const int kc = st & (kR256Chunks - 1);
    // next chunk: same offset, next k-chunk, or the next offset
    const int sn = st + 1 < nsteps ? st + 1 : st;
    const int kcn = sn & (kR256Chunks - 1);
    const int4 en = tab[sn / kR256Chunks];
    const size_t on = kcn == 0 ? src_col(en) : o;
    // row tiles 0-3, then the fragments of tiles 4-7, then the loads of the
    // next chunk (LDS reads must precede the DMA into the other buffer)
    f32x4 a[4][4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int q = 0; q < 4; q++) a[t][q] = sK[b][t][lane][q];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int w = 0; w < 4; w++)
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[t][q][w], bq[q][w], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int q = 0; q < 4; q++) a[t][q] = sK[b][4 + t][lane][q];
    __builtin_amdgcn_sched_barrier(0);
    f32x4 bn[4];
    load_chunk(en.x, kcn, b ^ 1);
    load_b(on, kcn, bn);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int w = 0; w < 4; w++)
          acc[4 + t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[t][q][w], bq[q][w], acc[4 + t], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; q++) bq[q] = bn[q];
    o = on;
    (void)kc;
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
  }
  if (!valid) return;
  const int fx = 2 * px + ((s >> 2) & 1), fy = 2 * py + ((s >> 1) & 1), fz = 2 * pz + (s & 1);
  const int A = g.levsta[l] + morton3(fx, fy, fz, l);
  float *Lp = (float *)g.L + ((size_t)rh * g.ncell + A) * 256 + 4 * half;
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      f32x4 v = {acc[t][4 * q], acc[t][4 * q + 1], acc[t][4 * q + 2], acc[t][4 * q + 3]};
      *reinterpret_cast<f32x4 *>(Lp + 32 * t + 8 * q) = v;
    }
}

// ---------------------------------------------------------------- plain kernel
// One thread per (column, row); 256 / R columns per block, all of one class.
template <int R>
__global__ void __launch_bounds__(256) k_m2l_grid_plain(GridArgs g)
{
  constexpr int CPB = 256 / R;
  const int lb = xcd_block(g, blockIdx.x);
  if (lb < 0) return;
  int l, s, grp;
  decode_block(g, lb, l, s, grp);
  const int P = 1 << (l - 1);
  const int PP = P + 2;
  const int ncols = P * P * P * g.nrhs;
  const int i = threadIdx.x % R;
  const int col = grp * CPB + threadIdx.x / R;
  if (col >= ncols) return;  // no barrier in this kernel
  const int pl = col / g.nrhs, rh = col - (col / g.nrhs) * g.nrhs;
  const int pz = pl % P, py = (pl / P) % P, px = pl / (P * P);
  const int base = ((px + 1) * PP + py + 1) * PP + pz + 1;
  const int4 *tab = g.tab + s * kOffsets;
  const real4 *K4 = (const real4 *)g.Ka;
  const real *Mg = (const real *)g.Mg0;
  const int slot0 = g.lvl_slot0[l];
  real acc0 = 0, acc1 = 0;
  for (int v = 0; v < kOffsets; v++) {
    const int4 e = tab[v];
    const int dx = ((e.z >> 4) & 3) - 1, dy = ((e.z >> 2) & 3) - 1, dz = (e.z & 3) - 1;
    const size_t slot = (size_t)slot0 + e.y * PP * PP * PP + (dx * PP + dy) * PP + dz + base;
    const real4 *Mp = (const real4 *)(Mg + (slot * g.nrhs + rh) * R);
    const real4 *Kp = K4 + (size_t)e.x * (R / 4) * R + i;
#pragma unroll 8
    for (int jb = 0; jb < R / 4; jb += 2) {
      const real4 ka = Kp[(size_t)jb * R], kb = Kp[(size_t)(jb + 1) * R];
      const real4 ma = Mp[jb], mb = Mp[jb + 1];
      acc0 += ka.x * ma.x;
      acc1 += kb.x * mb.x;
      acc0 += ka.y * ma.y;
      acc1 += kb.y * mb.y;
      acc0 += ka.z * ma.z;
      acc1 += kb.z * mb.z;
      acc0 += ka.w * ma.w;
      acc1 += kb.w * mb.w;
    }
  }
  const int fx = 2 * px + ((s >> 2) & 1), fy = 2 * py + ((s >> 1) & 1), fz = 2 * pz + (s & 1);
  const int A = g.levsta[l] + morton3(fx, fy, fz, l);
  g.L[((size_t)rh * g.ncell + A) * R + i] = acc0 + acc1;
}

// r = 32, real precision: a block is 64 columns (lanes) of one class at one
// level and wave w computes rows 8w..8w+7 of all of them.  K_v (column-major,
// the reference's layout) is the same for every lane of a wave, so its
// entries arrive by scalar loads into SGPRs: each FMA reads one SGPR (K) and
// one VGPR (a coefficient of the lane's source column), no LDS or vector
// load is spent on K, and the only vector loads are the lane's 32-coefficient
// source column per offset, issued one offset ahead.
__global__ void __launch_bounds__(256) k_m2l_grid_plain32(GridArgs g)
{
  const int lb = xcd_block(g, blockIdx.x);
  if (lb < 0) return;
  int l, s, grp;
  decode_block(g, lb, l, s, grp);
  const int P = 1 << (l - 1);
  const int PP = P + 2;
  const int ncols = P * P * P * g.nrhs;
  const int lane = threadIdx.x & 63;
  const int rg = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);  // row group
  const int col = grp * 64 + lane;
  const bool valid = col < ncols;
  const int cc = valid ? col : 0;  // columns past the end compute column 0 and are not stored
  const int pl = cc / g.nrhs, rh = cc - (cc / g.nrhs) * g.nrhs;
  const int pz = pl % P, py = (pl / P) % P, px = pl / (P * P);
  const int base = ((px + 1) * PP + py + 1) * PP + pz + 1;
  const int4 *__restrict__ tab = g.tab + s * kOffsets;
  const real *__restrict__ Kc = (const real *)g.Ka + rg * 8;
  const real *__restrict__ Mg = (const real *)g.Mg0;
  const size_t slot0 = (size_t)g.lvl_slot0[l] + base;
  const int cls = PP * PP * PP;
  auto src = [&](int v) {
    const int4 e = tab[v];
    const int dx = ((e.z >> 4) & 3) - 1, dy = ((e.z >> 2) & 3) - 1, dz = (e.z & 3) - 1;
    const size_t slot = slot0 + e.y * cls + (dx * PP + dy) * PP + dz;
    return reinterpret_cast<const real4 *>(Mg + (slot * g.nrhs + rh) * 32);
  };
  real acc[8];
#pragma unroll
  for (int i = 0; i < 8; i++) acc[i] = 0;
  auto load = [&](int v, real4 *m) {
    const real4 *p = src(v);
#pragma unroll
    for (int q = 0; q < 8; q++) m[q] = p[q];
  };
  auto mac = [&](int v, const real4 *m) {
    const real *__restrict__ Kk = Kc + (size_t)tab[v].x * 1024;
#pragma unroll
    for (int jq = 0; jq < 8; jq++) {
      const real mj[4] = {m[jq].x, m[jq].y, m[jq].z, m[jq].w};
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const real *__restrict__ kc = Kk + (4 * jq + t) * 32;  // column j: rows 8 rg .. 8 rg + 7
#pragma unroll
        for (int i = 0; i < 8; i++) acc[i] += kc[i] * mj[t];
      }
    }
  };
  // two register sets in turn: the next offset's column is in flight while
  // the current one is multiplied (no copy between the sets, so no wait for
  // the loads at the end of an iteration)
  real4 ma[8], mb[8];
  load(0, ma);
  static_assert(kOffsets % 2 == 1, "the loop below ends on an odd offset count");
  for (int v = 0; v + 1 < kOffsets; v += 2) {
    load(v + 1, mb);
    __builtin_amdgcn_sched_barrier(0);  // loads issued before the products
    mac(v, ma);
    load(v + 2, ma);
    __builtin_amdgcn_sched_barrier(0);
    mac(v + 1, mb);
  }
  mac(kOffsets - 1, ma);
  if (!valid) return;
  const int fx = 2 * px + ((s >> 2) & 1), fy = 2 * py + ((s >> 1) & 1), fz = 2 * pz + (s & 1);
  const int A = g.levsta[l] + morton3(fx, fy, fz, l);
  real *Lp = g.L + ((size_t)rh * g.ncell + A) * 32 + rg * 8;
#pragma unroll
  for (int i = 0; i < 8; i++) Lp[i] = acc[i];
}

// Any rank (r other than 32 and 256): as k_m2l_grid_plain with K in the
// reference's column-major layout (K_k(i, j) at k r^2 + j r + i) and scalar
// loads; 256 / r columns per block (r <= 256) or one column with
// ceil(r / 256) rows per thread.
__global__ void __launch_bounds__(256) k_m2l_grid_plain_any(GridArgs g)
{
  constexpr int RPT = (kMaxRank + 255) / 256;
  const int r = g.r;
  const int rt = r < 256 ? r : 256, cpb = 256 / rt;
  const int lb = xcd_block(g, blockIdx.x);
  if (lb < 0) return;
  int l, s, grp;
  decode_block(g, lb, l, s, grp);
  const int P = 1 << (l - 1);
  const int PP = P + 2;
  const int ncols = P * P * P * g.nrhs;
  const int cl = (int)threadIdx.x / rt, i0 = (int)threadIdx.x - cl * rt;
  const int col = grp * cpb + cl;
  if (cl >= cpb || col >= ncols) return;  // no barrier in this kernel
  const int pl = col / g.nrhs, rh = col - (col / g.nrhs) * g.nrhs;
  const int pz = pl % P, py = (pl / P) % P, px = pl / (P * P);
  const int base = ((px + 1) * PP + py + 1) * PP + pz + 1;
  const int4 *tab = g.tab + s * kOffsets;
  const real *Kc = (const real *)g.Ka;
  const real *Mg = (const real *)g.Mg0;
  const int slot0 = g.lvl_slot0[l];
  real acc[RPT];
#pragma unroll
  for (int u = 0; u < RPT; u++) acc[u] = 0;
  for (int v = 0; v < kOffsets; v++) {
    const int4 e = tab[v];
    const int dx = ((e.z >> 4) & 3) - 1, dy = ((e.z >> 2) & 3) - 1, dz = (e.z & 3) - 1;
    const size_t slot = (size_t)slot0 + e.y * PP * PP * PP + (dx * PP + dy) * PP + dz + base;
    const real *Mp = Mg + (slot * g.nrhs + rh) * r;
    const real *Kk = Kc + (size_t)e.x * r * r + i0;
    for (int j = 0; j < r; j++) {
      const real m = Mp[j];
#pragma unroll
      for (int u = 0; u < RPT; u++)
        if (i0 + u * rt < r) acc[u] += Kk[(size_t)j * r + u * rt] * m;
    }
  }
  const int fx = 2 * px + ((s >> 2) & 1), fy = 2 * py + ((s >> 1) & 1), fz = 2 * pz + (s & 1);
  const int A = g.levsta[l] + morton3(fx, fy, fz, l);
#pragma unroll
  for (int u = 0; u < RPT; u++)
    if (i0 + u * rt < r) g.L[((size_t)rh * g.ncell + A) * r + i0 + u * rt] = acc[u];
}

// ---------------------------------------------------------------- host side
static void offset_table(std::vector<int4> &tab)
{
  int Kt[343];
  ktable(Kt);
  tab.assign(8 * kOffsets, int4{0, 0, 0, 0});
  for (int s = 0; s < 8; s++) {
    const int sb[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1};
    int nv = 0;
    for (int vx = -2 - sb[0]; vx <= 3 - sb[0]; vx++)
      for (int vy = -2 - sb[1]; vy <= 3 - sb[1]; vy++)
        for (int vz = -2 - sb[2]; vz <= 3 - sb[2]; vz++) {
          if (vx >= -1 && vx <= 1 && vy >= -1 && vy <= 1 && vz >= -1 && vz <= 1) continue;
          const int v[3] = {vx, vy, vz};
          int sc = 0, packed = 0;
          for (int a = 0; a < 3; a++) {
            const int t = sb[a] + v[a];           // in [-2, 3]
            const int d = t >= 0 ? t / 2 : -1;    // floor(t / 2) in {-1, 0, 1}
            sc |= (t - 2 * d) << (2 - a);
            packed |= (d + 1) << (2 * (2 - a));
          }
          tab[s * kOffsets + nv] = int4{Kt[49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)], sc, packed, 0};
          nv++;
        }
  }
}

// (s', d) steps of the source-stationary kernel for field classes split into
// groups of cls: for every source class s' and non-zero parent shift d, the
// operator index of each field class s of the group (offset v = 2d + s' - s),
// or the zero operator when v is a near-field offset; steps in which every
// class of the group is near are left out.  Record: [0] = s' | packed shift
// << 3, [1 + i/2] = operators of the group's classes i (low 16 bits) and
// i + 1 (high 16 bits), rest unused.  Layout [group][kSteps][kStepWords].
static void step_table(std::vector<int> &steps, int cls, int *nsteps)
{
  int Kt[343];
  ktable(Kt);
  const int ncg = 8 / cls;
  steps.assign((size_t)ncg * kSteps * kStepWords, 0);
  for (int cg = 0; cg < ncg; cg++) {
    int st = 0;
    for (int sc = 0; sc < 8; sc++)
      for (int dxi = 0; dxi < 3; dxi++)
        for (int dyi = 0; dyi < 3; dyi++)
          for (int dzi = 0; dzi < 3; dzi++) {
            if (dxi == 1 && dyi == 1 && dzi == 1) continue;
            int *e = &steps[((size_t)cg * kSteps + st) * kStepWords];
            e[0] = sc | (((dxi << 4) | (dyi << 2) | dzi) << 3);
            bool far_any = false;
            for (int i = 0; i < cls; i++) {
              const int c = cg * cls + i;
              const int vx = 2 * (dxi - 1) + ((sc >> 2) & 1) - ((c >> 2) & 1);
              const int vy = 2 * (dyi - 1) + ((sc >> 1) & 1) - ((c >> 1) & 1);
              const int vz = 2 * (dzi - 1) + (sc & 1) - (c & 1);
              const bool near = vx >= -1 && vx <= 1 && vy >= -1 && vy <= 1 && vz >= -1 && vz <= 1;
              far_any |= !near;
              const int k = near ? kZeroOp : Kt[49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)];
              e[1 + (i >> 1)] |= k << (16 * (i & 1));
            }
            if (far_any) {
              st++;
            } else {
              for (int w = 0; w < kStepWords; w++) e[w] = 0;
            }
          }
    nsteps[cg] = st;
  }
}

// field classes per wave of the source-stationary variants (1 / 4, 7 / 5, 6),
// 0 otherwise, and their 32-column tiles per wave
static int ss_classes(int variant)
{
  return variant == 1 ? 8 : (variant == 4 || variant == 7 || variant == 8) ? 4 : (variant == 5 || variant == 6) ? 2 : 0;
}

static int ss_tiles(int variant) { return variant == 6 || variant == 7 || variant == 8 ? 2 : 1; }

int grid_m2l_create(GridM2L **out, int r, int maxlev, int nrhs, int form, const real *K, hipStream_t s,
                    int minlev, int variant, int variant_upper, int upper_split)
{
  *out = nullptr;
  // variant 0 on a deepest level of 64 K columns (parents x vectors) or more,
  // one block of 256 columns and both class groups per CU on 256 CUs: that
  // level with two tiles per wave and the class groups sharing the source
  // tiles (variant 8; fewer bytes moved per product), the levels above in a
  // launch of one tile per wave, their steps in 4 parts while they have
  // fewer blocks than 3 per CU (N = 1M uniform, depth 5, 16 vectors: 0.579
  // against 0.595 ms with variant 7 on the deepest level and 0.601 ms for
  // one launch of one tile per wave)
  if (variant == 0 && variant_upper == 0 && upper_split == 1 && form == BBFMM_M2L_BF16X2 && r == 32 &&
      maxlev > minlev && (long long)(1 << (3 * (maxlev - 1))) * nrhs >= 64 * 1024) {
    long long up_blocks = 0;
    for (int l = minlev; l < maxlev; l++) up_blocks += 2 * (((long long)(1 << (3 * (l - 1))) * nrhs + 127) / 128);
    variant = 8;
    variant_upper = 4;
    upper_split = up_blocks * 4 <= 3 * 256 ? 4 : 1;
  }
  if (upper_split < 1 || upper_split > 16)
    return set_error(BBFMM_E_ARG, "dense-grid M2L: step parts %d outside 1..16", upper_split);
  // step parts without an upper-level variant: the whole (single) launch
  int whole_split = 1;
  if (variant_upper == 0 && upper_split > 1) {
    whole_split = upper_split;
    upper_split = 1;
  }
  if (variant_upper != 0 && (!(ss_classes(variant_upper) && ss_classes(variant)) || variant_upper == 8))
    return set_error(BBFMM_E_ARG, "dense-grid M2L: a variant for the upper levels (%d) needs source-stationary "
                     "variants for both (1, 4 to 7 above the deepest level)", variant_upper);
  if (maxlev > kMaxLev) return set_error(BBFMM_E_ARG, "dense-grid M2L: maxlev %d > %d", maxlev, kMaxLev);
  if (minlev < kMinLev || minlev > maxlev)
    return set_error(BBFMM_E_ARG, "dense-grid M2L: first level %d outside [%d, %d]", minlev, kMinLev, maxlev);
  if (variant < 0 || variant > 8) return set_error(BBFMM_E_ARG, "dense-grid M2L: variant %d", variant);
  if ((variant == 1 || variant >= 4) && form != BBFMM_M2L_BF16X2)
    return set_error(BBFMM_E_ARG, "dense-grid M2L: the source-stationary variant is split-BF16 only");
  if (variant >= 2 && (form == BBFMM_M2L_PLAIN || r != 32))
    return set_error(BBFMM_E_ARG, "dense-grid M2L: tile variants are for the r = 32 MFMA forms");
  if (r < 1 || r > kMaxRank) return set_error(BBFMM_E_ARG, "dense-grid M2L: r=%d outside 1..%d", r, kMaxRank);
  if (form != BBFMM_M2L_PLAIN && r != 32 && r != 256)
    return set_error(BBFMM_E_ARG, "dense-grid M2L: the MFMA forms need r = 32 or 256 (got r=%d)", r);
#ifdef SINGLE
  if (form == BBFMM_M2L_BF16X2 && r != 32)
    return set_error(BBFMM_E_ARG, "the split-BF16 M2L form is built for r = 32 (got r=%d)", r);
  if (form == BBFMM_M2L_FP16X3 && r != 32)
    return set_error(BBFMM_E_ARG, "the split-FP16 M2L form is built for r = 32 (got r=%d)", r);
#else
  if (form != BBFMM_M2L_PLAIN) return set_error(BBFMM_E_ARG, "MFMA M2L forms need the fp32 build");
#endif
  GridM2L *g = new GridM2L();
  g->r = r;
  g->maxlev = maxlev;
  g->nrhs = nrhs;
  g->form = form;
  GridArgs &a = g->args;
  a.r = r;
  a.nrhs = nrhs;
  a.minlev = minlev;
  a.maxlev = maxlev;
  a.nsplit = 1;
  g->minlev = minlev;
  g->variant = variant;
  for (int l = 0; l <= maxlev + 1; l++) a.levsta[l] = ((1 << (3 * l)) - 1) / 7;
  a.ncell = a.levsta[maxlev + 1];
  size_t slots = 0;
  for (int l = minlev; l <= maxlev; l++) {
    a.lvl_slot0[l] = (int)slots;
    const size_t PP = ((size_t)1 << (l - 1)) + 2;
    slots += 8 * PP * PP * PP;
  }
  g->nslots = slots;
  {
    const size_t gb = slots * (size_t)nrhs * r * 2;  // split-BF16 grids: 2 bytes per value
    a.gbytes = gb < ((size_t)1 << 31) ? (int)gb : 0;
  }
  int tpw = 1, cpb;
  // variant 0 (automatic): source-stationary blocking (4 classes per wave)
  // once the deepest level has 8192 columns (parents x vectors) or more (at
  // depth 4 with 16 vectors: 0.177 ms against 0.228 ms for two class-stationary
  // tiles per wave);
  // 1 / 4 / 5: source-stationary with 8 / 4 / 2 field classes per wave and
  // one 32-column tile, 6 / 7: 2 / 4 classes and two tiles per wave;
  // 2 / 3: class-stationary, one / two 32-column tiles per wave
  const long long deep = (long long)(1 << (3 * (maxlev - 1))) * nrhs;
  a.ss = 0;
  if (form == BBFMM_M2L_BF16X2) {
    a.ss = variant == 0 && deep >= 8 * 1024 ? 4 : ss_classes(variant);
    a.ss_tpw = ss_tiles(variant);
    a.ss_pair = variant == 8;
  }
  if (form == BBFMM_M2L_PLAIN) {
    cpb = r == 32 ? 64 : 256 / (r < 256 ? r : 256);  // k_m2l_grid_plain32: 64 columns per block
  } else if (a.ss) {
    cpb = 4 * 32 * a.ss_tpw;  // 4 waves x 32 or 64 columns, one class group per block
  } else {
    // two tiles per wave once there are enough columns to fill the chip
    tpw = variant == 3 ? 2 : variant == 2 ? 1 : (r == 32 && deep >= 8 * 1024 ? 2 : 1);  // r = 256: one tile
    cpb = 4 * tpw * 32;
  }
  a.wave_cols = tpw * 32;
  if (whole_split > 1 && !a.ss) {
    delete g;
    return set_error(BBFMM_E_ARG, "dense-grid M2L: step parts need the split-BF16 source-stationary kernel");
  }
  // variant 0 on a deepest level of 8 K to 64 K columns (one launch, 4
  // classes and one tile per wave): each class group's step list in as many
  // parts as keep the blocks to one round at 3 per CU on 256 CUs (N = 262144
  // uniform, depth 4, 16 vectors: 146 column-group blocks, 5 parts)
  if (variant == 0 && variant_upper == 0 && whole_split == 1 && a.ss == 4 && a.ss_tpw == 1 && !a.ss_pair) {
    long long nb1 = 0;
    for (int l = minlev; l <= maxlev; l++) nb1 += 2 * (((long long)(1 << (3 * (l - 1))) * nrhs + cpb - 1) / cpb);
    whole_split = (int)std::max(1LL, std::min(16LL, 3 * 256 / nb1));
  }
  a.nsplit = whole_split;
  a.up_cell0 = a.levsta[minlev];
  a.up_ncell = a.levsta[maxlev + 1] - a.levsta[minlev];
  int nb = 0;
  for (int l = minlev; l <= maxlev; l++) {
    a.blk0[l] = nb;
    const long long cols = (long long)(1 << (3 * (l - 1))) * nrhs;
    nb += whole_split * (a.ss_pair ? 1 : a.ss ? 8 / a.ss : 8) * (int)((cols + cpb - 1) / cpb);
  }
  a.blk0[maxlev + 1] = nb;
  g->nblocks = nb;

  std::vector<int4> tab;
  offset_table(tab);
  std::vector<int> steps;
  step_table(steps, a.ss && !a.ss_pair ? a.ss : 8, a.nsteps);
  // levels above the deepest with their own class-group size: two launches
  const int ss_up = ss_classes(variant_upper), tpw_up = ss_tiles(variant_upper);
  size_t steps_up_off = 0;
  if (a.ss && ss_up && (ss_up != a.ss || tpw_up != a.ss_tpw || upper_split > 1) && maxlev > minlev) {
    g->nparts = 2;
    GridArgs &u = g->part[0], &d = g->part[1];
    u = a;
    d = a;
    u.maxlev = maxlev - 1;
    u.ss = ss_up;
    u.ss_tpw = tpw_up;
    u.ss_pair = 0;
    u.wave_cols = 32 * tpw_up;
    u.nsplit = upper_split;
    u.up_cell0 = a.levsta[minlev];
    u.up_ncell = a.levsta[maxlev] - a.levsta[minlev];
    const int cpb_up = 4 * 32 * tpw_up;
    int nbu = 0;
    for (int l = minlev; l <= maxlev - 1; l++) {
      u.blk0[l] = nbu;
      const long long cols = (long long)(1 << (3 * (l - 1))) * nrhs;
      nbu += upper_split * (8 / ss_up) * (int)((cols + cpb_up - 1) / cpb_up);
    }
    u.blk0[maxlev] = nbu;
    d.minlev = maxlev;
    d.blk0[maxlev] = 0;
    d.blk0[maxlev + 1] = a.blk0[maxlev + 1] - a.blk0[maxlev];
    g->part_blocks[0] = nbu;
    g->part_blocks[1] = d.blk0[maxlev + 1];
    std::vector<int> su;
    step_table(su, ss_up, u.nsteps);
    steps_up_off = steps.size();
    steps.insert(steps.end(), su.begin(), su.end());
  }
  int rc = BBFMM_OK;
  auto fail = [&](hipError_t e, const char *what) {
    rc = set_error(BBFMM_E_HIP, "dense-grid M2L: %s: %s", what, hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&g->d_tab, sizeof(int4) * tab.size())) != hipSuccess) fail(e, "hipMalloc");
  if (rc == BBFMM_OK && (e = hipMemcpyAsync(g->d_tab, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
    fail(e, "copy");
  if (rc == BBFMM_OK && (e = hipMalloc(&g->d_steps, sizeof(int) * steps.size())) != hipSuccess) fail(e, "hipMalloc");
  if (rc == BBFMM_OK && (e = hipMemcpyAsync(g->d_steps, steps.data(), sizeof(int) * steps.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
    fail(e, "copy");
  // operators in the form's layout
  if (rc == BBFMM_OK) {
    if (form == BBFMM_M2L_PLAIN && r != 256) {
      // r = 32 and any rank: the caller's column-major operators as they are
      if ((e = hipMalloc(&g->d_K, sizeof(real) * (size_t)kNumOps * r * r)) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpyAsync(g->d_K, K, sizeof(real) * (size_t)kNumOps * r * r, hipMemcpyHostToDevice, s)) !=
               hipSuccess)
        fail(e, "copy");
      if (rc == BBFMM_OK && (e = hipStreamSynchronize(s)) != hipSuccess) fail(e, "sync");
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_Mg0, sizeof(real) * slots * nrhs * r)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_Mg0, 0, sizeof(real) * slots * nrhs * r, s)) != hipSuccess) fail(e, "memset");
    } else if (form == BBFMM_M2L_PLAIN) {
      std::vector<real4> K4((size_t)kNumOps * r * r / 4);
      for (int k = 0; k < kNumOps; k++)
        for (int jb = 0; jb < r / 4; jb++)
          for (int i = 0; i < r; i++) {
            const real *Kk = K + (size_t)k * r * r;
            K4[((size_t)k * (r / 4) + jb) * r + i] = real4{Kk[(size_t)(4 * jb) * r + i], Kk[(size_t)(4 * jb + 1) * r + i],
                                                          Kk[(size_t)(4 * jb + 2) * r + i], Kk[(size_t)(4 * jb + 3) * r + i]};
          }
      if ((e = hipMalloc(&g->d_K, sizeof(real4) * K4.size())) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpyAsync(g->d_K, K4.data(), sizeof(real4) * K4.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
        fail(e, "copy");
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_Mg0, sizeof(real) * slots * nrhs * r)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_Mg0, 0, sizeof(real) * slots * nrhs * r, s)) != hipSuccess) fail(e, "memset");
    }
#ifdef SINGLE
    else if (form == BBFMM_M2L_BF16X2) {
      // [k][kstep][hi/lo][lane][8]: lane l holds row l%32, k = 16*kstep + 8*(l/32) + e
      // one extra all-zero operator (kZeroOp) after the 316
      std::vector<__bf16> Kb((size_t)(kNumOps + 1) * 2 * 2 * 64 * 8, (__bf16)0.0f);
      for (int k = 0; k < kNumOps; k++)
        for (int ks = 0; ks < 2; ks++)
          for (int lane = 0; lane < 64; lane++)
            for (int q = 0; q < 8; q++) {
              const int row = lane & 31, col = 16 * ks + 8 * (lane >> 5) + q;
              const float v = K[(size_t)k * 32 * 32 + (size_t)col * 32 + row];
              const __bf16 h = (__bf16)v;
              const __bf16 lo = (__bf16)(v - (float)h);
              Kb[((((size_t)k * 2 + ks) * 2 + 0) * 64 + lane) * 8 + q] = h;
              Kb[((((size_t)k * 2 + ks) * 2 + 1) * 64 + lane) * 8 + q] = lo;
            }
      if ((e = hipMalloc(&g->d_K, sizeof(__bf16) * Kb.size())) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpyAsync(g->d_K, Kb.data(), sizeof(__bf16) * Kb.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
        fail(e, "copy");
      a.kbytes = (int)(sizeof(__bf16) * Kb.size());
      const size_t nb2 = sizeof(__bf16) * slots * nrhs * r;
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_Mg0, nb2)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_Mg1, nb2)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_Mg0, 0, nb2, s)) != hipSuccess) fail(e, "memset");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_Mg1, 0, nb2, s)) != hipSuccess) fail(e, "memset");
    } else if (form == BBFMM_M2L_FP16X3) {
      // [k][kstep][term][lane][8]: lane l holds row l%32, k = 16*kstep + 8*(l/32) + e,
      // scaled by 2^(15 - e) for the largest |K| = f 2^e
      float kmax = 0.0f;
      for (size_t i = 0; i < (size_t)kNumOps * 32 * 32; i++) kmax = std::max(kmax, std::fabs((float)K[i]));
      int ke = 0;
      (void)std::frexp(kmax, &ke);
      const float ksc = std::ldexp(1.0f, 15 - ke);
      a.kscale_inv = std::ldexp(1.0f, ke - 15);
      std::vector<_Float16> Kh((size_t)kNumOps * 2 * 3 * 64 * 8);
      for (int k = 0; k < kNumOps; k++)
        for (int ks = 0; ks < 2; ks++)
          for (int lane = 0; lane < 64; lane++)
            for (int q = 0; q < 8; q++) {
              const int row = lane & 31, col = 16 * ks + 8 * (lane >> 5) + q;
              const float v = (float)K[(size_t)k * 32 * 32 + (size_t)col * 32 + row] * ksc;
              const _Float16 t1 = (_Float16)v;
              const float r1 = v - (float)t1;
              const _Float16 t2 = (_Float16)r1;
              const _Float16 t3 = (_Float16)(r1 - (float)t2);
              const size_t b = (((size_t)k * 2 + ks) * 3) * 64 * 8 + (size_t)lane * 8 + q;
              Kh[b] = t1;
              Kh[b + 64 * 8] = t2;
              Kh[b + 2 * 64 * 8] = t3;
            }
      if ((e = hipMalloc(&g->d_K, sizeof(_Float16) * Kh.size())) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpyAsync(g->d_K, Kh.data(), sizeof(_Float16) * Kh.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
        fail(e, "copy");
      // the host copy must outlive the asynchronous copy
      if (rc == BBFMM_OK && (e = hipStreamSynchronize(s)) != hipSuccess) fail(e, "sync");
      const size_t nb2 = sizeof(_Float16) * slots * nrhs * r;
      void **grids[3] = {&g->d_Mg0, &g->d_Mg1, &g->d_Mg2};
      for (void **gp : grids) {
        if (rc == BBFMM_OK && (e = hipMalloc(gp, nb2)) != hipSuccess) fail(e, "hipMalloc");
        if (rc == BBFMM_OK && (e = hipMemsetAsync(*gp, 0, nb2, s)) != hipSuccess) fail(e, "memset");
      }
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_lvl_max, sizeof(unsigned) * (kMaxLev + 1))) != hipSuccess)
        fail(e, "hipMalloc");
    } else if (r == 256) {
      // FP32, r = 256: [k][k-chunk][row tile][lane][16]: lane l holds row
      // 32*tile + l%32 and columns 32*chunk + 16*(l/32) + 0..15
      std::vector<float> Kf((size_t)kNumOps * 256 * 256);
      for (int k = 0; k < kNumOps; k++)
        for (int kc = 0; kc < 8; kc++)
          for (int t = 0; t < 8; t++)
            for (int lane = 0; lane < 64; lane++)
              for (int ks = 0; ks < 16; ks++) {
                const int row = 32 * t + (lane & 31), col = 32 * kc + 16 * (lane >> 5) + ks;
                Kf[((((size_t)k * 8 + kc) * 8 + t) * 64 + lane) * 16 + ks] =
                    K[(size_t)k * 256 * 256 + (size_t)col * 256 + row];
              }
      if ((e = hipMalloc(&g->d_K, sizeof(float) * Kf.size())) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpyAsync(g->d_K, Kf.data(), sizeof(float) * Kf.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
        fail(e, "copy");
      // the host copy must outlive the asynchronous copy
      if (rc == BBFMM_OK && (e = hipStreamSynchronize(s)) != hipSuccess) fail(e, "sync");
      const size_t nb2 = sizeof(float) * slots * nrhs * r;
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_Mg0, nb2)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_Mg0, 0, nb2, s)) != hipSuccess) fail(e, "memset");
    } else {
      // FP32: [k][lane][16]: lane l holds row l%32, k = 16*(l/32) + ks
      std::vector<float> Kf((size_t)kNumOps * 64 * 16);
      for (int k = 0; k < kNumOps; k++)
        for (int lane = 0; lane < 64; lane++)
          for (int ks = 0; ks < 16; ks++) {
            const int row = lane & 31, col = 16 * (lane >> 5) + ks;
            Kf[((size_t)k * 64 + lane) * 16 + ks] = K[(size_t)k * 32 * 32 + (size_t)col * 32 + row];
          }
      if ((e = hipMalloc(&g->d_K, sizeof(float) * Kf.size())) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpyAsync(g->d_K, Kf.data(), sizeof(float) * Kf.size(), hipMemcpyHostToDevice, s)) != hipSuccess)
        fail(e, "copy");
      const size_t nb2 = sizeof(float) * slots * nrhs * r;
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_Mg0, nb2)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_Mg0, 0, nb2, s)) != hipSuccess) fail(e, "memset");
    }
#endif
  }
  if (rc == BBFMM_OK && g->nparts == 2 && g->part[0].nsplit > 1) {
    const size_t nx = (size_t)(g->part[0].nsplit - 1) * nrhs * g->part[0].up_ncell * 32;
    if ((e = hipMalloc(&g->d_Lx, sizeof(float) * nx)) != hipSuccess) fail(e, "hipMalloc");
  }
  if (rc == BBFMM_OK && g->nparts == 1 && a.nsplit > 1) {  // the whole launch in step parts
    const size_t nx = (size_t)(a.nsplit - 1) * nrhs * a.up_ncell * 32;
    if ((e = hipMalloc(&g->d_Lx, sizeof(float) * nx)) != hipSuccess) fail(e, "hipMalloc");
  }
  if (rc == BBFMM_OK && (e = hipStreamSynchronize(s)) != hipSuccess) fail(e, "sync");
  if (rc != BBFMM_OK) {
    grid_m2l_destroy(g);
    return rc;
  }
  a.tab = g->d_tab;
  a.steps = g->d_steps;
  a.Ka = g->d_K;
  a.Mg0 = g->d_Mg0;
  a.Mg1 = g->d_Mg1;
  a.Mg2 = g->d_Mg2;
  a.lvl_max = g->d_lvl_max;
  a.Lx = g->nparts == 1 ? g->d_Lx : nullptr;
  if (g->nparts == 2) {
    // everything from the final args (operator and grid pointers and sizes),
    // then each part's own levels, blocks, class-group size, tiles and steps
    for (int pi = 0; pi < 2; pi++) {
      const GridArgs own = g->part[pi];
      GridArgs &q = g->part[pi];
      q = a;
      q.minlev = own.minlev;
      q.maxlev = own.maxlev;
      for (int l = 0; l < kMaxLev + 2; l++) q.blk0[l] = own.blk0[l];
      q.ss = own.ss;
      q.ss_tpw = own.ss_tpw;
      q.ss_pair = own.ss_pair;
      q.wave_cols = own.wave_cols;
      for (int c = 0; c < 8; c++) q.nsteps[c] = own.nsteps[c];
      q.steps = g->d_steps + (pi == 0 ? steps_up_off : 0);
      q.nsplit = own.nsplit;
      q.up_cell0 = own.up_cell0;
      q.up_ncell = own.up_ncell;
      q.Lx = pi == 0 ? g->d_Lx : nullptr;
    }
  }
  *out = g;
  return BBFMM_OK;
}

void grid_m2l_set_flat(GridM2L *g, int flat)
{
  if (g) g->flat = flat != 0;
}

void grid_m2l_destroy(GridM2L *g)
{
  if (!g) return;
  if (g->d_tab) (void)hipFree(g->d_tab);
  if (g->d_steps) (void)hipFree(g->d_steps);
  if (g->d_K) (void)hipFree(g->d_K);
  if (g->d_Mg0) (void)hipFree(g->d_Mg0);
  if (g->d_Mg1) (void)hipFree(g->d_Mg1);
  if (g->d_Mg2) (void)hipFree(g->d_Mg2);
  if (g->d_lvl_max) (void)hipFree(g->d_lvl_max);
  if (g->d_Lx) (void)hipFree(g->d_Lx);
  delete g;
}

#ifdef SINGLE
// grids below 2 GB take the scalar-offset buffer-load form of the kernel
template <int CLS>
static void launch_ss(const GridM2L *g, const GridArgs &a, int nblocks, hipStream_t s)
{
  const dim3 grid((nblocks + 7) & ~7);
  const bool buf = a.gbytes > 0 && !g->flat;
  if (a.ss_pair) {
    if (buf)
      hipLaunchKernelGGL(k_m2l_mfma_ss_pair<true>, grid, dim3(512), 0, s, a);
    else
      hipLaunchKernelGGL(k_m2l_mfma_ss_pair<false>, grid, dim3(512), 0, s, a);
  } else if (CLS <= 4 && a.ss_tpw == 2) {  // two tiles per wave: 2 or 4 classes
    constexpr int C = CLS <= 4 ? CLS : 4;
    if (buf)
      hipLaunchKernelGGL((k_m2l_mfma_ss<C, 2, true>), grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_m2l_mfma_ss<C, 2, false>), grid, dim3(256), 0, s, a);
  } else if (buf)
    hipLaunchKernelGGL((k_m2l_mfma_ss<CLS, 1, true>), grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_m2l_mfma_ss<CLS, 1, false>), grid, dim3(256), 0, s, a);
}
#endif

int grid_m2l_leaf_grids(const GridM2L *g, void **hi, void **lo, int *slot0)
{
#ifdef SINGLE
  if (g && g->form == BBFMM_M2L_BF16X2 && g->nrhs == 16 && g->args.minlev <= g->args.maxlev) {
    *hi = g->d_Mg0;
    *lo = g->d_Mg1;
    *slot0 = g->args.lvl_slot0[g->args.maxlev];
    return 1;
  }
#endif
  (void)g;
  (void)hi;
  (void)lo;
  (void)slot0;
  return 0;
}

int grid_m2l_run(GridM2L *g, const real *M, real *L, int ncell, hipStream_t s, hipEvent_t mid, int deep_in_grid)
{
  GridArgs a = g->args;
  if (ncell != a.ncell) return set_error(BBFMM_E_ARG, "dense-grid M2L: ncell %d != %d", ncell, a.ncell);
  a.L = L;
  const long long cells = (long long)(a.levsta[a.maxlev + 1] - a.levsta[a.minlev]) * a.nrhs;
  const int sb = (int)((cells * (a.r / 8) + 255) / 256);
  switch (g->form) {
  case BBFMM_M2L_PLAIN:
    if (a.r % 8 == 0)
      hipLaunchKernelGGL(k_grid_scatter<BBFMM_M2L_PLAIN>, dim3(sb), dim3(256), 0, s, a, M);
    else
      hipLaunchKernelGGL((k_grid_scatter<BBFMM_M2L_PLAIN, 1>), dim3((int)((cells * a.r + 255) / 256)), dim3(256), 0,
                         s, a, M);
    BB_HIP(hipGetLastError());
    if (mid) BB_HIP(hipEventRecord(mid, s));
    if (a.r == 32)
      hipLaunchKernelGGL(k_m2l_grid_plain32, dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    else if (a.r != 256)
      hipLaunchKernelGGL(k_m2l_grid_plain_any, dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL(k_m2l_grid_plain<256>, dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
#ifdef SINGLE
  case BBFMM_M2L_BF16X2:
    if (deep_in_grid) {  // the levels above the deepest only
      if (a.maxlev > a.minlev) {
        GridArgs u = a;
        u.maxlev = a.maxlev - 1;
        const long long ucells = (long long)(a.levsta[a.maxlev] - a.levsta[a.minlev]) * a.nrhs;
        hipLaunchKernelGGL(k_grid_scatter<BBFMM_M2L_BF16X2>, dim3((int)((ucells * (a.r / 8) + 255) / 256)), dim3(256),
                           0, s, u, M);
      }
    } else {
      hipLaunchKernelGGL(k_grid_scatter<BBFMM_M2L_BF16X2>, dim3(sb), dim3(256), 0, s, a, M);
    }
    BB_HIP(hipGetLastError());
    if (mid) BB_HIP(hipEventRecord(mid, s));
    if (a.ss) {  // one launch, or the upper levels then the deepest
      const int np = g->nparts;
      for (int pi = 0; pi < np; pi++) {
        GridArgs pa = np == 2 ? g->part[pi] : a;
        pa.L = L;
        const int nbk = np == 2 ? g->part_blocks[pi] : g->nblocks;
        if (nbk == 0) continue;
        if (pa.ss == 8)  // rounded up to whole rounds of the 8 XCDs; the extra blocks exit
          launch_ss<8>(g, pa, nbk, s);
        else if (pa.ss == 4)
          launch_ss<4>(g, pa, nbk, s);
        else
          launch_ss<2>(g, pa, nbk, s);
        if (pa.nsplit > 1) {  // the upper part's steps in parts: add parts 1.. to L
          const size_t n4 = (size_t)pa.nrhs * pa.up_ncell * 8;
          hipLaunchKernelGGL(k_m2l_add_parts, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s, pa);
        }
      }
    } else if (a.wave_cols == 64)
      hipLaunchKernelGGL((k_m2l_mfma<BBFMM_M2L_BF16X2, 2>), dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_m2l_mfma<BBFMM_M2L_BF16X2, 1>), dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
  case BBFMM_M2L_FP16X3:
    // per-level scales first (a fill kernel, not a memset: the solve may be captured)
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, a.lvl_max, kMaxLev + 1);
    hipLaunchKernelGGL(k_level_absmax, dim3(256, a.maxlev - a.minlev + 1), dim3(256), 0, s, a, M);
    hipLaunchKernelGGL(k_grid_scatter<BBFMM_M2L_FP16X3>, dim3(sb), dim3(256), 0, s, a, M);
    BB_HIP(hipGetLastError());
    if (mid) BB_HIP(hipEventRecord(mid, s));
    if (a.wave_cols == 64)
      hipLaunchKernelGGL((k_m2l_mfma<BBFMM_M2L_FP16X3, 2>), dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_m2l_mfma<BBFMM_M2L_FP16X3, 1>), dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
  case BBFMM_M2L_FP32MFMA:
    hipLaunchKernelGGL(k_grid_scatter<BBFMM_M2L_FP32MFMA>, dim3(sb), dim3(256), 0, s, a, M);
    BB_HIP(hipGetLastError());
    if (mid) BB_HIP(hipEventRecord(mid, s));
    if (a.r == 256)  // rounded up to whole rounds of the 8 XCDs; the extra blocks exit
      hipLaunchKernelGGL(k_m2l_fp32_r256, dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    else if (a.wave_cols == 64)
      hipLaunchKernelGGL((k_m2l_mfma<BBFMM_M2L_FP32MFMA, 2>), dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_m2l_mfma<BBFMM_M2L_FP32MFMA, 1>), dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
#endif
  default: return set_error(BBFMM_E_ARG, "dense-grid M2L: unknown form %d", g->form);
  }
}

// ---------------------------------------------------------------- list MFMA (sparse trees)
// For trees that store only occupied cells (batches of instances) the dense
// grids do not exist; the same (class, offset) products are formed on
// gathered columns.  The field cells of levels >= 2 are grouped by sibling
// class s; for class s and its 189 offsets v (the dense-grid table order)
// vsrc[F * 189 + j] names the source cell F + v, or the zero cell ncell when
// that cell holds no points.  A block (4 waves x 32 columns, columns = (field
// cell, vector) pairs of one class) multiplies K_v by the gathered source
// columns for every v: the operator is shared by the whole block, as in the
// dense-grid kernel.  M is copied per solve to [rhs][ncell + 1][r] (bf16 hi
// and lo for the split form) with a zero last cell.

struct ListArgs {
  int ncell, nrhs;
  int cls_off[9];   // class c cells: cls_cells[cls_off[c] .. cls_off[c + 1])
  int blk0[9];      // block count of class c: blk0[c + 1] - blk0[c]
  // logical block -> 8 * (block within its class) + class: the classes are
  // interleaved in proportion to their sizes, so that the blocks running at
  // one time (and the contiguous range each XCD takes) cover one region of
  // the forest for all 8 classes and share its source cells in L2
  const int *blk_tab;
  const int *cls_cells;
  const int *vsrc;  // [ncell * 189]
  const int4 *tab;  // [8][189] (operator, ...)
  // per wave (4 per block): bit v of words 0..2 set when some column of the
  // wave has a non-empty source at offset v
  const unsigned long long *mask;  // [4 * blocks][4]
  // one vector per cell (nrhs = 1): per wave, the source cells of its 32
  // columns at each offset, so that one offset's indices are one 128-byte
  // line (ncell for columns past the end); null otherwise (the columns of a
  // wave then hold at most 16 cells, read from vsrc)
  const int *vsw;  // [4 * blocks][189][32]
  const void *Ka;
  const void *M0, *M1;
  real *L;
  int kbytes;  // bytes of Ka
  int mbytes;  // bytes of M0 (and M1) when below 2 GB (scalar-offset buffer loads), else 0
};

struct ListM2L {
  int form = 0, nrhs = 1, ncell = 0, nblocks = 0;
  int flat = 0;  // 1: pointer-form loads even on arrays below 2 GB
  ListArgs args{};
  int4 *d_tab = nullptr;
  int *d_blk = nullptr;
  unsigned long long *d_mask = nullptr;
  int *d_vsw = nullptr;
  void *d_K = nullptr, *d_M0 = nullptr, *d_M1 = nullptr;
};

// Columns of wave w of block b (in the list kernel's logical block order):
// class s and the column range.  Shared by the kernel and its mask set-up.
__device__ __forceinline__ int2 list_block(const ListArgs &g, int lb)
{
  const int code = g.blk_tab[lb];
  return make_int2(code & 7, code >> 3);  // class, block within the class
}

// One wave per wave of the list kernel: the union over its valid columns of
// the non-empty offsets (vsrc != ncell), as a 189-bit mask, and the wave's
// column-contiguous copy of the source table.
__global__ void __launch_bounds__(64) k_m2l_list_mask(ListArgs g, unsigned long long *mask, int *vsw)
{
  const int gw = blockIdx.x;  // 4 * logical block + wave
  const int lb = gw >> 2, wave = gw & 3;
  const int2 sb = list_block(g, lb);
  const int s = sb.x, grp = sb.y;
  const int ncols = (g.cls_off[s + 1] - g.cls_off[s]) * g.nrhs;
  const int lane = threadIdx.x;
  const int col = (grp * 4 + wave) * 32 + (lane & 31);
  const bool valid = lane < 32 && col < ncols;
  const int F = valid ? g.cls_cells[g.cls_off[s] + col / g.nrhs] : 0;
  const int *vs = g.vsrc + (size_t)F * kOffsets;
  unsigned long long m[3] = {0, 0, 0};
  int *out = vsw ? vsw + (size_t)gw * kOffsets * 32 + lane : nullptr;
  for (int v = 0; v < kOffsets; v++) {
    const int S = valid ? vs[v] : g.ncell;
    const bool any = __ballot(S != g.ncell) != 0;
    if (any) m[v >> 6] |= 1ull << (v & 63);
    if (out && lane < 32) out[v * 32] = S;
  }
  if (lane < 4) mask[(size_t)gw * 4 + lane] = lane < 3 ? m[lane] : 0ull;
}

template <int FORM>
__global__ void __launch_bounds__(256) k_m2l_list_prep(ListArgs g, const real *__restrict__ M)
{
  const long long total = (long long)g.nrhs * g.ncell * 4;  // 8 coefficients per thread
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int ch = (int)(t & 3);
  const long long cr = t >> 2;
  const int A = (int)(cr % g.ncell), rh = (int)(cr / g.ncell);
  const real *src = M + ((size_t)rh * g.ncell + A) * 32 + ch * 8;
  const size_t dst = ((size_t)rh * (g.ncell + 1) + A) * 32 + ch * 8;
  if (FORM == BBFMM_M2L_BF16X2) {
    bf16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const float v = (float)src[e];
      const __bf16 h = (__bf16)v;
      hi[e] = h;
      lo[e] = (__bf16)(v - (float)h);
    }
    *reinterpret_cast<bf16x8 *>((__bf16 *)g.M0 + dst) = hi;
    *reinterpret_cast<bf16x8 *>((__bf16 *)g.M1 + dst) = lo;
  } else {
    float *d = (float *)g.M0 + dst;
#pragma unroll
    for (int e = 0; e < 8; e++) d[e] = (float)src[e];
  }
}

template <int FORM>
__global__ void __launch_bounds__(256) k_m2l_list_mfma(ListArgs g)
{
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // XCD-aware order: the launch is a multiple of 8 blocks and hardware block b
  // runs on XCD b % 8, which takes a contiguous range of logical blocks lb
  // (one region of the forest, all classes) so that neighbouring targets,
  // which gather the same sources, share one L2.
  const int nb = g.blk0[8];
  const int b = blockIdx.x;
  const int lb = (b & 7) * ((nb + 7) >> 3) + (b >> 3);
  if (lb >= nb) return;
  const int2 sb = list_block(g, lb);
  const int s = sb.x, grp = sb.y;
  const int ncols = (g.cls_off[s + 1] - g.cls_off[s]) * g.nrhs;
  const int half = lane >> 5;
  // columns past the end compute a copy of the class's first column and are not stored
  const int col = (grp * 4 + wave) * 32 + (lane & 31);
  const bool valid = col < ncols;
  const int cc = valid ? col : 0;
  const int F = g.cls_cells[g.cls_off[s] + cc / g.nrhs];
  const int rh = cc % g.nrhs;
  const int4 *tab = g.tab + s * kOffsets;
  const size_t rbase = (size_t)rh * (g.ncell + 1);
  const int gw = __builtin_amdgcn_readfirstlane(lb * 4 + wave);
  // source of column (lane % 32) at offset v: the wave-major copy, or vsrc
  const int *vsl = g.vsw ? g.vsw + (size_t)gw * kOffsets * 32 + (lane & 31) : g.vsrc + (size_t)F * kOffsets;
  const int vst = g.vsw ? 32 : 1;
  auto vs = [&](int v) { return vsl[v * vst]; };
  f32x16 acc;
#pragma unroll
  for (int e = 0; e < 16; e++) acc[e] = 0.0f;
  // Only the offsets at which some column of the wave has a non-empty source
  // are visited (the others would add K_v times the zero cell); they come in
  // increasing order from the wave's 189-bit mask.  Software pipeline: the
  // operator fragments and the gathered source column of the next offset are
  // in flight while the MFMAs of the current one run, and the source index
  // of the one after that a step earlier (two register stages, the loop
  // unrolled by two so that no stage is copied).
  struct Stage {
    f32x4 a[4], b[4];  // BF16X2: a[0..3] = hi0, lo0, hi1, lo1 of K_v; b likewise of the source
  };
  auto load = [&](int v, int S, Stage &t) {
    const int k = tab[v].x;
    if (FORM == BBFMM_M2L_BF16X2) {
      const f32x4 *Ka = (const f32x4 *)((const bf16x8 *)g.Ka + (size_t)k * 4 * 64 + lane);
#pragma unroll
      for (int q = 0; q < 4; q++) t.a[q] = Ka[q * 64];
      const size_t o = (rbase + S) * 32 + half * 8;
      t.b[0] = *(const f32x4 *)((const __bf16 *)g.M0 + o);
      t.b[1] = *(const f32x4 *)((const __bf16 *)g.M1 + o);
      t.b[2] = *(const f32x4 *)((const __bf16 *)g.M0 + o + 16);
      t.b[3] = *(const f32x4 *)((const __bf16 *)g.M1 + o + 16);
    } else {
      const f32x4 *Ka = (const f32x4 *)g.Ka + ((size_t)k * 64 + lane) * 4;
      const f32x4 *bp = (const f32x4 *)((const float *)g.M0 + (rbase + S) * 32 + half * 16);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        t.a[q] = Ka[q];
        t.b[q] = bp[q];
      }
    }
  };
  auto mma = [&](const Stage &t) {
    if (FORM == BBFMM_M2L_BF16X2) {
      typedef __attribute__((ext_vector_type(8))) __bf16 b8;
      const b8 a0h = __builtin_bit_cast(b8, t.a[0]), a0l = __builtin_bit_cast(b8, t.a[1]);
      const b8 a1h = __builtin_bit_cast(b8, t.a[2]), a1l = __builtin_bit_cast(b8, t.a[3]);
      const b8 b0h = __builtin_bit_cast(b8, t.b[0]), b0l = __builtin_bit_cast(b8, t.b[1]);
      const b8 b1h = __builtin_bit_cast(b8, t.b[2]), b1l = __builtin_bit_cast(b8, t.b[3]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, b0h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, b0l, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0l, b0h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, b1h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, b1l, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1l, b1h, acc, 0, 0, 0);
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int w = 0; w < 4; w++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(t.a[q][w], t.b[q][w], acc, 0, 0, 0);
    }
  };
  unsigned long long m0 = g.mask[(size_t)gw * 4], m1 = g.mask[(size_t)gw * 4 + 1], m2 = g.mask[(size_t)gw * 4 + 2];
  // next visited offset, or -1 past the last (scalar selects, no branches)
  auto pop = [&]() -> int {
    const bool a0 = m0 != 0, a1 = m1 != 0, a2 = m2 != 0;
    const int t0 = a0 ? __builtin_ctzll(m0) : 0, t1 = a1 ? __builtin_ctzll(m1) : 0,
              t2 = a2 ? __builtin_ctzll(m2) : 0;
    const int v = a0 ? t0 : a1 ? 64 + t1 : a2 ? 128 + t2 : -1;
    m0 = a0 ? m0 & (m0 - 1) : m0;
    m1 = !a0 && a1 ? m1 & (m1 - 1) : m1;
    m2 = !a0 && !a1 ? m2 & (m2 - 1) : m2;
    return v;
  };
  // Loads are unconditional (offset 0 stands in past the last) so that no
  // branch separates a load from the MFMAs that wait for it; the source
  // indices are read an iteration (two offsets) before their columns.
  Stage X, Y;
  int vy, vz, vw;
  const int vx = pop();
  vy = pop();
  vz = pop();
  vw = pop();
  if (vx >= 0) {
    load(vx, vs(vx), X);
    load(max(vy, 0), vs(max(vy, 0)), Y);
    int Sz = vs(max(vz, 0)), Sw = vs(max(vw, 0));
    // X holds an offset, Y the next one (vy, when vy >= 0); vz and vw follow,
    // with sources Sz and Sw
    for (;;) {
      const int vu = pop(), vt = pop();
      const int Su = vs(max(vu, 0)), St = vs(max(vt, 0));
      __builtin_amdgcn_sched_barrier(0);
      mma(X);
      load(max(vz, 0), Sz, X);
      if (vy < 0) break;
      mma(Y);
      load(max(vw, 0), Sw, Y);
      if (vz < 0) break;
      vy = vw;
      vz = vu;
      vw = vt;
      Sz = Su;
      Sw = St;
    }
  }
  if (!valid) return;
  float *Lp = (float *)g.L + ((size_t)rh * g.ncell + F) * 32 + 4 * half;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    f32x4 o = {acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]};
    *reinterpret_cast<f32x4 *>(Lp + 8 * q) = o;
  }
}

// slot j of offset v in class s's order (offset_table), or -1
static void slot_table(std::vector<int> &slot)
{
  slot.assign(8 * 343, -1);
  for (int s = 0; s < 8; s++) {
    const int sb[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1};
    int nv = 0;
    for (int vx = -2 - sb[0]; vx <= 3 - sb[0]; vx++)
      for (int vy = -2 - sb[1]; vy <= 3 - sb[1]; vy++)
        for (int vz = -2 - sb[2]; vz <= 3 - sb[2]; vz++) {
          if (vx >= -1 && vx <= 1 && vy >= -1 && vy <= 1 && vz >= -1 && vz <= 1) continue;
          slot[s * 343 + 49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)] = nv++;
        }
  }
}

void slot_operators(std::vector<int> &op)
{
  std::vector<int4> tab;
  offset_table(tab);
  op.resize(tab.size());
  for (size_t i = 0; i < tab.size(); i++) op[i] = tab[i].x;
}

int class_slot(int s, int vx, int vy, int vz)
{
  static const std::vector<int> slot = [] {  // built once, thread-safe
    std::vector<int> t;
    slot_table(t);
    return t;
  }();
  if (vx < -3 || vx > 3 || vy < -3 || vy > 3 || vz < -3 || vz > 3) return -1;
  return slot[s * 343 + 49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)];
}

// Split-BF16 list M2L with the operator shared by the block: the 4 waves of
// a block visit the union of their offset masks in step, and each offset's
// operator fragments (4 KB) are loaded once per block (one 1 KB fragment per
// wave, through registers into LDS, double-buffered) instead of once per
// wave; the vector-memory pipe, which the per-wave kernel saturates with
// 8 KB per offset and wave, then moves 5 KB.  Per step: the next offset's
// fragment and source column are loaded while the current offset's six
// products run; one barrier per step.  Same columns, order of products and
// results as k_m2l_list_mfma<BBFMM_M2L_BF16X2> (the extra offsets a wave
// visits add K_v times its zero sources).
template <bool BUF>
__global__ void __launch_bounds__(256) k_m2l_list_lds(ListArgs g)
{
  __shared__ bf16x8 sK[2][4][64];  // [buffer][fragment: k-step x hi/lo][lane]
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nb = g.blk0[8];
  const int b = blockIdx.x;
  const int lb = (b & 7) * ((nb + 7) >> 3) + (b >> 3);
  if (lb >= nb) return;  // whole block
  const int2 sb = list_block(g, lb);
  const int s = sb.x, grp = sb.y;
  const int ncols = (g.cls_off[s + 1] - g.cls_off[s]) * g.nrhs;
  const int half = lane >> 5;
  const int col = (grp * 4 + wave) * 32 + (lane & 31);
  const bool valid = col < ncols;
  const int cc = valid ? col : 0;
  const int F = g.cls_cells[g.cls_off[s] + cc / g.nrhs];
  const int rh = cc % g.nrhs;
  const int4 *tab = g.tab + s * kOffsets;
  const size_t rbase = (size_t)rh * (g.ncell + 1);
  const int gw = __builtin_amdgcn_readfirstlane(lb * 4 + wave);
  // source of column (lane % 32) at offset v: the wave-major copy, or vsrc
  const int *vsl = g.vsw ? g.vsw + (size_t)gw * kOffsets * 32 + (lane & 31) : g.vsrc + (size_t)F * kOffsets;
  const int vst = g.vsw ? 32 : 1;
  auto vs = [&](int v) { return vsl[v * vst]; };
  // the block's union of the 4 waves' masks
  const unsigned long long *mk = g.mask + (size_t)lb * 16;
  unsigned long long m0 = mk[0] | mk[4] | mk[8] | mk[12];
  unsigned long long m1 = mk[1] | mk[5] | mk[9] | mk[13];
  unsigned long long m2 = mk[2] | mk[6] | mk[10] | mk[14];
  auto pop = [&]() -> int {
    const bool a0 = m0 != 0, a1 = m1 != 0, a2 = m2 != 0;
    const int t0 = a0 ? __builtin_ctzll(m0) : 0, t1 = a1 ? __builtin_ctzll(m1) : 0,
              t2 = a2 ? __builtin_ctzll(m2) : 0;
    const int v = a0 ? t0 : a1 ? 64 + t1 : a2 ? 128 + t2 : -1;
    m0 = a0 ? m0 & (m0 - 1) : m0;
    m1 = !a0 && a1 ? m1 & (m1 - 1) : m1;
    m2 = !a0 && !a1 ? m2 & (m2 - 1) : m2;
    return v;
  };
  // this wave's fragment of offset v's operator, and a column's source.
  // BUF: buffer loads, the operator's index a scalar offset over a fixed
  // per-lane part, the source's offset one 32-bit lane value (grids below
  // 2 GB), so that no 64-bit address arithmetic goes into a step
  const __amdgpu_buffer_rsrc_t rK = __builtin_amdgcn_make_buffer_rsrc((void *)g.Ka, 0, g.kbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rM0 = __builtin_amdgcn_make_buffer_rsrc((void *)g.M0, 0, g.mbytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rM1 = __builtin_amdgcn_make_buffer_rsrc((void *)g.M1, 0, g.mbytes, 0x00020000);
  auto load_k = [&](int v) {
    const int k = tab[v].x;
    if (BUF)
      return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rK, (wave * 64 + lane) * 16, k * 4096, 0));
    return *((const f32x4 *)((const bf16x8 *)g.Ka + ((size_t)k * 4 + wave) * 64 + lane));
  };
  auto load_b = [&](int S, f32x4 *t) {
    if (BUF) {
      const int o = (((int)rbase + S) * 32 + half * 8) * 2;
      t[0] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rM0, o, 0, 0));
      t[1] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rM1, o, 0, 0));
      t[2] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rM0, o + 32, 0, 0));
      t[3] = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rM1, o + 32, 0, 0));
      return;
    }
    const size_t o = (rbase + S) * 32 + half * 8;
    t[0] = *(const f32x4 *)((const __bf16 *)g.M0 + o);
    t[1] = *(const f32x4 *)((const __bf16 *)g.M1 + o);
    t[2] = *(const f32x4 *)((const __bf16 *)g.M0 + o + 16);
    t[3] = *(const f32x4 *)((const __bf16 *)g.M1 + o + 16);
  };
  typedef __attribute__((ext_vector_type(8))) __bf16 b8;
  f32x16 acc;
#pragma unroll
  for (int e = 0; e < 16; e++) acc[e] = 0.0f;
  int v = pop();
  if (v >= 0) {  // uniform over the block
    f32x4 bc[4];
    {
      const f32x4 kf = load_k(v);
      load_b(vs(v), bc);
      *reinterpret_cast<f32x4 *>(&sK[0][wave][lane]) = kf;
    }
    int vn = pop();
    int Sn = vs(max(vn, 0));
    __syncthreads();
    for (int buf = 0;; buf ^= 1) {
      const b8 a0h = __builtin_bit_cast(b8, sK[buf][0][lane]), a0l = __builtin_bit_cast(b8, sK[buf][1][lane]);
      const b8 a1h = __builtin_bit_cast(b8, sK[buf][2][lane]), a1l = __builtin_bit_cast(b8, sK[buf][3][lane]);
      // next offset (offset 0 stands in past the last): operator fragment and
      // source column, issued before this offset's products (kept there)
      const f32x4 kn = load_k(max(vn, 0));
      f32x4 bn[4];
      load_b(Sn, bn);
      const int vnn = pop();
      const int Snn = vs(max(vnn, 0));
      __builtin_amdgcn_sched_barrier(0);
      const b8 b0h = __builtin_bit_cast(b8, bc[0]), b0l = __builtin_bit_cast(b8, bc[1]);
      const b8 b1h = __builtin_bit_cast(b8, bc[2]), b1l = __builtin_bit_cast(b8, bc[3]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, b0h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, b0l, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0l, b0h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, b1h, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, b1l, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1l, b1h, acc, 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (vn < 0) break;  // uniform over the block
      *reinterpret_cast<f32x4 *>(&sK[buf ^ 1][wave][lane]) = kn;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; q++) bc[q] = bn[q];
      vn = vnn;
      Sn = Snn;
    }
  }
  if (!valid) return;
  float *Lp = (float *)g.L + ((size_t)rh * g.ncell + F) * 32 + 4 * half;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    f32x4 o = {acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]};
    *reinterpret_cast<f32x4 *>(Lp + 8 * q) = o;
  }
}

int list_m2l_create(ListM2L **out, int r, int nrhs, int form, const real *K, int ncell, const int *cls_off,
                    const int *d_cls_cells, const int *d_vsrc, hipStream_t s)
{
  *out = nullptr;
#ifndef SINGLE
  (void)r; (void)nrhs; (void)form; (void)K; (void)ncell; (void)cls_off; (void)d_cls_cells; (void)d_vsrc; (void)s;
  return set_error(BBFMM_E_ARG, "the list MFMA M2L needs the fp32 build");
#else
  if (r != 32) return set_error(BBFMM_E_ARG, "the list MFMA M2L is built for r = 32 (got r=%d)", r);
  if (form != BBFMM_M2L_BF16X2 && form != BBFMM_M2L_FP32MFMA)
    return set_error(BBFMM_E_ARG, "the list MFMA M2L runs the split-BF16 or FP32 MFMA form");
  ListM2L *g = new ListM2L();
  g->form = form;
  g->nrhs = nrhs;
  g->ncell = ncell;
  ListArgs &a = g->args;
  a.ncell = ncell;
  a.nrhs = nrhs;
  int nb = 0;
  for (int c = 0; c <= 8; c++) a.cls_off[c] = cls_off[c];
  for (int c = 0; c < 8; c++) {
    a.blk0[c] = nb;
    const long long cols = (long long)(cls_off[c + 1] - cls_off[c]) * nrhs;
    nb += (int)((cols + 127) / 128);
  }
  a.blk0[8] = nb;
  g->nblocks = nb;
  std::vector<int4> tab;
  offset_table(tab);
  // proportional interleave of the classes' blocks (stable: ties by class)
  std::vector<std::pair<double, int>> order;
  order.reserve(nb);
  for (int c = 0; c < 8; c++) {
    const int n = a.blk0[c + 1] - a.blk0[c];
    for (int k = 0; k < n; k++) order.push_back({(k + 0.5) / n, 8 * k + c});
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<double, int> &x, const std::pair<double, int> &y) { return x.first < y.first; });
  std::vector<int> blk(nb);
  for (int i = 0; i < nb; i++) blk[i] = order[i].second;
  int rc = BBFMM_OK;
  auto fail = [&](hipError_t e, const char *what) {
    rc = set_error(BBFMM_E_HIP, "list M2L: %s: %s", what, hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&g->d_tab, sizeof(int4) * tab.size())) != hipSuccess) fail(e, "hipMalloc");
  if (rc == BBFMM_OK && (e = hipMemcpy(g->d_tab, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice)) != hipSuccess)
    fail(e, "copy");
  if (rc == BBFMM_OK && nb > 0) {
    if ((e = hipMalloc(&g->d_blk, sizeof(int) * (size_t)nb)) != hipSuccess) fail(e, "hipMalloc");
    else if ((e = hipMemcpy(g->d_blk, blk.data(), sizeof(int) * (size_t)nb, hipMemcpyHostToDevice)) != hipSuccess)
      fail(e, "copy");
  }
  if (rc == BBFMM_OK) {
    if (form == BBFMM_M2L_BF16X2) {
      std::vector<__bf16> Kb((size_t)kNumOps * 2 * 2 * 64 * 8);
      for (int k = 0; k < kNumOps; k++)
        for (int ks = 0; ks < 2; ks++)
          for (int lane = 0; lane < 64; lane++)
            for (int q = 0; q < 8; q++) {
              const int row = lane & 31, col = 16 * ks + 8 * (lane >> 5) + q;
              const float v = K[(size_t)k * 32 * 32 + (size_t)col * 32 + row];
              const __bf16 h = (__bf16)v;
              Kb[((((size_t)k * 2 + ks) * 2 + 0) * 64 + lane) * 8 + q] = h;
              Kb[((((size_t)k * 2 + ks) * 2 + 1) * 64 + lane) * 8 + q] = (__bf16)(v - (float)h);
            }
      if ((e = hipMalloc(&g->d_K, sizeof(__bf16) * Kb.size())) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpy(g->d_K, Kb.data(), sizeof(__bf16) * Kb.size(), hipMemcpyHostToDevice)) != hipSuccess)
        fail(e, "copy");
      const size_t nb2 = sizeof(__bf16) * (size_t)nrhs * (ncell + 1) * 32;
      a.kbytes = (int)(sizeof(__bf16) * Kb.size());
      a.mbytes = nb2 < ((size_t)1 << 31) ? (int)nb2 : 0;
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_M0, nb2)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_M1, nb2)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_M0, 0, nb2, s)) != hipSuccess) fail(e, "memset");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_M1, 0, nb2, s)) != hipSuccess) fail(e, "memset");
    } else {
      std::vector<float> Kf((size_t)kNumOps * 64 * 16);
      for (int k = 0; k < kNumOps; k++)
        for (int lane = 0; lane < 64; lane++)
          for (int ks = 0; ks < 16; ks++) {
            const int row = lane & 31, col = 16 * (lane >> 5) + ks;
            Kf[((size_t)k * 64 + lane) * 16 + ks] = K[(size_t)k * 32 * 32 + (size_t)col * 32 + row];
          }
      if ((e = hipMalloc(&g->d_K, sizeof(float) * Kf.size())) != hipSuccess) fail(e, "hipMalloc");
      else if ((e = hipMemcpy(g->d_K, Kf.data(), sizeof(float) * Kf.size(), hipMemcpyHostToDevice)) != hipSuccess)
        fail(e, "copy");
      const size_t nb2 = sizeof(float) * (size_t)nrhs * (ncell + 1) * 32;
      if (rc == BBFMM_OK && (e = hipMalloc(&g->d_M0, nb2)) != hipSuccess) fail(e, "hipMalloc");
      if (rc == BBFMM_OK && (e = hipMemsetAsync(g->d_M0, 0, nb2, s)) != hipSuccess) fail(e, "memset");
    }
  }
  if (rc == BBFMM_OK && (e = hipStreamSynchronize(s)) != hipSuccess) fail(e, "sync");
  if (rc != BBFMM_OK) {
    list_m2l_destroy(g);
    return rc;
  }
  a.cls_cells = d_cls_cells;
  a.vsrc = d_vsrc;
  a.tab = g->d_tab;
  a.blk_tab = g->d_blk;
  a.Ka = g->d_K;
  a.M0 = g->d_M0;
  a.M1 = g->d_M1;
  if (nb > 0) {
    if ((e = hipMalloc(&g->d_mask, sizeof(unsigned long long) * 16 * (size_t)nb)) != hipSuccess ||
        (nrhs == 1 && (e = hipMalloc(&g->d_vsw, sizeof(int) * 4 * (size_t)nb * kOffsets * 32)) != hipSuccess)) {
      fail(e, "hipMalloc");
    } else {
      hipLaunchKernelGGL(k_m2l_list_mask, dim3(4 * nb), dim3(64), 0, s, a, g->d_mask, g->d_vsw);
      if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(s)) != hipSuccess) fail(e, "mask set-up");
    }
    if (rc != BBFMM_OK) {
      list_m2l_destroy(g);
      return rc;
    }
  }
  a.mask = g->d_mask;
  a.vsw = g->d_vsw;
  *out = g;
  return BBFMM_OK;
#endif
}

void list_m2l_set_flat(ListM2L *g, int flat)
{
  if (g) g->flat = flat != 0;
}

void list_m2l_destroy(ListM2L *g)
{
  if (!g) return;
  for (void *p : {(void *)g->d_tab, (void *)g->d_blk, (void *)g->d_mask, (void *)g->d_vsw, g->d_K, g->d_M0, g->d_M1})
    if (p) (void)hipFree(p);
  delete g;
}

int list_m2l_run(ListM2L *g, const real *M, real *L, hipStream_t s, hipEvent_t mid)
{
  ListArgs a = g->args;
  a.L = L;
  const long long prep = (long long)a.nrhs * a.ncell * 4;
  const int pb = (int)((prep + 255) / 256);
#ifdef SINGLE
  if (g->form == BBFMM_M2L_BF16X2) {
    hipLaunchKernelGGL(k_m2l_list_prep<BBFMM_M2L_BF16X2>, dim3(pb), dim3(256), 0, s, a, M);
    BB_HIP(hipGetLastError());
    if (mid) BB_HIP(hipEventRecord(mid, s));
    if (g->nblocks && a.mbytes > 0 && !g->flat)  // grids below 2 GB: scalar-offset buffer loads
      hipLaunchKernelGGL(k_m2l_list_lds<true>, dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
    else if (g->nblocks)
      hipLaunchKernelGGL(k_m2l_list_lds<false>, dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_m2l_list_prep<BBFMM_M2L_FP32MFMA>, dim3(pb), dim3(256), 0, s, a, M);
    BB_HIP(hipGetLastError());
    if (mid) BB_HIP(hipEventRecord(mid, s));
    if (g->nblocks) hipLaunchKernelGGL(k_m2l_list_mfma<BBFMM_M2L_FP32MFMA>, dim3((g->nblocks + 7) & ~7), dim3(256), 0, s, a);
  }
  BB_HIP(hipGetLastError());
  return BBFMM_OK;
#else
  (void)pb; (void)M; (void)s; (void)mid;
  return set_error(BBFMM_E_ARG, "the list MFMA M2L needs the fp32 build");
#endif
}

} // namespace bb
