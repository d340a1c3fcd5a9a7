// m2l.hip -- list-driven M2L kernels (plain real-precision FMAs):
//   basic   (M2L_SCHEME_BA): one thread per (field cell, row), walks the
//            interaction list (the semantics of m2l_kern_basic.cu:218-384);
//   sibling (M2L_SCHEME_SI): one block per 8-sibling field cluster, the 8
//            source siblings of each neighbouring cluster staged in LDS
//            (the blocking of m2l_kern_sibling_blocking.cu / CPU7A);
//   cluster (M2L_SCHEME_CL): one block per chunk of 2^3 field clusters, the
//            operators staged in LDS and shared by its 64 cells (the
//            blocking of m2l_host_cluster_blocking.cu).
// Both write L (not accumulate) for every cell of levels minlev..maxlev.
// r = 32 and 256 have kernels with the rank fixed at compile time (K in the
// K4 layout below); every other rank runs the *_any kernels (K column-major).
//
// K4 layout: K4[k][j/4][i] = (K_k(i,4j'), K_k(i,4j'+1), K_k(i,4j'+2), K_k(i,4j'+3))
// so that the lanes of a row group read consecutive 16-byte (fp32) words.
#include "common.h"
#include "kernels.h"

namespace bb {

// The sources of field cell F come from its interaction list (inter, kidx),
// or, on sparse forests (SLOTS), by sibling-class slot: vsrc[F][189] names
// the source of each of the class's 189 offsets (ncell: empty) and
// slotop[class][189] its operator; the empty slots are skipped.
struct SlotArgs {
  const int *vsrc = nullptr, *cellcls = nullptr, *slotop = nullptr;
  const int *celllev = nullptr;  // not null: only the cells of levels lev_lo..lev_hi
  int lev_lo = 0, lev_hi = 1 << 30;
};

__device__ __forceinline__ bool slot_cell_skipped(const SlotArgs &sa, int F)
{
  return sa.celllev && (sa.celllev[F] < sa.lev_lo || sa.celllev[F] > sa.lev_hi);
}

template <bool SLOTS>
__device__ __forceinline__ bool m2l_source(const SlotArgs &sa, int F, int t, int ncell, const int *il,
                                           const int *kl, int &S, int &k)
{
  if (SLOTS) {
    S = sa.vsrc[(size_t)kM2LOffsetsPerClass * F + t];
    k = sa.slotop[sa.cellcls[F] * kM2LOffsetsPerClass + t];
    return S < ncell;
  }
  S = il[t];
  k = kl[t];
  return true;
}

template <int R, int BLOCK, bool SLOTS = false>
__global__ void __launch_bounds__(BLOCK) k_m2l_basic(int F0, int nF, int ncell, const int *__restrict__ iinter,
                                                   const int *__restrict__ inter, int ipitch,
                                                   const int *__restrict__ kidx,
                                                   const real4 *__restrict__ K4,
                                                   const real *__restrict__ M, real *__restrict__ L, SlotArgs sa)
{
  constexpr int CPB = BLOCK / R;  // field cells per block
  const int cl = threadIdx.x / R;
  const int i = threadIdx.x % R;
  const int F = F0 + blockIdx.x * CPB + cl;
  if (F >= F0 + nF) return;  // no barrier below
  if (SLOTS && slot_cell_skipped(sa, F)) return;
  const size_t roff = (size_t)blockIdx.y * ncell * R;
  const int ni = SLOTS ? kM2LOffsetsPerClass : iinter[F];
  const int *il = SLOTS ? nullptr : inter + (size_t)ipitch * F;
  const int *kl = SLOTS ? nullptr : kidx + (size_t)kPitchInteraction * F;
  real acc0 = 0, acc1 = 0;
  for (int t = 0; t < ni; t++) {
    int S, k;
    if (!m2l_source<SLOTS>(sa, F, t, ncell, il, kl, S, k)) continue;
    const real4 *Kp = K4 + (size_t)k * (R / 4) * R + i;
    const real4 *Mp = reinterpret_cast<const real4 *>(M + roff + (size_t)S * R);
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
  L[roff + (size_t)F * R + i] = acc0 + acc1;
}

// Morton helpers: local index within a level <-> grid coordinates.  The
// sibling index is 4*xbit + 2*ybit + zbit at every level (anotherFMMDistribute.c:239-253).
__device__ __forceinline__ void morton_decode(int m, int level, int &x, int &y, int &z)
{
  x = y = z = 0;
  for (int b = 0; b < level; b++) {
    const int s = (m >> (3 * b)) & 7;
    x |= ((s >> 2) & 1) << b;
    y |= ((s >> 1) & 1) << b;
    z |= (s & 1) << b;
  }
}

__device__ __forceinline__ int morton_encode(int x, int y, int z, int level)
{
  int m = 0;
  for (int b = 0; b < level; b++)
    m |= ((((x >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((z >> b) & 1)) << (3 * b);
  return m;
}

// Sibling blocking: one block per field cluster (the 8 children of a parent
// P at level-1).  Threads = R rows, each accumulating the row for all 8
// field siblings; the 8 source siblings of each of P's 26 neighbours are
// staged in LDS and every (field, source) pair whose offset is an M2L
// offset uses K[Ktable[offset]].
template <int R>
__global__ void __launch_bounds__(R < 64 ? 64 : R) k_m2l_sibling(int level, int plev0, int lev0, int ncell,
                                                                 const int *__restrict__ Ktable,
                                                                 const real4 *__restrict__ K4,
                                                                 const real *__restrict__ M,
                                                                 real *__restrict__ L)
{
  constexpr int CPB = R < 64 ? 64 / R : 1;  // clusters per block
  constexpr int BLOCK = R * CPB;
  __shared__ real4 sm[CPB][8][R / 4];
  __shared__ int skt[343];
  const int cl = threadIdx.x / R;
  const int i = threadIdx.x % R;
  const int plevel = level - 1;
  const int npar = 1 << (3 * plevel);
  const int pidx = blockIdx.x * CPB + cl;  // local parent index
  const bool active = pidx < npar;
  const size_t roff = (size_t)blockIdx.y * ncell * R;
  for (int o = threadIdx.x; o < 343; o += BLOCK) skt[o] = Ktable[o];
  int px = 0, py = 0, pz = 0;
  if (active) morton_decode(pidx, plevel, px, py, pz);
  const int G = 1 << plevel;
  real acc[8];
#pragma unroll
  for (int f = 0; f < 8; f++) acc[f] = 0;
  for (int d = 0; d < 27; d++) {
    const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
    const int sx = px + dx, sy = py + dy, sz = pz + dz;
    const bool ok = active && (dx | dy | dz) != 0 && sx >= 0 && sy >= 0 && sz >= 0 && sx < G && sy < G && sz < G;
    __syncthreads();
    if (ok) {
      const int SC = plev0 + morton_encode(sx, sy, sz, plevel);
      const real4 *src = reinterpret_cast<const real4 *>(M + roff + (size_t)(8 * SC + 1) * R);
      for (int o = i; o < 8 * (R / 4); o += R) sm[cl][o / (R / 4)][o % (R / 4)] = src[o];
    }
    __syncthreads();
    if (!ok) continue;
#pragma unroll
    for (int f = 0; f < 8; f++) {
      const int fx = (f >> 2) & 1, fy = (f >> 1) & 1, fz = f & 1;
      for (int s = 0; s < 8; s++) {
        const int vx = 2 * dx + ((s >> 2) & 1) - fx;
        const int vy = 2 * dy + ((s >> 1) & 1) - fy;
        const int vz = 2 * dz + (s & 1) - fz;
        const int k = skt[49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)];
        if (k < 0) continue;
        const real4 *Kp = K4 + (size_t)k * (R / 4) * R + i;
        real a = 0;
#pragma unroll 8
        for (int jb = 0; jb < R / 4; jb++) {
          const real4 kv = Kp[(size_t)jb * R];
          const real4 mv = sm[cl][s][jb];
          a += kv.x * mv.x + kv.y * mv.y + kv.z * mv.z + kv.w * mv.w;
        }
        acc[f] += a;
      }
    }
  }
  if (active) {
    const int A0 = lev0 + 8 * pidx;
#pragma unroll
    for (int f = 0; f < 8; f++) L[roff + (size_t)(A0 + f) * R + i] = acc[f];
  }
}

// ---------------------------------------------------------------- any rank
// Ranks other than 32 and 256 (the reference reaches them through its
// generic-r path, m2l_host_ij_blocking.cu:13-17,419-428, and rejects them in
// the basic and sibling hosts; r = n^3/2 for n = 3..8 and n^3/2 - 1 under
// DISABLE_INCREASE_CUTOFF_BY_ONE, bbfmm.c:1340-1343).  K is read in the
// reference's own column-major layout (K_k(i, j) at k r^2 + j r + i): the r
// rows of a field cell are consecutive lanes, so each column's loads are
// coalesced, and the source's M entries are the same address for all of
// them.  A block of BLOCK threads holds BLOCK / r cells (r <= BLOCK) or one
// cell with ceil(r / BLOCK) rows per thread.
template <int BLOCK>
__device__ __forceinline__ void any_rank_lane(int r, int &rt, int &cpb, int &cl, int &i0)
{
  rt = r < BLOCK ? r : BLOCK;
  cpb = BLOCK / rt;
  cl = (int)threadIdx.x / rt;
  i0 = (int)threadIdx.x - cl * rt;
}

template <int BLOCK, bool SLOTS = false>
__global__ void __launch_bounds__(BLOCK) k_m2l_basic_any(int r, int F0, int nF, int ncell,
                                                       const int *__restrict__ iinter,
                                                       const int *__restrict__ inter, int ipitch,
                                                       const int *__restrict__ kidx,
                                                       const real *__restrict__ Kc,
                                                       const real *__restrict__ M, real *__restrict__ L, SlotArgs sa)
{
  constexpr int RPT = (kMaxRank + BLOCK - 1) / BLOCK;  // rows per thread at most
  int rt, cpb, cl, i0;
  any_rank_lane<BLOCK>(r, rt, cpb, cl, i0);
  const int F = F0 + blockIdx.x * cpb + cl;
  if (cl >= cpb || F >= F0 + nF) return;  // no barrier below
  if (SLOTS && slot_cell_skipped(sa, F)) return;
  const size_t roff = (size_t)blockIdx.y * ncell * r;
  const int ni = SLOTS ? kM2LOffsetsPerClass : iinter[F];
  const int *il = SLOTS ? nullptr : inter + (size_t)ipitch * F;
  const int *kl = SLOTS ? nullptr : kidx + (size_t)kPitchInteraction * F;
  real acc[RPT];
#pragma unroll
  for (int u = 0; u < RPT; u++) acc[u] = 0;
  for (int t = 0; t < ni; t++) {
    int S, k;
    if (!m2l_source<SLOTS>(sa, F, t, ncell, il, kl, S, k)) continue;
    const real *Kk = Kc + (size_t)k * r * r + i0;
    const real *Ms = M + roff + (size_t)S * r;
    for (int j = 0; j < r; j++) {
      const real m = Ms[j];
#pragma unroll
      for (int u = 0; u < RPT; u++)
        if (i0 + u * rt < r) acc[u] += Kk[(size_t)j * r + u * rt] * m;
    }
  }
#pragma unroll
  for (int u = 0; u < RPT; u++)
    if (i0 + u * rt < r) L[roff + (size_t)F * r + i0 + u * rt] = acc[u];
}

// Sibling blocking at any rank: as k_m2l_sibling, a cluster's r threads
// accumulate the rows of its 8 field siblings; the 8 source siblings of each
// neighbouring cluster are staged in LDS ([cluster][sibling][r], dynamic).
__global__ void __launch_bounds__(256) k_m2l_sibling_any(int r, int level, int plev0, int lev0, int ncell,
                                                        const int *__restrict__ Ktable,
                                                        const real *__restrict__ Kc,
                                                        const real *__restrict__ M, real *__restrict__ L)
{
  constexpr int RPT = (kMaxRank + 255) / 256;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  real *sm_all = reinterpret_cast<real *>(smem_raw);
  __shared__ int skt[343];
  int rt, cpb, cl, i0;
  any_rank_lane<256>(r, rt, cpb, cl, i0);
  const int plevel = level - 1;
  const int npar = 1 << (3 * plevel);
  const int pidx = blockIdx.x * cpb + cl;  // local parent index
  const bool active = cl < cpb && pidx < npar;
  real *sm = sm_all + (size_t)(cl < cpb ? cl : 0) * 8 * r;
  const size_t roff = (size_t)blockIdx.y * ncell * r;
  for (int o = threadIdx.x; o < 343; o += 256) skt[o] = Ktable[o];
  int px = 0, py = 0, pz = 0;
  if (active) morton_decode(pidx, plevel, px, py, pz);
  const int G = 1 << plevel;
  real acc[8][RPT];
#pragma unroll
  for (int f = 0; f < 8; f++)
#pragma unroll
    for (int u = 0; u < RPT; u++) acc[f][u] = 0;
  for (int d = 0; d < 27; d++) {
    const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
    const int sx = px + dx, sy = py + dy, sz = pz + dz;
    const bool ok = active && (dx | dy | dz) != 0 && sx >= 0 && sy >= 0 && sz >= 0 && sx < G && sy < G && sz < G;
    __syncthreads();
    if (ok) {
      const int SC = plev0 + morton_encode(sx, sy, sz, plevel);
      const real *src = M + roff + (size_t)(8 * SC + 1) * r;
      for (int o = i0; o < 8 * r; o += rt) sm[o] = src[o];
    }
    __syncthreads();
    if (!ok) continue;
#pragma unroll
    for (int f = 0; f < 8; f++) {
      const int fx = (f >> 2) & 1, fy = (f >> 1) & 1, fz = f & 1;
      for (int s = 0; s < 8; s++) {
        const int vx = 2 * dx + ((s >> 2) & 1) - fx;
        const int vy = 2 * dy + ((s >> 1) & 1) - fy;
        const int vz = 2 * dz + (s & 1) - fz;
        const int k = skt[49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)];
        if (k < 0) continue;
        const real *Kk = Kc + (size_t)k * r * r + i0;
        const real *ms = sm + s * r;
        for (int j = 0; j < r; j++) {
          const real m = ms[j];
#pragma unroll
          for (int u = 0; u < RPT; u++)
            if (i0 + u * rt < r) acc[f][u] += Kk[(size_t)j * r + u * rt] * m;
        }
      }
    }
  }
  if (active) {
    const int A0 = lev0 + 8 * pidx;
#pragma unroll
    for (int f = 0; f < 8; f++)
#pragma unroll
      for (int u = 0; u < RPT; u++)
        if (i0 + u * rt < r) L[roff + (size_t)(A0 + f) * r + i0 + u * rt] = acc[f][u];
  }
}

// ---------------------------------------------------------------- cluster blocking
// M2L_SCHEME_CL (the blocking of m2l_host_cluster_blocking.cu:29-166): a
// block is a chunk of 2 x 2 x 2 field clusters, the 64 cells of a 4 x 4 x 4
// cube of the level.  The 316 operators pass through LDS in batches of
// opb = 4096 / r^2 (double-buffered: the next batch is loaded into registers
// while the current one is applied) and are shared by the chunk's 64 cells;
// each cell applies those whose offset is in its interaction list (the
// source inside the domain and its parent a neighbour of the cell's parent,
// anotherInteractionList.c:11-111).  Four threads per cell, each holding
// RPT consecutive rows; the rows a thread reads from LDS are the same
// address for the 16 cells of a wave.  K column-major (K(i, j) at
// k r^2 + j r + i), r <= 4 RPT.
constexpr int kClusterBatch = 4096;  // reals of operators per LDS buffer

template <int RPT>
__global__ void __launch_bounds__(256) k_m2l_cluster(int r, int level, int lev0, int ncell,
                                                    const int *__restrict__ opoff,
                                                    const real *__restrict__ Kc,
                                                    const real *__restrict__ M, real *__restrict__ L)
{
  constexpr int PER_THREAD = kClusterBatch / 256;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  real *sk = reinterpret_cast<real *>(smem_raw);  // [2][kClusterBatch]
  const int c = threadIdx.x >> 2, rg = threadIdx.x & 3;
  const int G = 1 << level;   // cells per axis
  const int nch = G >> 2;     // chunks per axis
  const int ch = blockIdx.x;
  const int cx = ch / (nch * nch), cy = (ch / nch) % nch, cz = ch % nch;
  const int fx = 4 * cx + (c >> 4), fy = 4 * cy + ((c >> 2) & 3), fz = 4 * cz + (c & 3);
  const size_t roff = (size_t)blockIdx.y * ncell * r;
  const int i0 = rg * RPT;
  const int r2 = r * r;
  const int opb = kClusterBatch / r2;  // operators per batch (>= 1 for r <= 64)
  const int nbt = (kNumOps + opb - 1) / opb;
  const size_t ktot = (size_t)kNumOps * r2;
  real acc[RPT];
#pragma unroll
  for (int u = 0; u < RPT; u++) acc[u] = 0;
  real pre[PER_THREAD];
  auto fetch = [&](int bt) {
    const size_t base = (size_t)bt * opb * r2;
#pragma unroll
    for (int v = 0; v < PER_THREAD; v++) {
      const int o = threadIdx.x + 256 * v;
      pre[v] = (o < opb * r2 && base + o < ktot) ? Kc[base + o] : (real)0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int v = 0; v < PER_THREAD; v++) sk[buf * kClusterBatch + threadIdx.x + 256 * v] = pre[v];
  };
  fetch(0);
  store(0);
  __syncthreads();
  for (int bt = 0; bt < nbt; bt++) {
    if (bt + 1 < nbt) fetch(bt + 1);
    const real *skb = sk + (bt & 1) * kClusterBatch;
    for (int q = 0; q < opb; q++) {
      const int k = bt * opb + q;
      if (k >= kNumOps) break;
      const int pv = opoff[k];  // (vx + 3) 49 + (vy + 3) 7 + (vz + 3)
      const int vx = pv / 49 - 3, vy = (pv / 7) % 7 - 3, vz = pv % 7 - 3;
      const int sx = fx + vx, sy = fy + vy, sz = fz + vz;
      if (sx < 0 || sy < 0 || sz < 0 || sx >= G || sy >= G || sz >= G) continue;
      const int dpx = (sx >> 1) - (fx >> 1), dpy = (sy >> 1) - (fy >> 1), dpz = (sz >> 1) - (fz >> 1);
      if (dpx < -1 || dpx > 1 || dpy < -1 || dpy > 1 || dpz < -1 || dpz > 1) continue;
      const real *Ms = M + roff + (size_t)(lev0 + morton_encode(sx, sy, sz, level)) * r;
      const real *kq = skb + q * r2 + i0;
      // the source's coefficients 16 at a time: the loads are issued together
      for (int j0 = 0; j0 < r; j0 += 16) {
        real m[16];
#pragma unroll
        for (int t = 0; t < 16; t++) m[t] = j0 + t < r ? Ms[j0 + t] : (real)0;
#pragma unroll
        for (int t = 0; t < 16; t++) {
          if (j0 + t >= r) break;
          const real *kc = kq + (j0 + t) * r;
#pragma unroll
          for (int u = 0; u < RPT; u++)
            if (i0 + u < r) acc[u] += kc[u] * m[t];
        }
      }
    }
    // the other buffer was last read in batch bt - 1, before the barrier that ended it
    if (bt + 1 < nbt) store((bt + 1) & 1);
    __syncthreads();
  }
  real *Lp = L + roff + (size_t)(lev0 + morton_encode(fx, fy, fz, level)) * r;
#pragma unroll
  for (int u = 0; u < RPT; u++)
    if (i0 + u < r) Lp[i0 + u] = acc[u];
}

// r = 32: the same blocking with the chunk's operators shared through the
// scalar unit instead of LDS.  Lane l of each wave is cell l of the chunk
// (all 32 rows); wave w applies the operators w, w + 4, ...; K_k is the same
// for the 64 lanes, so its entries come by scalar loads into SGPRs and each
// FMA reads one SGPR and one VGPR (the lane's source coefficient).  The four
// waves' partial sums are added through LDS in a fixed order.
__global__ void __launch_bounds__(256) k_m2l_cluster32(int level, int lev0, int ncell,
                                                      const int *__restrict__ opoff,
                                                      const real *__restrict__ Kc,
                                                      const real *__restrict__ M, real *__restrict__ L)
{
  __shared__ real red[3][32][64];  // partial sums of waves 1..3, [row][cell]
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int G = 1 << level, nch = G >> 2;
  const int ch = blockIdx.x;
  const int cx = ch / (nch * nch), cy = (ch / nch) % nch, cz = ch % nch;
  const int fx = 4 * cx + (lane >> 4), fy = 4 * cy + ((lane >> 2) & 3), fz = 4 * cz + (lane & 3);
  const size_t roff = (size_t)blockIdx.y * ncell * 32;
  real acc[32];
#pragma unroll
  for (int i = 0; i < 32; i++) acc[i] = 0;
  for (int k = w; k < kNumOps; k += 4) {
    const int pv = opoff[k];
    const int vx = pv / 49 - 3, vy = (pv / 7) % 7 - 3, vz = pv % 7 - 3;
    const int sx = fx + vx, sy = fy + vy, sz = fz + vz;
    if (sx < 0 || sy < 0 || sz < 0 || sx >= G || sy >= G || sz >= G) continue;
    const int dpx = (sx >> 1) - (fx >> 1), dpy = (sy >> 1) - (fy >> 1), dpz = (sz >> 1) - (fz >> 1);
    if (dpx < -1 || dpx > 1 || dpy < -1 || dpy > 1 || dpz < -1 || dpz > 1) continue;
    const real4 *Ms = reinterpret_cast<const real4 *>(M + roff + (size_t)(lev0 + morton_encode(sx, sy, sz, level)) * 32);
    real4 m[8];
#pragma unroll
    for (int q = 0; q < 8; q++) m[q] = Ms[q];
    const real *__restrict__ Kk = Kc + (size_t)k * 1024;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const real mj[4] = {m[q].x, m[q].y, m[q].z, m[q].w};
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const real *__restrict__ kc = Kk + (4 * q + t) * 32;  // column j
#pragma unroll
        for (int i = 0; i < 32; i++) acc[i] += kc[i] * mj[t];
      }
    }
  }
  if (w > 0)
#pragma unroll
    for (int i = 0; i < 32; i++) red[w - 1][i][lane] = acc[i];
  __syncthreads();
  if (w != 0) return;
  real *Lp = L + roff + (size_t)(lev0 + morton_encode(fx, fy, fz, level)) * 32;
#pragma unroll
  for (int i = 0; i < 32; i++) Lp[i] = ((acc[i] + red[0][i][lane]) + red[1][i][lane]) + red[2][i][lane];
}

int launch_m2l_cluster(const M2LArgs &a, hipStream_t s)
{
  if (a.minlev < 2) return set_error(BBFMM_E_ARG, "M2L cluster scheme needs minlev >= 2");
  if (!a.h_levsta) return set_error(BBFMM_E_ARG, "M2L cluster scheme needs a complete tree");
  if (a.r < 1 || a.r > kClusterMaxRank)
    return set_error(BBFMM_E_ARG, "M2L cluster scheme: r=%d outside 1..%d", a.r, kClusterMaxRank);
  if (!a.Kc || !a.opoff) return set_error(BBFMM_E_ARG, "M2L cluster scheme needs the column-major operators");
  static_assert(kClusterMaxRank * kClusterMaxRank <= kClusterBatch, "one operator per LDS buffer at least");
  const size_t sh = sizeof(real) * 2 * kClusterBatch;
  for (int level = a.minlev; level <= a.maxlev; level++) {
    const int nch = 1 << (3 * (level - 2));
    if (a.r == 32)
      hipLaunchKernelGGL(k_m2l_cluster32, dim3(nch, a.nrhs), dim3(256), 0, s, level, a.h_levsta[level], a.ncell,
                         a.opoff, a.Kc, a.M, a.L);
    else if (a.r <= 32)
      hipLaunchKernelGGL(k_m2l_cluster<8>, dim3(nch, a.nrhs), dim3(256), sh, s, a.r, level, a.h_levsta[level],
                         a.ncell, a.opoff, a.Kc, a.M, a.L);
    else
      hipLaunchKernelGGL(k_m2l_cluster<16>, dim3(nch, a.nrhs), dim3(256), sh, s, a.r, level, a.h_levsta[level],
                         a.ncell, a.opoff, a.Kc, a.M, a.L);
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

template <int BLOCK>
static int basic_any_t(const M2LArgs &a, hipStream_t s)
{
  const int rt = a.r < BLOCK ? a.r : BLOCK, cpb = BLOCK / rt;
  SlotArgs sa;
  sa.vsrc = a.vsrc;
  sa.cellcls = a.cellcls;
  sa.slotop = a.slotop;
  sa.celllev = a.celllev;
  sa.lev_lo = a.lev_lo;
  sa.lev_hi = a.lev_hi;
  if (!a.h_levsta) {
    // sparse forests: every cell, sources by class slot; cells of levels 0-1 have none (L = 0)
    if (!a.vsrc) return set_error(BBFMM_E_ARG, "M2L basic: a sparse forest needs its slot sources");
    hipLaunchKernelGGL((k_m2l_basic_any<BLOCK, true>), dim3((a.ncell + cpb - 1) / cpb, a.nrhs), dim3(BLOCK), 0, s,
                       a.r, 0, a.ncell, a.ncell, a.iinter, a.inter, a.inter_pitch, a.kidx, a.Kc, a.M, a.L, sa);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
  }
  for (int level = a.minlev; level <= a.maxlev; level++) {
    const int F0 = a.h_levsta[level], nF = a.h_levsta[level + 1] - F0;
    hipLaunchKernelGGL((k_m2l_basic_any<BLOCK>), dim3((nF + cpb - 1) / cpb, a.nrhs), dim3(BLOCK), 0, s, a.r, F0,
                       nF, a.ncell, a.iinter, a.inter, a.inter_pitch, a.kidx, a.Kc, a.M, a.L, sa);
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

static int basic_any(const M2LArgs &a, hipStream_t s)
{
  if (!a.Kc) return set_error(BBFMM_E_ARG, "M2L at r=%d needs the column-major operators", a.r);
  switch (a.variant) {
  case 0: return basic_any_t<256>(a, s);
  case 1: return basic_any_t<128>(a, s);
  case 2: return basic_any_t<512>(a, s);
  default: return set_error(BBFMM_E_ARG, "M2L basic: variant %d", a.variant);
  }
}

static int sibling_any(const M2LArgs &a, hipStream_t s)
{
  if (!a.Kc) return set_error(BBFMM_E_ARG, "M2L at r=%d needs the column-major operators", a.r);
  const int rt = a.r < 256 ? a.r : 256, cpb = 256 / rt;
  const size_t sh = sizeof(real) * (size_t)cpb * 8 * a.r;
  for (int level = a.minlev; level <= a.maxlev; level++) {
    const int npar = 1 << (3 * (level - 1));
    hipLaunchKernelGGL(k_m2l_sibling_any, dim3((npar + cpb - 1) / cpb, a.nrhs), dim3(256), sh, s, a.r, level,
                       a.h_levsta[level - 1], a.h_levsta[level], a.ncell, a.Ktable, a.Kc, a.M, a.L);
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

template <int R, int BLOCK>
static int basic_t(const M2LArgs &a, hipStream_t s)
{
  constexpr int CPB = BLOCK / R;
  SlotArgs sa;
  sa.vsrc = a.vsrc;
  sa.cellcls = a.cellcls;
  sa.slotop = a.slotop;
  sa.celllev = a.celllev;
  sa.lev_lo = a.lev_lo;
  sa.lev_hi = a.lev_hi;
  if (!a.h_levsta) {
    // no level table (a batch of sparse instances): one launch over every
    // cell, sources by class slot; cells of levels 0-1 have none (L = 0)
    if (!a.vsrc) return set_error(BBFMM_E_ARG, "M2L basic: a sparse forest needs its slot sources");
    hipLaunchKernelGGL((k_m2l_basic<R, BLOCK, true>), dim3((a.ncell + CPB - 1) / CPB, a.nrhs), dim3(BLOCK), 0, s,
                       0, a.ncell, a.ncell, a.iinter, a.inter, a.inter_pitch, a.kidx, a.K4, a.M, a.L, sa);
    BB_HIP(hipGetLastError());
    return BBFMM_OK;
  }
  for (int level = a.minlev; level <= a.maxlev; level++) {
    const int F0 = a.h_levsta[level], nF = a.h_levsta[level + 1] - F0;
    hipLaunchKernelGGL((k_m2l_basic<R, BLOCK>), dim3((nF + CPB - 1) / CPB, a.nrhs), dim3(BLOCK), 0, s, F0, nF,
                       a.ncell, a.iinter, a.inter, a.inter_pitch, a.kidx, a.K4, a.M, a.L, sa);
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

template <int R>
static int basic_r(const M2LArgs &a, hipStream_t s)
{
  switch (a.variant) {
  case 0: return basic_t<R, 256>(a, s);
  case 1: return basic_t<R, (R > 128 ? R : 128)>(a, s);
  case 2: return basic_t<R, 512>(a, s);
  default: return set_error(BBFMM_E_ARG, "M2L basic: variant %d", a.variant);
  }
}

int launch_m2l_basic(const M2LArgs &a, hipStream_t s)
{
  if (a.r < 1 || a.r > kMaxRank) return set_error(BBFMM_E_ARG, "M2L: r=%d outside 1..%d", a.r, kMaxRank);
  if (a.r == 32 && a.K4) return basic_r<32>(a, s);
  if (a.r == 256 && a.K4) return basic_r<256>(a, s);
  return basic_any(a, s);
}

template <int R>
static int sibling_t(const M2LArgs &a, hipStream_t s)
{
  constexpr int CPB = R < 64 ? 64 / R : 1;
  constexpr int BLOCK = R * CPB;
  for (int level = a.minlev; level <= a.maxlev; level++) {
    const int npar = 1 << (3 * (level - 1));
    hipLaunchKernelGGL((k_m2l_sibling<R>), dim3((npar + CPB - 1) / CPB, a.nrhs), dim3(BLOCK), 0, s, level,
                       a.h_levsta[level - 1], a.h_levsta[level], a.ncell, a.Ktable, a.K4, a.M, a.L);
    BB_HIP(hipGetLastError());
  }
  return BBFMM_OK;
}

int launch_m2l_sibling(const M2LArgs &a, hipStream_t s)
{
  if (a.minlev < 2) return set_error(BBFMM_E_ARG, "M2L sibling scheme needs minlev >= 2");
  if (a.r < 1 || a.r > kMaxRank) return set_error(BBFMM_E_ARG, "M2L: r=%d outside 1..%d", a.r, kMaxRank);
  if (a.r == 32 && a.K4) return sibling_t<32>(a, s);
  if (a.r == 256 && a.K4) return sibling_t<256>(a, s);
  return sibling_any(a, s);
}

} // namespace bb
