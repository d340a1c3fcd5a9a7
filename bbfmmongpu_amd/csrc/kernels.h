// kernels.h -- argument blocks and launchers of the device kernels.
#pragma once

#include "common.h"

#include <vector>

namespace bb {

// Everything the pass kernels read, by value (device pointers).
struct PassArgs {
  int N, n, r, ncell, maxlev, nleaf;
  real prefac3;         // (2/n)^3
  const real4 *pos;     // [N] leaf order
  const real4 *center;  // [ncell]
  const int *leafsta;   // [nleaf]
  const int *leafcnt;   // [nleaf]
  const int *leafof;    // [N] leaf (local index) of each leaf-ordered particle
  const int *leafcell;  // [nleaf] cell index of each leaf
  const real *leafihalf;  // [nleaf] 1 / (half the leaf's length)
  const int *celllev;   // [ncell] level of each cell
  const int *child;     // [ncell * 8] children (-1: none)
  const int *plist;     // cells with children, grouped by level
  const real *scale;    // [maxlev+1] (1/L_l)^homogen
  const real *Tkz, *Kw, *Cw;
  const real *Ucm;      // U column-major (n^3 x r)
  const real *Urm;      // U row-major
  const real *q;        // [rhs][N] leaf order
  real *phi;            // [rhs][N] leaf order
  real *S, *M, *L, *F;
  int potential = 0;    // 1: L2P writes the potential, 0: q * potential (the reference's output)
  int l2p_ref = 0;      // n % 4 == 3: L2P restates the reference's z weights (anotherDownwardPassX.c:304)
  // particle-row form of 16 vectors (rows = 1): charges and potentials as
  // rows of 16 in the padded leaf order of the MFMA P2P; P2M reads qt, L2P
  // writes the far-field potential into phit and vector 0's far field (with
  // the output's q factor) into farfield
  int rows = 0;
  const real *qt = nullptr;   // [Npad][16]
  real *phit = nullptr;       // [Npad][16]
  const int *pad_of = nullptr;  // [N] leaf-order index -> padded index
  real *farfield = nullptr;   // [N] leaf order
  // fused last downward level (launch_down_leaf): post-M2L skips cells from
  // cell_end on and L2L stops before parents of level l2l_end
  int cell_end = 0x7fffffff;
  int l2l_end = 0x7fffffff;
  // fused first upward level (launch_up_leaf): M2M starts at parents of this level
  int m2m_start = 0x7fffffff;
  // fused first upward level with the split-BF16 dense-grid M2L: the leaves'
  // M also go straight into the grids (hi and lo parts, [slot][16][32] bf16
  // from slot grid_slot0 of the deepest level); null: not written
  void *grid_hi = nullptr;
  void *grid_lo = nullptr;
  int grid_slot0 = 0;
  // sparse forests, n = 4, r = 32, fp32: P2M and post-M2L take 32 cells
  // (leaves) per wave as the columns of MFMA products
  int cols = 0;
  // sparse forests, n = 4, r = 32, fp32: post-M2L and L2L fused level by
  // level (launch_down_cols); null otherwise
  const int2 *dcp = nullptr;  // cells of levels >= 2 by (level, class): (cell, parent)
  const int2 *dwv = nullptr;  // waves: (first entry of dcp, class << 8 | count <= 32)
  const real *KU = nullptr;   // Kw .* U, column-major (64 x 32)
  const real *T8 = nullptr;   // (2/n)^3 L2L maps of the 8 child classes, [class][parent node][child node]
  // sparse forests, n = 4, r = 32, fp32: M2M on MFMA (k_m2m_cols); null otherwise
  const real *T8m = nullptr;    // the maps as [class][child node][parent node]
  const int *mplist = nullptr;  // plist with each level's parents ordered by their set of children
};

int launch_gather(const int *perm, const real *src, real *dst, int N, int nrhs, hipStream_t s);
// dst[j] = src[iperm[j]] per vector (iperm: user index -> leaf-order index)
int launch_scatter(const int *iperm, const real *src, real *dst, int N, int nrhs, hipStream_t s);
// particle-row form of 16 vectors: qt[ipad[j]][v] = q[v][j] for the nq
// (1..16) vectors of q, 0 for v >= nq, and phi[v][j] = phit[ipad[j]][v] for
// v < nq (ipad: user index -> padded leaf-order index)
int launch_gather_rows16(const int *ipad, const real *q, int nq, real *qt, int N, hipStream_t s);
// phin (may be null): near-field rows to add, then the output's q factor
// (qt rows, unless potential); without phin the rows of phit are final
int launch_scatter_rows16(const int *ipad, const real *phit, const real *phin, const real *qt, int potential,
                          real *phi, int nq, int N, hipStream_t s);
int launch_p2m(const PassArgs &a, hipStream_t s, int nrhs);
// h_ploff[l]: offset of level l's parents in a.plist (levels 0..maxlev)
int launch_m2m(const PassArgs &a, const int *h_ploff, hipStream_t s, int nrhs);
int launch_postm2l(const PassArgs &a, hipStream_t s, int nrhs);
int launch_l2l(const PassArgs &a, const int *h_ploff, hipStream_t s, int nrhs);
// sparse forests (a.dcp set): per level l >= 2, F = scale_l Kw .* (U L) plus
// the L2L of the parent's F, for every cell of the level; h_dwv_off[l] is the
// first wave of level l
int launch_down_cols(const PassArgs &a, const int *h_dwv_off, hipStream_t s, int nrhs);
int launch_l2p(const PassArgs &a, hipStream_t s, int nrhs);
// particle rows, n = 4, complete tree, maxlev >= 3: post-M2L of the leaves,
// the last L2L level and L2P in one kernel (set cell_end = first leaf cell
// and l2l_end = maxlev - 1 for the other launches)
int launch_down_leaf(const PassArgs &a, const int *h_ploff, hipStream_t s);
// particle rows, n = 4, complete tree, maxlev >= 3: P2M and the M2M into the
// leaves' parents in one kernel (set m2m_start = maxlev - 2 for the M2M)
int launch_up_leaf(const PassArgs &a, const int *h_ploff, hipStream_t s);

// ---- M2L ----
struct M2LArgs {
  int r, ncell, nrhs, minlev, maxlev;
  const int *h_levsta;     // host [maxlev+2]; NULL for a batch of sparse instances (basic scheme only)
  // basic / sibling schemes
  const int *iinter;       // [ncell]
  const int *inter;        // [ncell*inter_pitch]
  int inter_pitch = kPitchInteraction;
  const int *kidx;         // [ncell*192]
  const real4 *K4;         // [316][r/4][r] four consecutive columns of row i (r = 32, 256)
  const real *Kc = nullptr;  // [316][r][r] column-major (the reference's layout; other ranks)
  const int *Ktable;       // [343] (device)
  const real *M;           // [rhs][ncell][r]
  real *L;                 // [rhs][ncell][r]
  int variant = 0;         // basic scheme: 0 / 1 / 2 = 256 / 128 / 512 threads per block
  // sparse forests (h_levsta NULL): sources by sibling-class slot instead of
  // lists: vsrc [ncell][189] (ncell: empty slot), cellcls [ncell], slotop
  // [8][189] operator of each slot (device)
  const int *vsrc = nullptr, *cellcls = nullptr, *slotop = nullptr;
  // cluster scheme: offset of each operator, (vx + 3) 49 + (vy + 3) 7 + (vz + 3) [316] (device)
  const int *opoff = nullptr;
  // sparse forests: only the cells of levels lev_lo..lev_hi (celllev [ncell], device; NULL: all)
  const int *celllev = nullptr;
  int lev_lo = 0, lev_hi = 1 << 30;
};
constexpr int kClusterMaxRank = 64;  // the cluster kernel stages whole r x r operators in LDS
int launch_m2l_basic(const M2LArgs &a, hipStream_t s);
int launch_m2l_sibling(const M2LArgs &a, hipStream_t s);
// cluster blocking, r <= kClusterMaxRank, complete trees (a.Kc, a.opoff set)
int launch_m2l_cluster(const M2LArgs &a, hipStream_t s);
// operator offsets in Ktable order: op[k] = (vx + 3) 49 + (vy + 3) 7 + (vz + 3)
void operator_offsets(int *op);

// ---- dense-grid M2L (ghost-padded per-sibling-class grids) ----
struct GridM2L;  // defined in m2l_grid.hip
// minlev: first level handled by the grids (levels below are left to the
// caller); variant: 0 automatic, 1 / 4 / 5 source-stationary (split-BF16)
// with 8 / 4 / 2 field classes per wave, 2 / 3 one / two 32-column tiles per
// wave (r = 32 MFMA forms); variant_upper: 0, or a source-stationary variant
// for the levels above the deepest (then two launches); upper_split: that
// launch takes each class group's steps in this many parts (1..16), one
// block each, whose partial sums are added afterwards
int grid_m2l_create(GridM2L **g, int r, int maxlev, int nrhs, int form, const real *K_colmajor_host,
                    hipStream_t s, int minlev, int variant, int variant_upper = 0, int upper_split = 1);
void grid_m2l_destroy(GridM2L *g);
// flat = 1: the split-BF16 source-stationary kernel reads through pointers
// (its form for grids of 2 GB and more) whatever the grid size
void grid_m2l_set_flat(GridM2L *g, int flat);
// records `mid` (if not null) between the scatter into the grids and the M2L kernel
// deep_in_grid: the deepest level's multipoles are already in the grids
// (written by the fused first upward level), so only the levels above are
// scattered from M
int grid_m2l_run(GridM2L *g, const real *M, real *L, int ncell, hipStream_t s, hipEvent_t mid, int deep_in_grid = 0);
// the split-BF16 grids of the deepest level, for the fused first upward
// level to write into: returns 1 and sets hi, lo and the level's first slot
// when the grid is split-BF16 with 16 vectors and covers that level, else 0
int grid_m2l_leaf_grids(const GridM2L *g, void **hi, void **lo, int *slot0);

// ---- list MFMA M2L (trees storing only occupied cells) ----
struct ListM2L;  // defined in m2l_grid.hip
// cls_off[9] (host) / d_cls_cells (device): the field cells of levels >= 2
// grouped by sibling class; d_vsrc [ncell * 189] (device): source cell of
// each class offset (class_slot order), ncell for an empty source
int list_m2l_create(ListM2L **g, int r, int nrhs, int form, const real *K_colmajor_host, int ncell,
                    const int *cls_off, const int *d_cls_cells, const int *d_vsrc, hipStream_t s);
void list_m2l_destroy(ListM2L *g);
// flat = 1: pointer-form loads (the form for arrays of 2 GB and more) always
void list_m2l_set_flat(ListM2L *g, int flat);
int list_m2l_run(ListM2L *g, const real *M, real *L, hipStream_t s, hipEvent_t mid);
// slot of offset (vx, vy, vz) among the 189 M2L offsets of sibling class s, or -1
int class_slot(int s, int vx, int vy, int vz);
// operator index (Ktable order) of every slot: op[8 * 189], class-major
void slot_operators(std::vector<int> &op);

// ---- P2P ----
struct P2PArgs {
  int N, nleaf, nrhs;
  const real4 *pos;        // [N] leaf order (x,y,z,0)
  const int *leafsta, *leafcnt;
  const int *nbr;          // [nleaf*27] neighbour leaves (local index), -1 padded
  const int *nnbr;         // [nleaf]
  const real *q;           // [rhs][N]
  real *phi;               // [rhs][N] (accumulated into)
  int use_mfma = 0;        // fp32: 16 vectors at a time on MFMA
  int potential = 0;       // 1: add the potential, 0: q * potential (the reference's output)
  // padded particle copy (every leaf padded to a multiple of 4 with far-away,
  // zero-charge dummies); kernels fall back to the unpadded ones when null
  const real4 *ppos = nullptr;  // [Npad]
  const int2 *seg = nullptr;    // [nleaf][27] (padded start, padded count), self first, empties dropped
  const int *nseg = nullptr;    // [nleaf]
  const int *pad_of = nullptr;  // [N] leaf-order index -> padded index
  int npad = 0;                 // entries of ppos (and rows of qt)
  float *qt = nullptr;          // [Npad][16] charges of 16 vectors, source-major
  // particle-row form: the far-field potentials are in phit [Npad][16]; P2P
  // adds the near field and applies the output's q factor there (phi unused)
  float *phit = nullptr;
  // particle-row form beside the far field: P2P stores the near-field
  // potentials in rows of their own (the scatter combines them); phit unused
  float *phin = nullptr;
  // sparse forests (a few particles per leaf): a lane per target particle,
  // leafof[i] its leaf; used when set
  const int *leafof = nullptr;
  real4 *pq = nullptr;  // [N] scratch: positions with one vector's charge in w (one load per source)
  // split-BF16 accumulation (16 vectors, padded copy below 2^25 entries):
  // hi and lo bf16 parts of qt at [padded source / 4][vector][4 sources],
  // written from qt before the pass; used when set
  void *qsh = nullptr, *qsl = nullptr;
  // 1: the fp32 16-vector P2P reads through pointers (its form at 2^25 padded
  // particles and more) whatever the size
  int flat = 0;
};
int launch_p2p(const P2PArgs &a, hipStream_t s);

// ---- adjoint (position gradient) ----
struct GradArgs {
  int N, n, nleaf, ncell, m;  // m vectors: charges q in rows 0..m-1 of qg, dLoss/dphi in m..2m-1
  int v0 = 0, mb = 0;         // the block of vectors v0..v0+mb-1 a launch handles (mb <= 16)
  real prefac3;
  const int *leafcell;        // [nleaf]
  const real *leafihalf;      // [nleaf]
  const real4 *pos, *center;         // leaf order
  const int *leafsta, *leafcnt, *nbr, *nnbr;
  const real *Tkz;
  int l2p_ref = 0;                   // n % 4 == 3: the reference's L2P z weights (as PassArgs)
  const real *F;                     // [2m][ncell][n^3] local values after L2L
  const real *qg;                    // [2m][N] leaf order
  real *dx;                          // [3][N] leaf order (scratch)
};
// dx_out [N][3] (original order) = position gradient; iperm: original -> leaf order
int launch_gradient(const GradArgs &a, const int *iperm, real *dx_out, hipStream_t s);

// ---- FP64 operator precompute ----
int operators_fp64(int n, int r, double *U, double *K, double *sigma);
// same results with the kernel sampling, the Gram matrix and the 316
// products E_k = U^T K0_k U on the GPU (the eigenproblem stays on the host)
int operators_fp64_device(int n, int r, double *U, double *K, double *sigma);
// eigen-decomposition of a symmetric m x m matrix (column-major, destroyed):
// eigenvalues w descending, eigenvectors as the columns of V, each signed so
// that its largest-magnitude entry is positive
void jacobi_eig(int m, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V);

} // namespace bb
