// api.hip -- C ABI of the library: error state, the persistent plan, the
// drop-in entry points of the reference's CUDA translation unit
// (anotherFMMCompute, anotherFMMInteraction, anotherDownwardPass, m2l_drv,
// the event timers) and the tree builder.
#include "common.h"
#include "kernels.h"
#include "forest.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

// ---------------------------------------------------------------- errors
namespace bb {
static thread_local int g_err = BBFMM_OK;
static thread_local char g_msg[512] = "";

int set_error(int code, const char *fmt, ...)
{
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error()
{
  g_err = BBFMM_OK;
  g_msg[0] = 0;
}
} // namespace bb

extern "C" int bbfmm_amd_last_error(void) { return bb::g_err; }
extern "C" const char *bbfmm_amd_last_error_string(void) { return bb::g_msg; }

// m2l_host.h:24
extern "C" {
double kernel_exec_time = 0;
}

static int g_default_scheme = M2L_SCHEME_IJ;

extern "C" int bbfmm_amd_set_default_scheme(int scheme)
{
  if (scheme < M2L_SCHEME_BA || scheme > M2L_SCHEME_IJ)
    return bb::set_error(BBFMM_E_ARG, "unknown M2L scheme %d", scheme);
  g_default_scheme = scheme;
  return BBFMM_OK;
}

static bool g_verbose = getenv("BBFMM_AMD_VERBOSE") != nullptr;

// ---------------------------------------------------------------- plan
enum { ST_GATHER, ST_P2M, ST_M2M, ST_M2L_PREP, ST_M2L, ST_POST, ST_L2L, ST_L2P, ST_P2P, ST_SCATTER, ST_TOTAL };
// per timed solve: the ST_TOTAL + 1 events of the plan's stream, then the
// start and end of P2P on the side stream (the near field beside the far
// field); stage_ms[ST_TOTAL + 1] is that P2P's own duration
constexpr int kEvPerSet = ST_TOTAL + 3;
constexpr int kEvSideA = ST_TOTAL + 1, kEvSideB = ST_TOTAL + 2;

struct bbfmm_amd_plan {
  int N = 0, n = 0, n3 = 0, r = 0, nrhs = 1, maxlev = 0, ncell = 0, nleaf = 0, leaf0 = 0;
  // vectors of the cell arrays (S, F, M, L), the dense-grid M2L of the MFMA
  // forms and the particle-row form: 16 for plans of 8 to 15 vectors on a
  // complete tree (fp32 build), whose solves then run the 16-vector
  // particle-row kernels with the vectors past nrhs zero; nrhs otherwise
  int width = 1;
  bool complete = true;  // one complete octree (reference layout); false: a batch of sparse instances
  int ninst = 1;
  std::vector<int> h_ploff;  // [maxlev+1] offsets of each level's parents in d_plist
  real L = 1, homogen = 1;
  hipStream_t stream = nullptr;
  int scheme = M2L_SCHEME_BA, form = BBFMM_M2L_PLAIN, timing = 0;
  std::vector<int> h_levsta;  // [maxlev+2]
  std::vector<int> h_perm;    // leaf order -> original index
  double m2l_interactions = 0, p2p_pairs = 0;
  // device
  real4 *d_pos = nullptr, *d_center = nullptr, *d_K4 = nullptr;
  real *d_Kc = nullptr;  // column-major operators (ranks other than 32 and 256; the cluster scheme)
  int *d_opoff = nullptr;  // offset of each operator (cluster scheme)
  real4 *d_pq = nullptr;  // sparse forests: positions with one vector's charge in w (lane-per-particle P2P)
  int *d_perm = nullptr, *d_leafsta = nullptr, *d_leafcnt = nullptr, *d_leafof = nullptr;
  int *d_iperm = nullptr;  // original index -> leaf order
  int *d_celllev = nullptr, *d_child = nullptr, *d_plist = nullptr, *d_leafcell = nullptr;
  real *d_leafihalf = nullptr;
  int *d_nbr = nullptr, *d_nnbr = nullptr, *d_iinter = nullptr, *d_inter = nullptr, *d_kidx = nullptr;
  int *d_Ktable = nullptr, *d_levsta = nullptr;
  real *d_Ucm = nullptr, *d_Urm = nullptr, *d_Tkz = nullptr, *d_Kw = nullptr, *d_Cw = nullptr;
  real *d_scale = nullptr;
  real *d_q = nullptr, *d_phi = nullptr, *d_S = nullptr, *d_M = nullptr, *d_L = nullptr, *d_F = nullptr;
  real *d_qin = nullptr, *d_phiout = nullptr;  // staging for the host-array entry points
  real *d_farfield = nullptr;                  // far-field q.*phi of the last solve (vector 0)
  real *d_qt = nullptr;                        // [Npad][16] charges, source-major (MFMA P2P)
  real *d_phit = nullptr;                      // [Npad][16] potentials of the particle-row form
  int *d_ipad = nullptr;                       // [N] user index -> padded leaf-order index
  real4 *d_ppos = nullptr;                     // padded particle copy for P2P
  int2 *d_seg = nullptr;
  int *d_nseg = nullptr, *d_pad_of = nullptr;
  int npad = 0;
  int potential = 0;           // output: 0 = q * potential (reference), 1 = potential
  int l2p_ref = 0;             // n % 4 == 3: L2P with the reference's z weights (anotherDownwardPassX.c:304)
  int flat_loads = 0;          // BBFMM_OPT_ADDRESSING: pointer-form loads in every kernel that has both forms
  // host copies for moving particles inside their leaves (update_positions)
  std::vector<int> h_leafof, h_pad_of, h_leafxyz, h_leafdepth;  // [N], [N], [nleaf*3], [nleaf]
  std::vector<real> h_leafinv;                                   // [nleaf] 1 / leaf length
  real *d_grad = nullptr;      // [3][N] position-gradient scratch (allocated by the first gradient call)
  std::vector<real> h_K;                       // column-major operators (for the grid form)
  bb::GridM2L *grid = nullptr;
  int grid_form = -1;
  // list MFMA M2L of sparse forests: class lists and source table (device),
  // and the per-form operator layout (created on first use)
  int h_cls_off[9] = {};
  int *d_cls_cells = nullptr, *d_vsrc = nullptr;
  // the list M2L above a level switch: the class lists of the cells of
  // levels >= list_sw (host copies of every cell's level and class to build them)
  std::vector<int> h_celllev, h_cellcls;
  int h_cls_off_sw[9] = {};
  int *d_cls_cells_sw = nullptr;
  int list_sw = -1;
  int *d_cellcls = nullptr, *d_slotop = nullptr;  // class of each cell, operator of each class slot
  // downward pass of sparse forests on MFMA (n = 4, r = 32, fp32): the cells
  // of levels >= 2 grouped by (level, class) as (cell, parent), waves of up
  // to 32 of them per level, and the operators Kw .* U and the 8 L2L maps
  int2 *d_dcp = nullptr, *d_dwv = nullptr;
  real *d_KU = nullptr, *d_T8 = nullptr, *d_T8m = nullptr;
  int *d_mplist = nullptr;  // parents of each level ordered by their set of children (M2M on MFMA)
  std::vector<int> h_dwv_off;  // [maxlev + 2] first wave of each level
  bb::ListM2L *list = nullptr;
  int list_form = -1;
  // per-stage timing: a ring of `nsets` event sets, one set per solve,
  // read (and averaged) only when stage times are requested
  int nsets = 0;
  std::vector<hipEvent_t> ev;  // nsets * kEvPerSet
  long long solves = 0, timed_from = 0;
  double stage_ms[ST_TOTAL + 2] = {};
  bool side_timed = false;     // the timed solves ran P2P on the side stream
  // near field beside the far field (particle-row form, BBFMM_OPT_NEAR_STREAM):
  // P2P on a side stream forked after the M2L and joined before the
  // scatter, into near-field rows of its own
  int near_stream = 0;
  hipStream_t side = nullptr;
  // BBFMM_OPT_P2P_FORM: split-BF16 accumulation of the 16-vector P2P, and
  // the hi / lo charge parts it reads ([Npad][16] bf16 each)
  int p2p_split = 0;
  void *d_qsh = nullptr, *d_qsl = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  real *d_phin = nullptr;
  bool solved = false;
  // whole-solve hipGraph (BBFMM_OPT_GRAPH): captured on a private stream for
  // one (q, phi) pointer pair, replayed on the plan's stream
  int use_graph = 0;
  // M2L configuration picked by hand or by the autotuner: tile variant of the
  // scheme, and for the dense-grid schemes the first level they handle (the
  // basic scheme takes the levels below)
  int m2l_variant = 0, level_switch = bb::kMinLev;
  int m2l_variant_up = 0;  // source-stationary variant of the levels above the deepest (0: as the deepest)
  int m2l_upper_split = 1;  // parts of the upper levels' step lists (one block each)
  int grid_minlev = bb::kMinLev, grid_variant = 0, grid_variant_up = 0, grid_upper_split = 1;
  hipGraphExec_t gexec = nullptr;
  const real *gq = nullptr;
  real *gphi = nullptr;
  hipStream_t gstream = nullptr;
  bool graph_timed = false;  // the timed solves were graph launches (total time only)
};

namespace bb {

template <typename T>
static int dalloc(T **p, size_t count)
{
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(BBFMM_E_NOMEM, "hipMalloc of %zu bytes failed: %s", sizeof(T) * count,
                     hipGetErrorString(e));
  }
  return BBFMM_OK;
}

template <typename T>
static int upload(T **p, const T *h, size_t count, hipStream_t s)
{
  BB_TRY(dalloc(p, count));
  if (count) BB_HIP(hipMemcpyAsync(*p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
  return BBFMM_OK;
}

int check_device()
{
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return set_error(BBFMM_E_NODEVICE, "no GPU visible (%s)", hipGetErrorString(e));
  }
  return BBFMM_OK;
}

static void plan_free_device(bbfmm_amd_plan *p)
{
  void *ptrs[] = {p->d_pos, p->d_center, p->d_K4, p->d_Kc, p->d_opoff, p->d_perm, p->d_leafsta, p->d_leafcnt, p->d_leafof,
                  p->d_nbr, p->d_nnbr, p->d_iinter, p->d_inter, p->d_kidx, p->d_Ktable, p->d_levsta,
                  p->d_Ucm, p->d_Urm, p->d_Tkz, p->d_Kw, p->d_Cw, p->d_scale, p->d_q, p->d_phi,
                  p->d_S, p->d_M, p->d_L, p->d_F, p->d_qin, p->d_phiout, p->d_farfield, p->d_qt,
                  p->d_ppos, p->d_seg, p->d_nseg, p->d_pad_of, p->d_iperm, p->d_grad, p->d_phit, p->d_ipad, p->d_phin,
                  p->d_celllev, p->d_child, p->d_plist, p->d_leafcell, p->d_leafihalf, p->d_cls_cells,
                  p->d_cls_cells_sw,
                  p->d_vsrc, p->d_dcp, p->d_dwv, p->d_KU, p->d_T8, p->d_T8m, p->d_mplist, p->d_pq,
                  p->d_qsh,   p->d_qsl, p->d_cellcls, p->d_slotop};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  p->gexec = nullptr;
  if (p->gstream) (void)hipStreamDestroy(p->gstream);
  p->gstream = nullptr;
  if (p->side) (void)hipStreamDestroy(p->side);
  p->side = nullptr;
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  p->ev_fork = p->ev_join = nullptr;
  if (p->grid) grid_m2l_destroy(p->grid);
  p->grid = nullptr;
  if (p->list) list_m2l_destroy(p->list);
  p->list = nullptr;
  for (auto &e : p->ev)
    if (e) (void)hipEventDestroy(e);
  p->ev.clear();
  p->nsets = 0;
}

static int set_timing(bbfmm_amd_plan *p, int nsets)
{
  (void)hipStreamSynchronize(p->stream);
  for (auto &e : p->ev)
    if (e) (void)hipEventDestroy(e);
  p->ev.assign((size_t)nsets * kEvPerSet, nullptr);
  for (auto &e : p->ev) BB_HIP(hipEventCreate(&e));
  p->nsets = nsets;
  p->timed_from = p->solves;
  p->graph_timed = p->use_graph != 0;
  return BBFMM_OK;
}

static int read_timing(bbfmm_amd_plan *p)
{
  for (auto &v : p->stage_ms) v = 0;
  const long long done = p->solves - p->timed_from;
  const int sets = (int)(done < p->nsets ? done : p->nsets);
  if (sets <= 0) return BBFMM_OK;
  BB_HIP(hipStreamSynchronize(p->stream));
  for (int k = 0; k < sets; k++) {
    hipEvent_t *e = &p->ev[(size_t)k * kEvPerSet];
    if (p->graph_timed) {  // a graph launch: the total only
      float tot = 0;
      BB_HIP(hipEventElapsedTime(&tot, e[0], e[ST_TOTAL]));
      p->stage_ms[ST_TOTAL] += tot / sets;
      continue;
    }
    for (int st = 0; st < ST_TOTAL; st++) {
      float ms = 0;
      BB_HIP(hipEventElapsedTime(&ms, e[st], e[st + 1]));
      p->stage_ms[st] += ms / sets;
    }
    float tot = 0;
    BB_HIP(hipEventElapsedTime(&tot, e[0], e[ST_TOTAL]));
    p->stage_ms[ST_TOTAL] += tot / sets;
    if (p->side_timed) {
      float ms = 0;
      BB_HIP(hipEventElapsedTime(&ms, e[kEvSideA], e[kEvSideB]));
      p->stage_ms[ST_TOTAL + 1] += ms / sets;
    }
  }
  return BBFMM_OK;
}

static inline hipEvent_t stage_event(bbfmm_amd_plan *p, int st)
{
  return p->ev[(size_t)(p->solves % p->nsets) * kEvPerSet + st];
}

// Upload everything a solve needs from a forest (one complete tree in the
// reference's layout, or a batch of sparse instances).  Tkz/Kw/Cw are the
// caller's Chebyshev weights; K and U are in the reference's column-major
// file layouts.
static int plan_init_forest(bbfmm_amd_plan *p, const HostForest &f, const real3 *pos, int n, int r,
                            const real *K, const real *U, const real *Tkz, const real *Kw, const real *Cw,
                            real homogen, int nrhs, hipStream_t stream)
{
  BB_TRY(check_device());
  if (n < kMinOrder || n > kMaxOrder)
    return set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", n, kMinOrder, kMaxOrder);
  if (r < 1 || r > n * n * n) return set_error(BBFMM_E_ARG, "rank r=%d outside 1..n^3=%d", r, n * n * n);
  if (nrhs < 1) return set_error(BBFMM_E_ARG, "nrhs must be >= 1");
  if (f.maxdepth < kMinLev) return set_error(BBFMM_E_ARG, "maxlev=%d < %d", f.maxdepth, kMinLev);
  p->complete = f.complete;
  p->ninst = (int)f.inst_N.size();
  p->N = f.N;
  p->n = n;
  p->n3 = n * n * n;
  p->r = r;
  p->nrhs = nrhs;
  p->width = nrhs;
#ifdef SINGLE
  if (f.complete && nrhs >= 8 && nrhs < 16) p->width = 16;
#endif
  p->maxlev = f.maxdepth;
  p->ncell = f.ncell;
  p->nleaf = f.nleaf;
  p->leaf0 = f.complete ? f.levsta[f.maxdepth] : 0;
  p->L = f.celeng[0];
  p->homogen = homogen;
  p->stream = stream;
  if (f.complete) p->h_levsta = f.levsta;
  p->h_ploff = f.ploff;
  const int N = p->N, ncell = p->ncell, nleaf = p->nleaf, n3 = p->n3;

  // particles in leaf order
  p->h_perm = f.perm;
  std::vector<real4> hpos((size_t)N);
  for (int i = 0; i < N; i++) {
    const real3 x = pos[p->h_perm[i]];
    hpos[i] = real4{x.x, x.y, x.z, 0};
  }
  p->p2p_pairs = f.pairs;
  p->m2l_interactions = f.ninter;
  int Kt[343];
  ktable(Kt);
  // operators: K4 (four consecutive columns of a row; r = 32 and 256, the
  // ranks with compiled-in kernels), column-major as given (other ranks), U
  // both ways
  const bool k4 = r == 32 || r == 256;
  std::vector<real4> K4(k4 ? (size_t)kNumOps * r * r / 4 : 0);
  for (int k = 0; k < (k4 ? kNumOps : 0); k++)
    for (int jb = 0; jb < r / 4; jb++)
      for (int i = 0; i < r; i++) {
        const real *Kk = K + (size_t)k * r * r;
        real4 v;
        v.x = Kk[(size_t)(4 * jb + 0) * r + i];
        v.y = Kk[(size_t)(4 * jb + 1) * r + i];
        v.z = Kk[(size_t)(4 * jb + 2) * r + i];
        v.w = Kk[(size_t)(4 * jb + 3) * r + i];
        K4[((size_t)k * (r / 4) + jb) * r + i] = v;
      }
  p->h_K.assign(K, K + (size_t)kNumOps * r * r);
  std::vector<real> Urm((size_t)n3 * r);
  for (int l = 0; l < n3; l++)
    for (int i = 0; i < r; i++) Urm[(size_t)l * r + i] = U[l + (size_t)n3 * i];
  std::vector<real> scale((size_t)p->maxlev + 1);
  for (int l = 0; l <= p->maxlev; l++) scale[l] = std::pow((real)1 / f.celeng[l], homogen);

  BB_TRY(upload(&p->d_pos, hpos.data(), (size_t)N, stream));
  BB_TRY(upload(&p->d_center, f.center.data(), (size_t)ncell, stream));
  BB_TRY(upload(&p->d_celllev, f.celllev.data(), (size_t)ncell, stream));
  BB_TRY(upload(&p->d_child, f.child.data(), f.child.size(), stream));
  BB_TRY(upload(&p->d_plist, f.plist.data(), f.plist.size(), stream));
  BB_TRY(upload(&p->d_leafcell, f.leafcell.data(), (size_t)nleaf, stream));
  BB_TRY(upload(&p->d_leafihalf, f.leafihalf.data(), (size_t)nleaf, stream));
  BB_TRY(upload(&p->d_perm, p->h_perm.data(), (size_t)N, stream));
  {
    std::vector<int> iperm((size_t)N);
    for (int i = 0; i < N; i++) iperm[p->h_perm[i]] = i;
    BB_TRY(upload(&p->d_iperm, iperm.data(), (size_t)N, stream));
  }
  BB_TRY(upload(&p->d_leafsta, f.leafsta.data(), (size_t)nleaf, stream));
  BB_TRY(upload(&p->d_leafcnt, f.leafcnt.data(), (size_t)nleaf, stream));
  BB_TRY(upload(&p->d_leafof, f.leafof.data(), (size_t)N, stream));
  BB_TRY(upload(&p->d_nbr, f.nbr.data(), f.nbr.size(), stream));
  BB_TRY(upload(&p->d_nnbr, f.nnbr.data(), f.nnbr.size(), stream));
  BB_TRY(upload(&p->d_iinter, f.iinter.data(), f.iinter.size(), stream));
  if (f.complete) {
    BB_TRY(upload(&p->d_inter, f.inter.data(), f.inter.size(), stream));
    BB_TRY(upload(&p->d_kidx, f.kidx.data(), f.kidx.size(), stream));
  }
  BB_TRY(upload(&p->d_Ktable, Kt, (size_t)343, stream));
  if (f.complete) BB_TRY(upload(&p->d_levsta, p->h_levsta.data(), p->h_levsta.size(), stream));
  if (!f.complete) {
    for (int c = 0; c < 9; c++) p->h_cls_off[c] = f.cls_off[c];
    BB_TRY(upload(&p->d_cls_cells, f.cls_cells.data(), f.cls_cells.size(), stream));
    p->h_celllev = f.celllev;
    p->h_cellcls = f.cellcls;
    BB_TRY(upload(&p->d_vsrc, f.vsrc.data(), f.vsrc.size(), stream));
    BB_TRY(upload(&p->d_cellcls, f.cellcls.data(), f.cellcls.size(), stream));
    std::vector<int> sop;
    slot_operators(sop);
    BB_TRY(upload(&p->d_slotop, sop.data(), sop.size(), stream));
  }
  if (k4) BB_TRY(upload(&p->d_K4, K4.data(), K4.size(), stream));
  // column-major K: the any-rank kernels and the cluster scheme
  if (!k4 || r <= kClusterMaxRank) BB_TRY(upload(&p->d_Kc, K, (size_t)kNumOps * r * r, stream));
  {
    int op[kNumOps];
    operator_offsets(op);
    BB_TRY(upload(&p->d_opoff, op, (size_t)kNumOps, stream));
  }
  BB_TRY(upload(&p->d_Ucm, U, (size_t)n3 * r, stream));
  BB_TRY(upload(&p->d_Urm, Urm.data(), Urm.size(), stream));
  BB_TRY(upload(&p->d_Tkz, Tkz, (size_t)n * n, stream));
  BB_TRY(upload(&p->d_Kw, Kw, (size_t)n3, stream));
  BB_TRY(upload(&p->d_Cw, Cw, (size_t)2 * n * n, stream));
  BB_TRY(upload(&p->d_scale, scale.data(), scale.size(), stream));
  {
    // leaf grid coordinates (the bucketing of build_tree) for update_positions
    p->h_leafof = f.leafof;
    p->h_leafinv.resize(nleaf);
    p->h_leafxyz.resize((size_t)nleaf * 3);
    for (int k = 0; k < nleaf; k++) {
      const int A = f.leafcell[k];
      const real inv = 1 / f.celeng[f.celllev[A]];
      p->h_leafinv[k] = inv;
      const real4 cc = f.center[A];
      p->h_leafxyz[3 * k] = (int)((cc.x + p->L / 2) * inv);
      p->h_leafxyz[3 * k + 1] = (int)((cc.y + p->L / 2) * inv);
      p->h_leafxyz[3 * k + 2] = (int)((cc.z + p->L / 2) * inv);
    }
    p->h_leafdepth.resize(nleaf);
    for (int k = 0; k < nleaf; k++) p->h_leafdepth[k] = f.celllev[f.leafcell[k]];
  }
  {
    // padded copy for P2P: leaf ranges start at multiples of 4, dummies far away with q = 0
    const std::vector<int> &lsta = f.leafsta, &lcnt = f.leafcnt, &nbr = f.nbr, &nnbr = f.nnbr;
    std::vector<int> psta((size_t)nleaf), pcnt((size_t)nleaf), pad_of((size_t)N);
    int run = 0;
    for (int k = 0; k < nleaf; k++) {
      psta[k] = run;
      pcnt[k] = (lcnt[k] + 3) & ~3;
      for (int j = 0; j < lcnt[k]; j++) pad_of[lsta[k] + j] = run + j;
      run += pcnt[k];
    }
    p->npad = run;
    p->h_pad_of = pad_of;
    std::vector<real4> ppos((size_t)run);
    for (auto &x : ppos) {
      x.x = x.y = x.z = (real)1e6;
      x.w = 0;
    }
    for (int i = 0; i < N; i++) ppos[pad_of[i]] = hpos[i];
    BB_TRY(upload(&p->d_ppos, ppos.data(), ppos.size(), stream));
    if (p->width >= 16) {  // segment tables of the 16-vector P2P kernels
      std::vector<int2> seg((size_t)nleaf * 27, make_int2(0, 0));
      std::vector<int> nseg((size_t)nleaf, 0);
      for (int k = 0; k < nleaf; k++) {
        int m = 0;
        if (pcnt[k] > 0) seg[(size_t)k * 27 + m++] = make_int2(psta[k], pcnt[k]);
        for (int b = 0; b < nnbr[k]; b++) {
          const int B = nbr[(size_t)k * 27 + b];
          if (B != k && pcnt[B] > 0) seg[(size_t)k * 27 + m++] = make_int2(psta[B], pcnt[B]);
        }
        nseg[k] = m;
      }
      BB_TRY(upload(&p->d_seg, seg.data(), seg.size(), stream));
      BB_TRY(upload(&p->d_nseg, nseg.data(), nseg.size(), stream));
    }
    BB_TRY(upload(&p->d_pad_of, pad_of.data(), pad_of.size(), stream));
    std::vector<int> ipad((size_t)N);
    for (int i = 0; i < N; i++) ipad[p->h_perm[i]] = pad_of[i];
    BB_TRY(upload(&p->d_ipad, ipad.data(), ipad.size(), stream));
  }
  // sparse forests, n = 4, r = 32: tables of the fused post-M2L + L2L kernel
  std::vector<int2> dcp, dwv;
  std::vector<real> KU, T8, T8m;
  std::vector<int> mplist;
#ifdef SINGLE
  if (!f.complete && n == 4 && r == 32) {
    std::vector<int> parent((size_t)ncell, -1), cls((size_t)ncell, 0);
    for (int P = 0; P < ncell; P++)
      for (int c = 0; c < 8; c++) {
        const int ch = f.child[(size_t)8 * P + c];
        if (ch >= 0) {
          parent[ch] = P;
          cls[ch] = c;
        }
      }
    const int nl = f.maxdepth + 1;
    std::vector<std::vector<int>> bucket((size_t)nl * 8);
    for (int A = 0; A < ncell; A++)
      if (f.celllev[A] >= kMinLev) bucket[(size_t)f.celllev[A] * 8 + cls[A]].push_back(A);
    p->h_dwv_off.assign(nl + 1, 0);
    for (int l = 0; l < nl; l++) {
      p->h_dwv_off[l] = (int)dwv.size();
      for (int c = 0; c < 8; c++) {
        const std::vector<int> &b = bucket[(size_t)l * 8 + c];
        const int start = (int)dcp.size();
        for (int A : b) dcp.push_back(make_int2(A, parent[A]));
        for (int k = 0; k < (int)b.size(); k += 32)
          dwv.push_back(make_int2(start + k, (c << 8) | std::min(32, (int)b.size() - k)));
      }
    }
    p->h_dwv_off[nl] = (int)dwv.size();
    KU.resize((size_t)n3 * r);
    for (int i = 0; i < r; i++)
      for (int c = 0; c < n3; c++) KU[(size_t)c + (size_t)n3 * i] = Kw[c] * U[(size_t)c + (size_t)n3 * i];
    const real pf = (real)2 / (real)n, pf3 = pf * pf * pf;
    T8.resize((size_t)8 * n3 * n3);
    for (int c8 = 0; c8 < 8; c8++) {
      const int xb = c8 >> 2, yb = (c8 >> 1) & 1, zb = c8 & 1;
      for (int pn = 0; pn < n3; pn++)
        for (int cn = 0; cn < n3; cn++) {
          const int px = pn / 16, py = (pn / 4) % 4, pz = pn % 4, cx = cn / 16, cy = (cn / 4) % 4, cz = cn % 4;
          T8[((size_t)c8 * n3 + pn) * n3 + cn] =
              pf3 * Cw[xb * 16 + px * 4 + cx] * Cw[yb * 16 + py * 4 + cy] * Cw[zb * 16 + pz * 4 + cz];
        }
    }
    BB_TRY(upload(&p->d_dcp, dcp.data(), dcp.size(), stream));
    BB_TRY(upload(&p->d_dwv, dwv.data(), dwv.size(), stream));
    BB_TRY(upload(&p->d_KU, KU.data(), KU.size(), stream));
    BB_TRY(upload(&p->d_T8, T8.data(), T8.size(), stream));
    // M2M: the maps transposed per class ([class][child node][parent node]),
    // and each level's parents ordered by their set of children, so that the
    // 32 parents of a wave mostly share it and the wave skips the classes
    // none of them has
    T8m.resize(T8.size());
    for (int c8 = 0; c8 < 8; c8++)
      for (int pn = 0; pn < n3; pn++)
        for (int cn = 0; cn < n3; cn++)
          T8m[((size_t)c8 * n3 + cn) * n3 + pn] = T8[((size_t)c8 * n3 + pn) * n3 + cn];
    BB_TRY(upload(&p->d_T8m, T8m.data(), T8m.size(), stream));
    // (a stable counting sort by the 8-bit child mask, per level)
    mplist.resize(f.plist.size());
    for (size_t l = 0; l + 1 < f.ploff.size(); l++) {
      const int b = f.ploff[l], e = f.ploff[l + 1];
      std::vector<unsigned char> mk((size_t)(e - b));
      int cnt[257] = {0};
      for (int i = b; i < e; i++) {
        const int P = f.plist[i];
        int m = 0;
        for (int c = 0; c < 8; c++) m |= (f.child[(size_t)8 * P + c] >= 0) << c;
        mk[i - b] = (unsigned char)m;
        cnt[m + 1]++;
      }
      for (int m = 0; m < 256; m++) cnt[m + 1] += cnt[m];
      for (int i = b; i < e; i++) mplist[b + cnt[mk[i - b]]++] = f.plist[i];
    }
    if (!mplist.empty()) BB_TRY(upload(&p->d_mplist, mplist.data(), mplist.size(), stream));
  }
#endif
  if (!f.complete && (long long)N < 12LL * nleaf) BB_TRY(dalloc(&p->d_pq, (size_t)N));
  const size_t NR = (size_t)N * nrhs;
  BB_TRY(dalloc(&p->d_q, NR));
  BB_TRY(dalloc(&p->d_phi, NR));
  BB_TRY(dalloc(&p->d_farfield, (size_t)N));
  const int W = p->width;
  BB_TRY(dalloc(&p->d_S, (size_t)ncell * n3 * W));
  BB_TRY(dalloc(&p->d_F, (size_t)ncell * n3 * W));
  BB_TRY(dalloc(&p->d_M, (size_t)ncell * r * W));
  BB_TRY(dalloc(&p->d_L, (size_t)ncell * r * W));
  // zero: the vectors past nrhs stay zero in the forms that compute nrhs
  BB_HIP(hipMemsetAsync(p->d_M, 0, sizeof(real) * (size_t)ncell * r * W, stream));
  BB_HIP(hipMemsetAsync(p->d_L, 0, sizeof(real) * (size_t)ncell * r * W, stream));
  BB_HIP(hipMemsetAsync(p->d_S, 0, sizeof(real) * (size_t)ncell * n3 * W, stream));
  BB_HIP(hipMemsetAsync(p->d_F, 0, sizeof(real) * (size_t)ncell * n3 * W, stream));
  BB_HIP(hipStreamSynchronize(stream));
  return BBFMM_OK;
}

static int plan_init(bbfmm_amd_plan *p, const anotherTree *t, const real3 *pos, int n, int r, const real *K,
                     const real *U, const real *Tkz, const real *Kw, const real *Cw, real homogen, int nrhs,
                     hipStream_t stream)
{
  if (t->maxlev < kMinLev) return set_error(BBFMM_E_ARG, "maxlev=%d < %d", t->maxlev, kMinLev);
  HostForest f;
  BB_TRY(forest_from_tree(f, t));
  return plan_init_forest(p, f, pos, n, r, K, U, Tkz, Kw, Cw, homogen, nrhs, stream);
}

// Particle-row form of a solve: 16 vectors in an MFMA form (fp32 build); the
// charges and potentials move as rows of 16 in the padded leaf order that the
// MFMA P2P reads.
static bool rows16(const bbfmm_amd_plan *p)
{
#ifdef SINGLE
  return p->form != BBFMM_M2L_PLAIN && p->width == 16 && p->npad > 0;
#else
  (void)p;
  return false;
#endif
}

static PassArgs pass_args(const bbfmm_amd_plan *p, bool rows = false)
{
  PassArgs a;
  a.N = p->N;
  a.n = p->n;
  a.r = p->r;
  a.ncell = p->ncell;
  a.maxlev = p->maxlev;
  a.nleaf = p->nleaf;
  const real pf = (real)2 / (real)p->n;
  a.prefac3 = pf * pf * pf;
  a.pos = p->d_pos;
  a.center = p->d_center;
  a.leafsta = p->d_leafsta;
  a.leafcnt = p->d_leafcnt;
  a.leafof = p->d_leafof;
  a.leafcell = p->d_leafcell;
  a.leafihalf = p->d_leafihalf;
  a.celllev = p->d_celllev;
  a.child = p->d_child;
  a.plist = p->d_plist;
  a.scale = p->d_scale;
  a.Tkz = p->d_Tkz;
  a.Kw = p->d_Kw;
  a.Cw = p->d_Cw;
  a.Ucm = p->d_Ucm;
  a.Urm = p->d_Urm;
  a.q = p->d_q;
  a.phi = p->d_phi;
  a.S = p->d_S;
  a.M = p->d_M;
  a.L = p->d_L;
  a.F = p->d_F;
  a.potential = p->potential;
  a.l2p_ref = p->l2p_ref;
#ifdef SINGLE
  a.cols = !p->complete && p->n == 4 && p->r == 32;
  a.dcp = p->d_dcp;
  a.dwv = p->d_dwv;
  a.KU = p->d_KU;
  a.T8 = p->d_T8;
  a.T8m = p->d_T8m;
  a.mplist = p->d_mplist;
#endif
  if (rows) {
    a.rows = 1;
    a.qt = p->d_qt;
    a.phit = p->d_phit;
    a.pad_of = p->d_pad_of;
    a.farfield = p->d_farfield;
  }
  return a;
}

// the dense grids of the CL / IJ schemes (created on first use: allocates,
// uploads the operator layout and synchronises the stream)
static int ensure_grid(bbfmm_amd_plan *p)
{
  if (p->scheme != M2L_SCHEME_CL && p->scheme != M2L_SCHEME_IJ) return BBFMM_OK;
  if (p->scheme == M2L_SCHEME_CL && p->complete && p->r <= kClusterMaxRank) return BBFMM_OK;  // cluster kernel
  if (!p->complete) {
    // sparse forest: the list MFMA form of the IJ blocking
    if (p->scheme != M2L_SCHEME_IJ) return set_error(BBFMM_E_ARG, "scheme CL needs a complete tree");
    const int sw = p->level_switch > kMinLev ? p->level_switch : kMinLev;
    if (!p->list || p->list_form != p->form || p->list_sw != sw) {
      if (p->list) list_m2l_destroy(p->list);
      p->list = nullptr;
      const int *off = p->h_cls_off, *cells = p->d_cls_cells;
      if (sw > kMinLev) {
        // the class lists restricted to levels >= sw (same order)
        std::vector<int> cc;
        int o[9] = {};
        for (int c = 0; c < 8; c++) {
          o[c] = (int)cc.size();
          for (int A = 0; A < p->ncell; A++)
            if (p->h_celllev[A] >= sw && p->h_cellcls[A] == c) cc.push_back(A);
        }
        o[8] = (int)cc.size();
        if (p->d_cls_cells_sw) (void)hipFree(p->d_cls_cells_sw);
        p->d_cls_cells_sw = nullptr;
        BB_TRY(upload(&p->d_cls_cells_sw, cc.data(), cc.size(), p->stream));
        BB_HIP(hipStreamSynchronize(p->stream));  // the host list must outlive the copy
        for (int c = 0; c < 9; c++) p->h_cls_off_sw[c] = o[c];
        off = p->h_cls_off_sw;
        cells = p->d_cls_cells_sw;
      }
      BB_TRY(list_m2l_create(&p->list, p->r, p->nrhs, p->form, p->h_K.data(), p->ncell, off, cells, p->d_vsrc,
                             p->stream));
      p->list_form = p->form;
      p->list_sw = sw;
    }
    list_m2l_set_flat(p->list, p->flat_loads);
    return BBFMM_OK;
  }
  const int form = p->scheme == M2L_SCHEME_CL ? BBFMM_M2L_PLAIN : p->form;
  const int minlev = p->level_switch > kMinLev ? p->level_switch : kMinLev;
  const int variant = p->scheme == M2L_SCHEME_IJ ? p->m2l_variant : 0;
  const int variant_up = p->scheme == M2L_SCHEME_IJ ? p->m2l_variant_up : 0;
  const int upper_split = p->m2l_upper_split;
  if (!p->grid || p->grid_form != form || p->grid_minlev != minlev || p->grid_variant != variant ||
      p->grid_variant_up != variant_up || p->grid_upper_split != upper_split) {
    if (p->grid) grid_m2l_destroy(p->grid);
    p->grid = nullptr;
    // the MFMA forms on the plan's width (the particle-row form's vectors)
    const int gv = form != BBFMM_M2L_PLAIN ? p->width : p->nrhs;
    BB_TRY(grid_m2l_create(&p->grid, p->r, p->maxlev, gv, form, p->h_K.data(), p->stream, minlev, variant,
                           variant_up, upper_split));
    p->grid_form = form;
    p->grid_minlev = minlev;
    p->grid_variant = variant;
    p->grid_variant_up = variant_up;
    p->grid_upper_split = upper_split;
  }
  grid_m2l_set_flat(p->grid, p->flat_loads);
  return BBFMM_OK;
}

// everything a solve allocates on first use, so that a solve itself (and a
// captured graph of it) allocates and clears nothing
static int prepare_solve(bbfmm_amd_plan *p)
{
  BB_TRY(ensure_grid(p));
#ifdef SINGLE
  if (p->form != BBFMM_M2L_PLAIN && p->width >= 16 && !p->d_qt) {
    const size_t cnt = (size_t)(p->npad > 0 ? p->npad : 1) * 16;
    BB_TRY(dalloc(&p->d_qt, cnt));
    BB_HIP(hipMemsetAsync(p->d_qt, 0, sizeof(real) * cnt, p->stream));
  }
  if (rows16(p) && !p->d_phit) BB_TRY(dalloc(&p->d_phit, (size_t)p->npad * 16));
  if (p->p2p_split && p->form != BBFMM_M2L_PLAIN && p->width >= 16 && p->npad > 0 && !p->d_qsh) {
    const size_t bytes = sizeof(__bf16) * (size_t)p->npad * 16;
    BB_HIP(hipMalloc(&p->d_qsh, bytes));
    BB_HIP(hipMalloc(&p->d_qsl, bytes));
  }
  if (rows16(p) && p->near_stream) {
    if (!p->d_phin) BB_TRY(dalloc(&p->d_phin, (size_t)p->npad * 16));
    if (!p->side) BB_HIP(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
    if (!p->ev_fork) BB_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    if (!p->ev_join) BB_HIP(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
  }
#endif
  return BBFMM_OK;
}

// deep_in_grid: the deepest level's multipoles were written into the
// split-BF16 grids by the fused first upward level
static int run_m2l(bbfmm_amd_plan *p, const real *M, real *L, hipEvent_t mid = nullptr, bool deep_in_grid = false)
{
  M2LArgs m;
  m.r = p->r;
  m.ncell = p->ncell;
  m.nrhs = p->nrhs;
  m.minlev = kMinLev;
  m.maxlev = p->maxlev;
  m.h_levsta = p->complete ? p->h_levsta.data() : nullptr;
  m.iinter = p->d_iinter;
  m.inter = p->d_inter;
  m.kidx = p->d_kidx;
  m.K4 = p->d_K4;
  m.Kc = p->d_Kc;
  m.opoff = p->d_opoff;
  m.Ktable = p->d_Ktable;
  m.M = M;
  m.L = L;
  m.variant = p->scheme == M2L_SCHEME_BA ? p->m2l_variant : 0;
  if (!p->complete) {
    m.vsrc = p->d_vsrc;
    m.cellcls = p->d_cellcls;
    m.slotop = p->d_slotop;
  }
  if (!p->complete && p->scheme == M2L_SCHEME_IJ) {
    BB_TRY(ensure_grid(p));
    if (p->level_switch > kMinLev) {
      // the levels below the switch on the basic scheme (sources by class slot)
      M2LArgs b = m;
      b.variant = 0;
      b.celllev = p->d_celllev;
      b.lev_hi = p->level_switch - 1;
      BB_TRY(launch_m2l_basic(b, p->stream));
    }
    return list_m2l_run(p->list, M, L, p->stream, mid);
  }
  if (!p->complete && p->scheme != M2L_SCHEME_BA)
    return set_error(BBFMM_E_ARG, "a batch of sparse instances runs the basic or the IJ (list MFMA) scheme");
  switch (p->scheme) {
  case M2L_SCHEME_BA:
    if (mid) BB_HIP(hipEventRecord(mid, p->stream));
    return launch_m2l_basic(m, p->stream);
  case M2L_SCHEME_SI:
    if (mid) BB_HIP(hipEventRecord(mid, p->stream));
    return launch_m2l_sibling(m, p->stream);
  case M2L_SCHEME_CL:
    if (p->r <= kClusterMaxRank) {
      // levels below the switch on the basic scheme, as for IJ
      if (p->level_switch > kMinLev) {
        M2LArgs b = m;
        b.maxlev = p->level_switch - 1;
        b.variant = 0;
        BB_TRY(launch_m2l_basic(b, p->stream));
        m.minlev = p->level_switch;
      }
      if (mid) BB_HIP(hipEventRecord(mid, p->stream));
      return launch_m2l_cluster(m, p->stream);
    }
    // larger ranks: the operators do not fit LDS whole; the dense grids (real precision)
    [[fallthrough]];
  case M2L_SCHEME_IJ:
    BB_TRY(ensure_grid(p));
    if (p->level_switch > kMinLev) {
      // levels below the switch on the basic scheme (m2l_drv.cu:39-60 splits
      // the levels the same way)
      M2LArgs b = m;
      b.maxlev = p->level_switch - 1;
      b.variant = 0;
      BB_TRY(launch_m2l_basic(b, p->stream));
    }
    return grid_m2l_run(p->grid, M, L, p->ncell, p->stream, mid, deep_in_grid ? 1 : 0);
  default: return set_error(BBFMM_E_ARG, "unknown M2L scheme %d", p->scheme);
  }
}

// The solve pipeline on device arrays in leaf order (d_q -> d_phi).
// rows: the charges are in d_qt as particle rows (the particle-row form);
// otherwise in d_q [nrhs][N] leaf order, the potentials going to d_phi
static P2PArgs p2p_args(const bbfmm_amd_plan *p)
{
  P2PArgs pp;
  pp.N = p->N;
  pp.nleaf = p->nleaf;
  pp.nrhs = p->nrhs;
  pp.pos = p->d_pos;
  pp.leafsta = p->d_leafsta;
  pp.leafcnt = p->d_leafcnt;
  pp.nbr = p->d_nbr;
  pp.nnbr = p->d_nnbr;
  pp.q = p->d_q;
  pp.phi = p->d_phi;
  pp.use_mfma = p->form != BBFMM_M2L_PLAIN;
  pp.potential = p->potential;
  pp.ppos = p->d_ppos;
  pp.seg = p->d_seg;
  pp.nseg = p->d_nseg;
  pp.pad_of = p->d_pad_of;
  pp.npad = p->npad;
  pp.flat = p->flat_loads;
  if (p->p2p_split && p->d_qsh) {
    pp.qsh = p->d_qsh;
    pp.qsl = p->d_qsl;
  }
  // sparse forests with a few particles per leaf: a lane per target particle
  if (!p->complete && p->N < 12LL * p->nleaf) {
    pp.leafof = p->d_leafof;
    pp.pq = p->d_pq;
  }
  return pp;
}

// the particle-row form with P2P beside the far field
static bool near_beside(const bbfmm_amd_plan *p, bool rows)
{
  return rows && p->near_stream && p->side && p->d_phin;
}

static int solve_sorted(bbfmm_amd_plan *p, bool tm, bool rows)
{
  hipStream_t s = p->stream;
  PassArgs a = pass_args(p, rows);
  P2PArgs pp = p2p_args(p);
  const int nv = rows ? p->width : p->nrhs;  // vectors this solve computes
  pp.nrhs = nv;
#ifdef SINGLE
  if (pp.use_mfma && nv >= 16) {
    if (!p->d_qt) return set_error(BBFMM_E_ARG, "internal: solve not prepared");
    pp.qt = p->d_qt;
    if (rows) pp.phit = p->d_phit;
  }
#endif
  const bool beside = near_beside(p, rows);  // particle rows: fp32 build only
  // fork after the M2L: P2P (compute-bound) then shares the GPU with the
  // latency-bound downward pass
  auto fork_near = [&]() -> int {
#ifdef SINGLE
    if (beside) {
      pp.phit = nullptr;
      pp.phin = p->d_phin;
      BB_HIP(hipEventRecord(p->ev_fork, s));
      BB_HIP(hipStreamWaitEvent(p->side, p->ev_fork, 0));
      if (tm) BB_HIP(hipEventRecord(stage_event(p, kEvSideA), p->side));
      BB_TRY(launch_p2p(pp, p->side));
      if (tm) BB_HIP(hipEventRecord(stage_event(p, kEvSideB), p->side));
      BB_HIP(hipEventRecord(p->ev_join, p->side));
    }
#endif
    return BBFMM_OK;
  };
  if (tm) p->side_timed = beside;
  // the first upward level (P2M, M2M into the leaves' parents) and the last
  // downward level (post-M2L of the leaves, L2L into them, L2P) each in one
  // kernel: the leaves' S and F are then not stored (the F fetch recomputes
  // post-M2L from L)
  const bool fuse = rows && p->n == 4 && p->r == 32 && p->complete && p->maxlev >= 3;
  if (fuse) {
    a.cell_end = p->h_levsta[p->maxlev];
    a.l2l_end = p->maxlev - 1;
    a.m2m_start = p->maxlev - 2;
  }
  // the split-BF16 dense-grid M2L: the fused first upward level writes the
  // leaves' multipoles into the grids too, and the M2L scatters only the
  // levels above
  bool leaves_in_grid = false;
  if (fuse && p->scheme == M2L_SCHEME_IJ && p->grid &&
      grid_m2l_leaf_grids(p->grid, &a.grid_hi, &a.grid_lo, &a.grid_slot0))
    leaves_in_grid = true;
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_P2M), s));
  if (fuse)  // P2M and the first M2M level in one kernel
    BB_TRY(launch_up_leaf(a, p->h_ploff.data(), s));
  else
    BB_TRY(launch_p2m(a, s, nv));
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_M2M), s));
  BB_TRY(launch_m2m(a, p->h_ploff.data(), s, nv));
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_M2L_PREP), s));
  BB_TRY(run_m2l(p, p->d_M, p->d_L, tm ? stage_event(p, ST_M2L) : nullptr, leaves_in_grid));
  BB_TRY(fork_near());
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_POST), s));
  if (a.dcp) {
    // sparse forests: post-M2L and L2L level by level in one product per cell
    BB_TRY(launch_down_cols(a, p->h_dwv_off.data(), s, nv));
    if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_L2L), s));
  } else {
    BB_TRY(launch_postm2l(a, s, nv));
    if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_L2L), s));
    BB_TRY(launch_l2l(a, p->h_ploff.data(), s, nv));
  }
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_L2P), s));
  if (fuse)
    BB_TRY(launch_down_leaf(a, p->h_ploff.data(), s));
  else
    BB_TRY(launch_l2p(a, s, nv));
  if (!rows)  // the particle-row L2P writes the far field itself
    BB_HIP(hipMemcpyAsync(p->d_farfield, p->d_phi, sizeof(real) * p->N, hipMemcpyDeviceToDevice, s));
  // "p2p" stage: P2P itself, or with the near field beside the far field the
  // wait for it after the far field is done (its own time: ST_TOTAL + 1)
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_P2P), s));
  if (beside)
    BB_HIP(hipStreamWaitEvent(s, p->ev_join, 0));  // join
  else
    BB_TRY(launch_p2p(pp, s));
  return BBFMM_OK;
}

static void drop_graph(bbfmm_amd_plan *p)
{
  if (p->gexec) {
    (void)hipStreamSynchronize(p->stream);
    (void)hipGraphExecDestroy(p->gexec);
  }
  p->gexec = nullptr;
  p->gq = nullptr;
  p->gphi = nullptr;
}

static int solve_eager(bbfmm_amd_plan *p, const real *q_dev, real *phi_dev, bool tm);

// Capture gather, solve and scatter for (q_dev, phi_dev) on a private
// stream into one graph; replay it on the plan's stream.
static int solve_graph(bbfmm_amd_plan *p, const real *q_dev, real *phi_dev)
{
  if (!p->gexec || p->gq != q_dev || p->gphi != phi_dev) {
    drop_graph(p);
    BB_TRY(prepare_solve(p));
    BB_HIP(hipStreamSynchronize(p->stream));
    if (!p->gstream) BB_HIP(hipStreamCreateWithFlags(&p->gstream, hipStreamNonBlocking));
    hipStream_t user = p->stream;
    p->stream = p->gstream;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(p->gstream, hipStreamCaptureModeThreadLocal);
    int rc = e == hipSuccess ? solve_eager(p, q_dev, phi_dev, false)
                             : set_error(BBFMM_E_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    if (e == hipSuccess) {
      const hipError_t e2 = hipStreamEndCapture(p->gstream, &g);
      if (rc == BBFMM_OK && e2 != hipSuccess)
        rc = set_error(BBFMM_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e2));
    }
    p->stream = user;
    if (rc == BBFMM_OK) {
      const hipError_t e3 = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
      if (e3 != hipSuccess) {
        p->gexec = nullptr;
        rc = set_error(BBFMM_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e3));
      }
    }
    if (g) (void)hipGraphDestroy(g);
    BB_TRY(rc);
    p->gq = q_dev;
    p->gphi = phi_dev;
  }
  const bool tm = p->nsets > 0;
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_GATHER), p->stream));
  BB_HIP(hipGraphLaunch(p->gexec, p->stream));
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_TOTAL), p->stream));
  p->solves++;
  p->solved = true;
  return BBFMM_OK;
}

static int solve_dev(bbfmm_amd_plan *p, const real *q_dev, real *phi_dev)
{
  if (p->use_graph) return solve_graph(p, q_dev, phi_dev);
  BB_TRY(prepare_solve(p));
  BB_TRY(solve_eager(p, q_dev, phi_dev, p->nsets > 0));
  p->solves++;
  p->solved = true;
  return BBFMM_OK;
}

static int solve_eager(bbfmm_amd_plan *p, const real *q_dev, real *phi_dev, bool tm)
{
  hipStream_t s = p->stream;
  const bool rows = rows16(p);
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_GATHER), s));
  if (rows)
    BB_TRY(launch_gather_rows16(p->d_ipad, q_dev, p->nrhs, p->d_qt, p->N, s));
  else
    BB_TRY(launch_gather(p->d_perm, q_dev, p->d_q, p->N, p->nrhs, s));
  BB_TRY(solve_sorted(p, tm, rows));
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_SCATTER), s));
  if (rows) {
    const bool beside = near_beside(p, rows);
    BB_TRY(launch_scatter_rows16(p->d_ipad, p->d_phit, beside ? p->d_phin : nullptr, p->d_qt, p->potential, phi_dev,
                                 p->nrhs, p->N, s));
  }
  else
    BB_TRY(launch_scatter(p->d_iperm, p->d_phi, phi_dev, p->N, p->nrhs, s));
  if (tm) BB_HIP(hipEventRecord(stage_event(p, ST_TOTAL), s));
  return BBFMM_OK;
}

} // namespace bb

using namespace bb;

// ---------------------------------------------------------------- plan API
extern "C" int bbfmm_amd_maxlev(int N, int lmod) { return maxlev_rule(N, lmod); }

extern "C" void bbfmm_amd_weights(int n, real *Tkz, int *Ktable, real *Kweights, real *Cweights)
{
  weights(n, Tkz, Ktable, Kweights, Cweights);
}

extern "C" int bbfmm_amd_plan_create(bbfmm_amd_plan **plan, int N, const real3 *pos, real L, int n, int lmod,
                                     int r, const real *K, const real *U, int nrhs, void *stream)
{
  clear_error();
  if (!plan || (!pos && N > 0) || !K || !U || N < 0) return set_error(BBFMM_E_ARG, "null argument");
  *plan = nullptr;
  if (n < kMinOrder || n > kMaxOrder)
    return set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", n, kMinOrder, kMaxOrder);
  anotherTree *t = (anotherTree *)calloc(1, sizeof(anotherTree));
  const int maxlev = maxlev_rule(N > 0 ? N : 1, lmod);
  int rc = build_tree(t, L, maxlev, N, pos);
  if (rc != BBFMM_OK) {
    free_tree(t);
    return set_error(rc, "a particle lies outside the root cell [-L/2, L/2]^3");
  }
  std::vector<real> Tkz((size_t)n * n), Kw((size_t)n * n * n), Cw((size_t)2 * n * n);
  int Kt[343];
  weights(n, Tkz.data(), Kt, Kw.data(), Cw.data());
  bbfmm_amd_plan *p = new bbfmm_amd_plan();
  rc = plan_init(p, t, pos, n, r, K, U, Tkz.data(), Kw.data(), Cw.data(), (real)1, nrhs, (hipStream_t)stream);
  free_tree(t);
  if (rc != BBFMM_OK) {
    plan_free_device(p);
    delete p;
    return rc;
  }
  *plan = p;
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_batch_create(bbfmm_amd_plan **plan, int ninst, const int *N, const real3 *pos,
                                     const int *depth, int leafmax, real L, int n, int r, const real *K,
                                     const real *U, int nrhs, void *stream, int *depth_out)
{
  clear_error();
  if (!plan || !N || !pos || !K || !U || ninst < 1) return set_error(BBFMM_E_ARG, "null argument");
  *plan = nullptr;
  if (n < kMinOrder || n > kMaxOrder)
    return set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", n, kMinOrder, kMaxOrder);
  HostForest f;
  std::vector<int> dep;
  BB_TRY(forest_from_points(f, ninst, N, pos, depth, leafmax, L, &dep));
  std::vector<real> Tkz((size_t)n * n), Kw((size_t)n * n * n), Cw((size_t)2 * n * n);
  int Kt[343];
  weights(n, Tkz.data(), Kt, Kw.data(), Cw.data());
  bbfmm_amd_plan *p = new bbfmm_amd_plan();
  const int rc = plan_init_forest(p, f, pos, n, r, K, U, Tkz.data(), Kw.data(), Cw.data(), (real)1, nrhs,
                                  (hipStream_t)stream);
  if (rc != BBFMM_OK) {
    plan_free_device(p);
    delete p;
    return rc;
  }
  p->scheme = M2L_SCHEME_BA;
  if (depth_out)
    for (int i = 0; i < ninst; i++) depth_out[i] = dep[i];
  *plan = p;
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_plan_destroy(bbfmm_amd_plan *p)
{
  if (!p) return BBFMM_OK;
  (void)hipStreamSynchronize(p->stream);
  plan_free_device(p);
  delete p;
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_plan_set(bbfmm_amd_plan *p, int option, int value)
{
  clear_error();
  if (!p) return set_error(BBFMM_E_ARG, "null plan");
  switch (option) {
  case BBFMM_OPT_SCHEME:
    if (value < M2L_SCHEME_BA || value > M2L_SCHEME_IJ) return set_error(BBFMM_E_ARG, "unknown M2L scheme %d", value);
    if (!p->complete && value != M2L_SCHEME_BA && value != M2L_SCHEME_IJ)
      return set_error(BBFMM_E_ARG, "a batch of sparse instances runs the basic or the IJ (list MFMA) scheme");
    p->scheme = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_FORM:
    if (value < BBFMM_M2L_PLAIN || value > BBFMM_M2L_FP16X3) return set_error(BBFMM_E_ARG, "unknown M2L form %d", value);
#ifndef SINGLE
    if (value != BBFMM_M2L_PLAIN) return set_error(BBFMM_E_ARG, "MFMA forms are fp32-build only");
#endif
    p->form = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_TIMING:
    if (value < 0 || value > 4096) return set_error(BBFMM_E_ARG, "timing sets %d out of range", value);
    p->timing = value;
    return set_timing(p, value);
  case BBFMM_OPT_OUTPUT:
    if (value != 0 && value != 1) return set_error(BBFMM_E_ARG, "output %d: 0 (q .* potential) or 1 (potential)", value);
    p->potential = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_ADDRESSING:
    if (value != 0 && value != 1) return set_error(BBFMM_E_ARG, "addressing %d: 0 (automatic) or 1 (pointers)", value);
    p->flat_loads = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_L2P_REFERENCE:
    if (value != 0 && value != 1) return set_error(BBFMM_E_ARG, "L2P form %d: 0 (interpolant) or 1 (reference)", value);
    p->l2p_ref = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_M2L_VARIANT_UPPER:
    if (value != 0 && value != 1 && (value < 4 || value > 7))
      return set_error(BBFMM_E_ARG, "M2L upper-level variant %d: 0, 1 or 4 to 7", value);
    p->m2l_variant_up = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_M2L_UPPER_SPLIT:
    if (value < 1 || value > 16) return set_error(BBFMM_E_ARG, "M2L upper-level step parts %d: 1 to 16", value);
    p->m2l_upper_split = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_M2L_VARIANT:
    if (value < 0 || value > 8) return set_error(BBFMM_E_ARG, "M2L variant %d out of range", value);
    p->m2l_variant = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_NEAR_STREAM:
    if (value != 0 && value != 1) return set_error(BBFMM_E_ARG, "near stream %d: 0 or 1", value);
    p->near_stream = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_P2P_FORM:
    if (value != 0 && value != 1) return set_error(BBFMM_E_ARG, "P2P form %d: 0 (fp32) or 1 (split-BF16)", value);
    p->p2p_split = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_LEVEL_SWITCH:
    if (value < kMinLev || value > p->maxlev)
      return set_error(BBFMM_E_ARG, "level switch %d outside [%d, %d]", value, kMinLev, p->maxlev);
    p->level_switch = value;
    drop_graph(p);
    return BBFMM_OK;
  case BBFMM_OPT_GRAPH:
    if (value != 0 && value != 1) return set_error(BBFMM_E_ARG, "graph %d: 0 (eager) or 1 (captured)", value);
    drop_graph(p);
    p->use_graph = value;
    if (p->nsets > 0) return set_timing(p, p->nsets);
    return BBFMM_OK;
  default: return set_error(BBFMM_E_ARG, "unknown option %d", option);
  }
}

// M2L autotuner: times every configuration valid for the plan's tree and
// arithmetic form on the plan's own multipoles and keeps the fastest.
constexpr int kTuneCols = 5;  // scheme, variant, upper-level variant, level switch, ms

extern "C" int bbfmm_amd_plan_tune(bbfmm_amd_plan *p, int reps, double *table, int maxrows, int *nrows)
{
  clear_error();
  if (!p) return set_error(BBFMM_E_ARG, "null plan");
  if (reps < 1) reps = 3;
  struct Cand {
    int scheme, variant, sw, vup;
  };
  std::vector<Cand> cands;
  if (!p->complete) {
    for (int v = 0; v < 3; v++) cands.push_back({M2L_SCHEME_BA, v, kMinLev, 0});
#ifdef SINGLE
    if ((p->form == BBFMM_M2L_BF16X2 || p->form == BBFMM_M2L_FP32MFMA) && p->r == 32)
      for (int sw = kMinLev; sw <= std::min(p->maxlev, kMinLev + 2); sw++)  // list M2L from level sw, basic below
        cands.push_back({M2L_SCHEME_IJ, 0, sw, 0});
#endif
  } else {
    std::vector<int> sws{kMinLev};
    if (p->maxlev > kMinLev) sws.push_back(kMinLev + 1);
    if (p->form == BBFMM_M2L_PLAIN) {
      for (int v = 0; v < 3; v++) cands.push_back({M2L_SCHEME_BA, v, kMinLev, 0});
      cands.push_back({M2L_SCHEME_SI, 0, kMinLev, 0});
      if (p->r <= kClusterMaxRank) cands.push_back({M2L_SCHEME_CL, 0, kMinLev, 0});
      for (int sw : sws) cands.push_back({M2L_SCHEME_IJ, 0, sw, 0});
    } else if (p->form == BBFMM_M2L_BF16X2) {
      for (int sw : sws)
        for (int v = 1; v <= 5; v++) cands.push_back({M2L_SCHEME_IJ, v, sw, 0});
      // source-stationary with a class-group size of its own above the deepest level
      if (p->maxlev > kMinLev)
        for (int v : {1, 4, 5})
          for (int u : {1, 4, 5})
            if (u != v) cands.push_back({M2L_SCHEME_IJ, v, kMinLev, u});
    } else if (p->r == 32) {
      for (int sw : sws)
        for (int v = 2; v <= 3; v++) cands.push_back({M2L_SCHEME_IJ, v, sw, 0});
    } else {
      for (int sw : sws) cands.push_back({M2L_SCHEME_IJ, 0, sw, 0});
    }
  }
  const int s0 = p->scheme, v0 = p->m2l_variant, w0 = p->level_switch, u0 = p->m2l_variant_up;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BB_HIP(hipEventCreate(&e0));
  BB_HIP(hipEventCreate(&e1));
  int best = -1;
  double best_ms = 0;
  int rc = BBFMM_OK, rows = 0;
  for (size_t c = 0; c < cands.size() && rc == BBFMM_OK; c++) {
    p->scheme = cands[c].scheme;
    p->m2l_variant = cands[c].variant;
    p->level_switch = cands[c].sw;
    p->m2l_variant_up = cands[c].vup;
    rc = run_m2l(p, p->d_M, p->d_L);  // warm-up (creates the grids)
    if (rc != BBFMM_OK) break;
    if (hipEventRecord(e0, p->stream) != hipSuccess) rc = set_error(BBFMM_E_HIP, "event record");
    for (int k = 0; k < reps && rc == BBFMM_OK; k++) rc = run_m2l(p, p->d_M, p->d_L);
    if (rc != BBFMM_OK) break;
    float ms = 0;
    if (hipEventRecord(e1, p->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = set_error(BBFMM_E_HIP, "tuner timing failed");
      break;
    }
    const double per = ms / reps;
    if (table && rows < maxrows) {
      double *row = table + (size_t)kTuneCols * rows;
      row[0] = cands[c].scheme;
      row[1] = cands[c].variant;
      row[2] = cands[c].vup;
      row[3] = cands[c].sw;
      row[4] = per;
    }
    rows++;
    if (best < 0 || per < best_ms) {
      best = (int)c;
      best_ms = per;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (nrows) *nrows = rows;
  if (rc != BBFMM_OK || best < 0) {
    p->scheme = s0;
    p->m2l_variant = v0;
    p->level_switch = w0;
    p->m2l_variant_up = u0;
    return rc != BBFMM_OK ? rc : set_error(BBFMM_E_ARG, "no M2L configuration to tune");
  }
  p->scheme = cands[best].scheme;
  p->m2l_variant = cands[best].variant;
  p->level_switch = cands[best].sw;
  p->m2l_variant_up = cands[best].vup;
  drop_graph(p);
  BB_TRY(ensure_grid(p));
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_plan_info(const bbfmm_amd_plan *p, int *out, double *dout)
{
  if (!p) return set_error(BBFMM_E_ARG, "null plan");
  if (out) {
    out[0] = p->maxlev;
    out[1] = p->ncell;
    out[2] = p->N;
    out[3] = p->r;
    out[4] = p->n;
    out[5] = p->nrhs;
  }
  if (dout) {
    dout[0] = p->m2l_interactions;
    dout[1] = p->p2p_pairs;
  }
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_solve(bbfmm_amd_plan *p, const real *q_dev, real *phi_dev)
{
  clear_error();
  if (!p || !q_dev || !phi_dev) return set_error(BBFMM_E_ARG, "null argument");
  return solve_dev(p, q_dev, phi_dev);
}

extern "C" int bbfmm_amd_solve_host(bbfmm_amd_plan *p, const real *q, real *phi)
{
  clear_error();
  if (!p || !q || !phi) return set_error(BBFMM_E_ARG, "null argument");
  const size_t NR = (size_t)p->N * p->nrhs;
  if (!p->d_qin) BB_TRY(dalloc(&p->d_qin, NR));
  if (!p->d_phiout) BB_TRY(dalloc(&p->d_phiout, NR));
  BB_HIP(hipMemcpyAsync(p->d_qin, q, sizeof(real) * NR, hipMemcpyHostToDevice, p->stream));
  BB_TRY(solve_dev(p, p->d_qin, p->d_phiout));
  BB_HIP(hipMemcpyAsync(phi, p->d_phiout, sizeof(real) * NR, hipMemcpyDeviceToHost, p->stream));
  BB_HIP(hipStreamSynchronize(p->stream));
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_plan_update_positions(bbfmm_amd_plan *p, const real3 *pos)
{
  clear_error();
  if (!p || !pos) return set_error(BBFMM_E_ARG, "null argument");
  const int N = p->N;
  const real half = p->L / 2;
  std::vector<real4> hpos((size_t)N), ppos((size_t)(p->npad > 0 ? p->npad : 1));
  for (auto &x : ppos) {
    x.x = x.y = x.z = (real)1e6;
    x.w = 0;
  }
  for (int i = 0; i < N; i++) {
    const real3 x = pos[p->h_perm[i]];
    const int k = p->h_leafof[i];
    const real inv = p->h_leafinv[k];
    const int ncpd = 1 << p->h_leafdepth[k];
    int ix = (int)((x.x + half) * inv), iy = (int)((x.y + half) * inv), iz = (int)((x.z + half) * inv);
    if (!(x.x >= -half && x.x <= half && x.y >= -half && x.y <= half && x.z >= -half && x.z <= half))
      return set_error(BBFMM_E_RANGE, "a particle lies outside the root cell [-L/2, L/2]^3");
    ix = ix < ncpd - 1 ? ix : ncpd - 1;
    iy = iy < ncpd - 1 ? iy : ncpd - 1;
    iz = iz < ncpd - 1 ? iz : ncpd - 1;
    const int *c = &p->h_leafxyz[(size_t)3 * k];
    if (ix != c[0] || iy != c[1] || iz != c[2])
      return set_error(BBFMM_E_RANGE, "particle %d moved to another leaf: build a new plan", p->h_perm[i]);
    hpos[i] = real4{x.x, x.y, x.z, 0};
    ppos[p->h_pad_of[i]] = hpos[i];
  }
  BB_HIP(hipMemcpyAsync(p->d_pos, hpos.data(), sizeof(real4) * (size_t)N, hipMemcpyHostToDevice, p->stream));
  BB_HIP(hipMemcpyAsync(p->d_ppos, ppos.data(), sizeof(real4) * ppos.size(), hipMemcpyHostToDevice, p->stream));
  BB_HIP(hipStreamSynchronize(p->stream));
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_gradient(bbfmm_amd_plan *p, const real *q_dev, const real *g_dev, real *dq_dev,
                                  real *dx_dev)
{
  clear_error();
  if (!p || !q_dev || !g_dev || !dq_dev || !dx_dev) return set_error(BBFMM_E_ARG, "null argument");
  if (p->nrhs % 2) return set_error(BBFMM_E_ARG, "gradient: the plan needs an even vector count (2m), got %d", p->nrhs);
  const int m = p->nrhs / 2;
  const int N = p->N;
  hipStream_t s = p->stream;
  if (!p->d_grad) BB_TRY(dalloc(&p->d_grad, (size_t)3 * N));
  // [q; g] in leaf order, one solve of the 2m vectors in potential mode
  BB_TRY(launch_gather(p->d_perm, q_dev, p->d_q, N, m, s));
  BB_TRY(launch_gather(p->d_perm, g_dev, p->d_q + (size_t)m * N, N, m, s));
  BB_TRY(prepare_solve(p));
  const int saved_potential = p->potential;
  p->potential = 1;
  const int rc = solve_sorted(p, false, false);  // not a timed solve
  p->potential = saved_potential;
  BB_TRY(rc);
  // dLoss/dq = A g: the potentials of the g vectors
  BB_TRY(launch_scatter(p->d_iperm, p->d_phi + (size_t)m * N, dq_dev, N, m, s));
  GradArgs ga;
  ga.N = N;
  ga.n = p->n;
  ga.nleaf = p->nleaf;
  ga.ncell = p->ncell;
  ga.m = m;
  const PassArgs pa = pass_args(p);
  ga.prefac3 = pa.prefac3;
  ga.leafcell = p->d_leafcell;
  ga.leafihalf = p->d_leafihalf;
  ga.pos = p->d_pos;
  ga.center = p->d_center;
  ga.leafsta = p->d_leafsta;
  ga.leafcnt = p->d_leafcnt;
  ga.nbr = p->d_nbr;
  ga.nnbr = p->d_nnbr;
  ga.Tkz = p->d_Tkz;
  ga.l2p_ref = p->l2p_ref;
  ga.F = p->d_F;
  ga.qg = p->d_q;
  ga.dx = p->d_grad;
  BB_TRY(launch_gradient(ga, p->d_iperm, dx_dev, s));
  p->solved = true;
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_m2l(bbfmm_amd_plan *p, const real *M_dev, real *L_dev)
{
  clear_error();
  if (!p || !M_dev || !L_dev) return set_error(BBFMM_E_ARG, "null argument");
  if (p->width == p->nrhs) return run_m2l(p, M_dev, L_dev);
  // a plan wider than its vectors: the nrhs vectors through the plan's own
  // M and L (the vectors past nrhs are zero there)
  const size_t cnt = (size_t)p->nrhs * p->ncell * p->r;
  BB_HIP(hipMemcpyAsync(p->d_M, M_dev, sizeof(real) * cnt, hipMemcpyDeviceToDevice, p->stream));
  BB_TRY(run_m2l(p, p->d_M, p->d_L));
  BB_HIP(hipMemcpyAsync(L_dev, p->d_L, sizeof(real) * cnt, hipMemcpyDeviceToDevice, p->stream));
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_fetch(bbfmm_amd_plan *p, int which, real *dst)
{
  clear_error();
  if (!p || !dst) return set_error(BBFMM_E_ARG, "null argument");
  if (!p->solved) return set_error(BBFMM_E_ARG, "no solve has run on this plan");
  BB_HIP(hipStreamSynchronize(p->stream));
  switch (which) {
  case 0: BB_HIP(hipMemcpy(dst, p->d_M, sizeof(real) * (size_t)p->ncell * p->r, hipMemcpyDeviceToHost)); break;
  case 1: BB_HIP(hipMemcpy(dst, p->d_L, sizeof(real) * (size_t)p->ncell * p->r, hipMemcpyDeviceToHost)); break;
  case 2: {
    // F after post-M2L: recompute it from L (the solve's F also holds L2L)
    const PassArgs a = pass_args(p);
    BB_TRY(launch_postm2l(a, p->stream, 1));
    BB_HIP(hipMemcpyAsync(dst, p->d_F, sizeof(real) * (size_t)p->ncell * p->n3, hipMemcpyDeviceToHost, p->stream));
    BB_HIP(hipStreamSynchronize(p->stream));
    break;
  }
  case 3: {
    std::vector<real> tmp((size_t)p->N);
    BB_HIP(hipMemcpy(tmp.data(), p->d_farfield, sizeof(real) * p->N, hipMemcpyDeviceToHost));
    for (int i = 0; i < p->N; i++) dst[p->h_perm[i]] = tmp[i];
    break;
  }
  default: return set_error(BBFMM_E_ARG, "unknown buffer %d", which);
  }
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_stage_times(bbfmm_amd_plan *p, double *out, int nout)
{
  if (!p || !out) return 0;
  if (read_timing(p) != BBFMM_OK) return 0;
  const int m = nout < ST_TOTAL + 2 ? nout : ST_TOTAL + 2;
  for (int k = 0; k < m; k++) out[k] = p->stage_ms[k];
  return m;
}

extern "C" int bbfmm_amd_operators(int n, int r, real *U, real *K, double *sigma)
{
  clear_error();
  if (n < kMinOrder || n > kMaxOrder)
    return set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", n, kMinOrder, kMaxOrder);
  const int n3 = n * n * n;
  if (r < 1 || r > n3) return set_error(BBFMM_E_ARG, "rank r=%d out of range", r);
  std::vector<double> Ud((size_t)n3 * r), Kd((size_t)kNumOps * r * r), sg((size_t)n3);
  BB_TRY(operators_fp64(n, r, Ud.data(), Kd.data(), sg.data()));
  for (size_t i = 0; i < Ud.size(); i++) U[i] = (real)Ud[i];
  for (size_t i = 0; i < Kd.size(); i++) K[i] = (real)Kd[i];
  if (sigma)
    for (int i = 0; i < n3; i++) sigma[i] = sg[i];
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_operators_device(int n, int r, real *U, real *K, double *sigma)
{
  clear_error();
  if (n < kMinOrder || n > kMaxOrder)
    return set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", n, kMinOrder, kMaxOrder);
  const int n3 = n * n * n;
  if (r < 1 || r > n3) return set_error(BBFMM_E_ARG, "rank r=%d out of range", r);
  BB_TRY(check_device());
  std::vector<double> Ud((size_t)n3 * r), Kd((size_t)kNumOps * r * r), sg((size_t)n3);
  BB_TRY(operators_fp64_device(n, r, Ud.data(), Kd.data(), sg.data()));
  for (size_t i = 0; i < Ud.size(); i++) U[i] = (real)Ud[i];
  for (size_t i = 0; i < Kd.size(); i++) K[i] = (real)Kd[i];
  if (sigma)
    for (int i = 0; i < n3; i++) sigma[i] = sg[i];
  return BBFMM_OK;
}

// Operator files in the reference's binary layout (bbfmm.c:216-254 reads,
// bbfmm.c:1355-1403 writes): U file = int32 r, then n^3 x r reals
// column-major; K file = 316 blocks of r x r reals, column-major.
extern "C" int bbfmm_amd_write_operators(const char *Kpath, const char *Upath, int n, int r, const real *U,
                                         const real *K)
{
  clear_error();
  if (!Kpath || !Upath || !U || !K || n < 1 || r < 1) return set_error(BBFMM_E_ARG, "bad operator-file arguments");
  const size_t usize = (size_t)r * n * n * n, ksize = (size_t)kNumOps * r * r;
  FILE *f = fopen(Upath, "wb");
  if (!f) return set_error(BBFMM_E_ARG, "cannot open %s for writing", Upath);
  const bool okU = fwrite(&r, sizeof(int), 1, f) == 1 && fwrite(U, sizeof(real), usize, f) == usize;
  const bool closeU = fclose(f) == 0;
  if (!okU || !closeU) return set_error(BBFMM_E_ARG, "short write to %s", Upath);
  f = fopen(Kpath, "wb");
  if (!f) return set_error(BBFMM_E_ARG, "cannot open %s for writing", Kpath);
  const bool okK = fwrite(K, sizeof(real), ksize, f) == ksize;
  const bool closeK = fclose(f) == 0;
  if (!okK || !closeK) return set_error(BBFMM_E_ARG, "short write to %s", Kpath);
  return BBFMM_OK;
}

extern "C" int bbfmm_amd_read_operators(const char *Kpath, const char *Upath, int n, int *r, real *U, real *K)
{
  clear_error();
  if (!Kpath || !Upath || !r || n < 1) return set_error(BBFMM_E_ARG, "bad operator-file arguments");
  FILE *f = fopen(Upath, "rb");
  if (!f) return set_error(BBFMM_E_ARG, "cannot open %s", Upath);
  int cutoff = 0;
  if (fread(&cutoff, sizeof(int), 1, f) != 1 || cutoff < 1 || cutoff > n * n * n) {
    fclose(f);
    return set_error(BBFMM_E_ARG, "%s: bad rank header", Upath);
  }
  *r = cutoff;
  if (!U || !K) {  // size query
    fclose(f);
    return BBFMM_OK;
  }
  const size_t usize = (size_t)cutoff * n * n * n, ksize = (size_t)kNumOps * cutoff * cutoff;
  const bool okU = fread(U, sizeof(real), usize, f) == usize;
  fclose(f);
  if (!okU) return set_error(BBFMM_E_ARG, "%s: short file", Upath);
  f = fopen(Kpath, "rb");
  if (!f) return set_error(BBFMM_E_ARG, "cannot open %s", Kpath);
  const bool okK = fread(K, sizeof(real), ksize, f) == ksize;
  fclose(f);
  if (!okK) return set_error(BBFMM_E_ARG, "%s: short file", Kpath);
  return BBFMM_OK;
}

// ---------------------------------------------------------------- tree builder
extern "C" void anotherBuildFMMHierachy(real L, int n, int dof, int l, int cutoff, anotherTree **atree, int Nf,
                                        int Ns, real3 *field, real3 *source)
{
  (void)n;
  (void)dof;
  (void)cutoff;
  (void)Ns;
  (void)source;
  clear_error();
  anotherTree *t = (anotherTree *)calloc(1, sizeof(anotherTree));
  const int rc = build_tree(t, L, l, Nf, field);
  if (rc != BBFMM_OK) set_error(rc, "a particle lies outside the root cell [-L/2, L/2]^3");
  *atree = t;
}

extern "C" void anotherFMMCleanup(anotherTree **atree)
{
  if (atree && *atree) {
    free_tree(*atree);
    *atree = nullptr;
  }
}

// ---------------------------------------------------------------- event timers
extern "C" void initEventTimer(eventTimerType *t)
{
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    set_error(BBFMM_E_HIP, "hipEventCreate failed");
  t->start = (void *)a;
  t->stop = (void *)b;
  t->time = 0.0f;
}

extern "C" void startEventTimer(eventTimerType *t) { (void)hipEventRecord((hipEvent_t)t->start, 0); }

extern "C" void stopEventTimer(eventTimerType *t)
{
  (void)hipEventRecord((hipEvent_t)t->stop, 0);
  (void)hipEventSynchronize((hipEvent_t)t->stop);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, (hipEvent_t)t->start, (hipEvent_t)t->stop);
  t->time += ms;
}

extern "C" void finalizeEventTimer(eventTimerType *t)
{
  (void)hipEventDestroy((hipEvent_t)t->start);
  (void)hipEventDestroy((hipEvent_t)t->stop);
}

extern "C" void printEventTimer(FILE *fp, char *s, eventTimerType *t)
{
  fprintf(fp, "# %s: %s = %14.7e\n", "printEventTimer", s, t->time / 1000);
}

extern "C" double readEventTimer(eventTimerType *t) { return (double)t->time; }
extern "C" double getEventTimer(eventTimerType *t) { return (double)(t->time / 1000); }
extern "C" void setEventTimer(eventTimerType *t, double val) { t->time = (float)val; }

// ---------------------------------------------------------------- drop-in passes
namespace {

struct DropInPlan {
  bbfmm_amd_plan *p = nullptr;
  ~DropInPlan()
  {
    if (p) bbfmm_amd_plan_destroy(p);
  }
};

int dropin_plan(DropInPlan &d, anotherTree *t, const real3 *field, real *K, real *U, real *Tkz, real *Kw,
                real *Cw, real homogen, int cutoff, int n)
{
  d.p = new bbfmm_amd_plan();
  const int rc = plan_init(d.p, t, field, n, cutoff, K, U, Tkz, Kw, Cw, homogen, 1, nullptr);
  if (rc != BBFMM_OK) return rc;
  d.p->scheme = g_default_scheme;
  d.p->form = BBFMM_M2L_PLAIN;
  d.p->l2p_ref = 1;  // the drop-in reproduces the reference's field output, quirks included
  return BBFMM_OK;
}

} // namespace

extern "C" void anotherFMMCompute(anotherTree **atree, real3 *field, real3 *source, real *q, real *K, real *U,
                                  real *Tkz, int *Ktable, real *Kweights, real *Cweights, real homogen, int cutoff,
                                  int n, int dof, real *phi)
{
  (void)source;
  (void)Ktable;
  clear_error();
  if (dof != 1) {
    set_error(BBFMM_E_ARG, "dof=%d not supported (dof = 1)", dof);
    return;
  }
  DropInPlan d;
  if (dropin_plan(d, *atree, field, K, U, Tkz, Kweights, Cweights, homogen, cutoff, n) != BBFMM_OK) return;
  if (bbfmm_amd_solve_host(d.p, q, phi) != BBFMM_OK) return;
  if (g_verbose) {
    fprintf(stderr, "# anotherFMMCompute: N = %d, maxlev = %d, scheme = %d\n", d.p->N, d.p->maxlev, d.p->scheme);
  }
}

extern "C" void anotherFMMInteraction(anotherTree **atree, real *E, int *Ktable, real *U, real *Kweights, int n,
                                      int dof, int cutoff, real homogen)
{
  (void)Ktable;
  clear_error();
  anotherTree *t = *atree;
  if (dof != 1) {
    set_error(BBFMM_E_ARG, "dof=%d not supported (dof = 1)", dof);
    return;
  }
  if (!t->proxysval) {
    set_error(BBFMM_E_ARG, "proxysval is not set (run the upward pass first)");
    return;
  }
  // weights are only needed for the scaling here; Tkz/Cw are unused
  std::vector<real> Tkz((size_t)n * n), Kw((size_t)n * n * n), Cw((size_t)2 * n * n);
  int Kt[343];
  weights(n, Tkz.data(), Kt, Kw.data(), Cw.data());
  std::vector<real3> nopos((size_t)(t->Nf > 0 ? t->Nf : 1));
  for (auto &x : nopos) x.x = x.y = x.z = 0;
  DropInPlan d;
  if (dropin_plan(d, t, nopos.data(), E, U, Tkz.data(), Kweights, Cw.data(), homogen, cutoff, n) != BBFMM_OK)
    return;
  bbfmm_amd_plan *p = d.p;
  const size_t nm = (size_t)p->ncell * p->r;
  if (hipMemcpy(p->d_M, t->proxysval, sizeof(real) * nm, hipMemcpyHostToDevice) != hipSuccess) {
    set_error(BBFMM_E_HIP, "copy of M failed");
    return;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, p->stream);
  if (run_m2l(p, p->d_M, p->d_L) != BBFMM_OK) return;
  (void)hipEventRecord(e1, p->stream);
  const PassArgs a = pass_args(p);
  if (launch_postm2l(a, p->stream, 1) != BBFMM_OK) return;
  float ms = 0;
  (void)hipEventSynchronize(e1);
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  kernel_exec_time = ms * 1e-3;
  t->fieldval = (real *)calloc((size_t)p->ncell * p->n3, sizeof(real));
  if (hipMemcpy(t->fieldval, p->d_F, sizeof(real) * (size_t)p->ncell * p->n3, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error(BBFMM_E_HIP, "copy of F failed");
    return;
  }
  free(t->proxysval);
  t->proxysval = nullptr;
  if (g_verbose) {
    const double flop = (double)p->r * (2.0 * p->r - 1) * p->m2l_interactions;
    fprintf(stderr, "# anotherFMMInteraction: calc_performance: kernel = %f [Gflop/s]\n",
            flop / kernel_exec_time * 1e-9);
  }
}

extern "C" void anotherDownwardPass(anotherTree **atree, real3 *field, real3 *source, real *Cweights, real *Tkz,
                                    real *q, real *U, int cutoff, int n, int dof, real homogen, real *phi)
{
  (void)source;
  clear_error();
  anotherTree *t = *atree;
  if (dof != 1) {
    set_error(BBFMM_E_ARG, "dof=%d not supported (dof = 1)", dof);
    return;
  }
  if (!t->fieldval) {
    set_error(BBFMM_E_ARG, "fieldval is not set (run anotherFMMInteraction first)");
    return;
  }
  std::vector<real> Tk((size_t)n * n), Kw((size_t)n * n * n), Cw((size_t)2 * n * n);
  int Kt[343];
  weights(n, Tk.data(), Kt, Kw.data(), Cw.data());
  const int r = cutoff;
  std::vector<real> K((size_t)kNumOps * r * r, (real)0);  // M2L is not run here
  DropInPlan d;
  if (dropin_plan(d, t, field, K.data(), U, Tkz, Kw.data(), Cweights, homogen, cutoff, n) != BBFMM_OK) return;
  bbfmm_amd_plan *p = d.p;
  hipStream_t s = p->stream;
  if (hipMemcpy(p->d_F, t->fieldval, sizeof(real) * (size_t)p->ncell * p->n3, hipMemcpyHostToDevice) != hipSuccess) {
    set_error(BBFMM_E_HIP, "copy of F failed");
    return;
  }
  if (dalloc(&p->d_qin, (size_t)p->N) != BBFMM_OK || dalloc(&p->d_phiout, (size_t)p->N) != BBFMM_OK) return;
  if (hipMemcpy(p->d_qin, q, sizeof(real) * p->N, hipMemcpyHostToDevice) != hipSuccess) {
    set_error(BBFMM_E_HIP, "copy of q failed");
    return;
  }
  if (launch_gather(p->d_perm, p->d_qin, p->d_q, p->N, 1, s) != BBFMM_OK) return;
  const PassArgs a = pass_args(p);
  if (launch_l2l(a, p->h_ploff.data(), s, 1) != BBFMM_OK) return;
  if (launch_l2p(a, s, 1) != BBFMM_OK) return;
  P2PArgs pp;
  pp.N = p->N;
  pp.nleaf = p->nleaf;
  pp.nrhs = 1;
  pp.pos = p->d_pos;
  pp.leafsta = p->d_leafsta;
  pp.leafcnt = p->d_leafcnt;
  pp.nbr = p->d_nbr;
  pp.nnbr = p->d_nnbr;
  pp.q = p->d_q;
  pp.phi = p->d_phi;
  if (launch_p2p(pp, s) != BBFMM_OK) return;
  if (launch_scatter(p->d_iperm, p->d_phi, p->d_phiout, p->N, 1, s) != BBFMM_OK) return;
  if (hipMemcpy(phi, p->d_phiout, sizeof(real) * p->N, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error(BBFMM_E_HIP, "copy of phi failed");
    return;
  }
  free(t->fieldval);
  t->fieldval = nullptr;
}

// ---------------------------------------------------------------- m2l_drv
extern "C" int m2l_drv(int r, real *L, real *K, real *M, int minlev, int maxlev, real L0, real3 *center,
                       int *ineigh, int pitch_neighbors, int *neighbors, int *iinter, int pitch_interaction,
                       int *interaction, int scheme)
{
  (void)ineigh;
  (void)pitch_neighbors;
  (void)neighbors;
  clear_error();
  if (r < 1 || r > kMaxRank) {
    set_error(BBFMM_E_ARG, "m2l_drv: r=%d outside 1..%d", r, kMaxRank);
    return M2L_EXIT_FAIL;
  }
  if (scheme < M2L_SCHEME_BA || scheme > M2L_SCHEME_IJ) {
    set_error(BBFMM_E_ARG, "m2l_drv: unknown scheme %d", scheme);
    return M2L_EXIT_FAIL;
  }
  if (minlev < kMinLev || maxlev < minlev || maxlev > 9 || pitch_interaction < 189) {
    set_error(BBFMM_E_ARG, "m2l_drv: invalid level range %d..%d or pitch %d", minlev, maxlev, pitch_interaction);
    return M2L_EXIT_FAIL;
  }
  if (check_device() != BBFMM_OK) return M2L_EXIT_FAIL;
  bbfmm_amd_plan p;
  p.r = r;
  p.nrhs = 1;
  p.maxlev = maxlev;
  p.scheme = scheme;
  p.form = BBFMM_M2L_PLAIN;
  p.ncell = ((1 << (3 * (maxlev + 1))) - 1) / 7;
  p.h_levsta.resize(maxlev + 2);
  for (int l = 0; l <= maxlev + 1; l++) p.h_levsta[l] = ((1 << (3 * l)) - 1) / 7;
  const int ncell = p.ncell;
  int rc = BBFMM_OK;
  std::vector<int> inter((size_t)ncell * kPitchInteraction, 0), kidx((size_t)ncell * kPitchInteraction, 0);
  std::vector<int> iint(iinter, iinter + ncell);
  int Kt[343];
  ktable(Kt);
  for (int A = 0; A < ncell; A++) {
    if (iint[A] > kPitchInteraction) {
      set_error(BBFMM_E_ARG, "m2l_drv: cell %d has %d interactions", A, iint[A]);
      return M2L_EXIT_FAIL;
    }
    for (int i = 0; i < iint[A]; i++) inter[(size_t)kPitchInteraction * A + i] = interaction[(size_t)pitch_interaction * A + i];
  }
  {
    std::vector<int> ls(maxlev + 1), le(maxlev + 1);
    std::vector<real> cl(maxlev + 1);
    level_tables(maxlev, L0, ls.data(), le.data(), cl.data());
    anotherTree tv{};
    cell cv{};
    cv.center = center;
    cv.iinter = iint.data();
    cv.interaction = inter.data();
    cv.pitch_interaction = kPitchInteraction;
    tv.c = &cv;
    tv.maxlev = maxlev;
    tv.levsta = ls.data();
    tv.levend = le.data();
    tv.celeng = cl.data();
    kindex(&tv, Kt, kidx.data());
  }
  // K4 layout for the compiled-in ranks, the caller's column-major K otherwise
  const bool k4 = r == 32 || r == 256;
  std::vector<real4> K4(k4 ? (size_t)kNumOps * r * r / 4 : 0);
  for (int k = 0; k < (k4 ? kNumOps : 0); k++)
    for (int jb = 0; jb < r / 4; jb++)
      for (int i = 0; i < r; i++) {
        const real *Kk = K + (size_t)k * r * r;
        real4 v;
        v.x = Kk[(size_t)(4 * jb + 0) * r + i];
        v.y = Kk[(size_t)(4 * jb + 1) * r + i];
        v.z = Kk[(size_t)(4 * jb + 2) * r + i];
        v.w = Kk[(size_t)(4 * jb + 3) * r + i];
        K4[((size_t)k * (r / 4) + jb) * r + i] = v;
      }
  p.h_K.assign(K, K + (size_t)kNumOps * r * r);
  real *dL = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  do {
    if ((rc = upload(&p.d_iinter, iint.data(), iint.size(), nullptr)) != BBFMM_OK) break;
    if ((rc = upload(&p.d_inter, inter.data(), inter.size(), nullptr)) != BBFMM_OK) break;
    if ((rc = upload(&p.d_kidx, kidx.data(), kidx.size(), nullptr)) != BBFMM_OK) break;
    if ((rc = upload(&p.d_Ktable, Kt, (size_t)343, nullptr)) != BBFMM_OK) break;
    if (k4 && (rc = upload(&p.d_K4, K4.data(), K4.size(), nullptr)) != BBFMM_OK) break;
    if ((!k4 || r <= kClusterMaxRank) && (rc = upload(&p.d_Kc, K, (size_t)kNumOps * r * r, nullptr)) != BBFMM_OK)
      break;
    {
      int op[kNumOps];
      operator_offsets(op);
      if ((rc = upload(&p.d_opoff, op, (size_t)kNumOps, nullptr)) != BBFMM_OK) break;
    }
    if ((rc = upload(&p.d_M, M, (size_t)ncell * r, nullptr)) != BBFMM_OK) break;
    if ((rc = dalloc(&dL, (size_t)ncell * r)) != BBFMM_OK) break;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      rc = set_error(BBFMM_E_HIP, "hipEventCreate failed");
      break;
    }
    (void)hipEventRecord(e0, nullptr);
    M2LArgs m;
    m.r = r;
    m.ncell = ncell;
    m.nrhs = 1;
    m.minlev = minlev;
    m.maxlev = maxlev;
    m.h_levsta = p.h_levsta.data();
    m.iinter = p.d_iinter;
    m.inter = p.d_inter;
    m.kidx = p.d_kidx;
    m.K4 = p.d_K4;
    m.Kc = p.d_Kc;
    m.opoff = p.d_opoff;
    m.Ktable = p.d_Ktable;
    m.M = p.d_M;
    m.L = dL;
    if (scheme == M2L_SCHEME_BA) rc = launch_m2l_basic(m, nullptr);
    else if (scheme == M2L_SCHEME_SI) rc = launch_m2l_sibling(m, nullptr);
    else if (scheme == M2L_SCHEME_CL && r <= kClusterMaxRank) rc = launch_m2l_cluster(m, nullptr);
    else {
      rc = grid_m2l_create(&p.grid, r, maxlev, 1, BBFMM_M2L_PLAIN, p.h_K.data(), nullptr, minlev, 0);
      if (rc == BBFMM_OK) rc = grid_m2l_run(p.grid, p.d_M, dL, ncell, nullptr, nullptr);
    }
    if (rc != BBFMM_OK) break;
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess) {
      rc = set_error(BBFMM_E_HIP, "M2L kernel failed");
      break;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    kernel_exec_time = ms * 1e-3;
    const size_t off = (size_t)p.h_levsta[minlev] * r;
    const size_t cnt = (size_t)(p.h_levsta[maxlev + 1] - p.h_levsta[minlev]) * r;
    if (hipMemcpy(L + off, dL + off, sizeof(real) * cnt, hipMemcpyDeviceToHost) != hipSuccess)
      rc = set_error(BBFMM_E_HIP, "copy of L failed");
  } while (0);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (dL) (void)hipFree(dL);
  plan_free_device(&p);
  return rc == BBFMM_OK ? M2L_EXIT_SUCCESS : M2L_EXIT_FAIL;
}
