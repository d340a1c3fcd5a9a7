/*
 * bbfmm_amd.h -- C ABI of the MI355X (gfx950) bbFMM downward-pass library.
 *
 * Two shared libraries are built from the same sources, as the reference
 * builds two executables from one tree (real.h:6-62):
 *   libbbfmm_amd_s.so   real = float   (define SINGLE before including)
 *   libbbfmm_amd_d.so   real = double
 *
 * Part 1 is the drop-in surface: the entry points the reference's CUDA
 * translation unit (main.cu -> anotherFMMCompute.cu) exports, with the same
 * names, argument meaning and data layouts.  Part 2 is a persistent,
 * device-resident plan used by the batched forms, the benchmark and the
 * parity tests.  Every solve pass (P2M, M2M, M2L, post-M2L, L2L, L2P, P2P)
 * runs on the GPU; there is no CPU fallback.  Setup (tree, lists, Chebyshev
 * weights, operator precompute) runs on the host, as in the reference.
 */
#ifndef BBFMM_AMD_H
#define BBFMM_AMD_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef SINGLE
typedef float real;
#else
typedef double real;
#endif

/* vec234.h:18-24: real3 is an unaligned triple, real4 is 16-byte aligned */
typedef struct { real x, y, z; } real3;

/* cell.h:10-22 (non-PBC build) */
typedef struct {
  int *neighbors;
  int *interaction;
  real3 *center;
  int *fieldsta, *fieldend, *sourcesta, *sourceend;
  int *ineigh, *iinter;
  int pitch_neighbors, pitch_interaction;
} cell;

/* anotherTree.h:7-20 */
typedef struct {
  int ncell;
  int minlev;
  int maxlev;
  int *levsta;
  int *levend;
  real *celeng;
  cell *c;
  real *fieldval;
  real *sourceval;
  real *proxysval;
  int ncell0;
  int Nf, Ns, *fieldlist, *sourcelist;
} anotherTree;

/* m2l_drv.h:4-10 */
#define M2L_EXIT_SUCCESS 0
#define M2L_EXIT_FAIL 1
#define M2L_SCHEME_BA 0
#define M2L_SCHEME_SI 1
#define M2L_SCHEME_CL 2
#define M2L_SCHEME_IJ 3

/* ====================================================================== */
/* Part 1: drop-in entry points                                           */
/* ====================================================================== */

/* Replaces anotherFMMCompute (another.h:42-45; anotherFMMCompute.cu:8-45).
   Runs the whole solve (upward, M2L, post-M2L, L2L/L2P, P2P) on the GPU for
   the tree built by anotherBuildFMMHierachy; writes phi = q .* potential. */
void anotherFMMCompute(anotherTree **atree, real3 *field, real3 *source, real *q,
                       real *K, real *U, real *Tkz, int *Ktable,
                       real *Kweights, real *Cweights, real homogen,
                       int cutoff, int n, int dof, real *phi);

/* Replaces anotherFMMInteraction (another.h:47-49; anotherFMMInteraction.cu:29-126).
   Reads (*atree)->proxysval (M), runs M2L through m2l_drv and post-M2L on the
   GPU, allocates and fills (*atree)->fieldval, frees proxysval. */
void anotherFMMInteraction(anotherTree **atree, real *E, int *Ktable,
                           real *U, real *Kweights, int n, int dof,
                           int cutoff, real homogen);

/* Replaces the GPU anotherDownwardPass (another.h:52-54; anotherDownwardPass.cu:41-411).
   L2L and L2P from (*atree)->fieldval (freed on return), then P2P; writes phi. */
void anotherDownwardPass(anotherTree **atree, real3 *field, real3 *source,
                         real *Cweights, real *Tkz, real *q, real *U,
                         int cutoff, int n, int dof, real homogen, real *phi);

/* Replaces m2l_drv (m2l_drv.cu:14-93).  L[cell*r] for cells in levels
   minlev..maxlev from M[cell*r] and the 316 column-major r x r operators K.
   scheme: M2L_SCHEME_BA (basic), _SI (sibling), _CL (cluster), _IJ (dense
   grid).  Returns M2L_EXIT_SUCCESS, or M2L_EXIT_FAIL for an unsupported r
   or scheme, or a HIP failure (the reference aborts or returns FAIL). */
int m2l_drv(int r, real *L, real *K, real *M, int minlev, int maxlev,
            real L0, real3 *center,
            int *ineigh, int pitch_neighbors, int *neighbors,
            int *iinter, int pitch_interaction, int *interaction,
            int scheme);

/* m2l_host.h:24: seconds spent in M2L kernels by the last m2l_drv call */
extern double kernel_exec_time;

/* Tree builder (another.h:36-39; anotherBuildFMMHierachy.c:24-155) */
void anotherBuildFMMHierachy(real L, int n, int dof, int l, int cutoff,
                             anotherTree **atree, int Nf, int Ns,
                             real3 *field, real3 *source);
void anotherFMMCleanup(anotherTree **atree);

/* Event timers (eventTimer.h:16-32), on the library's HIP stream */
typedef struct {
  float time;    /* accumulated milliseconds */
  void *start;   /* hipEvent_t */
  void *stop;    /* hipEvent_t */
} eventTimerType;
void initEventTimer(eventTimerType *timer);
void startEventTimer(eventTimerType *timer);
void stopEventTimer(eventTimerType *timer);
void finalizeEventTimer(eventTimerType *timer);
void printEventTimer(FILE *fp, char *s, eventTimerType *timer);
double readEventTimer(eventTimerType *timer);   /* milliseconds */
double getEventTimer(eventTimerType *timer);    /* seconds */
void setEventTimer(eventTimerType *timer, double val);

/* Error reporting for the void entry points: 0 if the last call on this
   thread succeeded, else a BBFMM_E_* code; the string describes it. */
int bbfmm_amd_last_error(void);
const char *bbfmm_amd_last_error_string(void);

/* Default M2L scheme used by anotherFMMInteraction / anotherFMMCompute (the
   reference fixes it at compile time with -DSCHEME; default here: IJ). */
int bbfmm_amd_set_default_scheme(int scheme);

#define BBFMM_OK 0
#define BBFMM_E_ARG 1        /* invalid argument or unsupported n / r / scheme */
#define BBFMM_E_HIP 2        /* HIP runtime failure (message has the HIP error) */
#define BBFMM_E_NOMEM 3      /* allocation failure */
#define BBFMM_E_NODEVICE 4   /* no GPU visible */
#define BBFMM_E_RANGE 5      /* particle outside the root cell */

/* ====================================================================== */
/* Part 2: persistent plan (device-resident tree, operators and buffers)   */
/* ====================================================================== */
typedef struct bbfmm_amd_plan bbfmm_amd_plan;

/* M2L arithmetic forms */
#define BBFMM_M2L_PLAIN 0     /* real-precision FMAs (drop-in parity) */
#define BBFMM_M2L_BF16X2 1    /* split BF16 hi/lo operands, 3 MFMA products, FP32 accumulate */
#define BBFMM_M2L_FP32MFMA 2  /* FP32 MFMA */
#define BBFMM_M2L_FP16X3 3    /* split FP16 operands in three terms (K and each level's M
                                 scaled by powers of two), 6 MFMA products, FP32 accumulate;
                                 dense grids (complete trees), r = 32 */

/* Plan options (bbfmm_amd_plan_set) */
#define BBFMM_OPT_SCHEME 1    /* M2L_SCHEME_* */
#define BBFMM_OPT_FORM 2      /* BBFMM_M2L_* */
#define BBFMM_OPT_TIMING 3    /* k > 0: record per-stage HIP events on every solve into a
                                 ring of k event sets; 0: off (resets the averages) */
#define BBFMM_OPT_OUTPUT 4    /* 0: q .* potential (the reference's output, default);
                                 1: the potential */
#define BBFMM_OPT_GRAPH 5     /* 1: capture the whole solve (gather .. scatter) into one
                                 hipGraph per (q, phi) pointer pair on first use and
                                 replay it; stage times then report the total only */
#define BBFMM_OPT_M2L_VARIANT 6  /* tile shape of the M2L scheme: basic 0/1/2 = 256/128/512
                                    threads per block; IJ MFMA forms 0 automatic,
                                    1/4/5 source-stationary (split-BF16) with 8/4/2 field
                                    classes per wave, 6/7 the same with 2/4 classes and
                                    two 32-column tiles per wave, 8 as 7 with both class
                                    groups in one block sharing the source tiles through
                                    LDS (the 8-class step list), 2/3 class-stationary
                                    with one/two 32-column tiles per wave */
#define BBFMM_OPT_NEAR_STREAM 8  /* 16-vector MFMA solves: 1 runs P2P on a side stream
                                    beside the downward pass (forked after the M2L, joined
                                    before the scatter), 0 (default) after it; the results
                                    are the same bits.  At N = 1M the concurrent kernels
                                    compete for the CUs and the solve is slower */
#define BBFMM_OPT_P2P_FORM 9  /* 16-vector MFMA solves (fp32 build): 0 (default) the fp32
                                 P2P (1/r and the sum over sources in fp32); 1 1/r in fp32,
                                 the sum over sources as three split-BF16 products (hi*hi +
                                 hi*lo + lo*hi) with fp32 accumulation, relative error of a
                                 term below 3 * 2^-16.  No counterpart in the reference
                                 (its P2P is plain fp32) */
#define BBFMM_OPT_LEVEL_SWITCH 7 /* CL/IJ schemes: first level on the dense grids (on a
                                    batch of sparse instances: on the list M2L), the
                                    levels below on the basic scheme (default 2) */
#define BBFMM_OPT_L2P_REFERENCE 10 /* orders n % 4 == 3 (n = 3, 7): 1 evaluates L2P as the
                                      reference does, weighting the last z node with the
                                      second-to-last node's interpolation weight
                                      (anotherDownwardPassX.c:304); 0 (default) evaluates the
                                      Chebyshev interpolant.  The drop-in entry points
                                      (anotherFMMCompute, anotherDownwardPass) always take 1.
                                      Other orders: no effect */
#define BBFMM_OPT_M2L_VARIANT_UPPER 12 /* IJ split-BF16 with a source-stationary variant
                                          (1, 4 to 7): the class-group size and tiles
                                          per wave of the levels above the deepest, run
                                          as a launch of their own (1, 4, 5 = 8, 4, 2
                                          classes per wave, one tile; 6, 7 = 2, 4
                                          classes, two tiles); 0 (default): the deepest
                                          level's, or the automatic configuration */
#define BBFMM_OPT_M2L_UPPER_SPLIT 13 /* IJ split-BF16 source-stationary M2L: with
                                        BBFMM_OPT_M2L_VARIANT_UPPER set, the launch of
                                        the upper levels, else the single launch, takes
                                        each class group's steps in this many parts (1
                                        to 16, default 1), one block each, their partial
                                        sums added to L after it (a shorter critical
                                        path when the columns are few); variant 0
                                        chooses the parts itself */
#define BBFMM_OPT_ADDRESSING 11 /* 0 (default): kernels with two forms read through
                                   scalar-offset buffer loads while their arrays are
                                   below 2 GB and through pointers beyond; 1: through
                                   pointers always (the split-BF16 source-stationary
                                   and list M2L, the fp32 16-vector P2P).  Same results */

/* Create a plan for N particles at host positions pos (N real3, inside the
   root cell [-L/2,L/2]^3), Chebyshev order n, level modifier lmod (the -d
   option; maxlev follows bbfmm.c:151-160), and nrhs charge vectors per solve.
   K (316 r x r column-major) and U (n^3 x r column-major) are host arrays
   in the reference's file layout.  The tree and lists are built on the host
   and uploaded with the operators.  stream: hipStream_t (NULL = default). */
int bbfmm_amd_plan_create(bbfmm_amd_plan **plan, int N, const real3 *pos, real L,
                          int n, int lmod, int r, const real *K, const real *U,
                          int nrhs, void *stream);
int bbfmm_amd_plan_destroy(bbfmm_amd_plan *plan);

/* A plan over ninst independent problems solved together (no counterpart in
   the reference, which solves one problem per call): instance i has N[i]
   points, concatenated in pos (sum N real3, each inside [-L/2,L/2]^3), and
   its own octree of uniform depth depth[i] (depth NULL or depth[i] <= 0: the
   smallest depth >= 2 whose leaves hold at most leafmax points, at most 10),
   storing only the cells that hold points.  Each instance's field equals the
   reference's complete-octree field at the same depth (empty cells add
   exact zeros).  Every stage is one launch over all instances (M2M and L2L
   one per level); q and phi are nrhs x sum(N), instances concatenated.
   M2L: the basic scheme.  depth_out (optional, ninst ints) receives the
   depths used.  Destroy with bbfmm_amd_plan_destroy. */
int bbfmm_amd_batch_create(bbfmm_amd_plan **plan, int ninst, const int *N, const real3 *pos,
                           const int *depth, int leafmax, real L, int n, int r, const real *K,
                           const real *U, int nrhs, void *stream, int *depth_out);
int bbfmm_amd_plan_set(bbfmm_amd_plan *plan, int option, int value);

/* M2L autotuner (no counterpart in the reference, which fixes the scheme at
   compile time): runs every M2L configuration valid for the plan's tree and
   arithmetic form (scheme basic/sibling/cluster/ij, tile variant, level
   switch) reps times on the plan's current multipoles, keeps the fastest
   and writes up to maxrows rows of 5 doubles (scheme, variant, upper-level
   variant, level switch, ms per M2L) to table.  The arithmetic form is never
   changed. */
int bbfmm_amd_plan_tune(bbfmm_amd_plan *plan, int reps, double *table, int maxrows, int *nrows);

/* Tree facts: out[0]=maxlev, out[1]=ncell, out[2]=N, out[3]=r, out[4]=n,
   out[5]=nrhs.  dout[0]=M2L interactions (levels 2..maxlev, per vector),
   dout[1]=P2P pairs (self pairs included, per vector). */
int bbfmm_amd_plan_info(const bbfmm_amd_plan *plan, int *out, double *dout);

/* One solve on device memory: q_dev is nrhs x N (row per vector, original
   particle order); phi_dev receives q .* potential in the same layout.
   Asynchronous on the plan's stream. */
int bbfmm_amd_solve(bbfmm_amd_plan *plan, const real *q_dev, real *phi_dev);

/* Same from host arrays (PCIe copies included; synchronous). */
int bbfmm_amd_solve_host(bbfmm_amd_plan *plan, const real *q, real *phi);

/* Move the particles to new host positions pos (N real3, original order)
   that stay in their leaves: the tree, lists and operators are kept.
   Returns BBFMM_E_RANGE, leaving the plan unchanged, if a particle would
   change leaf (build a new plan then).  No counterpart in the reference
   (which rebuilds the tree for every call). */
int bbfmm_amd_plan_update_positions(bbfmm_amd_plan *plan, const real3 *pos);

/* Adjoint of the solve (no counterpart in the reference): for a plan made
   with nrhs = 2m vectors, charges q_dev (m x N) and g_dev = dLoss/dphi
   (m x N, phi the potentials), writes dq_dev = dLoss/dq (m x N, the
   potentials of the charges g, the FMM operator being symmetric) and
   dx_dev = dLoss/dpositions (N x 3, original order).  The expansions are
   recomputed by one solve of the 2m vectors [q; g]; the far part of dx
   differentiates the L2P interpolant (Chebyshev derivative) and the near
   part the 1/r kernel over the P2P neighbourhoods (in blocks of 16 vectors
   when m > 16). */
int bbfmm_amd_gradient(bbfmm_amd_plan *plan, const real *q_dev, const real *g_dev, real *dq_dev,
                       real *dx_dev);

/* M2L alone on the plan's tree: M_dev and L_dev are nrhs x ncell x r
   device arrays in global cell order; L is written for levels 2..maxlev. */
int bbfmm_amd_m2l(bbfmm_amd_plan *plan, const real *M_dev, real *L_dev);

/* Copy an intermediate of the last solve to host memory, in global cell
   order for vector 0: which = 0 M (ncell*r), 1 L (ncell*r),
   2 F after post-M2L (ncell*n^3), 3 far-field q.*phi (N, original order). */
int bbfmm_amd_fetch(bbfmm_amd_plan *plan, int which, real *dst);

/* Per-stage times (ms, HIP events on the plan's stream) averaged over the
   last min(solves since timing was set, k) timed solves: 0 gather, 1 P2M,
   2 M2M, 3 M2L preparation (dense-grid scatter), 4 M2L kernel, 5 post-M2L,
   6 L2L, 7 L2P, 8 P2P, 9 scatter, 10 total, 11 P2P on the side stream.
   With P2P beside the far field (BBFMM_OPT_NEAR_STREAM) entry 8 is the wait
   for it after the far field and entry 11 its own duration (0 otherwise).
   Synchronises the stream.  Returns the number of entries written. */
int bbfmm_amd_stage_times(bbfmm_amd_plan *plan, double *out, int nout);

/* FP64 direct sum on the GPU (DirectCalc3D, bbfmm.c:471-504 with start = 0):
   out[t] = q_i * sum_{j: x_j != x_i} q_j / |x_i - x_j| for i = targets[t]
   (all particles when targets is NULL, nt = N).  pos is N x 3 doubles. */
int bbfmm_amd_direct(int N, const double *pos, const double *q, int nt, const int *targets, double *out);

/* No counterpart in the reference: bbfmm_amd_direct for callers that cannot
   allocate device memory (code under stream capture, serving loops, harnesses
   that forbid hipMalloc).  pos (N x 3), q (N), targets (nt indices, or NULL
   for 0..nt-1) and out (nt) are DEVICE arrays; the call is asynchronous on
   `stream` (hipStream_t, NULL = default) and allocates nothing: the partial
   sums use a file-scope device array of 64 x BBFMM_DIRECT_NOALLOC_MAX_TARGETS
   doubles (8 MiB) that each GPU holds until the process ends.  nt above the
   maximum is rejected (BBFMM_E_ARG).  Calls that overlap on one GPU (other
   streams or threads) race on that array.  Same kernels and results as
   bbfmm_amd_direct. */
#define BBFMM_DIRECT_NOALLOC_MAX_TARGETS 16384
int bbfmm_amd_direct_noalloc(int N, const double *pos, const double *q, int nt, const int *targets,
                             double *out, void *stream);
/* FP64 direct-sum potentials and their gradients for nvec charge vectors
   (q: nvec x N): for target i = targets[t], pot[v*nt + t] =
   sum_{j: x_j != x_i} q_vj / r_ij and grad[(v*nt + t)*3 + c] = the c-th
   component of its gradient with respect to x_i.  The analytic reference of
   the adjoint; no counterpart in the reference. */
int bbfmm_amd_direct_field(int N, const double *pos, int nvec, const double *q, int nt, const int *targets,
                           double *pot, double *grad);

/* Host-side setup helpers (reference layouts) */
int bbfmm_amd_maxlev(int N, int lmod);                                   /* bbfmm.c:151-160 */
void bbfmm_amd_weights(int n, real *Tkz, int *Ktable, real *Kweights,
                       real *Cweights);                                  /* bbfmm.c:922-1028 */
/* Operator precompute in FP64 (bbfmm.c:1209-1421): U (n^3 x r) and K
   (316 r x r), column-major, cast to real.  sigma (n^3 values) optional. */
int bbfmm_amd_operators(int n, int r, real *U, real *K, double *sigma);
/* The same precompute with the kernel sampling, K0 K0^T and the 316 products
   U^T K0_k U in FP64 on the GPU (the eigenproblem on the host). */
int bbfmm_amd_operators_device(int n, int r, real *U, real *K, double *sigma);
/* Operator files in the reference's binary layout (FMMReadMatrices,
   bbfmm.c:216-254; written as ComputeKernelSVD writes them, bbfmm.c:1355-1403):
   U file = int32 r followed by n^3 x r reals, K file = 316 r x r reals, both
   column-major.  read: with U or K NULL only *r is returned (size query). */
int bbfmm_amd_write_operators(const char *Kpath, const char *Upath, int n, int r, const real *U, const real *K);
int bbfmm_amd_read_operators(const char *Kpath, const char *Upath, int n, int *r, real *U, real *K);

#ifdef __cplusplus
}
#endif

#endif /* BBFMM_AMD_H */
