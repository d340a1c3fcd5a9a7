/*
 * bbfmm_oracle.h -- CPU restatement of the reference bbFMM path (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This oracle is the checker for the GPU library in bbfmmongpu_amd/.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it.  Nothing in the product links or calls it.
 *
 * Precision: `real` is float when ORC_SINGLE is defined (the reference's
 * -DSINGLE build) and double otherwise, exactly like real.h:6-62.  Every
 * function carries the reference file:line it restates.
 */
#ifndef BBFMM_ORACLE_H
#define BBFMM_ORACLE_H

#ifdef ORC_SINGLE
typedef float real;
#else
typedef double real;
#endif

/* 12-byte (fp32) / 24-byte (fp64) unaligned triple: vec234.h:18-24 */
typedef struct { real x, y, z; } orc_real3;

#define ORC_PITCH_NEIGHBORS 64     /* FLOORUP(27,64): anotherBuildFMMHierachy.c:90-97 */
#define ORC_PITCH_INTERACTION 192  /* FLOORUP(189,64) */
#define ORC_PITCH_SOURCECLUSTERS 32 /* m2l_drv.h:12-14 */
#define ORC_NULL (-1)

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tree, lists and index tables (int work: bit-exact) ---- */
int orc_maxlev(int N, int lmod);
int orc_ncell(int maxlev);
void orc_levels(int maxlev, real L, int *levsta, int *levend, real *celeng);
void orc_centers(int maxlev, real L, orc_real3 *center);
int orc_distribute(int maxlev, real L, const orc_real3 *center, int N,
                   const orc_real3 *pos, int *list, int *sta, int *end);
void orc_lists(int maxlev, real L, const orc_real3 *center,
               int *ineigh, int *neighbors, int *iinter, int *interaction);
void orc_Ktable(int *Ktable);
void orc_Kindex(int maxlev, real L0, const orc_real3 *center, const int *iinter,
                const int *interaction, const int *Ktable, int *Kindex);
void orc_sourceclusters(int minlev, int maxlev, real L0, const orc_real3 *center,
                        const int *ineigh, const int *neighbors, int *sourceclusters);
void orc_sibling_tables(int *Kindex0, int *Kindex1);

/* ---- setup: Chebyshev weights and operator precompute ---- */
void orc_weights(int n, real *Tkz, int *Ktable, real *Kweights, real *Cweights);
void orc_kernel_matrix(int n, double *K0);
int orc_sym_eig(int m, double *A, double *w, double *V);
void orc_transfer_operators(int n, int r, const double *K0, const double *U, double *E);

/* ---- passes ---- */
void orc_upward(int maxlev, real L, int n, int r, const orc_real3 *center,
                const int *list, const int *sta, const int *end,
                const orc_real3 *pos, const real *q, const real *U,
                const real *Tkz, const real *Kweights, const real *Cweights,
                real *M);
void orc_m2l_basic(int r, int minlev, int maxlev, const real *K, const real *M,
                   const int *iinter, const int *interaction, const int *Kindex,
                   real *Lv);
void orc_m2l_sibling(int r, int minlev, int maxlev, const real *K, const real *M,
                     const int *sourceclusters, const int *Ktable, real *Lv);
void orc_postm2l(int maxlev, real L, int n, int r, real homogen, const real *U,
                 const real *Kweights, const real *Lv, real *F);
/* L2P at n % 4 == 3: 0 restates the reference (anotherDownwardPassX.c:304),
   1 evaluates the interpolant exactly */
void orc_set_l2p_exact(int on);
void orc_downward(int maxlev, real L, int n, const orc_real3 *center,
                  const int *list, const int *sta, const int *end,
                  const orc_real3 *pos, const real *q, const real *Tkz,
                  const real *Cweights, real *F, real *phi);
void orc_nearfield(int maxlev, int N, const int *list, const int *sta, const int *end,
                   const int *ineigh, const int *neighbors, const orc_real3 *pos,
                   const real *q, real *phi);
double orc_p2p_pairs(int maxlev, const int *sta, const int *end,
                     const int *ineigh, const int *neighbors);
void orc_direct(int N, const double *pos, const double *q, int nt,
                const int *targets, double *phi);

#ifdef __cplusplus
}
#endif

#endif /* BBFMM_ORACLE_H */
