/*
 * bbfmm_oracle.c -- CPU restatement of the reference bbFMM path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the GPU library.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load it; the product (bbfmmongpu_amd/) neither links nor calls it.
 *
 * Every function names the reference file:line it restates.  Arithmetic is
 * done in `real` in the same operation order as the reference wherever the
 * reference fixes one, so that the fp32 build can be compared with the
 * compiled reference (oracle/_ref) to round-off.
 *
 * Normalisations of reference quirks (documented, equivalent results):
 *  - proxysval is zeroed before U^T*Sw is accumulated into it (the reference
 *    accumulates into fresh malloc memory, anotherUpwardPass.c:50, atransx.c:242).
 *  - M2M includes all eight children; the reference gates on the parent's
 *    particle count, read from entries that are only set at maxlev
 *    (anotherUpwardPass.c:388, anotherFMMDistribute.c:318-319).  Empty
 *    children contribute zero, so the results are the same.
 *  - sta/end are filled with true ranges at every level (the reference sets
 *    them only at maxlev); leaves are identical.
 */
#include "bbfmm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORC_SINGLE
#define RSQRT_EXACT(x) (1.0f / sqrtf(x))
#define R_SQRT(x) sqrtf(x)
#define R_COS(x) cosf(x)
#define R_POW(x, y) powf(x, y)
#define R_RINT(x) rintf(x)
#else
#define RSQRT_EXACT(x) (1.0 / sqrt(x))
#define R_SQRT(x) sqrt(x)
#define R_COS(x) cos(x)
#define R_POW(x, y) pow(x, y)
#define R_RINT(x) rint(x)
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ZERO ((real)0)
#define ONE ((real)1)
#define TWO ((real)2)
#define MINLEV 2

static inline int pow8(int i) { return 1 << (3 * i); }
static inline int parent_of(int A) { return (A - 1) >> 3; }   /* cell.h:27 */
static inline int child_of(int A, int s) { return 8 * A + 1 + s; } /* cell.h:28 */

/* ------------------------------------------------------------------ */
/* Level rule: bbfmm.c:151-160.  `l` starts at the -d level modifier.  */
/* ------------------------------------------------------------------ */
int orc_maxlev(int N, int lmod)
{
  int l = lmod;
  l += (int)log((double)N) / log(8.0); /* (int)ln N, then / ln 8, truncated on assignment */
  if (l > 0) l = l - 1;
  if (l < MINLEV) l = MINLEV;
  return l;
}

int orc_ncell(int maxlev) { return (pow8(maxlev + 1) - 1) / 7; }

/* anotherBuildFMMHierachy.c:3-21 (globalIndex) and :44-55 */
void orc_levels(int maxlev, real L, int *levsta, int *levend, real *celeng)
{
  int num = 1, s = 0;
  real length = L;
  for (int level = 0; level <= maxlev; level++) {
    levsta[level] = s;
    levend[level] = s + num - 1;
    celeng[level] = length;
    s += num;
    num *= 8;
    length *= (real)0.5;
  }
}

/* anotherFMMDistribute.c:191-255 (comp_center) and :330-351 */
void orc_centers(int maxlev, real L, orc_real3 *center)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L, levsta, levend, celeng);
  center[0].x = ZERO; center[0].y = ZERO; center[0].z = ZERO;
  for (int level = 1; level <= maxlev; level++) {
    const real Lhalf = celeng[level] / 2;
    for (int A = levsta[level]; A <= levend[level]; A++) {
      const int P = parent_of(A);
      const int s = A - child_of(P, 0);
      center[A].x = (s < 4) ? center[P].x - Lhalf : center[P].x + Lhalf;
      center[A].y = (s % 4 < 2) ? center[P].y - Lhalf : center[P].y + Lhalf;
      center[A].z = (s % 2 == 0) ? center[P].z - Lhalf : center[P].z + Lhalf;
    }
  }
}

/*
 * Leaf bucketing: anotherFMMDistribute.c:258-298 (xyz index), :314-484.
 * Particles of a leaf are listed in ascending particle index (the list fill
 * at :462-473 is serial).  Returns 0, or -1 if a particle lies outside the
 * root cell.  sta/end are filled for every level (true ranges).
 */
int orc_distribute(int maxlev, real L, const orc_real3 *center, int N,
                   const orc_real3 *pos, int *list, int *sta, int *end)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L, levsta, levend, celeng);
  const int Asta = levsta[maxlev], Aend = levend[maxlev];
  const int ncpd = 1 << maxlev;
  const int nc = Aend - Asta + 1;
  const real L0half = celeng[0] / 2;
  const real iLm = 1 / celeng[maxlev];
  int *map = (int *)malloc(sizeof(int) * nc);
  int *cnt = (int *)calloc(nc, sizeof(int));
  int *leaf = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
  int rc = 0;
  for (int A = Asta; A <= Aend; A++) {
    const int ix = (int)((center[A].x + L0half) * iLm);
    const int iy = (int)((center[A].y + L0half) * iLm);
    const int iz = (int)((center[A].z + L0half) * iLm);
    map[(iz * ncpd + iy) * ncpd + ix] = A;
  }
  for (int i = 0; i < N; i++) {
    int ix = (int)((pos[i].x + L0half) * iLm);
    int iy = (int)((pos[i].y + L0half) * iLm);
    int iz = (int)((pos[i].z + L0half) * iLm);
    if (ix < 0 || iy < 0 || iz < 0 ||
        !(pos[i].x <= L0half) || !(pos[i].y <= L0half) || !(pos[i].z <= L0half)) {
      rc = -1;
      ix = iy = iz = 0;
    }
    ix = ix < ncpd - 1 ? ix : ncpd - 1;
    iy = iy < ncpd - 1 ? iy : ncpd - 1;
    iz = iz < ncpd - 1 ? iz : ncpd - 1;
    const int A = map[(iz * ncpd + iy) * ncpd + ix];
    leaf[i] = A - Asta;
    cnt[A - Asta]++;
  }
  int l = 0;
  for (int k = 0; k < nc; k++) {
    sta[Asta + k] = l;
    l += cnt[k];
    end[Asta + k] = l - 1;
    cnt[k] = sta[Asta + k];
  }
  for (int i = 0; i < N; i++) list[cnt[leaf[i]]++] = i;
  /* true ranges for non-leaf cells (children of A are contiguous) */
  for (int level = maxlev - 1; level >= 0; level--) {
    for (int A = levsta[level]; A <= levend[level]; A++) {
      sta[A] = sta[child_of(A, 0)];
      end[A] = end[child_of(A, 7)];
    }
  }
  free(map); free(cnt); free(leaf);
  return rc;
}

/* anotherInteractionList.c:3 (1.5 safety factor), :11-111 */
void orc_lists(int maxlev, real L, const orc_real3 *center,
               int *ineigh, int *neighbors, int *iinter, int *interaction)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L, levsta, levend, celeng);
  const int ncell = orc_ncell(maxlev);
  memset(ineigh, 0, sizeof(int) * ncell);
  memset(iinter, 0, sizeof(int) * ncell);
  neighbors[ORC_PITCH_NEIGHBORS * 0 + 0] = 0;
  ineigh[0] = 1;
  for (int level = 1; level <= maxlev; level++) {
    const real Lsafe = celeng[level] * (real)1.5;
    const real Lsafe2 = Lsafe * Lsafe;
    for (int A = levsta[level]; A <= levend[level]; A++) {
      int *nA = &neighbors[ORC_PITCH_NEIGHBORS * A];
      int *iA = &interaction[ORC_PITCH_INTERACTION * A];
      const orc_real3 f = center[A];
      const int P = parent_of(A);
      for (int i = 0; i < ineigh[P]; i++) {
        const int U = neighbors[ORC_PITCH_NEIGHBORS * P + i];
        for (int j = 0; j < 8; j++) {
          const int C = child_of(U, j);
          const real dx = center[C].x - f.x;
          const real dy = center[C].y - f.y;
          const real dz = center[C].z - f.z;
          if (dx * dx <= Lsafe2 && dy * dy <= Lsafe2 && dz * dz <= Lsafe2) {
            nA[ineigh[A]++] = C;
          } else {
            iA[iinter[A]++] = C;
          }
        }
      }
    }
  }
}

/* m2l_aux.c:3-26 (same table as bbfmm.c:933-950) */
void orc_Ktable(int *Ktable)
{
  int ncell = 0, ninteract = 0;
  for (int i = 0; i < 343; i++) Ktable[i] = ORC_NULL;
  for (int vx = -3; vx < 4; vx++)
    for (int vy = -3; vy < 4; vy++)
      for (int vz = -3; vz < 4; vz++) {
        if (abs(vx) > 1 || abs(vy) > 1 || abs(vz) > 1) Ktable[ncell] = ninteract++;
        ncell++;
      }
}

/* m2l_aux.c:1121-1132, :1135-1191 */
void orc_Kindex(int maxlev, real L0, const orc_real3 *center, const int *iinter,
                const int *interaction, const int *Ktable, int *Kindex)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L0, levsta, levend, celeng);
  for (int level = 0; level <= maxlev; level++) {
    const real L = L0 / (real)(1 << level);
    const real iL = ONE / L;
    for (int F = levsta[level]; F <= levend[level]; F++) {
      const orc_real3 fc = center[F];
      for (int i = 0; i < iinter[F]; i++) {
        const int S = interaction[ORC_PITCH_INTERACTION * F + i];
        const real dx = (center[S].x - fc.x) * iL;
        const real dy = (center[S].y - fc.y) * iL;
        const real dz = (center[S].z - fc.z) * iL;
        const int vx = (int)R_RINT(dx), vy = (int)R_RINT(dy), vz = (int)R_RINT(dz);
        Kindex[ORC_PITCH_INTERACTION * F + i] = Ktable[49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)];
      }
    }
  }
}

/* m2l_aux.c:982-1012 (cluster index) and :1015-1118 */
void orc_sourceclusters(int minlev, int maxlev, real L0, const orc_real3 *center,
                        const int *ineigh, const int *neighbors, int *sourceclusters)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L0, levsta, levend, celeng);
  for (int plevel = minlev - 1; plevel <= maxlev - 1; plevel++) {
    const int level = plevel + 1;
    const real length = L0 / (real)(1 << level);
    const real h = (length * 2) / TWO;
    for (int FC = levsta[plevel]; FC <= levend[plevel]; FC++) {
      const int F0 = FC - levsta[minlev - 1];
      int *sc = &sourceclusters[ORC_PITCH_SOURCECLUSTERS * F0];
      for (int i = 0; i < ORC_PITCH_SOURCECLUSTERS; i++) sc[i] = ORC_NULL;
      for (int i = 0; i < ineigh[FC]; i++) {
        const int SC = neighbors[ORC_PITCH_NEIGHBORS * FC + i];
        if (SC == FC) continue;
        const real dx = center[SC].x - center[FC].x;
        const real dy = center[SC].y - center[FC].y;
        const real dz = center[SC].z - center[FC].z;
        const int ix = dx < -h ? 0 : (dx < h ? 1 : 2);
        const int iy = dy < -h ? 0 : (dy < h ? 1 : 2);
        const int iz = dz < -h ? 0 : (dz < h ? 1 : 2);
        sc[ix + 3 * iy + 9 * iz] = child_of(SC, 0);
      }
    }
  }
}

/*
 * Sibling-blocking tables (reference tables: m2l_kern_sibling_blocking.cu:42-107,
 * anotherFMMInteraction.c:7836-7869), derived here from geometry:
 *  Kindex0[d] = Ktable row of the offset 2*(source-cluster offset), d = ix+3iy+9iz;
 *  Kindex1[kind] = 49dx+7dy+dz for the sibling-bit difference of the kind.
 * Kinds are grouped by the number of zero components of the difference
 * (0: 8 kinds, 1: 12, 2: 6, 3: 1); within a group they are ordered by the
 * lowest field sibling, then by decreasing source sibling.
 */
void orc_sibling_tables(int *Kindex0, int *Kindex1)
{
  for (int iz = 0; iz < 3; iz++)
    for (int iy = 0; iy < 3; iy++)
      for (int ix = 0; ix < 3; ix++)
        Kindex0[ix + 3 * iy + 9 * iz] =
            49 * (2 * (ix - 1) + 3) + 7 * (2 * (iy - 1) + 3) + (2 * (iz - 1) + 3);
  int key[27], val[27];
  int k = 0;
  for (int dx = -1; dx <= 1; dx++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dz = -1; dz <= 1; dz++) {
        const int nz = (dx == 0) + (dy == 0) + (dz == 0);
        const int fx = dx < 0, fy = dy < 0, fz = dz < 0; /* lowest field sibling */
        const int f = 4 * fx + 2 * fy + fz;
        const int s = 4 * (fx + dx) + 2 * (fy + dy) + (fz + dz);
        key[k] = nz * 1000 + f * 10 + (7 - s);
        val[k] = 49 * dx + 7 * dy + dz;
        k++;
      }
  for (int i = 0; i < 27; i++) { /* selection sort by key */
    int m = i;
    for (int j = i + 1; j < 27; j++) if (key[j] < key[m]) m = j;
    int t = key[i]; key[i] = key[m]; key[m] = t;
    t = val[i]; val[i] = val[m]; val[m] = t;
    Kindex1[i] = val[i];
  }
}

/* ------------------------------------------------------------------ */
/* Chebyshev helpers: bbfmm.c:1626-1633 (Tk), :1641-1679 (Sn, Clenshaw) */
/* ------------------------------------------------------------------ */
static void compute_Tk(real x, int n, real *vec)
{
  vec[0] = 1;
  vec[1] = x;
  for (int k = 2; k < n; k++) vec[k] = 2 * x * vec[k - 1] - vec[k - 2];
}

static real clenshaw_Sn(const real *vec, real *d, int n, real x)
{
  d[n] = d[n + 1] = 0;
  for (int j = n - 1; j > 0; j--) d[j] = 2 * x * d[j + 1] - d[j + 2] + vec[j];
  return x * d[1] - d[2] + (real)0.5 * vec[0];
}

/* bbfmm.c:1692-1733 (ComputeSn2): Snx[m*N+i] = S_n(x_m, point_i.x) */
static void compute_Sn2(const orc_real3 *point, const real *Tkz, int n, int N,
                        real *Snx, real *Sny, real *Snz)
{
  real d[64];
  for (int m = 0; m < n; m++) {
    const real *vec = &Tkz[m * n];
    for (int i = 0; i < N; i++) {
      Snx[m * N + i] = clenshaw_Sn(vec, d, n, point[i].x);
      Sny[m * N + i] = clenshaw_Sn(vec, d, n, point[i].y);
      Snz[m * N + i] = clenshaw_Sn(vec, d, n, point[i].z);
    }
  }
}

/* ComputeWeights: bbfmm.c:922-1028 */
void orc_weights(int n, real *Tkz, int *Ktable, real *Kweights, real *Cweights)
{
  const int n3 = n * n * n, Nc = 2 * n3;
  real *nodes = (real *)malloc(sizeof(real) * n);
  real *vec = (real *)malloc(sizeof(real) * n);
  orc_real3 *fieldt = (orc_real3 *)malloc(sizeof(orc_real3) * Nc);
  const real pi = (real)M_PI;
  orc_Ktable(Ktable);
  for (int m = 0; m < n; m++) nodes[m] = R_COS(pi * (real)(2 * m + 1) / (real)(2 * n));
  for (int m = 0; m < n; m++) {
    compute_Tk(nodes[m], n, vec);
    for (int k = 0; k < n; k++) Tkz[m * n + k] = vec[k];
  }
  int count = 0;
  for (int l1 = 0; l1 < n; l1++) {
    const real t1 = 1 / R_SQRT(1 - nodes[l1] * nodes[l1]);
    for (int l2 = 0; l2 < n; l2++) {
      const real t2 = t1 / R_SQRT(1 - nodes[l2] * nodes[l2]);
      for (int l3 = 0; l3 < n; l3++) Kweights[count++] = t2 / R_SQRT(1 - nodes[l3] * nodes[l3]);
    }
  }
  int k = 0;
  for (int i = 0; i < 2; i++) {
    const real vz = (i == 0) ? -1 : 1;
    for (int l1 = 0; l1 < n; l1++)
      for (int l2 = 0; l2 < n; l2++)
        for (int l3 = 0; l3 < n; l3++) {
          fieldt[k].x = (real)0.5 * (nodes[l1] + (real)-1);
          fieldt[k].y = (real)0.5 * (nodes[l2] + (real)-1);
          fieldt[k].z = (real)0.5 * (nodes[l3] + vz);
          k++;
        }
  }
  /* Only the z components at points 0..n-1 and n3..n3+n-1 are used
     (bbfmm.c:1007-1020); compute those directly with Clenshaw. */
  real d[66];
  count = 0;
  for (int l1 = 0; l1 < n; l1++)
    for (int l2 = 0; l2 < n; l2++)
      Cweights[count++] = clenshaw_Sn(&Tkz[l1 * n], d, n, fieldt[l2].z);
  for (int l1 = 0; l1 < n; l1++)
    for (int l2 = 0; l2 < n; l2++)
      Cweights[count++] = clenshaw_Sn(&Tkz[l1 * n], d, n, fieldt[n3 + l2].z);
  free(nodes); free(vec); free(fieldt);
}

/*
 * Sampled kernel matrix of ComputeKernelSVD (bbfmm.c:1246-1319), in FP64:
 * K0 is n3 x (316 n3) column-major, K0[(k*n3 + j)*n3 + i] = k(f_i, s_j)/(w_j w_i)
 * with field nodes in the unit cell at the origin and source nodes in the
 * unit cell centred at offset k (Ktable order).
 */
void orc_kernel_matrix(int n, double *K0)
{
  const int n3 = n * n * n;
  double *nodes = (double *)malloc(sizeof(double) * n);
  double *w = (double *)malloc(sizeof(double) * n3);
  for (int m = 0; m < n; m++) nodes[m] = cos(M_PI * (double)(2 * m + 1) / (double)(2 * n));
  int c = 0;
  for (int l1 = 0; l1 < n; l1++) {
    const double t1 = 1 / sqrt(1 - nodes[l1] * nodes[l1]);
    for (int l2 = 0; l2 < n; l2++) {
      const double t2 = t1 / sqrt(1 - nodes[l2] * nodes[l2]);
      for (int l3 = 0; l3 < n; l3++) w[c++] = t2 / sqrt(1 - nodes[l3] * nodes[l3]);
    }
  }
  long count = 0;
  for (int k1 = -3; k1 < 4; k1++)
    for (int k2 = -3; k2 < 4; k2++)
      for (int k3 = -3; k3 < 4; k3++) {
        if (!(abs(k1) > 1 || abs(k2) > 1 || abs(k3) > 1)) continue;
        for (int j = 0; j < n3; j++) {
          const double sx = k1 + 0.5 * nodes[j / (n * n)];
          const double sy = k2 + 0.5 * nodes[(j / n) % n];
          const double sz = k3 + 0.5 * nodes[j % n];
          for (int i = 0; i < n3; i++) {
            const double dx = sx - 0.5 * nodes[i / (n * n)];
            const double dy = sy - 0.5 * nodes[(i / n) % n];
            const double dz = sz - 0.5 * nodes[i % n];
            const double kern = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
            K0[count++] = kern / (w[j] * w[i]);
          }
        }
      }
  free(nodes); free(w);
}

/*
 * Symmetric eigen-decomposition by cyclic Jacobi rotations (textbook method).
 * A (m x m, column-major) is destroyed.  Eigenvalues w are returned in
 * decreasing order with eigenvectors in the columns of V; each eigenvector
 * is signed so that its largest-magnitude entry is positive.
 * Returns the number of sweeps, or -1 if it failed to converge.
 */
int orc_sym_eig(int m, double *A, double *w, double *V)
{
  for (int j = 0; j < m; j++)
    for (int i = 0; i < m; i++) V[j * m + i] = (i == j);
  int sweep;
  for (sweep = 0; sweep < 100; sweep++) {
    double off = 0, diag = 0;
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) {
        if (i == j) diag += A[j * m + i] * A[j * m + i];
        else off += A[j * m + i] * A[j * m + i];
      }
    if (off <= 1e-30 * diag) break;
    for (int p = 0; p < m - 1; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = A[q * m + p];
        if (fabs(apq) < 1e-300) continue;
        const double app = A[p * m + p], aqq = A[q * m + q];
        const double theta = (aqq - app) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        const double c = 1 / sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < m; k++) { /* columns p, q */
          const double akp = A[p * m + k], akq = A[q * m + k];
          A[p * m + k] = c * akp - s * akq;
          A[q * m + k] = s * akp + c * akq;
        }
        for (int k = 0; k < m; k++) { /* rows p, q */
          const double apk = A[k * m + p], aqk = A[k * m + q];
          A[k * m + p] = c * apk - s * aqk;
          A[k * m + q] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; k++) {
          const double vkp = V[p * m + k], vkq = V[q * m + k];
          V[p * m + k] = c * vkp - s * vkq;
          V[q * m + k] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < m; i++) w[i] = A[i * m + i];
  for (int i = 0; i < m; i++) { /* sort descending */
    int b = i;
    for (int j = i + 1; j < m; j++) if (w[j] > w[b]) b = j;
    if (b != i) {
      double t = w[i]; w[i] = w[b]; w[b] = t;
      for (int k = 0; k < m; k++) { t = V[i * m + k]; V[i * m + k] = V[b * m + k]; V[b * m + k] = t; }
    }
    int big = 0;
    for (int k = 1; k < m; k++) if (fabs(V[i * m + k]) > fabs(V[i * m + big])) big = k;
    if (V[i * m + big] < 0) for (int k = 0; k < m; k++) V[i * m + k] = -V[i * m + k];
  }
  return sweep < 100 ? sweep : -1;
}

/* E_k = U^T K_k U (bbfmm.c:1384-1398), U n3 x r column-major, E r x r column-major */
void orc_transfer_operators(int n, int r, const double *K0, const double *U, double *E)
{
  const int n3 = n * n * n;
#pragma omp parallel
  {
    double *KU = (double *)malloc(sizeof(double) * n3 * r);
#pragma omp for
    for (int k = 0; k < 316; k++) {
      const double *Kk = &K0[(long)k * n3 * n3];
      for (int c = 0; c < r; c++)
        for (int i = 0; i < n3; i++) {
          double s = 0;
          for (int j = 0; j < n3; j++) s += Kk[j * n3 + i] * U[c * n3 + j];
          KU[c * n3 + i] = s;
        }
      for (int c = 0; c < r; c++)
        for (int a = 0; a < r; a++) {
          double s = 0;
          for (int i = 0; i < n3; i++) s += U[a * n3 + i] * KU[c * n3 + i];
          E[(long)k * r * r + c * r + a] = s;
        }
    }
    free(KU);
  }
}

/* ------------------------------------------------------------------ */
/* Upward pass: anotherUpwardPass.c:20-571, atransx.c:207-247, xdot.c  */
/* ------------------------------------------------------------------ */
static real xdot(int n, const real *x, int incx, const real *y, int incy)
{
  real sum = ZERO;
  for (int i = 0; i < n; i++) sum += x[i * incx] * y[i * incy];
  return sum;
}

static void atransx(int m, int nn, const real *a, const real *x, real *y)
{
  for (int i = 0; i < nn; i++) {
    real sum = ZERO;
    const real *ai = &a[m * i];
    for (int j = 0; j < m; j++) sum += ai[j] * x[j];
    y[i] += sum;
  }
}

void orc_upward(int maxlev, real L, int n, int r, const orc_real3 *center,
                const int *list, const int *sta, const int *end,
                const orc_real3 *pos, const real *q, const real *U,
                const real *Tkz, const real *Kweights, const real *Cweights,
                real *M)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L, levsta, levend, celeng);
  const int n2 = n * n, n3 = n * n * n;
  const int ncell = orc_ncell(maxlev);
  const real prefac = TWO / (real)n;
  const real prefac3 = prefac * prefac * prefac;
  real *S_all = (real *)calloc((size_t)ncell * n3, sizeof(real));
  memset(M, 0, sizeof(real) * (size_t)ncell * r); /* normalised: zeroed */

  { /* P2M at maxlev: :61-330 */
    const int level = maxlev;
    const real halfL = celeng[level] / TWO;
    const real ihalfL = ONE / halfL;
#pragma omp parallel for schedule(static)
    for (int A = levsta[level]; A <= levend[level]; A++) {
      const int Ns = end[A] - sta[A] + 1;
      real *Sw = (real *)malloc(sizeof(real) * n3);
      orc_real3 *st = (orc_real3 *)malloc(sizeof(orc_real3) * (Ns > 0 ? Ns : 1));
      real *Sx = (real *)malloc(sizeof(real) * n * (Ns > 0 ? Ns : 1));
      real *Sy = (real *)malloc(sizeof(real) * n * (Ns > 0 ? Ns : 1));
      real *Sz = (real *)malloc(sizeof(real) * n * (Ns > 0 ? Ns : 1));
      real *qt = (real *)malloc(sizeof(real) * (Ns > 0 ? Ns : 1));
      for (int j = 0; j < Ns; j++) {
        const int k = list[sta[A] + j];
        st[j].x = ihalfL * (pos[k].x - center[A].x);
        st[j].y = ihalfL * (pos[k].y - center[A].y);
        st[j].z = ihalfL * (pos[k].z - center[A].z);
        qt[j] = q[k];
      }
      compute_Sn2(st, Tkz, n, Ns, Sx, Sy, Sz);
      real *S = &S_all[(size_t)n3 * A];
      int l = 0;
      for (int l1 = 0; l1 < n; l1++)
        for (int l2 = 0; l2 < n; l2++)
          for (int l3 = 0; l3 < n; l3++) {
            real sum = ZERO;
            for (int j = 0; j < Ns; j++) {
              const real wq = qt[j] * Sx[l1 * Ns + j] * Sy[l2 * Ns + j];
              sum += wq * Sz[l3 * Ns + j];
            }
            S[l] = prefac3 * sum;
            Sw[l] = S[l] * Kweights[l];
            l++;
          }
      atransx(n3, r, U, Sw, &M[(size_t)r * A]);
      free(Sw); free(st); free(Sx); free(Sy); free(Sz); free(qt);
    }
  }

  /* M2M: :345-560, all children included (normalised gate) */
  for (int level = maxlev - 1; level >= MINLEV; level--) {
#pragma omp parallel for schedule(static)
    for (int A = levsta[level]; A <= levend[level]; A++) {
      real *Sw = (real *)malloc(sizeof(real) * n3);
      real *Sy = (real *)calloc(2 * n3, sizeof(real));
      real *Sz = (real *)calloc(4 * n3, sizeof(real));
      real *S = &S_all[(size_t)n3 * A];
      for (int l = 0; l < n3; l++) S[l] = ZERO;
      for (int j = 0; j < 8; j++) { /* z */
        int l = (j / 2) * n3;
        const real *Sc = &S_all[(size_t)n3 * child_of(A, j)];
        const int ws = (j % 2 == 0) ? 0 : n2;
        for (int l1 = 0; l1 < n2; l1++)
          for (int l3 = 0; l3 < n; l3++) {
            Sz[l] += xdot(n, &Sc[l1 * n], 1, &Cweights[ws + l3 * n], 1);
            l++;
          }
      }
      for (int j = 0; j < 4; j++) { /* y */
        int l = (j / 2) * n3;
        const int ws = (j % 2 == 0) ? 0 : n2;
        for (int l1 = 0; l1 < n; l1++)
          for (int l2 = 0; l2 < n; l2++)
            for (int l3 = 0; l3 < n; l3++) {
              Sy[l] += xdot(n, &Sz[j * n3 + l1 * n2 + l3], n, &Cweights[ws + l2 * n], 1);
              l++;
            }
      }
      for (int j = 0; j < 2; j++) { /* x */
        int l = 0;
        const int ws = (j == 0) ? 0 : n2;
        for (int l1 = 0; l1 < n; l1++)
          for (int l2 = 0; l2 < n; l2++)
            for (int l3 = 0; l3 < n; l3++) {
              S[l] += xdot(n, &Sy[j * n3 + l2 * n + l3], n2, &Cweights[ws + l1 * n], 1);
              l++;
            }
      }
      for (int l = 0; l < n3; l++) {
        S[l] *= prefac3;
        Sw[l] = S[l] * Kweights[l];
      }
      atransx(n3, r, U, Sw, &M[(size_t)r * A]);
      free(Sw); free(Sy); free(Sz);
    }
  }
  free(S_all);
}

/* ------------------------------------------------------------------ */
/* M2L                                                                 */
/* ------------------------------------------------------------------ */
/* Basic scheme: L(F) = sum_S K[k(F,S)] M(S) in interaction-list order,
   overwriting L (m2l_kern_basic.cu:311-384, m2l_host_basic.cu:18-134).
   K is 316 column-major r x r blocks. */
void orc_m2l_basic(int r, int minlev, int maxlev, const real *K, const real *M,
                   const int *iinter, const int *interaction, const int *Kindex,
                   real *Lv)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, (real)1, levsta, levend, celeng);
#pragma omp parallel for schedule(static)
  for (int F = levsta[minlev]; F <= levend[maxlev]; F++) {
    real *LF = &Lv[(size_t)r * F];
    for (int i = 0; i < r; i++) {
      real acc = ZERO;
      for (int t = 0; t < iinter[F]; t++) {
        const int S = interaction[ORC_PITCH_INTERACTION * F + t];
        const real *Kk = &K[(size_t)Kindex[ORC_PITCH_INTERACTION * F + t] * r * r];
        const real *MS = &M[(size_t)r * S];
        for (int j = 0; j < r; j++) acc += Kk[j * r + i] * MS[j];
      }
      LF[i] = acc;
    }
  }
}

/* Sibling-blocked CPU scheme (anotherFMMInteraction.c:7616-7940 px1/2/4/8 and
   m2lx8421, driver :7952-8170): per field cluster, per source cluster d, per
   interaction kind, L_row += sum_j K_row,j M_j with the row sum formed first. */
void orc_m2l_sibling(int r, int minlev, int maxlev, const real *K, const real *M,
                     const int *sourceclusters, const int *Ktable, real *Lv)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, (real)1, levsta, levend, celeng);
  int Kindex0[27], Kindex1[27];
  orc_sibling_tables(Kindex0, Kindex1);
  /* field-sibling list of each kind, in ascending sibling order */
  int kind_f[27][8], kind_n[27];
  for (int kd = 0; kd < 27; kd++) {
    int v = Kindex1[kd];
    int dx = (v + 57) / 49 - 1; v -= 49 * dx;
    int dy = (v + 8) / 7 - 1; v -= 7 * dy;
    int dz = v;
    kind_n[kd] = 0;
    for (int f = 0; f < 8; f++) {
      const int fx = (f >> 2) & 1, fy = (f >> 1) & 1, fz = f & 1;
      const int sx = fx + dx, sy = fy + dy, sz = fz + dz;
      if (sx < 0 || sx > 1 || sy < 0 || sy > 1 || sz < 0 || sz > 1) continue;
      kind_f[kd][kind_n[kd]++] = f;
    }
  }
  for (int A = levsta[minlev]; A <= levend[maxlev]; A++)
    for (int i = 0; i < r; i++) Lv[(size_t)r * A + i] = ZERO;
  for (int level = minlev; level <= maxlev; level++) {
    const int nclu = (levend[level] - levsta[level] + 1) / 8;
#pragma omp parallel for schedule(static)
    for (int Fc = 0; Fc < nclu; Fc++) {
      const int A0 = levsta[level] + 8 * Fc;
      const int F0 = (pow8(level - 1) - pow8(minlev - 1)) / 7 + Fc;
      const int *B0 = &sourceclusters[ORC_PITCH_SOURCECLUSTERS * F0];
      for (int d = 0; d < 27; d++) {
        if (B0[d] == ORC_NULL) continue;
        for (int kd = 0; kd < 27; kd++) {
          const int kb = Ktable[Kindex0[d] + Kindex1[kd]];
          if (kb == ORC_NULL) continue;
          int v = Kindex1[kd];
          int dx = (v + 57) / 49 - 1; v -= 49 * dx;
          int dy = (v + 8) / 7 - 1; v -= 7 * dy;
          const int dsib = 4 * dx + 2 * dy + v;
          const real *Kk = &K[(size_t)kb * r * r];
          for (int t = 0; t < kind_n[kd]; t++) {
            const int f = kind_f[kd][t];
            real *LF = &Lv[(size_t)r * (A0 + f)];
            const real *MS = &M[(size_t)r * (B0[d] + f + dsib)];
            for (int i = 0; i < r; i++) {
              real sum = ZERO;
              for (int j = 0; j < r; j++) sum += Kk[j * r + i] * MS[j];
              LF[i] += sum;
            }
          }
        }
      }
    }
  }
}

/* postm2l: auxAnotherFMMInteraction.c:364-416 with the per-level scale of
   anotherFMMInteraction.cu:89-113: F = (1/L_l)^homogen Kw .* (U L). */
void orc_postm2l(int maxlev, real L, int n, int r, real homogen, const real *U,
                 const real *Kweights, const real *Lv, real *F)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L, levsta, levend, celeng);
  const int n3 = n * n * n;
  const int ncell = orc_ncell(maxlev);
  memset(F, 0, sizeof(real) * (size_t)ncell * n3);
  for (int level = MINLEV; level <= maxlev; level++) {
    const real iL = ONE / celeng[level];
    const real scale = R_POW(iL, homogen);
#pragma omp parallel for schedule(static)
    for (int A = levsta[level]; A <= levend[level]; A++)
      for (int i = 0; i < n3; i++) {
        real sum = ZERO;
        for (int j = 0; j < r; j++) sum += U[i + n3 * j] * Lv[(size_t)r * A + j];
        F[(size_t)n3 * A + i] = scale * Kweights[i] * sum;
      }
  }
}

/* ------------------------------------------------------------------ */
/* Downward pass: anotherDownwardPassX.c:40-211 (L2L), :226-424 (L2P)  */
/* ------------------------------------------------------------------ */
/* L2P of orders n % 4 == 3: 0 (default) restates the reference's reuse of
   its second-to-last z row; 1 evaluates the Chebyshev interpolant itself. */
static int orc_l2p_exact = 0;
void orc_set_l2p_exact(int on) { orc_l2p_exact = on != 0; }

void orc_downward(int maxlev, real L, int n, const orc_real3 *center,
                  const int *list, const int *sta, const int *end,
                  const orc_real3 *pos, const real *q, const real *Tkz,
                  const real *Cweights, real *F, real *phi)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, L, levsta, levend, celeng);
  const int n2 = n * n, n3 = n * n * n;
  const real prefac = TWO / (real)n;
  const real prefac3 = prefac * prefac * prefac;
  for (int level = MINLEV; level < maxlev; level++) {
#pragma omp parallel for schedule(static)
    for (int A = levsta[level]; A <= levend[level]; A++) {
      const real *FA = &F[(size_t)n3 * A];
      real *Fx = (real *)calloc(2 * n3, sizeof(real));
      real *Fy = (real *)calloc(4 * n3, sizeof(real));
      int present[8];
      for (int i = 0; i < 8; i++) {
        const int B = child_of(A, i);
        present[i] = (end[B] - sta[B] + 1) != 0;
      }
      for (int j = 0; j < 2; j++) { /* x */
        if (!(present[4 * j] | present[4 * j + 1] | present[4 * j + 2] | present[4 * j + 3])) continue;
        int l = j * n3;
        const int ws = (j == 0) ? 0 : n2;
        for (int l1 = 0; l1 < n; l1++)
          for (int l2 = 0; l2 < n; l2++)
            for (int l3 = 0; l3 < n; l3++) {
              Fx[l] = xdot(n, &FA[l2 * n + l3], n2, &Cweights[ws + l1], n);
              l++;
            }
      }
      for (int j = 0; j < 4; j++) { /* y */
        if (!(present[2 * j] | present[2 * j + 1])) continue;
        int l = j * n3;
        const int ws = (j % 2 == 0) ? 0 : n2;
        for (int l1 = 0; l1 < n; l1++)
          for (int l2 = 0; l2 < n; l2++)
            for (int l3 = 0; l3 < n; l3++) {
              Fy[l] = xdot(n, &Fx[(j / 2) * n3 + l1 * n2 + l3], n, &Cweights[ws + l2], n);
              l++;
            }
      }
      for (int j = 0; j < 8; j++) { /* z */
        if (!present[j]) continue;
        int l = 0;
        real *Fc = &F[(size_t)n3 * child_of(A, j)];
        const int ws = (j % 2 == 0) ? 0 : n2;
        for (int l1 = 0; l1 < n2; l1++)
          for (int l3 = 0; l3 < n; l3++) {
            Fc[l] += prefac3 * xdot(n, &Fy[(j / 2) * n3 + l1 * n], 1, &Cweights[ws + l3], n);
            l++;
          }
      }
      free(Fx); free(Fy);
    }
  }
  { /* L2P */
    const int level = maxlev;
    const real halfL = celeng[level] / TWO;
    const real ihalfL = ONE / halfL;
#pragma omp parallel for schedule(static)
    for (int A = levsta[level]; A <= levend[level]; A++) {
      const int Nf = end[A] - sta[A] + 1;
      if (Nf <= 0) continue;
      const real *FA = &F[(size_t)n3 * A];
      orc_real3 *ft = (orc_real3 *)malloc(sizeof(orc_real3) * Nf);
      real *Sx = (real *)malloc(sizeof(real) * n * Nf);
      real *Sy = (real *)malloc(sizeof(real) * n * Nf);
      real *Sz = (real *)malloc(sizeof(real) * n * Nf);
      real *sum = (real *)calloc(Nf, sizeof(real));
      for (int i = 0; i < Nf; i++) {
        const int k = list[sta[A] + i];
        ft[i].x = ihalfL * (pos[k].x - center[A].x);
        ft[i].y = ihalfL * (pos[k].y - center[A].y);
        ft[i].z = ihalfL * (pos[k].z - center[A].z);
      }
      compute_Sn2(ft, Tkz, n, Nf, Sx, Sy, Sz);
      int l = 0;
      for (int l1 = 0; l1 < n; l1++) {
        const real *x1 = &Sx[l1 * Nf];
        for (int l2 = 0; l2 < n; l2++) {
          const real *y2 = &Sy[l2 * Nf];
          for (int l3 = 0; l3 < n - 3; l3 += 4) {
            const real *z0 = &Sz[l3 * Nf], *z1 = &Sz[(l3 + 1) * Nf];
            const real *z2 = &Sz[(l3 + 2) * Nf], *z3 = &Sz[(l3 + 3) * Nf];
            for (int i = 0; i < Nf; i++)
              sum[i] += x1[i] * y2[i] * (z0[i] * FA[l] + z1[i] * FA[l + 1] + z2[i] * FA[l + 2] + z3[i] * FA[l + 3]);
            l += 4;
          }
          if (n % 4 == 1) {
            const real *z0 = &Sz[(n - 1) * Nf];
            for (int i = 0; i < Nf; i++) sum[i] += x1[i] * y2[i] * z0[i] * FA[l];
            l += 1;
          } else if (n % 4 == 2) {
            const real *z0 = &Sz[(n - 2) * Nf], *z1 = &Sz[(n - 1) * Nf];
            for (int i = 0; i < Nf; i++) sum[i] += x1[i] * y2[i] * (z0[i] * FA[l] + z1[i] * FA[l + 1]);
            l += 2;
          } else if (n % 4 == 3) {
            /* the reference uses its second z row twice here (anotherDownwardPassX.c:304);
               orc_set_l2p_exact(1) takes the last row instead (the interpolant itself) */
            const real *z0 = &Sz[(n - 3) * Nf], *z1 = &Sz[(n - 2) * Nf];
            const real *z2 = orc_l2p_exact ? &Sz[(n - 1) * Nf] : z1;
            for (int i = 0; i < Nf; i++)
              sum[i] += x1[i] * y2[i] * (z0[i] * FA[l] + z1[i] * FA[l + 1] + z2[i] * FA[l + 2]);
            l += 3;
          }
        }
      }
      for (int i = 0; i < Nf; i++) {
        const int k = list[sta[A] + i];
        phi[k] = q[k] * prefac3 * sum[i];
      }
      free(ft); free(Sx); free(Sy); free(Sz); free(sum);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Near field CPU8A: anotherNearField.c:137-206, :219-346.  phi holds  */
/* the far-field value on entry; q_i * sum_j q_j / r_ij is added.      */
/* ------------------------------------------------------------------ */
void orc_nearfield(int maxlev, int N, const int *list, const int *sta, const int *end,
                   const int *ineigh, const int *neighbors, const orc_real3 *pos,
                   const real *q, real *phi)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, (real)1, levsta, levend, celeng);
  real *x = (real *)malloc(sizeof(real) * (N > 0 ? N : 1));
  real *y = (real *)malloc(sizeof(real) * (N > 0 ? N : 1));
  real *z = (real *)malloc(sizeof(real) * (N > 0 ? N : 1));
  real *p = (real *)malloc(sizeof(real) * (N > 0 ? N : 1));
  real *qs = (real *)malloc(sizeof(real) * (N > 0 ? N : 1));
  for (int A = levsta[maxlev]; A <= levend[maxlev]; A++)
    for (int i = sta[A]; i <= end[A]; i++) {
      const int k = list[i];
      x[i] = pos[k].x; y[i] = pos[k].y; z[i] = pos[k].z;
      p[i] = phi[k]; qs[i] = q[k];
    }
#pragma omp parallel for schedule(static)
  for (int A = levsta[maxlev]; A <= levend[maxlev]; A++) {
    for (int i = sta[A]; i <= end[A]; i++) {
      const real fx = x[i], fy = y[i], fz = z[i];
      real v = ZERO;
      for (int b = 0; b < ineigh[A]; b++) {
        const int B = neighbors[ORC_PITCH_NEIGHBORS * A + b];
        for (int j = sta[B]; j <= end[B]; j++) {
          const real dx = x[j] - fx, dy = y[j] - fy, dz = z[j] - fz;
          const real rr = dx * dx + dy * dy + dz * dz;
          const real rinv = (rr != ZERO ? ONE / R_SQRT(rr) : ZERO);
          v += rinv * qs[j];
        }
      }
      p[i] += qs[i] * v;
    }
  }
  for (int A = levsta[maxlev]; A <= levend[maxlev]; A++)
    for (int i = sta[A]; i <= end[A]; i++) phi[list[i]] = p[i];
  free(x); free(y); free(z); free(p); free(qs);
}

/* Pair count of anotherNearField.c:326-339 / anotherDownwardPass.cu:387-400
   (self pairs included). */
double orc_p2p_pairs(int maxlev, const int *sta, const int *end,
                     const int *ineigh, const int *neighbors)
{
  int levsta[32], levend[32];
  real celeng[32];
  orc_levels(maxlev, (real)1, levsta, levend, celeng);
  double pairs = 0;
  for (int A = levsta[maxlev]; A <= levend[maxlev]; A++) {
    const int nf = end[A] - sta[A] + 1;
    for (int b = 0; b < ineigh[A]; b++) {
      const int B = neighbors[ORC_PITCH_NEIGHBORS * A + b];
      pairs += (double)nf * (double)(end[B] - sta[B] + 1);
    }
  }
  return pairs;
}

/* FP64 direct sum (DirectCalc3D with start=0 and EvaluateField:
   bbfmm.c:471-504, :880-913): phi_t = sum_{j: x_j != x_t} q_j / |x_t - x_j|
   for each target t (all particles when targets is NULL).  The caller
   multiplies by q_t to obtain the reference's output convention. */
void orc_direct(int N, const double *pos, const double *q, int nt,
                const int *targets, double *phi)
{
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < nt; t++) {
    const int i = targets ? targets[t] : t;
    const double fx = pos[3 * i], fy = pos[3 * i + 1], fz = pos[3 * i + 2];
    double s = 0;
    for (int j = 0; j < N; j++) {
      const double sx = pos[3 * j], sy = pos[3 * j + 1], sz = pos[3 * j + 2];
      if (sx != fx || sy != fy || sz != fz) {
        const double dx = sx - fx, dy = sy - fy, dz = sz - fz;
        s += q[j] / sqrt(dx * dx + dy * dy + dz * dz);
      }
    }
    phi[t] = s;
  }
}
