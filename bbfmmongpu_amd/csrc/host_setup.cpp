// host_setup.cpp -- host-side setup of the bbFMM plan: octree, particle
// buckets, neighbour/interaction lists, Chebyshev weights and the FP64
// operator precompute.  Setup runs once per geometry on the host, as in the
// reference (anotherBuildFMMHierachy.c, bbfmm.c); every solve pass runs on
// the GPU (see kernels.hip).
//
// Index rules are kept bit-exact with the reference:
//  - levsta(l) = (8^l-1)/7, child(A,s) = 8A+1+s, parent(A) = (A-1)>>3
//    (anotherBuildFMMHierachy.c:3-21, cell.h:27-28);
//  - centres by the recursive +-Lhalf rule (anotherFMMDistribute.c:191-255);
//  - leaf of a particle by truncation of (x+L/2)/L_maxlev, clamped
//    (anotherFMMDistribute.c:258-298), particles of a leaf in ascending index;
//  - lists: uncles in the parent's neighbour order x children 0..7, a cousin
//    is a neighbour when every |dc|^2 <= (1.5 L_l)^2 (anotherInteractionList.c:11-111).
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace bb {

int maxlev_rule(int N, int lmod)
{
  // bbfmm.c:151-160: l += (int)ln(N) / ln(8) (truncated on assignment)
  int l = lmod;
  l += (int)std::log((double)N) / std::log(8.0);
  if (l > 0) l = l - 1;
  if (l < kMinLev) l = kMinLev;
  return l;
}

void level_tables(int maxlev, real L, int *levsta, int *levend, real *celeng)
{
  int num = 1, first = 0;
  real length = L;
  for (int level = 0; level <= maxlev; level++) {
    levsta[level] = first;
    levend[level] = first + num - 1;
    celeng[level] = length;
    first += num;
    num *= 8;
    length *= (real)0.5;
  }
}

// Builds centres, leaf buckets and lists into the reference's anotherTree
// layout.  Returns BBFMM_OK or BBFMM_E_RANGE.
int build_tree(anotherTree *t, real L, int maxlev, int N, const real3 *pos)
{
  t->minlev = kMinLev;
  t->maxlev = maxlev;
  t->levsta = (int *)malloc(sizeof(int) * (maxlev + 1));
  t->levend = (int *)malloc(sizeof(int) * (maxlev + 1));
  t->celeng = (real *)malloc(sizeof(real) * (maxlev + 1));
  level_tables(maxlev, L, t->levsta, t->levend, t->celeng);
  const int ncell = t->levend[maxlev] + 1;
  t->ncell = ncell;
  t->ncell0 = 0;
  t->Nf = t->Ns = N;
  t->fieldval = t->sourceval = t->proxysval = nullptr;
  cell *c = (cell *)malloc(sizeof(cell));
  t->c = c;
  c->center = (real3 *)malloc(sizeof(real3) * ncell);
  c->fieldsta = (int *)malloc(sizeof(int) * ncell);
  c->fieldend = (int *)malloc(sizeof(int) * ncell);
  c->sourcesta = c->fieldsta;
  c->sourceend = c->fieldend;
  c->ineigh = (int *)calloc(ncell, sizeof(int));
  c->iinter = (int *)calloc(ncell, sizeof(int));
  c->pitch_neighbors = kPitchNeighbors;
  c->pitch_interaction = kPitchInteraction;
  c->neighbors = (int *)malloc(sizeof(int) * (size_t)ncell * kPitchNeighbors);
  c->interaction = (int *)malloc(sizeof(int) * (size_t)ncell * kPitchInteraction);
  t->fieldlist = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
  t->sourcelist = t->fieldlist;

  // centres
  real3 *ctr = c->center;
  ctr[0].x = ctr[0].y = ctr[0].z = (real)0;
  for (int level = 1; level <= maxlev; level++) {
    const real h = t->celeng[level] / 2;
    for (int A = t->levsta[level]; A <= t->levend[level]; A++) {
      const int P = (A - 1) >> 3;
      const int s = A - (8 * P + 1);
      ctr[A].x = (s & 4) ? ctr[P].x + h : ctr[P].x - h;
      ctr[A].y = (s & 2) ? ctr[P].y + h : ctr[P].y - h;
      ctr[A].z = (s & 1) ? ctr[P].z + h : ctr[P].z - h;
    }
  }

  // leaf buckets
  const int Asta = t->levsta[maxlev];
  const int nleaf = t->levend[maxlev] - Asta + 1;
  const int ncpd = 1 << maxlev;
  const real half = t->celeng[0] / 2;
  const real inv = 1 / t->celeng[maxlev];
  std::vector<int> xyz2leaf((size_t)nleaf), leafof((size_t)(N > 0 ? N : 1)), cnt((size_t)nleaf, 0);
  for (int k = 0; k < nleaf; k++) {
    const int A = Asta + k;
    const int ix = (int)((ctr[A].x + half) * inv);
    const int iy = (int)((ctr[A].y + half) * inv);
    const int iz = (int)((ctr[A].z + half) * inv);
    xyz2leaf[((size_t)iz * ncpd + iy) * ncpd + ix] = k;
  }
  int rc = BBFMM_OK;
  for (int i = 0; i < N; i++) {
    const real x = pos[i].x, y = pos[i].y, z = pos[i].z;
    int ix = (int)((x + half) * inv), iy = (int)((y + half) * inv), iz = (int)((z + half) * inv);
    if (!(x >= -half && x <= half && y >= -half && y <= half && z >= -half && z <= half)) {
      rc = BBFMM_E_RANGE;
      ix = iy = iz = 0;
    }
    ix = std::min(ix, ncpd - 1);
    iy = std::min(iy, ncpd - 1);
    iz = std::min(iz, ncpd - 1);
    const int k = xyz2leaf[((size_t)iz * ncpd + iy) * ncpd + ix];
    leafof[i] = k;
    cnt[k]++;
  }
  {
    int run = 0;
    for (int k = 0; k < nleaf; k++) {
      c->fieldsta[Asta + k] = run;
      run += cnt[k];
      c->fieldend[Asta + k] = run - 1;
      cnt[k] = c->fieldsta[Asta + k];
    }
    for (int i = 0; i < N; i++) t->fieldlist[cnt[leafof[i]]++] = i;
  }
  // contiguous particle ranges of the coarser cells (the reference leaves
  // these entries unset; here they hold the true subtree ranges)
  for (int level = maxlev - 1; level >= 0; level--)
    for (int A = t->levsta[level]; A <= t->levend[level]; A++) {
      c->fieldsta[A] = c->fieldsta[8 * A + 1];
      c->fieldend[A] = c->fieldend[8 * A + 8];
    }

  // neighbour and interaction lists
  int *nb = c->neighbors, *it = c->interaction;
  nb[0] = 0;
  c->ineigh[0] = 1;
  for (int level = 1; level <= maxlev; level++) {
    const real safe = t->celeng[level] * (real)1.5;
    const real safe2 = safe * safe;
    for (int A = t->levsta[level]; A <= t->levend[level]; A++) {
      const int P = (A - 1) >> 3;
      int ne = 0, ni = 0;
      for (int u = 0; u < c->ineigh[P]; u++) {
        const int Ucell = nb[(size_t)kPitchNeighbors * P + u];
        for (int j = 0; j < 8; j++) {
          const int Cc = 8 * Ucell + 1 + j;
          const real dx = ctr[Cc].x - ctr[A].x;
          const real dy = ctr[Cc].y - ctr[A].y;
          const real dz = ctr[Cc].z - ctr[A].z;
          if (dx * dx <= safe2 && dy * dy <= safe2 && dz * dz <= safe2)
            nb[(size_t)kPitchNeighbors * A + ne++] = Cc;
          else
            it[(size_t)kPitchInteraction * A + ni++] = Cc;
        }
      }
      c->ineigh[A] = ne;
      c->iinter[A] = ni;
    }
  }
  return rc;
}

void free_tree(anotherTree *t)
{
  if (!t) return;
  cell *c = t->c;
  if (c) {
    free(c->neighbors);
    free(c->interaction);
    free(c->center);
    free(c->fieldsta);
    free(c->fieldend);
    free(c->ineigh);
    free(c->iinter);
    free(c);
  }
  free(t->levsta);
  free(t->levend);
  free(t->celeng);
  free(t->fieldlist);
  free(t);
}

// 343-entry table from offset (vx,vy,vz) in [-3,3]^3 to operator index
// (lexicographic over the 316 offsets outside [-1,1]^3): m2l_aux.c:3-26.
void ktable(int *Kt)
{
  int idx = 0, op = 0;
  for (int vx = -3; vx <= 3; vx++)
    for (int vy = -3; vy <= 3; vy++)
      for (int vz = -3; vz <= 3; vz++, idx++)
        Kt[idx] = (std::abs(vx) > 1 || std::abs(vy) > 1 || std::abs(vz) > 1) ? op++ : -1;
}

// Inverse of the Ktable: the offset of operator k as (vx + 3) 49 + (vy + 3) 7 + (vz + 3).
void operator_offsets(int *op)
{
  int Kt[343];
  ktable(Kt);
  for (int idx = 0; idx < 343; idx++)
    if (Kt[idx] >= 0) op[Kt[idx]] = idx;
}

// Operator index of every interaction (m2l_aux.c:1135-1191): offset in units
// of the cell length, rounded to nearest.
void kindex(const anotherTree *t, const int *Kt, int *Kindex)
{
  const cell *c = t->c;
  for (int level = 0; level <= t->maxlev; level++) {
    const real Ll = t->celeng[0] / (real)(1 << level);
    const real iL = (real)1 / Ll;
    for (int F = t->levsta[level]; F <= t->levend[level]; F++)
      for (int i = 0; i < c->iinter[F]; i++) {
        const int S = c->interaction[(size_t)kPitchInteraction * F + i];
        const int vx = (int)std::rint((c->center[S].x - c->center[F].x) * iL);
        const int vy = (int)std::rint((c->center[S].y - c->center[F].y) * iL);
        const int vz = (int)std::rint((c->center[S].z - c->center[F].z) * iL);
        Kindex[(size_t)kPitchInteraction * F + i] = Kt[49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)];
      }
  }
}

// Chebyshev helpers (bbfmm.c:1626-1679)
static real cheb_S(const real *Tm, int n, real x)
{
  // S_n(x_m, x) = 1/2 + sum_{k>=1} T_k(x_m) T_k(x), Clenshaw recurrence
  real b1 = 0, b2 = 0;
  for (int j = n - 1; j > 0; j--) {
    const real b0 = 2 * x * b1 - b2 + Tm[j];
    b2 = b1;
    b1 = b0;
  }
  return x * b1 - b2 + (real)0.5 * Tm[0];
}

// ComputeWeights (bbfmm.c:922-1028)
void weights(int n, real *Tkz, int *Kt, real *Kw, real *Cw)
{
  std::vector<real> nodes(n);
  const real pi = (real)3.14159265358979323846;
  ktable(Kt);
  for (int m = 0; m < n; m++) nodes[m] = std::cos(pi * (real)(2 * m + 1) / (real)(2 * n));
  for (int m = 0; m < n; m++) {
    real *T = &Tkz[m * n];
    T[0] = 1;
    if (n > 1) T[1] = nodes[m];
    for (int k = 2; k < n; k++) T[k] = 2 * nodes[m] * T[k - 1] - T[k - 2];
  }
  int c = 0;
  for (int a = 0; a < n; a++) {
    const real w1 = 1 / std::sqrt(1 - nodes[a] * nodes[a]);
    for (int b = 0; b < n; b++) {
      const real w2 = w1 / std::sqrt(1 - nodes[b] * nodes[b]);
      for (int d = 0; d < n; d++) Kw[c++] = w2 / std::sqrt(1 - nodes[d] * nodes[d]);
    }
  }
  // Cw[side*n*n + parent_node*n + child_node] = S_n(x_parent, child node
  // mapped into the parent: (x_child + side) / 2), side -1 -> 0, +1 -> 1.
  for (int side = 0; side < 2; side++)
    for (int p = 0; p < n; p++)
      for (int ch = 0; ch < n; ch++) {
        const real xc = (real)0.5 * (nodes[ch] + (side ? (real)1 : (real)-1));
        Cw[side * n * n + p * n + ch] = cheb_S(&Tkz[p * n], n, xc);
      }
}

} // namespace bb
