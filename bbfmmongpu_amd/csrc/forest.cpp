// forest.cpp -- host construction of HostForest (see forest.h).
//
// Sparse instances follow the reference's uniform octree rules at their own
// depth (anotherBuildFMMHierachy.c, anotherFMMDistribute.c,
// anotherInteractionList.c): the same leaf bucketing, centres obtained by
// halving from the root, particles of a leaf in ascending input order, and
// the interaction criterion (children of the parent's neighbours that are
// not neighbours).  Only cells holding particles are stored; the cells the
// reference would add are empty, carry M = 0 and contribute nothing, so
// each instance's field equals the reference's complete-octree result at
// the same depth up to summation order.
#include "forest.h"
#include "kernels.h"

#include <algorithm>
#include <cstdint>
#include <cstdlib>

namespace bb {

static inline uint32_t morton_key(int x, int y, int z, int bits)
{
  uint32_t m = 0;
  for (int b = 0; b < bits; b++)
    m |= (uint32_t)((((x >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((z >> b) & 1)) << (3 * b);
  return m;
}

static inline void morton_coords(uint32_t m, int bits, int &x, int &y, int &z)
{
  x = y = z = 0;
  for (int b = 0; b < bits; b++) {
    const int s = (int)((m >> (3 * b)) & 7);
    x |= ((s >> 2) & 1) << b;
    y |= ((s >> 1) & 1) << b;
    z |= (s & 1) << b;
  }
}

int forest_from_tree(HostForest &f, const anotherTree *t)
{
  const cell *c = t->c;
  const int maxlev = t->maxlev;
  f = HostForest();
  f.complete = true;
  f.N = t->Nf;
  f.ncell = t->ncell;
  f.maxdepth = maxlev;
  f.levsta.assign(t->levsta, t->levsta + maxlev + 1);
  f.levsta.push_back(t->levend[maxlev] + 1);
  f.celeng.assign(t->celeng, t->celeng + maxlev + 1);
  const int ncell = f.ncell, leaf0 = t->levsta[maxlev];
  f.nleaf = t->levend[maxlev] - leaf0 + 1;
  f.center.resize(ncell);
  f.celllev.resize(ncell);
  f.child.assign((size_t)ncell * 8, -1);
  for (int l = 0; l <= maxlev; l++)
    for (int A = f.levsta[l]; A < f.levsta[l + 1]; A++) {
      f.center[A] = real4{c->center[A].x, c->center[A].y, c->center[A].z, 0};
      f.celllev[A] = l;
      if (l < maxlev)
        for (int j = 0; j < 8; j++) f.child[(size_t)8 * A + j] = 8 * A + 1 + j;
    }
  f.ploff.assign(maxlev + 1, 0);
  for (int l = 0; l < maxlev; l++) {
    f.ploff[l] = (int)f.plist.size();
    for (int A = f.levsta[l]; A < f.levsta[l + 1]; A++) f.plist.push_back(A);
  }
  f.ploff[maxlev] = (int)f.plist.size();
  // leaves: the tree's own leaf ordering and lists
  const real leaflen = t->celeng[0] / (real)(1 << maxlev);
  const real ihalf = (real)1 / (leaflen / (real)2);
  f.leafcell.resize(f.nleaf);
  f.leafsta.resize(f.nleaf);
  f.leafcnt.resize(f.nleaf);
  f.leafihalf.assign(f.nleaf, ihalf);
  f.leafof.assign(f.N, 0);
  for (int k = 0; k < f.nleaf; k++) {
    const int A = leaf0 + k;
    f.leafcell[k] = A;
    f.leafsta[k] = c->fieldsta[A];
    f.leafcnt[k] = c->fieldend[A] - c->fieldsta[A] + 1;
    for (int i = c->fieldsta[A]; i <= c->fieldend[A]; i++) f.leafof[i] = k;
  }
  f.perm.assign(t->fieldlist, t->fieldlist + f.N);
  f.nbr.assign((size_t)f.nleaf * 27, -1);
  f.nnbr.assign(f.nleaf, 0);
  for (int k = 0; k < f.nleaf; k++) {
    const int A = leaf0 + k;
    const int ne = c->ineigh[A];
    if (ne > 27) return set_error(BBFMM_E_ARG, "leaf %d has %d neighbours (> 27)", A, ne);
    f.nnbr[k] = ne;
    for (int b = 0; b < ne; b++) {
      const int B = c->neighbors[(size_t)c->pitch_neighbors * A + b] - leaf0;
      f.nbr[(size_t)k * 27 + b] = B;
      f.pairs += (double)f.leafcnt[k] * (double)f.leafcnt[B];
    }
  }
  f.iinter.assign(c->iinter, c->iinter + ncell);
  f.inter.assign((size_t)ncell * kPitchInteraction, 0);
  f.kidx.assign((size_t)ncell * kPitchInteraction, 0);
  for (int A = 0; A < ncell; A++) {
    if (f.iinter[A] > kPitchInteraction) return set_error(BBFMM_E_ARG, "cell %d has %d interactions", A, f.iinter[A]);
    for (int i = 0; i < f.iinter[A]; i++)
      f.inter[(size_t)kPitchInteraction * A + i] = c->interaction[(size_t)c->pitch_interaction * A + i];
    if (A >= f.levsta[kMinLev]) f.ninter += f.iinter[A];
  }
  int Kt[343];
  ktable(Kt);
  anotherTree tv = *t;
  cell cv = *c;
  cv.interaction = f.inter.data();
  cv.pitch_interaction = kPitchInteraction;
  tv.c = &cv;
  kindex(&tv, Kt, f.kidx.data());
  f.inst_N.assign(1, f.N);
  f.inst_depth.assign(1, maxlev);
  f.inst_p0.assign(1, 0);
  f.inst_leaf0.assign(1, 0);
  return BBFMM_OK;
}

namespace {
struct Level {
  std::vector<uint32_t> key;  // sorted Morton keys of the stored cells
  int first = 0;              // global index of the first cell
  int find(uint32_t k) const
  {
    const auto it = std::lower_bound(key.begin(), key.end(), k);
    return (it != key.end() && *it == k) ? first + (int)(it - key.begin()) : -1;
  }
};

// leaf keys of n points at depth d (bucketing of anotherFMMDistribute.c:258-298)
int bucket(const real3 *pos, int n, int d, real L, std::vector<uint32_t> &key)
{
  std::vector<int> levsta(d + 1), levend(d + 1);
  std::vector<real> celeng(d + 1);
  level_tables(d, L, levsta.data(), levend.data(), celeng.data());
  const real half = celeng[0] / 2;
  const real inv = 1 / celeng[d];
  const int ncpd = 1 << d;
  key.resize(n);
  for (int i = 0; i < n; i++) {
    const real x = pos[i].x, y = pos[i].y, z = pos[i].z;
    int ix = (int)((x + half) * inv), iy = (int)((y + half) * inv), iz = (int)((z + half) * inv);
    if (!(x >= -half && x <= half && y >= -half && y <= half && z >= -half && z <= half))
      return set_error(BBFMM_E_RANGE, "a particle lies outside the root cell [-L/2, L/2]^3");
    ix = std::min(ix, ncpd - 1);
    iy = std::min(iy, ncpd - 1);
    iz = std::min(iz, ncpd - 1);
    key[i] = morton_key(ix, iy, iz, d);
  }
  return BBFMM_OK;
}
}  // namespace

int forest_from_points(HostForest &f, int ninst, const int *Ns, const real3 *pos, const int *depth, int leafmax,
                       real L, std::vector<int> *depth_out)
{
  f = HostForest();
  f.complete = false;
  if (ninst < 1) return set_error(BBFMM_E_ARG, "no instances");
  int Ktab[343];
  ktable(Ktab);
  long long ptot = 0;
  for (int i = 0; i < ninst; i++) {
    if (Ns[i] < 1) return set_error(BBFMM_E_ARG, "instance %d has %d points", i, Ns[i]);
    ptot += Ns[i];
  }
  if (ptot > (1LL << 30)) return set_error(BBFMM_E_ARG, "too many points (%lld)", ptot);
  if (leafmax < 1) leafmax = 32;
  f.N = (int)ptot;
  // choose the depths and count the cells
  std::vector<int> dep(ninst);
  std::vector<std::vector<uint32_t>> keys(ninst);
  int p0 = 0;
  for (int i = 0; i < ninst; i++) {
    const real3 *P = pos + p0;
    int d = (depth && depth[i] > 0) ? depth[i] : kMinLev;
    if (d < kMinLev || d > kMaxForestDepth)
      return set_error(BBFMM_E_ARG, "instance %d: depth %d outside [%d, %d]", i, d, kMinLev, kMaxForestDepth);
    for (;;) {
      BB_TRY(bucket(P, Ns[i], d, L, keys[i]));
      if (depth && depth[i] > 0) break;
      std::vector<uint32_t> s = keys[i];
      std::sort(s.begin(), s.end());
      int run = 1, worst = 1;
      for (size_t j = 1; j < s.size(); j++) {
        run = s[j] == s[j - 1] ? run + 1 : 1;
        worst = std::max(worst, run);
      }
      if (worst <= leafmax || d == kMaxForestDepth) break;
      d++;
    }
    dep[i] = d;
    p0 += Ns[i];
  }
  int dmax = 0;
  for (int i = 0; i < ninst; i++) dmax = std::max(dmax, dep[i]);
  f.maxdepth = dmax;
  {
    std::vector<int> ls(dmax + 1), le(dmax + 1);
    f.celeng.resize(dmax + 1);
    level_tables(dmax, L, ls.data(), le.data(), f.celeng.data());
  }
  f.ploff.assign(dmax + 1, 0);
  std::vector<std::vector<int>> plev(dmax);  // parents per level across instances
  std::vector<int> cellclass;                // sibling class of each cell
  p0 = 0;
  for (int i = 0; i < ninst; i++) {
    const int d = dep[i], n = Ns[i];
    const int c0 = f.ncell, lf0 = f.nleaf;
    f.inst_N.push_back(n);
    f.inst_depth.push_back(d);
    f.inst_p0.push_back(p0);
    f.inst_leaf0.push_back(lf0);
    // particles in leaf order (stable: ascending input index within a leaf)
    std::vector<int> order(n);
    for (int j = 0; j < n; j++) order[j] = j;
    const std::vector<uint32_t> &k = keys[i];
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return k[a] < k[b]; });
    // stored cells per level
    std::vector<Level> lev(d + 1);
    for (int j = 0; j < n; j++) {
      const uint32_t kk = k[order[j]];
      if (lev[d].key.empty() || lev[d].key.back() != kk) lev[d].key.push_back(kk);
    }
    for (int l = d - 1; l >= 0; l--)
      for (uint32_t kk : lev[l + 1].key) {
        const uint32_t pk = kk >> 3;
        if (lev[l].key.empty() || lev[l].key.back() != pk) lev[l].key.push_back(pk);
      }
    int run = c0;
    for (int l = 0; l <= d; l++) {
      lev[l].first = run;
      run += (int)lev[l].key.size();
    }
    f.ncell = run;
    f.center.resize(run);
    f.celllev.resize(run);
    f.child.resize((size_t)run * 8, -1);
    // centres by halving from the root (as build_tree), child table
    f.center[c0] = real4{0, 0, 0, 0};
    f.celllev[c0] = 0;
    for (int l = 1; l <= d; l++) {
      const real h = f.celeng[l] / 2;
      for (size_t j = 0; j < lev[l].key.size(); j++) {
        const uint32_t kk = lev[l].key[j];
        const int A = lev[l].first + (int)j;
        const int P = lev[l - 1].find(kk >> 3);
        const int s = (int)(kk & 7);
        const real4 pc = f.center[P];
        f.center[A] = real4{(s & 4) ? pc.x + h : pc.x - h, (s & 2) ? pc.y + h : pc.y - h,
                            (s & 1) ? pc.z + h : pc.z - h, 0};
        f.celllev[A] = l;
        f.child[(size_t)8 * P + s] = A;
      }
    }
    for (int l = 0; l < d; l++)
      for (size_t j = 0; j < lev[l].key.size(); j++) plev[l].push_back(lev[l].first + (int)j);
    // leaves
    const int nl = (int)lev[d].key.size();
    const real leaflen = L / (real)(1 << d);
    const real ihalf = (real)1 / (leaflen / (real)2);
    int pj = 0;
    for (int q = 0; q < nl; q++) {
      const uint32_t kk = lev[d].key[q];
      const int sta = f.N > 0 ? p0 + pj : 0;
      while (pj < n && k[order[pj]] == kk) {
        f.perm.push_back(p0 + order[pj]);
        f.leafof.push_back(lf0 + q);
        pj++;
      }
      f.leafcell.push_back(lev[d].first + q);
      f.leafsta.push_back(sta);
      f.leafcnt.push_back(p0 + pj - sta);
      f.leafihalf.push_back(ihalf);
    }
    f.nleaf = lf0 + nl;
    // P2P neighbour leaves (self included)
    f.nbr.resize((size_t)f.nleaf * 27, -1);
    f.nnbr.resize(f.nleaf, 0);
    const int ncpd = 1 << d;
    for (int q = 0; q < nl; q++) {
      int x, y, z;
      morton_coords(lev[d].key[q], d, x, y, z);
      int m = 0;
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dz = -1; dz <= 1; dz++) {
            const int X = x + dx, Y = y + dy, Z = z + dz;
            if (X < 0 || Y < 0 || Z < 0 || X >= ncpd || Y >= ncpd || Z >= ncpd) continue;
            const int B = lev[d].find(morton_key(X, Y, Z, d));
            if (B < 0) continue;
            const int bl = B - lev[d].first + lf0;
            f.nbr[(size_t)(lf0 + q) * 27 + m++] = bl;
            f.pairs += (double)f.leafcnt[lf0 + q] * (double)f.leafcnt[bl];
          }
      f.nnbr[lf0 + q] = m;
    }
    // interaction lists: children of the parent's neighbours that are not neighbours
    f.iinter.resize(run, 0);
    f.inter.resize((size_t)run * kPitchInteraction, 0);
    f.kidx.resize((size_t)run * kPitchInteraction, 0);
    f.vsrc.resize((size_t)run * kM2LOffsetsPerClass, -1);
    cellclass.resize(run, 0);
    for (int l = 1; l <= d; l++)
      for (size_t j = 0; j < lev[l].key.size(); j++) cellclass[lev[l].first + j] = (int)(lev[l].key[j] & 7);
    for (int l = kMinLev; l <= d; l++) {
      const int np = 1 << (l - 1);
      for (size_t j = 0; j < lev[l].key.size(); j++) {
        const int F = lev[l].first + (int)j;
        int fx, fy, fz;
        morton_coords(lev[l].key[j], l, fx, fy, fz);
        const int px = fx >> 1, py = fy >> 1, pz = fz >> 1;
        int m = 0;
        for (int dx = -1; dx <= 1; dx++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
              const int X = px + dx, Y = py + dy, Z = pz + dz;
              if (X < 0 || Y < 0 || Z < 0 || X >= np || Y >= np || Z >= np) continue;
              const int U = lev[l - 1].find(morton_key(X, Y, Z, l - 1));
              if (U < 0) continue;
              for (int s = 0; s < 8; s++) {
                const int C = f.child[(size_t)8 * U + s];
                if (C < 0) continue;
                const int vx = 2 * X + ((s >> 2) & 1) - fx, vy = 2 * Y + ((s >> 1) & 1) - fy,
                          vz = 2 * Z + (s & 1) - fz;
                if (std::abs(vx) <= 1 && std::abs(vy) <= 1 && std::abs(vz) <= 1) continue;
                f.inter[(size_t)kPitchInteraction * F + m] = C;
                f.kidx[(size_t)kPitchInteraction * F + m] = Ktab[49 * (vx + 3) + 7 * (vy + 3) + (vz + 3)];
                m++;
                const int slot = class_slot(cellclass[F], vx, vy, vz);
                if (slot >= 0) f.vsrc[(size_t)kM2LOffsetsPerClass * F + slot] = C;
              }
            }
        f.iinter[F] = m;
        f.ninter += m;
      }
    }
    p0 += n;
  }
  for (int l = 0; l < dmax; l++) {
    f.ploff[l] = (int)f.plist.size();
    f.plist.insert(f.plist.end(), plev[l].begin(), plev[l].end());
  }
  f.ploff[dmax] = (int)f.plist.size();
  // class lists for the list MFMA M2L; empty sources point at the zero cell
  for (auto &x : f.vsrc)
    if (x < 0) x = f.ncell;
  f.cls_off.assign(9, 0);
  for (int A = 0; A < f.ncell; A++)
    if (f.celllev[A] >= kMinLev) f.cls_off[cellclass[A] + 1]++;
  for (int c = 0; c < 8; c++) f.cls_off[c + 1] += f.cls_off[c];
  f.cls_cells.assign(f.cls_off[8], 0);
  {
    std::vector<int> fill(f.cls_off.begin(), f.cls_off.end() - 1);
    for (int A = 0; A < f.ncell; A++)
      if (f.celllev[A] >= kMinLev) f.cls_cells[fill[cellclass[A]]++] = A;
  }
  if (depth_out) *depth_out = dep;
  return BBFMM_OK;
}

}  // namespace bb
