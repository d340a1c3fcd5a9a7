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
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <thread>

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
  int first = 0;              // index of the first cell (instance-local)
  int find(uint32_t k) const
  {
    const auto it = std::lower_bound(key.begin(), key.end(), k);
    return (it != key.end() && *it == k) ? first + (int)(it - key.begin()) : -1;
  }
};

// One instance built on its own (instance-local cell, leaf and particle
// numbering); forest_from_points concatenates them.
struct Inst {
  int n = 0, d = 0, ncell = 0, nleaf = 0, rc = BBFMM_OK;
  std::vector<real4> center;
  std::vector<int> celllev, child, cls;           // [ncell], [ncell * 8], [ncell]
  std::vector<std::vector<int>> plev;             // cells with children per level 0..d-1
  std::vector<int> perm, leafof;                  // [n]
  std::vector<int> leafcell, leafsta, leafcnt;    // [nleaf]
  std::vector<Level> lev;                         // stored cells per level (sorted keys)
  double pairs = 0, ninter = 0;
};

// Leaf keys of the reference's bucketing (anotherFMMDistribute.c:258-298) at
// depth kMaxForestDepth.  The cell length at depth d is L / 2^d (halved
// exactly level by level), so (x + L/2) / L_d = 2^(D - d) times its value at
// depth D in floating point, and truncation commutes with the exact shift:
// the key at depth d <= D is the depth-D key shifted right by 3 (D - d) bits,
// the clamp to 2^d - 1 included.
int bucket_max(const real3 *pos, int n, real L, std::vector<uint64_t> &ki)
{
  constexpr int D = kMaxForestDepth;
  std::vector<int> levsta(D + 1), levend(D + 1);
  std::vector<real> celeng(D + 1);
  level_tables(D, L, levsta.data(), levend.data(), celeng.data());
  const real half = celeng[0] / 2;
  const real inv = 1 / celeng[D];
  const int ncpd = 1 << D;
  ki.resize(n);
  for (int i = 0; i < n; i++) {
    const real x = pos[i].x, y = pos[i].y, z = pos[i].z;
    int ix = (int)((x + half) * inv), iy = (int)((y + half) * inv), iz = (int)((z + half) * inv);
    if (!(x >= -half && x <= half && y >= -half && y <= half && z >= -half && z <= half)) return BBFMM_E_RANGE;
    ix = std::min(ix, ncpd - 1);
    iy = std::min(iy, ncpd - 1);
    iz = std::min(iz, ncpd - 1);
    // key in the high word, index in the low word: one sort gives the leaf
    // order with ascending input index inside a leaf
    ki[i] = ((uint64_t)morton_key(ix, iy, iz, D) << 32) | (uint32_t)i;
  }
  return BBFMM_OK;
}

// Offset tables of the list construction, by the sibling class sf of the
// field cell: far_slot[(sf * 27 + u) * 8 + sc] = the class slot of the
// offset to child sc of uncle u (u = 9 (dx + 1) + 3 (dy + 1) + dz + 1), -1
// when near, and slot_uc its inverse (each of the 189 slots of a class is one
// (uncle, child)); near_uc[sf * 27 + e] = (uncle, child) of the neighbour at
// the offset e of the same numbering.
struct ListTabs {
  std::vector<int> far_slot;
  std::vector<int2> near_uc;
  std::vector<int2> slot_uc;  // [sf * 189 + slot]: the (uncle, child) of each class slot
  ListTabs() : far_slot(8 * 27 * 8, -1), near_uc(8 * 27), slot_uc(8 * kM2LOffsetsPerClass, make_int2(-1, -1))
  {
    for (int sf = 0; sf < 8; sf++) {
      const int fb[3] = {(sf >> 2) & 1, (sf >> 1) & 1, sf & 1};
      for (int u = 0; u < 27; u++) {
        const int du[3] = {u / 9 - 1, (u / 3) % 3 - 1, u % 3 - 1};
        for (int sc = 0; sc < 8; sc++) {
          const int cb[3] = {(sc >> 2) & 1, (sc >> 1) & 1, sc & 1};
          int v[3];
          bool near = true;
          for (int a = 0; a < 3; a++) {
            v[a] = 2 * du[a] + cb[a] - fb[a];
            near = near && v[a] >= -1 && v[a] <= 1;
          }
          const int slot = near ? -1 : class_slot(sf, v[0], v[1], v[2]);
          far_slot[((size_t)sf * 27 + u) * 8 + sc] = slot;
          if (slot >= 0) slot_uc[(size_t)sf * kM2LOffsetsPerClass + slot] = make_int2(u, sc);
        }
      }
      for (int e = 0; e < 27; e++) {
        const int de[3] = {e / 9 - 1, (e / 3) % 3 - 1, e % 3 - 1};
        int u = 0, sc = 0;
        for (int a = 0; a < 3; a++) {
          const int t = fb[a] + de[a];               // in [-1, 2]
          const int dd = t >= 0 ? t / 2 : -1;        // floor(t / 2)
          u = 3 * u + (dd + 1);
          sc = 2 * sc + (t - 2 * dd);
        }
        near_uc[sf * 27 + e] = make_int2(u, sc);
      }
    }
  }
};

void build_instance(Inst &I, const real3 *P, int n, int depth, int leafmax, real L, const std::vector<real> &celeng,
                    const ListTabs &tabs)
{
  constexpr int D = kMaxForestDepth;
  I.n = n;
  std::vector<uint64_t> ki;
  if ((I.rc = bucket_max(P, n, L, ki)) != BBFMM_OK) return;
  std::sort(ki.begin(), ki.end());
  int d = depth > 0 ? depth : kMinLev;
  if (depth <= 0) {
    // the smallest depth whose leaves hold at most leafmax points
    for (;; d++) {
      const int sh = 32 + 3 * (D - d);
      int run = 1, worst = 1;
      for (int j = 1; j < n; j++) {
        run = (ki[j] >> sh) == (ki[j - 1] >> sh) ? run + 1 : 1;
        worst = std::max(worst, run);
      }
      if (worst <= leafmax || d == D) break;
    }
  }
  I.d = d;
  const int sh = 32 + 3 * (D - d);
  if (d < D) {
    // leaf order at depth d: ascending input index inside a leaf (the
    // depth-D order is by sub-cell first)
    for (auto &x : ki) x = ((x >> sh) << 32) | (x & 0xffffffffu);
    std::sort(ki.begin(), ki.end());
  }
  const int shd = 32;  // the key now holds the depth-d leaf
  // stored cells per level
  std::vector<Level> lev(d + 1);
  for (int j = 0; j < n; j++) {
    const uint32_t kk = (uint32_t)(ki[j] >> shd);
    if (lev[d].key.empty() || lev[d].key.back() != kk) lev[d].key.push_back(kk);
  }
  for (int l = d - 1; l >= 0; l--)
    for (uint32_t kk : lev[l + 1].key) {
      const uint32_t pk = kk >> 3;
      if (lev[l].key.empty() || lev[l].key.back() != pk) lev[l].key.push_back(pk);
    }
  int run = 0;
  for (int l = 0; l <= d; l++) {
    lev[l].first = run;
    run += (int)lev[l].key.size();
  }
  const int nc = run;
  I.ncell = nc;
  I.center.resize(nc);
  I.celllev.resize(nc);
  I.child.assign((size_t)nc * 8, -1);
  I.cls.assign(nc, 0);
  // centres by halving from the root (as build_tree), child table
  I.center[0] = real4{0, 0, 0, 0};
  I.celllev[0] = 0;
  for (int l = 1; l <= d; l++) {
    const real h = celeng[l] / 2;
    const Level &up = lev[l - 1];
    size_t pj = 0;  // parents are visited in key order
    for (size_t j = 0; j < lev[l].key.size(); j++) {
      const uint32_t kk = lev[l].key[j];
      while (up.key[pj] != (kk >> 3)) pj++;
      const int A = lev[l].first + (int)j;
      const int Pc = up.first + (int)pj;
      const int s = (int)(kk & 7);
      const real4 pc = I.center[Pc];
      I.center[A] = real4{(s & 4) ? pc.x + h : pc.x - h, (s & 2) ? pc.y + h : pc.y - h,
                          (s & 1) ? pc.z + h : pc.z - h, 0};
      I.celllev[A] = l;
      I.cls[A] = s;
      I.child[(size_t)8 * Pc + s] = A;
    }
  }
  I.plev.resize(d);
  for (int l = 0; l < d; l++)
    for (size_t j = 0; j < lev[l].key.size(); j++) I.plev[l].push_back(lev[l].first + (int)j);
  // leaves and particles (leaf order)
  const int nl = (int)lev[d].key.size();
  I.nleaf = nl;
  I.perm.resize(n);
  I.leafof.resize(n);
  I.leafcell.resize(nl);
  I.leafsta.resize(nl);
  I.leafcnt.resize(nl);
  {
    int pj = 0;
    for (int q = 0; q < nl; q++) {
      const uint32_t kk = lev[d].key[q];
      const int sta = pj;
      while (pj < n && (uint32_t)(ki[pj] >> shd) == kk) {
        I.perm[pj] = (int)(uint32_t)ki[pj];
        I.leafof[pj] = q;
        pj++;
      }
      I.leafcell[q] = lev[d].first + q;
      I.leafsta[q] = sta;
      I.leafcnt[q] = pj - sta;
    }
  }
  I.lev = std::move(lev);
}

// Lists of instance I, written straight into the global tables of f (its
// cells from c0, its leaves from lf0).  Level by level from the root, every
// cell's 27 neighbours (same level, -1 where absent; offset
// e = 9 (dx + 1) + 3 (dy + 1) + dz + 1) follow from its parent's: the
// neighbour at offset e of child s of P is child (s + e) mod 2 of P's
// neighbour floor((s + e) / 2) (per axis).  For a child f = 2p + s of P, the
// 27 neighbours ("uncles") of P give its interaction list: the child c =
// 2(p + u) + s' of uncle u is at offset v = 2u + s' - s, and the far offsets
// (|v|_inf >= 2) are stored by class slot, every slot written once (the zero
// cell where empty); at the deepest level the near ones are the leaf's P2P
// neighbours, in the order of the leaf's own offsets.
void build_lists(Inst &I, HostForest &f, int c0, int lf0, const ListTabs &tabs)
{
  const std::vector<Level> &lev = I.lev;
  const int d = I.d;
  const int zero = f.ncell;
  std::vector<int> nb_up(27, -1), nb_cur;
  nb_up[13] = 0;  // the root is its own neighbour at offset 0
  std::fill(f.vsrc.begin() + (size_t)kM2LOffsetsPerClass * c0, f.vsrc.begin() + (size_t)kM2LOffsetsPerClass * (c0 + 1),
            zero);
  for (int l = 1; l <= d; l++) {
    const Level &up = lev[l - 1];
    const bool more = l < d;  // the cells of level l are parents of the next level
    if (more) nb_cur.assign((size_t)27 * lev[l].key.size(), -1);
    for (size_t j = 0; j < up.key.size(); j++) {
      const int Pc = up.first + (int)j;
      const int *unc = &nb_up[27 * j];
      for (int sf = 0; sf < 8; sf++) {
        const int F = I.child[(size_t)8 * Pc + sf];
        if (F < 0) continue;
        const int q = F - lev[l].first;
        int nbF[27];
        for (int e = 0; e < 27; e++) {
          const int2 uc = tabs.near_uc[sf * 27 + e];
          const int U = unc[uc.x];
          nbF[e] = U < 0 ? -1 : I.child[(size_t)8 * U + uc.y];
        }
        if (more) std::copy(nbF, nbF + 27, &nb_cur[(size_t)27 * q]);
        int *vs = &f.vsrc[(size_t)kM2LOffsetsPerClass * (c0 + F)];
        if (l < kMinLev) {  // no list above level 2: all slots empty
          std::fill(vs, vs + kM2LOffsetsPerClass, zero);
          continue;
        }
        const int2 *suc = &tabs.slot_uc[(size_t)sf * kM2LOffsetsPerClass];
        int m = 0;
        for (int k = 0; k < kM2LOffsetsPerClass; k++) {
          const int U = unc[suc[k].x];
          const int C = U < 0 ? -1 : I.child[(size_t)8 * U + suc[k].y];
          vs[k] = C < 0 ? zero : c0 + C;
          m += C >= 0;
        }
        f.iinter[c0 + F] = m;
        I.ninter += m;
        if (l == d) {  // F is a leaf: its P2P neighbours (self included)
          int nm = 0;
          for (int e = 0; e < 27; e++) {
            const int B = nbF[e];
            if (B < 0) continue;
            const int bl = B - lev[d].first;
            f.nbr[(size_t)(lf0 + q) * 27 + nm++] = lf0 + bl;
            I.pairs += (double)I.leafcnt[q] * (double)I.leafcnt[bl];
          }
          f.nnbr[lf0 + q] = nm;
        }
      }
    }
    if (more) nb_up.swap(nb_cur);
  }
}


// f(i) for i in [0, n) on up to `threads` threads, items taken in order
template <class Fn>
void parallel_items(int n, Fn f)
{
  int nt = (int)std::thread::hardware_concurrency();
  if (const char *e = getenv("BBFMM_AMD_SETUP_THREADS")) nt = atoi(e);
  nt = std::max(1, std::min(nt, std::min(n, 64)));
  if (nt == 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++)
    pool.emplace_back([&]() {
      for (int i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &th : pool) th.join();
}
}  // namespace

int forest_from_points(HostForest &f, int ninst, const int *Ns, const real3 *pos, const int *depth, int leafmax,
                       real L, std::vector<int> *depth_out)
{
  f = HostForest();
  f.complete = false;
  if (ninst < 1) return set_error(BBFMM_E_ARG, "no instances");
  const ListTabs tabs;  // (also builds class_slot's table before the threads run)
  long long ptot = 0;
  for (int i = 0; i < ninst; i++) {
    if (Ns[i] < 1) return set_error(BBFMM_E_ARG, "instance %d has %d points", i, Ns[i]);
    if (depth && depth[i] > 0 && (depth[i] < kMinLev || depth[i] > kMaxForestDepth))
      return set_error(BBFMM_E_ARG, "instance %d: depth %d outside [%d, %d]", i, depth[i], kMinLev, kMaxForestDepth);
    ptot += Ns[i];
  }
  if (ptot > (1LL << 30)) return set_error(BBFMM_E_ARG, "too many points (%lld)", ptot);
  if (leafmax < 1) leafmax = 32;
  f.N = (int)ptot;
  std::vector<real> celeng(kMaxForestDepth + 1);
  {
    std::vector<int> ls(kMaxForestDepth + 1), le(kMaxForestDepth + 1);
    level_tables(kMaxForestDepth, L, ls.data(), le.data(), celeng.data());
  }
  std::vector<long long> p0s(ninst + 1, 0);
  for (int i = 0; i < ninst; i++) p0s[i + 1] = p0s[i] + Ns[i];
  // instances on their own, largest first
  std::vector<Inst> in(ninst);
  std::vector<int> byz(ninst);
  for (int i = 0; i < ninst; i++) byz[i] = i;
  std::stable_sort(byz.begin(), byz.end(), [&](int a, int b) { return Ns[a] > Ns[b]; });
  parallel_items(ninst, [&](int t) {
    const int i = byz[t];
    build_instance(in[i], pos + p0s[i], Ns[i], depth ? depth[i] : 0, leafmax, L, celeng, tabs);
  });
  for (int i = 0; i < ninst; i++)
    if (in[i].rc != BBFMM_OK) return set_error(in[i].rc, "a particle lies outside the root cell [-L/2, L/2]^3");
  // global numbering: instances one after the other
  std::vector<int> c0(ninst + 1, 0), lf0(ninst + 1, 0);
  int dmax = 0;
  for (int i = 0; i < ninst; i++) {
    c0[i + 1] = c0[i] + in[i].ncell;
    lf0[i + 1] = lf0[i] + in[i].nleaf;
    dmax = std::max(dmax, in[i].d);
    f.inst_N.push_back(Ns[i]);
    f.inst_depth.push_back(in[i].d);
    f.inst_p0.push_back((int)p0s[i]);
    f.inst_leaf0.push_back(lf0[i]);
  }
  const int ncell = c0[ninst], nleaf = lf0[ninst];
  f.ncell = ncell;
  f.nleaf = nleaf;
  f.maxdepth = dmax;
  f.celeng.assign(celeng.begin(), celeng.begin() + dmax + 1);
  f.center.resize(ncell);
  f.celllev.resize(ncell);
  f.child.resize((size_t)ncell * 8);
  f.perm.resize(f.N);
  f.leafof.resize(f.N);
  f.leafcell.resize(nleaf);
  f.leafsta.resize(nleaf);
  f.leafcnt.resize(nleaf);
  f.leafihalf.resize(nleaf);
  f.nbr.assign((size_t)nleaf * 27, -1);
  f.nnbr.assign(nleaf, 0);
  f.iinter.assign(ncell, 0);
  f.vsrc.resize((size_t)ncell * kM2LOffsetsPerClass);  // filled by build_lists
  f.cellcls.resize(ncell);
  std::vector<int> &cellclass = f.cellcls;
  parallel_items(ninst, [&](int i) {
    Inst &I = in[i];
    const int c = c0[i], lf = lf0[i];
    const int p = (int)p0s[i];
    auto shift = [](int v, int o) { return v < 0 ? v : v + o; };
    for (int A = 0; A < I.ncell; A++) {
      f.center[c + A] = I.center[A];
      f.celllev[c + A] = I.celllev[A];
      cellclass[c + A] = I.cls[A];
    }
    for (size_t k = 0; k < I.child.size(); k++) f.child[(size_t)8 * c + k] = shift(I.child[k], c);
    for (int j = 0; j < I.n; j++) {
      f.perm[p + j] = p + I.perm[j];
      f.leafof[p + j] = lf + I.leafof[j];
    }
    const real ihalf = (real)1 / ((L / (real)(1 << I.d)) / (real)2);
    for (int q = 0; q < I.nleaf; q++) {
      f.leafcell[lf + q] = c + I.leafcell[q];
      f.leafsta[lf + q] = p + I.leafsta[q];
      f.leafcnt[lf + q] = I.leafcnt[q];
      f.leafihalf[lf + q] = ihalf;
    }
    build_lists(I, f, c, lf, tabs);
  });
  for (int i = 0; i < ninst; i++) {
    f.pairs += in[i].pairs;
    f.ninter += in[i].ninter;
  }
  // parents per level across instances
  f.ploff.assign(dmax + 1, 0);
  for (int l = 0; l < dmax; l++) {
    f.ploff[l] = (int)f.plist.size();
    for (int i = 0; i < ninst; i++)
      if (l < in[i].d)
        for (int A : in[i].plev[l]) f.plist.push_back(A + c0[i]);
  }
  f.ploff[dmax] = (int)f.plist.size();
  // class lists for the list MFMA M2L
  f.cls_off.assign(9, 0);
  for (int A = 0; A < ncell; A++)
    if (f.celllev[A] >= kMinLev) f.cls_off[cellclass[A] + 1]++;
  for (int c = 0; c < 8; c++) f.cls_off[c + 1] += f.cls_off[c];
  f.cls_cells.assign(f.cls_off[8], 0);
  {
    std::vector<int> fill(f.cls_off.begin(), f.cls_off.end() - 1);
    for (int A = 0; A < ncell; A++)
      if (f.celllev[A] >= kMinLev) f.cls_cells[fill[cellclass[A]]++] = A;
  }
  if (depth_out) {
    depth_out->resize(ninst);
    for (int i = 0; i < ninst; i++) (*depth_out)[i] = in[i].d;
  }
  return BBFMM_OK;
}

}  // namespace bb
