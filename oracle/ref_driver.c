/*
 * ref_driver.c -- runs the reference's own CPU bbFMM path (CPU8A build,
 * compiled from /root/reference by oracle/Makefile.ref) on given inputs and
 * dumps every intermediate, for pinning the oracle and as the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY (checker side).  Operators are injected from
 * files (the reference reads K/U from files as well, bbfmm.c:216-254), so no
 * LAPACK routine of the reference is reached.
 *
 * usage: ref_fmm DIR N n lmod [dump [reps]]
 *   DIR/pos.bin  N*3 real     DIR/q.bin  N real
 *   DIR/K.bin    316*r*r real (reference K-file format)
 *   DIR/U.bin    int r, n^3*r real (reference U-file format)
 * writes DIR/phi.bin (q*phi, the reference's output convention) and, with
 * dump=1, the tree, lists, M (proxysval), F (fieldval) and the far-field phi.
 * Prints one line of timings (seconds) to stdout.  With reps > 1 the solve
 * (upward pass .. near field, on the tree built once) runs reps times; the
 * first run is a warm-up, t_solve is the median of the others and t_solve_all
 * lists every run.  phi.bin holds the last run's output.
 */
#include "bbfmm.h"
#include <malloc.h>
#include <string.h>
#include <time.h>

static double now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void *read_file(const char *dir, const char *name, size_t bytes)
{
  char path[1024];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  void *p = malloc(bytes ? bytes : 1);
  if (fread(p, 1, bytes, f) != bytes) { fprintf(stderr, "short read %s\n", path); exit(2); }
  fclose(f);
  return p;
}

static void write_file(const char *dir, const char *name, const void *p, size_t bytes)
{
  char path[1024];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "wb");
  if (!f) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
  fwrite(p, 1, bytes, f);
  fclose(f);
}

int main(int argc, char **argv)
{
  if (argc < 5) { fprintf(stderr, "usage: %s DIR N n lmod [dump [reps]]\n", argv[0]); return 2; }
  const char *dir = argv[1];
  const int N = atoi(argv[2]);
  const int n = atoi(argv[3]);
  int l = atoi(argv[4]);
  const int dump = argc > 5 ? atoi(argv[5]) : 0;
  int reps = argc > 6 ? atoi(argv[6]) : 1;
  if (reps < 1) reps = 1;
  if (reps > 64) reps = 64;
  const int dof = 1, n3 = n * n * n;
  const real L = 1;

  /* Level rule of FMMSetup (bbfmm.c:151-160) */
  l += (int)log(N) / log(8.0);
  if (l > 0) l = l - 1;
  if (l < MINLEV) l = MINLEV;
  const int ncell = (POW8(l + 1) - 1) / 7;

  /* The reference reads uninitialised sta/end entries and accumulates into
     un-zeroed proxysval (anotherUpwardPass.c:50,388), so its results depend
     on what malloc returns.  glibc's M_PERTURB option fills every block that
     malloc returns with the complement of the given byte: with 0xff every
     block starts as zeros, as on the fresh pages of the runs it was written
     for, whatever the heap held before. */
  mallopt(M_PERTURB, 0xff);

  real3 *field = read_file(dir, "pos.bin", sizeof(real3) * N);
  real *q = read_file(dir, "q.bin", sizeof(real) * N);
  int cutoff;
  {
    char path[1024];
    snprintf(path, sizeof(path), "%s/U.bin", dir);
    FILE *f = fopen(path, "rb");
    if (!f || fread(&cutoff, sizeof(int), 1, f) != 1) { fprintf(stderr, "bad U\n"); return 2; }
    fclose(f);
  }
  char Kpath[1024], Upath[1024];
  snprintf(Kpath, sizeof(Kpath), "%s/K.bin", dir);
  snprintf(Upath, sizeof(Upath), "%s/U.bin", dir);

  real *Kweights = malloc(n3 * sizeof(real));
  real *Cweights = malloc(2 * n * n * sizeof(real));
  real *Tkz = malloc(n * n * sizeof(real));
  int Ktable[343];
  real homogen = 1;
  ComputeWeights(Tkz, Ktable, Kweights, Cweights, n);
  real *K = malloc(316 * cutoff * cutoff * sizeof(real));
  real *U = malloc(cutoff * n3 * sizeof(real));
  FMMReadMatrices(K, U, cutoff, n, dof, Kpath, Upath);
  real *phi = calloc(N, sizeof(real));

  double t0 = now();
  anotherTree *atree;
  anotherBuildFMMHierachy(L, n, dof, l, cutoff, &atree, N, N, field, field);
  double t1 = now();
  cell *c = atree->c;
  if (dump) {
    write_file(dir, "ref_center.bin", c->center, sizeof(real3) * ncell);
    write_file(dir, "ref_list.bin", atree->fieldlist, sizeof(int) * N);
    const int Asta = atree->levsta[l], nl = atree->levend[l] - Asta + 1;
    write_file(dir, "ref_sta.bin", c->fieldsta + Asta, sizeof(int) * nl);
    write_file(dir, "ref_end.bin", c->fieldend + Asta, sizeof(int) * nl);
    write_file(dir, "ref_ineigh.bin", c->ineigh, sizeof(int) * ncell);
    write_file(dir, "ref_iinter.bin", c->iinter, sizeof(int) * ncell);
    write_file(dir, "ref_neighbors.bin", c->neighbors, sizeof(int) * ncell * c->pitch_neighbors);
    write_file(dir, "ref_interaction.bin", c->interaction, sizeof(int) * ncell * c->pitch_interaction);
  }
  double tup = 0, tint = 0, tdown = 0, tnear = 0, tsolve[64];
  for (int rep = 0; rep < reps; rep++) {
    const int last = rep == reps - 1;
    memset(phi, 0, sizeof(real) * N);
    const double s1 = now();
    anotherUpwardPass(&atree, field, Cweights, Tkz, q, U, Kweights, cutoff, n, dof);
    const double s2 = now();
    if (dump && last) write_file(dir, "ref_M.bin", atree->proxysval, sizeof(real) * ncell * cutoff);
    anotherFMMInteraction(&atree, K, Ktable, U, Kweights, n, dof, cutoff, homogen);
    const double s3 = now();
    if (dump && last) write_file(dir, "ref_F.bin", atree->fieldval, sizeof(real) * ncell * n3);
    anotherDownwardPassX(&atree, field, Cweights, Tkz, q, U, cutoff, n, dof, homogen, phi);
    const double s4 = now();
    if (dump && last) write_file(dir, "ref_phifar.bin", phi, sizeof(real) * N);
    anotherNearField(&atree, dof, homogen, field, field, phi, q);
    const double s5 = now();
    tsolve[rep] = s5 - s1;
    if (last) {
      tup = s2 - s1;
      tint = s3 - s2;
      tdown = s4 - s3;
      tnear = s5 - s4;
    }
  }
  write_file(dir, "phi.bin", phi, sizeof(real) * N);
  anotherFMMCleanup(&atree);
  /* median of the runs after the warm-up (all runs when reps == 1) */
  const int first = reps > 1 ? 1 : 0, cnt = reps - first;
  double sorted[64];
  for (int k = 0; k < cnt; k++) sorted[k] = tsolve[first + k];
  for (int a = 1; a < cnt; a++)
    for (int b = a; b > 0 && sorted[b - 1] > sorted[b]; b--) {
      const double x = sorted[b];
      sorted[b] = sorted[b - 1];
      sorted[b - 1] = x;
    }
  const double med = cnt % 2 ? sorted[cnt / 2] : 0.5 * (sorted[cnt / 2 - 1] + sorted[cnt / 2]);
  printf("{\"maxlev\": %d, \"cutoff\": %d, \"reps\": %d, \"t_tree\": %.6e, \"t_upward\": %.6e, "
         "\"t_interaction\": %.6e, \"t_downwardx\": %.6e, \"t_nearfield\": %.6e, "
         "\"t_solve\": %.6e, \"t_solve_all\": [",
         l, cutoff, reps, t1 - t0, tup, tint, tdown, tnear, med);
  for (int k = 0; k < reps; k++) printf("%s%.6e", k ? ", " : "", tsolve[k]);
  printf("]}\n");
  return 0;
}
