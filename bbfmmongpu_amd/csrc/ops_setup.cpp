// ops_setup.cpp -- FP64 operator precompute (ComputeKernelSVD, bbfmm.c:1209-1421).
//
// K0 (n^3 x 316 n^3) samples the kernel between the Chebyshev nodes of a
// unit field cell at the origin and of the 316 unit source cells at offsets
// in [-3,3]^3 outside [-1,1]^3, divided by the node weights of both ends.
// U = the r leading left singular vectors of K0, obtained here as the
// eigenvectors of the Gram matrix K0 K0^T (cyclic Jacobi rotations with a
// threshold sweep), and E_k = U^T K0_k U.  Singular vectors are signed so
// that their largest entry is positive.
#include "common.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace bb {

static void sample_K0(int n, std::vector<double> &K0)
{
  const int n3 = n * n * n;
  std::vector<double> x(n), w(n3);
  for (int m = 0; m < n; m++) x[m] = std::cos(M_PI * (2.0 * m + 1) / (2.0 * n));
  for (int a = 0; a < n3; a++) {
    const int i = a / (n * n), j = (a / n) % n, k = a % n;
    w[a] = 1.0 / (std::sqrt(1 - x[i] * x[i]) * std::sqrt(1 - x[j] * x[j]) * std::sqrt(1 - x[k] * x[k]));
  }
  K0.assign((size_t)kNumOps * n3 * n3, 0.0);
  size_t col = 0;
  for (int ox = -3; ox <= 3; ox++)
    for (int oy = -3; oy <= 3; oy++)
      for (int oz = -3; oz <= 3; oz++) {
        if (std::abs(ox) <= 1 && std::abs(oy) <= 1 && std::abs(oz) <= 1) continue;
        for (int s = 0; s < n3; s++, col++) {
          const double sx = ox + 0.5 * x[s / (n * n)], sy = oy + 0.5 * x[(s / n) % n], sz = oz + 0.5 * x[s % n];
          double *Kc = &K0[col * n3];
          for (int f = 0; f < n3; f++) {
            const double dx = sx - 0.5 * x[f / (n * n)], dy = sy - 0.5 * x[(f / n) % n], dz = sz - 0.5 * x[f % n];
            Kc[f] = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz) / (w[s] * w[f]);
          }
        }
      }
}

// symmetric eigenproblem, cyclic Jacobi with Rutishauser's threshold strategy
void jacobi_eig(int m, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V)
{
  V.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++) V[(size_t)i * m + i] = 1;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0, tot = 0;
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) {
        const double v = A[(size_t)j * m + i] * A[(size_t)j * m + i];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-32 * tot) break;
    const double thresh = sweep < 3 ? 0.2 * std::sqrt(off) / ((double)m * m) : 0.0;
    for (int pp = 0; pp < m - 1; pp++)
      for (int qq = pp + 1; qq < m; qq++) {
        const double apq = A[(size_t)qq * m + pp];
        if (std::fabs(apq) <= thresh || apq == 0.0) continue;
        const double theta = (A[(size_t)qq * m + qq] - A[(size_t)pp * m + pp]) / (2 * apq);
        const double tt = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(tt * tt + 1), s = tt * c;
        double *cp = &A[(size_t)pp * m], *cq = &A[(size_t)qq * m];
        for (int k = 0; k < m; k++) {
          const double a = cp[k], b = cq[k];
          cp[k] = c * a - s * b;
          cq[k] = s * a + c * b;
        }
        for (int k = 0; k < m; k++) {
          double *row = &A[(size_t)k * m];
          const double a = row[pp], b = row[qq];
          row[pp] = c * a - s * b;
          row[qq] = s * a + c * b;
        }
        double *vp = &V[(size_t)pp * m], *vq = &V[(size_t)qq * m];
        for (int k = 0; k < m; k++) {
          const double a = vp[k], b = vq[k];
          vp[k] = c * a - s * b;
          vq[k] = s * a + c * b;
        }
      }
  }
  std::vector<int> order(m);
  for (int i = 0; i < m; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return A[(size_t)a * m + a] > A[(size_t)b * m + b]; });
  std::vector<double> Vs((size_t)m * m);
  w.resize(m);
  for (int c = 0; c < m; c++) {
    const int src = order[c];
    w[c] = A[(size_t)src * m + src];
    int big = 0;
    for (int k = 1; k < m; k++)
      if (std::fabs(V[(size_t)src * m + k]) > std::fabs(V[(size_t)src * m + big])) big = k;
    const double sg = V[(size_t)src * m + big] < 0 ? -1.0 : 1.0;
    for (int k = 0; k < m; k++) Vs[(size_t)c * m + k] = sg * V[(size_t)src * m + k];
  }
  V.swap(Vs);
}

int operators_fp64(int n, int r, double *U, double *K, double *sigma)
{
  const int n3 = n * n * n;
  std::vector<double> K0;
  sample_K0(n, K0);
  std::vector<double> G((size_t)n3 * n3, 0.0);
  const size_t ncol = (size_t)kNumOps * n3;
  for (size_t c = 0; c < ncol; c++) {
    const double *kc = &K0[c * n3];
    for (int j = 0; j < n3; j++) {
      const double kj = kc[j];
      double *g = &G[(size_t)j * n3];
      for (int i = 0; i <= j; i++) g[i] += kc[i] * kj;
    }
  }
  for (int j = 0; j < n3; j++)
    for (int i = 0; i < j; i++) G[(size_t)i * n3 + j] = G[(size_t)j * n3 + i];
  std::vector<double> w, V;
  jacobi_eig(n3, G, w, V);
  for (int i = 0; i < n3; i++) sigma[i] = std::sqrt(std::max(w[i], 0.0));
  for (int c = 0; c < r; c++)
    for (int i = 0; i < n3; i++) U[(size_t)c * n3 + i] = V[(size_t)c * n3 + i];
  std::vector<double> KU((size_t)n3 * r);
  for (int k = 0; k < kNumOps; k++) {
    const double *Kk = &K0[(size_t)k * n3 * n3];
    std::fill(KU.begin(), KU.end(), 0.0);
    for (int c = 0; c < r; c++)
      for (int j = 0; j < n3; j++) {
        const double u = U[(size_t)c * n3 + j];
        const double *kj = &Kk[(size_t)j * n3];
        double *o = &KU[(size_t)c * n3];
        for (int i = 0; i < n3; i++) o[i] += kj[i] * u;
      }
    double *E = &K[(size_t)k * r * r];
    for (int c = 0; c < r; c++)
      for (int a = 0; a < r; a++) {
        double s = 0;
        for (int i = 0; i < n3; i++) s += U[(size_t)a * n3 + i] * KU[(size_t)c * n3 + i];
        E[(size_t)c * r + a] = s;
      }
  }
  return BBFMM_OK;
}

} // namespace bb
