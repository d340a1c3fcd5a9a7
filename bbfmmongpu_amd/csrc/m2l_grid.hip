// m2l_grid.hip -- dense-grid M2L (schemes CL and IJ).
#include "common.h"
#include "kernels.h"

namespace bb {

struct GridM2L {
  int dummy;
};

int grid_m2l_create(GridM2L **g, int r, int maxlev, int nrhs, int form, const real *K, hipStream_t s)
{
  (void)r; (void)maxlev; (void)nrhs; (void)form; (void)K; (void)s;
  *g = nullptr;
  return set_error(BBFMM_E_ARG, "dense-grid M2L is not built yet");
}

void grid_m2l_destroy(GridM2L *g) { delete g; }

int grid_m2l_run(GridM2L *g, const real *M, real *L, int ncell, hipStream_t s)
{
  (void)g; (void)M; (void)L; (void)ncell; (void)s;
  return set_error(BBFMM_E_ARG, "dense-grid M2L is not built yet");
}

} // namespace bb
