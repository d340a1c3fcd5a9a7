// common.h -- shared types, constants and error plumbing of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>

#include "bbfmm_amd.h"

#ifdef SINGLE
typedef float4 real4;
#else
typedef double4 real4;
#endif

namespace bb {

constexpr int kMinLev = 2;              // another.h:24-26
constexpr int kPitchNeighbors = 64;     // FLOORUP(27, ALIGN_SIZE=64)
constexpr int kPitchInteraction = 192;  // FLOORUP(189, 64)
constexpr int kNumOps = 316;            // offsets in [-3,3]^3 outside [-1,1]^3
constexpr int kM2LOffsetsPerClass = 189;  // M2L offsets of one sibling class
// Chebyshev orders with compiled pass kernels.  The reference reads n at run
// time (main.c:58, bbfmm.c:1340-1343 sets the rank from it); every order in
// this range is instantiated.
constexpr int kMinOrder = 3, kMaxOrder = 8;
// largest SVD rank any kernel accepts (r <= n^3 <= 512)
constexpr int kMaxRank = 512;

// api.hip: BBFMM_E_NODEVICE unless a GPU is visible
int check_device();

// host_setup.cpp
int maxlev_rule(int N, int lmod);
void level_tables(int maxlev, real L, int *levsta, int *levend, real *celeng);
int build_tree(anotherTree *t, real L, int maxlev, int N, const real3 *pos);
void free_tree(anotherTree *t);
void ktable(int *Kt);
void kindex(const anotherTree *t, const int *Kt, int *Kindex);
void weights(int n, real *Tkz, int *Kt, real *Kw, real *Cw);

// error state (api.hip)
int set_error(int code, const char *fmt, ...);
void clear_error();

} // namespace bb

#define BB_HIP(call)                                                            \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      return bb::set_error(BBFMM_E_HIP, "%s failed: %s (%s:%d)", #call,         \
                           hipGetErrorString(e_), __FILE__, __LINE__);          \
  } while (0)

#define BB_TRY(expr)                                                            \
  do {                                                                          \
    int rc_ = (expr);                                                           \
    if (rc_ != BBFMM_OK) return rc_;                                            \
  } while (0)

// Runs the statement with the compile-time Chebyshev order NN equal to the run-time
// order n (kMinOrder..kMaxOrder); other orders return BBFMM_E_ARG.
#define BB_CASE_ORDER_(v, ...) \
  case v: {                    \
    constexpr int NN = v;      \
    __VA_ARGS__;               \
  } break;
#define BB_WITH_ORDER(n, ...)                                                               \
  do {                                                                                      \
    switch (n) {                                                                            \
      BB_CASE_ORDER_(3, __VA_ARGS__)                                                               \
      BB_CASE_ORDER_(4, __VA_ARGS__)                                                               \
      BB_CASE_ORDER_(5, __VA_ARGS__)                                                               \
      BB_CASE_ORDER_(6, __VA_ARGS__)                                                               \
      BB_CASE_ORDER_(7, __VA_ARGS__)                                                               \
      BB_CASE_ORDER_(8, __VA_ARGS__)                                                               \
    default:                                                                                \
      return bb::set_error(BBFMM_E_ARG, "Chebyshev order n=%d not supported (n = %d..%d)", \
                           (int)(n), bb::kMinOrder, bb::kMaxOrder);                         \
    }                                                                                       \
  } while (0)
