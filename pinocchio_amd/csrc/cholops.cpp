// cholops.cpp -- products and triangular solves with the packed tree factor M = U D U^T of
// rbd_cholesky_decompose: cholesky::Uv, Utv, Uiv, Utiv and Mv (cholesky.hpp:86-150,
// cholesky.hxx:64-465), in place on x, one thread per configuration, the tree's pattern
// read at run time (column c's strictly-upper entries are [colPtr[c], colPtr[c+1]) of the
// packed list, their rows in patRow).  FP64.
#include <hip/hip_runtime.h>

#include "../../include/rbd.h"
#include "generic.hpp"
#include "kernels_common.hpp"

namespace rbd {

namespace {
// x <- U x: x_r += U_rc x_c over each column c in increasing order (x_c is not yet changed:
// only columns after c write it)
__device__ void uv(int nv, const int32_t* cp, const int32_t* pr, const double* U, int64_t us, double* x, int64_t xs) {
  for (int c = 0; c < nv; ++c) {
    const double xc = x[c * xs];
    for (int k = cp[c]; k < cp[c + 1]; ++k) x[pr[k] * xs] += U[k * us] * xc;
  }
}
// x <- U^T x: x_c += sum_r U_rc x_r, columns in decreasing order (its rows r < c still hold
// the input)
__device__ void utv(int nv, const int32_t* cp, const int32_t* pr, const double* U, int64_t us, double* x, int64_t xs) {
  for (int c = nv - 1; c >= 0; --c) {
    double s = x[c * xs];
    for (int k = cp[c]; k < cp[c + 1]; ++k) s += U[k * us] * x[pr[k] * xs];
    x[c * xs] = s;
  }
}
// x <- U^-1 x: columns in decreasing order; x_c is final once the later columns are done
__device__ void uiv(int nv, const int32_t* cp, const int32_t* pr, const double* U, int64_t us, double* x, int64_t xs) {
  for (int c = nv - 1; c >= 0; --c) {
    const double xc = x[c * xs];
    for (int k = cp[c]; k < cp[c + 1]; ++k) x[pr[k] * xs] -= U[k * us] * xc;
  }
}
// x <- U^-T x: columns in increasing order, each one a dot product with final rows r < c
__device__ void utiv(int nv, const int32_t* cp, const int32_t* pr, const double* U, int64_t us, double* x, int64_t xs) {
  for (int c = 0; c < nv; ++c) {
    double s = x[c * xs];
    for (int k = cp[c]; k < cp[c + 1]; ++k) s -= U[k * us] * x[pr[k] * xs];
    x[c * xs] = s;
  }
}
}  // namespace

__global__ __launch_bounds__(256) void k_chol_apply(int64_t B, int nv, const int32_t* __restrict__ cp,
                                                    const int32_t* __restrict__ pr, int op, CView<double> gU,
                                                    CView<double> gD, View<double> gx) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double* U = gU.p + b * gU.bs;
  double* x = gx.p + b * gx.bs;
  switch (op) {
    case RBD_CHOL_UV: uv(nv, cp, pr, U, gU.rs, x, gx.rs); break;
    case RBD_CHOL_UTV: utv(nv, cp, pr, U, gU.rs, x, gx.rs); break;
    case RBD_CHOL_UIV: uiv(nv, cp, pr, U, gU.rs, x, gx.rs); break;
    case RBD_CHOL_UTIV: utiv(nv, cp, pr, U, gU.rs, x, gx.rs); break;
    case RBD_CHOL_MV:  // U D U^T x
      utv(nv, cp, pr, U, gU.rs, x, gx.rs);
      for (int c = 0; c < nv; ++c) x[c * gx.rs] *= gD.p[c * gD.rs + b * gD.bs];
      uv(nv, cp, pr, U, gU.rs, x, gx.rs);
      break;
  }
}

void launchCholApply(hipStream_t st, const GenericModel* g, int op, int64_t B, CView<double> U, CView<double> D,
                     View<double> x) {
  hipLaunchKernelGGL(k_chol_apply, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, B, g->nv, g->colPtr,
                     g->patRow, op, U, D, x);
}

}  // namespace rbd
