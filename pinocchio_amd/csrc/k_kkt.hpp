// k_kkt.hpp -- batched forward dynamics under linear acceleration constraints
// J ddq + gamma = 0 (forwardDynamics(model, data, q, v, tau, J, gamma, inv_damping),
// contact-dynamics.hxx:90-167):
//
//   M = crba(q), nle = rnea(q, v, 0), M = U D U^T                 (existing kernels)
//   r = M^-1 (tau - nle),  Y = M^-1 J^T                           (k_kkt_solve)
//   lambda = (J Y + damping I)^-1 (-J r - gamma)                   (k_kkt_finish)
//   ddq = r + Y lambda
//
// The reference forms J M^-1 J^T as S^T S with S = D^-1/2 U^-1 J^T and adds M^-1 J^T lambda
// by one more solve; here the nc + 1 full solves run side by side (two per wave, the FP64
// factor of a 32-configuration tile staged once in LDS), so the small dense part needs no
// further pass over the factor.  Arithmetic is FP64 from the factor on.
#pragma once

#include "k_crba.hpp"

namespace rbd {

constexpr int kKktMaxConstraints = 12;
constexpr int kKktWaves = (kKktMaxConstraints + 2) / 2;  // nc + 1 right-hand sides, two per wave
constexpr int kKktTile = 32;                              // configurations per workgroup of the solve

// LDS bytes of the solve's staged factor
template<class TP> constexpr int kktSolveLdsBytes() { return (nnzOf<TP>() + TP::T.nv) * kKktTile * 8; }

// Y[:, j] = M^-1 J[j, :]^T for j < nc and Y[:, nc] = M^-1 (tau - nle).  Lanes 0-31 of wave w
// take right-hand side 2w and lanes 32-63 right-hand side 2w + 1, both for the tile's 32
// configurations, so the two halves read the same LDS words.  J is nc x nv per
// configuration, column major (element (j, k) at row j + k nc of the view); Y is
// (nc + 1) nv FP64 rows per configuration (element (k, j) at row k + j nv).
template<class TP, class S>
__global__ __launch_bounds__(64 * kKktWaves) void k_kkt_solve(int64_t B, CView<double> gU, CView<double> gD, int nc,
                                                             CView<S> gJ, CView<S> gtau, CView<S> gnle,
                                                             View<double> gY) {
  constexpr int NV = TP::T.nv;
  constexpr int NNZ = nnzOf<TP>();
  using P = CholPattern<TP>;
  __shared__ double sU[(NNZ + NV) * kKktTile];  // U entries, then 1/D; [entry][configuration]
  const int64_t b0 = int64_t(blockIdx.x) * kKktTile;
  const int nb = (int)(B - b0 < kKktTile ? B - b0 : kKktTile);
  {
    // kStageU loads in flight per thread; configurations past the batch end repeat the
    // last one (their results are not stored)
    constexpr int kTot = (NNZ + NV) * kKktTile, kStep = 64 * kKktWaves;
    const int t = (int)threadIdx.x;
    static_for<0, (kTot + kStep * kStageU - 1) / (kStep * kStageU)>([&](auto I) {
      double x[kStageU];
      static_for<0, kStageU>([&](auto U) {
        const int e = (int(I) * kStageU + int(U)) * kStep + t;
        const int ec = e < kTot ? e : kTot - 1;
        const int k = ec / kKktTile, c = ec % kKktTile;
        const int64_t bc = b0 + (c < nb ? c : nb - 1);
        x[U] = k < NNZ ? gU.p[(int64_t)k * gU.rs + bc * gU.bs] : gD.p[(int64_t)(k - NNZ) * gD.rs + bc * gD.bs];
      });
      static_for<0, kStageU>([&](auto U) {
        const int e = (int(I) * kStageU + int(U)) * kStep + t;
        if (e < kTot) sU[e] = e / kKktTile < NNZ ? x[U] : 1.0 / x[U];
      });
    });
  }
  __syncthreads();
  // no barrier below: idle half-waves may leave
  const int L = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  const int c = L & (kKktTile - 1);
  const int j = 2 * w + (L >> 5);
  if (j > nc || c >= nb) return;
  const int64_t b = b0 + c;
  double x[NV];
  if (j < nc) {
#pragma unroll
    for (int k = 0; k < NV; ++k) x[k] = double(gJ.p[(int64_t)(j + k * nc) * gJ.rs + b * gJ.bs]);
  } else {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      x[k] = double(gtau.p[(int64_t)k * gtau.rs + b * gtau.bs]) - double(gnle.p[(int64_t)k * gnle.rs + b * gnle.bs]);
  }
  // x <- M^-1 x: Uiv, D^-1, Utiv (cholesky.hxx:193-317, 466-520).  A compiler memory
  // clobber at every elimination step keeps each factor read next to its use (otherwise
  // the reads are scheduled together and the factor is held in registers).
  const double* su = sU + c;
  static_for<0, NV>([&](auto KK) {
    constexpr int k = NV - 1 - KK;
    stepFence();
    asm volatile("" ::: "memory");
    static_for<k + 1, NV>([&](auto EE) {
      constexpr int e2 = EE;
      if constexpr (P::anc(k, e2)) {
        constexpr int pk = P::packed(k, e2);  // (forced to a compile-time constant)
        x[k] -= su[pk * kKktTile] * x[e2];
      }
    });
  });
  asm volatile("" ::: "memory");
#pragma unroll
  for (int k = 0; k < NV; ++k) x[k] *= su[(NNZ + k) * kKktTile];
  static_for<0, NV>([&](auto KK) {
    constexpr int k = KK;
    stepFence();
    asm volatile("" ::: "memory");
    static_for<k + 1, NV>([&](auto EE) {
      constexpr int e2 = EE;
      if constexpr (P::anc(k, e2)) {
        constexpr int pk = P::packed(k, e2);
        x[e2] -= su[pk * kKktTile] * x[k];
      }
    });
  });
  double* py = gY.p + b * gY.bs + (int64_t)(j * NV) * gY.rs;
#pragma unroll
  for (int k = 0; k < NV; ++k) stOut(&py[(int64_t)k * gY.rs], x[k]);
}

// lambda and ddq from Y; one lane per configuration, the packed lower triangle of the
// nc x nc Delassus matrix J M^-1 J^T and lambda in LDS (run-time nc).  Its Cholesky
// factorisation has no pivoting and no rank check, like the reference's LLT: a rank-deficient
// J with zero damping gives non-finite results.
template<class TP, class S>
__global__ __launch_bounds__(64) void k_kkt_finish(int64_t B, int nc, CView<S> gJ, CView<S> gg, double damping,
                                                  CView<double> gY, View<S> gddq, View<S> glam) {
  constexpr int NV = TP::T.nv;
  constexpr int kG = kKktMaxConstraints * (kKktMaxConstraints + 1) / 2;
  __shared__ double sG[kG * kTile];
  __shared__ double sl[kKktMaxConstraints * kTile];
  const int L = laneId();
  const int64_t b = int64_t(blockIdx.x) * kTile + L;
  if (b >= B) return;  // no barrier below
  double* G = sG + L;
  double* lam = sl + L;
  auto gi = [](int r, int c) { return (r * (r + 1) / 2 + c) * kTile; };  // (r, c), c <= r
  const double* py = gY.p + b * gY.bs;
  auto yv = [&](int j, int k) { return py[(int64_t)(k + j * NV) * gY.rs]; };
  double r[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) r[k] = yv(nc, k);
#pragma unroll 1
  for (int i = 0; i < nc; ++i) {
    double Ji[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) Ji[k] = double(gJ.p[(int64_t)(i + k * nc) * gJ.rs + b * gJ.bs]);
    double s = 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) s += Ji[k] * r[k];
    lam[i * kTile] = -s - double(gg.p[(int64_t)i * gg.rs + b * gg.bs]);
#pragma unroll 1
    for (int d = 0; d <= i; ++d) {
      double g = 0;
#pragma unroll
      for (int k = 0; k < NV; ++k) g += Ji[k] * yv(d, k);
      G[gi(i, d)] = d == i ? g + damping : g;
    }
  }
  // G = L L^T, then L L^T lambda = rhs
#pragma unroll 1
  for (int i = 0; i < nc; ++i) {
#pragma unroll 1
    for (int d = 0; d <= i; ++d) {
      double s = G[gi(i, d)];
#pragma unroll 1
      for (int k = 0; k < d; ++k) s -= G[gi(i, k)] * G[gi(d, k)];
      G[gi(i, d)] = d == i ? sqrt(s) : s / G[gi(d, d)];
    }
  }
#pragma unroll 1
  for (int i = 0; i < nc; ++i) {
    double s = lam[i * kTile];
#pragma unroll 1
    for (int k = 0; k < i; ++k) s -= G[gi(i, k)] * lam[k * kTile];
    lam[i * kTile] = s / G[gi(i, i)];
  }
#pragma unroll 1
  for (int i = nc - 1; i >= 0; --i) {
    double s = lam[i * kTile];
#pragma unroll 1
    for (int k = i + 1; k < nc; ++k) s -= G[gi(k, i)] * lam[k * kTile];
    lam[i * kTile] = s / G[gi(i, i)];
  }
  // ddq = r + Y lambda
#pragma unroll 1
  for (int i = 0; i < nc; ++i) {
    const double li = lam[i * kTile];
#pragma unroll
    for (int k = 0; k < NV; ++k) r[k] += li * yv(i, k);
    stOut(&glam.p[(int64_t)i * glam.rs + b * glam.bs], S(li));
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) stOut(&gddq.p[(int64_t)k * gddq.rs + b * gddq.bs], S(r[k]));
}

}  // namespace rbd
