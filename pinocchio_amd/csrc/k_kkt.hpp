// k_kkt.hpp -- batched forward dynamics under linear acceleration constraints
// J ddq + gamma = 0 (forwardDynamics(model, data, q, v, tau, J, gamma, inv_damping),
// contact-dynamics.hxx:90-167):
//
//   M = crba(q), nle = rnea(q, v, 0), M = U D U^T                 (existing kernels)
//   r = M^-1 (tau - nle),  Y = M^-1 J^T                           (k_kkt, phase 2)
//   G = J Y + damping I,  lambda = G^-1 (-J r - gamma)             (phases 3-4)
//   ddq = r + Y lambda                                             (phase 5)
//
// The reference forms J M^-1 J^T as S^T S with S = D^-1/2 U^-1 J^T and adds M^-1 J^T lambda
// by one more solve.  Here one workgroup owns a tile of T configurations: the FP64 factor
// and J of the tile are staged in LDS once, the nc + 1 solves run side by side (T lanes
// per right-hand side, so the lanes of different right-hand sides read the same LDS
// words), each solve applies J to its own result, and the small dense part and the final
// combination go through LDS, so nothing but the inputs and ddq, lambda touches HBM.
// Arithmetic is FP64 from the factor on; Y is kept in LDS in the input precision (FP32 for
// FP32 inputs, whose outputs are FP32 too).
#pragma once

#include "dpp.hpp"
#include "k_crba.hpp"

namespace rbd {

constexpr int kKktMaxConstraints = 12;

// T configurations per tile: small enough that two (FP32) or three (FP64) workgroups fit a
// CU's LDS, so one tile's staging overlaps another's arithmetic
template<class S> struct KktShape {
  static constexpr int T = sizeof(S) == 4 ? 16 : 8;
  static constexpr int RPW = 64 / T;  // right-hand sides per wave
  static constexpr int W = (kKktMaxConstraints + 1 + RPW - 1) / RPW;
  static constexpr int kU = 32;       // staging loads in flight per thread
};

// LDS bytes of the kernel: the factor (U entries then 1/D, FP64) and J (S) of a tile; after
// the solves the space holds Y = M^-1 [J^T, tau - nle] (S), J [Y r] and lambda (FP64)
template<class TP, class S> constexpr int kktLdsBytes() {
  constexpr int T = KktShape<S>::T, NV = TP::T.nv, NC = kKktMaxConstraints;
  const int staged = (nnzOf<TP>() + NV) * T * 8 + NC * NV * T * (int)sizeof(S);
  const int later = (NC + 1) * NV * T * (int)sizeof(S) + NC * (NC + 1) * T * 8 + NC * T * 8;
  return staged > later ? staged : later;
}

// Proximal iterations of constraintDynamics (constrained-dynamics.hxx:450-486) on top of
// the damped solve: with G = J M^-1 J^T + mu I (mu = the damping), iteration k solves
// lambda_k = G^-1 (b + mu lambda_{k-1}) (lambda_0 = 0, b = -J r - gamma) and stops after
// maxIter iterations or when |gamma - (J M^-1 J^T) lambda_k|_inf <= absAcc or
// |lambda_k - lambda_{k-1}|_inf <= relAcc -- the reference's absolute and relative
// residuals, written for lambda = -x.  maxIter = 1 is the plain damped solve.
struct KktProx {
  int maxIter;
  double absAcc, relAcc;
  int* iters;  // per configuration (may be null)
};

// J is nc x nv per configuration, column major (element (i, k) at row i + k nc of gJ).
// mode 0: forwardDynamics / constraintDynamics.
// mode 1: impulseDynamics (contact-dynamics.hxx:288-331) -- gtau holds v_before, which is
// taken as r unsolved, the right-hand side is -(1 + rcoef) J v_before (gnle and gg unused),
// and gddq receives dq_after = v_before + M^-1 J^T lambda.
// mode 2: the Delassus matrix of the constraints (computeDelassusMatrix, delassus.hxx:
// 118-300, with damping = its mu): J (M + mu J^T J)^-1 J^T = (I + mu D0)^-1 D0 with
// D0 = J M^-1 J^T, written to glam (element (i, j) at row i + j nc); gtau, gnle, gg and
// gddq unused.
// mode 3: a solve with the KKT matrix [[-mu I, J], [J^T, M]] (the matrix of the contact
// Cholesky, contact-cholesky.hpp:376-400, mu = damping): gg holds the constraint rows b1
// of the right-hand side and gtau its joint rows b2 (gnle zero); glam receives x1 =
// (D0 + mu I)^-1 (J M^-1 b2 - b1) and gddq x2 = M^-1 (b2 - J^T x1) -- mode 0 with
// gamma = -b1 and lambda = -x1.
// mode 4: the damped Delassus inverse (computeDampedDelassusMatrixInverse, delassus.hxx:
// 800-836): (D0 + mu I)^-1, times rcoef (1, or mu for the reference's `scaled`), into glam
// as in mode 2.
template<class TP, class S>
__global__ __launch_bounds__(64 * KktShape<S>::W) void k_kkt(int64_t B, CView<double> gU, CView<double> gD, int nc,
                                                            CView<S> gJ, CView<S> gtau, CView<S> gnle, CView<S> gg,
                                                            double damping, int mode, double rcoef, KktProx prox,
                                                            View<S> gddq, View<S> glam) {
  using K = KktShape<S>;
  constexpr int T = K::T, NT = 64 * K::W;
  constexpr int NV = TP::T.nv;
  constexpr int NNZ = nnzOf<TP>();
  constexpr int NC = kKktMaxConstraints;
  using P = CholPattern<TP>;
  __shared__ double smem[kktLdsBytes<TP, S>() / 8];
  double* sF = smem;                                     // factor: [entry][configuration]
  S* sJ = reinterpret_cast<S*>(smem + (NNZ + NV) * T);  // J: element (i, k) at [(i + k NC) T + c]
  // after the solves: Y (element (k, j) at [(k + j NV) T + c]), J [Y r], lambda
  S* sY = reinterpret_cast<S*>(smem);
  constexpr int kYDoubles = ((NC + 1) * NV * T * (int)sizeof(S) + 7) / 8;
  double* sP = smem + kYDoubles;
  double* sL = sP + NC * (NC + 1) * T;
  const int64_t b0 = int64_t(blockIdx.x) * T;
  const int nb = (int)(B - b0 < T ? B - b0 : T);
  const int t = (int)threadIdx.x;

  // ---- phase 1: stage the factor and J (K::kU loads in flight per thread; columns past
  // the batch end repeat the last configuration and are never stored)
  {
    constexpr int kTot = (NNZ + NV) * T;
    static_for<0, (kTot + NT * K::kU - 1) / (NT * K::kU)>([&](auto I) {
      double xs[K::kU];
      static_for<0, K::kU>([&](auto U) {
        const int e = (int(I) * K::kU + int(U)) * NT + t;
        const int ec = e < kTot ? e : kTot - 1;
        const int k = ec / T, c = ec % T;
        const int64_t bc = b0 + (c < nb ? c : nb - 1);
        xs[U] = k < NNZ ? gU.p[(int64_t)k * gU.rs + bc * gU.bs] : gD.p[(int64_t)(k - NNZ) * gD.rs + bc * gD.bs];
      });
      static_for<0, K::kU>([&](auto U) {
        const int e = (int(I) * K::kU + int(U)) * NT + t;
        if (e < kTot) sF[e] = e / T < NNZ ? xs[U] : 1.0 / xs[U];
      });
    });
    // J padded to NC rows per column with zeros, so the products below need no guard
    constexpr int jTot = NC * NV * T;
    static_for<0, (jTot + NT * K::kU - 1) / (NT * K::kU)>([&](auto I) {
      S xs[K::kU];
      static_for<0, K::kU>([&](auto U) {
        const int e = (int(I) * K::kU + int(U)) * NT + t;
        const int ec = e < jTot ? e : jTot - 1;
        const int rp = ec / T, c = ec % T, i = rp % NC, k = rp / NC;
        const int64_t bc = b0 + (c < nb ? c : nb - 1);
        xs[U] = i < nc ? gJ.p[(int64_t)(i + k * nc) * gJ.rs + bc * gJ.bs] : S(0);
      });
      static_for<0, K::kU>([&](auto U) {
        const int e = (int(I) * K::kU + int(U)) * NT + t;
        if (e < jTot) sJ[e] = xs[U];
      });
    });
  }
  __syncthreads();

  // ---- phase 2: right-hand side j of configuration c (j < nc: row j of J; j = nc: tau - nle,
  // or v_before, unsolved, for impulses; none in mode 2)
  const int L = t & 63, w = t >> 6;
  const int c = L % T;
  const int j = w * K::RPW + L / T;
  const bool active = (j < nc || (j == nc && mode != 2 && mode != 4)) && c < nb;
  const int64_t b = b0 + c;
  double x[NV];
  double p[NC];
#pragma unroll
  for (int i = 0; i < NC; ++i) p[i] = 0.0;
#pragma unroll
  for (int k = 0; k < NV; ++k) x[k] = 0.0;
  if (active) {
    if (j < nc) {
#pragma unroll
      for (int k = 0; k < NV; ++k) x[k] = double(sJ[(j + k * NC) * T + c]);
    } else if (mode == 1) {
#pragma unroll
      for (int k = 0; k < NV; ++k) x[k] = double(gtau.p[(int64_t)k * gtau.rs + b * gtau.bs]);
    } else {
#pragma unroll
      for (int k = 0; k < NV; ++k)
        x[k] = double(gtau.p[(int64_t)k * gtau.rs + b * gtau.bs]) - double(gnle.p[(int64_t)k * gnle.rs + b * gnle.bs]);
    }
  }
  if (active && !(mode == 1 && j == nc)) {
    // x <- M^-1 x: Uiv, D^-1, Utiv (cholesky.hxx:193-317, 466-520).  U^-1 runs by columns:
    // once x[jc] is final it updates its ancestors' entries, independent multiply-adds
    // instead of the reference's row dot products (one dependent chain per row).  A
    // compiler memory clobber at every step keeps each factor read next to its use.
    const double* su = sF + c;
    static_for<0, NV>([&](auto JJ) {
      constexpr int jc = NV - 1 - JJ;
      stepFence();
      asm volatile("" ::: "memory");
      static_for<0, jc>([&](auto II) {
        constexpr int i = II;
        if constexpr (P::anc(i, jc)) {
          constexpr int pk = P::packed(i, jc);  // (forced to a compile-time constant)
          x[i] -= su[pk * T] * x[jc];
        }
      });
    });
    asm volatile("" ::: "memory");
#pragma unroll
    for (int k = 0; k < NV; ++k) x[k] *= su[(NNZ + k) * T];
    static_for<0, NV>([&](auto KK) {
      constexpr int k = KK;
      stepFence();
      asm volatile("" ::: "memory");
      static_for<k + 1, NV>([&](auto EE) {
        constexpr int e2 = EE;
        if constexpr (P::anc(k, e2)) {
          constexpr int pk = P::packed(k, e2);
          x[e2] -= su[pk * T] * x[k];
        }
      });
    });
  }
  if (active) {
    // p = J x over the zero-padded rows, one column of J per step (compile-time LDS offsets)
    const S* sj = sJ + c;
    static_for<0, NV>([&](auto KK) {
      constexpr int k = KK;
      asm volatile("" ::: "memory");
#pragma unroll
      for (int i = 0; i < NC; ++i) p[i] += double(sj[(i + k * NC) * T]) * x[k];
    });
  }
  __syncthreads();  // J and the factor are dead from here on

  // ---- phase 3: column j of Y and of J [Y r] into LDS (element (i, j) of the latter at
  // [(i + j NC) T + c])
  if (active) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sY[(k + j * NV) * T + c] = S(x[k]);
#pragma unroll
    for (int i = 0; i < NC; ++i)
      if (i < nc) sP[(i + j * NC) * T + c] = p[i];
  }
  __syncthreads();

  // ---- phase 4: sixteen lanes per configuration, a group of adjacent lanes of one wave
  // (lane r < NC holds row r of G = J Y + damping I, padded to NC rows with a unit
  // diagonal; lanes NC.. ride along).  Right-looking LLT: at step d the pivot's lane
  // broadcasts L_dd and each L_ed to the group, and every lane updates its own row; lane d
  // keeps the broadcast L_ed, which is column d of L.  No pivoting and no rank check, like
  // the reference's LLT: a rank-deficient J with zero damping gives non-finite results.
  // Then lambda = G^-1 (-J r - gamma) by two triangular sweeps of broadcasts, and the
  // proximal iterations with the group's residuals reduced across its lanes.  The group is
  // a DPP row: broadcasts and reductions are VALU operand modifiers (dpp.hpp).
  // mode 2 (Delassus): J M^-1 J^T out instead, or for damping mu > 0
  // (I + mu J M^-1 J^T)^-1 J M^-1 J^T.
  static_assert(NT == 16 * T, "sixteen lanes per configuration");
  if (t < 16 * nb) {  // whole groups: the configuration's sixteen lanes
    const int r = t & 15, cg = t >> 4;
    const bool row = r < nc;
    const int64_t bt = b0 + cg;
    const bool delassus = mode == 2 || mode == 4;  // matrices out, no right-hand side
    double a[NC], pr[NC], lc[NC];
    double rhs = 0.0, gam = 0.0, rid = 1.0;
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      pr[e] = row && e < nc ? sP[(r + e * NC) * T + cg] : 0.0;
      const double ge = mode == 2 ? (e == r ? 1.0 : 0.0) + damping * pr[e] : pr[e] + (e == r ? damping : 0.0);
      a[e] = row ? (e <= r ? ge : 0.0) : (e == r ? 1.0 : 0.0);
      lc[e] = 0.0;
    }
    if (row && !delassus) {
      const double grow = mode == 1 ? 0.0 : double(gg.p[(int64_t)r * gg.rs + bt * gg.bs]);
      rhs = mode == 1 ? -(1.0 + rcoef) * sP[(r + nc * NC) * T + cg]
                      : -sP[(r + nc * NC) * T + cg] + (mode == 3 ? grow : -grow);
      if (prox.maxIter > 1) gam = double(gg.p[(int64_t)r * gg.rs + bt * gg.bs]);
    }
    static_for<0, NC>([&](auto D) {
      constexpr int d = D;
      const double rd = 1.0 / sqrt(bcast16<d>(a[d]));
      rid = r == d ? rd : rid;
      a[d] = r > d ? a[d] * rd : a[d];
      static_for<d + 1, NC>([&](auto E) {
        constexpr int e = E;
        const double le = bcast16<e>(a[d]);
        a[e] = r >= e ? a[e] - a[d] * le : a[e];
        lc[e] = r == d ? le : lc[e];
      });
    });
    // the diagonal lives in rid; zero there, row r of a is strictly lower and column r in lc
    // strictly below the diagonal, so each sweep step updates every lane alike
#pragma unroll
    for (int e = 0; e < NC; ++e) a[e] = e == r ? 0.0 : a[e];
    {
      // G^-1 b for b spread one row per lane: L y = b, then L^T x = y (a lane's value stops
      // changing once the sweep has passed its row: y_r and x_r are its final value
      // times 1 / L_rr)
      auto solve = [&](double bv) {
        double s = bv;
        static_for<0, NC>([&](auto I) {
          constexpr int i = I;
          s -= a[i] * bcast16<i>(s * rid);
        });
        s *= rid;
        static_for<0, NC>([&](auto I) {
          constexpr int i = NC - 1 - I;
          s -= lc[i] * bcast16<i>(s * rid);
        });
        return s * rid;
      };
      if (delassus) {
        // column j of the output: the Delassus column (mu = 0) or (I + mu G)^-1 of it
        // (mode 2); column j of (G + mu I)^-1, scaled (mode 4)
        for (int j = 0; j < nc; ++j) {
          double gj = 0.0;
#pragma unroll
          for (int e = 0; e < NC; ++e) gj = e == j ? pr[e] : gj;
          const double bj = mode == 4 ? (r == j ? 1.0 : 0.0) : gj;
          const double xj = mode == 4 ? rcoef * solve(bj) : (damping > 0.0 ? solve(bj) : bj);
          if (row) stOut(&glam.p[(int64_t)(r + j * nc) * glam.rs + bt * glam.bs], S(xj));
        }
      }
      double lam = delassus ? 0.0 : solve(rhs);
      double prev = 0.0;
      int it = 1;
      while (!delassus && it < prox.maxIter) {
        double g = gam;
        static_for<0, NC>([&](auto D) {
          constexpr int d = D;
          g -= pr[d] * bcast16<d>(lam);
        });
        const double ares = max16(row ? fabs(g) : 0.0), rres = max16(row ? fabs(lam - prev) : 0.0);
        if (ares <= prox.absAcc || rres <= prox.relAcc) break;
        prev = lam;
        lam = solve(rhs + damping * lam);
        ++it;
      }
      if (row && !delassus) {
        sL[r * T + cg] = lam;
        stOut(&glam.p[(int64_t)r * glam.rs + bt * glam.bs], S(mode == 3 ? -lam : lam));
      }
      if (r == 0 && prox.iters) prox.iters[bt] = it;
    }
  }
  __syncthreads();
  if (mode == 2 || mode == 4) return;  // (uniform over the workgroup)

  // ---- phase 5: ddq = r + Y lambda, fixed order
  for (int e = t; e < NV * T; e += NT) {
    const int k = e / T, cc = e % T;
    if (cc < nb) {
      double a = double(sY[(k + nc * NV) * T + cc]);
      for (int i = 0; i < nc; ++i) a += sL[i * T + cc] * double(sY[(k + i * NV) * T + cc]);
      stOut(&gddq.p[(int64_t)k * gddq.rs + (b0 + cc) * gddq.bs], S(a));
    }
  }
}

}  // namespace rbd
