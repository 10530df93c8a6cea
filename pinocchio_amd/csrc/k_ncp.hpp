// k_ncp.hpp -- batched frictional contact solvers over Coulomb friction cones:
// PGSContactSolverTpl::solve (pgs-solver.hxx:19-141) and ADMMContactSolverTpl::solve
// (admm-solver.hxx:22-335), one independent problem per batch entry, FP64.
//
// A problem: nc cones (1..4), n = 3 nc rows, the Delassus matrix G (n x n, element (i, j)
// at row i + j n of its view), the free velocity g, the friction coefficients mu (one per
// cone), and for ADMM the compliance R.  Each cone's rows are (tangent x, tangent y,
// normal).  The cone operations restate constraints/coulomb-friction-cone.hpp:72-140,
// 249-270 and contact-solver-utils.hpp:25-110.
//
// PGS is Gauss-Seidel over the cones, so the work within one problem is a chain of cone
// steps.  A quad of lanes holds a problem: lane k keeps the three rows of G of cone k in
// registers and every lane a copy of x; at cone k's step lane k forms the three velocity
// rows, updates the cone (normal, then tangential with the disc projection) and broadcasts
// its three new components (DPP quad broadcasts).  64 problems per 256-thread workgroup.
//
// ADMM spends its time in the dense solves and products, which spread over rows: sixteen
// lanes hold a problem, lane r row r of G (rows n..15 ride along with a unit diagonal).
// The largest eigenvalue by power iteration (math/eigenvalues.hpp:35-55), the LLT of
// G + diag(R + mu_prox + tau rho) right-looking with the pivot's column broadcast (lane r
// keeps row r and column r of the factor), the two triangular sweeps as chains of
// broadcasts, the cone projections on each cone's three lanes, and the residuals reduced
// across the sixteen lanes.  All cross-lane moves are DPP row broadcasts and mirrors
// (dpp.hpp): VALU operand modifiers, where LDS permutes would put an LDS round trip on
// every step of the chains.  16 problems per 256-thread workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpp.hpp"
#include "kernels_common.hpp"

namespace rbd {

constexpr int kNcpMaxCones = 4, kNcpMaxRows = 3 * kNcpMaxCones;

// per-problem outputs of both solvers, one row each of `stats` (a view of kNcpStats rows):
// iterations, absolute residual, relative residual, converged (0 / 1), rho, rho_power,
// largest eigenvalue, factorizations (the last four ADMM only)
constexpr int kNcpStats = 8;

struct AdmmParams {
  double mu_prox, tau, rho_power, rho_power_factor, ratio_primal_dual, abs_prec, rel_prec;
  int max_it, max_it_eig;
};

// the three projections of one cone's triplet t (tangent x, tangent y, normal) -----------
__device__ __forceinline__ void coneProject(double mu, const double (&x)[3], double (&o)[3]) {
  const double z = x[2], muz = mu * z;
  const double tn = sqrt(x[0] * x[0] + x[1] * x[1]);
  if (mu * tn <= -z) {
    o[0] = o[1] = o[2] = 0.0;
  } else if (tn <= muz) {
    o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
  } else {
    double r0 = (mu / tn) * x[0], r1 = (mu / tn) * x[1], r2 = 1.0;
    const double nr = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    r0 /= nr; r1 /= nr; r2 /= nr;
    const double sc = x[0] * r0 + x[1] * r1 + x[2] * r2;
    o[0] = sc * r0; o[1] = sc * r1; o[2] = sc * r2;
  }
}
__device__ __forceinline__ void dualConeProject(double mu, const double (&x)[3], double (&o)[3]) {
  const double z = x[2];
  const double tn = sqrt(x[0] * x[0] + x[1] * x[1]);
  if (tn <= -mu * z) {
    o[0] = o[1] = o[2] = 0.0;
  } else if (mu * tn <= z) {
    o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
  } else {
    double r0 = x[0], r1 = x[1], r2 = mu * tn;
    const double nr = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    r0 /= nr; r1 /= nr; r2 /= nr;
    const double sc = x[0] * r0 + x[1] * r1 + x[2] * r2;
    o[0] = sc * r0; o[1] = sc * r1; o[2] = sc * r2;
  }
}

// ---------------------------------------------------------------------------------------
// PGS
// ---------------------------------------------------------------------------------------
constexpr int kPgsLanes = 4;

// gG: n^2 rows, gg / gx: n rows, gmu: nc rows; gx holds the initial guess and receives the
// solution; gst: kNcpStats rows (iterations, absolute, relative residual, converged).
template<int NC>
__global__ __launch_bounds__(256) void k_pgs(int64_t B, CView<double> gG, CView<double> gg, CView<double> gmu,
                                            View<double> gx, double over_relax, int max_it, double abs_prec,
                                            double rel_prec, View<double> gst) {
  constexpr int N = 3 * NC;
  const int t = (int)threadIdx.x;
  const int k = t & (kPgsLanes - 1);
  const int64_t b = (int64_t)blockIdx.x * (256 / kPgsLanes) + (t >> 2);
  const bool live = b < B;  // the same for the problem's four lanes
  if (!live) return;
  const bool own = k < NC;
  // lane k: rows 3k..3k+2 of G, its g and mu; everyone: x
  double gr[3][N], gv[3], mu = 0.0, x[N];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < N; ++j) gr[i][j] = own ? gG.p[(int64_t)(3 * k + i + j * N) * gG.rs + b * gG.bs] : 0.0;
    gv[i] = own ? gg.p[(int64_t)(3 * k + i) * gg.rs + b * gg.bs] : 0.0;
  }
  if (own) mu = gmu.p[(int64_t)k * gmu.rs + b * gmu.bs];
#pragma unroll
  for (int j = 0; j < N; ++j) x[j] = gx.p[(int64_t)j * gx.rs + b * gx.bs];
  auto ninf = [&]() {
    double m = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) m = fmax(m, fabs(x[j]));
    return m;
  };
  double xprevNorm = ninf();
  double compl_ = 0.0, dualf = 0.0, prox = 0.0;
  bool absR = false, relR = false;
  int it = 0;
  for (; it < max_it; ++it) {
    double xp[N];
#pragma unroll
    for (int j = 0; j < N; ++j) xp[j] = x[j];
    double cc = 0.0, df = 0.0;  // this lane's cone
    static_for<0, NC>([&](auto C) {
      constexpr int c = C, r0 = 3 * c;
      if (k == c) {
        double vel[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < N; ++j) s += gr[i][j] * x[j];
          vel[i] = s + gv[i];
        }
        const double fzp = x[r0 + 2];
        double fz = fzp - (over_relax / gr[2][r0 + 2]) * vel[2];
        fz = fmax(0.0, fz);
        x[r0 + 2] = fz;
#pragma unroll
        for (int i = 0; i < 3; ++i) vel[i] += gr[i][r0 + 2] * (fz - fzp);
        const double minD = fmin(gr[0][r0], gr[1][r0 + 1]);
        const double t0p = x[r0], t1p = x[r0 + 1];
        double t0 = t0p - over_relax / minD * vel[0], t1 = t1p - over_relax / minD * vel[1];
        const double ftn = sqrt(t0 * t0 + t1 * t1);
        const double mufz = mu * fz;
        if (ftn > mufz) {
          t0 *= mufz / ftn;
          t1 *= mufz / ftn;
        }
        x[r0] = t0;
        x[r0 + 1] = t1;
        // the velocity after the tangential update is that update's own contribution
        // (as the reference writes it); complementarity and dual feasibility on it
        const double d0 = t0 - t0p, d1 = t1 - t1p;
#pragma unroll
        for (int i = 0; i < 3; ++i) vel[i] = gr[i][r0] * d0 + gr[i][r0 + 1] * d1;
        const double ncorr = mu * sqrt(vel[0] * vel[0] + vel[1] * vel[1]);
        cc = fabs(t0 * vel[0] + t1 * vel[1] + fz * (vel[2] + ncorr));
        const double w[3] = {vel[0], vel[1], vel[2] + ncorr};
        double o[3];
        dualConeProject(mu, w, o);
        const double e0 = o[0] - w[0], e1 = o[1] - w[1], e2 = o[2] - w[2];
        df = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
      }
#pragma unroll
      for (int e = 0; e < 3; ++e) x[r0 + e] = bcast4<c>(x[r0 + e]);
    });
    cc = max4(cc);
    df = max4(df);
    compl_ = cc;
    dualf = df;
    absR = compl_ <= abs_prec && dualf <= abs_prec;
    prox = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) prox = fmax(prox, fabs(x[j] - xp[j]));
    const double xn = ninf();
    relR = prox <= rel_prec * fmax(xn, xprevNorm);
    if (absR || relR) break;
    xprevNorm = xn;
  }
  static_for<0, NC>([&](auto C) {
    constexpr int c = C;
    if (k == c) {
#pragma unroll
      for (int e = 0; e < 3; ++e) gx.p[(int64_t)(3 * c + e) * gx.rs + b * gx.bs] = x[3 * c + e];
    }
  });
  if (k == 0) {
    const double st[4] = {double(it), fmax(compl_, dualf), prox, (absR || relR) ? 1.0 : 0.0};
#pragma unroll
    for (int e = 0; e < 4; ++e) gst.p[(int64_t)e * gst.rs + b * gst.bs] = st[e];
  }
}

// ---------------------------------------------------------------------------------------
// ADMM
// ---------------------------------------------------------------------------------------
constexpr int kAdmmLanes = 16;


// gG: n^2 rows; gg, gR: n rows; gmu: nc rows; gpg / gdg: initial primal / dual guesses (n
// rows, p == nullptr for none); outputs gy (primal solution), gz (dual solution), gs (the
// complementarity shift) n rows each, gst kNcpStats rows.
template<int NC>
__global__ __launch_bounds__(256) void k_admm(int64_t B, CView<double> gG, CView<double> gg, CView<double> gmu,
                                             CView<double> gR, CView<double> gpg, CView<double> gdg, AdmmParams P,
                                             View<double> gy, View<double> gz, View<double> gs, View<double> gst) {
  constexpr int N = 3 * NC;
  const int t = (int)threadIdx.x;
  const int r = t & (kAdmmLanes - 1);
  const int64_t b = (int64_t)blockIdx.x * (256 / kAdmmLanes) + (t >> 4);
  if (b >= B) return;  // the same for the problem's sixteen lanes
  const bool row = r < N;
  const int base = (r / 3) * 3, comp = r % 3;
  double grow[N];
#pragma unroll
  for (int j = 0; j < N; ++j) grow[j] = row ? gG.p[(int64_t)(r + j * N) * gG.rs + b * gG.bs] : 0.0;
  const double g = row ? gg.p[(int64_t)r * gg.rs + b * gg.bs] : 0.0;
  const double R = row ? gR.p[(int64_t)r * gR.rs + b * gR.bs] : 0.0;
  const double mu = row ? gmu.p[(int64_t)(r / 3) * gmu.rs + b * gmu.bs] : 0.0;
  // (G v)_r for v spread one entry per lane
  auto matvec = [&](double v) {
    double s = 0.0;
    static_for<0, N>([&](auto J) { s += grow[J] * bcast16<J>(v); });
    return s;
  };
  // the three entries of this lane's cone (lanes of other cones: their own cone's)
  auto triplet = [&](double v, double (&o)[3]) { triple3(v, comp, o); };
  auto projK = [&](double v) {
    double tv[3], o[3];
    triplet(v, tv);
    coneProject(mu, tv, o);
    return row ? (comp == 0 ? o[0] : comp == 1 ? o[1] : o[2]) : 0.0;
  };
  auto projKd = [&](double v) {
    double tv[3], o[3];
    triplet(v, tv);
    dualConeProject(mu, tv, o);
    return row ? (comp == 0 ? o[0] : comp == 1 ? o[1] : o[2]) : 0.0;
  };
  auto shiftOf = [&](double v) {  // (0, 0, mu |v_t|) on the cone's normal row
    double tv[3];
    triplet(v, tv);
    return row && comp == 2 ? mu * sqrt(tv[0] * tv[0] + tv[1] * tv[1]) : 0.0;
  };
  // largest eigenvalue of G (power iteration from 1 / sqrt(n), relative tolerance 1e-8)
  double pv = row ? 1.0 / sqrt(double(N)) : 0.0;
  double est = sqrt(sum16(pv * pv));
  for (int k = 0; k < P.max_it_eig; ++k) {
    const double prev = est;
    pv /= est;
    pv = matvec(pv);
    est = sqrt(sum16(pv * pv));
    if (fabs(prev - est) <= 1e-8 * fmax(fabs(prev), fabs(est))) break;
  }
  const double L = est;
  const double muR = -max16(row ? -R : -1.0e300);
  const double m = P.mu_prox + muR;
  const double cond = L / m;
  const double rhoInc = pow(cond, P.rho_power_factor);
  double rho = sqrt(L * m) * pow(cond, P.rho_power);
  // factor of G + diag(R + prox_value): lane r keeps row r (a) and column r (lc) of L
  double a[N], lc[N], rid = 1.0;
  auto factor = [&](double proxValue) {
#pragma unroll
    for (int e = 0; e < N; ++e) {
      a[e] = row ? (e <= r ? grow[e] + (e == r ? R + proxValue : 0.0) : 0.0) : 0.0;
      lc[e] = 0.0;
    }
    if (!row) rid = 1.0;
    // (selects, not branches: every lane runs every step)
    static_for<0, N>([&](auto D) {
      constexpr int d = D;
      const double rd = 1.0 / sqrt(bcast16<d>(a[d]));
      rid = r == d ? rd : rid;
      a[d] = r > d ? a[d] * rd : a[d];
      static_for<d + 1, N>([&](auto E) {
        constexpr int e = E;
        const double le = bcast16<e>(a[d]);
        a[e] = r >= e ? a[e] - a[d] * le : a[e];
        lc[e] = r == d ? le : lc[e];
      });
    });
    // the diagonal lives in rid: a zero there lets the sweeps update every lane alike
#pragma unroll
    for (int e = 0; e < N; ++e) a[e] = e == r ? 0.0 : a[e];
  };
  // (a: strictly lower row, zero from the diagonal on; lc: strictly lower column, zero
  // down to the diagonal -- so each step updates every lane with one multiply-add, and a
  // lane's value stops changing once the sweep has passed its row: y_r and x_r are its
  // final value times 1 / L_rr)
  auto solve = [&](double bv) {
    double s = bv;
    static_for<0, N>([&](auto I) {
      constexpr int i = I;
      s -= a[i] * bcast16<i>(s * rid);
    });
    s *= rid;
    static_for<0, N>([&](auto I) {
      constexpr int i = N - 1 - I;
      s -= lc[i] * bcast16<i>(s * rid);
    });
    return row ? s * rid : 0.0;
  };
  double proxValue = P.mu_prox + P.tau * rho;
  factor(proxValue);
  int cholUpdates = 1;
  double x = gpg.p && row ? gpg.p[(int64_t)r * gpg.rs + b * gpg.bs] : 0.0;
  double y = projK(x);
  double z, s = 0.0;
  if (gdg.p) {
    z = row ? gdg.p[(int64_t)r * gdg.rs + b * gdg.bs] : 0.0;
  } else {
    z = matvec(y) + (R + proxValue) * y;  // the damped operator, as applyOnTheRight
    z += -proxValue * y + g;
    s = shiftOf(z);
    z += s;
    z = projKd(z);
  }
  double compl_ = 0.0, prox = 0.0, primal = 0.0, dual = 0.0;
  bool absR = false;
  int it = 1;
  for (; it <= P.max_it; ++it) {
    const double xp = x, yp = y;
    s = shiftOf(z);
    x = solve(row ? -(g + s - (rho * P.tau) * y - P.mu_prox * x - z) : 0.0);
    y = projK(x - z / (P.tau * rho));
    z -= (P.tau * rho) * (x - y);
    const double dy = y - yp;
    primal = max16(fabs(x - y));
    prox = max16(fabs(dy));
    dual = max16(fabs((P.tau * rho) * dy + P.mu_prox * (x - xp)));
    double zy[3], yy[3];
    triplet(z, zy);
    triplet(y, yy);
    compl_ = max16(row ? fabs(zy[0] * yy[0] + zy[1] * yy[1] + zy[2] * yy[2]) : 0.0);
    absR = compl_ <= P.abs_prec && dual <= P.abs_prec && primal <= P.abs_prec;
    if (absR) break;
    bool refactor = false;
    if (primal > P.ratio_primal_dual * dual) {
      rho *= rhoInc;
      refactor = true;
    } else if (dual > P.ratio_primal_dual * primal) {
      rho /= rhoInc;
      refactor = true;
    }
    if (refactor) {
      proxValue = P.mu_prox + P.tau * rho;
      factor(proxValue);
      ++cholUpdates;
    }
  }
  if (row) {
    gy.p[(int64_t)r * gy.rs + b * gy.bs] = y;
    gz.p[(int64_t)r * gz.rs + b * gz.bs] = z;
    if (gs.p) gs.p[(int64_t)r * gs.rs + b * gs.bs] = s;
  }
  if (r == 0) {
    const double st[kNcpStats] = {double(it),
                                  fmax(primal, fmax(compl_, dual)),
                                  prox,
                                  absR ? 1.0 : 0.0,
                                  rho,
                                  log(rho / sqrt(L * m)) / log(cond),
                                  L,
                                  double(cholUpdates)};
#pragma unroll
    for (int e = 0; e < kNcpStats; ++e) gst.p[(int64_t)e * gst.rs + b * gst.bs] = st[e];
  }
}

}  // namespace rbd
