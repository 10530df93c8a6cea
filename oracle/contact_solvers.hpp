// contact_solvers.hpp -- CPU restatement of the reference's frictional contact solvers,
// test infrastructure only (the GPU kernels are checked against it; nothing in the product
// links it).
//
// A contact problem: nc Coulomb cones, problem size n = 3 nc, the Delassus matrix G
// (n x n, column major), the free velocity g, friction coefficients mu[c].  Each cone's
// three components are (tangent x, tangent y, normal).
//
//   Coulomb friction cone   constraints/coulomb-friction-cone.hpp:20-199 (project :72-99,
//                           computeNormalCorrection :130-140, computeContactComplementarity
//                           :170-175) and its dual, :202-290 (project :249-270)
//   solver utilities        contact-solver-utils.hpp:25-110 (cone projections, conic
//                           complementarity, complementarity shift)
//   PGS                     pgs-solver.hxx:19-141
//   ADMM                    admm-solver.hxx:22-335, admm-solver.hpp:120-127 (defaults),
//                           :301-316 (computeRho, computeRhoPower)
//   power iteration         math/eigenvalues.hpp:35-55 (largest eigenvalue, start vector
//                           1 / sqrt(n) in every component)
//   dense Delassus          delassus-operator-dense.hpp:54-98 (updateDamping = LLT of
//                           G + diag(d); applyOnTheRight = G x + d .* x)
#pragma once

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace orc {

template<class S> struct ConeOps {
  // projection onto {|f_t| <= mu f_n} (coulomb-friction-cone.hpp:72-99)
  static void project(S mu, const S* x, S* out) {
    using std::sqrt;
    const S z = x[2], muz = mu * z;
    const S tn = sqrt(x[0] * x[0] + x[1] * x[1]);
    if (mu * tn <= -z) {
      out[0] = out[1] = out[2] = S(0);
    } else if (tn <= muz) {
      out[0] = x[0]; out[1] = x[1]; out[2] = x[2];
    } else {
      S r[3] = {(mu / tn) * x[0], (mu / tn) * x[1], S(1)};
      const S nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      for (int k = 0; k < 3; ++k) r[k] /= nr;
      const S scale = x[0] * r[0] + x[1] * r[1] + x[2] * r[2];
      for (int k = 0; k < 3; ++k) out[k] = scale * r[k];
    }
  }
  // projection onto the dual cone {mu |v_t| <= v_n} (coulomb-friction-cone.hpp:249-270)
  static void dualProject(S mu, const S* x, S* out) {
    using std::sqrt;
    const S z = x[2];
    const S tn = sqrt(x[0] * x[0] + x[1] * x[1]);
    if (tn <= -mu * z) {
      out[0] = out[1] = out[2] = S(0);
    } else if (mu * tn <= z) {
      out[0] = x[0]; out[1] = x[1]; out[2] = x[2];
    } else {
      S r[3] = {x[0], x[1], mu * tn};
      const S nr = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      for (int k = 0; k < 3; ++k) r[k] /= nr;
      const S scale = x[0] * r[0] + x[1] * r[1] + x[2] * r[2];
      for (int k = 0; k < 3; ++k) out[k] = scale * r[k];
    }
  }
  // (0, 0, mu |v_t|) (coulomb-friction-cone.hpp:128-140)
  static S normalCorrection(S mu, const S* v) { return mu * std::sqrt(v[0] * v[0] + v[1] * v[1]); }
};

template<class S> struct PgsResult {
  bool converged;
  int it;
  S absolute_residual, relative_residual;
};

// PGSContactSolverTpl::solve (pgs-solver.hxx:19-141), x in/out.  Note the reference's
// velocity after the tangential update: it is replaced by the tangential update's own
// contribution (not accumulated), and the complementarity and dual feasibility of the
// cone are taken on that vector; restated as written.
template<class S>
PgsResult<S> pgsSolve(int nc, const S* G, const S* g, const S* mu, S* x_sol, S over_relax, int max_it,
                      S absolute_precision, S relative_precision) {
  using std::abs; using std::sqrt;
  if (!(over_relax < S(2) && over_relax > S(0))) throw std::invalid_argument("over_relax should lie in ]0,2[.");
  const int n = 3 * nc;
  auto Gat = [&](int i, int j) { return G[i + (size_t)j * n]; };
  std::vector<S> x(x_sol, x_sol + n), xprev(n);
  auto ninf = [&](const std::vector<S>& a) { S m = S(0); for (S e : a) m = std::max(m, abs(e)); return m; };
  S x_previous_norm_inf = ninf(x);
  S complementarity = S(0), dual_feasibility = S(0), proximal_metric = S(0);
  bool abs_reached = false, rel_reached = false;
  int it = 0;
  for (; it < max_it; ++it) {
    xprev = x;
    complementarity = S(0);
    dual_feasibility = S(0);
    for (int c = 0; c < nc; ++c) {
      const int r0 = 3 * c;
      S vel[3];
      for (int k = 0; k < 3; ++k) {
        S s = S(0);
        for (int j = 0; j < n; ++j) s += Gat(r0 + k, j) * x[j];
        vel[k] = s + g[r0 + k];
      }
      // normal update
      S& fz = x[r0 + 2];
      const S fz_prev = fz;
      fz -= S(over_relax / Gat(r0 + 2, r0 + 2)) * vel[2];
      fz = std::max(S(0), fz);
      for (int k = 0; k < 3; ++k) vel[k] += Gat(r0 + k, r0 + 2) * (fz - fz_prev);
      // tangential update, projected on the disc of radius mu fz
      const S min_D = std::min(Gat(r0, r0), Gat(r0 + 1, r0 + 1));
      const S ft_prev[2] = {x[r0], x[r0 + 1]};
      x[r0] -= over_relax / min_D * vel[0];
      x[r0 + 1] -= over_relax / min_D * vel[1];
      const S ftn = sqrt(x[r0] * x[r0] + x[r0 + 1] * x[r0 + 1]);
      const S mufz = mu[c] * fz;
      if (ftn > mufz) {
        x[r0] *= mufz / ftn;
        x[r0 + 1] *= mufz / ftn;
      }
      const S d0 = x[r0] - ft_prev[0], d1 = x[r0 + 1] - ft_prev[1];
      for (int k = 0; k < 3; ++k) vel[k] = Gat(r0 + k, r0) * d0 + Gat(r0 + k, r0 + 1) * d1;
      // feasibility of the cone
      const S ncorr = ConeOps<S>::normalCorrection(mu[c], vel);
      const S cc = abs(x[r0] * vel[0] + x[r0 + 1] * vel[1] + x[r0 + 2] * (vel[2] + ncorr));
      complementarity = std::max(complementarity, cc);
      vel[2] += ncorr;
      S vp[3];
      ConeOps<S>::dualProject(mu[c], vel, vp);
      const S e0 = vp[0] - vel[0], e1 = vp[1] - vel[1], e2 = vp[2] - vel[2];
      dual_feasibility = std::max(dual_feasibility, sqrt(e0 * e0 + e1 * e1 + e2 * e2));
    }
    abs_reached = complementarity <= absolute_precision && dual_feasibility <= absolute_precision;
    proximal_metric = S(0);
    for (int k = 0; k < n; ++k) proximal_metric = std::max(proximal_metric, abs(x[k] - xprev[k]));
    const S x_norm_inf = ninf(x);
    rel_reached = proximal_metric <= relative_precision * std::max(x_norm_inf, x_previous_norm_inf);
    if (abs_reached || rel_reached) break;
    x_previous_norm_inf = x_norm_inf;
  }
  std::copy(x.begin(), x.end(), x_sol);
  return {abs_reached || rel_reached, it, std::max(complementarity, dual_feasibility), proximal_metric};
}

template<class S> struct AdmmSettings {
  S mu_prox = S(1e-6), tau = S(0.5), rho_power = S(0.2), rho_power_factor = S(0.05), ratio_primal_dual = S(10);
  int max_it_largest_eigen_value = 20;
  int max_it = 1000;
  S absolute_precision = S(1e-6), relative_precision = S(1e-6);
};
template<class S> struct AdmmResult {
  bool converged;
  int it, cholesky_update_count;
  S absolute_residual, relative_residual, rho, rho_power, largest_eigen_value;
};

// Dense LLT of A = G + diag(d) (column major, lower factor into L)
template<class S> void denseLLT(int n, const S* G, const std::vector<S>& d, std::vector<S>& L) {
  using std::sqrt;
  L.assign((size_t)n * n, S(0));
  for (int j = 0; j < n; ++j) {
    S s = G[j + (size_t)j * n] + d[j];
    for (int k = 0; k < j; ++k) s -= L[j + (size_t)k * n] * L[j + (size_t)k * n];
    const S ljj = sqrt(s);
    L[j + (size_t)j * n] = ljj;
    for (int i = j + 1; i < n; ++i) {
      S t = G[i + (size_t)j * n];
      for (int k = 0; k < j; ++k) t -= L[i + (size_t)k * n] * L[j + (size_t)k * n];
      L[i + (size_t)j * n] = t / ljj;
    }
  }
}
template<class S> void denseLLTSolve(int n, const std::vector<S>& L, S* x) {
  for (int i = 0; i < n; ++i) {
    S s = x[i];
    for (int k = 0; k < i; ++k) s -= L[i + (size_t)k * n] * x[k];
    x[i] = s / L[i + (size_t)i * n];
  }
  for (int i = n - 1; i >= 0; --i) {
    S s = x[i];
    for (int k = i + 1; k < n; ++k) s -= L[k + (size_t)i * n] * x[k];
    x[i] = s / L[i + (size_t)i * n];
  }
}

// ADMMContactSolverTpl::solve (admm-solver.hxx:22-335) for a fresh solver on a fresh dense
// Delassus operator (zero damping at entry): the largest eigenvalue of G by power
// iteration, rho = sqrt(L m) (L / m)^rho_power with m = mu_prox + min(R), the factor of
// G + diag(R + mu_prox + tau rho), then the x / y / z updates with the De Saxce shift, the
// primal / dual / complementarity residuals, and rho scaled by (L / m)^rho_power_factor
// whenever one residual exceeds ratio_primal_dual times the other.  Outputs y (primal
// solution), z (dual solution), s (complementarity shift).
template<class S>
AdmmResult<S> admmSolve(int nc, const S* G, const S* g, const S* mu, const S* R, const S* primal_guess,
                        const S* dual_guess, const AdmmSettings<S>& st, S* y_out, S* z_out, S* s_out) {
  using std::abs; using std::sqrt; using std::pow; using std::log;
  const int n = 3 * nc;
  S mu_R = R[0];
  for (int k = 1; k < n; ++k) mu_R = std::min(mu_R, R[k]);
  if (!(st.tau <= S(1) && st.tau > S(0))) throw std::invalid_argument("tau should lie in ]0,1].");
  if (!(st.mu_prox >= S(0))) throw std::invalid_argument("mu_prox should be positive.");
  if (!(mu_R >= S(0))) throw std::invalid_argument("R should be a positive vector.");
  auto matvec = [&](const std::vector<S>& v, std::vector<S>& out) {
    for (int i = 0; i < n; ++i) {
      S s = S(0);
      for (int j = 0; j < n; ++j) s += G[i + (size_t)j * n] * v[j];
      out[i] = s;
    }
  };
  auto norm2 = [&](const std::vector<S>& v) { S s = S(0); for (S e : v) s += e * e; return sqrt(s); };
  // power iteration (math/eigenvalues.hpp:35-55), rel_tol 1e-8
  std::vector<S> pv(n, S(1) / sqrt(S(n))), pw(n);
  S est = norm2(pv);
  for (int k = 0; k < st.max_it_largest_eigen_value; ++k) {
    const S prev = est;
    for (S& e : pv) e /= est;
    matvec(pv, pw);
    pv = pw;
    est = norm2(pv);
    if (abs(prev - est) <= S(1e-8) * std::max(abs(prev), abs(est))) break;
  }
  const S L = est;
  const S m = st.mu_prox + mu_R;
  const S cond = L / m;
  const S rho_increment = pow(cond, st.rho_power_factor);
  S rho = sqrt(L * m) * pow(cond, st.rho_power);
  S prox_value = st.mu_prox + st.tau * rho;
  std::vector<S> damp(n), Lf;
  for (int k = 0; k < n; ++k) damp[k] = R[k] + prox_value;
  denseLLT(n, G, damp, Lf);
  int chol_updates = 1;
  std::vector<S> x(n, S(0)), y(n), z(n), s(n), rhs(n), xp(n), yp(n), tmp(n);
  if (primal_guess) std::copy(primal_guess, primal_guess + n, x.begin());
  auto coneProj = [&](const std::vector<S>& in, std::vector<S>& out) {
    for (int c = 0; c < nc; ++c) ConeOps<S>::project(mu[c], &in[3 * c], &out[3 * c]);
  };
  auto shift = [&](const std::vector<S>& v, std::vector<S>& out) {
    for (int c = 0; c < nc; ++c) {
      out[3 * c] = out[3 * c + 1] = S(0);
      out[3 * c + 2] = ConeOps<S>::normalCorrection(mu[c], &v[3 * c]);
    }
  };
  coneProj(x, y);
  if (dual_guess) {
    std::copy(dual_guess, dual_guess + n, z.begin());
  } else {
    matvec(y, z);
    for (int k = 0; k < n; ++k) z[k] += damp[k] * y[k];  // applyOnTheRight with the damping
    for (int k = 0; k < n; ++k) z[k] += -prox_value * y[k] + g[k];
    shift(z, s);
    for (int k = 0; k < n; ++k) z[k] += s[k];
    for (int c = 0; c < nc; ++c) {
      S o[3];
      ConeOps<S>::dualProject(mu[c], &z[3 * c], o);
      for (int e = 0; e < 3; ++e) z[3 * c + e] = o[e];
    }
  }
  S complementarity = S(0), proximal_metric = S(0), primal_feasibility = S(0), dual_feasibility = S(0);
  bool abs_reached = false;
  int it = 1;
  for (; it <= st.max_it; ++it) {
    xp = x;
    yp = y;
    shift(z, s);
    for (int k = 0; k < n; ++k) rhs[k] = -(g[k] + s[k] - (rho * st.tau) * y[k] - st.mu_prox * x[k] - z[k]);
    denseLLTSolve(n, Lf, rhs.data());
    x = rhs;
    for (int k = 0; k < n; ++k) tmp[k] = x[k] - z[k] / (st.tau * rho);
    coneProj(tmp, y);
    for (int k = 0; k < n; ++k) z[k] -= (st.tau * rho) * (x[k] - y[k]);
    primal_feasibility = S(0);
    proximal_metric = S(0);
    dual_feasibility = S(0);
    for (int k = 0; k < n; ++k) {
      primal_feasibility = std::max(primal_feasibility, abs(x[k] - y[k]));
      const S dy = y[k] - yp[k];
      proximal_metric = std::max(proximal_metric, abs(dy));
      dual_feasibility = std::max(dual_feasibility, abs((st.tau * rho) * dy + st.mu_prox * (x[k] - xp[k])));
    }
    complementarity = S(0);
    for (int c = 0; c < nc; ++c) {
      const S cc = abs(z[3 * c] * y[3 * c] + z[3 * c + 1] * y[3 * c + 1] + z[3 * c + 2] * y[3 * c + 2]);
      complementarity = std::max(complementarity, cc);
    }
    abs_reached = complementarity <= st.absolute_precision && dual_feasibility <= st.absolute_precision &&
                  primal_feasibility <= st.absolute_precision;
    if (abs_reached) break;
    bool refactor = false;
    if (primal_feasibility > st.ratio_primal_dual * dual_feasibility) {
      rho *= rho_increment;
      refactor = true;
    } else if (dual_feasibility > st.ratio_primal_dual * primal_feasibility) {
      rho /= rho_increment;
      refactor = true;
    }
    if (refactor) {
      prox_value = st.mu_prox + st.tau * rho;
      for (int k = 0; k < n; ++k) damp[k] = R[k] + prox_value;
      denseLLT(n, G, damp, Lf);
      ++chol_updates;
    }
  }
  std::copy(y.begin(), y.end(), y_out);
  std::copy(z.begin(), z.end(), z_out);
  if (s_out) std::copy(s.begin(), s.end(), s_out);
  AdmmResult<S> r{};
  r.converged = abs_reached;
  r.it = it;
  r.cholesky_update_count = chol_updates;
  r.absolute_residual = std::max(primal_feasibility, std::max(complementarity, dual_feasibility));
  r.relative_residual = proximal_metric;
  r.rho = rho;
  r.rho_power = log(rho / sqrt(L * m)) / log(cond);
  r.largest_eigen_value = L;
  return r;
}

}  // namespace orc
