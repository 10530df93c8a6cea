// oracle_capi.cpp -- C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Batched wrappers in the shape of rneaInParallel / abaInParallel
// (algorithm/parallel/rnea.hpp:39-83, algorithm/parallel/aba.hpp:40-84): an OpenMP
// `parallel for schedule(static)` over configurations with one Data workspace per thread.
// All arrays are column-major with one configuration per column (Eigen's default):
// q is nq x B, v/a/tau/ddq are nv x B, matrices are nv x nv per configuration stored
// back to back.  Inputs and outputs are double; the `_s` suffix selects the scalar type
// the algorithm runs in (0 = double, 1 = float, 2 = long double), so the float rounding
// of the reference algorithm itself can be measured.
#include <omp.h>

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "contact_solvers.hpp"
#include "rbd_oracle.hpp"

using namespace oracle;

namespace {
thread_local std::string g_err;

template<class S, class Fn>
int runBatch(const ModelDesc* d, int64_t B, int nthreads, Fn&& fn) {
  try {
    if (!d) throw std::invalid_argument("null model");
    if (B < 0) throw std::invalid_argument("negative batch");
    const Model<S> model(*d);
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    std::string err;
#pragma omp parallel num_threads(nt)
    {
      Data<S> data(model);
#pragma omp for schedule(static)
      for (int64_t b = 0; b < B; ++b) {
        try {
          fn(model, data, b);
        } catch (const std::exception& e) {
#pragma omp critical
          err = e.what();
        }
      }
    }
    if (!err.empty()) throw std::runtime_error(err);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

template<class S> struct Buf {
  std::vector<S> x;
  const S* load(const double* src, int n) {
    x.resize(n);
    for (int k = 0; k < n; ++k) x[k] = S(src[k]);
    return x.data();
  }
};
template<class S> inline void store(const S* src, double* dst, int n) {
  for (int k = 0; k < n; ++k) dst[k] = double(src[k]);
}
template<class S> inline void storeMat(const Mat<S>& m, double* dst) {
  for (size_t k = 0; k < m.a.size(); ++k) dst[k] = double(m.a[k]);
}

template<class S>
int rneaT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a, double* tau, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv, ba;
    static thread_local std::vector<S> out;
    out.resize(m.nv);
    rnea(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * m.nv, m.nv), ba.load(a + b * m.nv, m.nv), out.data());
    store(out.data(), tau + b * m.nv, m.nv);
  });
}
template<class S>
int abaT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* tau, double* ddq, int conv,
         int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv, bt;
    static thread_local std::vector<S> out;
    out.resize(m.nv);
    const S* qq = bq.load(q + b * m.nq, m.nq);
    const S* vv = bv.load(v + b * m.nv, m.nv);
    const S* tt = bt.load(tau + b * m.nv, m.nv);
    if (conv == 0) abaLocal(m, data, qq, vv, tt, out.data());
    else abaWorld(m, data, qq, vv, tt, out.data());
    store(out.data(), ddq + b * m.nv, m.nv);
  });
}
template<class S> int crbaT(const ModelDesc* d, int64_t B, const double* q, double* M, int symmetrize, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    if (!m.isCompact()) throw std::invalid_argument("CRBA requires a compact (DFS-ordered) model");
    static thread_local Buf<S> bq;
    crbaWorld(m, data, bq.load(q + b * m.nq, m.nq));
    if (symmetrize)
      for (int c = 0; c < m.nv; ++c)
        for (int r = c + 1; r < m.nv; ++r) data.M(r, c) = data.M(c, r);
    storeMat(data.M, M + b * m.nv * m.nv);
  });
}
template<class S>
int choleskyT(const ModelDesc* d, int64_t B, const double* M, double* U, double* D, double* Minv, double* x, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    const int nv = m.nv;
    const double* Mb = M + b * nv * nv;
    for (int k = 0; k < nv * nv; ++k) data.M.a[k] = S(Mb[k]);
    choleskyDecompose(m, data);
    if (U) storeMat(data.Ufac, U + b * nv * nv);
    if (D) store(data.D.data(), D + b * nv, nv);
    if (Minv) {
      Mat<S> mi(nv, nv);
      choleskyComputeMinv(m, data, mi);
      for (int c = 0; c < nv; ++c)
        for (int r = c + 1; r < nv; ++r) mi(r, c) = mi(c, r);
      storeMat(mi, Minv + b * nv * nv);
    }
    if (x) {
      std::vector<S> xs(nv);
      for (int k = 0; k < nv; ++k) xs[k] = S(x[b * nv + k]);
      choleskySolve(m, data, xs.data());
      store(xs.data(), x + b * nv, nv);
    }
  });
}
template<class S>
int rneaDerivT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a, double* dq,
               double* dv, double* M, double* tau, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv, ba;
    const int nv = m.nv;
    rneaDerivatives(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * nv, nv), ba.load(a + b * nv, nv));
    if (dq) storeMat(data.dtau_dq, dq + b * nv * nv);
    if (dv) storeMat(data.dtau_dv, dv + b * nv * nv);
    if (M) storeMat(data.M, M + b * nv * nv);
    if (tau) store(data.tau.data(), tau + b * nv, nv);
  });
}
template<class S>
int abaDerivT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* tau, double* dq,
              double* dv, double* Minv, double* ddq, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv, bt;
    const int nv = m.nv;
    Mat<S> Xq(nv, nv), Xv(nv, nv), Mi(nv, nv);
    abaDerivatives(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * nv, nv), bt.load(tau + b * nv, nv), Xq, Xv,
                   Mi);
    if (dq) storeMat(Xq, dq + b * nv * nv);
    if (dv) storeMat(Xv, dv + b * nv * nv);
    if (Minv) storeMat(Mi, Minv + b * nv * nv);
    if (ddq) store(data.ddq.data(), ddq + b * nv, nv);
  });
}
// forward kinematics: per configuration oMi of joints 1..nj-1 as 12 numbers each (R row
// major, then translation), and (optional) local velocities / accelerations, 6 per joint
template<class S>
int fkT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a, double* oM, double* vo,
        double* ao, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv, ba;
    const int nj = m.njoints;
    const S* vv = v ? bv.load(v + b * m.nv, m.nv) : nullptr;
    const S* aa = a ? ba.load(a + b * m.nv, m.nv) : nullptr;
    forwardKinematics(m, data, bq.load(q + b * m.nq, m.nq), vv, aa);
    for (int i = 1; i < nj; ++i) {
      double* o = oM + (b * (nj - 1) + (i - 1)) * 12;
      for (int k = 0; k < 9; ++k) o[k] = double(data.oMi[i].R.m[k]);
      for (int k = 0; k < 3; ++k) o[9 + k] = double(data.oMi[i].p[k]);
      S tmp[6];
      if (vo) {
        toVec(data.v[i], tmp);
        store(tmp, vo + (b * (nj - 1) + (i - 1)) * 6, 6);
      }
      if (ao) {
        toVec(data.a[i], tmp);
        store(tmp, ao + (b * (nj - 1) + (i - 1)) * 6, 6);
      }
    }
  });
}
template<class S> int jacT(const ModelDesc* d, int64_t B, const double* q, double* J, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq;
    computeJointJacobians(m, data, bq.load(q + b * m.nq, m.nq));
    storeMat(data.J, J + b * 6 * m.nv);
  });
}
template<class S> int jacDotT(const ModelDesc* d, int64_t B, const double* q, const double* v, double* J, double* dJ,
                               int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv;
    static thread_local Mat<S> dJm;
    if (dJm.rows != 6 || dJm.cols != m.nv) dJm = Mat<S>(6, m.nv);
    computeJointJacobiansTimeVariation(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * m.nv, m.nv), dJm);
    storeMat(data.J, J + b * 6 * m.nv);
    storeMat(dJm, dJ + b * 6 * m.nv);
  });
}
template<class S>
int frameJacDotT(const ModelDesc* d, int64_t B, const double* q, const double* v, int nf, const int32_t* parent,
                 const double* placement, int rf, double* dJf, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv;
    static thread_local Mat<S> dJm;
    static thread_local std::vector<S> out;
    if (dJm.rows != 6 || dJm.cols != m.nv) dJm = Mat<S>(6, m.nv);
    out.resize((size_t)6 * m.nv);
    computeJointJacobiansTimeVariation(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * m.nv, m.nv), dJm);
    for (int f = 0; f < nf; ++f) {
      if (parent[f] < 0 || parent[f] >= m.njoints) throw std::invalid_argument("frame parent out of range");
      SE3<S> X;
      for (int k = 0; k < 9; ++k) X.R.m[k] = S(placement[12 * f + k]);
      for (int k = 0; k < 3; ++k) X.p[k] = S(placement[12 * f + 9 + k]);
      frameJacobianTimeVariation(m, data, dJm, parent[f], X, rf, out.data());
      store(out.data(), dJf + (b * nf + f) * 6 * m.nv, 6 * m.nv);
    }
  });
}
template<class S>
int framesT(const ModelDesc* d, int64_t B, const double* q, int nf, const int32_t* parent, const double* placement,
            int rf, double* oMf, double* J, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq;
    static thread_local std::vector<S> Jf;
    computeJointJacobians(m, data, bq.load(q + b * m.nq, m.nq));
    Jf.resize((size_t)6 * m.nv);
    for (int f = 0; f < nf; ++f) {
      if (parent[f] < 0 || parent[f] >= m.njoints) throw std::invalid_argument("frame parent out of range");
      SE3<S> X;
      for (int k = 0; k < 9; ++k) X.R.m[k] = S(placement[12 * f + k]);
      for (int k = 0; k < 3; ++k) X.p[k] = S(placement[12 * f + 9 + k]);
      SE3<S> oMfr;
      frameJacobian(m, data, parent[f], X, rf, oMfr, Jf.data());
      if (oMf) {
        double* o = oMf + (b * nf + f) * 12;
        for (int k = 0; k < 9; ++k) o[k] = double(oMfr.R.m[k]);
        for (int k = 0; k < 3; ++k) o[9 + k] = double(oMfr.p[k]);
      }
      if (J) store(Jf.data(), J + (b * nf + f) * 6 * m.nv, 6 * m.nv);
    }
  });
}
template<class S>
int frameMotionsT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a, int nf,
                  const int32_t* parent, const double* placement, int rf, double* vf, double* af, double* acf, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv, ba;
    const S* aa = a ? ba.load(a + b * m.nv, m.nv) : nullptr;
    forwardKinematics(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * m.nv, m.nv), aa);
    for (int f = 0; f < nf; ++f) {
      if (parent[f] < 0 || parent[f] >= m.njoints) throw std::invalid_argument("frame parent out of range");
      SE3<S> X;
      for (int k = 0; k < 9; ++k) X.R.m[k] = S(placement[12 * f + k]);
      for (int k = 0; k < 3; ++k) X.p[k] = S(placement[12 * f + 9 + k]);
      Motion<S> vel, acc, cl;
      frameMotions(data, parent[f], X, rf, a != nullptr, vel, acc, cl);
      S tmp[6];
      toVec(vel, tmp);
      store(tmp, vf + (b * nf + f) * 6, 6);
      if (a) {
        toVec(acc, tmp);
        store(tmp, af + (b * nf + f) * 6, 6);
        toVec(cl, tmp);
        store(tmp, acf + (b * nf + f) * 6, 6);
      }
    }
  });
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_dims(const ModelDesc* d, int* nq, int* nv, int* compact) {
  try {
    const Model<double> m(*d);
    *nq = m.nq;
    *nv = m.nv;
    if (compact) *compact = m.isCompact() ? 1 : 0;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Topology helpers as the reference's Data holds them (data.hxx:196-315).
int orc_topology(const ModelDesc* d, int* idx_q, int* idx_v, int* nv_subtree, int* parents_fromRow,
                 int* nvSubtree_fromRow) {
  try {
    const Model<double> m(*d);
    for (int i = 0; i < m.njoints; ++i) {
      idx_q[i] = m.idx_q[i];
      idx_v[i] = m.idx_v[i];
      nv_subtree[i] = m.nvSubtree[i];
    }
    for (int k = 0; k < m.nv; ++k) {
      parents_fromRow[k] = m.parents_fromRow[k];
      nvSubtree_fromRow[k] = m.nvSubtree_fromRow[k];
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

int orc_rnea(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a, double* tau, int scalar,
             int nthreads) {
#define CALL(S) rneaT<S>(d, B, q, v, a, tau, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_forward_kinematics(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a,
                           double* oM, double* vout, double* aout, int scalar, int nthreads) {
#define CALL(S) fkT<S>(d, B, q, v, a, oM, vout, aout, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_frames(const ModelDesc* d, int64_t B, const double* q, int nf, const int32_t* parent,
               const double* placement, int rf, double* oMf, double* J, int scalar, int nthreads) {
#define CALL(S) framesT<S>(d, B, q, nf, parent, placement, rf, oMf, J, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_frame_motions(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a, int nf,
                      const int32_t* parent, const double* placement, int rf, double* vf, double* af, double* acf,
                      int scalar, int nthreads) {
#define CALL(S) frameMotionsT<S>(d, B, q, v, a, nf, parent, placement, rf, vf, af, acf, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_joint_jacobians_time_variation(const ModelDesc* d, int64_t B, const double* q, const double* v, double* J,
                                       double* dJ, int scalar, int nthreads) {
#define CALL(S) jacDotT<S>(d, B, q, v, J, dJ, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_frame_jacobians_time_variation(const ModelDesc* d, int64_t B, const double* q, const double* v, int nf,
                                       const int32_t* parent, const double* placement, int rf, double* dJf,
                                       int scalar, int nthreads) {
#define CALL(S) frameJacDotT<S>(d, B, q, v, nf, parent, placement, rf, dJf, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_joint_jacobians(const ModelDesc* d, int64_t B, const double* q, double* J, int scalar, int nthreads) {
#define CALL(S) jacT<S>(d, B, q, J, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_aba(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* tau, double* ddq,
            int convention, int scalar, int nthreads) {
#define CALL(S) abaT<S>(d, B, q, v, tau, ddq, convention, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_crba(const ModelDesc* d, int64_t B, const double* q, double* M, int symmetrize, int scalar, int nthreads) {
#define CALL(S) crbaT<S>(d, B, q, M, symmetrize, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

// M: full (or upper) nv x nv per config.  Any of U, D, Minv, x may be null; x is solved in place.
int orc_cholesky(const ModelDesc* d, int64_t B, const double* M, double* U, double* D, double* Minv, double* x,
                 int scalar, int nthreads) {
#define CALL(S) choleskyT<S>(d, B, M, U, D, Minv, x, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_rnea_derivatives(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* a,
                         double* dtau_dq, double* dtau_dv, double* M, double* tau, int scalar, int nthreads) {
#define CALL(S) rneaDerivT<S>(d, B, q, v, a, dtau_dq, dtau_dv, M, tau, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_aba_derivatives(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* tau,
                        double* ddq_dq, double* ddq_dv, double* Minv, double* ddq, int scalar, int nthreads) {
#define CALL(S) abaDerivT<S>(d, B, q, v, tau, ddq_dq, ddq_dv, Minv, ddq, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

int orc_integrate(const ModelDesc* d, int64_t B, const double* q, const double* v, double* qout) {
  return runBatch<double>(d, B, 1, [&](const Model<double>& m, Data<double>&, int64_t b) {
    integrate(m, q + b * m.nq, v + b * m.nv, qout + b * m.nq);
  });
}

// Kinetic energy 0.5 v^T M v and potential energy -sum m g.c (energy.hpp), per config.
int orc_energy(const ModelDesc* d, int64_t B, const double* q, const double* v, double* T, double* V, int scalar) {
  if (scalar == 2)
    return runBatch<long double>(d, B, 1, [&](const Model<long double>& m, Data<long double>& data, int64_t b) {
      Buf<long double> bq, bv;
      const long double* qq = bq.load(q + b * m.nq, m.nq);
      if (T) T[b] = double(kineticEnergy(m, data, qq, bv.load(v + b * m.nv, m.nv)));
      if (V) V[b] = double(potentialEnergy(m, data, qq));
    });
  return runBatch<double>(d, B, 1, [&](const Model<double>& m, Data<double>& data, int64_t b) {
    if (T) T[b] = kineticEnergy(m, data, q + b * m.nq, v + b * m.nv);
    if (V) V[b] = potentialEnergy(m, data, q + b * m.nq);
  });
}

}  // extern "C"

// ---- algorithmic operation counts (one configuration, flop_count.hpp) -----------------
#include "flop_count.hpp"

extern "C" int orc_flops(const ModelDesc* d, int alg, const double* q, const double* v, const double* x,
                         long long* out) {
  try {
    using F = oracle::Flop;
    const Model<F> m(*d);
    Data<F> data(m);
    std::vector<F> qq(q, q + m.nq), vv(v, v + m.nv), xx(x, x + m.nv), res(m.nv);
    long long& n = F::count();
    n = 0;
    switch (alg) {
      case 0: rnea(m, data, qq.data(), vv.data(), xx.data(), res.data()); break;
      case 1: abaWorld(m, data, qq.data(), vv.data(), xx.data(), res.data()); break;
      case 2: abaLocal(m, data, qq.data(), vv.data(), xx.data(), res.data()); break;
      case 3: crbaWorld(m, data, qq.data()); break;
      case 4: {
        crbaWorld(m, data, qq.data());
        n = 0;
        choleskyDecompose(m, data);
        break;
      }
      case 5: rneaDerivatives(m, data, qq.data(), vv.data(), xx.data()); break;
      case 6: {
        Mat<F> a(m.nv, m.nv), b(m.nv, m.nv), c(m.nv, m.nv);
        abaDerivatives(m, data, qq.data(), vv.data(), xx.data(), a, b, c);
        break;
      }
      case 7: {
        crbaWorld(m, data, qq.data());
        choleskyDecompose(m, data);
        n = 0;
        choleskySolve(m, data, xx.data());
        break;
      }
      default: throw std::invalid_argument("unknown algorithm id");
    }
    *out = n;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// ---- CPU baseline timing: the batch loop of rneaInParallel / abaInParallel, timed in C --
// alg: 0 rnea, 1 aba (WORLD), 2 crba, 3 crba + cholesky::decompose, 4 aba + crba +
// decompose (one step of the headline workload), 5 computeRNEADerivatives,
// 6 computeABADerivatives.  Each thread owns one Data (ModelPool semantics) and writes
// results into it; inputs are the caller's column-major double arrays.
#include <chrono>
namespace {
template<class S>
int cfdT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* tau, int nc, const double* J,
         const double* gamma, double damping, double* ddq, double* lambda, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    if (!m.isCompact()) throw std::invalid_argument("constrained dynamics requires a compact (DFS-ordered) model");
    static thread_local Buf<S> bq, bv, bt, bJ, bg;
    static thread_local std::vector<S> a, l;
    a.resize(m.nv);
    l.resize(nc);
    constrainedForwardDynamics(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * m.nv, m.nv),
                               bt.load(tau + b * m.nv, m.nv), nc, bJ.load(J + b * nc * m.nv, nc * m.nv),
                               bg.load(gamma + b * nc, nc), S(damping), a.data(), l.data());
    store(a.data(), ddq + b * m.nv, m.nv);
    store(l.data(), lambda + b * nc, nc);
  });
}
}  // namespace

extern "C" int orc_constrained_forward_dynamics(const ModelDesc* d, int64_t B, const double* q, const double* v,
                                                const double* tau, int nc, const double* J, const double* gamma,
                                                double damping, double* ddq, double* lambda, int scalar,
                                                int nthreads) {
  if (nc < 1) {
    g_err = "nc must be positive";
    return 1;
  }
#define CALL(S) cfdT<S>(d, B, q, v, tau, nc, J, gamma, damping, ddq, lambda, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

namespace {
template<class S>
int impT(const ModelDesc* d, int64_t B, const double* q, const double* v, int nc, const double* J, double rc,
         double damping, double* dqa, double* imp, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    if (!m.isCompact()) throw std::invalid_argument("impulse dynamics requires a compact (DFS-ordered) model");
    static thread_local Buf<S> bq, bv, bJ;
    static thread_local std::vector<S> a, l;
    a.resize(m.nv);
    l.resize(nc);
    impulseDynamics(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * m.nv, m.nv), nc,
                    bJ.load(J + b * nc * m.nv, nc * m.nv), S(rc), S(damping), a.data(), l.data());
    store(a.data(), dqa + b * m.nv, m.nv);
    store(l.data(), imp + b * nc, nc);
  });
}
}  // namespace

extern "C" int orc_impulse_dynamics(const ModelDesc* d, int64_t B, const double* q, const double* v, int nc,
                                    const double* J, double r_coeff, double damping, double* dq_after,
                                    double* impulse, int scalar, int nthreads) {
  if (nc < 1) {
    g_err = "nc must be positive";
    return 1;
  }
#define CALL(S) impT<S>(d, B, q, v, nc, J, r_coeff, damping, dq_after, impulse, nthreads)
  switch (scalar) {
    case 0: return CALL(double);
    case 1: return CALL(float);
    case 2: return CALL(long double);
  }
#undef CALL
  g_err = "bad scalar";
  return 1;
}

// median wall time of the OpenMP batch loop of constrainedForwardDynamics (FP64)
extern "C" int orc_time_constrained(const ModelDesc* d, int64_t B, const double* q, const double* v,
                                    const double* tau, int nc, const double* J, const double* gamma, double damping,
                                    int nthreads, int reps, double* seconds) {
  try {
    const Model<double> m(*d);
    if (!m.isCompact()) throw std::invalid_argument("model is not compact");
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    std::vector<double> times;
    for (int r = 0; r < reps; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nt)
      {
        Data<double> data(m);
        std::vector<double> a(m.nv), l(nc);
#pragma omp for schedule(static)
        for (int64_t b = 0; b < B; ++b)
          constrainedForwardDynamics(m, data, q + b * m.nq, v + b * m.nv, tau + b * m.nv, nc, J + b * nc * m.nv,
                                     gamma + b * nc, damping, a.data(), l.data());
      }
      times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(times.begin(), times.end());
    *seconds = times[times.size() / 2];
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

extern "C" int orc_time(const ModelDesc* d, int alg, int64_t B, const double* q, const double* v, const double* x,
                        int nthreads, int reps, double* seconds) {
  try {
    const Model<double> m(*d);
    if ((alg >= 2 && alg <= 6) && !m.isCompact()) throw std::invalid_argument("model is not compact");
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    std::vector<double> times;
    for (int r = 0; r < reps; ++r) {
      const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nt)
      {
        Data<double> data(m);
        std::vector<double> out(m.nv);
        Mat<double> A(m.nv, m.nv), Bm(m.nv, m.nv), C(m.nv, m.nv);
#pragma omp for schedule(static)
        for (int64_t b = 0; b < B; ++b) {
          const double* qb = q + b * m.nq;
          const double* vb = v + b * m.nv;
          const double* xb = x + b * m.nv;
          switch (alg) {
            case 0: rnea(m, data, qb, vb, xb, out.data()); break;
            case 1: abaWorld(m, data, qb, vb, xb, out.data()); break;
            case 2: crbaWorld(m, data, qb); break;
            case 3: crbaWorld(m, data, qb); choleskyDecompose(m, data); break;
            case 4:
              abaWorld(m, data, qb, vb, xb, out.data());
              crbaWorld(m, data, qb);
              choleskyDecompose(m, data);
              break;
            case 5: rneaDerivatives(m, data, qb, vb, xb); break;
            case 6: abaDerivatives(m, data, qb, vb, xb, A, Bm, C); break;
          }
        }
      }
      times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(times.begin(), times.end());
    *seconds = times[times.size() / 2];
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// ---- constraintDynamics with rigid constraints --------------------------------------------
// One rigid constraint as the C API receives it (the same layout as the engine's
// rbd_rigid_constraint): type (0: CONTACT_3D, 1: CONTACT_6D), joint1, joint2, reference
// frame (1: LOCAL, 2: LOCAL_WORLD_ALIGNED), the two joint placements (R row major, p),
// the corrector gains and the desired contact acceleration.
struct OrcConstraint {
  int32_t type, joint1, joint2, rf;
  double placement1[12], placement2[12];
  double Kp[6], Kd[6], desired[6];
};

namespace {
template<class S> std::vector<RigidConstraint<S>> toConstraints(int ncon, const OrcConstraint* oc, int& nrows) {
  std::vector<RigidConstraint<S>> cons(ncon);
  nrows = 0;
  for (int c = 0; c < ncon; ++c) {
    RigidConstraint<S>& r = cons[c];
    r.type = oc[c].type;
    r.joint1 = oc[c].joint1;
    r.joint2 = oc[c].joint2;
    r.rf = oc[c].rf;
    if (r.type != CONTACT_3D && r.type != CONTACT_6D) throw std::invalid_argument("bad constraint type");
    for (int k = 0; k < 9; ++k) {
      r.placement1.R.m[k] = S(oc[c].placement1[k]);
      r.placement2.R.m[k] = S(oc[c].placement2[k]);
    }
    for (int k = 0; k < 3; ++k) {
      r.placement1.p[k] = S(oc[c].placement1[9 + k]);
      r.placement2.p[k] = S(oc[c].placement2[9 + k]);
    }
    for (int k = 0; k < 6; ++k) {
      r.Kp[k] = S(oc[c].Kp[k]);
      r.Kd[k] = S(oc[c].Kd[k]);
      r.desired[k] = S(oc[c].desired[k]);
    }
    nrows += r.size();
  }
  return cons;
}

template<class S>
int cdT(const ModelDesc* d, int64_t B, const double* q, const double* v, const double* tau, int ncon,
        const OrcConstraint* oc, double absAcc, double relAcc, double mu, int maxIter, double* ddq, double* lambda,
        int32_t* iters, int nt) {
  int nrows = 0;
  const std::vector<RigidConstraint<S>> cons = toConstraints<S>(ncon, oc, nrows);
  const ProximalSettings<S> ps{S(absAcc), S(relAcc), S(mu), maxIter};
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    for (const auto& c : cons)
      if (c.joint1 < 0 || c.joint1 >= m.njoints || c.joint2 < 0 || c.joint2 >= m.njoints)
        throw std::invalid_argument("constraint joint out of range");
    static thread_local Buf<S> bq, bv, bt;
    static thread_local std::vector<S> a, l;
    a.resize(m.nv);
    l.resize(nrows);
    int it = 0;
    constraintDynamics(m, data, bq.load(q + b * m.nq, m.nq), bv.load(v + b * m.nv, m.nv),
                       bt.load(tau + b * m.nv, m.nv), ncon, cons.data(), ps, a.data(), l.data(), &it);
    store(a.data(), ddq + b * m.nv, m.nv);
    store(l.data(), lambda + b * nrows, nrows);
    if (iters) iters[b] = it;
  });
}
}  // namespace

extern "C" int orc_constraint_dynamics(const ModelDesc* d, int64_t B, const double* q, const double* v,
                                       const double* tau, int ncon, const OrcConstraint* cons, double absolute_accuracy,
                                       double relative_accuracy, double mu, int max_iter, double* ddq, double* lambda,
                                       int32_t* iterations, int scalar, int nthreads) {
  try {
    if (ncon < 1 || !cons) throw std::invalid_argument("at least one constraint");
    if (max_iter < 1) throw std::invalid_argument("max_iter must be positive");
#define CALL(S) \
  cdT<S>(d, B, q, v, tau, ncon, cons, absolute_accuracy, relative_accuracy, mu, max_iter, ddq, lambda, iterations, nthreads)
    switch (scalar) {
      case 0: return CALL(double);
      case 1: return CALL(float);
      case 2: return CALL(long double);
    }
#undef CALL
    g_err = "bad scalar";
    return 1;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// computeDelassusMatrix per configuration: G (nrows x nrows, column major) back to back
namespace {
template<class S>
int delT(const ModelDesc* d, int64_t B, const double* q, int ncon, const OrcConstraint* oc, double mu, double* G,
         int nt) {
  int nrows = 0;
  const std::vector<RigidConstraint<S>> cons = toConstraints<S>(ncon, oc, nrows);
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    for (const auto& c : cons)
      if (c.joint1 < 0 || c.joint1 >= m.njoints || c.joint2 < 0 || c.joint2 >= m.njoints)
        throw std::invalid_argument("constraint joint out of range");
    static thread_local Buf<S> bq;
    static thread_local std::vector<S> g;
    g.resize((size_t)nrows * nrows);
    delassusMatrix(m, data, bq.load(q + b * m.nq, m.nq), ncon, cons.data(), S(mu), g.data());
    store(g.data(), G + b * nrows * nrows, nrows * nrows);
  });
}
}  // namespace

extern "C" int orc_delassus(const ModelDesc* d, int64_t B, const double* q, int ncon, const OrcConstraint* cons,
                            double mu, double* G, int scalar, int nthreads) {
  try {
    if (ncon < 1 || !cons) throw std::invalid_argument("at least one constraint");
    switch (scalar) {
      case 0: return delT<double>(d, B, q, ncon, cons, mu, G, nthreads);
      case 1: return delT<float>(d, B, q, ncon, cons, mu, G, nthreads);
      case 2: return delT<long double>(d, B, q, ncon, cons, mu, G, nthreads);
    }
    g_err = "bad scalar";
    return 1;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// Frictional contact solvers over B independent problems of nc cones (n = 3 nc), each
// problem's arrays back to back: G n x n column major, g and x / y / z / s n, mu nc, R n.
// PGS: x in/out; it, residuals (absolute, relative) and converged per problem.
extern "C" int orc_pgs_solve(int64_t B, int nc, const double* G, const double* g, const double* mu, double* x,
                             double over_relax, int max_it, double absolute_precision, double relative_precision,
                             int32_t* it, double* residuals, uint8_t* converged, int nthreads) {
  try {
    if (nc < 1) throw std::invalid_argument("at least one cone");
    if (max_it < 1) throw std::invalid_argument("max_it should be greater than 0.");
    const int n = 3 * nc;
    std::string err;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int64_t b = 0; b < B; ++b) {
      try {
        const auto r = orc::pgsSolve<double>(nc, G + b * n * n, g + b * n, mu + b * nc, x + b * n, over_relax, max_it,
                                             absolute_precision, relative_precision);
        if (it) it[b] = r.it;
        if (residuals) {
          residuals[2 * b] = r.absolute_residual;
          residuals[2 * b + 1] = r.relative_residual;
        }
        if (converged) converged[b] = r.converged ? 1 : 0;
      } catch (const std::exception& e) {
#pragma omp critical
        err = e.what();
      }
    }
    if (!err.empty()) throw std::invalid_argument(err);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// ADMM: settings = {mu_prox, tau, rho_power, rho_power_factor, ratio_primal_dual,
// absolute_precision, relative_precision}; primal_guess / dual_guess may be null; outputs
// y (primal), z (dual), s (shift), it, stats (absolute residual, relative residual, rho,
// rho_power, largest eigenvalue) and the factor count per problem.
extern "C" int orc_admm_solve(int64_t B, int nc, const double* G, const double* g, const double* mu, const double* R,
                              const double* primal_guess, const double* dual_guess, const double* settings,
                              int max_it, int max_it_eig, double* y, double* z, double* s, int32_t* it,
                              double* stats, int32_t* chol_updates, uint8_t* converged, int nthreads) {
  try {
    if (nc < 1) throw std::invalid_argument("at least one cone");
    if (max_it < 1) throw std::invalid_argument("max_it should be greater than 0.");
    const int n = 3 * nc;
    orc::AdmmSettings<double> st;
    st.mu_prox = settings[0];
    st.tau = settings[1];
    st.rho_power = settings[2];
    st.rho_power_factor = settings[3];
    st.ratio_primal_dual = settings[4];
    st.absolute_precision = settings[5];
    st.relative_precision = settings[6];
    st.max_it = max_it;
    st.max_it_largest_eigen_value = max_it_eig;
    std::string err;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int64_t b = 0; b < B; ++b) {
      try {
        const auto r = orc::admmSolve<double>(nc, G + b * n * n, g + b * n, mu + b * nc, R + b * n,
                                              primal_guess ? primal_guess + b * n : nullptr,
                                              dual_guess ? dual_guess + b * n : nullptr, st, y + b * n, z + b * n,
                                              s ? s + b * n : nullptr);
        if (it) it[b] = r.it;
        if (stats) {
          double* o = stats + 5 * b;
          o[0] = r.absolute_residual;
          o[1] = r.relative_residual;
          o[2] = r.rho;
          o[3] = r.rho_power;
          o[4] = r.largest_eigen_value;
        }
        if (chol_updates) chol_updates[b] = r.cholesky_update_count;
        if (converged) converged[b] = r.converged ? 1 : 0;
      } catch (const std::exception& e) {
#pragma omp critical
        err = e.what();
      }
    }
    if (!err.empty()) throw std::invalid_argument(err);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// ContactCholeskyDecomposition compute + solve per configuration (dense restatement): rhs and
// x hold nrows + nv entries per configuration, back to back
namespace {
template<class S>
int kktT(const ModelDesc* d, int64_t B, const double* q, int ncon, const OrcConstraint* oc, double mu,
         const double* rhs, double* x, int nt) {
  int nrows = 0;
  const std::vector<RigidConstraint<S>> cons = toConstraints<S>(ncon, oc, nrows);
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    for (const auto& c : cons)
      if (c.joint1 < 0 || c.joint1 >= m.njoints || c.joint2 < 0 || c.joint2 >= m.njoints)
        throw std::invalid_argument("constraint joint out of range");
    const int n = nrows + m.nv;
    static thread_local Buf<S> bq, br;
    static thread_local std::vector<S> xs;
    xs.resize(n);
    contactKktSolve(m, data, bq.load(q + b * m.nq, m.nq), ncon, cons.data(), S(mu), br.load(rhs + b * n, n),
                    xs.data());
    store(xs.data(), x + b * n, n);
  });
}
}  // namespace

extern "C" int orc_contact_cholesky_solve(const ModelDesc* d, int64_t B, const double* q, int ncon,
                                          const OrcConstraint* cons, double mu, const double* rhs, double* x,
                                          int scalar, int nthreads) {
  try {
    if (ncon < 1 || !cons) throw std::invalid_argument("at least one constraint");
    switch (scalar) {
      case 0: return kktT<double>(d, B, q, ncon, cons, mu, rhs, x, nthreads);
      case 1: return kktT<float>(d, B, q, ncon, cons, mu, rhs, x, nthreads);
      case 2: return kktT<long double>(d, B, q, ncon, cons, mu, rhs, x, nthreads);
    }
    g_err = "bad scalar";
    return 1;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}

// RNEA and ABA (LOCAL) with external forces: fext holds 6 (njoints - 1) numbers per
// configuration, back to back (joint 1.., [linear; angular] in the joint's frame)
namespace {
template<class S>
int fextT(const ModelDesc* d, int64_t B, int alg, const double* q, const double* v, const double* x,
          const double* fext, double* out, int nt) {
  return runBatch<S>(d, B, nt, [&](const Model<S>& m, Data<S>& data, int64_t b) {
    static thread_local Buf<S> bq, bv, bx, bf;
    static thread_local std::vector<S> o;
    o.resize(m.nv);
    const int nf = 6 * (m.njoints - 1);
    const S* qq = bq.load(q + b * m.nq, m.nq);
    const S* vv = bv.load(v + b * m.nv, m.nv);
    const S* xx = bx.load(x + b * m.nv, m.nv);
    const S* ff = bf.load(fext + b * nf, nf);
    if (alg == 0) rnea(m, data, qq, vv, xx, o.data(), ff);
    else abaLocal(m, data, qq, vv, xx, o.data(), ff);
    store(o.data(), out + b * m.nv, m.nv);
  });
}
}  // namespace

extern "C" int orc_fext(const ModelDesc* d, int64_t B, int alg, const double* q, const double* v, const double* x,
                        const double* fext, double* out, int scalar, int nthreads) {
  switch (scalar) {
    case 0: return fextT<double>(d, B, alg, q, v, x, fext, out, nthreads);
    case 1: return fextT<float>(d, B, alg, q, v, x, fext, out, nthreads);
    case 2: return fextT<long double>(d, B, alg, q, v, x, fext, out, nthreads);
  }
  g_err = "bad scalar";
  return 1;
}
