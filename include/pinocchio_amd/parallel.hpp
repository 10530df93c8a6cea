// pinocchio_amd/parallel.hpp -- C++ drop-in for pinocchio's batched dynamics calls.
//
// Mirrors, over the C ABI of rbd.h, the call shapes of
//   rneaInParallel(num_threads, pool, q, v, a, tau)   algorithm/parallel/rnea.hpp:39-83
//   abaInParallel (num_threads, pool, q, v, tau, a)   algorithm/parallel/aba.hpp:40-84
// and adds the batched counterparts of crba, computeRNEADerivatives and
// computeABADerivatives that the reference only offers per configuration
// (crba.hpp:47-52, rnea-derivatives.hpp:110-260, aba-derivatives.hpp:42-318).
//
// Matrices are column-major views with one configuration per column, exactly the
// memory of an Eigen::MatrixXd (pass m.data(), m.rows(), m.cols(), m.outerStride()).
// With Eigen available, the Eigen overloads at the bottom take the matrices directly.
// Argument errors throw std::invalid_argument for the reference's conditions
// (rnea.hpp:53-68, aba.hpp:54-69) plus the tau shape that abaInParallel leaves unchecked.
// Calls are synchronous; num_threads is validated against pool.size() and otherwise
// ignored (one GPU does the whole batch).  No global OpenMP state is touched.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../rbd.h"

namespace pinocchio_amd {

inline void check(int rc) {
  if (rc == RBD_OK) return;
  const std::string msg = rbd_last_error();
  if (rc == RBD_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// A column-major matrix view: element (r, c) at data[r + c * ld].
template<class S> struct ColView {
  S* data;
  int64_t rows, cols, ld;
};
using CMat = ColView<const double>;
using MMat = ColView<double>;

// Holds one model on one GPU: the counterpart of ModelPoolTpl (pool/model.hpp:19-165),
// whose per-thread Model/Data copies are not needed on the GPU.
class ModelPool {
 public:
  explicit ModelPool(const rbd_model_desc& desc, size_t pool_size = 1, int device = 0) : size_(pool_size) {
    if (pool_size == 0) throw std::invalid_argument("The pool size should be greater than zero");
    check(rbd_ctx_create(device, &ctx_));
    const int rc = rbd_model_create(ctx_, &desc, &model_);
    if (rc != RBD_OK) {
      rbd_ctx_destroy(ctx_);
      check(rc);
    }
    int32_t nq, nv, nj, compact;
    check(rbd_model_dims(model_, &nq, &nv, &nj, &compact));
    nq_ = nq;
    nv_ = nv;
  }
  ~ModelPool() {
    rbd_model_destroy(model_);
    rbd_ctx_destroy(ctx_);
  }
  ModelPool(const ModelPool&) = delete;
  ModelPool& operator=(const ModelPool&) = delete;

  size_t size() const { return size_; }
  int nq() const { return nq_; }
  int nv() const { return nv_; }
  rbd_ctx* ctx() const { return ctx_; }
  rbd_model* model() const { return model_; }

 private:
  size_t size_;
  rbd_ctx* ctx_ = nullptr;
  rbd_model* model_ = nullptr;
  int nq_ = 0, nv_ = 0;
};

namespace detail {
inline void checkThreads(size_t num_threads, const ModelPool& pool) {
  if (pool.size() == 0) throw std::invalid_argument("The pool should have at least one element");
  if (num_threads > pool.size()) throw std::invalid_argument("The number of threads should be lower than the pool size");
}
template<class V> void checkRows(const V& m, int64_t rows, const char* what) {
  if (m.rows != rows) throw std::invalid_argument(std::string(what) + " has wrong number of rows");
}
template<class V> void checkCols(const V& m, int64_t cols, const char* what) {
  if (m.cols != cols) throw std::invalid_argument(std::string(what) + " has wrong number of columns");
}
// Runs one batched vector op in FP32 or FP64 on host matrices.
template<class Fn>
void vecOp(Fn fn, const ModelPool& pool, bool fp32, const CMat& x0, const CMat& x1, const CMat& x2, const MMat& out) {
  const int64_t B = x0.cols;
  if (!fp32) {
    check(fn(pool.ctx(), pool.model(), B, RBD_F64, x0.data, x0.ld, x1.data, x1.ld, x2.data, x2.ld, out.data, out.ld,
             RBD_HOST_PTRS));
    return;
  }
  // FP32 compute: round the inputs, widen the result
  auto pack = [B](const CMat& m) {
    std::vector<float> f((size_t)(m.rows * B));
    for (int64_t c = 0; c < B; ++c)
      for (int64_t r = 0; r < m.rows; ++r) f[(size_t)(r + c * m.rows)] = (float)m.data[r + c * m.ld];
    return f;
  };
  std::vector<float> f0 = pack(x0), f1 = pack(x1), f2 = pack(x2), fo((size_t)(out.rows * B));
  check(fn(pool.ctx(), pool.model(), B, RBD_F32, f0.data(), x0.rows, f1.data(), x1.rows, f2.data(), x2.rows,
           fo.data(), out.rows, RBD_HOST_PTRS));
  for (int64_t c = 0; c < B; ++c)
    for (int64_t r = 0; r < out.rows; ++r) out.data[r + c * out.ld] = fo[(size_t)(r + c * out.rows)];
}
}  // namespace detail

// tau.col(i) = rnea(q.col(i), v.col(i), a.col(i)) for every column (FP64 by default).
inline void rneaInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const CMat& v, const CMat& a,
                           const MMat& tau, bool fp32 = false) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(v, pool.nv(), "The velocity matrix");
  detail::checkRows(a, pool.nv(), "The acceleration matrix");
  detail::checkRows(tau, pool.nv(), "The output matrix");
  detail::checkCols(v, q.cols, "The velocity matrix");
  detail::checkCols(a, q.cols, "The acceleration matrix");
  detail::checkCols(tau, q.cols, "The output matrix");
  detail::vecOp(&rbd_rnea, pool, fp32, q, v, a, tau);
}

// a.col(i) = aba(q.col(i), v.col(i), tau.col(i)) for every column (FP64 by default).
inline void abaInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const CMat& v, const CMat& tau,
                          const MMat& a, bool fp32 = false) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(v, pool.nv(), "The velocity matrix");
  detail::checkRows(tau, pool.nv(), "The torque matrix");
  detail::checkRows(a, pool.nv(), "The output matrix");
  detail::checkCols(v, q.cols, "The velocity matrix");
  detail::checkCols(tau, q.cols, "The torque matrix");
  detail::checkCols(a, q.cols, "The output matrix");
  detail::vecOp(&rbd_aba, pool, fp32, q, v, tau, a);
}

// M_i = crba(q.col(i)) for every column, written as nv x nv blocks side by side in M
// (nv rows, nv * B columns), symmetric.
inline void crbaInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const MMat& M) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(M, pool.nv(), "The output matrix");
  detail::checkCols(M, q.cols * pool.nv(), "The output matrix");
  if (M.ld != pool.nv()) throw std::invalid_argument("The output matrix must be contiguous");
  check(rbd_crba(pool.ctx(), pool.model(), q.cols, RBD_F64, q.data, q.ld, M.data, (int64_t)pool.nv() * pool.nv(),
                 RBD_HOST_PTRS));
}

namespace detail {
inline std::vector<float> toF32(const CMat& m) {
  std::vector<float> f((size_t)(m.rows * m.cols));
  for (int64_t c = 0; c < m.cols; ++c)
    for (int64_t r = 0; r < m.rows; ++r) f[(size_t)(r + c * m.rows)] = (float)m.data[r + c * m.ld];
  return f;
}
inline void fromF32(const std::vector<float>& f, const MMat& m) {
  for (int64_t c = 0; c < m.cols; ++c)
    for (int64_t r = 0; r < m.rows; ++r) m.data[r + c * m.ld] = f[(size_t)(r + c * m.rows)];
}
// Runs a derivative op (rbd_rnea_derivatives / rbd_aba_derivatives) whose three outputs
// are nv x (nv B) blocks side by side.  FP64: straight into the caller's matrices when
// they are contiguous (ld = nv), else through packed copies; FP32: inputs rounded, outputs
// widened.
template<class Fn>
void derivOp(Fn fn, const ModelPool& pool, bool fp32, const CMat& x0, const CMat& x1, const CMat& x2, const MMat& o1,
             const MMat& o2, const MMat& o3) {
  const int64_t nv = pool.nv(), nn = nv * nv, B = x0.cols;
  if (!fp32) {
    if (o1.ld == nv && o2.ld == nv && o3.ld == nv) {
      check(fn(pool.ctx(), pool.model(), B, RBD_F64, x0.data, x0.ld, x1.data, x1.ld, x2.data, x2.ld, o1.data, o2.data,
               o3.data, nn, RBD_HOST_PTRS));
      return;
    }
    std::vector<double> d1((size_t)(nn * B)), d2(d1.size()), d3(d1.size());
    check(fn(pool.ctx(), pool.model(), B, RBD_F64, x0.data, x0.ld, x1.data, x1.ld, x2.data, x2.ld, d1.data(),
             d2.data(), d3.data(), nn, RBD_HOST_PTRS));
    for (auto [d, m] : {std::pair{&d1, &o1}, std::pair{&d2, &o2}, std::pair{&d3, &o3}})
      for (int64_t c = 0; c < m->cols; ++c)
        for (int64_t r = 0; r < nv; ++r) m->data[r + c * m->ld] = (*d)[(size_t)(r + c * nv)];
    return;
  }
  std::vector<float> f0 = toF32(x0), f1 = toF32(x1), f2 = toF32(x2);
  std::vector<float> e1((size_t)(nn * B)), e2(e1.size()), e3(e1.size());
  check(fn(pool.ctx(), pool.model(), B, RBD_F32, f0.data(), x0.rows, f1.data(), x1.rows, f2.data(), x2.rows,
           e1.data(), e2.data(), e3.data(), nn, RBD_HOST_PTRS));
  fromF32(e1, o1);
  fromF32(e2, o2);
  fromF32(e3, o3);
}
}  // namespace detail

// computeMinverse per column: M(q)^-1 as nv x nv blocks side by side (FP64 by default).
inline void computeMinverseInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const MMat& Minv,
                                      bool fp32 = false) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(Minv, pool.nv(), "The output matrix");
  detail::checkCols(Minv, q.cols * pool.nv(), "The output matrix");
  if (Minv.ld != pool.nv()) throw std::invalid_argument("The output matrix must be contiguous");
  const int64_t nn = (int64_t)pool.nv() * pool.nv();
  if (!fp32) {
    check(rbd_compute_minverse(pool.ctx(), pool.model(), q.cols, RBD_F64, q.data, q.ld, Minv.data, nn,
                               RBD_HOST_PTRS));
    return;
  }
  std::vector<float> fq = detail::toF32(q), fo((size_t)(nn * q.cols));
  check(rbd_compute_minverse(pool.ctx(), pool.model(), q.cols, RBD_F32, fq.data(), q.rows, fo.data(), nn,
                             RBD_HOST_PTRS));
  detail::fromF32(fo, Minv);
}

// forwardDynamics(model, data, q, v, tau, J, gamma, inv_damping) per column
// (contact-dynamics.hxx:90-167): J is nc x (nv * B), the configurations' nc x nv constraint
// Jacobians side by side (contiguous, ld = nc); gamma and lambda are nc x B, ddq nv x B.
// FP64 throughout.
inline void forwardDynamicsInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const CMat& v,
                                      const CMat& tau, const CMat& J, const CMat& gamma, double inv_damping,
                                      const MMat& ddq, const MMat& lambda) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(v, pool.nv(), "The velocity matrix");
  detail::checkRows(tau, pool.nv(), "The joint torque matrix");
  detail::checkRows(ddq, pool.nv(), "The output matrix");
  detail::checkCols(v, q.cols, "The velocity matrix");
  detail::checkCols(tau, q.cols, "The joint torque matrix");
  detail::checkCols(ddq, q.cols, "The output matrix");
  const int nc = (int)J.rows;
  detail::checkCols(J, q.cols * pool.nv(), "The constraint Jacobian");
  if (J.ld != J.rows) throw std::invalid_argument("The constraint Jacobian must be contiguous");
  detail::checkRows(gamma, nc, "The constraint drift");
  detail::checkCols(gamma, q.cols, "The constraint drift");
  detail::checkRows(lambda, nc, "The multiplier matrix");
  detail::checkCols(lambda, q.cols, "The multiplier matrix");
  check(rbd_constrained_forward_dynamics(pool.ctx(), pool.model(), q.cols, RBD_F64, q.data, q.ld, v.data, v.ld,
                                         tau.data, tau.ld, nc, J.data, (int64_t)nc * pool.nv(), gamma.data,
                                         gamma.ld, inv_damping, ddq.data, ddq.ld, lambda.data, lambda.ld,
                                         RBD_HOST_PTRS));
}

// constraintDynamics(model, data, q, v, tau, contact_models, contact_datas, settings) per
// column (constrained-dynamics.hxx:193-500): the constraints as rbd_rigid_constraint, the
// same for every column; lambda has the constraints' total dimension in rows (at most 12);
// iterations (optional, B entries) receives settings.iter per column.  FP64 throughout.
inline void constraintDynamicsInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const CMat& v,
                                         const CMat& tau, const std::vector<rbd_rigid_constraint>& contacts,
                                         const rbd_proximal_settings& settings, const MMat& ddq, const MMat& lambda,
                                         int32_t* iterations = nullptr) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(v, pool.nv(), "The velocity matrix");
  detail::checkRows(tau, pool.nv(), "The joint torque matrix");
  detail::checkRows(ddq, pool.nv(), "The output matrix");
  detail::checkCols(v, q.cols, "The velocity matrix");
  detail::checkCols(tau, q.cols, "The joint torque matrix");
  detail::checkCols(ddq, q.cols, "The output matrix");
  int nrows = 0;
  for (const auto& c : contacts) nrows += c.type == RBD_CONTACT_6D ? 6 : 3;
  detail::checkRows(lambda, nrows, "The multiplier matrix");
  detail::checkCols(lambda, q.cols, "The multiplier matrix");
  check(rbd_constraint_dynamics(pool.ctx(), pool.model(), q.cols, RBD_F64, q.data, q.ld, v.data, v.ld, tau.data,
                                tau.ld, (int)contacts.size(), contacts.data(), &settings, ddq.data, ddq.ld,
                                lambda.data, lambda.ld, iterations, RBD_HOST_PTRS));
}

// computeRNEADerivatives per column: three nv x (nv * B) outputs (blocks side by side),
// FP64 by default.
inline void computeRNEADerivativesInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const CMat& v,
                                             const CMat& a, const MMat& dtau_dq, const MMat& dtau_dv,
                                             const MMat& dtau_da, bool fp32 = false) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(v, pool.nv(), "The velocity matrix");
  detail::checkRows(a, pool.nv(), "The acceleration matrix");
  detail::checkCols(v, q.cols, "The velocity matrix");
  detail::checkCols(a, q.cols, "The acceleration matrix");
  for (const MMat* o : {&dtau_dq, &dtau_dv, &dtau_da}) {
    detail::checkRows(*o, pool.nv(), "An output matrix");
    detail::checkCols(*o, q.cols * pool.nv(), "An output matrix");
  }
  detail::derivOp(&rbd_rnea_derivatives, pool, fp32, q, v, a, dtau_dq, dtau_dv, dtau_da);
}

// computeABADerivatives per column: ddq_dq, ddq_dv, ddq_dtau (= Minv) as nv x (nv * B),
// FP64 by default.
inline void computeABADerivativesInParallel(size_t num_threads, ModelPool& pool, const CMat& q, const CMat& v,
                                            const CMat& tau, const MMat& ddq_dq, const MMat& ddq_dv,
                                            const MMat& ddq_dtau, bool fp32 = false) {
  detail::checkThreads(num_threads, pool);
  detail::checkRows(q, pool.nq(), "The configuration matrix");
  detail::checkRows(v, pool.nv(), "The velocity matrix");
  detail::checkRows(tau, pool.nv(), "The torque matrix");
  detail::checkCols(v, q.cols, "The velocity matrix");
  detail::checkCols(tau, q.cols, "The torque matrix");
  for (const MMat* o : {&ddq_dq, &ddq_dv, &ddq_dtau}) {
    detail::checkRows(*o, pool.nv(), "An output matrix");
    detail::checkCols(*o, q.cols * pool.nv(), "An output matrix");
  }
  auto fn = [](rbd_ctx* c, const rbd_model* m, int64_t B, int dt, const void* x0, int64_t l0, const void* x1,
               int64_t l1, const void* x2, int64_t l2, void* o1, void* o2, void* o3, int64_t lo, uint32_t fl) {
    return rbd_aba_derivatives(c, m, B, dt, x0, l0, x1, l1, x2, l2, o1, o2, o3, lo, nullptr, 0, fl);
  };
  detail::derivOp(fn, pool, fp32, q, v, tau, ddq_dq, ddq_dv, ddq_dtau);
}

}  // namespace pinocchio_amd

#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
namespace pinocchio_amd {
template<class D> CMat view(const Eigen::MatrixBase<D>& m) {
  return CMat{m.derived().data(), m.rows(), m.cols(), m.derived().outerStride()};
}
template<class D> MMat mview(const Eigen::MatrixBase<D>& m) {
  auto& mm = const_cast<Eigen::MatrixBase<D>&>(m).derived();
  return MMat{mm.data(), mm.rows(), mm.cols(), mm.outerStride()};
}
template<class Q, class V, class A, class T>
void rneaInParallel(size_t n, ModelPool& pool, const Eigen::MatrixBase<Q>& q, const Eigen::MatrixBase<V>& v,
                    const Eigen::MatrixBase<A>& a, const Eigen::MatrixBase<T>& tau) {
  rneaInParallel(n, pool, view(q), view(v), view(a), mview(tau));
}
template<class Q, class V, class T, class A>
void abaInParallel(size_t n, ModelPool& pool, const Eigen::MatrixBase<Q>& q, const Eigen::MatrixBase<V>& v,
                   const Eigen::MatrixBase<T>& tau, const Eigen::MatrixBase<A>& a) {
  abaInParallel(n, pool, view(q), view(v), view(tau), mview(a));
}
}  // namespace pinocchio_amd
#endif
