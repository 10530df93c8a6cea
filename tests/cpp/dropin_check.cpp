// dropin_check.cpp -- drives the C++ drop-in (include/pinocchio_amd/parallel.hpp) the way
// a pinocchio caller would: a ModelPool, Eigen-shaped column-major host matrices, the
// batched calls, and the reference's argument errors.  Built with the host compiler alone
// (the drop-in needs only the C ABI); tests/test_gpu_dropin.py feeds it inputs and checks
// its outputs against the oracle.
//
//   dropin_check DIR   reads DIR/{dims.txt, parents.i32, jtype.i32, placement.f64,
//                      inertia.f64, q.f64, v.f64, a.f64, tau.f64, J.f64, gamma.f64},
//                      writes DIR/out_*.f64
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pinocchio_amd/parallel.hpp"

template<class T> std::vector<T> readRaw(const std::string& path, size_t n) {
  std::vector<T> v(n);
  std::ifstream f(path, std::ios::binary);
  if (!f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T))))
    throw std::runtime_error("cannot read " + path);
  return v;
}
void writeRaw(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}
template<class F> bool throwsInvalid(F&& f) {
  try {
    f();
  } catch (const std::invalid_argument&) {
    return true;
  }
  return false;
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s DIR\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  int nj = 0, nq = 0, nv = 0, nc = 0;
  long B = 0;
  {
    std::ifstream d(dir + "/dims.txt");
    d >> nj >> nq >> nv >> B >> nc;
  }
  auto parents = readRaw<int32_t>(dir + "/parents.i32", nj);
  auto jtype = readRaw<int32_t>(dir + "/jtype.i32", nj);
  auto placement = readRaw<double>(dir + "/placement.f64", 12 * nj);
  auto inertia = readRaw<double>(dir + "/inertia.f64", 10 * nj);
  auto q = readRaw<double>(dir + "/q.f64", (size_t)nq * B);
  auto v = readRaw<double>(dir + "/v.f64", (size_t)nv * B);
  auto a = readRaw<double>(dir + "/a.f64", (size_t)nv * B);
  auto tau = readRaw<double>(dir + "/tau.f64", (size_t)nv * B);
  rbd_model_desc desc{nj, parents.data(), jtype.data(), placement.data(), inertia.data(), nullptr, nullptr};

  const size_t threads = 4;
  pinocchio_amd::ModelPool pool(desc, threads);
  using pinocchio_amd::CMat;
  using pinocchio_amd::MMat;
  const CMat Q{q.data(), nq, B, nq}, V{v.data(), nv, B, nv}, A{a.data(), nv, B, nv}, T{tau.data(), nv, B, nv};
  std::vector<double> tau_out((size_t)nv * B), ddq((size_t)nv * B), M((size_t)nv * nv * B);
  pinocchio_amd::rneaInParallel(threads, pool, Q, V, A, MMat{tau_out.data(), nv, B, nv});
  pinocchio_amd::abaInParallel(threads, pool, Q, V, T, MMat{ddq.data(), nv, B, nv});
  pinocchio_amd::crbaInParallel(threads, pool, Q, MMat{M.data(), nv, nv * B, nv});
  writeRaw(dir + "/out_tau.f64", tau_out);
  writeRaw(dir + "/out_ddq.f64", ddq);
  pinocchio_amd::abaInParallel(threads, pool, Q, V, T, MMat{ddq.data(), nv, B, nv}, /*fp32=*/true);
  writeRaw(dir + "/out_ddq32.f64", ddq);
  writeRaw(dir + "/out_M.f64", M);
  if (pool.nv() == nv) {
    std::vector<double> d1((size_t)nv * nv * B), d2(d1.size()), d3(d1.size());
    const MMat o1{d1.data(), nv, nv * B, nv}, o2{d2.data(), nv, nv * B, nv}, o3{d3.data(), nv, nv * B, nv};
    pinocchio_amd::computeRNEADerivativesInParallel(threads, pool, Q, V, A, o1, o2, o3);
    writeRaw(dir + "/out_dtau_dq.f64", d1);
    writeRaw(dir + "/out_dtau_dv.f64", d2);
    pinocchio_amd::computeABADerivativesInParallel(threads, pool, Q, V, T, o1, o2, o3);
    writeRaw(dir + "/out_ddq_dq.f64", d1);
    writeRaw(dir + "/out_ddq_dv.f64", d2);
    writeRaw(dir + "/out_minv.f64", d3);
    pinocchio_amd::computeRNEADerivativesInParallel(threads, pool, Q, V, A, o1, o2, o3, /*fp32=*/true);
    writeRaw(dir + "/out_dtau_dq32.f64", d1);
    pinocchio_amd::computeABADerivativesInParallel(threads, pool, Q, V, T, o1, o2, o3, /*fp32=*/true);
    writeRaw(dir + "/out_ddq_dq32.f64", d1);
    writeRaw(dir + "/out_minv32.f64", d3);
    // outputs with a leading dimension above nv (staged FP64 path)
    std::vector<double> w1((size_t)(nv + 3) * nv * B), w2(w1.size()), w3(w1.size());
    pinocchio_amd::computeABADerivativesInParallel(threads, pool, Q, V, T, MMat{w1.data(), nv, nv * B, nv + 3},
                                                   MMat{w2.data(), nv, nv * B, nv + 3},
                                                   MMat{w3.data(), nv, nv * B, nv + 3});
    std::vector<double> packed((size_t)nv * nv * B);
    for (long c = 0; c < nv * B; ++c)
      for (int r = 0; r < nv; ++r) packed[(size_t)(r + c * nv)] = w2[(size_t)(r + c * (nv + 3))];
    writeRaw(dir + "/out_ddq_dv_ld.f64", packed);
    pinocchio_amd::computeMinverseInParallel(threads, pool, Q, o1);
    writeRaw(dir + "/out_minv64.f64", d1);
    // constrained forwardDynamics: J is nc x (nv B), the configurations' blocks side by side
    auto J = readRaw<double>(dir + "/J.f64", (size_t)nc * nv * B);
    auto gamma = readRaw<double>(dir + "/gamma.f64", (size_t)nc * B);
    std::vector<double> cddq((size_t)nv * B), lam((size_t)nc * B);
    pinocchio_amd::forwardDynamicsInParallel(threads, pool, Q, V, T, CMat{J.data(), nc, nv * B, nc},
                                             CMat{gamma.data(), nc, B, nc}, 0.0, MMat{cddq.data(), nv, B, nv},
                                             MMat{lam.data(), nc, B, nc});
    writeRaw(dir + "/out_cddq.f64", cddq);
    writeRaw(dir + "/out_lambda.f64", lam);
    // constraintDynamics: a point of joint nj - 1 held on a world point (3-D, LOCAL,
    // Baumgarte gains) and 10 proximal iterations with mu = 1e-3
    rbd_rigid_constraint rc{};
    rc.type = RBD_CONTACT_3D;
    rc.joint1_id = nj - 1;
    rc.joint2_id = 0;
    rc.reference_frame = RBD_LOCAL;
    for (int k = 0; k < 12; ++k) rc.joint1_placement[k] = rc.joint2_placement[k] = (k == 0 || k == 4 || k == 8) ? 1.0 : 0.0;
    rc.joint2_placement[11] = 0.5;
    for (int k = 0; k < 6; ++k) {
      rc.Kp[k] = 4.0;
      rc.Kd[k] = 1.5;
    }
    const rbd_proximal_settings ps{1e-10, 1e-10, 1e-3, 10};
    std::vector<double> kddq((size_t)nv * B), klam((size_t)3 * B);
    std::vector<int32_t> kit((size_t)B);
    pinocchio_amd::constraintDynamicsInParallel(threads, pool, Q, V, T, {rc}, ps, MMat{kddq.data(), nv, B, nv},
                                                MMat{klam.data(), 3, B, 3}, kit.data());
    writeRaw(dir + "/out_kddq.f64", kddq);
    writeRaw(dir + "/out_klambda.f64", klam);
    std::vector<double> kitd(kit.begin(), kit.end());
    writeRaw(dir + "/out_kiters.f64", kitd);
  }
  // the reference's argument errors (algorithm/parallel/rnea.hpp:53-68, aba.hpp:54-69)
  const bool e1 = throwsInvalid([&] { pinocchio_amd::rneaInParallel(threads + 1, pool, Q, V, A, MMat{tau_out.data(), nv, B, nv}); });
  const bool e2 = throwsInvalid([&] { pinocchio_amd::rneaInParallel(threads, pool, CMat{q.data(), nq - 1, B, nq}, V, A, MMat{tau_out.data(), nv, B, nv}); });
  const bool e3 = throwsInvalid([&] { pinocchio_amd::abaInParallel(threads, pool, Q, V, CMat{tau.data(), nv, B - 1, nv}, MMat{ddq.data(), nv, B, nv}); });
  const bool e4 = throwsInvalid([&] { pinocchio_amd::ModelPool bad(desc, 0); });
  std::printf("errors %d %d %d %d\n", (int)e1, (int)e2, (int)e3, (int)e4);
  return (e1 && e2 && e3 && e4) ? 0 : 1;
}
