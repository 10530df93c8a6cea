// ops.hpp -- launcher table of one model's kernels (host side).
//
// A prebuilt topology fills its table with launchers of kernels compiled for that tree;
// every model also gets launchers of the run-time-topology kernels (k_generic.hpp), which
// fill the entries the prebuilt table lacks (all of them for a tree that is not prebuilt).
// Every launcher receives the model's run-time program; prebuilt launchers ignore it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbd {

template<class S> struct View;
template<class S> struct CView;
struct GenericModel;  // run-time topology program of one model (generic.hpp)
using GM = const GenericModel*;

template<class S>
using LaunchVec3 = void (*)(hipStream_t, GM, const S* mc, int64_t B, CView<S> in0, CView<S> in1, CView<S> in2,
                            View<S> out);
template<class S>
using LaunchCrba = void (*)(hipStream_t, GM, const S* mc, int64_t B, CView<S> q, View<S> M, int upper);
template<class S>
using LaunchChol = void (*)(hipStream_t, GM, int64_t B, CView<S> M, View<double> U, View<double> D);
template<class S>
using LaunchCholSolve = void (*)(hipStream_t, GM, int64_t B, CView<double> U, CView<double> D, View<S> x);
// waves: 1 or 4 (any value >= 2) waves share a tile of 64 configurations (0: chosen from the batch size)
template<class S>
using LaunchRneaD = void (*)(hipStream_t, GM, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a,
                             View<S> dq, View<S> dv, View<S> M, int mask, int waves);
// FP32 factor for the ABA-derivative pipeline
using LaunchChol32 = void (*)(hipStream_t, GM, int64_t B, CView<float> M, View<float> U, View<float> D);
template<class S>
using LaunchAbaDSolve = void (*)(hipStream_t, GM, int64_t B, CView<S> U, CView<S> D, View<S> Xq, View<S> Xv,
                                 View<S> Minv, int jobBegin);

// forward kinematics / joint Jacobians: outputs placements, local velocities and
// accelerations, WORLD Jacobian columns (any may be null)
template<class S>
using LaunchKin = void (*)(hipStream_t, GM, const S* mc, int64_t B, CView<S> q, CView<S> v, CView<S> a, View<S> oM,
                           View<S> vo, View<S> ao, View<S> J);

// forward dynamics under J ddq + gamma = 0 from the FP64 tree factor of M and nle: ddq,
// lambda; mode 1: impulse dynamics with v_before in `tau` (nle and gamma unused); mode 2:
// the Delassus matrix into `lambda` (k_kkt.hpp); prox: proximal iterations of
// constraintDynamics
struct KktProx;
template<class S>
using LaunchKkt = void (*)(hipStream_t, GM, int64_t B, CView<double> U, CView<double> D, int nc, CView<S> J,
                           CView<S> tau, CView<S> nle, CView<S> gamma, double damping, int mode, double rcoef,
                           const KktProx& prox, View<S> ddq, View<S> lambda);

// frictional contact solvers over B problems of nc cones (1..4), FP64 (k_ncp.hpp)
struct AdmmParams;
void launchPgs(hipStream_t st, int nc, int64_t B, CView<double> G, CView<double> g, CView<double> mu, View<double> x,
               double over_relax, int max_it, double abs_prec, double rel_prec, View<double> stats);
void launchAdmm(hipStream_t st, int nc, int64_t B, CView<double> G, CView<double> g, CView<double> mu, CView<double> R,
                CView<double> primal_guess, CView<double> dual_guess, const AdmmParams& P, View<double> y,
                View<double> z, View<double> s, View<double> stats);

struct TopoOps {
  const char* name;
  int njoints;
  const int* parent;
  const int* type;
  int nq, nv, compact;
  LaunchVec3<float> rnea_f32;
  LaunchVec3<double> rnea_f64;
  LaunchVec3<float> aba_f32;
  LaunchVec3<double> aba_f64;
  LaunchCrba<float> crba_f32;
  LaunchCrba<double> crba_f64;
  LaunchChol<float> chol_f32;
  LaunchChol<double> chol_f64;
  LaunchCholSolve<float> cholsolve_f32;
  LaunchCholSolve<double> cholsolve_f64;
  LaunchRneaD<float> rnead_f32;
  LaunchRneaD<double> rnead_f64;
  LaunchChol32 chol32_f32;
  LaunchAbaDSolve<float> abad_solve_f32;
  LaunchAbaDSolve<double> abad_solve_f64;
  LaunchKin<float> kin_f32;
  LaunchKin<double> kin_f64;
  LaunchKkt<float> kkt_f32;
  LaunchKkt<double> kkt_f64;
  int nnzU;  // strictly-upper structural non-zeros of M
};

struct MultiGroup;
struct MultiTile;
// A compiled mix of topologies for the one-launch mixed-robot kernel (k_multi).
struct MixOps {
  const char* name;
  int ntopo;
  const TopoOps* topo[8];
  void (*launch)(hipStream_t, unsigned grid, const MultiGroup* groups, const MultiTile* tiles);
};

}  // namespace rbd
