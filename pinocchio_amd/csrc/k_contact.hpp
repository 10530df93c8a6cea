// k_contact.hpp -- rows of the rigid constraints of constraintDynamics
// (constrained-dynamics.hxx:193-420) from the kinematics pass's outputs.
//
// Per configuration and constraint (RigidConstraintModel::calc / ::jacobian,
// contact-info.hpp:426-700): oMc1 = oMi[j1] M1, oMc2 = oMi[j2] M2, c1Mc2 = oMc1^-1 oMc2;
// world spatial velocity and velocity-product acceleration of the joints, ov = oMi.act(v),
// oa = oMi.act(a) from the local v, a of the pass at zero joint acceleration (the same
// quantities the reference's forward step accumulates as ov_parent + ..., oa_parent +
// ov_parent x ov); the placement and velocity errors (log6 of c1Mc2 for 6-D rows), the
// Baumgarte correction -Kp e - Kd de, the drift terms of the contact frames, and the
// constraint Jacobian from the WORLD joint Jacobian columns, each signed by which joint's
// support holds it.  The rows go out as J (element (i, k) at row i + k nrows, the layout
// the KKT kernel stages) and gamma = -rhs (the KKT kernel solves J ddq + gamma = 0).
//
// One thread per configuration, batch-fastest inputs and outputs (coalesced).  All the
// arithmetic is FP64 whatever the input precision: it is a few hundred operations per
// constraint, and log6 and the drift differences keep their digits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "k_frames.hpp"  // reference-frame codes (kRfLocal, kRfLocalWorldAligned)
#include "kernels_common.hpp"
#include "spatial.hpp"

namespace rbd {

constexpr int kContact3D = 0, kContact6D = 1;

// One rigid constraint on the device.  sup (nv bytes per constraint, separate array):
// bit 0 = the dof is in joint1's support, bit 1 = in joint2's.
struct RigidConstraintDev {
  int type, j1, j2, rf, row, gains;
  double M1[12], M2[12];  // joint placements: R row major, then p
  double Kp[6], Kd[6], des[6];
};

using D3 = V3<double>;
using DR = R3<double>;
using DX = Xform<double>;
using DM = Mot<double>;

__device__ __forceinline__ DX xformFrom(const double* x) {
  return {{x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], x[8]}, {x[9], x[10], x[11]}};
}
__device__ __forceinline__ DR transposeR(const DR& R) { return {R.m00, R.m10, R.m20, R.m01, R.m11, R.m21, R.m02, R.m12, R.m22}; }
__device__ __forceinline__ DX inverseX(const DX& X) {
  const DR Rt = transposeR(X.R);
  return {Rt, -mul(Rt, X.p)};
}

// log6 of a rigid motion (log.hxx:58-230), away from rotations by pi: theta from the
// trace (Taylor expansions within cbrt(eps) of the identity), w = theta / sin(theta)
// unSkew(R), linear part alpha p - w x p / 2 + beta (w . p) w
__device__ inline DM log6(const DX& M) {
  const double prec = 6.0554544523933395e-06;  // cbrt(DBL_EPSILON)
  const DR& R = M.R;
  const double tr = R.m00 + R.m11 + R.m22;
  const double cv = (tr - 1.0) * 0.5;
  const double theta = tr <= 3.0 - prec ? acos(fmin(fmax(cv, -1.0), 1.0))
                                        : sqrt(2.0 * (1.0 - cv) + 2.220446049250313e-16 * 2.220446049250313e-16);
  const D3 a{0.5 * (R.m21 - R.m12), 0.5 * (R.m02 - R.m20), 0.5 * (R.m10 - R.m01)};
  const double t2 = dot(a, a);
  const double st = sin(theta), ct = cos(theta);
  const double t = theta >= prec ? theta / st : 1.0 + t2 / 6.0 + t2 * t2 * 3.0 / 40.0;
  const D3 w = t * a;
  const bool small = tr >= 3.0 - prec;
  const double alpha = small ? 1.0 - t2 / 12.0 - t2 * t2 / 720.0 : theta * st / (2.0 * (1.0 - ct));
  const double beta = small ? 1.0 / 12.0 + t2 / 720.0 : 1.0 / (theta * theta) - st / (2.0 * theta * (1.0 - ct));
  return {alpha * M.p - 0.5 * cross(w, M.p) + (beta * dot(w, M.p)) * w, w};
}

template<class S> __device__ __forceinline__ double ld(const CView<S>& g, int64_t row, int64_t b) {
  return double(g.p[row * g.rs + b * g.bs]);
}

// gM: 12 (nj - 1) rows (oMi of joints 1.., R row major then p); gV, gA: 6 (nj - 1) rows
// (local v, a); gJw: 6 nv rows (WORLD joint Jacobian, element r + 6 dof).
template<class S>
__global__ __launch_bounds__(256) void k_contact_terms(int64_t B, int nv, int ncon, int nrows,
                                                       const RigidConstraintDev* __restrict__ con,
                                                       const uint8_t* __restrict__ sup, CView<S> gM, CView<S> gV,
                                                       CView<S> gA, CView<S> gJw, View<S> gJc, View<S> gg) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  auto joint = [&](int j, DX& oM, DM& ov, DM& oa) {
    if (j == 0) {
      oM = {{1, 0, 0, 0, 1, 0, 0, 0, 1}, {0, 0, 0}};
      ov = oa = zeroMot<double>();
      return;
    }
    double x[12];
    for (int k = 0; k < 12; ++k) x[k] = ld(gM, (int64_t)(j - 1) * 12 + k, b);
    oM = xformFrom(x);
    const int64_t r = (int64_t)(j - 1) * 6;
    const DM v{{ld(gV, r, b), ld(gV, r + 1, b), ld(gV, r + 2, b)}, {ld(gV, r + 3, b), ld(gV, r + 4, b), ld(gV, r + 5, b)}};
    const DM a{{ld(gA, r, b), ld(gA, r + 1, b), ld(gA, r + 2, b)}, {ld(gA, r + 3, b), ld(gA, r + 4, b), ld(gA, r + 5, b)}};
    ov = act(oM, v);
    oa = act(oM, a);
  };
  for (int c = 0; c < ncon; ++c) {
    const RigidConstraintDev& cm = con[c];
    DX oMi1, oMi2;
    DM ov1, oa1, ov2, oa2;
    joint(cm.j1, oMi1, ov1, oa1);
    joint(cm.j2, oMi2, ov2, oa2);
    const DX oMc1 = compose(oMi1, xformFrom(cm.M1));
    const DX oMc2 = compose(oMi2, xformFrom(cm.M2));
    const DX c1Mc2 = compose(inverseX(oMc1), oMc2);
    const DM vc1 = cm.j1 > 0 ? actInv(oMc1, ov1) : zeroMot<double>();
    const DM vc2 = cm.j2 > 0 ? actInv(oMc2, ov2) : zeroMot<double>();
    const DM vc2in1 = act(c1Mc2, vc2);
    const bool six = cm.type == kContact6D;
    DM perr, verr, aerr = zeroMot<double>();
    if (six) {
      const DM l = log6(c1Mc2);
      perr = {-l.l, -l.a};
      verr = vc1 - vc2in1;
    } else {
      perr = {-c1Mc2.p, {0, 0, 0}};
      verr = {vc1.l - mul(c1Mc2.R, vc2.l), {0, 0, 0}};
    }
    if (cm.gains) {
      aerr.l = {-cm.Kp[0] * perr.l.x - cm.Kd[0] * verr.l.x, -cm.Kp[1] * perr.l.y - cm.Kd[1] * verr.l.y,
                -cm.Kp[2] * perr.l.z - cm.Kd[2] * verr.l.z};
      if (six)
        aerr.a = {-cm.Kp[3] * perr.a.x - cm.Kd[3] * verr.a.x, -cm.Kp[4] * perr.a.y - cm.Kd[4] * verr.a.y,
                  -cm.Kp[5] * perr.a.z - cm.Kd[5] * verr.a.z};
    }
    DM acc1, acc2;
    if (cm.rf == kRfLocalWorldAligned) {
      acc1.l = oa1.l + cross(oa1.a, oMc1.p);
      if (!six) acc1.l = acc1.l + cross(ov1.a, ov1.l + cross(ov1.a, oMc1.p));
      acc1.a = oa1.a;
      if (!six) {
        acc2.l = oa2.l + cross(oa2.a, oMc2.p) + cross(ov2.a, ov2.l + cross(ov2.a, oMc2.p));
        acc2.a = {0, 0, 0};
      } else {
        acc2.l = oa2.l + cross(oa2.a, oMc1.p);  // c1's origin, as the reference takes it
        acc2.a = oa2.a;
      }
      aerr = {mul(oMc1.R, aerr.l), mul(oMc1.R, aerr.a)};
    } else {
      acc1 = actInv(oMc1, oa1);
      if (!six) {
        acc1.l = acc1.l + cross(vc1.a, vc1.l);
        acc1.a = {0, 0, 0};
        acc2.l = mulT(oMc1.R, oa2.l + cross(oa2.a, oMc2.p) + cross(ov2.a, ov2.l + cross(ov2.a, oMc2.p)));
        acc2.a = {0, 0, 0};
      } else {
        acc1 = acc1 + crossMM(verr, vc2in1);
        acc2 = actInv(oMc1, oa2);
      }
    }
    const int dim = six ? 6 : 3;
    for (int k = 0; k < dim; ++k) {
      const double r = -comp(acc1, k) + comp(acc2, k) + comp(aerr, k) + cm.des[k];
      gg.p[(int64_t)(cm.row + k) * gg.rs + b * gg.bs] = S(-r);
    }
    // Jacobian columns
    const uint8_t* sc = sup + (size_t)c * nv;
    for (int k = 0; k < nv; ++k) {
      const int s = sc[k];
      double col[6] = {0, 0, 0, 0, 0, 0};
      if (s) {
        const int sign = s == 1 ? 1 : (s == 2 ? -1 : 0);
        const int64_t r = (int64_t)k * 6;
        const DM Jc{{ld(gJw, r, b), ld(gJw, r + 1, b), ld(gJw, r + 2, b)},
                    {ld(gJw, r + 3, b), ld(gJw, r + 4, b), ld(gJw, r + 5, b)}};
        if (!six) {
          D3 l;
          if (cm.rf == kRfLocal) {
            if (sign == 0) l = actInv(oMc1, Jc).l - mul(c1Mc2.R, actInv(oMc2, Jc).l);
            else if (sign == 1) l = actInv(oMc1, Jc).l;
            else l = -mul(c1Mc2.R, actInv(oMc2, Jc).l);
          } else {
            if (sign == 0) l = cross(oMc2.p - oMc1.p, Jc.a);
            else l = double(sign) * (Jc.l - cross(sign == 1 ? oMc1.p : oMc2.p, Jc.a));
          }
          col[0] = l.x;
          col[1] = l.y;
          col[2] = l.z;
        } else {
          const DM m = cm.rf == kRfLocal ? actInv(oMc1, Jc) : DM{Jc.l - cross(oMc1.p, Jc.a), Jc.a};
          for (int e = 0; e < 6; ++e) col[e] = double(sign) * comp(m, e);
        }
      }
      for (int e = 0; e < dim; ++e) gJc.p[(int64_t)(cm.row + e + k * nrows) * gJc.rs + b * gJc.bs] = S(col[e]);
    }
  }
}

}  // namespace rbd
