// k_frames.hpp -- frame placements and frame Jacobians from the joint placements oMi and
// the WORLD joint Jacobian that k_kin writes:
//   oMf = oMi[parent] * placement                       (updateFramePlacements, frames.hxx:15-35)
//   J_f = support columns of the WORLD Jacobian, in     (getFrameJacobian, frames.hxx:156-175;
//         WORLD, LOCAL_WORLD_ALIGNED or LOCAL            translateJointJacobian, jacobian.hxx:175-300)
// One thread per (configuration, frame): blockIdx.y is the frame, threads run along the
// batch so every load and store is coalesced in the batch-fastest workspace.
#pragma once

#include "kernels_common.hpp"

namespace rbd {

enum { kRfWorld = 0, kRfLocal = 1, kRfLocalWorldAligned = 2 };

// oMi: 12 (j - 1) + k rows (R row major, then p); Jw: r + 6 c rows; frame tables: parent
// joint and placement (12 per frame, S); support: njoints x nv flags (1 where column c is a
// dof of joint j or of one of its ancestors).  Outputs: oMf 12 f + k rows, Jf f 6 nv + r + 6 c
// rows; either may be absent.
template<class S>
__global__ __launch_bounds__(256) void k_frames(int64_t B, int nv, const int32_t* __restrict__ fpar,
                                                const S* __restrict__ fpl, const uint8_t* __restrict__ support,
                                                CView<S> oMi, CView<S> Jw, int rf, View<S> oMf, View<S> Jf) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (b >= B) return;
  const int p = fpar[f];
  const S* X = fpl + 12 * f;
  S R[9], t[3];
  if (p > 0) {
    S Ri[9], ti[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) Ri[k] = oMi.p[(int64_t)(12 * (p - 1) + k) * oMi.rs + b * oMi.bs];
#pragma unroll
    for (int k = 0; k < 3; ++k) ti[k] = oMi.p[(int64_t)(12 * (p - 1) + 9 + k) * oMi.rs + b * oMi.bs];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c) R[3 * r + c] = Ri[3 * r] * X[c] + Ri[3 * r + 1] * X[3 + c] + Ri[3 * r + 2] * X[6 + c];
      t[r] = Ri[3 * r] * X[9] + Ri[3 * r + 1] * X[10] + Ri[3 * r + 2] * X[11] + ti[r];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = X[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = X[9 + k];
  }
  if (oMf.p) {
#pragma unroll
    for (int k = 0; k < 9; ++k) oMf.p[(int64_t)(12 * f + k) * oMf.rs + b * oMf.bs] = R[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) oMf.p[(int64_t)(12 * f + 9 + k) * oMf.rs + b * oMf.bs] = t[k];
  }
  if (!Jf.p) return;
  const uint8_t* sup = support + (int64_t)p * nv;
  const int64_t base = (int64_t)f * 6 * nv;
  for (int c = 0; c < nv; ++c) {
    S o[6] = {S(0), S(0), S(0), S(0), S(0), S(0)};
    if (sup[c]) {
      S l[3], a[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        l[k] = Jw.p[(int64_t)(6 * c + k) * Jw.rs + b * Jw.bs];
        a[k] = Jw.p[(int64_t)(6 * c + 3 + k) * Jw.rs + b * Jw.bs];
      }
      if (rf != kRfWorld) {  // linear part at the frame origin: l - t x a
        const S cx = t[1] * a[2] - t[2] * a[1], cy = t[2] * a[0] - t[0] * a[2], cz = t[0] * a[1] - t[1] * a[0];
        l[0] -= cx;
        l[1] -= cy;
        l[2] -= cz;
      }
      if (rf == kRfLocal) {  // then rotated into the frame: R^T
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          o[k] = R[k] * l[0] + R[3 + k] * l[1] + R[6 + k] * l[2];
          o[3 + k] = R[k] * a[0] + R[3 + k] * a[1] + R[6 + k] * a[2];
        }
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          o[k] = l[k];
          o[3 + k] = a[k];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) Jf.p[(base + 6 * c + r) * Jf.rs + b * Jf.bs] = o[r];
  }
}

// Frame motions (getFrameVelocity / getFrameAcceleration / getFrameClassicalAcceleration,
// frames.hxx:71-149) from the joint placements oMi and the joints' LOCAL velocities and
// accelerations that k_kin writes (6 (j - 1) + k rows, [linear; angular]):
//   LOCAL                placement^-1 . m
//   WORLD                oMi . m
//   LOCAL_WORLD_ALIGNED  (R_i (m.lin + m.ang x placement.p), R_i m.ang)
// classical: acc.lin += vel.ang x vel.lin in the requested frame.  A frame on the universe
// (parent 0) does not move.  One thread per (configuration, frame) as in k_frames; aj, af
// and acf may be absent.
template<class S>
__device__ __forceinline__ void frameMotion(const S* R, const S* ti, const S* Xp, const S* Xt, int rf, const S* m,
                                            S* o) {
  if (rf == kRfLocal) {  // placement^-1 . m: (R_p^T (lin - t_p x ang), R_p^T ang)
    const S l0 = m[0] - (Xt[1] * m[5] - Xt[2] * m[4]);
    const S l1 = m[1] - (Xt[2] * m[3] - Xt[0] * m[5]);
    const S l2 = m[2] - (Xt[0] * m[4] - Xt[1] * m[3]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      o[k] = Xp[k] * l0 + Xp[3 + k] * l1 + Xp[6 + k] * l2;
      o[3 + k] = Xp[k] * m[3] + Xp[3 + k] * m[4] + Xp[6 + k] * m[5];
    }
    return;
  }
  S l[3] = {m[0], m[1], m[2]};
  if (rf == kRfLocalWorldAligned) {  // lin + ang x t_p
    l[0] += m[4] * Xt[2] - m[5] * Xt[1];
    l[1] += m[5] * Xt[0] - m[3] * Xt[2];
    l[2] += m[3] * Xt[1] - m[4] * Xt[0];
  }
  S w[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    w[r] = R[3 * r] * m[3] + R[3 * r + 1] * m[4] + R[3 * r + 2] * m[5];
    o[r] = R[3 * r] * l[0] + R[3 * r + 1] * l[1] + R[3 * r + 2] * l[2];
    o[3 + r] = w[r];
  }
  if (rf == kRfWorld) {  // oMi . m: R lin + t_i x (R ang)
    o[0] += ti[1] * w[2] - ti[2] * w[1];
    o[1] += ti[2] * w[0] - ti[0] * w[2];
    o[2] += ti[0] * w[1] - ti[1] * w[0];
  }
}

template<class S>
__global__ __launch_bounds__(256) void k_frame_motions(int64_t B, const int32_t* __restrict__ fpar,
                                                       const S* __restrict__ fpl, CView<S> oMi, CView<S> vj, CView<S> aj,
                                                       int rf, View<S> vf, View<S> af, View<S> acf) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (b >= B) return;
  const int p = fpar[f];
  const S* X = fpl + 12 * f;
  S vo[6] = {S(0), S(0), S(0), S(0), S(0), S(0)}, ao[6] = {S(0), S(0), S(0), S(0), S(0), S(0)};
  if (p > 0) {
    S R[9], ti[3], m[6];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = oMi.p[(int64_t)(12 * (p - 1) + k) * oMi.rs + b * oMi.bs];
#pragma unroll
    for (int k = 0; k < 3; ++k) ti[k] = oMi.p[(int64_t)(12 * (p - 1) + 9 + k) * oMi.rs + b * oMi.bs];
#pragma unroll
    for (int k = 0; k < 6; ++k) m[k] = vj.p[(int64_t)(6 * (p - 1) + k) * vj.rs + b * vj.bs];
    frameMotion(R, ti, X, X + 9, rf, m, vo);
    if (aj.p) {
#pragma unroll
      for (int k = 0; k < 6; ++k) m[k] = aj.p[(int64_t)(6 * (p - 1) + k) * aj.rs + b * aj.bs];
      frameMotion(R, ti, X, X + 9, rf, m, ao);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) vf.p[(int64_t)(6 * f + k) * vf.rs + b * vf.bs] = vo[k];
  if (af.p) {
#pragma unroll
    for (int k = 0; k < 6; ++k) af.p[(int64_t)(6 * f + k) * af.rs + b * af.bs] = ao[k];
  }
  if (acf.p) {  // + vel.ang x vel.lin
    ao[0] += vo[4] * vo[2] - vo[5] * vo[1];
    ao[1] += vo[5] * vo[0] - vo[3] * vo[2];
    ao[2] += vo[3] * vo[1] - vo[4] * vo[0];
#pragma unroll
    for (int k = 0; k < 6; ++k) acf.p[(int64_t)(6 * f + k) * acf.rs + b * acf.bs] = ao[k];
  }
}

// computeJointJacobiansTimeVariation (jacobian.hxx:673-740) from k_kin's oMi, LOCAL joint
// velocities and WORLD Jacobian: column c of joint j = dofJoint[c] gets
// dJ_c = ov_j x J_c, ov_j = oMi_j . v_j (motion action).  One thread per (configuration,
// column); J is copied to Jout when that is present.
template<class S>
__global__ __launch_bounds__(256) void k_jacobian_dot(int64_t B, const int32_t* __restrict__ dofJoint, CView<S> oMi,
                                                      CView<S> vj, CView<S> Jw, View<S> dJ, View<S> Jout) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (b >= B) return;
  const int j = dofJoint[c];
  S R[9], t[3], m[6], Jc[6];
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = oMi.p[(int64_t)(12 * (j - 1) + k) * oMi.rs + b * oMi.bs];
#pragma unroll
  for (int k = 0; k < 3; ++k) t[k] = oMi.p[(int64_t)(12 * (j - 1) + 9 + k) * oMi.rs + b * oMi.bs];
#pragma unroll
  for (int k = 0; k < 6; ++k) m[k] = vj.p[(int64_t)(6 * (j - 1) + k) * vj.rs + b * vj.bs];
#pragma unroll
  for (int k = 0; k < 6; ++k) Jc[k] = Jw.p[(int64_t)(6 * c + k) * Jw.rs + b * Jw.bs];
  S w[3], l[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    w[r] = R[3 * r] * m[3] + R[3 * r + 1] * m[4] + R[3 * r + 2] * m[5];
    l[r] = R[3 * r] * m[0] + R[3 * r + 1] * m[1] + R[3 * r + 2] * m[2];
  }
  l[0] += t[1] * w[2] - t[2] * w[1];
  l[1] += t[2] * w[0] - t[0] * w[2];
  l[2] += t[0] * w[1] - t[1] * w[0];
  S o[6];  // (w x J.lin + l x J.ang, w x J.ang)
  o[0] = w[1] * Jc[2] - w[2] * Jc[1] + l[1] * Jc[5] - l[2] * Jc[4];
  o[1] = w[2] * Jc[0] - w[0] * Jc[2] + l[2] * Jc[3] - l[0] * Jc[5];
  o[2] = w[0] * Jc[1] - w[1] * Jc[0] + l[0] * Jc[4] - l[1] * Jc[3];
  o[3] = w[1] * Jc[5] - w[2] * Jc[4];
  o[4] = w[2] * Jc[3] - w[0] * Jc[5];
  o[5] = w[0] * Jc[4] - w[1] * Jc[3];
#pragma unroll
  for (int k = 0; k < 6; ++k) dJ.p[(int64_t)(6 * c + k) * dJ.rs + b * dJ.bs] = o[k];
  if (Jout.p) {
#pragma unroll
    for (int k = 0; k < 6; ++k) Jout.p[(int64_t)(6 * c + k) * Jout.rs + b * Jout.bs] = Jc[k];
  }
}

// getFrameJacobianTimeVariation (frames.hxx:293-368) from oMi, the joints' LOCAL velocities,
// the WORLD Jacobian J and its time variation dJ (k_jacobian_dot): the support columns of
// dJ taken into the frame like k_frames takes J, then
//   LOCAL                dJ_c -= v_f x (oMf^-1 . J_c),  v_f = placement^-1 . v_parent
//   LOCAL_WORLD_ALIGNED  dJ_c.lin -= (ov.lin + ov.ang x oMf.p) x J_c.ang,  ov = oMi . v_parent
// Output rows f 6 nv + r + 6 c, zero outside the support.
template<class S>
__global__ __launch_bounds__(256) void k_frame_jacobian_dot(int64_t B, int nv, const int32_t* __restrict__ fpar,
                                                            const S* __restrict__ fpl, const uint8_t* __restrict__ support,
                                                            CView<S> oMi, CView<S> vj, CView<S> Jw, CView<S> dJw, int rf,
                                                            View<S> out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (b >= B) return;
  const int p = fpar[f];
  const S* X = fpl + 12 * f;
  S Ri[9], ti[3], vp[6];
  if (p > 0) {
#pragma unroll
    for (int k = 0; k < 9; ++k) Ri[k] = oMi.p[(int64_t)(12 * (p - 1) + k) * oMi.rs + b * oMi.bs];
#pragma unroll
    for (int k = 0; k < 3; ++k) ti[k] = oMi.p[(int64_t)(12 * (p - 1) + 9 + k) * oMi.rs + b * oMi.bs];
#pragma unroll
    for (int k = 0; k < 6; ++k) vp[k] = vj.p[(int64_t)(6 * (p - 1) + k) * vj.rs + b * vj.bs];
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) Ri[k] = (k % 4 == 0) ? S(1) : S(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) ti[k] = S(0);
#pragma unroll
    for (int k = 0; k < 6; ++k) vp[k] = S(0);
  }
  S R[9], t[3];  // oMf = oMi * placement
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) R[3 * r + c] = Ri[3 * r] * X[c] + Ri[3 * r + 1] * X[3 + c] + Ri[3 * r + 2] * X[6 + c];
    t[r] = Ri[3 * r] * X[9] + Ri[3 * r + 1] * X[10] + Ri[3 * r + 2] * X[11] + ti[r];
  }
  S vf[6], u[3];  // v_f = placement^-1 . v_parent; u = ov.lin + ov.ang x oMf.p
  frameMotion(R, ti, X, X + 9, kRfLocal, vp, vf);
  {
    S ov[6];
    frameMotion(Ri, ti, X, X + 9, kRfWorld, vp, ov);
    u[0] = ov[0] + ov[4] * t[2] - ov[5] * t[1];
    u[1] = ov[1] + ov[5] * t[0] - ov[3] * t[2];
    u[2] = ov[2] + ov[3] * t[1] - ov[4] * t[0];
  }
  const uint8_t* sup = support + (int64_t)p * nv;
  const int64_t base = (int64_t)f * 6 * nv;
  for (int c = 0; c < nv; ++c) {
    S o[6] = {S(0), S(0), S(0), S(0), S(0), S(0)};
    if (sup[c]) {
      S m[6], J[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        m[k] = dJw.p[(int64_t)(6 * c + k) * dJw.rs + b * dJw.bs];
        J[k] = Jw.p[(int64_t)(6 * c + k) * Jw.rs + b * Jw.bs];
      }
      if (rf == kRfWorld) {
#pragma unroll
        for (int k = 0; k < 6; ++k) o[k] = m[k];
      } else if (rf == kRfLocalWorldAligned) {  // lin - t x ang - u x J.ang
        o[0] = m[0] - (t[1] * m[5] - t[2] * m[4]) - (u[1] * J[5] - u[2] * J[4]);
        o[1] = m[1] - (t[2] * m[3] - t[0] * m[5]) - (u[2] * J[3] - u[0] * J[5]);
        o[2] = m[2] - (t[0] * m[4] - t[1] * m[3]) - (u[0] * J[4] - u[1] * J[3]);
#pragma unroll
        for (int k = 3; k < 6; ++k) o[k] = m[k];
      } else {  // oMf^-1 . m - v_f x (oMf^-1 . J)
        S ml[3], jl[3], ja[3];
        ml[0] = m[0] - (t[1] * m[5] - t[2] * m[4]);
        ml[1] = m[1] - (t[2] * m[3] - t[0] * m[5]);
        ml[2] = m[2] - (t[0] * m[4] - t[1] * m[3]);
        const S j0 = J[0] - (t[1] * J[5] - t[2] * J[4]);
        const S j1 = J[1] - (t[2] * J[3] - t[0] * J[5]);
        const S j2 = J[2] - (t[0] * J[4] - t[1] * J[3]);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          o[k] = R[k] * ml[0] + R[3 + k] * ml[1] + R[6 + k] * ml[2];
          o[3 + k] = R[k] * m[3] + R[3 + k] * m[4] + R[6 + k] * m[5];
          jl[k] = R[k] * j0 + R[3 + k] * j1 + R[6 + k] * j2;
          ja[k] = R[k] * J[3] + R[3 + k] * J[4] + R[6 + k] * J[5];
        }
        // v_f x (jl, ja) = (w x jl + vl x ja, w x ja), v_f = (vl, w)
        o[0] -= vf[4] * jl[2] - vf[5] * jl[1] + vf[1] * ja[2] - vf[2] * ja[1];
        o[1] -= vf[5] * jl[0] - vf[3] * jl[2] + vf[2] * ja[0] - vf[0] * ja[2];
        o[2] -= vf[3] * jl[1] - vf[4] * jl[0] + vf[0] * ja[1] - vf[1] * ja[0];
        o[3] -= vf[4] * ja[2] - vf[5] * ja[1];
        o[4] -= vf[5] * ja[0] - vf[3] * ja[2];
        o[5] -= vf[3] * ja[1] - vf[4] * ja[0];
      }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) out.p[(base + 6 * c + r) * out.rs + b * out.bs] = o[r];
  }
}

}  // namespace rbd
