// spatial.hpp -- spatial algebra for the per-lane tree sweeps (device and host).
//
// Conventions follow the reference (spatial/*.hpp): 6-vectors are [linear; angular];
// an Xform (R, p) maps frame-i quantities to the parent frame (SE3 liMi, se3-tpl.hpp);
// `act` / `actInv` are SE3::act / SE3::actInv on motions and forces
// (motion-dense.hpp:264-288, force-dense.hpp:214-243); the articulated inertia is a
// symmetric 6x6 stored as 21 packed entries.  Everything is written for registers:
// small structs of scalars, no indexing that survives inlining.
#pragma once

#include <hip/hip_runtime.h>

#define RBD_HD __host__ __device__ __forceinline__

namespace rbd {

template<class T> struct V3 {
  T x, y, z;
};
template<class T> RBD_HD V3<T> mk3(T a, T b, T c) { return V3<T>{a, b, c}; }
template<class T> RBD_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template<class T> RBD_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template<class T> RBD_HD V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template<class T> RBD_HD V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template<class T> RBD_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template<class T> RBD_HD V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template<class T> RBD_HD T comp(V3<T> a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// 3x3 rotation, row major.
template<class T> struct R3 {
  T m00, m01, m02, m10, m11, m12, m20, m21, m22;
};
template<class T> RBD_HD V3<T> mul(const R3<T>& R, V3<T> x) {
  return {R.m00 * x.x + R.m01 * x.y + R.m02 * x.z, R.m10 * x.x + R.m11 * x.y + R.m12 * x.z,
          R.m20 * x.x + R.m21 * x.y + R.m22 * x.z};
}
template<class T> RBD_HD V3<T> mulT(const R3<T>& R, V3<T> x) {
  return {R.m00 * x.x + R.m10 * x.y + R.m20 * x.z, R.m01 * x.x + R.m11 * x.y + R.m21 * x.z,
          R.m02 * x.x + R.m12 * x.y + R.m22 * x.z};
}
template<class T> RBD_HD R3<T> mul(const R3<T>& A, const R3<T>& B) {
  R3<T> r;
  r.m00 = A.m00 * B.m00 + A.m01 * B.m10 + A.m02 * B.m20;
  r.m01 = A.m00 * B.m01 + A.m01 * B.m11 + A.m02 * B.m21;
  r.m02 = A.m00 * B.m02 + A.m01 * B.m12 + A.m02 * B.m22;
  r.m10 = A.m10 * B.m00 + A.m11 * B.m10 + A.m12 * B.m20;
  r.m11 = A.m10 * B.m01 + A.m11 * B.m11 + A.m12 * B.m21;
  r.m12 = A.m10 * B.m02 + A.m11 * B.m12 + A.m12 * B.m22;
  r.m20 = A.m20 * B.m00 + A.m21 * B.m10 + A.m22 * B.m20;
  r.m21 = A.m20 * B.m01 + A.m21 * B.m11 + A.m22 * B.m21;
  r.m22 = A.m20 * B.m02 + A.m21 * B.m12 + A.m22 * B.m22;
  return r;
}
template<class T> RBD_HD V3<T> col(const R3<T>& R, int k) {
  return k == 0 ? V3<T>{R.m00, R.m10, R.m20} : (k == 1 ? V3<T>{R.m01, R.m11, R.m21} : V3<T>{R.m02, R.m12, R.m22});
}

// Placement Rp followed by a rotation of angle (s, c) about axis `ax` (0,1,2):
// R = Rp * Rot_ax(q).  Only the two columns orthogonal to the axis change.
template<class T> RBD_HD R3<T> rotAbout(const R3<T>& Rp, int ax, T s, T c) {
  R3<T> R = Rp;
  if (ax == 0) {  // columns 1, 2
    R.m01 = c * Rp.m01 + s * Rp.m02; R.m02 = c * Rp.m02 - s * Rp.m01;
    R.m11 = c * Rp.m11 + s * Rp.m12; R.m12 = c * Rp.m12 - s * Rp.m11;
    R.m21 = c * Rp.m21 + s * Rp.m22; R.m22 = c * Rp.m22 - s * Rp.m21;
  } else if (ax == 1) {  // columns 2, 0 : Ry = [[c,0,s],[0,1,0],[-s,0,c]]
    R.m00 = c * Rp.m00 - s * Rp.m02; R.m02 = s * Rp.m00 + c * Rp.m02;
    R.m10 = c * Rp.m10 - s * Rp.m12; R.m12 = s * Rp.m10 + c * Rp.m12;
    R.m20 = c * Rp.m20 - s * Rp.m22; R.m22 = s * Rp.m20 + c * Rp.m22;
  } else {  // columns 0, 1 : Rz = [[c,-s,0],[s,c,0],[0,0,1]]
    R.m00 = c * Rp.m00 + s * Rp.m01; R.m01 = c * Rp.m01 - s * Rp.m00;
    R.m10 = c * Rp.m10 + s * Rp.m11; R.m11 = c * Rp.m11 - s * Rp.m10;
    R.m20 = c * Rp.m20 + s * Rp.m21; R.m21 = c * Rp.m21 - s * Rp.m20;
  }
  return R;
}

// Eigen Quaternion::toRotationMatrix for [x, y, z, w] (not renormalised), as the
// reference's free-flyer calc (joint-free-flyer.hpp:333-351).
template<class T> RBD_HD R3<T> quatToRot(T x, T y, T z, T w) {
  const T tx = T(2) * x, ty = T(2) * y, tz = T(2) * z;
  const T twx = tx * w, twy = ty * w, twz = tz * w;
  const T txx = tx * x, txy = ty * x, txz = tz * x;
  const T tyy = ty * y, tyz = tz * y, tzz = tz * z;
  R3<T> R;
  R.m00 = T(1) - (tyy + tzz); R.m01 = txy - twz;          R.m02 = txz + twy;
  R.m10 = txy + twz;          R.m11 = T(1) - (txx + tzz); R.m12 = tyz - twx;
  R.m20 = txz - twy;          R.m21 = tyz + twx;          R.m22 = T(1) - (txx + tyy);
  return R;
}

template<class T> struct Mot {  // motion [linear; angular]
  V3<T> l, a;
};
template<class T> struct Frc {  // force [linear; angular]
  V3<T> l, a;
};
template<class T> RBD_HD Mot<T> operator+(const Mot<T>& x, const Mot<T>& y) { return {x.l + y.l, x.a + y.a}; }
template<class T> RBD_HD Mot<T> operator-(const Mot<T>& x, const Mot<T>& y) { return {x.l - y.l, x.a - y.a}; }
template<class T> RBD_HD Frc<T> operator+(const Frc<T>& x, const Frc<T>& y) { return {x.l + y.l, x.a + y.a}; }
template<class T> RBD_HD Frc<T> operator-(const Frc<T>& x, const Frc<T>& y) { return {x.l - y.l, x.a - y.a}; }
template<class T> RBD_HD Mot<T> scale(T s, const Mot<T>& x) { return {s * x.l, s * x.a}; }
template<class T> RBD_HD Frc<T> scale(T s, const Frc<T>& x) { return {s * x.l, s * x.a}; }
template<class T> RBD_HD T dot(const Mot<T>& m, const Frc<T>& f) { return dot(m.l, f.l) + dot(m.a, f.a); }
template<class T> RBD_HD Mot<T> zeroMot() { return {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}}; }
template<class T> RBD_HD Frc<T> zeroFrc() { return {{T(0), T(0), T(0)}, {T(0), T(0), T(0)}}; }
// component k of a 6-vector
template<class T> RBD_HD T comp(const Mot<T>& m, int k) { return k < 3 ? comp(m.l, k) : comp(m.a, k - 3); }
template<class T> RBD_HD T comp(const Frc<T>& m, int k) { return k < 3 ? comp(m.l, k) : comp(m.a, k - 3); }

// x ^ y (motion cross motion), motion-dense.hpp
template<class T> RBD_HD Mot<T> crossMM(const Mot<T>& x, const Mot<T>& y) {
  return {cross(x.a, y.l) + cross(x.l, y.a), cross(x.a, y.a)};
}
// v.cross(f) (motion cross force, dual action)
template<class T> RBD_HD Frc<T> crossMF(const Mot<T>& v, const Frc<T>& f) {
  return {cross(v.a, f.l), cross(v.a, f.a) + cross(v.l, f.l)};
}

template<class T> struct Xform {
  R3<T> R;
  V3<T> p;
};
template<class T> RBD_HD Xform<T> compose(const Xform<T>& A, const Xform<T>& B) {
  return {mul(A.R, B.R), A.p + mul(A.R, B.p)};
}
template<class T> RBD_HD Mot<T> act(const Xform<T>& X, const Mot<T>& m) {
  const V3<T> w = mul(X.R, m.a);
  return {mul(X.R, m.l) + cross(X.p, w), w};
}
template<class T> RBD_HD Mot<T> actInv(const Xform<T>& X, const Mot<T>& m) {
  return {mulT(X.R, m.l - cross(X.p, m.a)), mulT(X.R, m.a)};
}
template<class T> RBD_HD Frc<T> act(const Xform<T>& X, const Frc<T>& f) {
  const V3<T> l = mul(X.R, f.l);
  return {l, mul(X.R, f.a) + cross(X.p, l)};
}
template<class T> RBD_HD Frc<T> actInv(const Xform<T>& X, const Frc<T>& f) {
  return {mulT(X.R, f.l), mulT(X.R, f.a - cross(X.p, f.l))};
}

// Rigid-body inertia (mass m, lever c, rotational inertia about the CoM Ic = (xx, xy,
// yy, xz, yz, zz)), inertia.hpp.
template<class T> struct Body {
  T m;
  V3<T> c;
  T Ixx, Ixy, Iyy, Ixz, Iyz, Izz;
};
template<class T> RBD_HD V3<T> symMul(T xx, T xy, T yy, T xz, T yz, T zz, V3<T> v) {
  return {xx * v.x + xy * v.y + xz * v.z, xy * v.x + yy * v.y + yz * v.z, xz * v.x + yz * v.y + zz * v.z};
}
// f = Y v  (inertia.hpp:729-735)
template<class T> RBD_HD Frc<T> inertiaMul(const Body<T>& Y, const Mot<T>& v) {
  const V3<T> fl = Y.m * (v.l - cross(Y.c, v.a));
  return {fl, symMul(Y.Ixx, Y.Ixy, Y.Iyy, Y.Ixz, Y.Iyz, Y.Izz, v.a) + cross(Y.c, fl)};
}

// Packed symmetric 6x6: entries (r, c) with r <= c in row order:
// 0:(0,0) 1:(0,1) 2:(0,2) 3:(0,3) 4:(0,4) 5:(0,5) 6:(1,1) 7:(1,2) 8:(1,3) 9:(1,4) 10:(1,5)
// 11:(2,2) 12:(2,3) 13:(2,4) 14:(2,5) 15:(3,3) 16:(3,4) 17:(3,5) 18:(4,4) 19:(4,5) 20:(5,5)
RBD_HD constexpr int symIdx(int r, int c) {
  return r <= c ? (r * 6 - (r * (r - 1)) / 2 + (c - r)) : (c * 6 - (c * (c - 1)) / 2 + (r - c));
}
template<class T> struct Sym6 {
  T a[21];
};
template<class T> RBD_HD T sget(const Sym6<T>& S, int r, int c) { return S.a[symIdx(r, c)]; }

// Body inertia as a 6x6 (inertia.hpp:481-497): [[m I, -m[c]x], [m[c]x, Ic - m[c]x^2]]
template<class T> RBD_HD Sym6<T> bodyMatrix(const Body<T>& Y) {
  Sym6<T> S;
  const T m = Y.m;
  const V3<T> mc = m * Y.c;
  const T cx = Y.c.x, cy = Y.c.y, cz = Y.c.z;
  S.a[0] = m; S.a[1] = T(0); S.a[2] = T(0);
  // -m [c]x rows 0..2, cols 3..5: [[0, m cz, -m cy], [-m cz, 0, m cx], [m cy, -m cx, 0]]
  S.a[3] = T(0); S.a[4] = mc.z; S.a[5] = -mc.y;
  S.a[6] = m; S.a[7] = T(0);
  S.a[8] = -mc.z; S.a[9] = T(0); S.a[10] = mc.x;
  S.a[11] = m;
  S.a[12] = mc.y; S.a[13] = -mc.x; S.a[14] = T(0);
  // Ic - m [c]x^2 = Ic + m (|c|^2 I - c c^T)
  S.a[15] = Y.Ixx + m * (cy * cy + cz * cz);
  S.a[16] = Y.Ixy - m * cx * cy;
  S.a[17] = Y.Ixz - m * cx * cz;
  S.a[18] = Y.Iyy + m * (cx * cx + cz * cz);
  S.a[19] = Y.Iyz - m * cy * cz;
  S.a[20] = Y.Izz + m * (cx * cx + cy * cy);
  return S;
}

template<class T> RBD_HD Frc<T> symMul6(const Sym6<T>& S, const Mot<T>& v) {
  const T x[6] = {v.l.x, v.l.y, v.l.z, v.a.x, v.a.y, v.a.z};
  T y[6];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 6; ++c) s += S.a[symIdx(r, c)] * x[c];
    y[r] = s;
  }
  return {{y[0], y[1], y[2]}, {y[3], y[4], y[5]}};
}
// column k of S as a force
template<class T> RBD_HD Frc<T> symCol(const Sym6<T>& S, int k) {
  return {{S.a[symIdx(0, k)], S.a[symIdx(1, k)], S.a[symIdx(2, k)]},
          {S.a[symIdx(3, k)], S.a[symIdx(4, k)], S.a[symIdx(5, k)]}};
}
template<class T> RBD_HD void symAdd(Sym6<T>& A, const Sym6<T>& B) {
#pragma unroll
  for (int k = 0; k < 21; ++k) A.a[k] += B.a[k];
}
// A -= s * u u^T
template<class T> RBD_HD void symRank1Sub(Sym6<T>& A, const Frc<T>& u, T s) {
  const T x[6] = {u.l.x, u.l.y, u.l.z, u.a.x, u.a.y, u.a.z};
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    const T sr = s * x[r];
#pragma unroll
    for (int c = r; c < 6; ++c) A.a[symIdx(r, c)] -= sr * x[c];
  }
}

// Articulated-inertia transport to the parent frame: X^* I X^{-1} (aba.hxx:21-76).
// With I = [[A, B], [B^T, C]], P = [p]x:  A' = R A R^T, B0 = R B R^T, B' = B0 - A' P,
// C' = R C R^T + P B0 + (P B0)^T - P A' P.
template<class T> RBD_HD Sym6<T> transportUp(const Xform<T>& X, const Sym6<T>& I) {
  const R3<T>& R = X.R;
  // RA = R A (A symmetric), then A' = RA R^T
  T A[3][3], B[3][3], C[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      A[r][c] = I.a[symIdx(r, c)];
      B[r][c] = I.a[symIdx(r, 3 + c)];
      C[r][c] = I.a[symIdx(3 + r, 3 + c)];
    }
  const T Rm[3][3] = {{R.m00, R.m01, R.m02}, {R.m10, R.m11, R.m12}, {R.m20, R.m21, R.m22}};
  T RA[3][3], RB[3][3], RC[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      RA[r][c] = Rm[r][0] * A[0][c] + Rm[r][1] * A[1][c] + Rm[r][2] * A[2][c];
      RB[r][c] = Rm[r][0] * B[0][c] + Rm[r][1] * B[1][c] + Rm[r][2] * B[2][c];
      RC[r][c] = Rm[r][0] * C[0][c] + Rm[r][1] * C[1][c] + Rm[r][2] * C[2][c];
    }
  T Ap[3][3], B0[3][3], Cp[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Ap[r][c] = RA[r][0] * Rm[c][0] + RA[r][1] * Rm[c][1] + RA[r][2] * Rm[c][2];
      B0[r][c] = RB[r][0] * Rm[c][0] + RB[r][1] * Rm[c][1] + RB[r][2] * Rm[c][2];
      Cp[r][c] = RC[r][0] * Rm[c][0] + RC[r][1] * Rm[c][1] + RC[r][2] * Rm[c][2];
    }
  const T px = X.p.x, py = X.p.y, pz = X.p.z;
  // P = [[0,-pz,py],[pz,0,-px],[-py,px,0]]
  const T P[3][3] = {{T(0), -pz, py}, {pz, T(0), -px}, {-py, px, T(0)}};
  T Bp[3][3], PB0[3][3], AP[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      AP[r][c] = Ap[r][0] * P[0][c] + Ap[r][1] * P[1][c] + Ap[r][2] * P[2][c];
      PB0[r][c] = P[r][0] * B0[0][c] + P[r][1] * B0[1][c] + P[r][2] * B0[2][c];
    }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Bp[r][c] = B0[r][c] - AP[r][c];
  Sym6<T> O;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = r; c < 3; ++c) {
      // P A' P = P (A' P)
      const T PAP = P[r][0] * AP[0][c] + P[r][1] * AP[1][c] + P[r][2] * AP[2][c];
      O.a[symIdx(r, c)] = Ap[r][c];
      O.a[symIdx(3 + r, 3 + c)] = Cp[r][c] + PB0[r][c] + PB0[c][r] - PAP;
    }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) O.a[symIdx(r, 3 + c)] = Bp[r][c];
  return O;
}

// Composite rigid-body inertia in "origin" form (m, h = m c, Io = rotational inertia
// about the frame origin).  Additive, so children accumulate without divisions.
template<class T> struct CInertia {
  T m;
  V3<T> h;
  T xx, xy, yy, xz, yz, zz;
};
template<class T> RBD_HD CInertia<T> toComposite(const Body<T>& Y) {
  const T m = Y.m, cx = Y.c.x, cy = Y.c.y, cz = Y.c.z;
  return {m, m * Y.c,
          Y.Ixx + m * (cy * cy + cz * cz), Y.Ixy - m * cx * cy, Y.Iyy + m * (cx * cx + cz * cz),
          Y.Ixz - m * cx * cz, Y.Iyz - m * cy * cz, Y.Izz + m * (cx * cx + cy * cy)};
}
template<class T> RBD_HD void cadd(CInertia<T>& A, const CInertia<T>& B) {
  A.m += B.m; A.h = A.h + B.h;
  A.xx += B.xx; A.xy += B.xy; A.yy += B.yy; A.xz += B.xz; A.yz += B.yz; A.zz += B.zz;
}
// f = Y v for the origin form: f_l = m v_l - h x w ; f_a = Io w + h x v_l
template<class T> RBD_HD Frc<T> cmul(const CInertia<T>& Y, const Mot<T>& v) {
  return {Y.m * v.l - cross(Y.h, v.a), symMul(Y.xx, Y.xy, Y.yy, Y.xz, Y.yz, Y.zz, v.a) + cross(Y.h, v.l)};
}
// Express a composite inertia of frame i in the parent frame: m' = m,
// h' = R h + m p, Io' = R Io R^T + m(|p|^2 I - p p^T) + 2 (p.h') I - p h'^T - h' p^T + ... written
// via the body about the new origin: Io' = R Io R^T - m [p]x^2 - [p]x[Rh]x - [Rh]x[p]x.
template<class T> RBD_HD CInertia<T> transportUp(const Xform<T>& X, const CInertia<T>& Y) {
  const R3<T>& R = X.R;
  // R Io R^T
  const T I[3][3] = {{Y.xx, Y.xy, Y.xz}, {Y.xy, Y.yy, Y.yz}, {Y.xz, Y.yz, Y.zz}};
  const T Rm[3][3] = {{R.m00, R.m01, R.m02}, {R.m10, R.m11, R.m12}, {R.m20, R.m21, R.m22}};
  T RI[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) RI[r][c] = Rm[r][0] * I[0][c] + Rm[r][1] * I[1][c] + Rm[r][2] * I[2][c];
  T O[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = r; c < 3; ++c) O[r][c] = RI[r][0] * Rm[c][0] + RI[r][1] * Rm[c][1] + RI[r][2] * Rm[c][2];
  const V3<T> g = mul(R, Y.h);  // rotated first moment
  const V3<T> p = X.p;
  const T m = Y.m;
  // -m[p]x^2 = m(|p|^2 I - p p^T);  -([p]x[g]x + [g]x[p]x) = 2(p.g) I - p g^T - g p^T
  const T pp = dot(p, p), pg = dot(p, g);
  const T px[3] = {p.x, p.y, p.z}, gx[3] = {g.x, g.y, g.z};
  T Z[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = r; c < 3; ++c)
      Z[r][c] = O[r][c] - m * px[r] * px[c] - px[r] * gx[c] - gx[r] * px[c] + (r == c ? m * pp + T(2) * pg : T(0));
  return {m, g + m * p, Z[0][0], Z[0][1], Z[1][1], Z[0][2], Z[1][2], Z[2][2]};
}
// Rotate / translate a composite inertia to the world (same as transportUp with oMi).
}  // namespace rbd
