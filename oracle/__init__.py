"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  The oracle restates the reference's FP64
algorithms (see rbd_oracle.hpp for the file:line of each).

Python-side shapes follow Eigen's convention of one configuration per column: q is
(nq, B), v/a/tau are (nv, B).  Matrix outputs are returned as (B, nv, nv) arrays whose
[b] entry is the configuration's nv x nv matrix.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

SCALARS = {"f64": 0, "f32": 1, "f80": 2}


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(x):
    return x.ctypes.data_as(ctypes.c_void_p) if x is not None else None


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _cols(x, rows):
    """(rows, B) -> Fortran-ordered float64 copy (config-contiguous)."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    assert x.shape[0] == rows, (x.shape, rows)
    return np.asfortranarray(x)


def _mats(buf, B, nv):
    return buf.reshape(B, nv, nv).transpose(0, 2, 1).copy()


def dims(model):
    nq, nv, compact = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    d = model.cstruct()
    _check(lib().orc_dims(ctypes.byref(d), ctypes.byref(nq), ctypes.byref(nv), ctypes.byref(compact)))
    return nq.value, nv.value, bool(compact.value)


def topology(model):
    nj, nv = model.njoints, model.nv
    out = [np.zeros(nj, np.int32) for _ in range(3)] + [np.zeros(nv, np.int32) for _ in range(2)]
    d = model.cstruct()
    _check(lib().orc_topology(ctypes.byref(d), *[_p(o) for o in out]))
    return dict(zip(["idx_q", "idx_v", "nvSubtree", "parents_fromRow", "nvSubtree_fromRow"], out))


def rnea(model, q, v, a, scalar="f64", nthreads=0):
    q, v, a = _cols(q, model.nq), _cols(v, model.nv), _cols(a, model.nv)
    B = q.shape[1]
    tau = np.zeros((model.nv, B), order="F")
    d = model.cstruct()
    _check(lib().orc_rnea(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(a), _p(tau),
                          SCALARS[scalar], nthreads))
    return np.ascontiguousarray(tau)


def forward_kinematics(model, q, v=None, a=None, scalar="f64", nthreads=0):
    """(oMi, v, a): oMi (B, nj-1, 12) = world placement of joints 1.. (R row major, t);
    v, a (B, nj-1, 6) local spatial velocity / acceleration (None unless v / v and a given)."""
    q = _cols(q, model.nq)
    B = q.shape[1]
    nj = model.njoints
    vv = _cols(v, model.nv) if v is not None else None
    aa = _cols(a, model.nv) if a is not None else None
    oM = np.zeros((B, nj - 1, 12))
    vo = np.zeros((B, nj - 1, 6)) if v is not None else None
    ao = np.zeros((B, nj - 1, 6)) if (v is not None and a is not None) else None
    d = model.cstruct()
    _check(lib().orc_forward_kinematics(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(vv), _p(aa), _p(oM), _p(vo),
                                        _p(ao), SCALARS[scalar], nthreads))
    return oM, vo, ao


def joint_jacobians(model, q, scalar="f64", nthreads=0):
    """(B, 6, nv) world-frame joint Jacobians (computeJointJacobians)."""
    q = _cols(q, model.nq)
    B = q.shape[1]
    J = np.zeros((B, model.nv, 6))
    d = model.cstruct()
    _check(lib().orc_joint_jacobians(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(J), SCALARS[scalar], nthreads))
    return J.transpose(0, 2, 1).copy()


def joint_jacobians_time_variation(model, q, v, scalar="f64", nthreads=0):
    """(J, dJ), each (B, 6, nv): computeJointJacobiansTimeVariation (WORLD frame)."""
    q, v = _cols(q, model.nq), _cols(v, model.nv)
    B = q.shape[1]
    J = np.zeros((B, model.nv, 6))
    dJ = np.zeros((B, model.nv, 6))
    d = model.cstruct()
    _check(lib().orc_joint_jacobians_time_variation(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(J), _p(dJ),
                                                    SCALARS[scalar], nthreads))
    return J.transpose(0, 2, 1).copy(), dJ.transpose(0, 2, 1).copy()


REFERENCE_FRAMES = {"world": 0, "local": 1, "local_world_aligned": 2}


def frames(model, q, parents, placements, reference_frame="world", scalar="f64", nthreads=0):
    """updateFramePlacements + getFrameJacobian per column for frames given as (parent joint,
    placement [R row-major 9, p 3]).  Returns oMf (B, nf, 12) and J (B, nf, 6, nv)."""
    q = _cols(q, model.nq)
    B = q.shape[1]
    par = np.ascontiguousarray(parents, dtype=np.int32)
    pl = np.ascontiguousarray(placements, dtype=np.float64).reshape(len(par), 12)
    nf = len(par)
    oMf = np.zeros((B, nf, 12))
    J = np.zeros((B, nf, model.nv, 6))
    d = model.cstruct()
    _check(lib().orc_frames(ctypes.byref(d), ctypes.c_int64(B), _p(q), int(nf), _p(par), _p(pl),
                            REFERENCE_FRAMES[reference_frame], _p(oMf), _p(J), SCALARS[scalar], nthreads))
    return oMf, J.transpose(0, 1, 3, 2).copy()


def frame_jacobians_time_variation(model, q, v, parents, placements, reference_frame="local", scalar="f64",
                                   nthreads=0):
    """getFrameJacobianTimeVariation after computeJointJacobiansTimeVariation, frames as in
    frames(): dJ (B, nf, 6, nv)."""
    q, v = _cols(q, model.nq), _cols(v, model.nv)
    B = q.shape[1]
    par = np.ascontiguousarray(parents, dtype=np.int32)
    pl = np.ascontiguousarray(placements, dtype=np.float64).reshape(len(par), 12)
    nf = len(par)
    dJ = np.zeros((B, nf, model.nv, 6))
    d = model.cstruct()
    _check(lib().orc_frame_jacobians_time_variation(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), int(nf), _p(par),
                                                    _p(pl), REFERENCE_FRAMES[reference_frame], _p(dJ),
                                                    SCALARS[scalar], nthreads))
    return dJ.transpose(0, 1, 3, 2).copy()


def frame_motions(model, q, v, parents, placements, a=None, reference_frame="local", scalar="f64", nthreads=0):
    """getFrameVelocity / getFrameAcceleration / getFrameClassicalAcceleration after
    forwardKinematics(q, v[, a]) for frames given as (parent joint, placement [R row-major 9,
    p 3]).  Returns vf, af, acf (B, nf, 6) [linear, angular]; af and acf are None without a."""
    q, v = _cols(q, model.nq), _cols(v, model.nv)
    aa = _cols(a, model.nv) if a is not None else None
    B = q.shape[1]
    par = np.ascontiguousarray(parents, dtype=np.int32)
    pl = np.ascontiguousarray(placements, dtype=np.float64).reshape(len(par), 12)
    nf = len(par)
    vf = np.zeros((B, nf, 6))
    af = np.zeros((B, nf, 6)) if a is not None else None
    acf = np.zeros((B, nf, 6)) if a is not None else None
    d = model.cstruct()
    _check(lib().orc_frame_motions(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(aa), int(nf), _p(par), _p(pl),
                                   REFERENCE_FRAMES[reference_frame], _p(vf), _p(af), _p(acf), SCALARS[scalar],
                                   nthreads))
    return vf, af, acf


def aba(model, q, v, tau, convention="world", scalar="f64", nthreads=0):
    q, v, tau = _cols(q, model.nq), _cols(v, model.nv), _cols(tau, model.nv)
    B = q.shape[1]
    ddq = np.zeros((model.nv, B), order="F")
    d = model.cstruct()
    conv = {"local": 0, "world": 1}[convention]
    _check(lib().orc_aba(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(tau), _p(ddq), conv,
                         SCALARS[scalar], nthreads))
    return np.ascontiguousarray(ddq)


def crba(model, q, symmetrize=True, scalar="f64", nthreads=0):
    q = _cols(q, model.nq)
    B, nv = q.shape[1], model.nv
    M = np.zeros(B * nv * nv)
    d = model.cstruct()
    _check(lib().orc_crba(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(M), int(symmetrize), SCALARS[scalar],
                          nthreads))
    return _mats(M, B, nv)


def cholesky(model, M, rhs=None, scalar="f64", nthreads=0):
    """Tree-sparse UDU^T of each M[b]; returns (U, D, Minv, x) with x = M^{-1} rhs."""
    M = np.asarray(M, dtype=np.float64)
    B, nv = M.shape[0], model.nv
    Mc = np.ascontiguousarray(M.transpose(0, 2, 1)).ravel()
    U = np.zeros(B * nv * nv)
    D = np.zeros(B * nv)
    Minv = np.zeros(B * nv * nv)
    x = None
    if rhs is not None:
        x = np.asfortranarray(np.asarray(rhs, dtype=np.float64).reshape(nv, B)).copy(order="F")
    d = model.cstruct()
    _check(lib().orc_cholesky(ctypes.byref(d), ctypes.c_int64(B), _p(Mc), _p(U), _p(D), _p(Minv), _p(x),
                              SCALARS[scalar], nthreads))
    return _mats(U, B, nv), D.reshape(B, nv), _mats(Minv, B, nv), (np.ascontiguousarray(x) if x is not None else None)


def rnea_derivatives(model, q, v, a, scalar="f64", nthreads=0):
    """Returns (dtau_dq, dtau_dv, M_upper, tau); M has only its upper triangle, as the reference."""
    q, v, a = _cols(q, model.nq), _cols(v, model.nv), _cols(a, model.nv)
    B, nv = q.shape[1], model.nv
    dq, dv, M = (np.zeros(B * nv * nv) for _ in range(3))
    tau = np.zeros((nv, B), order="F")
    d = model.cstruct()
    _check(lib().orc_rnea_derivatives(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(a), _p(dq), _p(dv),
                                      _p(M), _p(tau), SCALARS[scalar], nthreads))
    return _mats(dq, B, nv), _mats(dv, B, nv), _mats(M, B, nv), np.ascontiguousarray(tau)


def aba_derivatives(model, q, v, tau, scalar="f64", nthreads=0):
    """Returns (ddq_dq, ddq_dv, Minv, ddq)."""
    q, v, tau = _cols(q, model.nq), _cols(v, model.nv), _cols(tau, model.nv)
    B, nv = q.shape[1], model.nv
    dq, dv, Mi = (np.zeros(B * nv * nv) for _ in range(3))
    ddq = np.zeros((nv, B), order="F")
    d = model.cstruct()
    _check(lib().orc_aba_derivatives(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(tau), _p(dq), _p(dv),
                                     _p(Mi), _p(ddq), SCALARS[scalar], nthreads))
    return _mats(dq, B, nv), _mats(dv, B, nv), _mats(Mi, B, nv), np.ascontiguousarray(ddq)


def integrate(model, q, v):
    q, v = _cols(q, model.nq), _cols(v, model.nv)
    B = q.shape[1]
    out = np.zeros((model.nq, B), order="F")
    d = model.cstruct()
    _check(lib().orc_integrate(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(out)))
    return np.ascontiguousarray(out)


def energy(model, q, v, scalar="f64"):
    q, v = _cols(q, model.nq), _cols(v, model.nv)
    B = q.shape[1]
    T, V = np.zeros(B), np.zeros(B)
    d = model.cstruct()
    _check(lib().orc_energy(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(T), _p(V), SCALARS[scalar]))
    return T, V


FLOP_ALGS = {"rnea": 0, "aba_world": 1, "aba_local": 2, "crba": 3, "cholesky": 4, "rnea_derivatives": 5,
             "aba_derivatives": 6, "cholesky_solve": 7}


def flops(model, alg, q=None, v=None, x=None):
    """Algorithmic floating-point operations of one configuration of `alg`, counted by
    running the oracle's restatement with a counting scalar (flop_count.hpp)."""
    from pinocchio_amd import models as M
    if q is None:
        q, v, x, _ = M.sample_inputs(model, 1, 4242)
        q, v, x = q[:, 0], v[:, 0], x[:, 0]
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = ctypes.c_longlong()
    d = model.cstruct()
    _check(lib().orc_flops(ctypes.byref(d), FLOP_ALGS[alg], _p(q), _p(v), _p(x), ctypes.byref(out)))
    return out.value


def constrained_forward_dynamics(model, q, v, tau, J, gamma, inv_damping=0.0, scalar="f64", nthreads=0):
    """forwardDynamics(model, data, q, v, tau, J, gamma, inv_damping) per column
    (contact-dynamics.hxx:90-167).  J: (B, nc, nv); gamma: (nc, B).  Returns (ddq, lambda)
    as (nv, B) and (nc, B)."""
    q, v, tau = _cols(q, model.nq), _cols(v, model.nv), _cols(tau, model.nv)
    B, nv = q.shape[1], model.nv
    J = np.asarray(J, dtype=np.float64)
    assert J.ndim == 3 and J.shape[0] == B and J.shape[2] == nv, J.shape
    nc = J.shape[1]
    Jc = np.ascontiguousarray(J.transpose(0, 2, 1))  # per configuration column major
    gamma = _cols(gamma, nc)
    ddq = np.zeros((nv, B), order="F")
    lam = np.zeros((nc, B), order="F")
    d = model.cstruct()
    _check(lib().orc_constrained_forward_dynamics(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(tau),
                                                  int(nc), _p(Jc), _p(gamma), ctypes.c_double(inv_damping),
                                                  _p(ddq), _p(lam), SCALARS[scalar], nthreads))
    return np.ascontiguousarray(ddq), np.ascontiguousarray(lam)


def impulse_dynamics(model, q, v_before, J, r_coeff=0.0, inv_damping=0.0, scalar="f64", nthreads=0):
    """impulseDynamics(model, data, q, v_before, J, r_coeff, inv_damping) per column
    (contact-dynamics.hxx:264-331).  J: (B, nc, nv).  Returns (dq_after (nv, B), impulse (nc, B))."""
    q, v = _cols(q, model.nq), _cols(v_before, model.nv)
    B, nv = q.shape[1], model.nv
    J = np.asarray(J, dtype=np.float64)
    assert J.ndim == 3 and J.shape[0] == B and J.shape[2] == nv, J.shape
    nc = J.shape[1]
    Jc = np.ascontiguousarray(J.transpose(0, 2, 1))
    dqa = np.zeros((nv, B), order="F")
    imp = np.zeros((nc, B), order="F")
    d = model.cstruct()
    _check(lib().orc_impulse_dynamics(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), int(nc), _p(Jc),
                                      ctypes.c_double(r_coeff), ctypes.c_double(inv_damping), _p(dqa), _p(imp),
                                      SCALARS[scalar], nthreads))
    return np.ascontiguousarray(dqa), np.ascontiguousarray(imp)


def time_constrained(model, q, v, tau, J, gamma, inv_damping=0.0, nthreads=0, reps=5):
    """Median wall time (s) of the OpenMP batch loop of constrained forward dynamics (FP64)."""
    q, v, tau = _cols(q, model.nq), _cols(v, model.nv), _cols(tau, model.nv)
    B, nv = q.shape[1], model.nv
    J = np.asarray(J, dtype=np.float64)
    nc = J.shape[1]
    Jc = np.ascontiguousarray(J.transpose(0, 2, 1))
    gamma = _cols(gamma, nc)
    sec = ctypes.c_double()
    d = model.cstruct()
    _check(lib().orc_time_constrained(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(tau), int(nc), _p(Jc),
                                      _p(gamma), ctypes.c_double(inv_damping), int(nthreads), int(reps),
                                      ctypes.byref(sec)))
    return sec.value


TIME_ALGS = {"rnea": 0, "aba": 1, "crba": 2, "crba_cholesky": 3, "step": 4, "rnea_derivatives": 5,
             "aba_derivatives": 6}


def time_batch(model, alg, q, v, x, nthreads=0, reps=5):
    """Median wall time (s) of the OpenMP batch loop over the columns of q, v, x."""
    q, v, x = _cols(q, model.nq), _cols(v, model.nv), _cols(x, model.nv)
    B = q.shape[1]
    sec = ctypes.c_double()
    d = model.cstruct()
    _check(lib().orc_time(ctypes.byref(d), TIME_ALGS[alg], ctypes.c_int64(B), _p(q), _p(v), _p(x), int(nthreads),
                          int(reps), ctypes.byref(sec)))
    return sec.value


CONTACT_3D, CONTACT_6D = 0, 1
LOCAL, LOCAL_WORLD_ALIGNED = 1, 2


class _Constraint(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("joint1", ctypes.c_int32), ("joint2", ctypes.c_int32),
                ("rf", ctypes.c_int32), ("placement1", ctypes.c_double * 12), ("placement2", ctypes.c_double * 12),
                ("Kp", ctypes.c_double * 6), ("Kd", ctypes.c_double * 6), ("desired", ctypes.c_double * 6)]


def constraint_table(constraints):
    """ctypes array of rigid constraints from dicts with keys type, joint1, joint2,
    reference_frame, placement1, placement2 (12 numbers: R row major, p), Kp, Kd, desired
    (3 or 6 numbers; missing ones are zero)."""
    arr = (_Constraint * len(constraints))()
    for k, c in enumerate(constraints):
        a = arr[k]
        a.type = int(c["type"])
        a.joint1 = int(c.get("joint1", 0))
        a.joint2 = int(c.get("joint2", 0))
        a.rf = int(c.get("reference_frame", LOCAL))
        eye = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
        for name in ("placement1", "placement2"):
            vals = np.asarray(c.get(name, eye), dtype=np.float64).ravel()
            assert vals.size == 12
            getattr(a, name)[:] = list(vals)
        for name in ("Kp", "Kd", "desired"):
            vals = np.zeros(6)
            x = np.asarray(c.get(name, []), dtype=np.float64).ravel()
            vals[:x.size] = x
            getattr(a, name)[:] = list(vals)
    return arr


def constraint_dynamics(model, q, v, tau, constraints, mu=0.0, max_iter=1, absolute_accuracy=1e-12,
                        relative_accuracy=1e-12, scalar="f64", nthreads=0):
    """constraintDynamics(model, data, q, v, tau, contact_models, contact_datas, settings) per
    column (constrained-dynamics.hxx:193-500).  constraints: list of dicts (constraint_table).
    Returns (ddq (nv, B), lambda (nrows, B), iterations (B,))."""
    q, v, tau = _cols(q, model.nq), _cols(v, model.nv), _cols(tau, model.nv)
    B, nv = q.shape[1], model.nv
    arr = constraint_table(constraints)
    nrows = sum(6 if c["type"] == CONTACT_6D else 3 for c in constraints)
    ddq = np.zeros((nv, B), order="F")
    lam = np.zeros((nrows, B), order="F")
    its = np.zeros(B, dtype=np.int32)
    d = model.cstruct()
    _check(lib().orc_constraint_dynamics(ctypes.byref(d), ctypes.c_int64(B), _p(q), _p(v), _p(tau), len(constraints),
                                         arr, ctypes.c_double(absolute_accuracy), ctypes.c_double(relative_accuracy),
                                         ctypes.c_double(mu), int(max_iter), _p(ddq), _p(lam),
                                         its.ctypes.data_as(ctypes.c_void_p), SCALARS[scalar], nthreads))
    return np.ascontiguousarray(ddq), np.ascontiguousarray(lam), its


def delassus(model, q, constraints, mu=0.0, scalar="f64", nthreads=0):
    """computeDelassusMatrix(model, data, q, contact_models, contact_datas, G, mu) per column
    (delassus.hxx:118-300): J (M + mu J^T J)^-1 J^T.  Returns (B, nrows, nrows)."""
    q = _cols(q, model.nq)
    B = q.shape[1]
    arr = constraint_table(constraints)
    nrows = sum(6 if c["type"] == CONTACT_6D else 3 for c in constraints)
    G = np.zeros(B * nrows * nrows)
    d = model.cstruct()
    _check(lib().orc_delassus(ctypes.byref(d), ctypes.c_int64(B), _p(q), len(constraints), arr, ctypes.c_double(mu),
                              _p(G), SCALARS[scalar], nthreads))
    return _mats(G, B, nrows)


def _problems(G, g, mu, nc):
    """(B, n, n) row-indexed Delassus matrices, (B, n) vectors, (B, nc) or (nc,) friction
    coefficients -> column-major contiguous copies."""
    n = 3 * nc
    G = np.asarray(G, dtype=np.float64)
    if G.ndim == 2:
        G = G[None]
    B = G.shape[0]
    assert G.shape == (B, n, n), (G.shape, n)
    Gc = np.ascontiguousarray(G.transpose(0, 2, 1))
    g = np.ascontiguousarray(np.broadcast_to(np.asarray(g, dtype=np.float64), (B, n)))
    mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (B, nc)))
    return B, Gc, g, mu


def pgs_solve(G, g, mu, x0=None, over_relax=1.0, max_it=1000, absolute_precision=1e-6, relative_precision=1e-6,
              nthreads=0):
    """PGSContactSolverTpl::solve (pgs-solver.hxx:19-141) on B problems: G (B, n, n), g (B, n),
    mu (B, nc) or (nc,), x0 (B, n) initial guess (zero if None).  Returns dict x, it,
    absolute_residual, relative_residual, converged."""
    nc = np.asarray(mu).shape[-1]
    B, Gc, g, mu = _problems(G, g, mu, nc)
    n = 3 * nc
    x = np.zeros((B, n)) if x0 is None else np.array(np.broadcast_to(x0, (B, n)), dtype=np.float64)
    it = np.zeros(B, dtype=np.int32)
    res = np.zeros((B, 2))
    conv = np.zeros(B, dtype=np.uint8)
    _check(lib().orc_pgs_solve(ctypes.c_int64(B), nc, _p(Gc), _p(g), _p(mu), _p(x), ctypes.c_double(over_relax),
                               int(max_it), ctypes.c_double(absolute_precision), ctypes.c_double(relative_precision),
                               _p(it), _p(res), _p(conv), nthreads))
    return dict(x=x, it=it, absolute_residual=res[:, 0], relative_residual=res[:, 1], converged=conv.astype(bool))


def admm_solve(G, g, mu, R=None, primal_guess=None, dual_guess=None, mu_prox=1e-6, tau=0.5, rho_power=0.2,
               rho_power_factor=0.05, ratio_primal_dual=10.0, max_it_largest_eigen_value=20, max_it=1000,
               absolute_precision=1e-6, relative_precision=1e-6, nthreads=0):
    """ADMMContactSolverTpl::solve (admm-solver.hxx:22-335) on B problems, each with a fresh
    solver and a fresh dense Delassus operator.  R (B, n) compliance (zero if None).
    Returns dict y (primal solution), z (dual), s (shift), it, absolute_residual,
    relative_residual, rho, rho_power, largest_eigen_value, cholesky_update_count,
    converged."""
    nc = np.asarray(mu).shape[-1]
    B, Gc, g, mu = _problems(G, g, mu, nc)
    n = 3 * nc
    R = np.zeros((B, n)) if R is None else np.ascontiguousarray(np.broadcast_to(np.asarray(R, float), (B, n)))
    pg = None if primal_guess is None else np.ascontiguousarray(np.broadcast_to(primal_guess, (B, n)), dtype=float)
    dg = None if dual_guess is None else np.ascontiguousarray(np.broadcast_to(dual_guess, (B, n)), dtype=float)
    st = np.array([mu_prox, tau, rho_power, rho_power_factor, ratio_primal_dual, absolute_precision,
                   relative_precision], dtype=np.float64)
    y, z, s = (np.zeros((B, n)) for _ in range(3))
    it = np.zeros(B, dtype=np.int32)
    stats = np.zeros((B, 5))
    chol = np.zeros(B, dtype=np.int32)
    conv = np.zeros(B, dtype=np.uint8)
    _check(lib().orc_admm_solve(ctypes.c_int64(B), nc, _p(Gc), _p(g), _p(mu), _p(R), _p(pg), _p(dg), _p(st),
                                int(max_it), int(max_it_largest_eigen_value), _p(y), _p(z), _p(s), _p(it),
                                _p(stats), _p(chol), _p(conv), nthreads))
    return dict(y=y, z=z, s=s, it=it, absolute_residual=stats[:, 0], relative_residual=stats[:, 1], rho=stats[:, 2],
                rho_power=stats[:, 3], largest_eigen_value=stats[:, 4], cholesky_update_count=chol,
                converged=conv.astype(bool))


def contact_cholesky_solve(model, q, constraints, mu, rhs, scalar="f64", nthreads=0):
    """ContactCholeskyDecomposition compute + solve per column (contact-cholesky.hpp:292-400):
    x = K^-1 rhs with K = [[-mu I, J], [J^T, M]]; rhs (nrows + nv, B).  Returns x (same shape)."""
    q = _cols(q, model.nq)
    B = q.shape[1]
    arr = constraint_table(constraints)
    nrows = sum(6 if c["type"] == CONTACT_6D else 3 for c in constraints)
    r = _cols(rhs, nrows + model.nv)
    x = np.zeros((nrows + model.nv, B), order="F")
    d = model.cstruct()
    _check(lib().orc_contact_cholesky_solve(ctypes.byref(d), ctypes.c_int64(B), _p(q), len(constraints), arr,
                                            ctypes.c_double(mu), _p(r), _p(x), SCALARS[scalar], nthreads))
    return np.ascontiguousarray(x)


def _fext(model, alg, q, v, x, fext, scalar, nthreads):
    q, v, x = _cols(q, model.nq), _cols(v, model.nv), _cols(x, model.nv)
    B = q.shape[1]
    f = _cols(fext, 6 * (model.njoints - 1))
    out = np.zeros((model.nv, B), order="F")
    d = model.cstruct()
    _check(lib().orc_fext(ctypes.byref(d), ctypes.c_int64(B), alg, _p(q), _p(v), _p(x), _p(f), _p(out),
                          SCALARS[scalar], nthreads))
    return np.ascontiguousarray(out)


def rnea_fext(model, q, v, a, fext, scalar="f64", nthreads=0):
    """rnea(model, data, q, v, a, fext) per column (rnea.hpp:60-79): fext (6 (njoints - 1), B),
    each joint's external force [linear; angular] in its local frame."""
    return _fext(model, 0, q, v, a, fext, scalar, nthreads)


def aba_fext(model, q, v, tau, fext, scalar="f64", nthreads=0):
    """aba(model, data, q, v, tau, fext) per column (aba.hpp:63-85), fext as in rnea_fext."""
    return _fext(model, 1, q, v, tau, fext, scalar, nthreads)
