"""pinocchio_amd -- MI355X batched rigid-body dynamics behind pinocchio's batch API.

Python binding of libpinocchio_amd.so (C ABI in include/rbd.h).  Two levels:

* the drop-in mirror of the reference's Python API for the batched path
  (bindings/python/algorithm/parallel/{rnea,aba}.cpp): ``ModelPool``,
  ``rneaInParallel(num_threads, pool, q, v, a)``, ``abaInParallel(num_threads, pool, q, v, tau)``
  taking and returning numpy matrices with one configuration per column;
* ``Engine``: device-resident calls on torch tensors (no host round trip), used by the
  benchmarks and the GPU parity tests.

There is no CPU fallback: if the library is missing or no GPU is present the calls
raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import models  # noqa: F401
from .models import ModelDesc

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpinocchio_amd.so")

RBD_F32, RBD_F64 = 0, 1
RBD_HOST_PTRS, RBD_ASYNC, RBD_LAYOUT_BATCH, RBD_UPPER_ONLY = 0x1, 0x2, 0x4, 0x8
RBD_NO_GRAPH, RBD_NO_AUTOTUNE = 0x10, 0x20
_ERRORS = {1: ValueError, 2: NotImplementedError, 3: RuntimeError, 4: MemoryError}

_lib = None


def lib():
    """Load libpinocchio_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C pinocchio_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        L.rbd_last_error.restype = ctypes.c_char_p
        L.rbd_version.restype = ctypes.c_char_p
        L.rbd_ctx_stream.restype = ctypes.c_void_p
        L.rbd_ctx_stream.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        msg = lib().rbd_last_error().decode()
        raise _ERRORS.get(rc, RuntimeError)(msg)


def _i64(x):
    return ctypes.c_int64(int(x))


class Context:
    """A HIP stream on one GPU (rbd_ctx)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().rbd_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    @property
    def stream(self) -> int:
        return lib().rbd_ctx_stream(self.handle)

    def synchronize(self):
        _check(lib().rbd_ctx_synchronize(self.handle))

    def capture(self):
        """Record the engine calls made on this context inside a ``with`` block into a HIP
        graph (rbd_capture_begin / _end) instead of running them::

            with ctx.capture() as cap:
                eng.aba(Q, V, T, out=ddq, async_=True)
            cap.graph.launch()        # replays; the inputs are read again each time

        Calls inside the block take preallocated device outputs (``out=``) and must have run
        once, uncaptured, with the same sizes."""
        return _Capture(self)

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rbd_ctx_destroy(self.handle)
            self.handle = None


class Graph:
    """A captured sequence of engine calls (rbd_graph)."""

    def __init__(self, ctx: Context, handle):
        self.ctx, self.handle = ctx, handle

    def launch(self, async_=False):
        _check(lib().rbd_graph_launch(self.ctx.handle, self.handle, ctypes.c_uint32(RBD_ASYNC if async_ else 0)))

    def close(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rbd_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class _Capture:
    def __init__(self, ctx: Context):
        self.ctx, self.graph = ctx, None

    def __enter__(self):
        _check(lib().rbd_capture_begin(self.ctx.handle))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        rc = lib().rbd_capture_end(self.ctx.handle, ctypes.byref(h))
        if exc_type is not None:  # the block failed: drop whatever was recorded
            if rc == 0:
                Graph(self.ctx, h).close()
            return False
        _check(rc)
        self.graph = Graph(self.ctx, h)
        return False


RBD_MODEL_RUNTIME_TOPOLOGY = 0x1


class Model:
    """A model uploaded to a context (rbd_model).

    Any tree of revolute X/Y/Z joints with free-flyer roots is accepted.  Trees the
    library has prebuilt kernels for run on those (``prebuilt`` is True); every other tree,
    or any tree with ``runtime_topology=True``, runs on the run-time-topology kernels."""

    def __init__(self, ctx: Context, desc: ModelDesc, runtime_topology: bool = False):
        self.ctx = ctx
        self.desc = desc
        self._cdesc = desc.cstruct()
        h = ctypes.c_void_p()
        flags = RBD_MODEL_RUNTIME_TOPOLOGY if runtime_topology else 0
        _check(lib().rbd_model_create_ex(ctx.handle, ctypes.byref(self._cdesc), ctypes.c_uint32(flags),
                                         ctypes.byref(h)))
        self.handle = h
        nq, nv, nj, compact, pre = (ctypes.c_int32() for _ in range(5))
        _check(lib().rbd_model_dims(h, ctypes.byref(nq), ctypes.byref(nv), ctypes.byref(nj), ctypes.byref(compact)))
        _check(lib().rbd_model_kernels(h, ctypes.byref(pre)))
        self.nq, self.nv, self.njoints, self.compact = nq.value, nv.value, nj.value, bool(compact.value)
        self.prebuilt = bool(pre.value)

    def cholesky_pattern(self):
        n = ctypes.c_int32()
        _check(lib().rbd_model_cholesky_pattern(self.handle, ctypes.byref(n), None, None))
        rows = np.zeros(n.value, np.int32)
        cols = np.zeros(n.value, np.int32)
        _check(lib().rbd_model_cholesky_pattern(self.handle, ctypes.byref(n), rows.ctypes.data_as(ctypes.c_void_p),
                                                cols.ctypes.data_as(ctypes.c_void_p)))
        return rows, cols

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rbd_model_destroy(self.handle)
            self.handle = None


# ------------------------------------------------------------------------------------
# Device-resident engine over torch tensors
# ------------------------------------------------------------------------------------
def _dtype_code(t):
    import torch
    if t.dtype == torch.float32:
        return RBD_F32
    if t.dtype == torch.float64:
        return RBD_F64
    raise TypeError(f"unsupported dtype {t.dtype}")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Engine:
    """Batched dynamics on device tensors.

    layout="col": tensors are (B, rows), one configuration per row of memory (the
    Eigen column-per-configuration matrix, column major).  layout="batch": tensors are
    (rows, B) with the batch index fastest (the engine's native layout).  Matrix outputs
    are (B, nv, nv) for "col" (element [b, c, r] = M_b(r, c)) and (nv*nv, B) for "batch"
    (row r + c*nv).
    """

    def __init__(self, model: Model, layout: str = "col"):
        assert layout in ("col", "batch")
        self.model = model
        self.layout = layout

    # helpers -----------------------------------------------------------------------
    def _flags(self, async_=False):
        f = RBD_LAYOUT_BATCH if self.layout == "batch" else 0
        return f | (RBD_ASYNC if async_ else 0)

    def _vec(self, t, rows, name):
        # a dimension of extent 1 may carry any stride
        if self.layout == "col":
            ok = t.dim() == 2 and t.shape[1] == rows and (t.stride(1) == 1 or rows == 1)
            assert ok, f"{name}: expected (B, {rows}) rows-contiguous"
            return t.shape[0], (t.stride(0) if t.shape[0] > 1 else rows)
        ok = t.dim() == 2 and t.shape[0] == rows and (t.stride(1) == 1 or t.shape[1] == 1)
        assert ok, f"{name}: expected ({rows}, B) contiguous"
        return t.shape[1], (t.stride(0) if rows > 1 else max(t.shape[1], 1))

    def _new_vec(self, like, rows, B):
        import torch
        if self.layout == "col":
            return torch.empty((B, rows), dtype=like.dtype, device=like.device)
        return torch.empty((rows, B), dtype=like.dtype, device=like.device)

    def _new_mat(self, like, B, dtype=None, zero=False):
        # every kernel writes every entry of its matrices (structural zeros included),
        # so fresh outputs need no fill unless only one triangle will be written
        import torch
        nv = self.model.nv
        dt = dtype or like.dtype
        alloc = torch.zeros if zero else torch.empty
        if self.layout == "col":
            return alloc((B, nv, nv), dtype=dt, device=like.device)
        return alloc((nv * nv, B), dtype=dt, device=like.device)

    def _mat_ld(self, t):
        nv = self.model.nv
        return nv * nv if self.layout == "col" else t.stride(0)

    # algorithms --------------------------------------------------------------------
    def rnea(self, q, v, a, out=None, async_=False):
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        _, lda = self._vec(a, m.nv, "a")
        out = self._new_vec(q, m.nv, B) if out is None else out
        _, ldo = self._vec(out, m.nv, "tau")
        _check(lib().rbd_rnea(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v), _i64(ldv),
                              _ptr(a), _i64(lda), _ptr(out), _i64(ldo), ctypes.c_uint32(self._flags(async_))))
        return out

    def aba(self, q, v, tau, out=None, async_=False):
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        _, ldt = self._vec(tau, m.nv, "tau")
        out = self._new_vec(q, m.nv, B) if out is None else out
        _, ldo = self._vec(out, m.nv, "ddq")
        _check(lib().rbd_aba(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v), _i64(ldv),
                             _ptr(tau), _i64(ldt), _ptr(out), _i64(ldo), ctypes.c_uint32(self._flags(async_))))
        return out

    def nonlinear_effects(self, q, v, out=None, async_=False):
        """nonLinearEffects: C(q, v) v + g(q) per configuration (rnea with a = 0)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        out = self._new_vec(q, m.nv, B) if out is None else out
        _, ldo = self._vec(out, m.nv, "nle")
        _check(lib().rbd_nonlinear_effects(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                           _ptr(v), _i64(ldv), _ptr(out), _i64(ldo),
                                           ctypes.c_uint32(self._flags(async_))))
        return out

    def generalized_gravity(self, q, out=None, async_=False):
        """computeGeneralizedGravity: g(q) per configuration (rnea with v = a = 0)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        out = self._new_vec(q, m.nv, B) if out is None else out
        _, ldo = self._vec(out, m.nv, "g")
        _check(lib().rbd_generalized_gravity(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                             _ptr(out), _i64(ldo), ctypes.c_uint32(self._flags(async_))))
        return out

    def rnea_fext(self, q, v, a, fext, out=None, async_=False):
        """rnea(model, data, q, v, a, fext) per configuration (rnea.hpp:60-79): fext has
        6 (njoints - 1) rows, each joint's external force [linear; angular] in its frame."""
        return self._fext("rbd_rnea_fext", q, v, a, fext, out, async_)

    def aba_fext(self, q, v, tau, fext, out=None, async_=False):
        """aba(model, data, q, v, tau, fext) per configuration (aba.hpp:63-85)."""
        return self._fext("rbd_aba_fext", q, v, tau, fext, out, async_)

    def static_torque(self, q, fext, out=None, async_=False):
        """computeStaticTorque (rnea.hpp:163): g(q) with external forces, rnea(q, 0, 0, fext);
        fext as in rnea_fext."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldf = self._vec(fext, 6 * (m.njoints - 1), "fext")
        out = self._new_vec(q, m.nv, B) if out is None else out
        _, ldo = self._vec(out, m.nv, "tau")
        _check(lib().rbd_static_torque(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                       _ptr(fext), _i64(ldf), _ptr(out), _i64(ldo),
                                       ctypes.c_uint32(self._flags(async_))))
        return out

    def _fext(self, fn, q, v, x, fext, out, async_):
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        _, ldx = self._vec(x, m.nv, "a/tau")
        _, ldf = self._vec(fext, 6 * (m.njoints - 1), "fext")
        out = self._new_vec(q, m.nv, B) if out is None else out
        _, ldo = self._vec(out, m.nv, "out")
        _check(getattr(lib(), fn)(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v),
                                  _i64(ldv), _ptr(x), _i64(ldx), _ptr(fext), _i64(ldf), _ptr(out), _i64(ldo),
                                  ctypes.c_uint32(self._flags(async_))))
        return out

    def _mat_arg(self, t, B, name):
        nv = self.model.nv
        if self.layout == "col":
            assert t.shape == (B, nv, nv) and t.is_contiguous(), f"{name}: expected contiguous (B, nv, nv)"
            return nv * nv
        assert t.dim() == 2 and t.shape[0] == nv * nv and t.shape[1] == B and t.stride(1) == 1, f"{name}: (nv*nv, B)"
        return t.stride(0) if B > 0 else max(B, 1)

    def crba(self, q, out=None, upper_only=False, async_=False):
        """Joint-space inertia matrix per configuration (symmetric unless upper_only)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        out = self._new_mat(q, B, zero=upper_only) if out is None else out
        ldm = self._mat_arg(out, B, "M")
        flags = self._flags(async_) | (RBD_UPPER_ONLY if upper_only else 0)
        _check(lib().rbd_crba(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(out), _i64(ldm),
                              ctypes.c_uint32(flags)))
        return out

    def cholesky_decompose(self, M, async_=False):
        """FP64 factor M = U D U^T; returns (U_packed, D): (B, nnz) and (B, nv) for "col",
        (nnz, B) and (nv, B) for "batch"."""
        import torch
        m = self.model
        B = M.shape[0] if self.layout == "col" else M.shape[1]
        ldm = self._mat_arg(M, B, "M")
        nnz = len(self.model.cholesky_pattern()[0])
        if self.layout == "col":
            U = torch.empty((B, max(nnz, 1)), dtype=torch.float64, device=M.device)
            D = torch.empty((B, m.nv), dtype=torch.float64, device=M.device)
            ldu, ldd = max(nnz, 1), m.nv
        else:
            U = torch.empty((max(nnz, 1), B), dtype=torch.float64, device=M.device)
            D = torch.empty((m.nv, B), dtype=torch.float64, device=M.device)
            ldu = ldd = max(B, 1)
        _check(lib().rbd_cholesky_decompose(m.ctx.handle, m.handle, _i64(B), _dtype_code(M), _ptr(M), _i64(ldm),
                                            _ptr(U), _i64(ldu), _ptr(D), _i64(ldd),
                                            ctypes.c_uint32(self._flags(async_))))
        return U, D

    def cholesky_solve(self, U, D, x, async_=False):
        """x <- M^-1 x in place with a factor from cholesky_decompose."""
        m = self.model
        B, ldx = self._vec(x, m.nv, "x")
        ldu = U.stride(0) if self.layout == "col" else U.stride(0)
        ldd = D.stride(0)
        _check(lib().rbd_cholesky_solve(m.ctx.handle, m.handle, _i64(B), _ptr(U), _i64(ldu), _ptr(D), _i64(ldd),
                                        _dtype_code(x), _ptr(x), _i64(ldx), ctypes.c_uint32(self._flags(async_))))
        return x

    def cholesky_apply(self, op, U, D, x, async_=False):
        """In place on x (float64) with a factor from cholesky_decompose: op "Uv" (U x), "Utv"
        (U^T x), "Uiv" (U^-1 x), "Utiv" (U^-T x) or "Mv" (M x = U D U^T x) -- cholesky::Uv,
        Utv, Uiv, Utiv, Mv (cholesky.hxx:64-465)."""
        codes = {"Uv": 0, "Utv": 1, "Uiv": 2, "Utiv": 3, "Mv": 4}
        if op not in codes:
            raise ValueError(f"op must be one of {sorted(codes)}")
        m = self.model
        B, ldx = self._vec(x, m.nv, "x")
        _check(lib().rbd_cholesky_apply(m.ctx.handle, m.handle, _i64(B), codes[op], _ptr(U), _i64(U.stride(0)),
                                        _ptr(D), _i64(D.stride(0)), _ptr(x), _i64(ldx),
                                        ctypes.c_uint32(self._flags(async_))))
        return x

    def rnea_derivatives(self, q, v, a, async_=False):
        """(dtau_dq, dtau_dv, M) per configuration (computeRNEADerivatives)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        _, lda = self._vec(a, m.nv, "a")
        outs = [self._new_mat(q, B) for _ in range(3)]
        ldo = self._mat_arg(outs[0], B, "out")
        _check(lib().rbd_rnea_derivatives(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v),
                                          _i64(ldv), _ptr(a), _i64(lda), *[_ptr(o) for o in outs], _i64(ldo),
                                          ctypes.c_uint32(self._flags(async_))))
        return tuple(outs)

    def aba_derivatives(self, q, v, tau, out=None, ddq_out=None, async_=False):
        """(ddq_dq, ddq_dv, Minv, ddq) per configuration (computeABADerivatives); `out`: three
        preallocated matrix outputs, `ddq_out`: a preallocated ddq (else allocated)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        _, ldt = self._vec(tau, m.nv, "tau")
        outs = [self._new_mat(q, B) for _ in range(3)] if out is None else list(out)
        ldo = self._mat_arg(outs[0], B, "out")
        ddq = self._new_vec(q, m.nv, B) if ddq_out is None else ddq_out
        _, ldd = self._vec(ddq, m.nv, "ddq")
        _check(lib().rbd_aba_derivatives(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v),
                                         _i64(ldv), _ptr(tau), _i64(ldt), *[_ptr(o) for o in outs], _i64(ldo),
                                         _ptr(ddq), _i64(ldd), ctypes.c_uint32(self._flags(async_))))
        return outs[0], outs[1], outs[2], ddq

    def forward_kinematics(self, q, v=None, a=None, async_=False):
        """(oMi, v_local, a_local): oMi rows 12 per joint (R row major, then t) for joints
        1.., velocities / accelerations 6 per joint in the joint frames (None unless given)."""
        import torch
        m = self.model
        nj = m.njoints
        B, ldq = self._vec(q, m.nq, "q")
        oM = self._new_vec(q, 12 * (nj - 1), B)
        _, ldo = self._vec(oM, 12 * (nj - 1), "oMi")
        vo = self._new_vec(q, 6 * (nj - 1), B) if v is not None else None
        ao = self._new_vec(q, 6 * (nj - 1), B) if (v is not None and a is not None) else None
        ldv = self._vec(v, m.nv, "v")[1] if v is not None else 0
        lda = self._vec(a, m.nv, "a")[1] if a is not None else 0
        ldva = self._vec(vo, 6 * (nj - 1), "v out")[1] if vo is not None else 0
        _check(lib().rbd_forward_kinematics(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                            _ptr(v) if v is not None else None, _i64(ldv),
                                            _ptr(a) if ao is not None else None, _i64(lda), _ptr(oM), _i64(ldo),
                                            _ptr(vo) if vo is not None else None,
                                            _ptr(ao) if ao is not None else None, _i64(ldva),
                                            ctypes.c_uint32(self._flags(async_))))
        return oM, vo, ao

    def joint_jacobians(self, q, async_=False):
        """WORLD-frame joint Jacobians, 6 x nv per configuration (element r + 6 c)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        J = self._new_vec(q, 6 * m.nv, B)
        _, ldj = self._vec(J, 6 * m.nv, "J")
        _check(lib().rbd_joint_jacobians(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(J),
                                         _i64(ldj), ctypes.c_uint32(self._flags(async_))))
        return J

    def joint_jacobians_time_variation(self, q, v, jacobians=True):
        """computeJointJacobiansTimeVariation: (J, dJ), WORLD-frame 6 x nv per configuration
        (element r + 6 c); J is None when jacobians=False.  Synchronous."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        dJ = self._new_vec(q, 6 * m.nv, B)
        J = self._new_vec(q, 6 * m.nv, B) if jacobians else None
        _, ldd = self._vec(dJ, 6 * m.nv, "dJ")
        ldj = self._vec(J, 6 * m.nv, "J")[1] if jacobians else 0
        _check(lib().rbd_joint_jacobians_time_variation(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q),
                                                        _i64(ldq), _ptr(v), _i64(ldv), _ptr(J) if jacobians else None,
                                                        _i64(ldj), _ptr(dJ), _i64(ldd),
                                                        ctypes.c_uint32(self._flags(False))))
        return J, dJ

    REFERENCE_FRAMES = {"world": 0, "local": 1, "local_world_aligned": 2}

    def frames(self, q, frame_joint, frame_placement, reference_frame="world", placements=True, jacobians=True):
        """Frame placements oMf (12 rows per frame: R row major, p) and 6 x nv frame Jacobians
        (rows f 6 nv + r + 6 c) for frames given by parent joint and placement (12 numbers
        each) -- updateFramePlacements and getFrameJacobian per configuration.  Synchronous."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        fj = np.ascontiguousarray(frame_joint, dtype=np.int32).ravel()
        nf = len(fj)
        fp = np.ascontiguousarray(frame_placement, dtype=np.float64).reshape(nf, 12)
        oMf = self._new_vec(q, 12 * nf, B) if placements else None
        J = self._new_vec(q, 6 * m.nv * nf, B) if jacobians else None
        ldo = self._vec(oMf, 12 * nf, "oMf")[1] if placements else 0
        ldj = self._vec(J, 6 * m.nv * nf, "J")[1] if jacobians else 0
        _check(lib().rbd_frames(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), int(nf),
                                fj.ctypes.data_as(ctypes.c_void_p), fp.ctypes.data_as(ctypes.c_void_p),
                                int(self.REFERENCE_FRAMES[reference_frame]), _ptr(oMf) if placements else None,
                                _i64(ldo), _ptr(J) if jacobians else None, _i64(ldj),
                                ctypes.c_uint32(self._flags(False))))
        return oMf, J

    def frame_jacobians_time_variation(self, q, v, frame_joint, frame_placement, reference_frame="local"):
        """getFrameJacobianTimeVariation after computeJointJacobiansTimeVariation(q, v): each
        frame's 6 x nv dJ in the reference frame (rows f 6 nv + r + 6 c).  Synchronous."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        fj = np.ascontiguousarray(frame_joint, dtype=np.int32).ravel()
        nf = len(fj)
        fp = np.ascontiguousarray(frame_placement, dtype=np.float64).reshape(nf, 12)
        dJ = self._new_vec(q, 6 * m.nv * nf, B)
        _, ldd = self._vec(dJ, 6 * m.nv * nf, "dJ")
        _check(lib().rbd_frame_jacobians_time_variation(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q),
                                                        _i64(ldq), _ptr(v), _i64(ldv), int(nf),
                                                        fj.ctypes.data_as(ctypes.c_void_p),
                                                        fp.ctypes.data_as(ctypes.c_void_p),
                                                        int(self.REFERENCE_FRAMES[reference_frame]), _ptr(dJ),
                                                        _i64(ldd), ctypes.c_uint32(self._flags(False))))
        return dJ

    def frame_motions(self, q, v, frame_joint, frame_placement, a=None, reference_frame="local"):
        """getFrameVelocity, getFrameAcceleration and getFrameClassicalAcceleration after
        forwardKinematics(q, v[, a]) for frames given by parent joint and placement (12 numbers
        each): (vf, af, acf), 6 rows per frame ([linear; angular], row 6 f + k); af and acf
        are None without a.  Synchronous."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        lda = self._vec(a, m.nv, "a")[1] if a is not None else 0
        fj = np.ascontiguousarray(frame_joint, dtype=np.int32).ravel()
        nf = len(fj)
        fp = np.ascontiguousarray(frame_placement, dtype=np.float64).reshape(nf, 12)
        vf = self._new_vec(q, 6 * nf, B)
        af = self._new_vec(q, 6 * nf, B) if a is not None else None
        acf = self._new_vec(q, 6 * nf, B) if a is not None else None
        ldm = self._vec(vf, 6 * nf, "vf")[1]
        _check(lib().rbd_frame_motions(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v),
                                       _i64(ldv), _ptr(a) if a is not None else None, _i64(lda), int(nf),
                                       fj.ctypes.data_as(ctypes.c_void_p), fp.ctypes.data_as(ctypes.c_void_p),
                                       int(self.REFERENCE_FRAMES[reference_frame]), _ptr(vf),
                                       _ptr(af) if af is not None else None, _ptr(acf) if acf is not None else None,
                                       _i64(ldm), ctypes.c_uint32(self._flags(False))))
        return vf, af, acf

    def compute_minverse(self, q, out=None, async_=False):
        """M(q)^-1 per configuration (computeMinverse), FP32 or FP64 like q."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        out = self._new_mat(q, B) if out is None else out
        ldm = self._mat_arg(out, B, "Minv")
        _check(lib().rbd_compute_minverse(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                          _ptr(out), _i64(ldm), ctypes.c_uint32(self._flags(async_))))
        return out

    def constrained_forward_dynamics(self, q, v, tau, J, gamma, inv_damping=0.0, out=None, lambda_out=None,
                                     async_=False):
        """forwardDynamics(model, data, q, v, tau, J, gamma, inv_damping) per configuration
        (contact-dynamics.hxx:90-167): ddq under J ddq + gamma = 0 and the multipliers lambda.
        J holds nc * nv rows per configuration, element (i, k) of that configuration's
        nc x nv matrix at row i + k * nc (Eigen's column-major memory); gamma and lambda have
        nc rows.  Returns (ddq, lambda)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        _, ldt = self._vec(tau, m.nv, "tau")
        rows = J.shape[1] if self.layout == "col" else J.shape[0]
        nc = rows // m.nv
        if nc * m.nv != rows:
            raise ValueError(f"J: {rows} rows per configuration is not a multiple of nv = {m.nv}")
        _, ldJ = self._vec(J, nc * m.nv, "J")
        _, ldg = self._vec(gamma, nc, "gamma")
        out = self._new_vec(q, m.nv, B) if out is None else out
        lambda_out = self._new_vec(q, nc, B) if lambda_out is None else lambda_out
        _, lda = self._vec(out, m.nv, "ddq")
        _, ldl = self._vec(lambda_out, nc, "lambda")
        _check(lib().rbd_constrained_forward_dynamics(
            m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v), _i64(ldv), _ptr(tau),
            _i64(ldt), int(nc), _ptr(J), _i64(ldJ), _ptr(gamma), _i64(ldg), ctypes.c_double(inv_damping), _ptr(out),
            _i64(lda), _ptr(lambda_out), _i64(ldl), ctypes.c_uint32(self._flags(async_))))
        return out, lambda_out

    def impulse_dynamics(self, q, v_before, J, r_coeff=0.0, inv_damping=0.0, out=None, impulse_out=None,
                         async_=False):
        """impulseDynamics(model, data, q, v_before, J, r_coeff, inv_damping) per configuration
        (contact-dynamics.hxx:264-331): the velocity after impact and the impulses.  J as in
        constrained_forward_dynamics.  Returns (dq_after, impulse)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v_before, m.nv, "v_before")
        rows = J.shape[1] if self.layout == "col" else J.shape[0]
        nc = rows // m.nv
        if nc * m.nv != rows:
            raise ValueError(f"J: {rows} rows per configuration is not a multiple of nv = {m.nv}")
        _, ldJ = self._vec(J, nc * m.nv, "J")
        out = self._new_vec(q, m.nv, B) if out is None else out
        impulse_out = self._new_vec(q, nc, B) if impulse_out is None else impulse_out
        _, lda = self._vec(out, m.nv, "dq_after")
        _, ldi = self._vec(impulse_out, nc, "impulse")
        _check(lib().rbd_impulse_dynamics(
            m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v_before), _i64(ldv), int(nc),
            _ptr(J), _i64(ldJ), ctypes.c_double(r_coeff), ctypes.c_double(inv_damping), _ptr(out), _i64(lda),
            _ptr(impulse_out), _i64(ldi), ctypes.c_uint32(self._flags(async_))))
        return out, impulse_out

    def constraint_dynamics(self, q, v, tau, constraints, settings=None, out=None, lambda_out=None,
                            iterations_out=None, async_=False):
        """constraintDynamics(model, data, q, v, tau, contact_models, contact_datas, settings)
        per configuration (constrained-dynamics.hxx:193-500).  constraints: a list of
        RigidConstraintModel; settings: ProximalSettings (None: the reference's defaults).
        Returns (ddq, lambda) -- lambda holds the constraints' rows in order -- and fills
        iterations_out (int32, one per configuration) when given."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        _, ldv = self._vec(v, m.nv, "v")
        _, ldt = self._vec(tau, m.nv, "tau")
        table = _constraint_table(constraints)
        nrows = sum(c.size() for c in constraints)
        out = self._new_vec(q, m.nv, B) if out is None else out
        lambda_out = self._new_vec(q, nrows, B) if lambda_out is None else lambda_out
        _, lda = self._vec(out, m.nv, "ddq")
        _, ldl = self._vec(lambda_out, nrows, "lambda")
        ps = None if settings is None else ctypes.byref(settings._c())
        it = None if iterations_out is None else _ptr(iterations_out)
        _check(lib().rbd_constraint_dynamics(
            m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq), _ptr(v), _i64(ldv), _ptr(tau),
            _i64(ldt), len(constraints), table, ps, _ptr(out), _i64(lda), _ptr(lambda_out), _i64(ldl), it,
            ctypes.c_uint32(self._flags(async_))))
        return out, lambda_out

    def delassus(self, q, constraints, mu=0.0, out=None, async_=False):
        """computeDelassusMatrix(model, data, q, contact_models, contact_datas, G, mu) per
        configuration (delassus.hxx:118-300): J (M + mu J^T J)^-1 J^T for the constraints'
        Jacobian (RigidConstraintModel list, LOCAL or LOCAL_WORLD_ALIGNED rows).  Returns G
        with nrows^2 rows per configuration (element (i, j) at row i + j nrows)."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        table = _constraint_table(constraints)
        nrows = sum(c.size() for c in constraints)
        out = self._new_vec(q, nrows * nrows, B) if out is None else out
        _, ldG = self._vec(out, nrows * nrows, "G")
        _check(lib().rbd_delassus(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                  len(constraints), table, ctypes.c_double(mu), _ptr(out), _i64(ldG),
                                  ctypes.c_uint32(self._flags(async_))))
        return out

    def damped_delassus_inverse(self, q, constraints, mu, scaled=False, out=None, async_=False):
        """computeDampedDelassusMatrixInverse(model, data, q, contact_models, contact_datas, Ginv,
        mu, scaled) per configuration (delassus.hxx:800-836): (J M^-1 J^T + mu I)^-1, times mu
        when scaled; layout as delassus()."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        table = _constraint_table(constraints)
        nrows = sum(c.size() for c in constraints)
        out = self._new_vec(q, nrows * nrows, B) if out is None else out
        _, ldG = self._vec(out, nrows * nrows, "Ginv")
        _check(lib().rbd_damped_delassus_inverse(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                                 len(constraints), table, ctypes.c_double(mu), int(bool(scaled)),
                                                 _ptr(out), _i64(ldG), ctypes.c_uint32(self._flags(async_))))
        return out

    def contact_cholesky_solve(self, q, constraints, mu, rhs, out=None, async_=False):
        """ContactCholeskyDecomposition compute(model, data, contact_models, contact_datas, mu)
        then solve(rhs) per configuration (contact-cholesky.hpp:292-400): x = K^-1 rhs with
        K = [[-mu I, J], [J^T, M]]; rhs and x have nrows + nv rows per configuration, the
        constraint rows first."""
        m = self.model
        B, ldq = self._vec(q, m.nq, "q")
        table = _constraint_table(constraints)
        nrows = sum(c.size() for c in constraints)
        _, ldr = self._vec(rhs, nrows + m.nv, "rhs")
        out = self._new_vec(q, nrows + m.nv, B) if out is None else out
        _, ldx = self._vec(out, nrows + m.nv, "x")
        _check(lib().rbd_contact_cholesky_solve(m.ctx.handle, m.handle, _i64(B), _dtype_code(q), _ptr(q), _i64(ldq),
                                                len(constraints), table, ctypes.c_double(mu), _ptr(rhs), _i64(ldr),
                                                _ptr(out), _i64(ldx), ctypes.c_uint32(self._flags(async_))))
        return out

    @staticmethod
    def mats_to_numpy(t, layout, nv):
        """Device matrix output -> (B, nv, nv) numpy array with [b] the configuration's matrix."""
        a = t.double().cpu().numpy()
        if layout == "col":
            return a.transpose(0, 2, 1)  # [b, c, r] -> [b, r, c]
        B = a.shape[1]
        return a.reshape(nv, nv, B).transpose(2, 1, 0)  # [c, r, b] -> [b, r, c]


# ------------------------------------------------------------------------------------
# Rigid constraints (RigidConstraintModel, ProximalSettings) for constraint_dynamics
# ------------------------------------------------------------------------------------
CONTACT_3D, CONTACT_6D = 0, 1
WORLD, LOCAL, LOCAL_WORLD_ALIGNED = 0, 1, 2
_IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0)


class _RigidConstraintC(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("joint1_id", ctypes.c_int32), ("joint2_id", ctypes.c_int32),
                ("reference_frame", ctypes.c_int32), ("joint1_placement", ctypes.c_double * 12),
                ("joint2_placement", ctypes.c_double * 12), ("Kp", ctypes.c_double * 6), ("Kd", ctypes.c_double * 6),
                ("desired_acceleration", ctypes.c_double * 6)]


class _ProximalSettingsC(ctypes.Structure):
    _fields_ = [("absolute_accuracy", ctypes.c_double), ("relative_accuracy", ctypes.c_double),
                ("mu", ctypes.c_double), ("max_iter", ctypes.c_int32)]


class RigidConstraintModel:
    """A rigid constraint (algorithm/contact-info.hpp:95-420): CONTACT_3D holds the point of
    joint1's frame on the point of joint2's frame, CONTACT_6D the two frames together; joint
    0 is the universe.  Placements are 12 numbers (R row major, then p) or 4 x 4 matrices;
    Kp, Kd: Baumgarte corrector gains (3 or 6 numbers); desired_acceleration: the contact
    data's contact_acceleration_desired."""

    def __init__(self, type, joint1_id, joint1_placement=_IDENTITY, joint2_id=0, joint2_placement=_IDENTITY,
                 reference_frame=LOCAL, Kp=(), Kd=(), desired_acceleration=()):
        self.type, self.joint1_id, self.joint2_id = int(type), int(joint1_id), int(joint2_id)
        self.reference_frame = int(reference_frame)
        self.joint1_placement = _placement12(joint1_placement)
        self.joint2_placement = _placement12(joint2_placement)
        self.Kp, self.Kd = _pad6(Kp), _pad6(Kd)
        self.desired_acceleration = _pad6(desired_acceleration)

    def size(self):
        return 6 if self.type == CONTACT_6D else 3


class ProximalSettings:
    """ProximalSettingsTpl (algorithm/proximal.hpp:20-60): the reference's defaults are
    accuracy 1e-12, mu 0 and one iteration."""

    def __init__(self, accuracy=1e-12, mu=0.0, max_iter=1, relative_accuracy=None):
        self.absolute_accuracy = float(accuracy)
        self.relative_accuracy = float(accuracy if relative_accuracy is None else relative_accuracy)
        self.mu, self.max_iter = float(mu), int(max_iter)

    def _c(self):
        return _ProximalSettingsC(self.absolute_accuracy, self.relative_accuracy, self.mu, self.max_iter)


def _placement12(x):
    a = np.asarray(x, dtype=np.float64)
    if a.shape == (4, 4):
        a = np.concatenate([a[:3, :3].ravel(), a[:3, 3]])
    a = a.ravel()
    if a.size != 12:
        raise ValueError("a placement is 12 numbers (R row major, p) or a 4 x 4 matrix")
    return a


def _pad6(x):
    a = np.zeros(6)
    v = np.asarray(x, dtype=np.float64).ravel()
    if v.size > 6:
        raise ValueError("at most 6 gains")
    a[:v.size] = v
    return a


def _constraint_table(constraints):
    arr = (_RigidConstraintC * len(constraints))()
    for k, c in enumerate(constraints):
        a = arr[k]
        a.type, a.joint1_id, a.joint2_id, a.reference_frame = c.type, c.joint1_id, c.joint2_id, c.reference_frame
        a.joint1_placement[:] = list(c.joint1_placement)
        a.joint2_placement[:] = list(c.joint2_placement)
        a.Kp[:] = list(c.Kp)
        a.Kd[:] = list(c.Kd)
        a.desired_acceleration[:] = list(c.desired_acceleration)
    return arr


# ------------------------------------------------------------------------------------
# Mixed-robot batches (one launch over instances of several models)
# ------------------------------------------------------------------------------------
# ------------------------------------------------------------------------------------
# Frictional contact solvers (PGS, ADMM) over batches of problems
# ------------------------------------------------------------------------------------
RBD_NCP_STATS = 8


class CoulombFrictionCone:
    """CoulombFrictionConeTpl (constraints/coulomb-friction-cone.hpp:20-199): {|f_t| <= mu f_n}
    on a contact's (tangent x, tangent y, normal) components."""

    def __init__(self, mu: float):
        if mu < 0:
            raise ValueError("mu must be positive")
        self.mu = float(mu)


class _PgsSettingsC(ctypes.Structure):
    _fields_ = [("max_it", ctypes.c_int32), ("absolute_precision", ctypes.c_double),
                ("relative_precision", ctypes.c_double), ("over_relax", ctypes.c_double)]


class _AdmmSettingsC(ctypes.Structure):
    _fields_ = [("mu_prox", ctypes.c_double), ("tau", ctypes.c_double), ("rho_power", ctypes.c_double),
                ("rho_power_factor", ctypes.c_double), ("ratio_primal_dual", ctypes.c_double),
                ("max_it_largest_eigen_value", ctypes.c_int32), ("max_it", ctypes.c_int32),
                ("absolute_precision", ctypes.c_double), ("relative_precision", ctypes.c_double)]


class _ContactSolverBase:
    """ContactSolverBaseTpl (contact-solver-base.hpp:20-120) over a batch: the problems are the
    rows of (B, ...) float64 device tensors -- G (B, n, n), g (B, n) -- and every per-problem
    result (iteration count, residuals) is an array of B values."""

    def __init__(self, ctx: Context, problem_dim: int):
        if problem_dim % 3 or not 3 <= problem_dim <= 12:
            raise ValueError("the problem size must be 3 ncones with 1 <= ncones <= 4")
        self.ctx = ctx
        self.problem_dim = int(problem_dim)
        self.max_it = 1000
        self.absolute_precision = 1e-6
        self.relative_precision = 1e-6
        self.stats = None

    def getProblemSize(self):
        return self.problem_dim

    def setMaxIterations(self, max_it):
        if max_it <= 0:
            raise ValueError("max_it should be greater than 0.")
        self.max_it = int(max_it)

    def getMaxIterations(self):
        return self.max_it

    def setAbsolutePrecision(self, p):
        if p < 0:
            raise ValueError("absolute_precision should be positive.")
        self.absolute_precision = float(p)

    def getAbsolutePrecision(self):
        return self.absolute_precision

    def setRelativePrecision(self, p):
        if p < 0:
            raise ValueError("relative_precision should be positive.")
        self.relative_precision = float(p)

    def getRelativePrecision(self):
        return self.relative_precision

    def _stat(self, k):
        if self.stats is None:
            raise RuntimeError("solve() has not run")
        return self.stats[:, k]

    def getIterationCount(self):
        return self._stat(0).to(_torch_int32())

    def getAbsoluteConvergenceResidual(self):
        return self._stat(1)

    def getRelativeConvergenceResidual(self):
        return self._stat(2)

    def _problem(self, G, g, cones):
        import torch
        n = self.problem_dim
        nc = n // 3
        if G.dim() == 3:
            G = G.transpose(1, 2)  # element (i, j) at i + j n
        G = G.reshape(G.shape[0], n * n).to(torch.float64).contiguous()
        B = G.shape[0]
        g = g.to(torch.float64).contiguous()
        if tuple(g.shape) != (B, n):
            raise ValueError(f"g: expected ({B}, {n}), got {tuple(g.shape)}")
        if isinstance(cones, (list, tuple)):
            if len(cones) != nc:
                raise ValueError(f"{nc} cones expected, got {len(cones)}")
            mu = torch.tensor([[c.mu for c in cones]], dtype=torch.float64, device=G.device).expand(B, nc)
        else:
            mu = cones
        mu = mu.to(torch.float64).contiguous()
        if tuple(mu.shape) != (B, nc):
            raise ValueError(f"friction coefficients: expected ({B}, {nc}), got {tuple(mu.shape)}")
        return B, nc, G, g, mu


def _torch_int32():
    import torch
    return torch.int32


class PGSContactSolver(_ContactSolverBase):
    """PGSContactSolverTpl (pgs-solver.hpp:16-59) over a batch of problems."""

    def __init__(self, ctx: Context, problem_dim: int):
        super().__init__(ctx, problem_dim)

    def solve(self, G, g, cones, x, over_relax=1.0, async_=False):
        """PGSContactSolverTpl::solve(G, g, cones, x, over_relax) (pgs-solver.hxx:19-141) on each
        problem: x (B, n) float64 holds the initial guess and receives the solution.  cones:
        a list of CoulombFrictionCone shared by every problem, or a (B, ncones) tensor of
        friction coefficients.  Returns a bool tensor: converged per problem."""
        import torch
        B, nc, G, g, mu = self._problem(G, g, cones)
        n = self.problem_dim
        if x.dtype != torch.float64 or not x.is_contiguous() or tuple(x.shape) != (B, n):
            raise ValueError(f"x: a contiguous float64 ({B}, {n}) tensor")
        self.stats = torch.empty((B, RBD_NCP_STATS), dtype=torch.float64, device=G.device)
        st = _PgsSettingsC(self.max_it, self.absolute_precision, self.relative_precision, over_relax)
        _check(lib().rbd_pgs_solve(self.ctx.handle, _i64(B), nc, _ptr(G), _i64(n * n), _ptr(g), _i64(n), _ptr(mu),
                                   _i64(nc), _ptr(x), _i64(n), ctypes.byref(st), _ptr(self.stats),
                                   _i64(RBD_NCP_STATS), ctypes.c_uint32(RBD_ASYNC if async_ else 0)))
        return self.stats[:, 3] != 0


class ADMMContactSolver(_ContactSolverBase):
    """ADMMContactSolverTpl (admm-solver.hpp:120-127 parameters) over a batch of problems;
    each problem is solved by a fresh solver on a fresh dense Delassus operator."""

    def __init__(self, ctx: Context, problem_dim: int, mu_prox=1e-6, tau=0.5, rho_power=0.2, rho_power_factor=0.05,
                 ratio_primal_dual=10.0, max_it_largest_eigen_value_solver=20):
        super().__init__(ctx, problem_dim)
        self.mu_prox, self.tau = float(mu_prox), float(tau)
        self.rho_power, self.rho_power_factor = float(rho_power), float(rho_power_factor)
        self.ratio_primal_dual = float(ratio_primal_dual)
        self.max_it_largest_eigen_value = int(max_it_largest_eigen_value_solver)
        self.y = self.z = self.s = None

    def solve(self, G, g, cones, R=None, primal_guess=None, dual_guess=None, async_=False):
        """ADMMContactSolverTpl::solve(delassus, g, cones, R, primal_guess, dual_guess)
        (admm-solver.hxx:22-335) on each problem.  R: (B, n) compliance (zero if None).
        Returns a bool tensor: converged per problem."""
        import torch
        B, nc, G, g, mu = self._problem(G, g, cones)
        n = self.problem_dim
        dev = G.device
        R = torch.zeros((B, n), dtype=torch.float64, device=dev) if R is None else R.to(torch.float64).contiguous()
        pg = None if primal_guess is None else primal_guess.to(torch.float64).contiguous()
        dg = None if dual_guess is None else dual_guess.to(torch.float64).contiguous()
        self.y, self.z, self.s = (torch.empty((B, n), dtype=torch.float64, device=dev) for _ in range(3))
        self.stats = torch.empty((B, RBD_NCP_STATS), dtype=torch.float64, device=dev)
        st = _AdmmSettingsC(self.mu_prox, self.tau, self.rho_power, self.rho_power_factor, self.ratio_primal_dual,
                            self.max_it_largest_eigen_value, self.max_it, self.absolute_precision,
                            self.relative_precision)
        _check(lib().rbd_admm_solve(self.ctx.handle, _i64(B), nc, _ptr(G), _i64(n * n), _ptr(g), _i64(n), _ptr(mu),
                                    _i64(nc), _ptr(R), _i64(n), None if pg is None else _ptr(pg),
                                    None if dg is None else _ptr(dg), _i64(n), ctypes.byref(st), _ptr(self.y),
                                    _ptr(self.z), _ptr(self.s), _i64(n), _ptr(self.stats), _i64(RBD_NCP_STATS),
                                    ctypes.c_uint32(RBD_ASYNC if async_ else 0)))
        return self.stats[:, 3] != 0

    def setTau(self, tau):
        self.tau = float(tau)

    def getTau(self):
        return self.tau

    def setRhoPowerFactor(self, f):
        self.rho_power_factor = float(f)

    def getRhoPowerFactor(self):
        return self.rho_power_factor

    def setRatioPrimalDual(self, r):
        if r <= 0:
            raise ValueError("The ratio primal/dual should be positive strictly")
        self.ratio_primal_dual = float(r)

    def getRatioPrimalDual(self):
        return self.ratio_primal_dual

    def setProximalValue(self, mu):
        if mu < 0:
            raise ValueError("mu_prox should be positive.")
        self.mu_prox = float(mu)

    def getProximalValue(self):
        return self.mu_prox

    def getPrimalSolution(self):
        return self.y

    def getDualSolution(self):
        return self.z

    def getComplementarityShift(self):
        return self.s

    def getRho(self):
        return self._stat(4)

    def getRhoPower(self):
        return self._stat(5)

    def getLargestEigenValue(self):
        return self._stat(6)

    def getCholeskyUpdateCount(self):
        return self._stat(7).to(_torch_int32())


class _BatchGroup(ctypes.Structure):
    _fields_ = [("model", ctypes.c_void_p), ("count", ctypes.c_int64), ("q", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("a", ctypes.c_void_p), ("tau", ctypes.c_void_p),
                ("tau_out", ctypes.c_void_p), ("ddq_out", ctypes.c_void_p)]


class MixedBatch:
    """RNEA + ABA over groups of instances of different models in one kernel launch.

    ``groups`` is a list of (Model, q, v, a, tau) with FP32 device tensors in the
    batch-fastest layout ((nq, count) and (nv, count), contiguous).  Outputs (tau_out,
    ddq_out) are allocated once; the first call autotunes and captures a HIP graph that
    later calls replay (inputs may be rewritten in place between calls)."""

    def __init__(self, ctx: Context, groups, graph=True, autotune=True):
        import torch
        self.ctx = ctx
        self.groups = []
        arr = (_BatchGroup * len(groups))()
        for k, (m, q, v, a, tau) in enumerate(groups):
            n = q.shape[1]
            for t, rows, name in ((q, m.nq, "q"), (v, m.nv, "v"), (a, m.nv, "a"), (tau, m.nv, "tau")):
                if t.dtype != torch.float32 or tuple(t.shape) != (rows, n) or not t.is_contiguous():
                    raise ValueError(f"group {k}: {name} must be a contiguous float32 ({rows}, {n}) tensor")
            tau_out = torch.empty((m.nv, n), dtype=torch.float32, device=q.device)
            ddq_out = torch.empty((m.nv, n), dtype=torch.float32, device=q.device)
            self.groups.append((m, q, v, a, tau, tau_out, ddq_out))
            arr[k] = _BatchGroup(m.handle.value, n, q.data_ptr(), v.data_ptr(), a.data_ptr(), tau.data_ptr(),
                                 tau_out.data_ptr(), ddq_out.data_ptr())
        self._arr = arr
        self._flags = (0 if graph else RBD_NO_GRAPH) | (0 if autotune else RBD_NO_AUTOTUNE)

    def run(self, async_=False):
        _check(lib().rbd_multi_rnea_aba(self.ctx.handle, ctypes.c_int32(len(self.groups)), self._arr,
                                        ctypes.c_uint32(self._flags | (RBD_ASYNC if async_ else 0))))
        return [(g[5], g[6]) for g in self.groups]

    def tuning(self) -> dict:
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().rbd_multi_tuning(self.ctx.handle, buf, ctypes.c_int32(len(buf))))
        return json.loads(buf.value.decode()) if buf.value else {}


# ------------------------------------------------------------------------------------
# Drop-in mirror of the reference's Python batch API
# ------------------------------------------------------------------------------------
class ModelPool:
    """Counterpart of pinocchio.ModelPool (multibody/pool/model.hpp:19-165): holds the
    model once on the GPU instead of one copy per OpenMP thread; `size` is kept for the
    reference's argument checks.  `model` is a ModelDesc or a pinocchio Model (the
    reference's Python ``pinocchio.Model``, or any object with its attributes; see
    ``models.from_pinocchio``)."""

    def __init__(self, model, size: int = 1, device: int = 0):
        if size < 1:
            raise ValueError("The pool size must be positive")
        desc = model if isinstance(model, ModelDesc) else models.from_pinocchio(model)
        self.ctx = Context(device)
        self.model = Model(self.ctx, desc)
        self._size = size

    def size(self) -> int:
        return self._size


def _host_call(fn_name, num_threads, pool, q, v, x, dtype, out=None):
    """Shared argument checks of algorithm/parallel/{rnea,aba}.hpp:53-69 and the call.
    With `out` (an nv x B float64 array) the result is written into it, as the
    reference's output-argument overloads do, and None is returned."""
    m = pool.model
    if pool.size() == 0:
        raise ValueError("The pool should have at least one element")
    if num_threads > pool.size():
        raise ValueError("The number of threads should be lower than the pool size")
    q = np.asfortranarray(np.asarray(q, dtype=dtype))
    v = np.asfortranarray(np.asarray(v, dtype=dtype))
    x = np.asfortranarray(np.asarray(x, dtype=dtype))
    if q.ndim != 2 or q.shape[0] != m.nq:
        raise ValueError("The configuration matrix has wrong number of rows")
    if v.ndim != 2 or v.shape[0] != m.nv:
        raise ValueError("The velocity matrix has wrong number of rows")
    if x.ndim != 2 or x.shape[0] != m.nv:
        raise ValueError("The third input matrix has wrong number of rows")
    B = q.shape[1]
    if v.shape[1] != B or x.shape[1] != B:
        raise ValueError("The input matrices do not have the same number of columns")
    if out is not None:
        if not isinstance(out, np.ndarray) or out.ndim != 2 or out.shape[0] != m.nv:
            raise ValueError("The output matrix has wrong number of rows")
        if out.shape[1] != B:
            raise ValueError("The output matrix does not have the same number of columns")
        if not out.flags.writeable:
            raise ValueError("The output matrix is not writeable")
    res = np.zeros((m.nv, B), dtype=dtype, order="F")
    code = RBD_F64 if dtype == np.float64 else RBD_F32
    if B > 0:
        f = getattr(lib(), fn_name)
        _check(f(pool.ctx.handle, m.handle, _i64(B), code, q.ctypes.data_as(ctypes.c_void_p), _i64(m.nq),
                 v.ctypes.data_as(ctypes.c_void_p), _i64(m.nv), x.ctypes.data_as(ctypes.c_void_p), _i64(m.nv),
                 res.ctypes.data_as(ctypes.c_void_p), _i64(m.nv), ctypes.c_uint32(RBD_HOST_PTRS)))
    if out is not None:
        out[...] = res
        return None
    return res.astype(np.float64)


def rneaInParallel(num_threads, pool, q, v, a, tau=None, dtype=np.float64):
    """tau = rnea(q[:, i], v[:, i], a[:, i]) for each column (algorithm/parallel/rnea.hpp:39-83).

    Both overloads of the reference's binding (bindings/python/algorithm/parallel/
    rnea.cpp:16-24, 54-55): without `tau` the result is returned; with `tau` (an nv x B
    array) it is written there and None is returned.  FP64 by default, like the
    reference; dtype=np.float32 computes in FP32."""
    return _host_call("rbd_rnea", num_threads, pool, q, v, a, dtype, tau)


def abaInParallel(num_threads, pool, q, v, tau, a=None, dtype=np.float64):
    """ddq = aba(q[:, i], v[:, i], tau[:, i]) for each column (algorithm/parallel/aba.hpp:40-84).

    Both overloads of the reference's binding (bindings/python/algorithm/parallel/
    aba.cpp:16-24, 54-55): with `a` (an nv x B array) the result is written there and None
    is returned.  FP64 by default, like the reference; dtype=np.float32 computes in FP32."""
    return _host_call("rbd_aba", num_threads, pool, q, v, tau, dtype, a)
