"""CPU checks of the boundary: the C-ABI library loads and exports every entry point
include/rbd.h declares; the host-side model builders and sampling are reproducible.
No compute call is made here (no GPU)."""
import ctypes
import os
import re

import numpy as np

import pinocchio_amd as pa
from pinocchio_amd import models as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rbd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = pa.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert L.rbd_version().startswith(b"pinocchio_amd")


def test_error_without_context_is_reported():
    L = pa.lib()
    rc = L.rbd_ctx_create(0, None)  # null output pointer: argument error, no device touched
    assert rc == 1 and b"null" in L.rbd_last_error()


def test_models_reproducible_and_shaped():
    a, b = M.humanoid36(), M.humanoid36()
    np.testing.assert_array_equal(a.placement, b.placement)
    assert (a.nq, a.nv, a.njoints) == (37, 36, 32)
    for name, (nq, nv) in {"manipulator6": (6, 6), "arm7": (7, 7), "quadruped18": (19, 18),
                           "random60": (61, 60)}.items():
        m = M.SAMPLE_MODELS[name]()
        assert (m.nq, m.nv) == (nq, nv)
    m = M.random60()
    assert np.all(m.parents[1:] < np.arange(1, m.njoints))
    # parameter ranges of the benchmark models
    assert np.all((m.inertia[1:, 0] >= 0.5) & (m.inertia[1:, 0] <= 5.0))
    assert np.all(np.abs(m.inertia[1:, 1:4]) <= 0.1)
    assert np.all(np.abs(m.placement[2:, 9:]) <= 0.3)
    for j in range(1, m.njoints):
        I = np.array([[m.inertia[j, 4], m.inertia[j, 5], m.inertia[j, 7]],
                      [m.inertia[j, 5], m.inertia[j, 6], m.inertia[j, 8]],
                      [m.inertia[j, 7], m.inertia[j, 8], m.inertia[j, 9]]])
        ev = np.linalg.eigvalsh(I)
        assert ev.min() >= 0.01 - 1e-12 and ev.max() <= 0.1 + 1e-12
        R = m.placement[j, :9].reshape(3, 3)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)


def test_inputs_distribution():
    m = M.humanoid36()
    q, v, a, tau = M.sample_inputs(m, 20000, 99, tau_sigma=20.0)
    np.testing.assert_allclose(np.linalg.norm(q[3:7], axis=0), 1.0, atol=1e-12)
    assert np.abs(q[7:]).max() <= np.pi
    assert abs(v.std() - 1.0) < 0.02 and abs(tau.std() - 20.0) < 0.4
    q2, *_ = M.sample_inputs(m, 20000, 99)
    np.testing.assert_array_equal(q, q2)


def test_cpp_dropin_header_compiles_without_hip_or_eigen():
    """The C++ drop-in needs only the C ABI: a host compiler with no HIP or Eigen headers
    on the include path compiles a caller (tests/cpp/dropin_check.cpp)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    assert gxx, "g++ missing"
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
