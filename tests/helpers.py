"""Shared helpers for the parity tests: seeded inputs, FP32 rounding, error metric."""
import numpy as np

from pinocchio_amd import models as M


def f32_round(*xs):
    """Round inputs to float32 and back, so the FP64 oracle sees exactly what the FP32
    kernels see (input rounding is not counted as error)."""
    return tuple(np.asarray(x, dtype=np.float32).astype(np.float64) for x in xs)


def rel_err_cols(x, ref):
    """Per-configuration ||x - ref||_inf / max(||ref||_inf, 1e-30); arrays are (n, B)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return np.abs(x - ref).max(axis=0) / np.maximum(np.abs(ref).max(axis=0), 1e-30)


def rel_err_mats(x, ref):
    """Per-configuration matrix error for (B, n, n) arrays."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    d = np.abs(x - ref).reshape(len(ref), -1).max(axis=1)
    return d / np.maximum(np.abs(ref).reshape(len(ref), -1).max(axis=1), 1e-30)


def inputs(model, B, k=0, tau_sigma=1.0):
    return M.sample_inputs(model, B, M.INPUT_SEED_BASE + 100 + k, tau_sigma=tau_sigma)


def two_link_arm(m1=1.7, m2=0.9, l1c=0.35, L1=0.6, l2c=0.25, off=0.4):
    """Two revolute-Z joints, point masses on the links, a rotated and offset first
    placement (Rx(90 deg), t = (0.1, 0.2, 0.3)) and a second placement offset along the
    first link and rotated about z by `off`: a planar arm in a vertical plane."""
    Rx90 = np.array([[1.0, 0, 0], [0, 0, -1], [0, 1, 0]])
    c, s = np.cos(off), np.sin(off)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    placement = np.zeros((3, 12))
    placement[0, :9] = np.eye(3).ravel()
    placement[1] = np.concatenate([Rx90.ravel(), [0.1, 0.2, 0.3]])
    placement[2] = np.concatenate([Rz.ravel(), [L1, 0, 0]])
    inertia = np.zeros((3, 10))
    inertia[1, :4] = [m1, l1c, 0, 0]
    inertia[2, :4] = [m2, l2c, 0, 0]
    return M.ModelDesc("two_link", [0, 0, 1], [0, M.JT_RZ, M.JT_RZ], placement, inertia)


def two_link_closed_form(q, v, a, m1=1.7, m2=0.9, l1c=0.35, L1=0.6, l2c=0.25, off=0.4, g=9.81):
    """Planar two-link dynamics with point masses; gravity acts along -y of the arm's plane
    (world -z seen through the Rx(90 deg) placement).  Returns tau and the world position
    of the second mass."""
    q1, q2 = q
    p2 = off + q2
    c1, c12, cp, sp = np.cos(q1), np.cos(q1 + p2), np.cos(p2), np.sin(p2)
    M11 = m1 * l1c**2 + m2 * (L1**2 + l2c**2 + 2 * L1 * l2c * cp)
    M12 = m2 * (l2c**2 + L1 * l2c * cp)
    M22 = m2 * l2c**2
    h = m2 * L1 * l2c * sp
    g1 = g * (m1 * l1c * c1 + m2 * (L1 * c1 + l2c * c12))
    g2 = g * m2 * l2c * c12
    tau1 = M11 * a[0] + M12 * a[1] - h * (2 * v[0] * v[1] + v[1]**2) + g1
    tau2 = M12 * a[0] + M22 * a[1] + h * v[0]**2 + g2
    # mass 2 in the arm's plane, then through the first placement
    x = L1 * np.cos(q1) + l2c * np.cos(q1 + p2)
    y = L1 * np.sin(q1) + l2c * np.sin(q1 + p2)
    Rx90 = np.array([[1.0, 0, 0], [0, 0, -1], [0, 1, 0]])
    pos2 = np.array([0.1, 0.2, 0.3])[:, None] + Rx90 @ np.stack([x, y, np.zeros_like(x)])
    return np.stack([tau1, tau2]), pos2
