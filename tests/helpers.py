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
