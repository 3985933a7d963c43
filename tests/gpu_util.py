"""Helpers for the GPU parity tests (calls go through liboks_hip.so)."""
import numpy as np

from openkspace_amd.device import DevCase, fft_centered, recon_dev  # noqa: F401


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
