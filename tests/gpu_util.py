"""Helpers for the GPU parity tests (calls go through liboks_hip.so)."""
import numpy as np

from openkspace_amd.device import DevCase, fft_centered, recon_dev  # noqa: F401


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def assert_parity(got, r32, r64, tol, label="", iterative=False):
    """FP32-path parity bar.

    The backend computes its small dense solves (noise Cholesky, GRAPPA ridge
    solve, SENSE R x R) in FP64 where the reference restatement uses FP32, so
    its output can be closer to the FP64 twin than the FP32 restatement is.
    Direct paths pass iff rel(got, FP64) <= tol and
    rel(got, FP32) <= tol + rel(FP32, FP64).
    Iterative FP32 paths (FISTA) accumulate rounding over the iterations, so
    the FP32 restatement itself can sit further than tol from FP64; they pass
    iff the backend is no further from FP64 than the restatement is (plus
    tol), and within tol + 2 rel(FP32, FP64) of the restatement.
    """
    e32, e64, e_ref = rel(got, r32), rel(got, r64), rel(r32, r64)
    print(f"{label}: rel vs fp32 oracle {e32:.2e}, vs fp64 oracle {e64:.2e} (fp32 oracle vs fp64 {e_ref:.2e})")
    if iterative:
        assert e64 <= tol + e_ref, (e32, e64, e_ref)
        assert e32 <= tol + 2 * e_ref, (e32, e64, e_ref)
    else:
        assert e64 <= tol, (e32, e64, e_ref)
        assert e32 <= tol + e_ref, (e32, e64, e_ref)
    return e32, e64
