"""Pin the CPU oracle against numpy FP64 golden vectors (tests/golden/golden.npz).

FP64 twin: tight tolerances (1e-10 relative); FP32 restatement: 1e-5.
"""
import numpy as np
import pytest

import oracle as O


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 16, 23, 46, 160, 256, 320, 368, 384, 640])
def test_dft_every_radix(golden, n):
    x = golden[f"dft{n}_in"]
    assert rel(O.fft1d(x, False, "f64"), golden[f"dft{n}_fwd"]) < 1e-12
    assert rel(O.fft1d(x, True, "f64"), golden[f"dft{n}_inv"]) < 1e-12
    assert rel(O.fft1d(x, True, "f32"), golden[f"dft{n}_inv"]) < 2e-6


@pytest.mark.parametrize("name", ["ifft2", "ifft2odd", "ifft3", "ifft1", "ifft2big"])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6)])
def test_centred_ifft(golden, name, prec, tol):
    x = golden[f"{name}_in"]
    axes = tuple(int(a) for a in golden[f"{name}_axes"])
    assert rel(O.ifft_centered(x, axes, prec), golden[f"{name}_out"]) < tol


def test_unitary_pair(golden):
    x = golden["cfft2_in"]
    assert rel(O.centred_fft2(x, False, "f64"), golden["cfft2_fwd"]) < 1e-12
    assert rel(O.centred_fft2(x, True, "f64"), golden["cfft2_inv"]) < 1e-12
    assert rel(O.centred_fft2(x, False, "f32"), golden["cfft2_fwd"]) < 2e-6


def test_prewhiten_calibration(golden):
    W = O.prewhiten_calibrate(golden["pw_noise"], "f64")
    assert rel(W, golden["pw_W"]) < 1e-10
    L = O.cholesky(golden["pw_psi"], "f64")
    assert rel(L, golden["pw_L"]) < 1e-12
    assert rel(O.invert_lower(L, "f64"), golden["pw_W"]) < 1e-10
    W32 = O.prewhiten_calibrate(golden["pw_noise"], "f32")
    assert rel(W32, golden["pw_W"]) < 1e-4


def test_prewhiten_apply_identity_cov(golden):
    # Invariant: whitened calibration noise has covariance exactly I (FP64).
    noise = golden["pw_noise"]
    W = O.prewhiten_calibrate(noise, "f64")
    y = O.prewhiten_apply(W, noise, "f64")
    cov = y @ y.conj().T / (noise.shape[1] - 1)
    assert np.abs(cov - np.eye(noise.shape[0])).max() < 1e-10
    # and apply == W @ s for a lower-triangular W
    assert rel(y, W @ noise) < 1e-13


@pytest.mark.parametrize("enc,rec", [(16, 8), (640, 320), (384, 128)])
def test_os_removal(golden, enc, rec):
    x = golden[f"os{enc}_{rec}_in"]
    assert rel(O.os_remove(x, rec, "f64"), golden[f"os{enc}_{rec}_out"]) < 1e-12
    assert rel(O.os_remove(x, rec, "f32"), golden[f"os{enc}_{rec}_out"]) < 2e-6


def test_grappa_calibrate_matches_numpy_solve(golden):
    r, kky, kkx = (int(v) for v in golden["gr_params"])
    st, w = O.grappa_calibrate(golden["gr_acs"], r, kky, kkx, float(golden["gr_ridge"]), "f64")
    assert st == 0
    # ridge is rounded to f32 exactly as the reference's `ridge: f32` field
    assert rel(w, golden["gr_weights"]) < 1e-6
    st, w32 = O.grappa_calibrate(golden["gr_acs"], r, kky, kkx, float(golden["gr_ridge"]), "f32")
    assert st == 0 and rel(w32, golden["gr_weights"]) < 5e-3


def test_grappa_synthesize_matches_numpy(golden):
    r, kky, kkx = (int(v) for v in golden["gr_params"])
    k = O.grappa_synthesize(golden["gr_weights"], golden["gr_k_in"], r, kky, kkx, "f64")
    assert rel(k, golden["gr_k_out"]) < 1e-13


def test_sense_unfold_and_gfactor(golden):
    r = int(golden["se_params"][0])
    st, rho, g, failed = O.sense_unfold(golden["se_aliased"], golden["se_maps"], r,
                                        float(golden["se_ridge"]), True, "f64")
    assert st == 0 and failed == 0
    assert rel(rho, golden["se_rho"]) < 1e-6
    assert rel(g, golden["se_g"]) < 1e-6
    st, rho32, g32, _ = O.sense_unfold(golden["se_aliased"], golden["se_maps"], r,
                                       float(golden["se_ridge"]), True, "f32")
    assert rel(rho32, golden["se_rho"]) < 1e-4 and rel(g32, golden["se_g"]) < 1e-4


def test_walsh_converges_to_eigh(golden):
    win = int(golden["wa_params"][0])
    maps = O.walsh_maps(golden["wa_img"], win, 300, "f64")
    assert rel(maps, golden["wa_maps"]) < 1e-6


def test_haar(golden):
    assert rel(O.haar(golden["haar_in"], False, "f64"), golden["haar_out"]) < 1e-14
    assert rel(O.haar(golden["haar_out"], True, "f64"), golden["haar_in"]) < 1e-14
