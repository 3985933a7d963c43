"""ESPIRiT in the oracle (recon/espirit.rs), pinned without the reference:

* the reference's own tests restated: eigenpairs of a diagonal matrix
  (recon/espirit.rs:368), sensitivity-ratio recovery (:387) and use of the
  maps by the SENSE unfold (:516);
* top eigenpairs against numpy.linalg.eigh on a well-separated spectrum;
* constant complex coil weights: inside the object the maps are proportional
  to the weights, including the phase of the ratio (a conjugated kernel or a
  reversed shift would flip it);
* the FP32 restatement tracks its FP64 twin.
"""
import numpy as np
import pytest

import oracle as O


def centred_fft2(img):
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(img, axes=(-2, -1))), axes=(-2, -1))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_eigpairs_recover_diagonal(prec):  # recon/espirit.rs:368
    m = np.diag([3.0, 1.0, 5.0, 2.0]).astype(np.complex128)
    lam, vec = O.eigpairs(m, 3, 50, prec)
    assert len(lam) == 3
    assert np.allclose(lam, [5.0, 3.0, 2.0], atol=1e-3)


def test_eigpairs_match_numpy_eigh():
    rng = np.random.default_rng(3)
    n = 12
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    ev = 10.0 * 0.5 ** np.arange(n)
    a = (q * ev) @ q.conj().T
    lam, vec = O.eigpairs(a, 4, 400, "f64")
    w, u = np.linalg.eigh(a)
    w, u = w[::-1], u[:, ::-1]
    assert np.allclose(lam, w[:4], rtol=1e-5)
    for k in range(4):
        assert abs(abs(np.vdot(u[:, k], vec[k])) - 1.0) < 1e-4


def _two_coil_case():
    nc, ny, nx = 2, 32, 32
    yy, xx = np.meshgrid(np.arange(ny) - ny / 2.0, np.arange(nx), indexing="ij")
    s1 = np.exp(-(yy ** 2 + (xx - nx / 3.0) ** 2) / 250.0)
    s2 = np.exp(-(yy ** 2 + (xx - 2.0 * nx / 3.0) ** 2) / 250.0)
    ph = np.zeros((ny, nx))
    ph[6:26, 6:26] = 1.0
    coil = np.stack([s1 * ph, s2 * ph]).astype(np.complex64)
    k = centred_fft2(coil)
    y0, x0 = (ny - 16) // 2, (nx - 16) // 2
    return coil, k[:, y0:y0 + 16, x0:x0 + 16], s1, s2


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_espirit_recovers_sensitivity_ratio(prec):  # recon/espirit.rs:387
    _, acs, s1, s2 = _two_coil_case()
    maps, kept = O.espirit_maps(acs, 32, 32, 5, 0.02, 40, prec)
    assert kept >= 1
    m0, m1 = np.abs(maps[0, 8:24, 8:24]), np.abs(maps[1, 8:24, 8:24])
    ok = m0 >= 1e-4
    want = (s2 / s1)[8:24, 8:24][ok]
    err = np.mean(np.abs(m1[ok] / m0[ok] - want) / np.maximum(want, 1e-6))
    assert err < 0.05, err


def test_espirit_maps_feed_sense_unfold():  # recon/espirit.rs:516
    nc, ny, nx, r = 2, 32, 32, 2
    yy, xx = np.meshgrid(np.arange(ny) - ny / 2.0, np.arange(nx), indexing="ij")
    ph = np.zeros((ny, nx))
    ph[8:24, 8:24] = 1.0
    coil = np.stack([np.exp(-(yy ** 2 + (xx - (c + 1.0) * nx / 3.0) ** 2) / 200.0) * ph for c in range(nc)])
    k = centred_fft2(coil.astype(np.complex64))
    acs = k[:, 8:24, 8:24]
    maps, _ = O.espirit_maps(acs, ny, nx, 5, 0.02, 40)
    assert maps.shape == (nc, ny, nx)
    half = ny // r
    al = 0.5 * (coil[:, :half] + coil[:, half:])
    aliased = np.concatenate([al, al], axis=1).astype(np.complex64)
    st, rho, _, _ = O.sense_unfold(aliased, maps, r, 1e-4, False, "f32")
    assert st == 0 and np.all(np.isfinite(rho)) and np.sum(np.abs(rho) ** 2) > 0


def test_constant_complex_weights_give_proportional_maps():
    # coil c sees the object times a constant complex weight w_c: ESPIRiT's
    # dominant eigenvector inside the object is w / |w| up to a common phase,
    # and the phase convention rotates the largest-magnitude coil to real.
    ny = nx = 32
    rng = np.random.default_rng(7)
    obj = np.zeros((ny, nx))
    obj[8:24, 6:26] = 1.0 + 0.3 * rng.random((16, 20))
    w = np.array([1.0, 0.6 * np.exp(0.7j), 0.3 * np.exp(-1.9j)])
    coil = (w[:, None, None] * obj[None]).astype(np.complex64)
    k = centred_fft2(coil)
    acs = k[:, 8:24, 8:24]
    for prec in ("f32", "f64"):
        maps, _ = O.espirit_maps(acs, ny, nx, 5, 0.02, 40, prec)
        ratio = maps[1:, 12:20, 10:22] / maps[0, 12:20, 10:22]
        want = (w[1:] / w[0])[:, None, None]
        assert np.max(np.abs(ratio - want)) < 2e-3, (prec, np.max(np.abs(ratio - want)))
        assert np.max(np.abs(maps[0, 12:20, 10:22].imag)) < 1e-5  # coil 0 is the largest: real


def test_espirit_f32_tracks_f64():
    _, acs, _, _ = _two_coil_case()
    m32, k32 = O.espirit_maps(acs, 32, 32, 5, 0.02, 30, "f32")
    m64, k64 = O.espirit_maps(acs, 32, 32, 5, 0.02, 30, "f64")
    assert k32 == k64
    rel = np.linalg.norm(m32 - m64) / np.linalg.norm(m64)
    print(f"espirit f32 vs f64: {rel:.2e} ({k32} kernels)")
    assert rel < 1e-3
