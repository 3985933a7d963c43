"""Dense brute-force pins of the oracle's iterative routines.

FISTA (recon/cs.rs:59-139) and the Walsh power iteration
(recon/sensitivity.rs:37-142) are checked here on tiny problems, where every
linear operator is an explicit FP64 matrix built from its definition, not
from the oracle's FFT or Haar code:

* the centred unitary DFT of an even length N is C[k, n] =
  exp(-2 pi i (k - N/2)(n - N/2) / N) / sqrt(N) (fftshift and ifftshift are
  both a roll by N/2 when N is even), and the 2-D transform of a row-major
  [ny, nx] array is kron(C_ny, C_nx);
* the 1-level Haar transform with the LL/HL/LH/HH quadrant layout of
  recon/wavelet.rs:32-75 is kron(H_ny, H_nx) with the orthonormal 1-D
  analysis matrix H that puts (x[2i] + x[2i+1]) / sqrt 2 at i and
  (x[2i] - x[2i+1]) / sqrt 2 at N/2 + i.

Each test also checks that the mistakes a restatement could plausibly make
(a wrong momentum factor, extrapolating from z instead of x, ISTA, a frozen
t, a flipped gradient, an e_0 power-iteration start, a missing phase fix) move the result far beyond the tolerance the
oracle is held to, so an oracle carrying any of them fails.
"""
import numpy as np
import pytest

import oracle as O


def centred_dft(n):
    k = np.arange(n) - n // 2
    return np.exp(-2j * np.pi * np.outer(k, k) / n) / np.sqrt(n)


def haar_1d(n):
    h = np.zeros((n, n))
    s = 1 / np.sqrt(2)
    for i in range(n // 2):
        h[i, 2 * i] = h[i, 2 * i + 1] = s
        h[n // 2 + i, 2 * i] = s
        h[n // 2 + i, 2 * i + 1] = -s
    return h


def soft_details(c, lam, ny, nx):
    """Complex soft threshold on every coefficient outside the LL quadrant."""
    c = c.reshape(ny, nx).copy()
    keep = np.zeros((ny, nx), bool)
    keep[: ny // 2, : nx // 2] = True
    m = np.abs(c)
    shrink = np.where(m > lam, (m - lam) / np.where(m > 0, m, 1), 0.0)
    return np.where(keep, c, c * shrink).ravel()


def dense_fista(y, mask, iters, lam, variant="reference"):
    """The FISTA iteration with dense operators; returns x_1 .. x_iters.

    variant names a plausible restatement mistake; "reference" is the
    recurrence of recon/cs.rs:95-136 (x = z = A^H y, t = 1; x+ = prox(z -
    A^H(A z - y)); t+ = (1 + sqrt(1 + 4 t^2)) / 2; z = x+ + (t - 1)/t+ (x+ - x)).
    """
    ny, nx = y.shape
    lam = float(np.float32(lam))  # the reference's lambda is an f32 (recon/cs.rs:62)
    F = np.kron(centred_dft(ny), centred_dft(nx))
    H = np.kron(haar_1d(ny), haar_1d(nx))
    M = np.diag(mask.ravel().astype(float))
    A = M @ F
    yv = (y * mask).ravel()
    atb = F.conj().T @ yv
    x = atb.copy()
    z = x.copy()
    t = 1.0
    out = []
    for _ in range(iters):
        g = A.conj().T @ (A @ z - yv)
        v = z + g if variant == "gradient_sign" else z - g
        x_new = H.T @ soft_details(H @ v, lam, ny, nx)
        t_new = 0.5 * (1 + np.sqrt(1 + 4 * t * t))
        alpha = (t - 1) / t if variant == "alpha_over_t" else (t - 1) / t_new
        if variant == "ista":
            alpha = 0.0
        base = z if variant == "extrapolate_from_z" else x
        z = x_new + alpha * (x_new - base)
        x = x_new
        if variant != "t_frozen":
            t = t_new
        out.append(x.reshape(ny, nx))
    return out


def _cs_problem(ny, nx, seed, density=0.45):
    rng = np.random.default_rng(seed)
    y = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    mask = (rng.random((ny, nx)) < density).astype(np.uint8)
    mask[ny // 2, nx // 2] = 1
    return y * mask, mask


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("ny,nx,lam,seed", [(4, 4, 0.05, 1), (4, 6, 0.2, 2), (6, 4, 0.5, 3)])
def test_fista_iterates_match_dense_fp64(ny, nx, lam, seed):
    y, mask = _cs_problem(ny, nx, seed)
    want = dense_fista(y, mask, 8, lam)
    for k in range(1, 9):
        st, got = O.fista_cs(y, mask, k, lam, "f64")
        assert st == 0
        assert _rel(got, want[k - 1]) < 1e-12, (k, _rel(got, want[k - 1]))


def test_fista_large_lambda_is_linear_ll_projection():
    # With lambda above every detail magnitude the prox is the projection P
    # onto the LL band, so x_{k+1} = P(z_k - A^H A z_k + A^H y) is linear in y:
    # iterate it with P as a plain matrix and compare, and check linearity of
    # the oracle's output in y.
    ny, nx = 4, 4
    y, mask = _cs_problem(ny, nx, 7)
    lam = 1e9
    F = np.kron(centred_dft(ny), centred_dft(nx))
    H = np.kron(haar_1d(ny), haar_1d(nx))
    keep = np.zeros((ny, nx))
    keep[: ny // 2, : nx // 2] = 1
    P = H.T @ np.diag(keep.ravel()) @ H
    A = np.diag(mask.ravel().astype(float)) @ F
    B = P @ (np.eye(ny * nx) - A.conj().T @ A)
    c = P @ (F.conj().T @ (y * mask).ravel())
    x = z = F.conj().T @ (y * mask).ravel()
    t = 1.0
    for k in range(1, 9):
        x_new = B @ z + c
        t_new = 0.5 * (1 + np.sqrt(1 + 4 * t * t))
        z = x_new + (t - 1) / t_new * (x_new - x)
        x, t = x_new, t_new
        _, got = O.fista_cs(y, mask, k, lam, "f64")
        assert _rel(got.ravel(), x) < 1e-12, k
    _, a = O.fista_cs(y, mask, 6, lam, "f64")
    _, b = O.fista_cs(2.5 * y, mask, 6, lam, "f64")
    assert _rel(b, 2.5 * a) < 1e-13


@pytest.mark.parametrize("variant", ["alpha_over_t", "extrapolate_from_z", "ista", "t_frozen",
                                     "gradient_sign"])
def test_fista_planted_mistakes_are_detected(variant):
    # An oracle with the planted mistake would produce the variant's iterates;
    # they sit far outside the 1e-12 the oracle is held to above.
    y, mask = _cs_problem(4, 4, 1)
    good = dense_fista(y, mask, 8, 0.05)
    bad = dense_fista(y, mask, 8, 0.05, variant)
    assert max(_rel(b, g) for b, g in zip(bad, good)) > 1e-4, variant
    dist = max(_rel(O.fista_cs(y, mask, k, 0.05, "f64")[1], bad[k - 1]) for k in range(1, 9))
    assert dist > 1e-4, (variant, dist)


# --- Walsh maps ---------------------------------------------------------------

def dense_walsh(img, window, iters, variant="reference"):
    """Walsh maps per voxel in FP64 (recon/sensitivity.rs:37-142)."""
    nc, ny, nx = img.shape
    out = np.zeros_like(img, dtype=np.complex128)
    for y in range(ny):
        for x in range(nx):
            ys = slice(max(y - window, 0), min(y + window + 1, ny))
            xs = slice(max(x - window, 0), min(x + window + 1, nx))
            s = img[:, ys, xs].reshape(nc, -1).astype(np.complex128)
            R = s @ s.conj().T
            v = img[:, y, x].astype(np.complex128)
            if variant == "e0_start" or not np.any(v):
                v = np.zeros(nc, np.complex128)
                v[0] = 1
            else:
                v = v / np.linalg.norm(v)
            lam = 0.0
            for _ in range(iters):
                w = R @ v
                n = np.linalg.norm(w)
                if n < 1e-20:
                    break
                v = w / n
                lam = n
            if variant != "no_phase_fix":
                m = np.abs(v)
                i = int(np.argmax(m))
                if m[i] > 0:
                    v = v * (np.conj(v[i]) / m[i])
            out[:, y, x] = v * np.sqrt(lam)
    return out


def _walsh_img(nc, ny, nx, seed):
    rng = np.random.default_rng(seed)
    img = rng.standard_normal((nc, ny, nx)) + 1j * rng.standard_normal((nc, ny, nx))
    img[:, 2, 3] = 0  # centre vector zero: the e_0 start
    return img


@pytest.mark.parametrize("iters", [1, 2, 6])
@pytest.mark.parametrize("window", [0, 1, 3])
def test_walsh_iterations_match_dense_fp64(iters, window):
    img = _walsh_img(4, 6, 7, 11)
    got = O.walsh_maps(img, window, iters, "f64")
    want = dense_walsh(img, window, iters)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_walsh_window0_closed_form():
    # With a 1x1 window R = s s^H: from the centre-voxel start R v = s |s|, so
    # after any number of iterations v = s/|s|, lambda = |s|^2 and the map is
    # s times the unit phase that makes its largest coil real and positive.
    img = _walsh_img(5, 4, 5, 12)
    got = O.walsh_maps(img, 0, 3, "f64")
    m = np.abs(img)
    i = m.argmax(0)
    ref = np.take_along_axis(img, i[None], 0)[0]
    phase = np.where(np.abs(ref) > 0, np.conj(ref) / np.where(np.abs(ref) > 0, np.abs(ref), 1), 0)
    want = img * phase[None]
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("variant", ["e0_start", "no_phase_fix"])
def test_walsh_planted_mistakes_are_detected(variant):
    img = _walsh_img(4, 6, 7, 11)
    good = dense_walsh(img, 1, 1)
    bad = dense_walsh(img, 1, 1, variant)
    assert np.abs(bad - good).max() > 1e-3 * np.abs(good).max()
