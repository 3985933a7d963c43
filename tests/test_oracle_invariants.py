"""Closed-form and invariant checks that pin the oracle's compositions where no
golden vector reaches (FISTA, strategy compositions, GRAPPA index rules)."""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def test_fista_lambda0_is_zero_filled():
    # With lambda = 0 the gradient at the zero-filled start is exactly zero
    # (F F^H M y = M y), so FISTA must return the zero-filled image.
    rng = np.random.default_rng(1)
    ny, nx = 12, 16
    k = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx)))
    mask = (rng.random((ny, nx)) < 0.4).astype(np.uint8)
    st, x = O.fista_cs(k * mask, mask, 7, 0.0, "f64")
    zf = O.centred_fft2(k * mask, True, "f64")
    assert st == 0 and rel(x, zf) < 1e-12


def test_fista_full_mask_is_one_prox():
    # Full mask: v = z - (z - atb) = atb every iteration, so x = H^-1 S_l H atb.
    rng = np.random.default_rng(2)
    ny, nx = 8, 10
    k = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    mask = np.ones((ny, nx), np.uint8)
    lam = 0.3
    st, x = O.fista_cs(k, mask, 5, lam, "f64")
    atb = O.centred_fft2(k, True, "f64")
    c = O.haar(atb, False, "f64")
    c = O.soft_threshold(c, lam, "f64")
    want = O.haar(c, True, "f64")
    assert rel(x, want) < 1e-12


def test_soft_threshold_closed_form():
    y = np.array([[1 + 1j, 0.1, 3j, -2], [0.05j, 4, 1, 1]], np.complex128)
    out = O.soft_threshold(y, 0.5, "f64")
    exp = y.copy()
    hy, hx = 1, 2
    for i in range(2):
        for j in range(4):
            if i < hy and j < hx:
                continue
            m = abs(y[i, j])
            exp[i, j] = 0 if m <= 0.5 else y[i, j] * (m - 0.5) / m
    assert np.abs(out - exp).max() < 1e-15


def test_fista_f32_tracks_f64():
    case = synth.make_case(nc=1, nz=1, ny=32, nx=32, kind="cs", seed=5)
    k = case["kspace"][0, 0]
    m = case["mask"][0]
    _, x32 = O.fista_cs(k, m, 40, 0.01, "f32")
    _, x64 = O.fista_cs(k, m, 40, 0.01, "f64")
    assert rel(x32, x64) < 1e-5


def _small_case(os_x=1, kind="full", **kw):
    p = dict(nc=3, nz=2, ny=12, nx=16 * os_x, os_x=os_x, kind=kind, seed=11, ellipses=4)
    p.update(kw)
    return synth.make_case(**p)


def test_ifft_rss_textbook_matches_numpy():
    case = _small_case(os_x=2)
    d = abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0)
    st, img, flags, _ = O.recon("ifft_rss", d, case, "f64")
    assert st == 0 and flags & abi.F_CROPPED
    W = O.prewhiten_calibrate(case["noise"].astype(np.complex128), "f64")
    k = np.einsum("ij,jzyx->izyx", W, case["kspace"].astype(np.complex128))
    im = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(k, axes=(2, 3))), axes=(2, 3))
    rss = np.sqrt((np.abs(im) ** 2).sum(0))
    nx = k.shape[3]
    exp = rss[:, :, nx // 4: nx // 4 + nx // 2]
    assert img.shape == exp.shape and rel(img, exp) < 1e-12


def test_oversampling_quirk_is_half_amplitude_half_fov():
    # remove_oversampling=true re-centres the 320-sample lines in the 640 grid
    # (io/ismrmrd/reader.rs:265,392-399): the image is 0.5 x the sinc-
    # interpolated central half of the recon FOV.  On the even samples the
    # interpolation is exact: Q[2j] = 0.5 * T[nx_rec/4 + j].
    case = _small_case(os_x=2)
    st, T, _, _ = O.recon("ifft_rss", abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0), case, "f64")
    st2, Q, flags, _ = O.recon("ifft_rss", abi.IfftRssDesc.default(remove_oversampling=1, partial_fourier=0), case, "f64")
    assert st == 0 and st2 == 0 and not (flags & abi.F_OS_SKIPPED)
    rec = T.shape[2]
    assert np.abs(Q[:, :, 0::2] - 0.5 * T[:, :, rec // 4: rec // 4 + rec // 2]).max() < 1e-12 * np.abs(T).max()


def test_prewhiten_skipped_flag():
    case = _small_case()
    case["noise"] = np.ascontiguousarray(case["noise"][:, :2])  # fewer samples than coils -> None
    st, img, flags, _ = O.recon("ifft_rss", abi.IfftRssDesc.default(partial_fourier=0), case, "f32")
    assert st == 0 and flags & abi.F_PREWHITEN_SKIPPED


def test_c2_grappa_index_rules():
    # Mask convention of the GRAPPA workload at ny=384, R=4, 24 ACS rows.
    ny, r = 384, 4
    rows = synth.grappa_mask(ny, r, 24)
    mask = np.repeat(rows[None, :, None], 8, 2).astype(np.uint8)
    assert O.detect_pattern(mask) == (4, 180, 205, 0, 380)
    fills = O.grappa_fills(rows.astype(np.uint8), r, 4)
    targets = fills[:, 0] + 4 + fills[:, 1]
    assert len(np.unique(targets)) == 261 and len(targets) == 261  # no double writes
    missing = np.where(~rows)[0]
    assert len(missing) == 270
    assert len(set(missing) - set(targets)) == 9


def test_grappa_last_writer_wins_order():
    # Adjacent ACS rows make targets reachable from two ky0; the fill list
    # must be in ascending ky0 so the later ky0 wins (recon/grappa.rs:363-420).
    s = np.zeros(40, np.uint8)
    s[::2] = 1
    s[17:24] = 1
    fills = O.grappa_fills(s, 2, 4)
    assert np.all(np.diff(fills[:, 0]) >= 0)


def test_grappa_strategy_f32_vs_f64():
    case = _small_case(kind="grappa", nc=4, nz=1, ny=48, nx=24, r=2, acs=16)
    d = abi.GrappaDesc.default()
    st, a, flags, _ = O.recon("grappa", d, case, "f32")
    st2, b, _, _ = O.recon("grappa", d, case, "f64")
    assert st == 0 and st2 == 0 and not (flags & abi.F_NO_PATTERN)
    assert rel(a, b) < 1e-3


def test_sense_strategy_fallback_indivisible():
    case = _small_case(kind="grappa", nc=4, nz=1, ny=50, nx=16, r=3, acs=16)
    st, img, flags, _ = O.recon("sense", abi.SenseDesc.default(), case, "f64")
    assert st == 0 and flags & abi.F_SENSE_FALLBACK
    st2, ref, _, _ = O.recon("ifft_rss", abi.IfftRssDesc.default(partial_fourier=0), case, "f64")
    assert rel(img, ref) < 1e-12


def test_cs_strategy_odd_fallback():
    case = _small_case(kind="full", nc=2, nz=1, ny=11, nx=16)
    st, img, flags, _ = O.recon("cs", abi.CsDesc.default(iters=3), case, "f64")
    assert st == 0 and flags & abi.F_CS_FALLBACK


def test_pf_gate_flags_small_even_fully_sampled_grid():
    # recon/partial_fourier.rs:84-108: a fully sampled even ny has ny/2 rows
    # below DC and ny/2-1 above; for ny <= 20 the asymmetry is >= 1.10, so the
    # reference routes it to homodyne.
    case = _small_case()
    st, _, flags, _ = O.recon("ifft_rss", abi.IfftRssDesc.default(), case, "f32")
    assert st == 0 and flags & abi.F_HOMODYNE
    assert not O.pf_detect(np.ones((1, 368, 4), np.uint8), 184)
