"""The oracle's partial-Fourier homodyne restatement (recon/partial_fourier.rs),
pinned without running the reference:

* the reference's own tests: the plan of a 6/8 acquisition
  (partial_fourier.rs:383-398) and the weight checks of
  homodyne_weights_sum_plausible (:401-420);
* weight invariants: H(dc + d) + H(dc - d) = 2 wherever both rows are
  acquired (the symmetric band and the inner half of the ramps), L symmetric
  about DC and zero outside [dc - sym, dc + sym];
* an exact answer: for positive real coil images whose k-space is zero on the ramp
  rows and the Nyquist row, H(k) + H(-k) = 2 on every remaining row, so the
  homodyne image of the clipped k-space is the RSS of the real images
  (odd and even sizes, which also pins NumPy's shift pair of ifft2_plane);
* the strategy dispatch: ifft-rss on a partial-Fourier mask returns the
  homodyne image (cropped like the plain path) and the F_HOMODYNE flag.
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth


def _mask_rows(ny, lo, hi=None, nz=1, nx=4):
    m = np.zeros((nz, ny, nx), np.uint8)
    m[:, lo:(ny if hi is None else hi + 1)] = 1
    return m


def test_reference_plan_6_8():  # partial_fourier.rs:383-398
    p = O.pf_plan(_mask_rows(64, 8), 32)
    assert (p["ky_lo"], p["ky_hi"], p["ky_dc"], p["sym_half"]) == (8, 63, 32, 24)
    assert p["ramp"] == 6  # max(24 / 4, 2).min(23)


def test_reference_weights_plausible():  # partial_fourier.rs:401-420
    p = O.pf_plan(_mask_rows(128, 16, nx=2), 64)
    _, w = O.pf_weights(p)
    assert np.all(w[:16] == 0.0)
    assert abs(w[64] - 1.0) < 1e-4
    assert abs(w[127] - 2.0) < 1e-4


@pytest.mark.parametrize("ny,lo,hi", [(64, 8, 63), (128, 16, 127), (65, 8, 64), (64, 0, 50), (96, 20, 95),
                                      (200, 0, 140)])
def test_weight_invariants(ny, lo, hi):
    dc = ny // 2
    p = O.pf_plan(_mask_rows(ny, lo, hi), dc)
    assert p is not None
    low, H = O.pf_weights(p)
    sym = p["sym_half"]
    assert H[dc] == 1.0 and low[dc] == pytest.approx(1.0, abs=1e-6)
    for d in range(1, sym + 1):  # both rows acquired
        assert H[dc + d] + H[dc - d] == pytest.approx(2.0, abs=2e-7), d
    for d in range(sym + p["ramp"] + 1, ny):  # long tail 2, short tail 0
        for ky in (dc + d, dc - d):
            if 0 <= ky < ny:
                assert H[ky] in (0.0, 2.0)
    assert np.all(H[:p["ky_lo"]] == 0) and np.all(H[p["ky_hi"] + 1:] == 0)
    assert np.all(low[: max(dc - sym, 0)] == 0) and np.all(low[dc + sym + 1:] == 0)
    for d in range(0, sym + 1):
        if dc + d < ny:
            assert low[dc + d] == pytest.approx(low[dc - d], abs=1e-6)


def _np_centred_fft2(x):
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x, axes=(-2, -1)), axes=(-2, -1)), axes=(-2, -1))


def _np_centred_ifft2(k):
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(k, axes=(-2, -1)), axes=(-2, -1)), axes=(-2, -1))


def _real_pf_case(nc, nz, ny, nx, lo, seed):
    """Positive real coil images with their k-space zeroed on the ramp rows
    (both sides of DC) and the self-mirrored Nyquist row; returns (clipped
    k-space, plan, expected RSS)."""
    rng = np.random.default_rng(seed)
    dc = ny // 2
    p = O.pf_plan(_mask_rows(ny, lo), dc)
    sym, r = p["sym_half"], p["ramp"]
    # positive images keep the low-pass phase estimate away from zero: the
    # f32 Hann weights are symmetric only to an ulp, so where the low-pass
    # image crosses zero its phase is undefined and so is the demodulation
    x = 6.0 + rng.standard_normal((nc, nz, ny, nx))
    k = _np_centred_fft2(x)
    f = np.arange(ny) - dc
    zero = (np.abs(f) > sym - r) & (np.abs(f) <= sym + r)
    if ny % 2 == 0:
        zero[0] = True  # frequency -ny/2 is its own mirror
    k[:, :, zero] = 0
    xr = _np_centred_ifft2(k)
    assert np.max(np.abs(xr.imag)) < 1e-12  # still real: the zeroed set is mirror-symmetric
    expect = np.sqrt((xr.real ** 2).sum(0))
    # a constant phase per coil and slice: the demodulation must remove it
    k = k * np.exp(1j * rng.uniform(-np.pi, np.pi, (nc, nz, 1, 1)))
    k[:, :, :lo] = 0  # the missing short tail
    return k, p, expect


@pytest.mark.parametrize("ny,nx,lo", [(64, 48, 8), (65, 33, 8), (96, 40, 20), (81, 64, 18)])
def test_real_object_exact(ny, nx, lo):
    k, p, expect = _real_pf_case(3, 2, ny, nx, lo, seed=ny + nx)
    # exact up to the f32 weights, whose Hann values are symmetric about DC
    # only to an ulp: the low-pass phase then carries a ~1e-8 residue
    got = O.homodyne(k, p, "f64")
    assert np.max(np.abs(got - expect)) < 1e-7 * np.max(expect)
    g32 = O.homodyne(k.astype(np.complex64), p, "f32")
    assert np.linalg.norm(g32 - expect) / np.linalg.norm(expect) < 1e-5


def _np_homodyne(k, plan, pre, post):
    """homodyne_reconstruct written with numpy, the rotation amounts of the
    plane transform given as functions of the axis length."""
    low, H = O.pf_weights(plan)
    ny, nx = k.shape[-2:]

    def plane(a):
        a = np.roll(np.roll(a, -pre(ny), axis=-2), -pre(nx), axis=-1)
        a = np.fft.ifft2(a, axes=(-2, -1))
        return np.roll(np.roll(a, -post(ny), axis=-2), -post(nx), axis=-1)

    h = plane(k * H[:, None])
    lp = plane(k * low[:, None])
    nrm = np.abs(lp)
    safe = np.where(nrm > 1e-12, nrm, 1.0)
    dm = np.where(nrm > 1e-12, np.conj(lp) / safe, 1.0)
    return np.sqrt(((h * dm).real ** 2).sum(0))


def test_exact_answer_discriminates_the_shift_pair():
    # the exact answer holds for NumPy's pair (floor before, ceil after), the
    # pair ifft2_plane uses; recon/shift.rs's pair (ceil before, floor after)
    # gives a shifted, phase-ramped plane for odd sizes, which it rejects
    k, p, expect = _real_pf_case(2, 1, 65, 33, 8, seed=7)
    numpy_pair = _np_homodyne(k, p, lambda n: n // 2, lambda n: (n + 1) // 2)
    shift_rs_pair = _np_homodyne(k, p, lambda n: (n + 1) // 2, lambda n: n // 2)
    assert np.max(np.abs(numpy_pair - expect)) < 1e-7 * expect.max()
    assert np.max(np.abs(shift_rs_pair - expect)) > 1e-2 * expect.max()
    assert np.max(np.abs(O.homodyne(k, p, "f64") - expect)) < 1e-7 * expect.max()


def test_exact_answer_discriminates_the_demodulation():
    # without the demodulation the coil phases leak into the real part
    k, p, expect = _real_pf_case(3, 1, 64, 40, 8, seed=11)
    low, H = O.pf_weights(p)
    h = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(k * H[:, None], axes=(-2, -1)), axes=(-2, -1)),
                        axes=(-2, -1))
    assert np.max(np.abs(np.sqrt((h.real ** 2).sum(0)) - expect)) > 1e-2 * expect.max()


def test_strategy_dispatch_and_crop():
    case = synth.make_case(nc=3, nz=2, ny=64, nx=48, kind="full", seed=5, ellipses=3)
    m = case["mask"].copy()
    m[:, :10] = False
    case["mask"] = m
    case["kspace"] = np.where(m[None], case["kspace"], 0).astype(np.complex64)
    case["hdr"]["recon_x"] = 40
    d = abi.IfftRssDesc.default(remove_oversampling=0)
    st, img, flags, _ = O.recon("ifft_rss", d, case, "f64")
    assert st == 0 and flags & abi.F_HOMODYNE and flags & abi.F_CROPPED
    W = O.prewhiten_calibrate(case["noise"], "f64")
    kw = O.prewhiten_apply(W, case["kspace"].reshape(3, -1), "f64").reshape(case["kspace"].shape)
    full = O.homodyne(kw, O.pf_plan(m, 32), "f64")
    assert img.shape == (2, 64, 40)
    assert np.allclose(img, full[:, :, 4:44], rtol=0, atol=1e-12 * full.max())
    # partial_fourier off: the plain path
    st2, img2, flags2, _ = O.recon("ifft_rss", abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0),
                                   case, "f64")
    assert st2 == 0 and not flags2 & abi.F_HOMODYNE and not np.allclose(img2, img)
