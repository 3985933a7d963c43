"""GPU parity of the output writers' pixel rendering (cli/main.rs:1095-1153)
against the oracle.

Bars: the windows are selected elements of each slice and must be bit-equal;
the bytes are equal everywhere except where the value before rounding lies
within 1e-3 of a half-integer (the device forms norm^0.9 with the hardware
log2 / exp2 and divides by a per-slice reciprocal: a few ulp from the host's
powf and division), and there by at most one.
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import DeviceArray, OksError, lib
from openkspace_amd.device import render_slices

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def _prerounding(v, win, mode):
    lo, hi = win[:, 0][:, None, None].astype(np.float64), win[:, 1][:, None, None].astype(np.float64)
    norm = np.clip((v.astype(np.float64) - lo) / (hi - lo), 0.0, 1.0)
    return (norm if mode == abi.RENDER_LINEAR else norm ** 0.9) * 255.0


def _check(ctx, v, lo=0.5, hi=99.5, mode=0, label=""):
    dv = DeviceArray.from_host(ctx, np.ascontiguousarray(v, dtype=np.float32))
    try:
        got, win = render_slices(ctx, dv, lo, hi, mode)
    finally:
        dv.free()
    ref, rwin = O.render_slices(v, lo, hi, mode)
    assert np.array_equal(win, rwin), label
    diff = got.astype(int) - ref.astype(int)
    pre = _prerounding(v, rwin, mode)
    near = np.abs(pre - np.floor(pre) - 0.5) < 1e-3
    bad = (diff != 0) & ~near
    print(f"{label}: {np.count_nonzero(diff)} of {diff.size} bytes differ, all within 1e-3 of a half: {not bad.any()}")
    assert not bad.any() and np.abs(diff).max(initial=0) <= 1


@pytest.mark.parametrize("shape", [(3, 17, 23), (2, 64, 48), (5, 1, 1), (1, 256, 256)])
def test_random_volumes(ctx, shape):
    v = np.random.default_rng(sum(shape)).standard_normal(shape).astype(np.float32)
    _check(ctx, v, label=f"random {shape}")
    _check(ctx, v, 5.0, 95.0, label=f"random {shape} 5-95")


def test_magnitude_volume_with_ties(ctx):
    # an ifft-rss magnitude with a zeroed background: many equal values
    case = synth.make_case(nc=4, nz=4, ny=96, nx=80, kind="full", seed=5, ellipses=5)
    _, img, _, _ = O.recon("ifft_rss", abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0), case, "f32")
    img = img.copy()
    img[:, :20] = 0.0
    _check(ctx, img, label="magnitude with ties")
    _check(ctx, img, 0.0, 100.0, label="magnitude full range")


def test_c1_image_volume(ctx):
    v = np.abs(np.random.default_rng(7).standard_normal((16, 368, 320))).astype(np.float32)
    _check(ctx, v, label="16 x 368 x 320")


def test_gfactor_linear(ctx):
    rng = np.random.default_rng(3)
    g = (1.0 + np.abs(rng.standard_normal((3, 40, 50))) * 0.7).astype(np.float32)
    _check(ctx, g, mode=abi.RENDER_LINEAR, label="g-factor linear")
    _check(ctx, np.ones((2, 8, 8), np.float32), mode=abi.RENDER_LINEAR, label="g-factor flat")


def test_invalid_windows(ctx):
    dv = DeviceArray.from_host(ctx, np.zeros((1, 4, 4), np.float32))
    try:
        for lo, hi in [(50.0, 50.0), (-1.0, 50.0), (100.0, 100.0), (10.0, 101.0)]:
            with pytest.raises(OksError) as e:
                render_slices(ctx, dv, lo, hi)
            assert e.value.code == abi.OKS_E_INVALID_ARG
    finally:
        dv.free()


@pytest.mark.parametrize("shape", [(2, 101, 99), (3, 256, 256)])
def test_select_split_over_workgroups(ctx, shape):
    # slices of 8192 pixels or more are split over several workgroups per
    # digit pass (101 x 99 also leaves the float4 path: 9999 is odd); heavy
    # ties and negative values exercise the per-digit counts and the carry of
    # the remaining rank from pass to pass
    rng = np.random.default_rng(shape[1])
    v = np.round(rng.standard_normal(shape) * 4.0).astype(np.float32) / 4.0
    _check(ctx, v, label=f"ties {shape}")
    _check(ctx, v, 0.0, 100.0, label=f"ties {shape} full range")
    _check(ctx, np.abs(v) + 1.0, mode=abi.RENDER_LINEAR, label=f"linear {shape}")
