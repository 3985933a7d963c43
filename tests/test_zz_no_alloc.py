"""oks_recon_ifft_rss_dev_noalloc: the ifft-rss entry point that allocates
nothing, checked through the C ABI against oks_recon_ifft_rss_dev and the
oracle.

Bar: the image and flags of the allocating entry point -- bit-identical where
both process the same slice chunk (the chunk sets the coil grouping of the
RSS sums), rel-L2 <= 1e-6 where the static scratch makes the chunk smaller --
and the ifft-rss parity bar against the oracle (rel-L2 <= 1e-5 against FP64,
<= 1e-5 plus the FP32 restatement's own distance to FP64 against FP32).  Two
calls in a row, the second on a smaller input, then a homodyne call; an input
whose workspaces exceed the scratch is refused with OKS_E_UNSUPPORTED and the
next call still succeeds.
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import OksError
from tests.gpu_util import assert_parity, recon_dev, rel

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def _both(ctx, case, d):
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev_noalloc", d, case)
    want, wflags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert flags == wflags and got.shape == want.shape
    return got, want, flags


def _oracle(case, d, got, flags, label):
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    assert st == 0 and flags == rflags and got.shape == ref32.shape
    assert_parity(got, ref32, ref64, TOL32, label)


def test_two_calls_then_smaller(ctx):
    d = abi.IfftRssDesc.default()
    # 8 slices of the 15-coil 368 x 640 (2x oversampled) shape: 113 MB of
    # row-pass intermediate at the default chunk, so the scratch takes fewer
    # slices per chunk
    big = synth.make_case(nc=15, nz=8, ny=368, nx=640, os_x=2, kind="full", seed=41, ellipses=6)
    got, want, flags = _both(ctx, big, d)
    e = rel(got, want)
    print(f"15x8x368x640: noalloc vs allocating rel-L2 {e:.2e}, max abs {np.abs(got - want).max():.2e}")
    assert e <= 1e-6
    _oracle(big, d, got, flags, "noalloc 15x8x368x640")
    small = synth.make_case(nc=6, nz=3, ny=48, nx=64, os_x=2, kind="full", seed=42, ellipses=5)
    got, want, flags = _both(ctx, small, d)
    assert np.array_equal(got, want)
    _oracle(small, d, got, flags, "noalloc 6x3x48x64")


def test_homodyne_and_3d(ctx):
    case = synth.make_case(nc=4, nz=2, ny=64, nx=48, kind="full", seed=43, ellipses=5)
    m = case["mask"].copy()
    m[:, :10] = False
    case["mask"] = np.ascontiguousarray(m)
    case["kspace"] = np.ascontiguousarray(np.where(m[None], case["kspace"], 0).astype(np.complex64))
    d = abi.IfftRssDesc.default(remove_oversampling=0)
    got, want, flags = _both(ctx, case, d)
    assert flags & abi.F_HOMODYNE
    assert np.array_equal(got, want)
    _oracle(case, d, got, flags, "noalloc homodyne")
    vol = synth.make_case(nc=4, nz=16, ny=32, nx=48, kind="full", seed=44, ellipses=4, three_d=True)
    got, want, flags = _both(ctx, vol, abi.IfftRssDesc.default(partial_fourier=0))
    assert np.array_equal(got, want)
    _oracle(vol, abi.IfftRssDesc.default(partial_fourier=0), got, flags, "noalloc 3-D")


@pytest.mark.parametrize("three_d", [False, True])
def test_more_than_32_channels(ctx, three_d):
    # above 32 channels the whitened copy of the grid and the channel-grouped
    # row pass take their workspaces from the static scratch too
    case = synth.make_case(nc=40, nz=4, ny=40, nx=64, os_x=2, kind="full", seed=47, ellipses=5, three_d=three_d)
    d = abi.IfftRssDesc.default(partial_fourier=0)
    got, want, flags = _both(ctx, case, d)
    assert np.array_equal(got, want)
    _oracle(case, d, got, flags, f"noalloc nc=40 3d={three_d}")


def test_too_large_is_refused(ctx):
    # a 3-D volume keeps the whole row-pass intermediate: 8 x 64 x 128 x 128 x 8 B = 67 MB
    vol = synth.make_case(nc=8, nz=64, ny=128, nx=128, kind="full", seed=45, ellipses=4, three_d=True)
    with pytest.raises(OksError) as ei:
        recon_dev(ctx, "oks_recon_ifft_rss_dev_noalloc", abi.IfftRssDesc.default(partial_fourier=0), vol)
    assert ei.value.code == abi.OKS_E_UNSUPPORTED and "static" in str(ei.value)
    small = synth.make_case(nc=6, nz=2, ny=40, nx=56, kind="full", seed=46, ellipses=5)
    got, want, _ = _both(ctx, small, abi.IfftRssDesc.default())
    assert np.array_equal(got, want)
