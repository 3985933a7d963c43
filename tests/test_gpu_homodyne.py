"""GPU parity of the partial-Fourier homodyne path of ifft-rss
(recon/strategy.rs:163-187, recon/partial_fourier.rs:287-356) against the
oracle.

Bar (FP32 path): rel-L2 <= 1e-5 against the FP64 restatement and <= 1e-5
plus the FP32 restatement's own distance to FP64 against the FP32 one; the
image shape and the flags are exact.  Cases: even and odd ky/kx lengths (the
plane transform uses NumPy's shift pair), either clipped side, pre-whitening,
readout-oversampling removal, the recon-matrix crop, and the exact answer of
positive real coil images (see test_oracle_homodyne).
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from tests.gpu_util import assert_parity, recon_dev, rel
from tests.test_oracle_homodyne import _real_pf_case

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def _pf_case(nc, nz, ny, nx, lo=None, hi=None, os_x=1, seed=3, recon_x=None):
    case = synth.make_case(nc=nc, nz=nz, ny=ny, nx=nx, os_x=os_x, kind="full", seed=seed, ellipses=5)
    m = case["mask"].copy()
    if lo:
        m[:, :lo] = False
    if hi is not None:
        m[:, hi + 1:] = False
    case["mask"] = np.ascontiguousarray(m)
    case["kspace"] = np.ascontiguousarray(np.where(m[None], case["kspace"], 0).astype(np.complex64))
    if recon_x is not None:
        case["hdr"]["recon_x"] = recon_x
    return case


def _check(ctx, case, d, label):
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and rflags & abi.F_HOMODYNE
    assert flags == rflags and got.shape == ref32.shape
    assert_parity(got, ref32, ref64, TOL32, label)


@pytest.mark.parametrize("ny,nx,lo,hi", [(64, 48, 10, None), (65, 33, 12, None), (96, 80, None, 70),
                                         (128, 128, 20, None), (81, 96, 17, None), (368, 64, 60, None)])
@pytest.mark.parametrize("prewhiten", [0, 1])
def test_homodyne_shapes(ctx, ny, nx, lo, hi, prewhiten):
    case = _pf_case(4, 2, ny, nx, lo=lo, hi=hi, seed=ny + nx)
    d = abi.IfftRssDesc.default(remove_oversampling=0, prewhiten=prewhiten)
    _check(ctx, case, d, f"homodyne {ny}x{nx} pw={prewhiten}")


@pytest.mark.parametrize("remove_os", [0, 1])
def test_homodyne_oversampled_readout_and_crop(ctx, remove_os):
    case = _pf_case(6, 3, 96, 128, lo=18, os_x=2, seed=9)
    d = abi.IfftRssDesc.default(remove_oversampling=remove_os)
    _check(ctx, case, d, f"homodyne os={remove_os}")


@pytest.mark.parametrize("remove_os", [0, 1])
def test_homodyne_more_than_32_channels(ctx, remove_os):
    case = _pf_case(40, 2, 64, 64, lo=12, os_x=2, seed=40)
    d = abi.IfftRssDesc.default(remove_oversampling=remove_os, prewhiten=1)
    _check(ctx, case, d, f"homodyne nc=40 os={remove_os}")


def test_homodyne_c1_geometry(ctx):
    # 15 coils, 368 ky (6/8 partial Fourier), 640 readout with 2x oversampling
    case = synth.make_config("c1_ifft_rss", nz=4)
    m = case["mask"].copy()
    m[:, :92] = False
    case["mask"] = np.ascontiguousarray(m)
    case["kspace"] = np.ascontiguousarray(np.where(m[None], case["kspace"], 0).astype(np.complex64))
    O.set_threads(8)
    try:
        _check(ctx, case, abi.IfftRssDesc.default(remove_oversampling=1), "homodyne c1 geometry")
    finally:
        O.set_threads(1)


@pytest.mark.parametrize("ny,nx,lo", [(64, 48, 8), (65, 33, 8)])
def test_homodyne_exact_answer(ctx, ny, nx, lo):
    k, p, expect = _real_pf_case(3, 2, ny, nx, lo, seed=ny + nx)
    m = np.zeros((2, ny, nx), np.uint8)
    m[:, lo:] = 1
    case = dict(kspace=np.ascontiguousarray(k.astype(np.complex64)), mask=m, noise=None,
                hdr=dict(encoded_x=nx, encoded_y=ny, encoded_z=1, recon_x=nx, recon_y=ny, recon_z=1,
                         is_3d_encoding=0))
    d = abi.IfftRssDesc.default(remove_oversampling=0, prewhiten=0)
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert flags & abi.F_HOMODYNE
    err = rel(got, expect)
    print(f"homodyne exact answer {ny}x{nx}: rel {err:.2e}")
    assert err < TOL32


def test_homodyne_off_takes_the_plain_path(ctx):
    case = _pf_case(4, 2, 64, 48, lo=10)
    d = abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0)
    _, ref, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert not flags & abi.F_HOMODYNE and not rflags & abi.F_HOMODYNE
    assert rel(got, ref) < 2e-5
