"""GPU parity of the function-level entry points against the oracle: RSS
(recon/coil.rs:24-27), the noise whitener and its application
(recon/prewhiten.rs:48-142), center_crop_3d (recon/crop.rs:13-27) and the
oversampling remover (recon/oversampling.rs:50-154).

Bars: crop windows are integer work and must be bit-exact; the FP32 paths
meet rel-L2 <= 1e-5 against the FP64 twin and <= 1e-5 plus the FP32
restatement's own distance to FP64 against the FP32 restatement.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi
from openkspace_amd._lib import DeviceArray, check, lib
from tests.gpu_util import assert_parity, rel

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def _cplx(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("nc,inner", [(1, 7), (15, 368 * 320), (32, 1000)])
def test_rss(ctx, nc, inner):
    x = _cplx(np.random.default_rng(nc), nc, inner)
    dx, do = DeviceArray.from_host(ctx, x), DeviceArray(ctx, (inner,), np.float32)
    check(ctx.ptr, lib().oks_rss_dev(ctx.ptr, dx.addr, nc, inner, do.addr))
    ctx.synchronize()
    got = do.download()
    assert_parity(got, O.rss(x, "f32"), O.rss(x, "f64"), TOL, f"rss nc={nc}")


@pytest.mark.parametrize("nc,ns", [(4, 256), (15, 256), (32, 300), (33, 256), (48, 256), (64, 1000)])
def test_prewhiten_calibrate_and_apply(ctx, nc, ns):
    rng = np.random.default_rng(100 + nc)
    A = _cplx(rng, nc, nc) / np.sqrt(2 * nc)
    noise = (A @ _cplx(rng, nc, ns)).astype(np.complex64)
    dn, dw = DeviceArray.from_host(ctx, noise), DeviceArray(ctx, (nc, nc), np.complex64)
    check(ctx.ptr, lib().oks_prewhiten_calibrate_dev(ctx.ptr, dn.addr, nc, ns, dw.addr))
    ctx.synchronize()
    W = dw.download()
    W32, W64 = O.prewhiten_calibrate(noise, "f32"), O.prewhiten_calibrate(noise, "f64")
    assert_parity(W, W32, W64, TOL, f"whitener nc={nc}")
    assert np.all(W[np.triu_indices(nc, 1)] == 0)  # lower triangular
    data = _cplx(rng, nc, 5000)
    dd = DeviceArray.from_host(ctx, data)
    check(ctx.ptr, lib().oks_prewhiten_apply_dev(ctx.ptr, dw.addr, dd.addr, nc, 5000))
    ctx.synchronize()
    got = dd.download()
    assert_parity(got, O.prewhiten_apply(W64, data, "f32"), O.prewhiten_apply(W64, data, "f64"), TOL,
                  f"whiten apply nc={nc}")
    # the backend's W whitens the calibration noise to the identity
    # covariance, up to W's FP32 rounding
    wn = O.prewhiten_apply(W, noise, "f64")
    cov = wn @ wn.conj().T / (ns - 1)
    assert np.max(np.abs(cov - np.eye(nc))) < 1e-5


def test_prewhiten_calibrate_errors(ctx):
    noise = np.ones((4, 2), np.complex64)
    dn, dw = DeviceArray.from_host(ctx, noise), DeviceArray(ctx, (4, 4), np.complex64)
    assert lib().oks_prewhiten_calibrate_dev(ctx.ptr, dn.addr, 4, 2, dw.addr) == abi.OKS_E_INCONSISTENT


@pytest.mark.parametrize("shape,target", [((16, 368, 640), (0, 368, 320)), ((5, 9, 8), (3, 4, 5)),
                                          ((2, 7, 6), (9, 9, 9)), ((4, 6, 6), (1, 1, 1)), ((3, 5, 7), (-1, 2, 0))])
def test_crop_windows_exact(ctx, shape, target):
    v = np.random.default_rng(7).standard_normal(shape).astype(np.float32)
    dv, do = DeviceArray.from_host(ctx, v), DeviceArray(ctx, (v.size,), np.float32)
    shp = np.zeros(3, np.int64)
    check(ctx.ptr, lib().oks_crop_dev(ctx.ptr, dv.addr, *shape, *target, do.addr, shp.ctypes.data))
    ctx.synchronize()
    want = O.crop3d(v, target, "f32")
    assert tuple(shp) == want.shape
    got = do.download()[: want.size].reshape(want.shape)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("enc,rec", [(640, 320), (256, 128), (384, 128), (96, 12), (368, 46)])
def test_os_remove(ctx, enc, rec):
    lines = _cplx(np.random.default_rng(enc + rec), 37, enc)
    dl, do = DeviceArray.from_host(ctx, lines), DeviceArray(ctx, (37, rec), np.complex64)
    check(ctx.ptr, lib().oks_os_remove_dev(ctx.ptr, dl.addr, 37, enc, rec, do.addr))
    ctx.synchronize()
    got = do.download()
    assert_parity(got, O.os_remove(lines, rec, "f32"), O.os_remove(lines, rec, "f64"), TOL, f"os {enc}->{rec}")


def test_os_remove_dc_known_answer(ctx):
    # recon/oversampling.rs:237: a DC spike of 10 at the centre of 16 encoded
    # samples becomes one spike of 10 * 8 / 16 at the centre of 8
    enc, rec = 16, 8
    lines = np.zeros((1, enc), np.complex64)
    lines[0, enc // 2] = 10.0
    dl, do = DeviceArray.from_host(ctx, lines), DeviceArray(ctx, (1, rec), np.complex64)
    check(ctx.ptr, lib().oks_os_remove_dev(ctx.ptr, dl.addr, 1, enc, rec, do.addr))
    ctx.synchronize()
    got = do.download()[0]
    want = np.zeros(rec, np.complex64)
    want[rec // 2] = 10.0 * rec / enc
    assert np.max(np.abs(got - want)) < 1e-5


@pytest.mark.parametrize("enc,rec", [(640, 0), (320, 320), (300, 200), (900, 100)])
def test_os_remove_none_cases(ctx, enc, rec):
    # OversamplingRemover::new returns None (recon/oversampling.rs:50-70):
    # reported as unsupported, as the oracle reports it
    lines = np.zeros((2, enc), np.complex64)
    dl, do = DeviceArray.from_host(ctx, lines), DeviceArray(ctx, (2, max(rec, 1)), np.complex64)
    assert lib().oks_os_remove_dev(ctx.ptr, dl.addr, 2, enc, rec, do.addr) == abi.OKS_E_UNSUPPORTED
    assert O.os_remove(lines, rec, "f32") is None


@pytest.mark.parametrize("fn,desc,limit", [("oks_recon_ifft_rss_dev", abi.IfftRssDesc.default(), 64),
                                           ("oks_recon_grappa_dev", abi.GrappaDesc.default(), 64),
                                           ("oks_recon_sense_dev", abi.SenseDesc.default(), 32),
                                           ("oks_recon_cs_dev", abi.CsDesc.default(iters=2), 64)])
def test_channel_limit_refused(ctx, fn, desc, limit):
    # A documented limit of this backend (the reference has none): ifft-rss,
    # GRAPPA and CS take up to 64 channels, SENSE up to 32; larger grids are
    # refused with OKS_E_UNSUPPORTED before any work.
    from openkspace_amd import synth
    from openkspace_amd._lib import OksError
    from tests.gpu_util import recon_dev
    case = synth.make_case(nc=limit + 1, nz=1, ny=32, nx=32, kind="full", seed=3, ellipses=3)
    with pytest.raises(OksError) as e:
        recon_dev(ctx, fn, desc, case)
    assert e.value.code == abi.OKS_E_UNSUPPORTED and str(limit) in str(e.value)
