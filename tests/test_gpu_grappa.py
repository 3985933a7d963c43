"""GPU parity of GRAPPA calibration, synthesis and the GrappaRss strategy.

Bars: index rules (which rows are filled / left alone) are bit-exact; the
split-BF16 synthesis is within 1e-3 (rel-L2) of the FP64 oracle, and within
1e-3 plus the FP32 restatement's own distance to FP64 of the FP32 restatement
(the restatement's FP32 Cholesky + explicit inverse of the ridge normal
equations is itself ~1e-2 from FP64 at the C2 shape; the backend solves them
in FP64).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import DeviceArray, OksError, check, lib
from tests.gpu_util import assert_parity, recon_dev, rel

pytestmark = pytest.mark.gpu
TOL = 1e-3


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def gpu_calibrate(ctx, acs, r, kky, kkx, ridge):
    acs = np.ascontiguousarray(acs, dtype=np.complex64)
    nc, ny_acs, nx = acs.shape
    d_acs = DeviceArray.from_host(ctx, acs)
    d_w = DeviceArray(ctx, (r - 1, nc, nc * kky * kkx), np.complex64)
    st = lib().oks_grappa_calibrate_dev(ctx.ptr, d_acs.addr, nc, ny_acs, nx, r, kky, kkx, C.c_float(ridge), d_w.addr)
    check(ctx.ptr, st)
    w = d_w.download()
    d_acs.free()
    d_w.free()
    return w


def gpu_synthesize(ctx, w, k, r, kky, kkx):
    k = np.ascontiguousarray(k, dtype=np.complex64)
    d_k = DeviceArray.from_host(ctx, k)
    d_w = DeviceArray.from_host(ctx, np.ascontiguousarray(w, dtype=np.complex64))
    check(ctx.ptr, lib().oks_grappa_synthesize_dev(ctx.ptr, d_w.addr, d_k.addr, *k.shape, r, kky, kkx))
    ctx.synchronize()
    out = d_k.download()
    d_k.free()
    d_w.free()
    return out


def test_synthesize_golden(ctx, golden):
    r, kky, kkx = (int(v) for v in golden["gr_params"])
    k_in, k_out = golden["gr_k_in"], golden["gr_k_out"]
    got = gpu_synthesize(ctx, golden["gr_weights"], k_in, r, kky, kkx)
    # index rules bit-exact: sampled rows untouched, unreachable rows still zero
    filled_ref = np.abs(k_out[0, 0]).sum(1) != 0
    filled_got = np.abs(got[0, 0]).sum(1) != 0
    assert np.array_equal(filled_ref, filled_got)
    sampled = np.abs(k_in[0, 0]).sum(1) != 0
    assert np.array_equal(got[:, :, sampled, :], k_in[:, :, sampled, :].astype(np.complex64))
    err = rel(got, k_out)
    print(f"synthesis vs numpy FP64: {err:.2e}")
    assert err < 1e-4


def test_calibrate_golden(ctx, golden):
    r, kky, kkx = (int(v) for v in golden["gr_params"])
    w = gpu_calibrate(ctx, golden["gr_acs"], r, kky, kkx, float(golden["gr_ridge"]))
    # compare through the synthesis the weights drive (well-conditioned output)
    k_gpu = O.grappa_synthesize(w, golden["gr_k_in"], r, kky, kkx, "f64")
    err_k = rel(k_gpu, golden["gr_k_out"])
    print(f"weights rel {rel(w, golden['gr_weights']):.2e}, synthesized k-space rel {err_k:.2e}")
    assert err_k < TOL


def test_calibrate_errors(ctx):
    acs = np.ones((2, 10, 16), np.complex64)
    with pytest.raises(OksError) as e:
        gpu_calibrate(ctx, acs, 2, 3, 5, 1e-3)
    assert e.value.code == abi.OKS_E_GRAPPA_BAD_CONFIG
    with pytest.raises(OksError) as e:
        gpu_calibrate(ctx, acs, 4, 4, 5, 1e-3)   # needs (4-1)*4+1+4 = 17 rows
    assert e.value.code == abi.OKS_E_GRAPPA_ACS_TOO_SMALL


def test_calibrate_cholesky_failure(ctx):
    acs = np.ones((2, 24, 16), np.complex64)
    acs[0, 3, 4] = np.nan
    with pytest.raises(OksError) as e:
        gpu_calibrate(ctx, acs, 2, 4, 5, 1e-3)
    assert e.value.code == abi.OKS_E_GRAPPA_CHOLESKY


def _gcase(**kw):
    p = dict(nc=8, nz=2, ny=64, nx=48, kind="grappa", r=2, acs=16, seed=31, ellipses=6)
    p.update(kw)
    return synth.make_case(**p)


@pytest.mark.parametrize("r,acs", [(2, 16), (3, 18), (4, 24)])
def test_grappa_strategy_small(ctx, r, acs):
    case = _gcase(r=r, acs=acs, ny=96)
    d = abi.GrappaDesc.default()
    st, r32, rflags, _ = O.recon("grappa", d, case, "f32")
    _, r64, _, _ = O.recon("grappa", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_grappa_dev", d, case)
    assert st == 0 and got.shape == r32.shape and flags == rflags
    assert_parity(got, r32, r64, TOL, f"grappa R={r}")


@pytest.mark.parametrize("nc,r,kky,kkx,acs", [(40, 2, 4, 5, 16), (64, 3, 2, 3, 18)])
def test_grappa_more_than_32_channels(ctx, nc, r, kky, kkx, acs):
    # above 32 channels the fill runs as the per-row MFMA product with the
    # target channels in tiles of 32 (the grouped synthesis holds 32)
    case = _gcase(nc=nc, r=r, acs=acs, seed=nc)
    d = abi.GrappaDesc.default(prewhiten=1, kernel_ky=kky, kernel_kx=kkx)
    st, r32, rflags, _ = O.recon("grappa", d, case, "f32")
    _, r64, _, _ = O.recon("grappa", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_grappa_dev", d, case)
    assert st == 0 and got.shape == r32.shape and flags == rflags
    assert_parity(got, r32, r64, TOL, f"grappa nc={nc} R={r}")


def test_grappa_no_pattern_and_bad_config(ctx):
    case = _gcase(kind="full")
    d = abi.GrappaDesc.default()
    _, ref, rflags, _ = O.recon("grappa", d, case, "f32")
    got, flags, _ = recon_dev(ctx, "oks_recon_grappa_dev", d, case)
    assert flags == rflags and flags & abi.F_NO_PATTERN
    assert rel(got, ref) < 2e-5
    case = _gcase()
    d = abi.GrappaDesc.default(kernel_ky=3)
    _, ref, rflags, _ = O.recon("grappa", d, case, "f32")
    got, flags, _ = recon_dev(ctx, "oks_recon_grappa_dev", d, case)
    assert flags == rflags and flags & abi.F_GRAPPA_CALIB_FAILED
    assert rel(got, ref) < 2e-5


def test_grappa_oversampled_grid(ctx):
    # remove_oversampling on a 2x readout: the reference re-centres the
    # 1/2-length lines in the full grid before GRAPPA (io/ismrmrd/reader.rs:392-399)
    case = _gcase(nx=64, os_x=2, r=2, acs=16, ny=64)
    d = abi.GrappaDesc.default()
    st, r32, rflags, _ = O.recon("grappa", d, case, "f32")
    _, r64, _, _ = O.recon("grappa", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_grappa_dev", d, case)
    assert st == 0 and flags == rflags and not flags & abi.F_OS_SKIPPED
    assert_parity(got, r32, r64, TOL, "grappa os")


@pytest.mark.parametrize("nc,nx", [(12, 256), (5, 320)])
def test_grappa_oversampled_grid_compiled_pairs(ctx, nc, nx):
    # readouts of 256 and 320 samples take the oversampling chain with
    # compile-time transform pairs (256 -> 128, 320 -> 160); 12 coils give a
    # workgroup wider than 8 wavefronts
    case = _gcase(nc=nc, nx=nx, os_x=2, r=2, acs=16, ny=40)
    d = abi.GrappaDesc.default()
    st, r32, rflags, _ = O.recon("grappa", d, case, "f32")
    _, r64, _, _ = O.recon("grappa", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_grappa_dev", d, case)
    assert st == 0 and flags == rflags and not flags & abi.F_OS_SKIPPED
    assert_parity(got, r32, r64, TOL, f"grappa os nc={nc} nx={nx}")


def test_grappa_c2_shape(ctx):
    # The C2 slice shape (32 coils, 384 x 384, R=4, 24 ACS rows, kernel 4x5,
    # ridge 1e-3) on 2 of the 8 slices: every slice is calibrated and
    # synthesised independently, so 2 slices exercise the full per-slice path.
    case = synth.make_config("c2_grappa", nz=2)
    d = abi.GrappaDesc.default()
    O.set_threads(16)
    st, r32, rflags, _ = O.recon("grappa", d, case, "f32")
    _, r64, _, _ = O.recon("grappa", d, case, "f64")
    O.set_threads(1)
    got, flags, _ = recon_dev(ctx, "oks_recon_grappa_dev", d, case)
    assert st == 0 and flags == rflags and got.shape == (2, 384, 384)
    assert_parity(got, r32, r64, TOL, "C2 grappa")


@pytest.mark.parametrize("nc,r,kky,kkx", [(32, 4, 4, 5), (15, 4, 4, 5), (17, 3, 2, 3), (4, 8, 2, 7), (1, 6, 4, 1),
                                          (9, 5, 4, 5)])
def test_synthesis_forms(nc, r, kky, kkx):
    # The grouped synthesis (default: one workgroup per source row group, the
    # source tile resident in LDS, three target rows per pass) and the
    # row-per-workgroup form (option grappa_synth = 0) fill the same rows and
    # leave every other sample bit-equal; both are within the split-BF16 bar
    # of the FP64 oracle.  Covers both coil paddings (16, 32), several passes
    # (R = 5, 6, 8) and single- and wide-tap kernels.
    from openkspace_amd._lib import Context
    rng = np.random.default_rng(nc * 100 + r * 10 + kkx)
    ny, nx, nz = 8 * r + 40, 70, 2
    k = (rng.standard_normal((nc, nz, ny, nx)) + 1j * rng.standard_normal((nc, nz, ny, nx))).astype(np.complex64)
    rows = (np.arange(ny) % r == 1)
    k[:, :, ~rows, :] = 0
    n_src = nc * kky * kkx
    w = ((rng.standard_normal((r - 1, nc, n_src)) + 1j * rng.standard_normal((r - 1, nc, n_src)))
         / np.sqrt(n_src)).astype(np.complex64)
    outs = []
    for form in (1, 0):
        c = Context(0)
        try:
            c.set_option("grappa_synth", form)
            outs.append(gpu_synthesize(c, w, k, r, kky, kkx))
        finally:
            c.close()
    grouped, rowwise = outs
    want = O.grappa_synthesize(w, k, r, kky, kkx, "f64")
    filled = np.any(want != k, axis=(0, 3))  # [nz, ny] rows the reference writes
    keep = np.broadcast_to(~filled[None, :, :, None], k.shape)
    for got in (grouped, rowwise):
        assert np.array_equal(got[keep], k[keep])
        assert rel(got, want) < TOL
    assert filled.any()
    assert rel(grouped, rowwise) < 1e-4
