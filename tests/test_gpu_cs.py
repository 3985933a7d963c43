"""GPU parity of FISTA L1-wavelet CS (fista_cs_single_coil, CsRss).

Bar (iterative FP32 path): the backend's rel-L2 to the FP64 twin is at most
the FP32 restatement's own distance to FP64 plus 1e-5, and its distance to the
FP32 restatement at most 1e-5 plus twice that (FISTA accumulates FP32 rounding
over the iterations, in the restatement as on the GPU).  The
HIP-graph replay of the iteration loop is checked bit-for-bit against the
eager launch sequence on every replay.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import DeviceArray, OksError, check, lib
from tests.gpu_util import assert_parity, recon_dev, rel

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def gpu_fista(ctx, kzf, mask, iters, lam):
    kzf = np.ascontiguousarray(kzf, dtype=np.complex64)
    if kzf.ndim == 2:
        kzf = kzf[None]
    planes, ny, nx = kzf.shape
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    mp = 1 if mask.ndim == 2 else mask.shape[0]
    d_k = DeviceArray.from_host(ctx, kzf)
    d_m = DeviceArray.from_host(ctx, mask)
    d_o = DeviceArray(ctx, kzf.shape, np.complex64)
    check(ctx.ptr, lib().oks_fista_cs_dev(ctx.ptr, d_k.addr, d_m.addr, planes, mp, ny, nx, iters, C.c_float(lam),
                                          d_o.addr))
    ctx.synchronize()
    out = d_o.download()
    for d in (d_k, d_m, d_o):
        d.free()
    return out


def test_fista_reference_phantom(ctx):  # recon/cs.rs:259 on the GPU
    ny = nx = 16
    truth = np.zeros((ny, nx), np.complex64)
    truth[5, 4], truth[10, 11], truth[3, 12], truth[12, 3] = 1, 0.8, 0.6, 0.5
    k = O.centred_fft2(truth, False)
    rows = np.array([(i % 2 == 0) or (ny // 2 - 2 <= i < ny // 2 + 2) for i in range(ny)])
    mask = np.repeat(rows[:, None], nx, 1).astype(np.uint8)
    kzf = k * mask
    zf_err = np.linalg.norm(O.centred_fft2(kzf, True) - truth)
    got = gpu_fista(ctx, kzf, mask, 200, 0.02)[0]
    assert np.linalg.norm(got - truth) < 0.8 * zf_err
    _, r32 = O.fista_cs(kzf, mask, 200, 0.02, "f32")
    _, r64 = O.fista_cs(kzf, mask, 200, 0.02, "f64")
    assert_parity(got, r32, r64, 1e-5, "fista 16x16 200it", iterative=True)


@pytest.mark.parametrize("ny,nx,iters", [(32, 32, 40), (64, 48, 30), (320, 320, 120)])
def test_fista_vs_oracle(ctx, ny, nx, iters):
    case = synth.make_case(nc=1, nz=1, ny=ny, nx=nx, kind="cs", seed=ny + nx, ellipses=6)
    kzf, mask = case["kspace"][0, 0], case["mask"][0]
    got = gpu_fista(ctx, kzf, mask, iters, 0.01)[0]
    _, r32 = O.fista_cs(kzf, mask, iters, 0.01, "f32")
    _, r64 = O.fista_cs(kzf, mask, iters, 0.01, "f64")
    assert_parity(got, r32, r64, TOL, f"fista {ny}x{nx} {iters}it", iterative=True)


def test_fista_closed_forms(ctx):
    rng = np.random.default_rng(4)
    k = (rng.standard_normal((12, 16)) + 1j * rng.standard_normal((12, 16))).astype(np.complex64)
    m = (rng.random((12, 16)) < 0.4).astype(np.uint8)
    # lambda = 0: the zero-filled image is a fixed point
    got = gpu_fista(ctx, k * m, m, 5, 0.0)[0]
    assert rel(got, O.centred_fft2(k * m, True, "f64")) < 1e-6
    # full mask: one prox of the zero-filled image
    full = np.ones_like(m)
    got = gpu_fista(ctx, k, full, 4, 0.3)[0]
    want = O.haar(O.soft_threshold(O.haar(O.centred_fft2(k, True, "f64"), False, "f64"), 0.3, "f64"), True, "f64")
    assert rel(got, want) < 1e-6


def test_fista_odd_rejected(ctx):
    with pytest.raises(OksError) as e:
        gpu_fista(ctx, np.zeros((10, 15), np.complex64), np.ones((10, 15), np.uint8), 3, 0.01)
    assert e.value.code == abi.OKS_E_CS_ODD


def test_graph_replays_match_eager(ctx):
    case = synth.make_case(nc=3, nz=2, ny=64, nx=64, kind="cs", seed=9, ellipses=5)
    d = abi.CsDesc.default(iters=20)
    ctx.set_option("graphs", 0)
    eager, _, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    ctx.set_option("graphs", 1)
    for rep in range(4):
        got, _, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
        assert np.array_equal(got, eager), f"replay {rep} differs"
        ctx.synchronize()


def _cs_case(**kw):
    p = dict(nc=4, nz=2, ny=64, nx=64, kind="cs", seed=17, ellipses=6)
    p.update(kw)
    return synth.make_case(**p)


@pytest.mark.parametrize("prewhiten", [0, 1])
def test_cs_strategy_small(ctx, prewhiten):
    case = _cs_case()
    d = abi.CsDesc.default(iters=30, prewhiten=prewhiten)
    st, r32, rflags, _ = O.recon("cs", d, case, "f32")
    _, r64, _, _ = O.recon("cs", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    assert st == 0 and got.shape == r32.shape and flags == rflags
    assert_parity(got, r32, r64, TOL, f"cs pw={prewhiten}", iterative=True)


@pytest.mark.parametrize("nc,ny,os_x", [(40, 64, 1), (64, 64, 2), (48, 33, 2)])
def test_cs_more_than_32_channels(ctx, nc, ny, os_x):
    # CS solves every channel's planes on their own; above 32 channels only
    # the whitening (and the odd-ky fallback's row pass) change form
    case = _cs_case(nc=nc, ny=ny, nx=32 * os_x, os_x=os_x, seed=nc)
    d = abi.CsDesc.default(iters=10, prewhiten=1)
    st, r32, rflags, _ = O.recon("cs", d, case, "f32")
    _, r64, _, _ = O.recon("cs", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    assert st == 0 and got.shape == r32.shape and flags == rflags
    assert_parity(got, r32, r64, TOL, f"cs nc={nc} ny={ny} os={os_x}", iterative=True)


@pytest.mark.parametrize("three_d,os_x", [(False, 2), (True, 1), (True, 2)])
def test_cs_fallback_fused(ctx, three_d, os_x):
    # odd ny: whitened, (oversampling-removed,) kz-transformed IFFT + RSS in one pass
    case = synth.make_case(nc=4, nz=4 if three_d else 2, ny=33, nx=20 * os_x, os_x=os_x, kind="cs",
                           three_d=three_d, seed=13, ellipses=4)
    d = abi.CsDesc.default(iters=5, prewhiten=1, remove_oversampling=1)
    st, r32, rflags, _ = O.recon("cs", d, case, "f32")
    _, r64, _, _ = O.recon("cs", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    assert st == 0 and flags == rflags and flags & abi.F_CS_FALLBACK
    assert_parity(got, r32, r64, TOL, f"cs fallback 3d={three_d} os={os_x}")


def test_cs_strategy_3d_and_odd_fallback(ctx):
    case = synth.make_case(nc=3, nz=6, ny=32, nx=32, kind="cs", three_d=True, seed=3, ellipses=4)
    d = abi.CsDesc.default(iters=15)
    st, r32, rflags, _ = O.recon("cs", d, case, "f32")
    _, r64, _, _ = O.recon("cs", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    assert flags == rflags
    assert_parity(got, r32, r64, TOL, "cs 3d", iterative=True)
    case = _cs_case(ny=30, nx=33, kind="full")
    st, r32, rflags, _ = O.recon("cs", d, case, "f32")
    got, flags, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    assert flags == rflags and flags & abi.F_CS_FALLBACK
    assert rel(got, r32) < 2e-5


def test_cs_c3_shape(ctx):
    # The C3 slice shape (8 coils, 320 x 320, R=6 variable-density mask with a
    # 24x24 centre, 120 iterations, lambda 0.01) on 2 of the 32 slices.
    case = synth.make_config("c3_cs", nz=2)
    d = abi.CsDesc.default(iters=120, lambda_=0.01)
    O.set_threads(16)
    st, r32, rflags, _ = O.recon("cs", d, case, "f32")
    _, r64, _, _ = O.recon("cs", d, case, "f64")
    O.set_threads(1)
    got, flags, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    assert st == 0 and flags == rflags and got.shape == (2, 320, 320)
    assert_parity(got, r32, r64, TOL, "C3 cs", iterative=True)


@pytest.mark.parametrize("iters", [1, 2, 3, 7])
def test_fista_iteration_count_parity(ctx, iters):
    # the iterates live in two buffers that swap roles every iteration; the
    # last iterate must reach the output for odd and even counts alike
    case = synth.make_case(nc=1, nz=1, ny=32, nx=48, kind="cs", seed=40 + iters, ellipses=5)
    kzf, mask = case["kspace"][0, 0], case["mask"][0]
    got = gpu_fista(ctx, kzf, mask, iters, 0.01)[0]
    _, r32 = O.fista_cs(kzf, mask, iters, 0.01, "f32")
    _, r64 = O.fista_cs(kzf, mask, iters, 0.01, "f64")
    assert_parity(got, r32, r64, TOL, f"fista {iters}it", iterative=True)


@pytest.mark.parametrize("iters", [6, 7])
def test_cs_chunked_planes_match(ctx, iters):
    # planes processed in chunks (each with its own buffer rotation) give the
    # same bits as one chunk
    case = _cs_case(nc=3, nz=3)
    d = abi.CsDesc.default(iters=iters)
    ctx.set_option("graphs", 0)
    try:
        whole, _, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
        ctx.set_option("cs_chunk", 2)
        chunked, _, _ = recon_dev(ctx, "oks_recon_cs_dev", d, case)
    finally:
        ctx.set_option("cs_chunk", 0)
        ctx.set_option("graphs", 1)
    assert np.array_equal(chunked, whole)


@pytest.mark.parametrize("ny,nx", [(256, 256), (320, 320), (384, 384), (368, 640), (640, 368), (160, 96)])
def test_launch_forms_bit_identical(ny, nx):
    # Two column-pass families: the register four-step transforms (default)
    # and the LDS Stockham transforms (cs_reg = 0).  Within a family the
    # persistent prefetching column pass, the LDS slot maps, the strip widths
    # and the workgroup sizes only move time: every form gives the family
    # default's bits.
    from openkspace_amd._lib import Context
    case = synth.make_case(nc=3, nz=2, ny=ny, nx=nx, kind="cs", seed=ny + nx, ellipses=4)
    d = abi.CsDesc.default(iters=7)
    families = [
        [{}, {"cs_pipe": 1}, {"cs_coop": 1}, {"cs_upd_map": 1}, {"cs_upd_map": 2}, {"cs_sx": 8},
         {"cs_threads": 1024}],
        [{"cs_reg": 0}, {"cs_reg": 0, "cs_coop": 1}, {"cs_reg": 0, "cs_pipe": 1},
         {"cs_reg": 0, "cs_pipe": 1, "cs_cols_map": 1}, {"cs_reg": 0, "cs_cols_map": 2, "cs_upd_map": 1},
         {"cs_reg": 0, "cs_upd_map": 2}, {"cs_reg": 0, "cs_sx": 8}, {"cs_reg": 0, "cs_threads": 1024}],
    ]
    for forms in families:
        imgs = []
        for opts in forms:
            c = Context(0)
            try:
                c.set_option("tuning", 0)
                for k, v in opts.items():
                    c.set_option(k, v)
                got, _, _ = recon_dev(c, "oks_recon_cs_dev", d, case)
                imgs.append(got)
            finally:
                c.close()
        for opts, img in zip(forms[1:], imgs[1:]):
            assert np.array_equal(img, imgs[0]), opts


@pytest.mark.parametrize("ny,nx", [(320, 320), (368, 640), (64, 48)])
def test_lds_column_pass_vs_oracle(ctx, ny, nx):
    # the LDS Stockham column pass (cs_reg = 0) keeps its own parity
    case = synth.make_case(nc=1, nz=1, ny=ny, nx=nx, kind="cs", seed=ny * 3 + nx, ellipses=6)
    kzf, mask = case["kspace"][0, 0], case["mask"][0]
    ctx.set_option("cs_reg", 0)
    try:
        got = gpu_fista(ctx, kzf, mask, 30, 0.01)[0]
    finally:
        ctx.set_option("cs_reg", 1)
    _, r32 = O.fista_cs(kzf, mask, 30, 0.01, "f32")
    _, r64 = O.fista_cs(kzf, mask, 30, 0.01, "f64")
    assert_parity(got, r32, r64, TOL, f"fista lds {ny}x{nx}", iterative=True)
