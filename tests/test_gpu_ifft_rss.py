"""GPU parity of the centred FFT passes and the ifft-rss strategy against the oracle.

Bar (FP32 path): rel-L2 <= 1e-5 against the FP64 twin, and <= 1e-5 plus the
FP32 restatement's own distance to FP64 against the FP32 restatement (the
backend's noise whitener is FP64, the restatement's FP32).  Without
pre-whitening the plain 1e-5 bar against the FP32 restatement applies.
Index windows (crop shapes) and fallback flags are exact.
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from tests.gpu_util import assert_parity, fft_centered, recon_dev, rel

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ["ifft2", "ifft2odd", "ifft3", "ifft1", "ifft2big"])
def test_centred_ifft_golden(ctx, golden, name):
    x = golden[f"{name}_in"]
    axes = tuple(int(a) for a in golden[f"{name}_axes"])
    got = fft_centered(ctx, x, axes, inverse=True, norm=1)
    err = rel(got, golden[f"{name}_out"])
    print(f"{name}: rel vs numpy FP64 = {err:.2e}")
    assert err < 2e-6


@pytest.mark.parametrize("n", [8, 16, 23, 46, 160, 256, 320, 368, 384, 640])
@pytest.mark.parametrize("inverse", [True, False])
def test_centred_1d_every_radix(ctx, n, inverse):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    got = fft_centered(ctx, x, (1,), inverse=inverse, norm=0)
    f = np.fft.ifft if inverse else np.fft.fft
    want = np.roll(f(np.roll(x.astype(np.complex128), -((n + 1) // 2), axis=1), axis=1), -(n // 2), axis=1)
    if inverse:
        want = want * n
    err = rel(got, want)
    assert err < 2e-6, err


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_strided_axis(ctx, axis):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((5, 24, 40)) + 1j * rng.standard_normal((5, 24, 40))).astype(np.complex64)
    got = fft_centered(ctx, x, (axis,), inverse=True, norm=1)
    want = O.ifft_centered(x, (axis,), "f64")
    assert rel(got, want) < 2e-6


def _case(**kw):
    p = dict(nc=6, nz=3, ny=48, nx=64, os_x=2, kind="full", seed=21, ellipses=6)
    p.update(kw)
    return synth.make_case(**p)


@pytest.mark.parametrize("remove_os", [0, 1])
@pytest.mark.parametrize("prewhiten", [0, 1])
def test_ifft_rss_small(ctx, remove_os, prewhiten):
    case = _case()
    d = abi.IfftRssDesc.default(remove_oversampling=remove_os, prewhiten=prewhiten)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and got.shape == ref32.shape
    assert flags == rflags
    e32, _ = assert_parity(got, ref32, ref64, TOL32, f"os={remove_os} pw={prewhiten}")
    if not prewhiten:
        assert e32 < TOL32


def test_ifft_rss_3d_and_zcrop(ctx):
    case = synth.make_case(nc=4, nz=10, ny=32, nx=40, os_x=2, kind="full", three_d=True, seed=5, ellipses=4)
    case["hdr"]["recon_z"] = 6
    d = abi.IfftRssDesc.default()
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and got.shape == ref32.shape == (6, 32, 20)
    assert_parity(got, ref32, ref64, TOL32, "3d")


def test_ifft_rss_small_even_grid_goes_homodyne(ctx):
    # a fully sampled ny = 16 has 8 rows below DC and 7 above: the reference's
    # gate routes it to homodyne (recon/partial_fourier.rs:84-108)
    case = _case(ny=16)
    d = abi.IfftRssDesc.default()
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and flags == rflags and flags & abi.F_HOMODYNE
    assert_parity(got, ref32, ref64, TOL32, "homodyne ny=16")


def test_ifft_rss_prewhiten_skipped(ctx):
    case = _case()
    case["noise"] = np.ascontiguousarray(case["noise"][:, :3])
    d = abi.IfftRssDesc.default()
    _, ref, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert flags & abi.F_PREWHITEN_SKIPPED and rflags & abi.F_PREWHITEN_SKIPPED
    assert rel(got, ref) < TOL32


@pytest.mark.parametrize("remove_os", [0, 1])
def test_ifft_rss_c1_full_size(ctx, remove_os):
    case = synth.make_config("c1_ifft_rss")
    d = abi.IfftRssDesc.default(remove_oversampling=remove_os)
    O.set_threads(8)
    st, ref32, _, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    O.set_threads(1)
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert got.shape == (16, 368, 320)
    assert_parity(got, ref32, ref64, TOL32, f"C1 os={remove_os}")


def test_host_buffer_entry_point(ctx):
    from openkspace_amd.strategy import IfftRss
    case = _case()
    vol = IfftRss(ctx=ctx, remove_oversampling=0).reconstruct(case)
    d = abi.IfftRssDesc.default(remove_oversampling=0)
    _, ref, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    assert vol.strategy == "ifft-rss" and vol.flags == rflags
    assert_parity(vol.data, ref, ref64, TOL32, "host entry")


@pytest.mark.parametrize("rows_pipe,cols_pipe", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("nc,nx,prewhiten", [(32, 384, 1), (17, 256, 1), (5, 128, 0), (15, 640, 1)])
def test_ifft_rss_pass_forms(ctx, rows_pipe, cols_pipe, nc, nx, prewhiten):
    # both row-pass forms (whitening before or after the kx IFFT; persistent
    # pipelined or one row per workgroup) and both column-pass forms meet the
    # same bar; nc = 17 and 5 leave the coil loops and thread groups ragged
    case = _case(nc=nc, nz=2, ny=40, nx=nx, os_x=2)
    d = abi.IfftRssDesc.default(remove_oversampling=0, prewhiten=prewhiten)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    ctx.set_option("rows_pipe", rows_pipe)
    ctx.set_option("cols_pipe", cols_pipe)
    try:
        got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    finally:
        ctx.set_option("rows_pipe", 1)
        ctx.set_option("cols_pipe", 1)
    assert st == 0 and flags == rflags and got.shape == ref32.shape
    assert_parity(got, ref32, ref64, TOL32, f"rows_pipe={rows_pipe} cols_pipe={cols_pipe} nc={nc} nx={nx}")


@pytest.mark.parametrize("nc,nx,remove_os", [(15, 640, 0), (15, 640, 1), (5, 256, 1), (20, 320, 0), (9, 368, 0)])
def test_ifft_rss_row_pass_widths(ctx, nc, nx, remove_os):
    # the wavefronts per workgroup of the one-row-per-workgroup row passes
    # (automatic, or 8 .. 16) change the launch only: every width gives the
    # same bits, and those meet the parity bar
    case = _case(nc=nc, nz=2, ny=24, nx=nx, os_x=2)
    d = abi.IfftRssDesc.default(remove_oversampling=remove_os)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    ctx.set_option("rows_pipe", 0)
    outs = {}
    try:
        for w in (0, 8, 11, 16):
            ctx.set_option("rows_waves", w)
            outs[w], flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
            assert flags == rflags
    finally:
        ctx.set_option("rows_waves", 0)
        ctx.set_option("rows_pipe", 1)
    for w, got in outs.items():
        assert np.array_equal(got, outs[8]), f"rows_waves={w}"
    assert st == 0
    assert_parity(outs[0], ref32, ref64, TOL32, f"nc={nc} nx={nx} os={remove_os}")


def test_rows_waves_option_range(ctx):
    for bad in (-1, 1, 7, 17):
        with pytest.raises(Exception):
            ctx.set_option("rows_waves", bad)


def test_ifft_rss_column_pass_shapes(ctx):
    # columns per workgroup and workgroup size of the column passes change the
    # launch only: the same bits for every shape
    case = _case(nc=7, nz=2, ny=40, nx=128, os_x=2)
    d = abi.IfftRssDesc.default(remove_oversampling=0)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    outs = {}
    try:
        for sx, th, pipe in [(0, 0, 1), (0, 0, 0), (8, 256, 1), (4, 256, 0), (2, 512, 1), (16, 256, 0)]:
            ctx.set_option("cols_sx", sx)
            ctx.set_option("cols_threads", th)
            ctx.set_option("cols_pipe", pipe)
            outs[(sx, th, pipe)], flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
            assert flags == rflags
    finally:
        ctx.set_option("cols_sx", 0)
        ctx.set_option("cols_threads", 0)
        ctx.set_option("cols_pipe", 1)
    base = outs[(0, 0, 1)]
    for k, got in outs.items():
        assert np.array_equal(got, base), f"cols shape {k}"
    assert_parity(base, ref32, ref64, TOL32, "column shapes")
    for name, bad in (("cols_sx", 3), ("cols_sx", 32), ("cols_threads", 128)):
        with pytest.raises(Exception):
            ctx.set_option(name, bad)


@pytest.mark.parametrize("ny", [64, 160, 256, 320, 368, 384, 512, 640])
def test_rss_column_pass_register_form(ctx, ny):
    # The RSS column pass with register four-step transforms (default for
    # these lengths) and the LDS Stockham form (option cols_reg = 0) both meet
    # the bar and agree to FP32 rounding; 7 coils over the RSS coil groups.
    case = _case(nc=7, nz=3, ny=ny, nx=48, os_x=2)
    d = abi.IfftRssDesc.default(remove_oversampling=0)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    outs = []
    try:
        for form in (1, 0):
            ctx.set_option("cols_reg", form)
            got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
            assert flags == rflags and got.shape == ref32.shape
            assert_parity(got, ref32, ref64, TOL32, f"ny={ny} cols_reg={form}")
            outs.append(got)
    finally:
        ctx.set_option("cols_reg", 1)
    assert rel(outs[0], outs[1]) < 2e-6


@pytest.mark.parametrize("three_d,mode", [(True, abi.FFT_2D), (False, abi.FFT_3D), (True, abi.FFT_3D),
                                          (False, abi.FFT_2D)])
def test_forced_fft_mode(ctx, three_d, mode):
    # FftMode::TwoD / ThreeD override the header's 3-D flag
    # (recon/strategy.rs:168-172): a 3-D acquisition forced to 2-D is
    # transformed slice by slice, a multi-slice one forced to 3-D gets the kz
    # transform; both against the oracle, and different from Auto where the
    # forced mode differs from the header's
    case = synth.make_case(nc=4, nz=6, ny=32, nx=40, os_x=2, kind="full", three_d=three_d, seed=12, ellipses=4)
    d = abi.IfftRssDesc.default(fft_mode=mode, partial_fourier=0)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and flags == rflags and got.shape == ref32.shape
    assert_parity(got, ref32, ref64, TOL32, f"fft_mode={mode} 3d={three_d}")
    auto, _, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", abi.IfftRssDesc.default(partial_fourier=0), case)
    if (mode == abi.FFT_3D) != three_d:
        assert rel(got, auto) > 1e-3
    else:
        assert np.array_equal(got, auto)


@pytest.mark.parametrize("rows_pipe", [1, 0])
@pytest.mark.parametrize("nc,nx,prewhiten", [(20, 64, 1), (5, 40, 1), (6, 128, 0)])
def test_ifft_rss_3d_zero_rows(ctx, rows_pipe, nc, nx, prewhiten):
    # 3-D tail with zero (z, y) rows: whole ky lines never acquired (as in a
    # ky-undersampled volume), single rows zero in some partitions only, and
    # a line zero in every partition but one.  The pipelined row pass flags
    # all-zero rows and the kz/ky passes skip them; the image equals the
    # oracle's.  nc = 20 and 5 take the 32- and 16-channel kernels; nx = 40
    # (runtime transform length) and rows_pipe 0 run the one-row-per-workgroup
    # kernel, which flags nothing, so nothing is skipped.
    case = synth.make_case(nc=nc, nz=6, ny=32, nx=nx, os_x=1, kind="full", three_d=True, seed=11, ellipses=4)
    k, m = case["kspace"], case["mask"]
    rng = np.random.default_rng(3)
    dead_y = [y for y in range(32) if y % 3 == 1 and not 12 <= y < 20]
    k[:, :, dead_y, :] = 0
    m[:, dead_y, :] = 0
    for z, y in zip(rng.integers(0, 6, 10), rng.integers(0, 32, 10)):
        k[:, z, y, :] = 0
        m[z, y, :] = 0
    k[:, 1:, 2, :] = 0
    m[1:, 2, :] = 0
    k[3, 2, 5, 7] = 0.5 + 0.25j  # one nonzero sample in one coil keeps the row live
    m[2, 5, 7] = 1
    d = abi.IfftRssDesc.default(remove_oversampling=0, prewhiten=prewhiten)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    ctx.set_option("rows_pipe", rows_pipe)
    try:
        got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    finally:
        ctx.set_option("rows_pipe", 1)
    assert st == 0 and flags == rflags and got.shape == ref32.shape
    assert_parity(got, ref32, ref64, TOL32, f"3-D zero rows nc={nc} nx={nx} pipe={rows_pipe}")


def test_ifft_rss_3d_all_zero(ctx):
    case = synth.make_case(nc=4, nz=4, ny=16, nx=32, os_x=1, kind="full", three_d=True, seed=2, ellipses=2)
    case["kspace"][:] = 0
    d = abi.IfftRssDesc.default(remove_oversampling=0, prewhiten=0, partial_fourier=0)
    got, _, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert got.shape == (4, 16, 32) and not np.any(got)


@pytest.mark.parametrize("n", [64, 128, 160, 256, 320, 384, 512])
@pytest.mark.parametrize("inner", [20, 48])
def test_strided_inverse_register_form(ctx, n, inner):
    # the inverse transform along a middle axis runs the register four-step
    # complex-output column pass at these lengths; inner = 20 leaves the last
    # 16-column strip ragged.  The LDS Stockham form (cols_reg = 0) meets the
    # same bar.
    rng = np.random.default_rng(n + inner)
    x = (rng.standard_normal((3, n, inner)) + 1j * rng.standard_normal((3, n, inner))).astype(np.complex64)
    want = O.ifft_centered(x, (1,), "f64")
    got = fft_centered(ctx, x, (1,), inverse=True, norm=1)
    assert rel(got, want) < 2e-6
    ctx.set_option("cols_reg", 0)
    try:
        lds = fft_centered(ctx, x, (1,), inverse=True, norm=1)
    finally:
        ctx.set_option("cols_reg", 1)
    assert rel(lds, want) < 2e-6


@pytest.mark.parametrize("nz", [6, 64])
def test_ifft_rss_3d_zero_rows_ragged_crop(ctx, nz):
    # the x crop leaves 100 columns, so column strips of 16 straddle ky lines
    # in the kz pass (flags read per column) and the last strip of the ky
    # pass is ragged; nz = 64 runs the register kz pass, nz = 6 the LDS one
    case = synth.make_case(nc=6, nz=nz, ny=32, nx=128, os_x=1, kind="full", three_d=True, seed=nz, ellipses=4)
    case["hdr"]["recon_x"] = 100
    k, m = case["kspace"], case["mask"]
    dead_y = [y for y in range(32) if y % 3 != 0 and not 12 <= y < 20]
    k[:, :, dead_y, :] = 0
    m[:, dead_y, :] = 0
    rng = np.random.default_rng(nz)
    for z, y in zip(rng.integers(0, nz, 8), rng.integers(0, 32, 8)):
        k[:, z, y, :] = 0
        m[z, y, :] = 0
    d = abi.IfftRssDesc.default(remove_oversampling=0)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and flags == rflags and got.shape == ref32.shape == (nz, 32, 100)
    assert_parity(got, ref32, ref64, TOL32, f"3-D zero rows, crop to 100 columns, nz={nz}")


@pytest.mark.parametrize("nc", [33, 48, 64])
@pytest.mark.parametrize("three_d,remove_os", [(False, 0), (False, 1), (True, 0), (True, 1)])
def test_ifft_rss_more_than_32_channels(ctx, nc, three_d, remove_os):
    # Above the 32 channels the fused row pass holds, the whitening runs as a
    # pass of its own and the row pass per group of at most 32 channels.
    case = _case(nc=nc, nz=4 if three_d else 3, ny=40, nx=64, three_d=three_d, seed=nc)
    d = abi.IfftRssDesc.default(remove_oversampling=remove_os, prewhiten=1)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and got.shape == ref32.shape and flags == rflags
    assert_parity(got, ref32, ref64, TOL32, f"nc={nc} 3d={three_d} os={remove_os}")


@pytest.mark.parametrize("nc,remove_os", [(64, 0), (64, 1), (33, 0)])
def test_ifft_rss_many_channels_long_readout(ctx, nc, remove_os):
    # a row of 32 channels x 640 samples does not fit in LDS: the channel
    # groups shrink until one does (22 + 21 + 21 at 64 channels)
    case = _case(nc=nc, nz=2, ny=32, nx=640, seed=nc + 640)
    d = abi.IfftRssDesc.default(remove_oversampling=remove_os, prewhiten=1)
    st, ref32, rflags, _ = O.recon("ifft_rss", d, case, "f32")
    _, ref64, _, _ = O.recon("ifft_rss", d, case, "f64")
    got, flags, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", d, case)
    assert st == 0 and got.shape == ref32.shape and flags == rflags
    assert_parity(got, ref32, ref64, TOL32, f"nc={nc} nx=640 os={remove_os}")
