"""GPU parity of the readout placement (read_kspace_with_mask,
io/ismrmrd/reader.rs:225-468, with the strategies' preprocess closure,
recon/strategy.rs:153-161: whitening, navigator phase correction,
readout-oversampling removal) against the oracle's restatement.

Bars: without preprocessing the placement moves samples and must be
bit-exact, mask and placed count included, also where later readouts
overwrite earlier ones; with whitening or oversampling removal the FP32
arithmetic meets rel-L2 <= 1e-5 against the FP64 restatement (and the FP32
one, plus its own distance to FP64) while the mask and count stay exact.
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import DeviceArray, OksError, lib
from openkspace_amd.device import place_acquisitions
from tests.gpu_util import assert_parity

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    from openkspace_amd._lib import Context
    c = Context(0)
    yield c
    c.close()


def _run(ctx, a, d, e, W=None, os=False, pc=False):
    dd = DeviceArray.from_host(ctx, d)
    dw = DeviceArray.from_host(ctx, W) if W is not None else None
    try:
        grid, mask, n = place_acquisitions(ctx, a, dd, e, W=dw, remove_oversampling=os, phase_correct=pc)
        ctx.synchronize()
        g = grid.download()
        grid.free()
        return g, mask, n
    finally:
        dd.free()
        if dw is not None:
            dw.free()


def _acq(ns, nc, ky, cs=0, dpre=0, flags=0, off=0):
    return abi.acq_records(1, flags=flags, number_of_samples=ns, active_channels=nc, discard_pre=dpre,
                           center_sample=cs, kspace_encode_step_1=ky, data_offset=off)


@pytest.mark.parametrize("kw", [dict(), dict(reverse_every=3, shuffle_seed=1, noise_scans=5),
                                dict(discard=(3, 2), reverse_every=2, shuffle_seed=4)])
def test_placement_exact(ctx, kw):
    case = synth.make_case(nc=5, nz=6, ny=40, nx=36, kind="grappa", r=3, acs=10, seed=3, ellipses=3, three_d=True)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], **kw)
    g, m, n = _run(ctx, a, d, e)
    rg, rm, rn = O.place(a, d, e)
    assert n == rn and np.array_equal(m, rm) and np.array_equal(g, rg)


def test_multislice_exact(ctx):
    case = synth.make_case(nc=4, nz=5, ny=24, nx=20, kind="full", seed=4, ellipses=2)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], slice_axis=True, shuffle_seed=7, reverse_every=4)
    g, m, n = _run(ctx, a, d, e)
    assert n == 5 * 24 and np.array_equal(g, case["kspace"]) and m.all()


def test_overlapping_writers_keep_order(ctx):
    # line ky 4 of a 32-wide grid (centre 16): A [14, 18); B [15, 23) has its
    # centre 19 free, so it is placed and overwrites A's 15..17; C, D and E
    # find their centres filled and are skipped; F sits on another line
    vals = [np.full(16, v + 1j * v, np.complex64) for v in range(1, 6)]
    data = np.concatenate(vals)
    a = np.concatenate([
        _acq(4, 1, 4, cs=2, off=0),
        _acq(8, 1, 4, cs=1, off=16),
        _acq(8, 1, 4, cs=4, off=32),
        _acq(6, 1, 4, cs=5, off=48),
        _acq(16, 1, 4, cs=0, dpre=12, off=64),
        _acq(10, 1, 5, cs=5, off=0),
    ])
    e = (32, 8, 1, 0, 4, 0)
    g, m, n = _run(ctx, a, data, e)
    rg, rm, rn = O.place(a, data, e)
    assert n == rn == 3 and np.array_equal(m, rm) and np.array_equal(g, rg)
    # a chain of ten: readout i covers [1 + 3i, 7 + 3i) of a 64-wide line, so
    # each overlaps the one before and finds its own centre free: ten
    # launch generations, the later samples winning every overlap
    chain = np.concatenate([_acq(6, 1, 2, cs=31 - 3 * i, off=16 * (i % 5)) for i in range(10)])
    e2 = (64, 4, 1, 0, 2, 0)
    g2, m2, n2 = _run(ctx, chain, data, e2)
    r2 = O.place(chain, data, e2)
    assert n2 == r2[2] == 10 and np.array_equal(m2, r2[1]) and np.array_equal(g2, r2[0])


@pytest.mark.parametrize("nc", [4, 16, 32])
def test_whitening_fused(ctx, nc):
    case = synth.make_case(nc=nc, nz=3, ny=32, nx=48, kind="grappa", r=2, acs=8, seed=nc, ellipses=3, three_d=True)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], reverse_every=2, shuffle_seed=2)
    rng = np.random.default_rng(nc)
    W = np.tril(rng.standard_normal((nc, nc)) + 1j * rng.standard_normal((nc, nc))).astype(np.complex64) / nc
    g, m, n = _run(ctx, a, d, e, W=W)
    r32, rm, rn = O.place(a, d, e, W=W, prec="f32")
    r64 = O.place(a, d, e, W=W, prec="f64")[0]
    assert n == rn and np.array_equal(m, rm)
    assert_parity(g, r32, r64, TOL, f"place whiten nc={nc}")


@pytest.mark.parametrize("enc,rec", [(640, 320), (256, 128), (96, 12), (384, 128)])
def test_oversampling_fused(ctx, enc, rec):
    case = synth.make_case(nc=6, nz=2, ny=20, nx=enc, kind="grappa", r=2, acs=6, seed=enc, ellipses=3, three_d=True)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], recon_x=rec, reverse_every=3, shuffle_seed=1)
    g, m, n = _run(ctx, a, d, e, os=True)
    r32, rm, rn = O.place(a, d, e, remove_os=True, prec="f32")
    r64 = O.place(a, d, e, remove_os=True, prec="f64")[0]
    assert n == rn and np.array_equal(m, rm)
    assert_parity(g, r32, r64, TOL, f"place os {enc}->{rec}")


def test_mixed_acquisitions(ctx):
    # channel counts other than W's are not whitened; lines whose sample count
    # is not encoded_x keep their oversampling (recon/prewhiten.rs:118-142,
    # recon/oversampling.rs:103-154)
    nc, enc, rec = 8, 128, 64
    case = synth.make_case(nc=nc, nz=2, ny=16, nx=enc, kind="full", seed=5, ellipses=3, three_d=True)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], recon_x=rec, shuffle_seed=3, reverse_every=2)
    short = synth.make_case(nc=nc, nz=2, ny=16, nx=100, kind="full", seed=6, ellipses=2, three_d=True)
    a2, d2, _ = synth.acquisitions(short["kspace"], short["mask"])
    a2["data_offset"] += d.size
    a2["kspace_encode_step_1"] = (a2["kspace_encode_step_1"] + 3) % 16
    few = a2[::3].copy()
    few["active_channels"] = 5
    # the short readouts come first and win their rows' centres; the long
    # ones fill the remaining rows
    acqs = np.concatenate([few, a2[1::3], a])
    data = np.concatenate([d, d2])
    rng = np.random.default_rng(11)
    W = np.tril(rng.standard_normal((nc, nc)) + 1j * rng.standard_normal((nc, nc))).astype(np.complex64) / nc
    g, m, n = _run(ctx, acqs, data, e, W=W, os=True)
    r32, rm, rn = O.place(acqs, data, e, W=W, remove_os=True, prec="f32")
    r64 = O.place(acqs, data, e, W=W, remove_os=True, prec="f64")[0]
    assert n == rn == 32 and np.array_equal(m, rm)
    assert_parity(g, r32, r64, TOL, "place mixed")


def test_c1_shaped_whiten_and_oversampling(ctx):
    # the brain-volume readout geometry: 16 coils, 640 encoded -> 320 samples
    nc, nz, ny, enc, rec = 16, 2, 368, 640, 320
    case = synth.make_case(nc=nc, nz=nz, ny=ny, nx=enc, kind="full", seed=21, ellipses=4, three_d=True)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], recon_x=rec, shuffle_seed=5, noise_scans=8)
    W = O.prewhiten_calibrate(case["noise"], "f32")
    g, m, n = _run(ctx, a, d, e, W=W, os=True)
    r32, rm, rn = O.place(a, d, e, W=W, remove_os=True, prec="f32")
    r64 = O.place(a, d, e, W=W, remove_os=True, prec="f64")[0]
    assert n == rn == nz * ny and np.array_equal(m, rm)
    assert_parity(g, r32, r64, TOL, "place c1-shaped")


@pytest.mark.parametrize("kw", [dict(segments=3), dict(segments=2, reverse_every=2, discard=(2, 1), shuffle_seed=1),
                                dict(segments=4, nav_fallback=True, noise_scans=3)])
def test_phase_correction(ctx, kw):
    case = synth.make_case(nc=5, nz=4, ny=32, nx=40, kind="grappa", r=2, acs=8, seed=7, ellipses=3, three_d=True)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], **kw)
    g, m, n = _run(ctx, a, d, e, pc=True)
    r32, rm, rn = O.place(a, d, e, prec="f32", phase_correct=True)
    r64 = O.place(a, d, e, prec="f64", phase_correct=True)[0]
    assert n == rn and np.array_equal(m, rm)
    assert_parity(g, r32, r64, TOL, f"place phasecorr {kw}")


def test_phase_correction_without_navigators_is_exact_copy(ctx):
    case = synth.make_case(nc=3, nz=2, ny=16, nx=12, kind="full", seed=2, ellipses=2, three_d=True)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], reverse_every=2)
    g, m, n = _run(ctx, a, d, e, pc=True)
    assert np.array_equal(g, O.place(a, d, e)[0])


def test_c1_geometry_navigators_whitening_oversampling(ctx):
    # multi-slice 2-D, 15 coils, 640 encoded -> 320, 3-echo trains with
    # navigators per (slice, echo), reversed even echoes, noise scans
    nc, nz, ny, enc, rec = 15, 2, 368, 640, 320
    case = synth.make_case(nc=nc, nz=nz, ny=ny, nx=enc, kind="full", seed=23, ellipses=4)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], slice_axis=True, recon_x=rec, segments=3,
                                 reverse_every=2, noise_scans=4)
    W = O.prewhiten_calibrate(case["noise"], "f32")
    g, m, n = _run(ctx, a, d, e, W=W, os=True, pc=True)
    r32, rm, rn = O.place(a, d, e, W=W, remove_os=True, prec="f32", phase_correct=True)
    r64 = O.place(a, d, e, W=W, remove_os=True, prec="f64", phase_correct=True)[0]
    assert n == rn == nz * ny and np.array_equal(m, rm)
    assert_parity(g, r32, r64, TOL, "place c1 navigators")


def test_errors(ctx):
    data = np.zeros(64, np.complex64)
    dd = DeviceArray.from_host(ctx, data)
    try:
        bad = _acq(16, 2, 1, cs=8, off=40)  # 32 samples from offset 40 overrun 64
        with pytest.raises(OksError) as ei:
            place_acquisitions(ctx, bad, dd, (16, 4, 1, 0, 2, 0))
        assert ei.value.code == abi.OKS_E_INCONSISTENT
        noise = _acq(16, 1, 0, flags=abi.ACQ_IS_NOISE_MEASUREMENT)
        with pytest.raises(ValueError):
            place_acquisitions(ctx, noise, dd, (16, 4, 1, 0, 2, 0))
        nav = abi.acq_records(1, flags=abi.ACQ_IS_PHASECORR_DATA, number_of_samples=16, active_channels=4,
                              data_offset=10)  # 64 samples from offset 10 overrun 64
        with pytest.raises(OksError) as ei:
            place_acquisitions(ctx, np.concatenate([nav, _acq(16, 1, 1, cs=8)]), dd, (16, 4, 1, 0, 2, 0),
                               phase_correct=True)
        assert ei.value.code == abi.OKS_E_INCONSISTENT
        W = DeviceArray(ctx, (33, 33), np.complex64)
        try:
            with pytest.raises(OksError) as ei:
                place_acquisitions(ctx, _acq(16, 1, 1), dd, (16, 4, 1, 0, 2, 0), W=W)
            assert ei.value.code == abi.OKS_E_UNSUPPORTED
        finally:
            W.free()
    finally:
        dd.free()


def test_placement_then_ifft_rss_end_to_end(ctx):
    # the image pass of a file on the GPU: navigators, noise scans, reversed
    # echoes and 2x oversampled readouts placed with the closure fused, then
    # ifft-rss on the placed grid with the closure off (it has run), against
    # the oracle doing the same two steps
    from tests.gpu_util import recon_dev
    nc, nz, ny, enc, rec = 6, 2, 48, 128, 64
    case = synth.make_case(nc=nc, nz=nz, ny=ny, nx=enc, kind="full", seed=31, ellipses=4)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], slice_axis=True, recon_x=rec, segments=2,
                                 reverse_every=2, noise_scans=3)
    W = O.prewhiten_calibrate(case["noise"], "f64").astype(np.complex64)
    g, m, n = _run(ctx, a, d, e, W=W, os=True, pc=True)
    hdr = dict(encoded_x=enc, encoded_y=ny, encoded_z=1, recon_x=rec, recon_y=ny, recon_z=1, is_3d_encoding=0)
    desc = abi.IfftRssDesc.default(remove_oversampling=0, prewhiten=0, partial_fourier=0)
    got, _, _ = recon_dev(ctx, "oks_recon_ifft_rss_dev", desc,
                          dict(kspace=g, mask=m.astype(np.uint8), noise=None, hdr=hdr))
    refs = []
    for prec in ("f32", "f64"):
        rg, rm, _ = O.place(a, d, e, W=W, remove_os=True, prec=prec, phase_correct=True)
        # the strategy's input tensor is complex64 (the FP64 twin's placed grid
        # is rounded to it, 1e-7, before its FP64 reconstruction)
        st, img, _, _ = O.recon("ifft_rss", desc, dict(kspace=np.ascontiguousarray(rg.astype(np.complex64)),
                                                       mask=rm.astype(np.uint8), noise=None, hdr=hdr), prec)
        assert st == 0
        refs.append(img)
    assert got.shape == refs[0].shape == (nz, ny, rec)
    assert_parity(got, refs[0], refs[1], TOL, "placement + ifft-rss")
