"""The oracle's read_kspace_with_mask restatement (io/ismrmrd/reader.rs:225-468),
pinned without the reference (which has no placement test of its own):

* round trips: readouts generated from a known grid (any order, reversed
  lines with their samples stored reversed, discard padding, interleaved
  noise scans, multi-slice 2-D) place back onto that grid and mask exactly;
* the index rules on hand-built cases: DC offsets and out-of-range lines,
  centre_sample 0 (kept samples centred), first writer wins at the readout
  centre, a later overlapping writer overwriting the overlap, fewer channels
  than the grid, the sizing pass;
* the preprocess closure: whitening placed = whitening applied to the placed
  grid, oversampling removal placed = the removed lines placed at the centre.
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth


def _case3d():
    return synth.make_case(nc=3, nz=4, ny=16, nx=12, kind="grappa", r=2, acs=6, seed=3, ellipses=3, three_d=True)


def _rows(case):
    return case["mask"].any(-1)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("kw", [dict(), dict(reverse_every=3, shuffle_seed=1),
                                dict(reverse_every=2, noise_scans=4, shuffle_seed=2)])
def test_round_trip(prec, kw):
    case = _case3d()
    k, rows = case["kspace"], _rows(case)
    a, d, e = synth.acquisitions(k, case["mask"], **kw)
    g, mk, n = O.place(a, d, e, prec=prec)
    assert n == rows.sum()
    assert np.array_equal(g, np.where(rows[None, :, :, None], k, 0))
    assert np.array_equal(mk, np.repeat(rows[:, :, None], k.shape[3], 2))


def test_discard_padding_widens_grid():
    case = _case3d()
    k, rows = case["kspace"], _rows(case)
    a, d, e = synth.acquisitions(k, case["mask"], discard=(3, 2), reverse_every=2, shuffle_seed=4)
    g, mk, _ = O.place(a, d, e)
    nxg = 12 + 5  # the sizing pass takes the raw sample count
    x0 = nxg // 2 - 6
    want = np.zeros((3, 4, 16, nxg), np.complex64)
    want[..., x0:x0 + 12] = np.where(rows[None, :, :, None], k, 0)
    assert np.array_equal(g, want)
    assert np.array_equal(mk[..., x0:x0 + 12], np.repeat(rows[:, :, None], 12, 2)) and not mk[..., :x0].any()


def test_multislice_uses_slice_axis():
    case = synth.make_case(nc=2, nz=3, ny=10, nx=8, kind="full", seed=4, ellipses=2)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], slice_axis=True, shuffle_seed=3)
    g, mk, n = O.place(a, d, e)
    assert g.shape == (2, 3, 10, 8) and n == 30
    assert np.array_equal(g, case["kspace"]) and mk.all()


def _acq(ns, nc, ky, kz=0, sl=0, cs=0, dpre=0, dpost=0, flags=0, off=0):
    return abi.acq_records(1, flags=flags, number_of_samples=ns, active_channels=nc, discard_pre=dpre,
                           discard_post=dpost, center_sample=cs, kspace_encode_step_1=ky, kspace_encode_step_2=kz,
                           slice=sl, data_offset=off)


def test_dc_offsets_and_out_of_range():
    case = _case3d()
    k = case["kspace"]
    a, d, e = synth.acquisitions(k, case["mask"])
    e2 = list(e)
    e2[4] = e[4] - 3  # ky_center three below the array centre: rows move up by three
    g, mk, n = O.place(a, d, e2)
    rows = _rows(case)
    assert n == rows[:, :13].sum()  # rows 13..15 would land outside [0, 16)
    assert np.array_equal(g[:, :, 3:], np.where(rows[None, :, :13, None], k[:, :, :13], 0))


def test_centre_sample_zero_centres_the_kept_samples():
    # ns 10, discards 2 + 2 -> keep 6; grid 16 wide; centre_sample 0 -> dst (16 - 6) / 2 = 5
    data = (np.arange(10) + 1j).astype(np.complex64)
    a = _acq(ns=10, nc=1, ky=4, dpre=2, dpost=2)
    g, mk, n = O.place(a, data, (16, 8, 1, 0, 4, 0))
    assert n == 1 and g.shape == (1, 1, 8, 16)
    want = np.zeros(16, np.complex64)
    want[5:11] = data[2:8]
    assert np.array_equal(g[0, 0, 4], want) and np.array_equal(mk[0, 4], want != 0)


def test_first_writer_wins_and_later_overlap_overwrites():
    ones, twos, threes = (np.full(8, v, np.complex64) for v in (1, 2, 3))
    data = np.concatenate([ones, twos, threes])
    # line ky 4 of a 16-wide grid.  A: 8 samples, centre 4 -> [4, 12), placed;
    # B: same line and centre -> its centre 8 is filled, skipped;
    # C: centre 0, discard_pre 6 -> 2 kept samples centred at [7, 9), skipped
    a = np.concatenate([_acq(8, 1, 4, cs=4, off=0), _acq(8, 1, 4, cs=4, off=8), _acq(8, 1, 4, dpre=6, off=16)])
    g, mk, n = O.place(a, data, (16, 8, 1, 0, 4, 0))
    assert n == 1 and np.array_equal(g[0, 0, 4, 4:12], ones)
    # centre 7 -> [1, 9): its centre 5 lies inside A, skipped
    a2 = np.concatenate([_acq(8, 1, 4, cs=4, off=0), _acq(8, 1, 4, cs=7, off=8)])
    assert O.place(a2, data, (16, 8, 1, 0, 4, 0))[2] == 1
    # A: 4 samples -> [6, 10); B: centre 1 -> [7, 15), centre 11 still free:
    # placed, and it overwrites A's samples 7..9 (the reference writes in order)
    a3 = np.concatenate([_acq(4, 1, 4, cs=2, off=0), _acq(8, 1, 4, cs=1, off=8)])
    g3, mk3, n3 = O.place(a3, data, (16, 8, 1, 0, 4, 0))
    assert n3 == 2
    assert np.array_equal(g3[0, 0, 4, 6:7], ones[:1]) and np.array_equal(g3[0, 0, 4, 7:15], twos)
    assert mk3[0, 4, 6:15].all() and not mk3[0, 4, :6].any()


def test_fewer_channels_than_grid_and_sizing():
    data = np.concatenate([np.full(2 * 6, 1, np.complex64), np.full(6, 2, np.complex64)])
    # acquisition 0: 2 channels, 6 samples, ky 1; acquisition 1: 1 channel, ky 2; encoded 10 x 4 x 1
    a = np.concatenate([_acq(6, 2, 1, cs=3, off=0), _acq(6, 1, 2, cs=3, off=12)])
    g, mk, n = O.place(a, data, (10, 4, 1, 0, 2, 0))
    assert g.shape == (2, 1, 4, 10) and n == 2
    assert np.all(g[:, 0, 1, 2:8] == 1) and np.all(g[0, 0, 2, 2:8] == 2) and np.all(g[1, 0, 2] == 0)
    # the sizing pass: the largest observed index and sample count beat a smaller header
    a2 = _acq(20, 1, 9)
    shp = O.place(a2, np.zeros(20, np.complex64), (10, 4, 1, 0, 2, 0))[0].shape
    assert shp == (1, 1, 10, 20)
    # no image acquisition: the reference errors
    noise = _acq(6, 1, 0, flags=abi.ACQ_IS_NOISE_MEASUREMENT)
    assert O.place(noise, np.zeros(6, np.complex64), (10, 4, 1, 0, 2, 0)) is None


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_whitening_commutes_with_placement(prec):
    case = _case3d()
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], reverse_every=2, shuffle_seed=5)
    rng = np.random.default_rng(9)
    W = np.tril(rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))).astype(np.complex64)
    g, _, _ = O.place(a, d, e, prec=prec)
    gw, _, _ = O.place(a, d, e, W=W, prec=prec)
    want = O.prewhiten_apply(W, g.reshape(3, -1), prec).reshape(g.shape)
    assert np.array_equal(gw, want)


def test_oversampling_removal_places_the_removed_lines():
    case = synth.make_case(nc=2, nz=1, ny=8, nx=32, kind="full", seed=6, ellipses=2)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], recon_x=16)
    g, mk, n = O.place(a, d, e, remove_os=True, prec="f64")
    assert g.shape == (2, 1, 8, 32) and n == 8
    rec = O.os_remove(case["kspace"][:, 0].reshape(-1, 32), 16, "f64").reshape(2, 8, 16)
    assert np.array_equal(g[:, 0, :, 8:24], rec)  # centre 16 -> [32/2 - 16/2, + 16)
    assert not g[:, 0, :, :8].any() and not g[:, 0, :, 24:].any()
    assert np.array_equal(mk[0], np.pad(np.ones((8, 16), bool), ((0, 0), (8, 8))))


def test_library_sizing_pass_matches_oracle():
    # oks_place_shape is host arithmetic: it runs without a GPU
    from openkspace_amd.device import place_shape
    case = _case3d()
    for kw in (dict(), dict(discard=(3, 2)), dict(noise_scans=3)):
        a, d, e = synth.acquisitions(case["kspace"], case["mask"], **kw)
        assert place_shape(a, e) == O.place(a, d, e)[0].shape
    ms = synth.make_case(nc=2, nz=3, ny=10, nx=8, kind="full", seed=4, ellipses=2)
    a, d, e = synth.acquisitions(ms["kspace"], ms["mask"], slice_axis=True)
    assert place_shape(a, e) == (2, 3, 10, 8)
    assert place_shape(_acq(20, 1, 9), (10, 4, 1, 0, 2, 0)) == (1, 1, 10, 20)
    assert place_shape(_acq(6, 1, 0, flags=abi.ACQ_IS_NOISE_MEASUREMENT), (10, 4, 1, 0, 2, 0)) is None
