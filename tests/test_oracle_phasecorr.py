"""The oracle's navigator phase correction (PhaseCorrector,
recon/phasecorr.rs:46-170) inside the placement closure, pinned without the
reference:

* the reference's own tests: an empty corrector is a no-op
  (phasecorr.rs:189-196) and a constant navigator phase of pi/3 is removed
  from an image line carrying it (:198-228);
* round trips: readouts carrying a per-(slice, segment) linear phase error,
  with navigators that measure it, place back onto the error-free grid
  (3-D, multi-slice, reversed lines with discards, the (65535, segment)
  fallback key), while placement without correction does not;
* the corrector's rules in closed form: two navigators average to the mean
  phase, a zero navigator gives the unit correction, a navigator of another
  length is dropped from its bin, a noise scan that also carries the
  phase-correction flag is a noise scan, and a short navigator corrects only
  its first samples.
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth

PC = abi.ACQ_IS_PHASECORR_DATA


def _rec(n=1, **kw):
    return abi.acq_records(n, **kw)


def _case3d():
    return synth.make_case(nc=3, nz=2, ny=16, nx=12, kind="grappa", r=2, acs=6, seed=3, ellipses=3, three_d=True)


def test_reference_empty_corrector_is_noop():  # phasecorr.rs:189-196
    case = _case3d()
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], reverse_every=2)
    assert O.phase_corrections(a, d) == {}
    g0, m0, n0 = O.place(a, d, e)
    g1, m1, n1 = O.place(a, d, e, phase_correct=True)
    assert n0 == n1 and np.array_equal(g0, g1) and np.array_equal(m0, m1)


def test_reference_removes_constant_phase_offset():  # phasecorr.rs:198-228
    ns, phi = 8, np.float32(np.pi / 3)
    nav = np.full(ns, np.cos(phi) + 1j * np.sin(phi), np.complex64)
    img = (nav * np.arange(1, ns + 1)).astype(np.complex64)
    a = np.concatenate([_rec(flags=PC, number_of_samples=ns, active_channels=1, data_offset=0),
                        _rec(number_of_samples=ns, active_channels=1, center_sample=ns // 2,
                             kspace_encode_step_1=2, data_offset=ns)])
    g, _, n = O.place(a, np.concatenate([nav, img]), (ns, 4, 1, 0, 2, 0), phase_correct=True)
    assert n == 1
    row = g[0, 0, 2]
    assert np.all(np.abs(row.real - np.arange(1, ns + 1)) < 1e-4) and np.all(np.abs(row.imag) < 1e-4)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-6), ("f32", 2e-6)])
@pytest.mark.parametrize("kw", [dict(segments=3), dict(segments=2, reverse_every=2, discard=(2, 1), shuffle_seed=1),
                                dict(segments=4, nav_fallback=True, noise_scans=3)])
def test_round_trip_3d(prec, tol, kw):
    case = _case3d()
    k, rows = case["kspace"], case["mask"].any(-1)
    a, d, e = synth.acquisitions(k, case["mask"], **kw)
    g, mk, n = O.place(a, d, e, prec=prec, phase_correct=True)
    g0 = O.place(a, d, e, prec=prec)[0]
    x0 = g.shape[3] // 2 - 6
    want = np.where(rows[None, :, :, None], k, 0)
    scale = np.abs(want).max()
    assert np.abs(g[..., x0:x0 + 12] - want).max() < tol * scale
    assert np.abs(g0[..., x0:x0 + 12] - want).max() > 0.1 * scale


def test_round_trip_multislice_bins_per_slice():
    case = synth.make_case(nc=2, nz=3, ny=10, nx=8, kind="full", seed=4, ellipses=2)
    a, d, e = synth.acquisitions(case["kspace"], case["mask"], slice_axis=True, segments=2, shuffle_seed=3)
    corr = O.phase_corrections(a, d)
    assert sorted(corr) == [(z, s) for z in range(3) for s in range(2)]
    g, mk, n = O.place(a, d, e, phase_correct=True)
    assert n == 30 and np.abs(g - case["kspace"]).max() < 2e-6 * np.abs(case["kspace"]).max()


def test_two_navigators_average_to_the_mean_phase():
    ns, p1, p2 = 6, 0.3, 1.1
    nav = np.concatenate([np.full(ns, 2 * np.exp(1j * p1)), np.full(ns, 2 * np.exp(1j * p2))]).astype(np.complex64)
    a = _rec(2, flags=PC, number_of_samples=ns, active_channels=1, slice=1, segment=4, data_offset=[0, ns])
    corr = O.phase_corrections(a, nav, prec="f64")
    assert list(corr) == [(1, 4)]
    assert np.allclose(corr[(1, 4)], np.exp(-1j * (p1 + p2) / 2), atol=1e-6)


def test_corrector_rules():
    ns = 6
    ones = np.ones(ns, np.complex64)
    twisted = np.full(ns, np.exp(0.7j), np.complex64)
    # a zero navigator: |avg| < 1e-20 gives the unit correction
    z = O.phase_corrections(_rec(flags=PC, number_of_samples=ns, active_channels=1), np.zeros(ns, np.complex64))
    assert np.array_equal(z[(0, 0)], ones)
    # a later navigator of another length is dropped from the bin
    a = _rec(2, flags=PC, number_of_samples=[ns, 4], active_channels=1, data_offset=[0, ns])
    c = O.phase_corrections(a, np.concatenate([twisted, np.full(4, 1j, np.complex64)]), prec="f64")
    assert np.allclose(c[(0, 0)], np.exp(-0.7j), atol=1e-6) and len(c[(0, 0)]) == ns
    # a noise scan with the phase-correction flag is a noise scan
    a = _rec(flags=PC | abi.ACQ_IS_NOISE_MEASUREMENT, number_of_samples=ns, active_channels=1)
    assert O.phase_corrections(a, twisted) == {}


def test_short_navigator_corrects_leading_samples_only():
    ns, nn = 8, 5
    nav = np.full(nn, np.exp(0.5j), np.complex64)
    img = np.full(ns, np.exp(0.5j), np.complex64)
    a = np.concatenate([_rec(flags=PC, number_of_samples=nn, active_channels=1),
                        _rec(number_of_samples=ns, active_channels=1, center_sample=ns // 2,
                             kspace_encode_step_1=2, data_offset=nn)])
    g, _, _ = O.place(a, np.concatenate([nav, img]), (ns, 4, 1, 0, 2, 0), prec="f64", phase_correct=True)
    row = g[0, 0, 2]
    assert np.allclose(row[:nn], 1.0, atol=1e-6) and np.allclose(row[nn:], np.exp(0.5j), atol=1e-6)
