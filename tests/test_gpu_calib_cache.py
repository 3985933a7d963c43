"""The resident calibration cache: GRAPPA weights per slice (served path) and
SENSE coil maps per call (context option), keyed by a device hash of the
calibration rows after whitening.  A hit must give the bits a fresh
calibration gives, flags included, and changed calibration data must miss.
"""
import numpy as np
import pytest

from openkspace_amd import abi, synth
from openkspace_amd._lib import Context
from openkspace_amd.server import Server
from tests.gpu_util import recon_dev
from tests.test_gpu_server import Slice

pytestmark = pytest.mark.gpu


def _serve_all(slices, cache_bytes, passes):
    images = []
    with Server(slots=2, max_batch=8, calib_cache_bytes=cache_bytes) as s:
        for p in range(passes):
            t0 = s.now() + 0.02
            for sl in slices:
                s.submit(Server.request(sl.desc, sl.ks, sl.out.addr, None, 100 * p + sl.tag, t0))
            got = s.collect(len(slices))
            assert all(c.status == 0 for c in got.values()), s.last_error()
            images.append({sl.tag: (sl.image(got[100 * p + sl.tag]), got[100 * p + sl.tag].flags) for sl in slices})
        st = s.stats()
    return images, st


def test_grappa_weights_cache_hits_are_bit_identical():
    ctx = Context(0)
    gd = abi.GrappaDesc.default(remove_oversampling=0)
    cases = [synth.make_case(nc=8, nz=1, ny=96, nx=64, kind="grappa", r=4, acs=24, seed=300 + i, ellipses=5)
             for i in range(4)]
    bad = cases[3]["kspace"].copy()
    bad[0, 0, 48, 30] = np.nan  # this slice's calibration fails: the flag must survive a cache hit
    cases[3] = dict(cases[3], kspace=bad)
    slices = [Slice(ctx, c, gd, i) for i, c in enumerate(cases)]
    try:
        ref, st0 = _serve_all(slices, 0, 1)
        got, st = _serve_all(slices, 256 << 20, 3)
        assert st0.calib_hits == 0
        assert st.calib_hits >= len(slices) and st.calib_misses >= len(slices), (st.calib_hits, st.calib_misses)
        assert st.calib_bytes > 0
        for p in range(3):
            for sl in slices:
                img, fl = got[p][sl.tag]
                r_img, r_fl = ref[0][sl.tag]
                assert fl == r_fl, (p, sl.tag)
                assert np.array_equal(img, r_img, equal_nan=True), (p, sl.tag)
        assert got[2][3][1] & abi.F_GRAPPA_CALIB_FAILED and not got[2][0][1] & abi.F_GRAPPA_CALIB_FAILED
    finally:
        for sl in slices:
            sl.free()
        ctx.close()


@pytest.mark.parametrize("source", [abi.MAPS_WALSH, abi.MAPS_ESPIRIT])
def test_sense_maps_cache(source):
    case = synth.make_case(nc=6, nz=2, ny=96, nx=48, kind="grappa", r=2, acs=24, seed=41, ellipses=6)
    d = abi.SenseDesc.default(map_source=source, compute_gfactor=1)
    plain, cached = Context(0), Context(0)
    try:
        cached.set_option("calib_cache_mb", 64)
        want, wf, wg = recon_dev(plain, "oks_recon_sense_dev", d, case)
        a, fa, ga = recon_dev(cached, "oks_recon_sense_dev", d, case)
        assert cached.get_option("calib_hits") == 0
        b, fb, gb = recon_dev(cached, "oks_recon_sense_dev", d, case)
        assert cached.get_option("calib_hits") == 1
        for img, fl, g in ((a, fa, ga), (b, fb, gb)):
            assert fl == wf and np.array_equal(img, want) and np.array_equal(g, wg)
        # other calibration data: a miss, and the other data's own image
        other = dict(case, kspace=(case["kspace"] * np.complex64(0.5 + 0.25j)).astype(np.complex64))
        c, _, _ = recon_dev(cached, "oks_recon_sense_dev", d, other)
        assert cached.get_option("calib_hits") == 1
        w2, _, _ = recon_dev(plain, "oks_recon_sense_dev", d, other)
        assert np.array_equal(c, w2)
    finally:
        plain.close()
        cached.close()


def test_grappa_cache_on_plain_context_and_changed_acs():
    case = synth.make_case(nc=6, nz=3, ny=96, nx=64, kind="grappa", r=3, acs=24, seed=55, ellipses=5)
    d = abi.GrappaDesc.default()
    plain, cached = Context(0), Context(0)
    try:
        cached.set_option("calib_cache_mb", 64)
        want, wf, _ = recon_dev(plain, "oks_recon_grappa_dev", d, case)
        for _ in range(2):
            got, fl, _ = recon_dev(cached, "oks_recon_grappa_dev", d, case)
            assert fl == wf and np.array_equal(got, want)
        assert cached.get_option("calib_hits") == 3
        k = case["kspace"].copy()
        k[2, 1, 48, 10] += np.complex64(1e-3)  # one ACS sample of slice 1 changes
        other = dict(case, kspace=k)
        got, _, _ = recon_dev(cached, "oks_recon_grappa_dev", d, other)
        assert cached.get_option("calib_hits") == 3
        w2, _, _ = recon_dev(plain, "oks_recon_grappa_dev", d, other)
        assert np.array_equal(got, w2)
    finally:
        plain.close()
        cached.close()
