"""GPU tests of the served path (oks_server_*): continuous batching.

* Each request's image is the same alone (a batch of one, the plain
  single-request path) and inside a mixed stream where it shares a multi-slice
  launch with other requests: within 1e-5 rel-L2 for the FP32 strategies and
  1e-3 for split-BF16 GRAPPA (the bars of the strategies themselves); the
  measured distance is printed.
* Served images match the oracle (FP32 restatement and FP64 twin) to the
  strategy bars.
* Flags are per request inside a batch (a request without noise scans, a
  GRAPPA slice whose calibration fails), admission refuses what exceeds the
  HBM budget, a bounded queue answers BUSY, and arrival times are honoured.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import Context, DeviceArray, OksError, lib
from openkspace_amd.server import Server, footprint
from tests.gpu_util import assert_parity, rel

pytestmark = pytest.mark.gpu
TOL = {"ifft_rss": 1e-5, "grappa": 1e-3, "cs": 1e-5}


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


class Slice:
    """One device-resident single-slice request: k-space, noise, output."""

    def __init__(self, ctx, case, desc, tag):
        self.case, self.desc, self.tag = case, desc, tag
        self.oracle = abi.STRATEGY_OF[type(desc)][2]
        k = np.ascontiguousarray(case["kspace"], np.complex64)
        self.k = DeviceArray.from_host(ctx, k)
        nz = case.get("noise")
        self.noise = DeviceArray.from_host(ctx, np.ascontiguousarray(nz, np.complex64)) if nz is not None else None
        self.mask = np.ascontiguousarray(case["mask"], np.uint8)
        nc, z, ny, nx = k.shape
        self.ks = abi.KSpace(self.k.addr, self.mask.ctypes.data, nc, z, ny, nx,
                             self.noise.addr if self.noise is not None else None,
                             nz.shape[1] if nz is not None else 0, abi.Header(**case["hdr"]))
        self.out = DeviceArray(ctx, (z * ny * nx,), np.float32)

    def request(self, not_before=0.0):
        return Server.request(self.desc, self.ks, self.out.addr, None, self.tag, not_before)

    def image(self, c):
        return self.out.download((c.nz, c.ny, c.nx))

    def free(self):
        for d in (self.k, self.noise, self.out):
            if d is not None:
                d.free()


def _mix(ctx):
    """Three strategies x two shapes x two distinct slices, single-slice each."""
    specs = [
        (abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0), dict(nc=4, ny=64, nx=96, kind="full")),
        (abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0),
         dict(nc=5, ny=48, nx=128, os_x=2, kind="full")),
        (abi.GrappaDesc.default(), dict(nc=6, ny=64, nx=48, kind="grappa", r=4, acs=24)),
        (abi.GrappaDesc.default(), dict(nc=4, ny=96, nx=64, kind="grappa", r=3, acs=24)),
        (abi.CsDesc.default(iters=20), dict(nc=3, ny=64, nx=64, kind="cs")),
        (abi.CsDesc.default(iters=20), dict(nc=2, ny=48, nx=80, kind="cs")),
    ]
    slices, tag = [], 0
    for i, (desc, p) in enumerate(specs):
        for rep in range(2):
            case = synth.make_case(nz=1, seed=1000 + 10 * i + rep, ellipses=5, **p)
            slices.append(Slice(ctx, case, desc, tag))
            tag += 1
    return slices


def _serve(slices, **cfg):
    with Server(**cfg) as s:
        t0 = s.now() + 0.05  # everything arrives together
        for sl in slices:
            s.submit(sl.request(t0))
        got = s.collect(len(slices))
        st = s.stats()
    return got, st


def test_alone_vs_mixed_and_oracle(ctx):
    slices = _mix(ctx)
    try:
        alone = {}
        with Server(slots=1, max_batch=1) as s:
            for sl in slices:
                s.submit(sl.request())
                c = s.collect(1)[sl.tag]
                assert c.status == 0 and c.batch == 1
                alone[sl.tag] = (sl.image(c), c.flags)
        # replicate every slice 3x so the mix holds same-key batches
        mixed = slices * 3
        with Server(slots=3, max_batch=8) as s:
            t0 = s.now() + 0.05
            for j, sl in enumerate(mixed):
                s.submit(Server.request(sl.desc, sl.ks, sl.out.addr, None, 1000 + j, t0))
            got = s.collect(len(mixed))
            st = s.stats()
        assert st.completed == len(mixed) and st.failed == 0
        assert max(c.batch for c in got.values()) > 1, "no request shared a launch"
        assert st.batches < len(mixed)
        # The outputs of replicas overwrite each other's buffer with the same
        # request: compare the last-written image of each slice.
        for sl in slices:
            c = [got[1000 + j] for j, m in enumerate(mixed) if m is sl][-1]
            img, fl = sl.image(c), c.flags
            a_img, a_fl = alone[sl.tag]
            e = rel(img, a_img)
            print(f"{sl.oracle} {sl.case['kspace'].shape}: in-mix (batch {c.batch}) vs alone rel-L2 {e:.2e}")
            assert fl == a_fl
            assert e <= TOL[sl.oracle]
        # served images against the oracle, one slice per strategy/shape
        for sl in slices[::2]:
            st32, r32, f32, _ = O.recon(sl.oracle, sl.desc, sl.case, "f32")
            _, r64, _, _ = O.recon(sl.oracle, sl.desc, sl.case, "f64")
            img, fl = alone[sl.tag]
            assert st32 == 0 and fl == f32
            assert_parity(img, r32, r64, TOL[sl.oracle], f"served {sl.oracle} {sl.case['kspace'].shape}",
                          iterative=sl.oracle == "cs")
    finally:
        for sl in slices:
            sl.free()


def test_flags_are_per_request_in_a_batch(ctx):
    desc = abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0)
    cases = [synth.make_case(nc=4, nz=1, ny=64, nx=64, kind="full", seed=70 + i, ellipses=4) for i in range(4)]
    cases[2] = dict(cases[2], noise=None)  # no noise scans -> pre-whitening skipped for this one only
    gd = abi.GrappaDesc.default()
    gcases = [synth.make_case(nc=4, nz=1, ny=64, nx=48, kind="grappa", r=4, acs=24, seed=90 + i, ellipses=4)
              for i in range(3)]
    # A non-finite ACS sample makes this slice's normal matrix non-finite, so its
    # Cholesky fails (recon/prewhiten.rs:160-164) and it stays undersampled; an
    # all-zero slice would not fail, the ridge is floored at f32::EPSILON
    # (recon/grappa.rs:262-275).  Its batch mates must not be affected.
    bad = gcases[1]["kspace"].copy()
    bad[0, 0, 32, 24] = np.nan
    gcases[1] = dict(gcases[1], kspace=bad)
    slices = [Slice(ctx, c, desc, i) for i, c in enumerate(cases)]
    slices += [Slice(ctx, c, gd, 10 + i) for i, c in enumerate(gcases)]
    try:
        got, _ = _serve(slices, slots=1, max_batch=8)
        for sl in slices:
            c = got[sl.tag]
            assert c.status == 0 and c.batch > 1
            _, ref, rflags, _ = O.recon(sl.oracle, sl.desc, sl.case, "f32")
            assert c.flags == rflags, (sl.tag, c.flags, rflags)
            img = sl.image(c)
            if sl.tag != 11:
                assert rel(img, ref) < 2e-3
        assert got[2].flags & abi.F_PREWHITEN_SKIPPED and not got[1].flags & abi.F_PREWHITEN_SKIPPED
        assert got[11].flags & abi.F_GRAPPA_CALIB_FAILED and not got[10].flags & abi.F_GRAPPA_CALIB_FAILED
    finally:
        for sl in slices:
            sl.free()


def test_admission_budget_queue_and_arrivals(ctx):
    desc = abi.IfftRssDesc.default(remove_oversampling=0, partial_fourier=0)
    case = synth.make_case(nc=4, nz=1, ny=64, nx=64, kind="full", seed=5, ellipses=4)
    sl = Slice(ctx, case, desc, 0)
    try:
        fp = footprint(sl.request())
        assert fp >= 4 * 64 * 64 * 8
        with Server(slots=2, max_batch=8, hbm_budget=fp - 1) as s:
            with pytest.raises(OksError) as e:
                s.submit(sl.request())
            assert e.value.code == abi.OKS_E_OOM
        # a budget of one request's footprint: no batching, every request served
        with Server(slots=2, max_batch=8, hbm_budget=fp + fp // 2) as s:
            t0 = s.now() + 0.02
            for j in range(6):
                s.submit(Server.request(desc, sl.ks, sl.out.addr, None, j, t0))
            got = s.collect(6)
            st = s.stats()
            assert all(c.status == 0 and c.batch == 1 for c in got.values())
            assert st.peak_inflight_bytes <= fp + fp // 2
        # bounded queue, arrivals in the future, then dropped at destroy
        with Server(slots=1, max_batch=4, max_pending=2) as s:
            far = s.now() + 3600.0
            s.submit(sl.request(far))
            s.submit(sl.request(far))
            with pytest.raises(OksError) as e:
                s.submit(sl.request(far))
            assert e.value.code == abi.OKS_E_BUSY
            assert s.poll(8, 0) == []
        # arrival times are honoured
        with Server(slots=1, max_batch=4) as s:
            t0 = s.now()
            for j in range(4):
                s.submit(Server.request(desc, sl.ks, sl.out.addr, None, j, t0 + 0.05 * (j + 1)))
            got = s.collect(4)
            for j, c in got.items():
                assert c.t_arrival >= t0 + 0.05 * (j + 1) - 1e-6
                assert c.t_start >= c.t_arrival and c.t_done >= c.t_start
        with Server(slots=1, max_batch=1) as s:
            bad = sl.request()
            bad.strategy = 9
            with pytest.raises(OksError) as e:
                s.submit(bad)
            assert e.value.code == abi.OKS_E_INVALID_ARG
    finally:
        sl.free()


def test_partial_fourier_requests_served(ctx):
    # ifft-rss requests acquired at 6/8 partial Fourier take the homodyne path
    # (recon/strategy.rs:163-187) in a batch as alone
    desc = abi.IfftRssDesc.default(remove_oversampling=0)
    slices = []
    for i in range(3):
        case = synth.make_case(nc=4, nz=1, ny=64, nx=96, kind="full", seed=2000 + i, ellipses=5)
        m = case["mask"].copy()
        m[:, :16] = False
        case["mask"] = np.ascontiguousarray(m)
        case["kspace"] = np.ascontiguousarray(np.where(m[None], case["kspace"], 0).astype(np.complex64))
        slices.append(Slice(ctx, case, desc, i))
    try:
        alone = {}
        with Server(slots=1, max_batch=1) as s:
            for sl in slices:
                s.submit(sl.request())
                c = s.collect(1)[sl.tag]
                assert c.status == 0 and c.flags & abi.F_HOMODYNE
                alone[sl.tag] = (sl.image(c), c.flags)
        mixed = slices * 3
        with Server(slots=2, max_batch=8) as s:
            t0 = s.now() + 0.05
            for j, sl in enumerate(mixed):
                s.submit(Server.request(sl.desc, sl.ks, sl.out.addr, None, 100 + j, t0))
            got = s.collect(len(mixed))
        assert max(c.batch for c in got.values()) > 1
        for sl in slices:
            c = [got[100 + j] for j, m in enumerate(mixed) if m is sl][-1]
            a_img, a_fl = alone[sl.tag]
            assert c.flags == a_fl and rel(sl.image(c), a_img) <= TOL["ifft_rss"]
            st32, r32, f32, _ = O.recon("ifft_rss", desc, sl.case, "f32")
            _, r64, _, _ = O.recon("ifft_rss", desc, sl.case, "f64")
            assert st32 == 0 and f32 == a_fl
            assert_parity(a_img, r32, r64, TOL["ifft_rss"], "served homodyne")
    finally:
        for sl in slices:
            sl.free()


def test_more_than_32_channel_requests_run_alone(ctx):
    # the batched kernels hold at most 32 channels: larger requests (up to
    # 64) run alone on the plain path beside batches of small ones; above 64
    # a request is refused at submission
    specs = [(abi.IfftRssDesc.default(), dict(nc=40, ny=48, nx=64, os_x=2, kind="full")),
             (abi.CsDesc.default(iters=10), dict(nc=36, ny=48, nx=48, kind="cs")),
             (abi.GrappaDesc.default(kernel_ky=2, kernel_kx=3), dict(nc=34, ny=64, nx=48, kind="grappa", r=2, acs=16)),
             (abi.IfftRssDesc.default(), dict(nc=8, ny=48, nx=64, os_x=2, kind="full"))]
    slices = [Slice(ctx, synth.make_case(nz=1, seed=700 + i, ellipses=5, **kw), desc, i)
              for i, (desc, kw) in enumerate(specs)]
    try:
        mixed = slices + slices
        with Server(slots=2, max_batch=8) as s:
            t0 = s.now() + 0.05
            for j, sl in enumerate(mixed):
                s.submit(Server.request(sl.desc, sl.ks, sl.out.addr, None, 100 + j, t0))
            got = s.collect(len(mixed))
            big = slices[0]
            ks = abi.KSpace(big.k.addr, big.mask.ctypes.data, 65, 1, 48, 64, None, 0, abi.Header(**big.case["hdr"]))
            with pytest.raises(OksError) as e:
                s.submit(Server.request(big.desc, ks, big.out.addr, None, 999, 0.0))
            assert e.value.code == abi.OKS_E_INVALID_ARG and "64" in str(e.value)
        for j, sl in enumerate(mixed):
            c = got[100 + j]
            nc = sl.case["kspace"].shape[0]
            assert c.status == 0
            if nc > 32:
                assert c.batch == 1
            st, r32, fl, _ = O.recon(sl.oracle, sl.desc, sl.case, "f32")
            _, r64, _, _ = O.recon(sl.oracle, sl.desc, sl.case, "f64")
            assert st == 0 and c.flags == fl
            assert_parity(sl.image(c), r32, r64, TOL[sl.oracle], f"served {sl.oracle} nc={nc}",
                          iterative=sl.oracle == "cs")
    finally:
        for sl in slices:
            sl.free()
