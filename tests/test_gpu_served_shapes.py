"""The served mixed stream at its own request shapes, on the GPU.

Every (shape, strategy) pair of the served workload -- 256 x 256 x 8 coils,
320 x 320 x 16, 368 x 640 x 15 (2x readout oversampling, recon 320) and
384 x 384 x 32; ifft-rss, GRAPPA R=4 (24 ACS rows, 4 x 5 kernel, ridge 1e-3)
and CS (60 iterations, lambda 0.01) -- with the descriptors the benchmark
serves.  Two distinct slices per pair are served alone (one slot, batches of
one: the plain single-request path) and then replicated three times inside
one mixed stream, all arriving together, so same-key requests share batches.

Bars: in-mix against alone within 1e-5 rel-L2 for the FP32 strategies and
1e-3 for split-BF16 GRAPPA, with flags equal; each pair's first slice against
the oracle at the strategy's bar (FP32 restatement and FP64 twin).
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import Context
from openkspace_amd.server import Server
from tests.gpu_util import assert_parity, rel
from tests.test_gpu_server import Slice

pytestmark = pytest.mark.gpu
TOL = {"ifft_rss": 1e-5, "grappa": 1e-3, "cs": 1e-5}
DESCS = {"ifft_rss": abi.IfftRssDesc.default(remove_oversampling=0),
         "grappa": abi.GrappaDesc.default(remove_oversampling=0),
         "cs": abi.CsDesc.default(remove_oversampling=0, iters=60)}
PAIRS = [(si, name) for si in range(len(synth.SERVE_SHAPES)) for name in ("ifft_rss", "grappa", "cs")]


@pytest.fixture(scope="module")
def served():
    ctx = Context(0)
    slices = {}
    tag = 0
    for pair in PAIRS:
        slices[pair] = []
        for j in range(2):
            case = synth.serve_pool_case(pair[0], pair[1], j)
            slices[pair].append(Slice(ctx, case, DESCS[pair[1]], tag))
            tag += 1
    every = [sl for p in PAIRS for sl in slices[p]]
    try:
        alone = {}
        with Server(slots=1, max_batch=1) as s:
            for sl in every:
                s.submit(sl.request())
                c = s.collect(1)[sl.tag]
                assert c.status == 0 and c.batch == 1, s.last_error()
                alone[sl.tag] = (sl.image(c), c.flags)
        print(f"served alone: {len(every)} requests", flush=True)
        mixed = every * 3
        order = np.random.default_rng(5).permutation(len(mixed))
        with Server(slots=4, max_batch=32) as s:
            t0 = s.now() + 0.05
            for j in order:
                sl = mixed[j]
                s.submit(Server.request(sl.desc, sl.ks, sl.out.addr, None, 10000 + int(j), t0))
            got = s.collect(len(mixed))
            st = s.stats()
        print(f"served mixed: {st.completed} requests in {st.batches} batches, failed {st.failed}", flush=True)
        assert st.completed == len(mixed) and st.failed == 0
        inmix = {}
        for j, sl in enumerate(mixed):
            inmix.setdefault(sl.tag, []).append(got[10000 + j])
        yield dict(slices=slices, alone=alone, inmix=inmix, got=got)
    finally:
        for sl in every:
            sl.free()
        ctx.close()


@pytest.mark.parametrize("pair", PAIRS, ids=[f"{synth.SERVE_SHAPES[si][0]}x{synth.SERVE_SHAPES[si][1]}-{n}"
                                              for si, n in PAIRS])
def test_in_mix_equals_alone(served, pair):
    name = pair[1]
    batched = 0
    for sl in served["slices"][pair]:
        comps = served["inmix"][sl.tag]
        batched = max(batched, max(c.batch for c in comps))
        # replicas of one request write the same buffer: it holds the last one's image
        c = max(comps, key=lambda c: c.t_done)
        img = sl.image(c)
        a_img, a_fl = served["alone"][sl.tag]
        e = rel(img, a_img)
        print(f"{name} {sl.case['kspace'].shape}: in-mix (batch {c.batch}) vs alone rel-L2 {e:.2e}")
        assert all(x.status == 0 and x.flags == a_fl for x in comps)
        assert e <= TOL[name]
    assert batched > 1, "no request of this pair shared a batch"


@pytest.mark.parametrize("pair", PAIRS, ids=[f"{synth.SERVE_SHAPES[si][0]}x{synth.SERVE_SHAPES[si][1]}-{n}"
                                              for si, n in PAIRS])
def test_served_vs_oracle(served, pair):
    name = pair[1]
    sl = served["slices"][pair][0]
    O.set_threads(16)
    try:
        st32, r32, f32, _ = O.recon(name, sl.desc, sl.case, "f32")
        _, r64, _, _ = O.recon(name, sl.desc, sl.case, "f64")
    finally:
        O.set_threads(1)
    img, fl = served["alone"][sl.tag]
    assert st32 == 0 and fl == f32 and img.shape == r32.shape
    assert_parity(img, r32, r64, TOL[name], f"served {name} {sl.case['kspace'].shape}", iterative=name == "cs")
