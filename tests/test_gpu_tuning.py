"""The committed gfx950 launch-shape table on the GPU: a context reads it, a
call at a tuned shape takes its options, and the image equals the untuned
call's bit for bit (no tuned option changes the computed bits)."""
import os

import numpy as np
import pytest

from openkspace_amd import abi, synth
from openkspace_amd._lib import Context
from tests.gpu_util import recon_dev

pytestmark = pytest.mark.gpu
TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openkspace_amd", "tuning",
                     "gfx950.json")


@pytest.mark.skipif(not os.path.exists(TABLE), reason="no committed gfx950 table")
@pytest.mark.parametrize("strategy,si", [("ifft_rss", 2), ("ifft_rss", 3), ("cs", 0), ("cs", 1)])
def test_tuned_equals_untuned(strategy, si):
    case = synth.serve_pool_case(si, strategy, 0)
    if strategy == "cs":
        desc, fn = abi.CsDesc.default(remove_oversampling=0, iters=12), "oks_recon_cs_dev"
    else:
        desc, fn = abi.IfftRssDesc.default(remove_oversampling=0), "oks_recon_ifft_rss_dev"
    tuned, plain = Context(0), Context(0)
    try:
        assert tuned.arch() == "gfx950"
        assert tuned.get_option("tuning_entries") > 0
        plain.set_option("tuning", 0)
        a, fa, _ = recon_dev(tuned, fn, desc, case)
        assert tuned.get_option("tuned_calls") >= 1
        b, fb, _ = recon_dev(plain, fn, desc, case)
        assert plain.get_option("tuned_calls") == 0
        assert fa == fb and np.array_equal(a, b)
    finally:
        tuned.close()
        plain.close()
