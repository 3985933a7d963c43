"""Pins of the seeded workload masks.

The variable-density CS masks are random draws: their SHA-256 digests are
fixed here so a change of generator, seed or numpy stream shows up. The
regular masks are integer rules and are checked by their defining property.
"""
import numpy as np

from openkspace_amd import synth

# SHA-256 of the uint8 mask bytes.
C3_MASKS_SHA = "322b02a9dce8e2183936fad814f220ea0f1e59dfa805bb83584f8c4d0f766e2f"
CS_MASK_320_SHA = "a3729221bd5154e10005f5adb08549a81a40c3e7c75aba5ff8c9acdd69bfdefc"
SERVE_CS_MASK_SHA = [
    "5d29454568262849433992be4ffd6123ebfbf53ecac3ca4c51663037edb7036d",  # 256 x 256, 8 coils
    "b0aa76d41b4e49c00a6717eaa77c75d788e2dcf37351a6b8591bf1ad42cd83d1",  # 320 x 320, 16 coils
    "9500528947bb6cc407912e312a92e9dde9ce4302d8a1cdaaba136dc9aaed9fca",  # 368 x 640, 15 coils
    "b9f4f21226780bd55976dc741bcf0318eaa2b89da1f4a4121104ffd63185d5cd",  # 384 x 384, 32 coils
]


def _check_cs_mask(m):
    ny, nx = m.shape
    assert m[ny // 2 - 12: ny // 2 + 12, nx // 2 - 12: nx // 2 + 12].all()
    assert abs(m.mean() - 1 / 6) < 0.01


def test_cs_mask_draw_sha():
    m = synth.cs_mask(np.random.default_rng(synth.SEED_BASE + 2), 320, 320)
    _check_cs_mask(m)
    assert synth.mask_sha256(m) == CS_MASK_320_SHA


def test_c3_masks_sha():
    case = synth.make_config("c3_cs")
    assert case["mask"].shape == (32, 320, 320)
    for z in range(32):
        _check_cs_mask(case["mask"][z])
    assert synth.mask_sha256(case["mask"]) == C3_MASKS_SHA


def test_served_cs_masks_sha():
    for si, (ny, nx, nc, os_x) in enumerate(synth.SERVE_SHAPES):
        case = synth.serve_pool_case(si, "cs", 0)
        assert case["mask"].shape == (1, ny, nx)
        _check_cs_mask(case["mask"][0])
        assert synth.mask_sha256(case["mask"]) == SERVE_CS_MASK_SHA[si], si


def test_c2_mask_rule():
    rows = synth.grappa_mask(384, 4, 24)
    ky = np.arange(384)
    acs = (ky >= 180) & (ky < 204)
    assert np.array_equal(rows, (ky % 4 == 0) | acs)
