"""The output writers (cli/main.rs:1033-1153): the oracle's PNG pixel
restatement and the NIfTI header, pinned without the reference:

* the reference's own test nifti_round_trip_header (cli/main.rs:1160-1200)
  on both the oracle's and the library's header (host arithmetic, no GPU);
* percentile(): the window is the element round(pct / 100 * (n - 1)) of the
  sorted slice, checked against numpy's sort, including a rank that lands
  exactly on .5 (Rust's round goes away from zero, numpy's to even);
* the byte map in closed form: window edges to 0 and 255, a mid-window value
  to round(0.5^0.9 * 255) = 137, the linear g-factor window [1, max(P99, 2)].
"""
import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi


def _check_header(b, nz, ny, nx):  # cli/main.rs:1160-1200
    assert len(b) == 352
    assert int.from_bytes(b[0:4], "little", signed=True) == 348
    dims = np.frombuffer(b[40:56], "<i2")
    assert list(dims) == [3, nx, ny, nz, 1, 1, 1, 1]
    assert np.frombuffer(b[70:74], "<i2").tolist() == [16, 32]
    assert np.frombuffer(b[76:108], "<f4").tolist() == [1, 1, 1, 1, 0, 0, 0, 0]
    assert np.frombuffer(b[108:116], "<f4").tolist() == [352.0, 1.0]
    assert b[344:348] == b"n+1\x00" and b[348:352] == b"\x00" * 4


def test_reference_nifti_header():
    _check_header(O.nifti_header(4, 8, 16), 4, 8, 16)


def test_library_nifti_header_matches():
    from openkspace_amd.device import nifti_header
    from openkspace_amd._lib import OksError
    for shape in [(4, 8, 16), (16, 368, 320), (160, 256, 256), (1, 1, 1)]:
        b = nifti_header(*shape)
        _check_header(b, *shape)
        assert b == O.nifti_header(*shape)
    with pytest.raises(OksError) as e:
        nifti_header(1, 1, 40000)
    assert e.value.code == abi.OKS_E_UNSUPPORTED
    assert O.nifti_header(1, 1, 40000) is None


def _rust_rank(pct, n):
    kf = np.float32(np.float32(pct) / np.float32(100.0)) * np.float32(n - 1)
    k = int(np.floor(np.float64(kf) + 0.5))  # half away from zero for kf >= 0
    return min(max(k, 0), n - 1)


@pytest.mark.parametrize("shape,lo,hi", [((3, 17, 23), 0.5, 99.5), ((2, 64, 48), 5.0, 95.0), ((1, 9, 7), 0.0, 100.0)])
def test_window_is_the_sorted_element(shape, lo, hi):
    v = np.random.default_rng(sum(shape)).standard_normal(shape).astype(np.float32)
    _, win = O.render_slices(v, lo, hi)
    n = shape[1] * shape[2]
    for z in range(shape[0]):
        s = np.sort(v[z].ravel())
        assert win[z, 0] == s[_rust_rank(lo, n)]
        assert win[z, 1] == max(s[_rust_rank(hi, n)], win[z, 0] + np.finfo(np.float32).eps)


def test_rank_rounds_half_away_from_zero():
    # n = 6: 50 % of 5 is exactly 2.5 -> index 3 (numpy's round would give 2)
    v = np.array([[[5.0, 1.0, 4.0], [2.0, 6.0, 3.0]]], np.float32)
    _, win = O.render_slices(v, 0.0, 50.0)
    assert win[0].tolist() == [1.0, 4.0]


def test_byte_map_closed_form():
    v = np.array([[[0.0, 10.0, 5.0, -3.0, 20.0, 2.5]]], np.float32)
    b, win = O.render_slices(v, 0.0, 100.0)  # window [min, max] = [-3, 20]
    assert win[0].tolist() == [-3.0, 20.0]
    b2, _ = O.render_slices(np.array([[[0.0, 1.0, 0.5, 0.25]]], np.float32), 0.0, 100.0)
    assert b2[0, 0].tolist() == [0, 255, round(0.5 ** 0.9 * 255), round(0.25 ** 0.9 * 255)]
    assert b[0, 0, 3] == 0 and b[0, 0, 4] == 255


def test_linear_gfactor_window():
    g = np.array([[[1.0, 1.5, 1.25, 0.5, 1.9, 1.0]]], np.float32)
    b, win = O.render_slices(g, mode=abi.RENDER_LINEAR)
    assert win[0].tolist() == [1.0, 2.0]  # P99 = 1.9 < 2
    assert b[0, 0].tolist() == [0, 128, 64, 0, round(0.9 * 255), 0]  # 127.5 rounds away from zero
    g2 = np.array([[[1.0, 4.0, 3.0, 2.0]]], np.float32)
    _, win2 = O.render_slices(g2, mode=abi.RENDER_LINEAR)
    assert win2[0].tolist() == [1.0, 4.0]
