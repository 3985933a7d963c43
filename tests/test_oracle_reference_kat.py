"""The reference's own unit tests (known answers), restated against the oracle.

Each test names the reference test it restates (file:line).
"""
import numpy as np
import pytest

import oracle as O


def c(v):
    return np.array(v, dtype=np.complex64)


def test_fftshift_1d_even():  # recon/shift.rs:63
    assert np.array_equal(O.shift(c([0, 1, 2, 3]), 0, False).real, [2, 3, 0, 1])


def test_fftshift_1d_odd():  # recon/shift.rs:70
    assert np.array_equal(O.shift(c([0, 1, 2, 3, 4]), 0, False).real, [2, 3, 4, 0, 1])


def test_shift_roundtrip_odd():  # recon/shift.rs:78
    a = c([0, 1, 2, 3, 4])
    assert np.array_equal(O.shift(O.shift(a, 0, False), 0, True), a)


def test_ifft2_impulse_gives_constant():  # recon/fft.rs:92
    n = 8
    k = np.zeros((n, n), np.complex64)
    k[n // 2, n // 2] = 1
    out = O.ifft_centered(k, (0, 1))
    assert np.abs(np.abs(out) - 1 / (n * n)).max() < 1e-6


def test_ifft2_parseval():  # recon/fft.rs:115
    n = 16
    k = np.zeros((n, n), np.complex64)
    k[0, 0] = 1
    k[1, 2] = 0.5 - 0.25j
    before = np.sum(np.abs(k) ** 2)
    after = np.sum(np.abs(O.ifft_centered(k, (0, 1))) ** 2)
    assert abs(after - before / n ** 2) < 1e-5


def test_ifft3_impulse_gives_constant():  # recon/fft.rs:138
    nz, ny, nx = 4, 8, 8
    k = np.zeros((nz, ny, nx), np.complex64)
    k[nz // 2, ny // 2, nx // 2] = 1
    out = O.ifft_centered(k, (0, 1, 2))
    assert np.abs(np.abs(out) - 1 / (nz * ny * nx)).max() < 1e-6


def test_ifft1_decouples_from_2d():  # recon/fft.rs:161
    nz, ny, nx = 4, 6, 6
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    k = ((z + 1.0) * 0.5) * ((y - 2.0) + 1j * (x * 0.25))
    k = k.astype(np.complex64)
    a = O.ifft_centered(k, (0, 1, 2))
    b = O.ifft_centered(O.ifft_centered(k, (0,)), (1, 2))
    assert np.abs(a - b).max() < 1e-4


def test_rss_pythagorean():  # recon/coil.rs:35
    a = np.array([3, 4, 5], np.complex64).reshape(3, 1, 1)
    assert abs(O.rss(a)[0, 0] - np.sqrt(50.0)) < 1e-6


def test_rss_complex_magnitude():  # recon/coil.rs:47
    a = np.array([3 + 4j, 12j], np.complex64).reshape(2, 1, 1)
    assert abs(O.rss(a)[0, 0] - 13.0) < 1e-6


def test_crop_even_centered():  # recon/crop.rs:35
    v = np.zeros((1, 4, 8), np.float32)
    v[0, 1:3, 2:6] = 1
    out = O.crop3d(v, (1, 2, 4))
    assert out.shape == (1, 2, 4) and np.all(out == 1)


def test_crop_target_larger():  # recon/crop.rs:49
    assert O.crop3d(np.ones((2, 3, 5), np.float32), (10, 10, 10)).shape == (2, 3, 5)


def _lcg():  # recon/prewhiten.rs:238-245
    state = [0xDEADBEEFCAFEF00D]
    mask = (1 << 64) - 1

    def nxt():
        state[0] = (state[0] * 6364136223846793005 + 1442695040888963407) & mask
        u = np.float32(np.uint32(state[0] >> 33)) / np.float32(4294967295.0)
        return np.float32(u - np.float32(0.5))
    return nxt


def test_whitens_correlated_noise_to_identity():  # recon/prewhiten.rs:219
    nc, ns = 4, 4096
    L = np.zeros((nc, nc), np.complex64)
    for i in range(nc):
        L[i, i] = 1.0 + 0.5 * i
        for j in range(i):
            L[i, j] = complex(0.1 * (i + j), -0.05 * i)
    rng = _lcg()
    data = np.zeros((nc, ns), np.complex64)
    for n in range(ns):
        n0 = np.array([complex(rng(), rng()) for _ in range(nc)], np.complex64)
        for i in range(nc):
            data[i, n] = np.sum(L[i, : i + 1] * n0[: i + 1])
    W = O.prewhiten_calibrate(data, "f32")
    assert W is not None
    y = O.prewhiten_apply(W, data, "f32")
    cov = (y @ y.conj().T) / (ns - 1)
    assert np.abs(cov.real - np.eye(nc)).max() < 0.08
    assert np.abs(cov.imag).max() < 0.08


def test_os_constructor_rules():  # recon/oversampling.rs:217,224,229
    x = np.zeros((1, 256), np.complex64)
    assert O.os_remove(x, 256) is None
    assert O.os_remove(np.zeros((1, 256), np.complex64), 512) is None
    assert O.os_remove(np.zeros((1, 300), np.complex64), 256) is None
    assert O.os_remove(np.zeros((1, 512), np.complex64), 256) is not None


def test_os_preserves_dc_only_signal():  # recon/oversampling.rs:237
    enc, rec = 16, 8
    line = np.zeros((1, enc), np.complex64)
    line[0, enc // 2] = 10
    out = O.os_remove(line, rec)
    exp = 10 * rec / enc
    assert abs(out[0, rec // 2] - exp) < 1e-4
    others = np.delete(out[0], rec // 2)
    assert np.abs(others).max() < 1e-4


def test_pattern_r2_with_acs():  # recon/grappa.rs:578
    m = np.zeros(64, np.uint8)
    m[::2] = 1
    m[24:40] = 1
    r, s, e, _, _ = O.pattern_detect_1d(m)
    assert r == 2 and 23 <= s <= 24 and 40 <= e <= 41


def test_pattern_r3():  # recon/grappa.rs:598
    m = np.zeros(96, np.uint8)
    m[::3] = 1
    m[40:60] = 1
    r, s, e, _, _ = O.pattern_detect_1d(m)
    assert r == 3 and s <= 40 and e >= 60 and e - s <= 22


def test_pattern_rejects_fully_sampled():  # recon/grappa.rs:617
    assert O.pattern_detect_1d(np.ones(64, np.uint8)) is None


def _phantom_kspace(nc, ny, nx):  # recon/grappa.rs:555-574
    k = np.zeros((nc, 1, ny, nx), np.complex64)
    for c_ in range(nc):
        for y in range(ny):
            for x in range(nx):
                yy = np.float32(y - ny / 2.0)
                xx = np.float32(x - nx / 2.0)
                r2 = (yy * yy + xx * xx) / np.float32(ny * nx)
                base = np.exp(np.float32(-3.0) * r2)
                phase = np.float32(c_ * 0.4 + 0.02 * (yy + xx))
                amp = np.float32(1.0 + 0.3 * (c_ - nc / 2.0) * yy / ny)
                px = base * amp
                k[c_, 0, y, x] = complex(px * np.cos(phase), px * np.sin(phase))
    return k


def test_grappa_reconstructs_r2_phantom():  # recon/grappa.rs:623
    nc, ny, nx = 4, 40, 32
    truth = _phantom_kspace(nc, ny, nx)
    acs = truth[:, 0, 8:32, :]
    st, w = O.grappa_calibrate(acs, 2, 4, 5, 1e-3, "f32")
    assert st == 0
    keep = np.array([(y % 2 == 0) or (8 <= y < 32) for y in range(ny)])
    us = truth * keep[None, None, :, None]
    out = O.grappa_synthesize(w, us, 2, 4, 5, "f32")
    fill_lo, fill_hi = 3, ny - (7 - 3) - 1
    num = den = 0.0
    for y in range(ny):
        if keep[y] or y < fill_lo or y > fill_hi:
            continue
        t = truth[:, 0, y, 2:nx - 2]
        g = out[:, 0, y, 2:nx - 2]
        num += np.sum(np.abs(t - g) ** 2)
        den += np.sum(np.abs(t) ** 2)
    assert np.sqrt(num / den) < 0.1


def test_grappa_nc1_leaves_sampled_rows():  # recon/grappa.rs:717
    acs = np.ones((1, 12, 16), np.complex64)
    st, w = O.grappa_calibrate(acs, 2, 4, 5, 1e-3, "f32")
    if st != 0:
        return
    us = np.zeros((1, 1, 20, 16), np.complex64)
    us[0, 0, ::2, :] = 1
    out = O.grappa_synthesize(w, us, 2, 4, 5, "f32")
    assert np.array_equal(out[0, 0, ::2, :], us[0, 0, ::2, :])


def test_sense_unfolds_r2_phantom():  # recon/sense.rs:255
    nc, ny, nx, r = 4, 32, 24, 2
    truth = np.zeros((ny, nx), np.float32)
    truth[6:26, 4:20] = 1
    truth[14:18, 10:14] = 0.3
    maps = np.zeros((nc, ny, nx), np.complex64)
    for ci, (fy, fx) in enumerate([(0.25, 0.25), (0.25, 0.75), (0.75, 0.25), (0.75, 0.75)]):
        yy, xx = np.meshgrid(np.arange(ny) - fy * ny, np.arange(nx) - fx * nx, indexing="ij")
        mag = np.exp(-(yy ** 2 + xx ** 2) / 200.0)
        maps[ci] = mag * np.exp(1j * 0.2 * ci)
    coil = maps * truth[None]
    k = np.fft.fft(coil, axis=1)
    k[:, np.arange(ny) % r != 0, :] = 0
    aliased = np.fft.ifft(k, axis=1).astype(np.complex64)
    st, rho, _, _ = O.sense_unfold(aliased, maps, r, 1e-5)
    assert st == 0
    nrmse = np.sqrt(np.sum((truth - np.abs(rho)) ** 2) / np.sum(truth ** 2))
    assert nrmse < 0.1


def test_sense_r1_passthrough():  # recon/sense.rs:365
    nc, ny, nx = 2, 8, 6
    maps = np.ones((nc, ny, nx), np.complex64)
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    aliased = np.stack([np.sin(y * 0.1 + x * 0.05 + c_) for c_ in range(nc)]).astype(np.complex64)
    st, rho, _, _ = O.sense_unfold(aliased, maps, 1, 0.0)
    assert np.abs(rho - aliased.mean(0)).max() < 1e-4


def test_sense_gfactor_unit_orthogonal_maps():  # recon/sense.rs:406,436
    nc, ny, nx, r = 2, 8, 4, 2
    maps = np.zeros((nc, ny, nx), np.complex64)
    maps[0, : ny // 2] = 1
    maps[1, ny // 2:] = 1
    _, _, g, _ = O.sense_unfold(np.zeros_like(maps), maps, r, 0.0, gfactor=True)
    assert np.abs(g - 1).max() < 1e-5


def test_sense_nc1_divides():  # recon/sense.rs:467
    s = np.full((1, 4, 3), 2.0 + 0j, np.complex64)
    a = np.full((1, 4, 3), 3.0 - 1j, np.complex64)
    _, rho, _, _ = O.sense_unfold(a, s, 1, 0.0)
    assert np.abs(rho - a[0] / s[0]).max() < 1e-5


def test_sense_indivisible_errors():  # recon/sense.rs:80-82
    st, _, _, _ = O.sense_unfold(np.zeros((2, 10, 4), np.complex64), np.zeros((2, 10, 4), np.complex64), 3, 0.0)
    assert st == 3


def test_walsh_recovers_sensitivity_ratio():  # recon/sensitivity.rs:153
    ny = nx = 32
    yy, x1 = np.meshgrid(np.arange(ny) - ny / 2.0, np.arange(nx) - nx / 3.0, indexing="ij")
    _, x2 = np.meshgrid(np.arange(ny), np.arange(nx) - 2.0 * nx / 3.0, indexing="ij")
    s1 = np.exp(-(yy ** 2 + x1 ** 2) / 200.0)
    s2 = np.exp(-(yy ** 2 + x2 ** 2) / 200.0)
    ph = np.zeros((ny, nx))
    ph[8:24, 8:24] = 1
    coil = np.stack([s1 * ph * np.exp(0j), s2 * ph * np.exp(0.7j)]).astype(np.complex64)
    maps = O.walsh_maps(coil, 3, 8)
    m0 = np.abs(maps[0, 10:22, 10:22])
    m1 = np.abs(maps[1, 10:22, 10:22])
    ok = m0 >= 1e-4
    err = np.abs(m1[ok] / m0[ok] - (s2 / s1)[10:22, 10:22][ok]).mean()
    assert err < 0.1


def test_walsh_phase_real_positive():  # recon/sensitivity.rs:219
    phi = 0.35 * np.pi
    coil = np.full((1, 8, 8), np.exp(1j * phi), np.complex64)
    maps = O.walsh_maps(coil, 1, 6)
    assert np.all(maps.real > 0.5) and np.abs(maps.imag).max() < 1e-4


def test_fft_pair_is_unitary():  # recon/cs.rs:225
    i, j = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    x = ((i + j) + 1j * (i - j) * 0.3).astype(np.complex64)
    y = O.centred_fft2(x, False)
    assert abs(np.sum(np.abs(x) ** 2) - np.sum(np.abs(y) ** 2)) < 1e-3
    assert np.abs(O.centred_fft2(y, True) - x).max() < 1e-4


def test_cs_recovers_sparse_phantom():  # recon/cs.rs:259
    ny = nx = 16
    truth = np.zeros((ny, nx), np.complex64)
    truth[5, 4], truth[10, 11], truth[3, 12], truth[12, 3] = 1, 0.8, 0.6, 0.5
    k = O.centred_fft2(truth, False)
    rows = np.array([(i % 2 == 0) or (ny // 2 - 2 <= i < ny // 2 + 2) for i in range(ny)])
    mask = np.repeat(rows[:, None], nx, 1).astype(np.uint8)
    kzf = k * mask
    zf = O.centred_fft2(kzf, True)
    zf_err = np.linalg.norm(zf - truth)
    st, rec = O.fista_cs(kzf, mask, 200, 0.02)
    assert st == 0 and np.linalg.norm(rec - truth) < 0.8 * zf_err


def test_haar_roundtrip_and_energy():  # recon/wavelet.rs:106,124
    i, j = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    x = (i + 1j * j * 0.3).astype(np.complex64)
    assert np.abs(O.haar(O.haar(x), True) - x).max() < 1e-5
    x = np.zeros((4, 4), np.complex64)
    x[0, 0], x[1, 2], x[3, 3] = 1, 1j, 0.5 - 0.5j
    assert abs(np.sum(np.abs(O.haar(x)) ** 2) - np.sum(np.abs(x) ** 2)) < 1e-5


@pytest.mark.parametrize("case", ["sym", "regular", "pf"])
def test_pf_gate(case):  # recon/partial_fourier.rs:364,371,383
    m = np.zeros((1, 64, 4), np.uint8)
    if case == "sym":
        m[:] = 1
        assert not O.pf_detect(m, 32)
    elif case == "regular":
        m[0, ::2] = 1
        assert not O.pf_detect(m, 32)
    else:
        m[0, 8:] = 1
        assert O.pf_detect(m, 32)
