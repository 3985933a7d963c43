"""Generate tests/golden/golden.npz: numpy FP64 golden vectors for the oracle.

Each vector is computed from numpy's own routines (pocketfft, linalg), not
from the oracle, so that the oracle is pinned against an independent
implementation:

* centred IFFT (recon/fft.rs:19-56): numpy ifftn between the reference's
  rotations (ifftshift = rotate left by ceil(n/2), fftshift = rotate left by
  floor(n/2), recon/shift.rs:11-26); for even n this is exactly
  np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x))), asserted below;
* raw un-normalised 1-D DFTs at every radix the workloads use;
* unitary centred FFT pair (recon/cs.rs:141-218): numpy norm="ortho";
* noise covariance, Cholesky and inverse (recon/prewhiten.rs:48-111);
* readout-oversampling removal (recon/oversampling.rs:122-147);
* GRAPPA weights from numpy.linalg.solve of the ridge normal equations
  (recon/grappa.rs:156-308) and their application;
* SENSE unfold / g-factor from per-pixel numpy.linalg.inv (recon/sense.rs);
* Walsh maps from numpy.linalg.eigh (recon/sensitivity.rs:37-142);
* 1-level Haar from a reshape-based transform (recon/wavelet.rs:32-75).

Run: python tests/golden/make_golden.py   (deterministic; seed 7)
"""
import os

import numpy as np

rng = np.random.default_rng(7)
G = {}


def crandn(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def ref_ifftn(x, axes):
    y = x
    for ax in axes:
        n = y.shape[ax]
        y = np.roll(y, -((n + 1) // 2), axis=ax)
    y = np.fft.ifftn(y, axes=axes)
    for ax in axes:
        n = y.shape[ax]
        y = np.roll(y, -(n // 2), axis=ax)
    return y


# --- centred IFFT -----------------------------------------------------------
for name, shape, axes in [("ifft2", (2, 3, 8, 10), (2, 3)), ("ifft2odd", (2, 5, 7), (1, 2)),
                          ("ifft3", (2, 4, 6, 8), (1, 2, 3)), ("ifft1", (3, 6, 4), (1,)),
                          ("ifft2big", (1, 1, 46, 40), (2, 3))]:
    x = crandn(*shape)
    y = ref_ifftn(x, axes)
    if all(shape[a] % 2 == 0 for a in axes):
        y2 = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x, axes=axes), axes=axes), axes=axes)
        assert np.allclose(y, y2)
    G[f"{name}_in"] = x
    G[f"{name}_out"] = y
    G[f"{name}_axes"] = np.array(axes)

for n in [2, 3, 4, 5, 8, 16, 23, 46, 160, 256, 320, 368, 384, 640]:
    x = crandn(n)
    G[f"dft{n}_in"] = x
    G[f"dft{n}_fwd"] = np.fft.fft(x)
    G[f"dft{n}_inv"] = np.fft.ifft(x) * n

# --- unitary centred FFT pair -------------------------------------------------
x = crandn(8, 12)
G["cfft2_in"] = x
G["cfft2_fwd"] = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x), norm="ortho"))
G["cfft2_inv"] = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(x), norm="ortho"))

# --- prewhitening ---------------------------------------------------------------
nc, ns = 5, 300
A = crandn(nc, nc) / np.sqrt(2 * nc)
noise = A @ crandn(nc, ns)
psi = noise @ noise.conj().T / (ns - 1)
L = np.linalg.cholesky(psi)
W = np.linalg.inv(L)
G["pw_noise"] = noise
G["pw_psi"] = psi
G["pw_L"] = L
G["pw_W"] = W

# --- oversampling removal ------------------------------------------------------
for enc, rec in [(16, 8), (640, 320), (384, 128)]:
    lines = crandn(3, enc)
    img = np.fft.fftshift(np.fft.ifft(np.fft.ifftshift(lines, axes=1), axis=1), axes=1)
    c0 = (enc - rec) // 2
    crop = img[:, c0:c0 + rec]
    out = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(crop, axes=1), axis=1), axes=1)
    G[f"os{enc}_{rec}_in"] = lines
    G[f"os{enc}_{rec}_out"] = out

# --- GRAPPA ---------------------------------------------------------------------
nc, ny_acs, nx, r, kky, kkx, ridge = 4, 24, 20, 2, 4, 5, 1e-3
acs = crandn(nc, ny_acs, nx) * np.exp(-np.linspace(-2, 2, nx) ** 2)[None, None, :]
span = (kky - 1) * r + 1
kc = kky // 2 - 1
n_ky = min(ny_acs - span, ny_acs - (kc * r + r - 1) - 1)
n_kx = nx - kkx + 1
rows_A, rows_B = [], []
for ky0 in range(n_ky):
    for kx0 in range(n_kx):
        src = acs[:, ky0: ky0 + span: r, kx0: kx0 + kkx]          # [nc, kky, kkx]
        rows_A.append(src.reshape(-1))
        tgt = acs[:, ky0 + kc * r + 1: ky0 + kc * r + r, kx0 + kkx // 2]  # [nc, r-1]
        rows_B.append(tgt.T.reshape(-1))
Am = np.array(rows_A)
Bm = np.array(rows_B)
AhA = Am.conj().T @ Am
lam = ridge * np.mean(np.real(np.diag(AhA)))
X = np.linalg.solve(AhA + lam * np.eye(AhA.shape[0]), Am.conj().T @ Bm)
Wg = X.T.reshape(r - 1, nc, -1)
G["gr_acs"] = acs
G["gr_params"] = np.array([r, kky, kkx])
G["gr_ridge"] = np.array(ridge)
G["gr_weights"] = Wg
# application on an undersampled slice
ny = 32
k = crandn(nc, 1, ny, nx)
sampled = (np.arange(ny) % r == 0)
sampled[12:20] = True
k[:, :, ~sampled, :] = 0
out = k.copy()
for ky0 in range(ny):
    if ky0 + span > ny:
        break
    if not all(sampled[ky0 + j * r] for j in range(kky)):
        continue
    for d in range(1, r):
        ty = ky0 + kc * r + d
        if sampled[ty]:
            continue
        for kx0 in range(nx - kkx + 1):
            src = k[:, 0, ky0: ky0 + span: r, kx0: kx0 + kkx].reshape(-1)
            out[:, 0, ty, kx0 + kkx // 2] = Wg[d - 1] @ src
G["gr_k_in"] = k
G["gr_k_out"] = out

# --- SENSE ----------------------------------------------------------------------
nc, ny, nx, r, ridge = 4, 12, 6, 3, 1e-3
maps = crandn(nc, ny, nx)
aliased = crandn(nc, ny, nx)
nyr = ny // r
rho = np.zeros((ny, nx), complex)
g = np.zeros((ny, nx))
for x in range(nx):
    for y0 in range(nyr):
        Cm = maps[:, y0::nyr, x][:, :r]
        chc = Cm.conj().T @ Cm + ridge * np.eye(r)
        inv = np.linalg.inv(chc)
        sol = inv @ (Cm.conj().T @ aliased[:, y0, x] * r)
        for kk in range(r):
            rho[y0 + kk * nyr, x] = sol[kk]
            g[y0 + kk * nyr, x] = np.sqrt(np.real(inv[kk, kk]) * np.real(chc[kk, kk]))
G["se_maps"] = maps
G["se_aliased"] = aliased
G["se_params"] = np.array([r])
G["se_ridge"] = np.array(ridge)
G["se_rho"] = rho
G["se_g"] = g

# --- Walsh maps (dominant eigenpair per window) ------------------------------------
nc, ny, nx, win = 3, 9, 8, 2
base = crandn(ny, nx) + 2.0
sens = crandn(nc, 1, 1) + 0.2 * crandn(nc, ny, nx)
img = sens * base[None] + 0.01 * crandn(nc, ny, nx)
out = np.zeros_like(img)
for y in range(ny):
    for x in range(nx):
        blk = img[:, max(0, y - win): y + win + 1, max(0, x - win): x + win + 1].reshape(nc, -1)
        R = blk @ blk.conj().T
        w, V = np.linalg.eigh(R)
        v = V[:, -1]
        i = np.argmax(np.abs(v))
        v = v * (np.conj(v[i]) / np.abs(v[i]))
        out[:, y, x] = v * np.sqrt(w[-1])
G["wa_img"] = img
G["wa_params"] = np.array([win])
G["wa_maps"] = out

# --- Haar -----------------------------------------------------------------------
x = crandn(6, 8)
a, b = x[0::2, 0::2], x[0::2, 1::2]
c, d = x[1::2, 0::2], x[1::2, 1::2]
top = np.concatenate([(a + b + c + d) / 2, (a - b + c - d) / 2], axis=1)
bot = np.concatenate([(a + b - c - d) / 2, (a - b - c + d) / 2], axis=1)
G["haar_in"] = x
G["haar_out"] = np.concatenate([top, bot], axis=0)

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
np.savez_compressed(out, **G)
print("wrote", out, sum(v.nbytes for v in G.values()), "bytes raw")
