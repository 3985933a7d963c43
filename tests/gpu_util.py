"""Helpers for the GPU parity tests: device buffers via torch (plumbing only),
calls through the C ABI of liboks_hip.so."""
import ctypes as C

import numpy as np

from openkspace_amd import abi
from openkspace_amd._lib import Context, check, lib


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


def to_dev(x):
    torch = torch_dev()
    x = np.ascontiguousarray(x)
    if np.iscomplexobj(x):
        t = torch.from_numpy(x.astype(np.complex64).view(np.float32).copy()).cuda()
    else:
        t = torch.from_numpy(x.copy()).cuda()
    torch.cuda.synchronize()
    return t


def c64_from_dev(t, shape):
    torch = torch_dev()
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.complex64).reshape(shape)


def fft_centered(ctx, x, axes, inverse=True, norm=1):
    """oks_fft_centered_dev on a host array (copied in and out)."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    t = to_dev(x)
    shape = np.array(x.shape, dtype=np.int64)
    ax = np.array(axes, dtype=np.int32)
    check(ctx.ptr, lib().oks_fft_centered_dev(ctx.ptr, t.data_ptr(), shape.ctypes.data, x.ndim,
                                              ax.ctypes.data, len(axes), int(inverse), norm))
    ctx.synchronize()
    return c64_from_dev(t, x.shape)


def recon_dev(ctx, fn_name, desc, case):
    """Run a strategy with device-resident k-space/noise/image (mask on host)."""
    torch = torch_dev()
    k = to_dev(case["kspace"])
    noise = to_dev(case["noise"]) if case.get("noise") is not None else None
    mask = np.ascontiguousarray(case["mask"], dtype=np.uint8)
    nc, nz, ny, nx = case["kspace"].shape
    out = torch.zeros(nz * ny * nx, dtype=torch.float32, device="cuda")
    want_g = bool(getattr(desc, "compute_gfactor", 0))
    g = torch.zeros(nz * ny * nx, dtype=torch.float32, device="cuda") if want_g else None
    torch.cuda.synchronize()
    h = abi.Header(**case["hdr"])
    ks = abi.KSpace(k.data_ptr(), mask.ctypes.data, nc, nz, ny, nx,
                    noise.data_ptr() if noise is not None else None,
                    case["noise"].shape[1] if noise is not None else 0, h)
    img = abi.Image(out.data_ptr(), g.data_ptr() if g is not None else None, 0, 0, 0, 0)
    check(ctx.ptr, getattr(lib(), fn_name)(ctx.ptr, C.byref(desc), C.byref(ks), C.byref(img)))
    ctx.synchronize()
    flags = img.flags | ctx.take_flags()
    shape = (img.nz, img.ny, img.nx)
    n = int(np.prod(shape))
    o = out[:n].cpu().numpy().reshape(shape)
    gg = g[:n].cpu().numpy().reshape(shape) if g is not None else None
    return o, flags, gg
