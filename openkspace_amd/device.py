"""Device-resident inputs for the benchmark and the GPU tests: buffers are
allocated through the backend's own C ABI (oks_device_malloc / oks_memcpy)."""
import ctypes as C

import numpy as np

from . import abi
from ._lib import DeviceArray, check, lib


def fft_centered(ctx, x, axes, inverse=True, norm=1):
    """oks_fft_centered_dev on a host array (copied in and out)."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    d = DeviceArray.from_host(ctx, x)
    shape = np.array(x.shape, dtype=np.int64)
    ax = np.array(axes, dtype=np.int32)
    check(ctx.ptr, lib().oks_fft_centered_dev(ctx.ptr, d.addr, shape.ctypes.data, x.ndim,
                                              ax.ctypes.data, len(axes), int(inverse), norm))
    ctx.synchronize()
    out = d.download()
    d.free()
    return out


class DevCase:
    """Device-resident k-space / noise / image buffers for one synthetic case."""

    def __init__(self, ctx, case, gfactor=False):
        self.ctx = ctx
        self.case = case
        self.k = DeviceArray.from_host(ctx, np.ascontiguousarray(case["kspace"], dtype=np.complex64))
        self.noise = (DeviceArray.from_host(ctx, np.ascontiguousarray(case["noise"], dtype=np.complex64))
                      if case.get("noise") is not None else None)
        self.mask = np.ascontiguousarray(case["mask"], dtype=np.uint8)
        nc, nz, ny, nx = case["kspace"].shape
        self.out = DeviceArray(ctx, (nz * ny * nx,), np.float32)
        self.g = DeviceArray(ctx, (nz * ny * nx,), np.float32) if gfactor else None
        self.ks = abi.KSpace(self.k.addr, self.mask.ctypes.data, nc, nz, ny, nx,
                             self.noise.addr if self.noise is not None else None,
                             case["noise"].shape[1] if self.noise is not None else 0,
                             abi.Header(**case["hdr"]))
        self.img = abi.Image(self.out.addr, self.g.addr if self.g is not None else None, 0, 0, 0, 0)

    def run(self, fn_name, desc):
        check(self.ctx.ptr, getattr(lib(), fn_name)(self.ctx.ptr, C.byref(desc), C.byref(self.ks),
                                                    C.byref(self.img)))

    def result(self):
        self.ctx.synchronize()
        shape = (self.img.nz, self.img.ny, self.img.nx)
        o = self.out.download(shape)
        g = self.g.download(shape) if self.g is not None else None
        return o, g

    def free(self):
        for b in (self.k, self.noise, self.out, self.g):
            if b is not None:
                b.free()


def recon_dev(ctx, fn_name, desc, case):
    """Run a strategy with device-resident k-space/noise/image (mask on host)."""
    dc = DevCase(ctx, case, gfactor=bool(getattr(desc, "compute_gfactor", 0)))
    try:
        dc.run(fn_name, desc)
        o, g = dc.result()
        flags = dc.img.flags | ctx.take_flags()
        return o, flags, g
    finally:
        dc.free()


def place_shape(acqs, enc):
    """Sizing pass of read_kspace_with_mask (io/ismrmrd/reader.rs:235-287) on an
    acquisition table of abi.ACQ_DTYPE: the grid shape (nc, nz, ny, nx), or None
    where the reference errors (no image acquisition)."""
    a = np.ascontiguousarray(acqs, dtype=abi.ACQ_DTYPE)
    e = abi.Encoding(*[int(v) for v in enc])
    shape = np.zeros(4, np.int64)
    st = lib().oks_place_shape(a.ctypes.data, len(a), C.byref(e), shape.ctypes.data)
    if st == abi.OKS_E_MISSING_FIELD:
        return None
    check(None, st)
    return tuple(int(v) for v in shape)


def place_acquisitions(ctx, acqs, data, enc, W=None, remove_oversampling=False, phase_correct=False):
    """read_kspace_with_mask (io/ismrmrd/reader.rs:225-468) with the strategies'
    preprocess closure (recon/strategy.rs:153-161) on device-resident samples.

    acqs: host acquisition table (abi.ACQ_DTYPE); data: DeviceArray of complex64
    samples; enc: (encoded_x, encoded_y, encoded_z, recon_x, ky_center,
    kz_center); W: DeviceArray [w_nc, w_nc] or None; phase_correct: build the
    navigator corrector from the table's ACQ_IS_PHASECORR_DATA acquisitions
    and apply it.  Returns (grid DeviceArray [nc, nz, ny, nx], mask bool
    [nz, ny, nx], placed count)."""
    a = np.ascontiguousarray(acqs, dtype=abi.ACQ_DTYPE)
    shape = place_shape(a, enc)
    if shape is None:
        raise ValueError("no image acquisitions found")
    e = abi.Encoding(*[int(v) for v in enc])
    grid = DeviceArray(ctx, shape, np.complex64)
    mask = np.zeros(shape[1:], np.uint8)
    placed = C.c_int64(0)
    try:
        check(ctx.ptr, lib().oks_place_acquisitions_dev(
            ctx.ptr, a.ctypes.data, len(a), data.addr, int(np.prod(data.shape)), C.byref(e),
            W.addr if W is not None else None, W.shape[0] if W is not None else 0, int(bool(phase_correct)),
            int(bool(remove_oversampling)), grid.addr, mask.ctypes.data, C.byref(placed)))
    except Exception:
        grid.free()
        raise
    return grid, mask.astype(bool), int(placed.value)


def render_slices(ctx, vol, pct_low=0.5, pct_high=99.5, mode=0):
    """The 8-bit PNG pixels of every slice of a device volume (DeviceArray
    float32 [nz, ny, nx]) as the CLI's writers compute them
    (cli/main.rs:1095-1153).  Returns (pixels uint8 [nz, ny, nx], windows
    float32 [nz, 2])."""
    nz, ny, nx = vol.shape
    out = DeviceArray(ctx, (nz, ny, nx), np.uint8)
    win = np.zeros((nz, 2), np.float32)
    try:
        check(ctx.ptr, lib().oks_render_slices_dev(ctx.ptr, vol.addr, nz, ny, nx, float(pct_low), float(pct_high),
                                                    int(mode), out.addr, win.ctypes.data))
        return out.download(), win
    finally:
        out.free()


def nifti_header(nz, ny, nx):
    """write_nifti_volume's 352 leading bytes (cli/main.rs:1033-1073)."""
    out = np.zeros(352, np.uint8)
    check(None, lib().oks_nifti_header(nz, ny, nx, out.ctypes.data))
    return out.tobytes()
