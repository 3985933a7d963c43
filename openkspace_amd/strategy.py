"""Host-side mirror of the reference `ReconStrategy` plugin API over the C ABI.

    trait ReconStrategy { fn name(&self) -> &'static str;
                          fn reconstruct(&self, file: &IsmrmrdFile) -> IoResult<ImageVolume>; }
    (recon/strategy.rs:58-67)

`reconstruct` takes the in-memory equivalent of the file (k-space grid, mask,
noise scans, header fields; see include/oks.h `oks_kspace`) and returns an
`ImageVolume` with `data` [slice, y, x] float32, `strategy` and `gfactor`
(recon/strategy.rs:26-34).  The configuration fields and their defaults are
the reference strategy structs'.  Errors raise `OksError` carrying the status
code that mirrors the reference error enums.
"""
import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import abi
from ._lib import Context, OksError, check, lib


@dataclass
class ImageVolume:
    data: np.ndarray
    strategy: str
    gfactor: Optional[np.ndarray] = None
    flags: int = 0


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class _Strategy:
    _desc_cls = None
    _name = None
    _fn = None

    def __init__(self, ctx: Optional[Context] = None, **fields):
        self.desc = self._desc_cls.default(**fields)
        self.ctx = ctx

    def name(self):
        return self._name

    def reconstruct(self, case) -> ImageVolume:
        ctx = self.ctx or default_context()
        k = np.ascontiguousarray(case["kspace"], dtype=np.complex64)
        m = np.ascontiguousarray(case["mask"], dtype=np.uint8)
        noise = case.get("noise")
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.complex64)
        ks = abi.make_kspace_struct(k, m, noise, case["hdr"])
        nz, ny, nx = m.shape
        out = np.zeros(nz * ny * nx, np.float32)
        want_g = bool(getattr(self.desc, "compute_gfactor", 0))
        g = np.zeros(nz * ny * nx, np.float32) if want_g else None
        img = abi.Image(out.ctypes.data, g.ctypes.data if g is not None else None, 0, 0, 0, 0)
        fn = getattr(lib(), self._fn)
        check(ctx.ptr, fn(ctx.ptr, C.byref(self.desc), C.byref(ks), C.byref(img)))
        shape = (img.nz, img.ny, img.nx)
        n = int(np.prod(shape))
        gv = g[:n].reshape(shape) if g is not None else None
        return ImageVolume(out[:n].reshape(shape), self._name, gv, int(img.flags))


class IfftRss(_Strategy):
    """IfftRss (recon/strategy.rs:86-254)."""
    _desc_cls = abi.IfftRssDesc
    _name = "ifft-rss"
    _fn = "oks_recon_ifft_rss"


class GrappaRss(_Strategy):
    """GrappaRss (recon/strategy.rs:275-454)."""
    _desc_cls = abi.GrappaDesc
    _name = "grappa-rss"
    _fn = "oks_recon_grappa"


class SenseRss(_Strategy):
    """SenseRss (recon/strategy.rs:492-734)."""
    _desc_cls = abi.SenseDesc
    _name = "sense"
    _fn = "oks_recon_sense"


class CsRss(_Strategy):
    """CsRss (recon/strategy.rs:753-892)."""
    _desc_cls = abi.CsDesc
    _name = "cs"
    _fn = "oks_recon_cs"


__all__ = ["ImageVolume", "IfftRss", "GrappaRss", "SenseRss", "CsRss", "OksError", "Context"]
