"""Python handle on the served path (oks_server_* in include/oks.h).

The server is native: a C++ request loop with continuous batching, one worker
thread, context and HIP stream per slot.  This wrapper only builds request
structs and reads completions.
"""
import ctypes as C

from . import abi
from ._lib import OksError, lib

_P = C.c_void_p


class Server:
    def __init__(self, device=0, slots=4, max_batch=32, hbm_budget=0, max_pending=0, time_stage=0,
                 graphs=True, calib_cache_bytes=0):
        cfg = abi.ServerConfig()
        lib().oks_server_default(C.byref(cfg))
        cfg.device, cfg.slots, cfg.max_batch = device, slots, max_batch
        cfg.hbm_budget, cfg.max_pending = int(hbm_budget), int(max_pending)
        cfg.time_stage = int(time_stage)
        cfg.no_graphs = 0 if graphs else 1
        cfg.calib_cache_bytes = int(calib_cache_bytes)
        self.ptr = _P()
        st = lib().oks_server_create(C.byref(cfg), C.byref(self.ptr))
        if st != 0:
            raise OksError(st, "oks_server_create failed")
        self.config = cfg

    def _check(self, st):
        if st != 0:
            msg = lib().oks_server_last_error(self.ptr)
            raise OksError(st, msg.decode() if msg else "")

    def last_error(self):
        msg = lib().oks_server_last_error(self.ptr)
        return msg.decode() if msg else ""

    def now(self):
        return lib().oks_server_now(self.ptr)

    @staticmethod
    def request(desc, ks, out_addr, gfactor_addr=None, tag=0, not_before=0.0):
        """An oks_request over a descriptor and a device-resident KSpace struct."""
        sid = abi.STRATEGY_OF[type(desc)][0]
        return abi.Request(tag, sid, C.cast(C.pointer(desc), _P), ks, out_addr, gfactor_addr, float(not_before))

    def submit(self, req):
        self._check(lib().oks_server_submit(self.ptr, C.byref(req)))

    def poll(self, max_n=256, timeout_ms=-1):
        buf = (abi.Completion * max_n)()
        n = C.c_int32(0)
        self._check(lib().oks_server_poll(self.ptr, buf, max_n, timeout_ms, C.byref(n)))
        return [buf[i] for i in range(n.value)]

    def drain(self, timeout_ms=-1):
        self._check(lib().oks_server_drain(self.ptr, timeout_ms))

    def collect(self, count, timeout_ms=600000):
        """Wait for `count` completions; returns {tag: Completion}."""
        got = {}
        while len(got) < count:
            cs = self.poll(min(1024, count - len(got)), timeout_ms)
            if not cs:
                raise OksError(abi.OKS_E_BUSY, f"timed out with {len(got)} of {count} completions")
            for c in cs:
                got[int(c.tag)] = c
        return got

    def stats(self):
        st = abi.ServerStats()
        self._check(lib().oks_server_get_stats(self.ptr, C.byref(st)))
        return st

    def close(self):
        if self.ptr:
            lib().oks_server_destroy(self.ptr)
            self.ptr = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def footprint(req):
    b = C.c_int64(0)
    st = lib().oks_request_footprint(C.byref(req), C.byref(b))
    if st != 0:
        raise OksError(st, "oks_request_footprint failed")
    return int(b.value)
