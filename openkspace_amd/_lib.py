"""ctypes binding of liboks_hip.so (the C ABI declared in include/oks.h).

There is no fallback: if the library is missing or no HIP device is present
the calls raise.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboks_hip.so")

_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_lib = None


class OksError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oks status {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OksError(-1, f"{LIB_PATH} is not built; run python -m openkspace_amd.build")
    L = C.CDLL(LIB_PATH)
    L.oks_ctx_create.argtypes = [C.c_int, C.POINTER(_P)]
    L.oks_ctx_destroy.argtypes = [_P]
    L.oks_ctx_destroy.restype = None
    L.oks_last_error.argtypes = [_P]
    L.oks_last_error.restype = C.c_char_p
    L.oks_ctx_set_stream.argtypes = [_P, _P]
    L.oks_ctx_stream.argtypes = [_P]
    L.oks_ctx_stream.restype = _P
    L.oks_ctx_synchronize.argtypes = [_P]
    L.oks_ctx_take_flags.argtypes = [_P, C.POINTER(C.c_uint32)]
    L.oks_output_shape.argtypes = [_P, _I32, _P]
    L.oks_output_shape.restype = None
    L.oks_detect_pattern.argtypes = [_P, _I64, _I64, _I64, _P]
    for name in ("ifft_rss", "grappa", "sense", "cs"):
        for suf in ("", "_dev", "_dev_noalloc"):
            fn = getattr(L, f"oks_recon_{name}{suf}", None)
            if fn is not None:
                fn.argtypes = [_P, _P, _P, _P]
    L.oks_device_malloc.argtypes = [_P, C.c_size_t, C.POINTER(_P)]
    L.oks_device_free.argtypes = [_P, _P]
    L.oks_memcpy.argtypes = [_P, _P, _P, C.c_size_t, _I32]
    L.oks_memset.argtypes = [_P, _P, C.c_int, C.c_size_t]
    L.oks_event_create.argtypes = [_P, C.POINTER(_P)]
    L.oks_event_record.argtypes = [_P, _P]
    L.oks_event_elapsed_ms.argtypes = [_P, _P, _P, C.POINTER(C.c_float)]
    L.oks_event_destroy.argtypes = [_P, _P]
    L.oks_grappa_calibrate_dev.argtypes = [_P, _P] + [_I64] * 6 + [C.c_float, _P]
    L.oks_grappa_synthesize_dev.argtypes = [_P, _P, _P] + [_I64] * 7
    L.oks_ctx_set_option.argtypes = [_P, C.c_char_p, _I64]
    L.oks_ctx_get_option.argtypes = [_P, C.c_char_p, C.POINTER(_I64)]
    L.oks_ctx_arch.argtypes = [_P, C.c_char_p, _I32]
    L.oks_tune_lookup.argtypes = [C.c_char_p, C.c_char_p, _P, _I32, C.c_char_p, C.POINTER(_I64)]
    L.oks_fista_cs_dev.argtypes = [_P, _P, _P] + [_I64] * 5 + [C.c_float, _P]
    L.oks_sense_unfold_dev.argtypes = [_P, _P, _P] + [_I64] * 4 + [C.c_float, _P, _P]
    L.oks_sense_unfold_volume_dev.argtypes = [_P, _P, _P] + [_I64] * 5 + [C.c_float, _P, _P]
    L.oks_walsh_maps_dev.argtypes = [_P, _P] + [_I64] * 6 + [_P]
    L.oks_espirit_maps_dev.argtypes = [_P, _P] + [_I64] * 6 + [C.c_float, _I64, _P, _P]
    L.oks_fft_centered_dev.argtypes = [_P, _P, _P, _I32, _P, _I32, _I32, _I32]
    L.oks_rss_dev.argtypes = [_P, _P, _I64, _I64, _P]
    L.oks_crop_dev.argtypes = [_P, _P] + [_I64] * 6 + [_P, _P]
    L.oks_place_shape.argtypes = [_P, _I64, _P, _P]
    L.oks_place_acquisitions_dev.argtypes = [_P, _P, _I64, _P, _I64, _P, _P, _I32, _I32, _I32, _P, _P, _P]
    L.oks_os_remove_dev.argtypes = [_P, _P] + [_I64] * 3 + [_P]
    L.oks_prewhiten_calibrate_dev.argtypes = [_P, _P, _I32, _I64, _P]
    L.oks_prewhiten_apply_dev.argtypes = [_P, _P, _P, _I32, _I64]
    L.oks_ctx_stage_time.argtypes = [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_I64)]
    L.oks_server_default.argtypes = [_P]
    L.oks_server_default.restype = None
    L.oks_server_create.argtypes = [_P, C.POINTER(_P)]
    L.oks_server_destroy.argtypes = [_P]
    L.oks_server_destroy.restype = None
    L.oks_server_last_error.argtypes = [_P]
    L.oks_server_last_error.restype = C.c_char_p
    L.oks_server_now.argtypes = [_P]
    L.oks_server_now.restype = C.c_double
    L.oks_server_submit.argtypes = [_P, _P]
    L.oks_server_poll.argtypes = [_P, _P, _I32, _I32, C.POINTER(_I32)]
    L.oks_server_drain.argtypes = [_P, _I32]
    L.oks_server_get_stats.argtypes = [_P, _P]
    L.oks_request_footprint.argtypes = [_P, C.POINTER(_I64)]
    L.oks_render_slices_dev.argtypes = [_P, _P, _I64, _I64, _I64, C.c_float, C.c_float, _I32, _P, _P]
    L.oks_nifti_header.argtypes = [_I64, _I64, _I64, _P]
    _lib = L
    return L


def tune_lookup(arch, op, key, name):
    """The tuning table's value of option `name` for (op, key) on `arch`, or
    None when the table, the entry or the option is absent."""
    import numpy as np
    k = np.ascontiguousarray(key, dtype=np.int64)
    v = _I64(0)
    st = lib().oks_tune_lookup(arch.encode(), op.encode(), k.ctypes.data, len(k), name.encode(), C.byref(v))
    if st == 3:  # OKS_E_MISSING_FIELD
        return None
    if st != 0:
        raise OksError(st, "malformed tuning table")
    return int(v.value)


def check(ctx, st):
    if st != 0:
        msg = lib().oks_last_error(ctx)
        raise OksError(st, msg.decode() if msg else "")
    return st


class Context:
    """Owns one oks_ctx (a HIP stream, FFT plans and workspaces on one device)."""

    def __init__(self, device=0):
        self.ptr = _P()
        st = lib().oks_ctx_create(device, C.byref(self.ptr))
        if st != 0:
            raise OksError(st, "oks_ctx_create failed")

    def close(self):
        if self.ptr:
            lib().oks_ctx_destroy(self.ptr)
            self.ptr = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        check(self.ptr, lib().oks_ctx_set_stream(self.ptr, stream_ptr))

    @property
    def stream(self):
        return lib().oks_ctx_stream(self.ptr)

    def synchronize(self):
        check(self.ptr, lib().oks_ctx_synchronize(self.ptr))

    def set_option(self, name, value):
        check(self.ptr, lib().oks_ctx_set_option(self.ptr, name.encode(), int(value)))

    def get_option(self, name):
        v = _I64(0)
        check(self.ptr, lib().oks_ctx_get_option(self.ptr, name.encode(), C.byref(v)))
        return int(v.value)

    def arch(self):
        buf = C.create_string_buffer(64)
        check(self.ptr, lib().oks_ctx_arch(self.ptr, buf, 64))
        return buf.value.decode()

    def stage_time(self):
        """(ms, units, launches) of the stage selected by set_option("time_stage", ...)."""
        ms, units, n = C.c_double(0), C.c_double(0), _I64(0)
        check(self.ptr, lib().oks_ctx_stage_time(self.ptr, C.byref(ms), C.byref(units), C.byref(n)))
        return float(ms.value), float(units.value), int(n.value)

    def take_flags(self):
        f = C.c_uint32(0)
        check(self.ptr, lib().oks_ctx_take_flags(self.ptr, C.byref(f)))
        return int(f.value)


class DeviceArray:
    """A device allocation owned through the backend's C ABI (no other HIP stack)."""

    def __init__(self, ctx, shape, dtype):
        import numpy as np
        self.ctx = ctx
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = _P()
        check(ctx.ptr, lib().oks_device_malloc(ctx.ptr, max(self.nbytes, 1), C.byref(self.ptr)))

    @classmethod
    def from_host(cls, ctx, arr):
        import numpy as np
        arr = np.ascontiguousarray(arr)
        d = cls(ctx, arr.shape, arr.dtype)
        d.upload(arr)
        return d

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, arr):
        import numpy as np
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        check(self.ctx.ptr, lib().oks_memcpy(self.ctx.ptr, self.ptr, arr.ctypes.data, self.nbytes, 0))

    def download(self, shape=None):
        import numpy as np
        out = np.empty(self.shape, dtype=self.dtype)
        check(self.ctx.ptr, lib().oks_memcpy(self.ctx.ptr, out.ctypes.data, self.ptr, self.nbytes, 1))
        return out if shape is None else out.reshape(-1)[: int(np.prod(shape))].reshape(shape)

    def zero(self):
        check(self.ctx.ptr, lib().oks_memset(self.ctx.ptr, self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().oks_device_free(self.ctx.ptr, self.ptr)
            self.ptr = _P()

    def __del__(self):
        try:
            if self.ctx.ptr:
                self.free()
        except Exception:
            pass


class Event:
    def __init__(self, ctx):
        self.ctx = ctx
        self.ptr = _P()
        check(ctx.ptr, lib().oks_event_create(ctx.ptr, C.byref(self.ptr)))

    def record(self):
        check(self.ctx.ptr, lib().oks_event_record(self.ctx.ptr, self.ptr))

    def elapsed_ms(self, stop):
        ms = C.c_float(0)
        check(self.ctx.ptr, lib().oks_event_elapsed_ms(self.ctx.ptr, self.ptr, stop.ptr, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self.ctx.ptr and self.ptr:
                lib().oks_event_destroy(self.ctx.ptr, self.ptr)
        except Exception:
            pass
