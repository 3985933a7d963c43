"""The C ABI without a GPU: liboks_hip.so loads and exports every entry point
include/oks.h declares, the ctypes mirrors have the C layouts, and the
host-only entry points (descriptor defaults, output shape, sampling-pattern
detection, request footprint) agree with the oracle's integer rules."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from openkspace_amd import abi, synth
from openkspace_amd._lib import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oks.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(oks_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    assert os.path.exists(LIB_PATH), "build the library first (__graft_entry__.build())"
    L = C.CDLL(LIB_PATH)
    names = declared()
    assert len(names) >= 45
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


STRUCTS = {
    "oks_header": abi.Header, "oks_kspace": abi.KSpace, "oks_image": abi.Image,
    "oks_ifft_rss_desc": abi.IfftRssDesc, "oks_grappa_desc": abi.GrappaDesc, "oks_sense_desc": abi.SenseDesc,
    "oks_cs_desc": abi.CsDesc, "oks_pattern": abi.Pattern, "oks_server_config": abi.ServerConfig,
    "oks_request": abi.Request, "oks_completion": abi.Completion, "oks_server_stats": abi.ServerStats,
    "oks_acq": abi.Acq, "oks_encoding": abi.Encoding,
}


def test_acquisition_table_dtype_matches_struct():
    assert abi.ACQ_DTYPE.itemsize == C.sizeof(abi.Acq) == 40
    for name in abi.ACQ_DTYPE.names:
        assert abi.ACQ_DTYPE.fields[name][1] == getattr(abi.Acq, name).offset, name


def test_struct_layouts_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "oks.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            cf = "lambda" if f == "lambda_" else f
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {cf}));')
    lines.append('printf("max %d %d\\n", OKS_SERVER_MAX_SLOTS, OKS_SERVER_MAX_BATCH);')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            parts = ln.split()
            if parts[0] != "max":
                got[(parts[0], parts[1])] = int(parts[2])
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
    last = [ln for ln in out if ln.startswith("max")][0].split()
    assert (int(last[1]), int(last[2])) == (abi.SERVER_MAX_SLOTS, abi.SERVER_MAX_BATCH)


def test_descriptor_defaults_match_reference_defaults():
    L = lib()
    for cls, fn in ((abi.IfftRssDesc, "oks_ifft_rss_default"), (abi.GrappaDesc, "oks_grappa_default"),
                    (abi.SenseDesc, "oks_sense_default"), (abi.CsDesc, "oks_cs_default")):
        d = cls()
        getattr(L, fn).argtypes = [C.c_void_p]
        getattr(L, fn)(C.byref(d))
        ref = cls.default()
        assert bytes(d) == bytes(ref), fn


def _ks(case):
    k, m = case["kspace"], np.ascontiguousarray(case["mask"], np.uint8)
    return abi.make_kspace_struct(k, m, case["noise"], case["hdr"]), m


def test_output_shape_and_pattern_detection():
    L = lib()
    case = synth.make_case(nc=2, nz=1, ny=48, nx=64, os_x=2, kind="grappa", r=4, acs=24, seed=3, ellipses=3)
    ks, m = _ks(case)
    shape = np.zeros(3, np.int64)
    L.oks_output_shape(C.byref(ks), 1, shape.ctypes.data)
    assert tuple(shape) == (1, 48, 32)
    L.oks_output_shape(C.byref(ks), 0, shape.ctypes.data)
    assert tuple(shape) == (1, 48, 64)
    # detection on the C2 mask convention: R=4, ACS [180,205) (recon/grappa.rs:50-129)
    rows = synth.grappa_mask(384, 4, 24).astype(np.uint8)
    mask = np.ascontiguousarray(np.repeat(rows[None, :, None], 8, axis=2))
    p = abi.Pattern()
    assert L.oks_detect_pattern(mask.ctypes.data, 1, 384, 8, C.byref(p)) == 0
    assert (p.r, p.acs_start, p.acs_end) == (4, 180, 205)
    ref = O.detect_pattern(mask)
    assert ref is not None and (ref[0], ref[1], ref[2]) == (p.r, p.acs_start, p.acs_end)
    full = np.ones((1, 40, 8), np.uint8)
    assert L.oks_detect_pattern(full.ctypes.data, 1, 40, 8, C.byref(p)) == abi.OKS_E_UNSUPPORTED


def test_request_footprint_rules():
    from openkspace_amd.server import Server, footprint
    out_addr = 1  # never dereferenced: the footprint is host arithmetic
    small = synth.make_case(nc=4, nz=1, ny=64, nx=64, kind="full", seed=1, ellipses=2)
    ks, m = _ks(small)
    f_ifft = footprint(Server.request(abi.IfftRssDesc.default(), ks, out_addr))
    f_cs = footprint(Server.request(abi.CsDesc.default(), ks, out_addr))
    grid = 4 * 64 * 64 * 8
    assert f_ifft >= grid and f_cs >= f_ifft + 3 * grid
    g = synth.make_case(nc=4, nz=1, ny=64, nx=64, kind="grappa", r=4, acs=24, seed=2, ellipses=2)
    ks2, m2 = _ks(g)
    f_gr = footprint(Server.request(abi.GrappaDesc.default(), ks2, out_addr))
    n_src = 4 * 4 * 5
    assert f_gr >= f_ifft + n_src * n_src * 40
