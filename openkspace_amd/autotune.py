"""Launch-shape autotuner for the gfx950 backend.

    python -m openkspace_amd.autotune [--out openkspace_amd/tuning/gfx950.json] [--reps 5]

For every (op, shape) the library keys its table by -- "ifft_rss"
[nc, ny, nx, tx, 1] for the IFFT + RSS tail and "cs" [ny, nx] for the FISTA
passes -- the tuner times the strategy call through the C ABI with each
candidate set of launch options (explicit oks_ctx_set_option values, so the
table being written is not consulted) and keeps the fastest.  Candidates are
searched one option group at a time, starting from the built-in defaults.
Every option it sets leaves the computed bits unchanged; the tuner checks
that each candidate's image equals the default's bit for bit and drops any
that does not.  Timing: HIP events around each call on the context's stream,
the median of `reps` calls after two warm-up calls.  Inputs are random data
of the workload's shapes (timing does not depend on the values), with the
variable-density CS mask of the workloads.
"""
import argparse
import itertools
import json
import os
import time

import numpy as np

from . import abi, synth
from ._lib import Context, Event, OksError
from .device import DevCase

HERE = os.path.dirname(os.path.abspath(__file__))

# (ny, nx, nc, readout oversampling, slices per call) of the shapes tuned:
# the served stream's four (batched 16 per call), and the multi-slice
# workloads' C1, C2 and C3 slices.
SHAPES = [(256, 256, 8, 1, 16), (320, 320, 16, 1, 16), (368, 640, 15, 2, 16), (384, 384, 32, 1, 16),
          (384, 384, 32, 1, 8), (320, 320, 8, 1, 32)]

IFFT_GROUPS = [("rows_waves", [0, 8, 10, 12, 14, 16]),
               ("cols_sx", [0, 4, 8, 16]),
               ("cols_threads", [0, 256, 512])]
CS_GROUPS = [(("cs_sx", "cs_threads"), list(itertools.product([4, 8, 16, 32], [256, 512, 1024]))),
             ("cs_pipe", [0, 1]),
             ("cs_coop", [0, 1]),
             ("cs_cols_map", [0, 1, 2]),
             ("cs_upd_map", [0, 1, 2])]


def _case(ny, nx, nc, os_x, nz, kind, seed):
    rng = np.random.default_rng(seed)
    k = (rng.standard_normal((nc, nz, ny, nx)) + 1j * rng.standard_normal((nc, nz, ny, nx))).astype(np.complex64)
    if kind == "cs":
        mask = np.stack([synth.cs_mask(rng, ny, nx) for _ in range(nz)]).astype(np.uint8)
        k *= mask[None]
    else:
        mask = np.ones((nz, ny, nx), np.uint8)
    noise = (rng.standard_normal((nc, 256)) + 1j * rng.standard_normal((nc, 256))).astype(np.complex64)
    hdr = dict(encoded_x=nx, encoded_y=ny, encoded_z=1, recon_x=nx // os_x, recon_y=ny, recon_z=1, is_3d_encoding=0)
    return dict(kspace=np.ascontiguousarray(k), mask=mask, noise=noise, hdr=hdr)


def _time(ctx, dc, fn, desc, reps):
    for _ in range(2):
        dc.run(fn, desc)
    ctx.synchronize()
    ms = []
    for _ in range(reps):
        a, b = Event(ctx), Event(ctx)
        a.record()
        dc.run(fn, desc)
        b.record()
        ctx.synchronize()
        ms.append(a.elapsed_ms(b))
    return float(np.median(ms))


def tune_one(op, shape, reps, log):
    ny, nx, nc, os_x, nz = shape
    if op == "cs":
        case = _case(ny, nx, nc, os_x, nz, "cs", 7)
        fn, desc, groups = "oks_recon_cs_dev", abi.CsDesc.default(remove_oversampling=0, iters=10), CS_GROUPS
        key = [ny, nx]
    else:
        case = _case(ny, nx, nc, os_x, nz, "full", 7)
        fn, desc, groups = "oks_recon_ifft_rss_dev", abi.IfftRssDesc.default(remove_oversampling=0,
                                                                            partial_fourier=0), IFFT_GROUPS
        key = [nc, ny, nx, nx // os_x, 1]
    best = {}
    tried = {}

    def run(opts):
        ctx = Context(0)
        try:
            ctx.set_option("tuning", 0)
            for name, v in opts.items():
                ctx.set_option(name, v)
            dc = DevCase(ctx, case)
            try:
                ms = _time(ctx, dc, fn, desc, reps)
                img, _ = dc.result()
            finally:
                dc.free()
            return ms, img
        except OksError:
            return None, None
        finally:
            ctx.close()

    base_ms, base_img = run({})
    best_ms = base_ms
    for names, values in groups:
        names = names if isinstance(names, tuple) else (names,)
        for v in values:
            v = v if isinstance(v, tuple) else (v,)
            opts = dict(best, **dict(zip(names, v)))
            ms, img = run(opts)
            label = ",".join(f"{k}={opts[k]}" for k in sorted(opts))
            if ms is None or not np.array_equal(img, base_img):
                tried[label] = None
                continue
            tried[label] = round(ms, 4)
            if ms < best_ms * 0.995:
                best_ms, best = ms, opts
    log(f"{op} {key}: default {base_ms:.3f} ms -> {best_ms:.3f} ms with {best}")
    return {"op": op, "key": key, "opts": best, "ms": round(best_ms, 4), "default_ms": round(base_ms, 4),
            "slices_per_call": nz, "tried": tried}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--out", default=None)
    p.add_argument("--reps", type=int, default=5)
    p.add_argument("--ops", default="ifft_rss,cs")
    a = p.parse_args()
    ctx = Context(0)
    arch = ctx.arch()
    ctx.close()
    out = a.out or os.path.join(HERE, "tuning", f"{arch}.json")
    entries = []
    t0 = time.time()
    seen = set()
    for op in a.ops.split(","):
        for shape in SHAPES:
            k = (op, shape[:4]) if op == "ifft_rss" else (op, shape[:2])
            if k in seen:
                continue
            seen.add(k)
            entries.append(tune_one(op, shape, a.reps, lambda m: print(m, flush=True)))
    table = {"arch": arch, "tool": "python -m openkspace_amd.autotune",
             "timing": f"median of {a.reps} HIP-event-timed strategy calls after 2 warm-ups",
             "seconds": round(time.time() - t0, 1), "entries": entries}
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
