"""Benchmark of the MI355X ReconStrategy backend (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|...]

A step is one reconstruction call over one batch of synthetic k-space already
resident in HBM.  Rank 0 prints one JSON line with slices/s (whole job),
p50/p99 step latency, rel-L2 of the output against the FP32 oracle and its
FP64 twin, the roofline of the timed launch sequence and the CPU baseline
(the oracle, single-threaded, on a bounded sample of the same workload).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "slices/sec + p50/p99 latency at stated rel-L2 vs FP64; served vs plain GPU"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", default="c1")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-check", action="store_true")
    return p.parse_args()


def workload_c1():
    from openkspace_amd import abi, synth
    case = synth.make_config("c1_ifft_rss")
    nc, nz, ny, nx = case["kspace"].shape
    desc = abi.IfftRssDesc.default(remove_oversampling=0)
    tx = nx // 2
    algo_bytes = nz * (nc * ny * nx * 8 + ny * tx * 4)  # per launch sequence (16 slices)
    return dict(
        name="c1_ifft_rss",
        config={"workload": "c1_ifft_rss", "slices": nz, "coils": nc, "ky": ny, "kx": nx,
                "readout_oversampling": 2, "recon_matrix": [ny, tx], "strategy": "IfftRss",
                "remove_oversampling": False, "prewhiten": True, "noise_samples": case["noise"].shape[1]},
        case=case, desc=desc, fn="oks_recon_ifft_rss_dev", oracle="ifft_rss", slices=nz,
        bound="hbm", algo=algo_bytes, kernel="ifft_rss pass sequence (whitener + row + column)",
        cpu_unit=lambda c, z: _slice_case(c, z), dtype="c64/f32",
    )


def _slice_case(case, z):
    k = np.ascontiguousarray(case["kspace"][:, z:z + 1])
    m = np.ascontiguousarray(case["mask"][z:z + 1])
    return dict(kspace=k, mask=m, noise=case["noise"], hdr=dict(case["hdr"]))


def workload_c2():
    from openkspace_amd import abi, synth
    case = synth.make_config("c2_grappa")
    nc, nz, ny, nx = case["kspace"].shape
    desc = abi.GrappaDesc.default()
    # apply work actually performed: fills x (nx - kkx + 1) x nc x n_src complex MACs per slice
    import oracle as O
    rows = synth.grappa_mask(ny, 4, 24).astype(np.uint8)
    n_fill = len(O.grappa_fills(rows, 4, 4))
    flops = nz * 8 * n_fill * (nx - 5 + 1) * nc * nc * 4 * 5
    return dict(
        name="c2_grappa",
        config={"workload": "c2_grappa", "slices": nz, "coils": nc, "ky": ny, "kx": nx, "R": 4, "acs_lines": 24,
                "kernel": [4, 5], "ridge": 1e-3, "strategy": "GrappaRss", "prewhiten": True,
                "calibration": "split-bf16 MFMA Gram + FP64 blocked ridge solve", "apply": "split-bf16 MFMA"},
        case=case, desc=desc, fn="oks_recon_grappa_dev", oracle="grappa", slices=nz,
        bound="mfma", algo=flops, kernel="GrappaRss launch sequence (calibrate + apply + ifft/rss)",
        cpu_unit=lambda c, z: _slice_case(c, z), dtype="c64/f32 (split-bf16 MFMA, fp64 solve)",
    )


WORKLOADS = {"c1": workload_c1, "c2": workload_c2}


def cpu_baseline(wl, seconds):
    import oracle as O
    O.set_threads(1)
    case = wl["case"]
    nz = case["kspace"].shape[1]
    done, t0 = 0, time.perf_counter()
    while done < nz:
        sub = wl["cpu_unit"](case, done)
        st, _, _, _ = O.recon(wl["oracle"], wl["desc"], sub, "f32")
        assert st == 0
        done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "slices/s", "cores": 1, "kind": "port",
            "sample": f"{done} of {nz} slices of {wl['name']} through the FP32 oracle, single thread, "
                      f"{dt:.1f} s"}


def main():
    a = _args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    from openkspace_amd._lib import Context, Event
    from openkspace_amd.device import DevCase

    wl = WORKLOADS[a.workload]()
    ctx = Context(0)
    dc = DevCase(ctx, wl["case"])
    ctx.synchronize()

    def step():
        dc.run(wl["fn"], wl["desc"])

    for _ in range(a.warmup):
        step()
    ctx.synchronize()
    evs = [(Event(ctx), Event(ctx)) for _ in range(a.steps)]
    t0 = time.perf_counter()
    for s, e in evs:
        s.record()
        step()
        e.record()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([s.elapsed_ms(e) for s, e in evs])  # device time per step, ms
    slices = wl["slices"] * a.steps
    value = slices / wall
    out = {
        "metric": METRIC, "value": value, "unit": "slices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": wl["dtype"], "data": "synthetic",
        "config": wl["config"],
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean())},
    }
    avg_s = float(lat.mean()) * 1e-3
    if wl["bound"] == "hbm":
        achieved = wl["algo"] / avg_s / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": wl["kernel"],
                           "algorithmic_bytes": wl["algo"], "avg_launch_ms": avg_s * 1e3}
    else:
        achieved = wl["algo"] / avg_s / 1e12
        out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": wl["kernel"],
                           "algorithmic_flops": wl["algo"], "avg_launch_ms": avg_s * 1e3}
    if not a.no_check:
        import oracle as O
        O.set_threads(min(16, os.cpu_count() or 1))
        got, _ = dc.result()
        _, r32, _, _ = O.recon(wl["oracle"], wl["desc"], wl["case"], "f32")
        _, r64, _, _ = O.recon(wl["oracle"], wl["desc"], wl["case"], "f64")
        out["rel_l2"] = {"vs_fp32_oracle": float(np.linalg.norm(got - r32) / np.linalg.norm(r32)),
                         "vs_fp64_oracle": float(np.linalg.norm(got - r64) / np.linalg.norm(r64)),
                         "fp32_oracle_vs_fp64": float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
