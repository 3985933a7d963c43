"""Benchmark of the MI355X ReconStrategy backend (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload served|c1|c2]

served (default): the served mixed stream -- single-slice requests, shapes
uniform over {256^2 x 8, 320^2 x 16, 640 x 368 x 15 (2x readout
oversampling), 384^2 x 32}, strategies 50 % ifft-rss / 30 % GRAPPA R=4 / 20 %
CS (60 iterations), k-space resident in HBM.  A step is one pass of the
stream (all requests arriving at once) through the continuous-batching
server; `value` is served slices/s.  The line also carries p50/p99 latency
under Poisson arrivals at 0.25, 0.5 and 0.75 of the served rate, the plain
GPU path (the same requests one at a time, one launch chain per slice, k-space
resident) and the plain host-buffer path (the blocking entry points, copies
included) with their slices/s, per-strategy served and plain rates, the
alone-vs-served and oracle rel-L2 per strategy and shape, the CS roofline over
the step (41 algorithmic bytes per pixel-iteration), every request's bytes
over the step, and the CPU baselines (the FP32 oracle on one thread and on
all usable host threads, on bounded samples of the stream, with the CPU model).

c1 / c2 / c3 / c4: one multi-slice reconstruction call (ifft-rss 16 x 15 x
640 x 368, GRAPPA 8 x 32 x 384^2 R=4, CS 32 x 8 x 320^2 120 iterations, 3-D
SENSE 32 x 160 x 256^2 R=3, which falls back to IFFT + RSS) per step.
c1os: C1 with readout-oversampling removal on (the reference-exact form).
c4u: the SENSE unfold + g-factor of the C4 volume at ky = 258 (divisible by
R = 3), one launch per volume.  c1h: C1 with 6/8 partial Fourier along ky,
reconstructed by homodyne.  c1p: the C1 acquisition's 5888 readouts
placed into the k-space grid with fused whitening and oversampling removal.
render: the PNG writers' pixels of a 160 x 256 x 256 volume.  hbm_copy: the
copy-bandwidth ceiling.
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
FP64_PEAK_TFLOPS = 78.6


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", default="served")
    p.add_argument("--requests", type=int, default=512, help="served: requests per stream pass")
    p.add_argument("--slots", type=int, default=4, help="served: batches in flight")
    p.add_argument("--max-batch", type=int, default=32, help="served: requests per batch")
    p.add_argument("--pool", type=int, default=3, help="served: distinct slices per shape and strategy")
    p.add_argument("--host-requests", type=int, default=128, help="served: requests timed on the host-buffer path")
    p.add_argument("--loads", default="0.25,0.5,0.75",
                   help="served: Poisson rates as fractions of the served rate (latency sweep; 0.5 is the headline)")
    p.add_argument("--stage", default="cs_iterations", help="served: stage timed for the roofline")
    p.add_argument("--no-graphs", action="store_true", help="launch CS iterations eagerly (no HIP graph)")
    p.add_argument("--ctx-option", action="append", default=[], metavar="NAME=VALUE",
                   help="c1-c4: a launch-shape option of the context (oks_ctx_set_option)")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--cpu-threads", type=int, default=16, help="threads of the all-cores CPU baseline (capped by the "
                   "host's usable cores)")
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


def workload_c1os():
    """C1 with the reference-exact composition: readout-oversampling removal on
    (IfftRss default), so every line runs the 640 -> 320 chain and is placed
    back into the 640-wide grid before the 2-D IFFT, RSS and the crop to 320."""
    w = workload_c1()
    from openkspace_amd import abi
    w["desc"] = abi.IfftRssDesc.default(remove_oversampling=1)
    w["name"] = "c1os_ifft_rss"
    w["config"] = dict(w["config"], workload="c1os_ifft_rss", remove_oversampling=True)
    w["kernel"] = "ifft_rss pass sequence (whitener + oversampling-chain row pass + column pass)"
    return w


def workload_c1w():
    """C1's geometry at 64 channels: above the 32 channels the fused row pass
    holds, the whitening runs as a pass of its own and the row pass in two
    groups of 32 channels."""
    from openkspace_amd import abi, synth
    case = synth.make_config("c1_ifft_rss", nc=64)
    nc, nz, ny, nx = case["kspace"].shape
    tx = nx // 2
    w = workload_c1()
    w.update(name="c1w_ifft_rss", case=case, algo=nz * (nc * ny * nx * 8 + ny * tx * 4),
             kernel="ifft_rss pass sequence (whitener + whitening pass + grouped row passes + column pass)")
    w["config"] = dict(w["config"], workload="c1w_ifft_rss", coils=nc, noise_samples=case["noise"].shape[1])
    w["desc"] = abi.IfftRssDesc.default(remove_oversampling=0)
    return w


def workload_c1h():
    """C1 acquired with 6/8 partial Fourier along ky (rows below 92 of 368
    missing): ifft-rss dispatches to the homodyne reconstruction."""
    from openkspace_amd import abi, synth
    case = synth.make_config("c1_ifft_rss")
    m = case["mask"].copy()
    m[:, :92] = False
    case["mask"] = np.ascontiguousarray(m)
    case["kspace"] = np.ascontiguousarray(np.where(m[None], case["kspace"], 0).astype(np.complex64))
    nc, nz, ny, nx = case["kspace"].shape
    desc = abi.IfftRssDesc.default(remove_oversampling=0)
    tx = nx // 2
    algo_bytes = nz * (nc * ny * nx * 8 + ny * tx * 4)
    return dict(
        name="c1h_homodyne",
        config={"workload": "c1h_homodyne", "slices": nz, "coils": nc, "ky": ny, "kx": nx,
                "readout_oversampling": 2, "recon_matrix": [ny, tx], "strategy": "IfftRss",
                "partial_fourier": "rows [92, 368) acquired (6/8), homodyne", "prewhiten": True,
                "remove_oversampling": False},
        case=case, desc=desc, fn="oks_recon_ifft_rss_dev", oracle="ifft_rss", slices=nz,
        bound="hbm", algo=algo_bytes, kernel="homodyne launch sequence (whitening, kx pass, ky pass + demod + RSS)",
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


def workload_c3():
    from openkspace_amd import abi, synth
    case = synth.make_config("c3_cs")
    nc, nz, ny, nx = case["kspace"].shape
    iters = 120
    desc = abi.CsDesc.default(iters=iters, lambda_=0.01)
    # per coil-slice-iteration: read z, x, y and the mask, write x and z
    algo_bytes = nc * nz * iters * CS_BYTES_PER_PIXEL_ITER * ny * nx
    return dict(
        name="c3_cs",
        config={"workload": "c3_cs", "slices": nz, "coils": nc, "ky": ny, "kx": nx, "mask": "variable-density R=6, "
                "24x24 centre", "iterations": iters, "lambda": 0.01, "wavelet": "Haar, 1 level", "strategy": "CsRss",
                "prewhiten": True, "iteration_loop": "HIP graph"},
        case=case, desc=desc, fn="oks_recon_cs_dev", oracle="cs", slices=nz,
        bound="hbm", algo=algo_bytes, kernel="CsRss launch sequence (FISTA graph + RSS)",
        cpu_unit=lambda c, z: _slice_case(c, z), dtype="c64/f32",
    )


def workload_c4():
    from openkspace_amd import abi, synth
    case = synth.make_config("c4_sense3d")
    nc, nz, ny, nx = case["kspace"].shape
    desc = abi.SenseDesc.default(map_source=abi.MAPS_ESPIRIT, espirit_kernel=6, espirit_threshold=0.02,
                                 espirit_iters=30, ridge=1e-4, compute_gfactor=1)
    # Ny = 256 is not divisible by R = 3: the reference falls back to the 3-D
    # IFFT + RSS (recon/strategy.rs:631-647).  Bytes: read the grid, write the volume.
    algo_bytes = nc * nz * ny * nx * 8 + nz * ny * nx * 4
    return dict(
        name="c4_sense3d",
        config={"workload": "c4_sense3d", "volumes": 1, "coils": nc, "kx": nx, "ky": ny, "kz": nz, "R": 3,
                "acs_lines": 32, "strategy": "SenseRss", "maps": "ESPIRiT (kernel 6, threshold 0.02, 30 it)",
                "ridge": 1e-4, "gfactor": True, "path_taken": "fallback 3-D IFFT + RSS (Ny % R != 0)"},
        case=case, desc=desc, fn="oks_recon_sense_dev", oracle="sense", slices=nz,
        bound="hbm", algo=algo_bytes,
        kernel="SenseRss launch sequence (whitened kx IFFT flagging zero rows, kz IFFT of the live ky lines, "
               "ky IFFT + RSS)",
        cpu_unit=None, dtype="c64/f32",
    )


def workload_c4e():
    """C4 with a ky count divisible by R = 3 (258 instead of 256), so SenseRss
    does not fall back: kz IFFT, ESPIRiT maps of every kz partition (kernel 6,
    threshold 0.02, 30 iterations, 32 coils, recon/espirit.rs:52-282), the
    aliased 2-D IFFTs and the per-pixel FP64 R x R unfold with g-factor
    (recon/strategy.rs:542-733).  No oracle leg: the FP32 restatement's
    ESPIRiT takes minutes per kz partition on the host; GPU parity of this
    composition is tested at reduced sizes."""
    from openkspace_amd import abi, synth
    case = synth.make_config("c4_sense3d", ny=258)
    nc, nz, ny, nx = case["kspace"].shape
    desc = abi.SenseDesc.default(map_source=abi.MAPS_ESPIRIT, espirit_kernel=6, espirit_threshold=0.02,
                                 espirit_iters=30, ridge=1e-4, compute_gfactor=1)
    algo_bytes = nc * nz * ny * nx * 8 + nz * ny * nx * 4
    return dict(
        name="c4e_sense3d",
        config={"workload": "c4e_sense3d", "volumes": 1, "coils": nc, "kx": nx, "ky": ny, "kz": nz, "R": 3,
                "acs_lines": 32, "strategy": "SenseRss", "maps": "ESPIRiT (kernel 6, threshold 0.02, 30 it)",
                "ridge": 1e-4, "gfactor": True, "path_taken": "maps + unfold (Ny % R == 0)"},
        case=case, desc=desc, fn="oks_recon_sense_dev", oracle=None, slices=nz,
        bound="hbm", algo=algo_bytes, kernel="SenseRss launch sequence (kz IFFT, ESPIRiT, 2-D IFFT, unfold)",
        cpu_unit=None, dtype="c64/f32 (ESPIRiT and the R x R solves fp64)",
    )


WORKLOADS = {"c1": workload_c1, "c2": workload_c2, "c3": workload_c3, "c4": workload_c4, "c1h": workload_c1h,
             "c1os": workload_c1os, "c1w": workload_c1w, "c4e": workload_c4e}

# HBM bytes the dominant kernels move per algorithmic byte, measured with the
# FETCH_SIZE / WRITE_SIZE counters scaled by a copy kernel of known size
# (profiles/hbm_traffic_counters.md): CS 57.5 B per pixel-iteration (33
# algorithmic), ifft-rss 60.8 MB per 15 x 368 x 640 slice (28.73 MB algorithmic),
# the 3-D tail of a 32 x 160 x 256 x 256 volume with R = 3 ky lines 7.175 GB
# (2.726 GB algorithmic; the row pass reads every row, the later passes the
# live ky lines only).
TRAFFIC_PER_ALGO_BYTE = {"c1": 60.8e6 / 28.73e6, "c3": 57.5 / 41.0, "served": 57.5 / 41.0,
                         "c4": 7.175e9 / 2.7263e9}
# FISTA's algorithmic bytes per coil-pixel and iteration: read z, x, the
# measured k-space and the mask, write x and z (recon/cs.rs:99-136)
CS_BYTES_PER_PIXEL_ITER = 5 * 8 + 1

# The stage whose launches each multi-slice workload's roofline is taken on.
WORKLOAD_STAGE = {"c1": "ifft_rss", "c1os": "ifft_rss", "c1w": "ifft_rss", "c4": "ifft_rss", "c2": "grappa_apply",
                  "c3": "cs_iterations", "c4e": "espirit"}
STAGE_KERNELS = {"ifft_rss": "IFFT + crop + RSS tail: k_rows_pw* (+ k_cols_reg_c along kz for 3-D) + "
                             "k_cols_rss_reg (+ k_rss_finalize)",
                 "grappa_apply": "GRAPPA synthesis GEMM k_grappa_synth (split-bf16 MFMA, one workgroup per "
                                 "source row group and kx tile)",
                 "cs_iterations": "FISTA iteration pair k_cs_cols_reg + k_cs_update, one HIP graph per call",
                 "espirit": "ESPIRiT maps (k_espirit_gram, k_espirit_eig, k_espirit_q, k_espirit_voxel): FP64 "
                            "vector arithmetic, against the FP64 peak"}

# Stages a roofline can be taken on: (oks_stage id, bound, unit, peak, scale).
STAGES = {
    "ifft_rss": (1, "hbm", "GB/s", HBM_PEAK_GBS, 1e9),
    "grappa_apply": (3, "mfma", "TFLOP/s", BF16_PEAK_TFLOPS, 1e12),
    "cs_iterations": (4, "hbm", "GB/s", HBM_PEAK_GBS, 1e9),
    "espirit": (6, "mfma", "TFLOP/s", FP64_PEAK_TFLOPS, 1e12),
}


def serve_descs():
    """Descriptors of the served stream: reference defaults, CS at 60
    iterations, readout-oversampling removal off (the x crop to the recon
    matrix is the fused crop of the IFFT)."""
    from openkspace_amd import abi
    return {"ifft_rss": abi.IfftRssDesc.default(remove_oversampling=0),
            "grappa": abi.GrappaDesc.default(remove_oversampling=0),
            "cs": abi.CsDesc.default(remove_oversampling=0, iters=60)}


def run_hbm_copy(a):
    """Copy-bandwidth ceiling and the calibration kernel of the byte counters:
    k_prewhiten with one channel and W = [1] reads and writes 64 Mi complex64
    (512 MiB each way) per launch."""
    import ctypes as C
    from openkspace_amd._lib import Context, DeviceArray, Event, check, lib
    M = 64 * 1024 * 1024
    ctx = Context(0)
    buf = DeviceArray(ctx, (M,), np.complex64)
    buf.zero()
    w = DeviceArray.from_host(ctx, np.ones(1, np.complex64))
    L = lib()
    for _ in range(a.warmup):
        check(ctx.ptr, L.oks_prewhiten_apply_dev(ctx.ptr, w.addr, buf.addr, 1, M))
    e0, e1 = Event(ctx), Event(ctx)
    e0.record()
    for _ in range(a.steps):
        check(ctx.ptr, L.oks_prewhiten_apply_dev(ctx.ptr, w.addr, buf.addr, 1, M))
    e1.record()
    ctx.synchronize()
    ms = e0.elapsed_ms(e1) / a.steps
    gbs = 2 * 8 * M / (ms * 1e-3) / 1e9
    print(json.dumps({"metric": "copy bandwidth (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
                      "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
                      "config": {"workload": "hbm_copy", "bytes_read_per_launch": 8 * M,
                                 "bytes_written_per_launch": 8 * M}}))
    buf.free()
    w.free()
    ctx.close()


def c4u_volume(nc=32, nz=160, ny=258, nx=256, r=3, seed=None):
    """The C4 acquisition at a divisible ky count (258 = 3 x 86): per kz the
    coil images of a 10-ellipse phantom folded R times along ky (the aliased
    image the unfold reads, rows 0..ny/R, repeated periodically), and the
    smooth coil sensitivities used as the maps."""
    from openkspace_amd import synth
    rng = np.random.default_rng(synth.SEED_BASE + 3 if seed is None else seed)
    maps2 = synth.coil_maps(rng, nc, ny, nx).astype(np.complex64)
    obj = synth._ellipses(rng, ny, nx, 1, 10).astype(np.float32)
    coil = maps2 * obj[None]
    nr = ny // r
    fold = sum(coil[:, k * nr:(k + 1) * nr] for k in range(r))
    al2 = np.concatenate([fold] * r, axis=1).astype(np.complex64)
    scale = (1.0 + 0.25 * np.cos(np.arange(nz) * 2 * np.pi / nz)).astype(np.float32)
    aliased = al2[:, None] * scale[None, :, None, None]
    maps = np.broadcast_to(maps2[:, None], (nc, nz, ny, nx))
    return np.ascontiguousarray(aliased), np.ascontiguousarray(maps)


def run_c4u(a):
    """SENSE unfold of the C4 volume at a divisible ky count: 32 coils x 160
    kz x 258 ky x 256 kx, R = 3, ridge 1e-4, g-factor on, the R x R solves in
    FP64, one launch per volume with maps and aliased data resident in HBM.
    The ESPIRiT maps feeding it are not part of this line (the maps are the
    generator's coil sensitivities)."""
    from openkspace_amd._lib import Context, DeviceArray, Event, check, lib
    nc, nz, ny, nx, r, ridge = 32, 160, 258, 256, 3, 1e-4
    aliased, maps = c4u_volume(nc, nz, ny, nx, r)
    ctx = Context(0)
    da, dm = DeviceArray.from_host(ctx, aliased), DeviceArray.from_host(ctx, maps)
    dr = DeviceArray(ctx, (nz, ny, nx), np.complex64)
    dg = DeviceArray(ctx, (nz, ny, nx), np.float32)
    L = lib()

    def step():
        check(ctx.ptr, L.oks_sense_unfold_volume_dev(ctx.ptr, da.addr, dm.addr, nc, nz, ny, nx, r, C.c_float(ridge),
                                                      dr.addr, dg.addr))

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
    lat = np.array([s.elapsed_ms(e) for s, e in evs])
    # per kz: aliased rows 0..ny/R and the maps read, rho (complex) and g written
    algo = nz * (nc * (ny // r) * nx * 8 + nc * ny * nx * 8 + ny * nx * 8 + ny * nx * 4)
    avg_s = float(lat.mean()) * 1e-3
    out = {
        "metric": METRIC, "value": nz * a.steps / wall, "unit": "slices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64/f32 (R x R solves f64)", "data": "synthetic",
        "config": {"workload": "c4u_sense_unfold", "volumes": 1, "coils": nc, "kx": nx, "ky": ny, "kz": nz, "R": r,
                   "ridge": ridge, "gfactor": True,
                   "note": "C4 at a ky count divisible by R (258); unfold + g-factor only, maps given"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean())},
        "roofline": {"bound": "hbm", "achieved": algo / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": algo / avg_s / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sense_unfold<3> (one launch per volume)", "algorithmic_bytes": algo,
                     "avg_launch_ms": avg_s * 1e3},
    }
    if not a.no_check or not a.no_cpu:
        import oracle as O
        rho, g = dr.download(), dg.download()
    if not a.no_check:
        errs = {}
        for z in (0, nz // 2):
            _, o32, g32, _ = O.sense_unfold(aliased[:, z], maps[:, z], r, ridge, True, "f32")
            _, o64, g64, _ = O.sense_unfold(aliased[:, z], maps[:, z], r, ridge, True, "f64")
            errs[f"kz={z}"] = {
                "rho_vs_fp32_oracle": float(np.linalg.norm(rho[z] - o32) / np.linalg.norm(o32)),
                "rho_vs_fp64_oracle": float(np.linalg.norm(rho[z] - o64) / np.linalg.norm(o64)),
                "g_vs_fp64_oracle": float(np.linalg.norm(g[z] - g64) / np.linalg.norm(g64)),
                "fp32_oracle_vs_fp64": float(np.linalg.norm(o32 - o64) / np.linalg.norm(o64))}
        out["rel_l2"] = errs
    if not a.no_cpu:
        O.set_threads(1)
        done, t0 = 0, time.perf_counter()
        while done < nz:
            O.sense_unfold(aliased[:, done], maps[:, done], r, ridge, True, "f32")
            done += 1
            if time.perf_counter() - t0 > a.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / dt, "unit": "slices/s", "cores": 1, "kind": "port",
                               "sample": f"{done} of {nz} kz planes through the FP32 oracle unfold + g-factor, "
                                         f"single thread, {dt:.1f} s"}
    for d in (da, dm, dr, dg):
        d.free()
    ctx.close()
    print(json.dumps(out))


def run_c1p(a):
    """Readout placement of the C1 acquisition: 16 slices x 368 readouts of 15
    coils x 640 samples (2x readout oversampling), in file order (one readout
    after another, channel-major), whitened by the device-calibrated noise
    whitener and cut to 320 samples by the oversampling removal as they are
    placed into the [15, 16, 368, 640] grid (read_kspace_with_mask with the
    strategies' preprocess closure).  Samples and grid resident in HBM; the
    acquisition table is host memory, as the header stream of a file is."""
    from openkspace_amd import abi, synth
    from openkspace_amd._lib import Context, DeviceArray, Event, check, lib
    case = synth.make_config("c1_ifft_rss")
    nc, nz, ny, nx = case["kspace"].shape
    rec = nx // 2
    acqs, data, enc = synth.acquisitions(case["kspace"], case["mask"], slice_axis=True, recon_x=rec)
    noise = case["noise"]
    e = abi.Encoding(*enc)
    ctx = Context(0)
    for opt in a.ctx_option:
        name, val = opt.split("=")
        ctx.set_option(name, int(val))
    dd = DeviceArray.from_host(ctx, data)
    dn = DeviceArray.from_host(ctx, noise)
    dw = DeviceArray(ctx, (nc, nc), np.complex64)
    L = lib()
    check(ctx.ptr, L.oks_prewhiten_calibrate_dev(ctx.ptr, dn.addr, nc, noise.shape[1], dw.addr))
    grid = DeviceArray(ctx, (nc, nz, ny, nx), np.complex64)
    mask = np.zeros((nz, ny, nx), np.uint8)
    placed = C.c_int64(0)

    def step():
        check(ctx.ptr, L.oks_place_acquisitions_dev(ctx.ptr, acqs.ctypes.data, len(acqs), dd.addr, data.size,
                                                     C.byref(e), dw.addr, nc, 0, 1, grid.addr, mask.ctypes.data,
                                                     C.byref(placed)))

    for _ in range(a.warmup):
        step()
    ctx.synchronize()
    ctx.set_option("time_stage", abi.STAGE_PLACE)
    evs = [(Event(ctx), Event(ctx)) for _ in range(a.steps)]
    t0 = time.perf_counter()
    for s, ev in evs:
        s.record()
        step()
        ev.record()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kms, kbytes, nl = ctx.stage_time()
    ctx.set_option("time_stage", abi.STAGE_NONE)
    lat = np.array([s.elapsed_ms(ev) for s, ev in evs])
    assert placed.value == nz * ny
    avg_s = kms / nl * 1e-3
    algo = kbytes / nl
    out = {
        "metric": METRIC, "value": nz * a.steps / wall, "unit": "slices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64/f32", "data": "synthetic",
        "config": {"workload": "c1_placement", "slices": nz, "coils": nc, "readouts": len(acqs),
                   "samples_per_readout": nx, "recon_x": rec, "prewhiten": True, "remove_oversampling": True,
                   "note": "readouts -> [coil, slice, ky, kx] grid with fused whitening and oversampling "
                           "removal; the grid zero fill is in the step, not in the kernel time"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean())},
        "roofline": {"bound": "hbm", "achieved": algo / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": algo / avg_s / 1e9 / HBM_PEAK_GBS, "traffic": None, "kernel": "k_place",
                     "algorithmic_bytes": algo, "avg_launch_ms": avg_s * 1e3},
    }
    if not a.no_check or not a.no_cpu:
        import oracle as O
        W = dw.download()
    if not a.no_check:
        g = grid.download()
        errs = {}
        for z in (0, nz // 2):
            sel = acqs[acqs["slice"] == z].copy()
            sel["slice"] = 0
            sub = (enc[0], enc[1], 1, enc[3], enc[4], 0)
            r32, m32, _ = O.place(sel, data, sub, W=W, remove_os=True, prec="f32")
            r64 = O.place(sel, data, sub, W=W, remove_os=True, prec="f64")[0]
            errs[f"slice={z}"] = {
                "vs_fp32_oracle": float(np.linalg.norm(g[:, z] - r32[:, 0]) / np.linalg.norm(r32)),
                "vs_fp64_oracle": float(np.linalg.norm(g[:, z] - r64[:, 0]) / np.linalg.norm(r64)),
                "fp32_oracle_vs_fp64": float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64)),
                "mask_exact": bool(np.array_equal(mask[z].astype(bool), m32[0]))}
        out["rel_l2"] = errs
    if not a.no_cpu:
        O.set_threads(1)
        done, t0 = 0, time.perf_counter()
        while done < nz:
            sel = acqs[acqs["slice"] == done].copy()
            sel["slice"] = 0
            O.place(sel, data, (enc[0], enc[1], 1, enc[3], enc[4], 0), W=W, remove_os=True, prec="f32")
            done += 1
            if time.perf_counter() - t0 > a.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / dt, "unit": "slices/s", "cores": 1, "kind": "port",
                               "sample": f"{done} of {nz} slices placed by the FP32 oracle with whitening and "
                                         f"oversampling removal, single thread, {dt:.1f} s"}
    for d in (dd, dn, dw, grid):
        d.free()
    ctx.close()
    print(json.dumps(out))


def run_render(a):
    """The PNG writers' pixels (percentile window 0.5-99.5, gamma 0.9) of a
    160 x 256 x 256 magnitude volume resident in HBM: per slice two exact
    order statistics by radix select, then the byte map; one call per step."""
    from openkspace_amd._lib import Context, DeviceArray, Event, check, lib
    nz, ny, nx = 160, 256, 256
    rng = np.random.default_rng(11)
    vol = np.abs(rng.standard_normal((nz, ny, nx)) + 1j * rng.standard_normal((nz, ny, nx))).astype(np.float32)
    ctx = Context(0)
    dv = DeviceArray.from_host(ctx, vol)
    do = DeviceArray(ctx, (nz, ny, nx), np.uint8)
    L = lib()

    def step():
        check(ctx.ptr, L.oks_render_slices_dev(ctx.ptr, dv.addr, nz, ny, nx, C.c_float(0.5), C.c_float(99.5), 0,
                                                do.addr, None))

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
    lat = np.array([s.elapsed_ms(e) for s, e in evs])
    algo = nz * ny * nx * (4 + 1)  # the volume read once, the bytes written once
    avg_s = float(lat.mean()) * 1e-3
    out = {
        "metric": METRIC, "value": nz * a.steps / wall, "unit": "slices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 -> u8", "data": "synthetic",
        "config": {"workload": "render_png_pixels", "slices": nz, "ky": ny, "kx": nx, "pct_low": 0.5,
                   "pct_high": 99.5, "note": "write_png_windowed's pixels for every slice; the PNG container "
                                             "stays with the host"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean())},
        "roofline": {"bound": "hbm", "achieved": algo / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": algo / avg_s / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_select + k_window (one call)", "algorithmic_bytes": algo,
                     "avg_launch_ms": avg_s * 1e3},
    }
    if not a.no_check or not a.no_cpu:
        import oracle as O
    if not a.no_check:
        got = do.download()
        ref, _ = O.render_slices(vol[:8])
        out["bytes_equal_first_8_slices"] = float(np.mean(got[:8] == ref))
    if not a.no_cpu:
        O.set_threads(1)
        done, t0 = 0, time.perf_counter()
        while done < nz:
            O.render_slices(vol[done:done + 1])
            done += 1
            if time.perf_counter() - t0 > a.cpu_seconds:
                break
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / dt, "unit": "slices/s", "cores": 1, "kind": "port",
                               "sample": f"{done} of {nz} slices through the oracle's sort-based writer, "
                                         f"single thread, {dt:.1f} s"}
    dv.free()
    do.free()
    ctx.close()
    print(json.dumps(out))


def _pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def run_served(a):
    from openkspace_amd import abi, synth
    from openkspace_amd._lib import Context, DeviceArray, check, lib
    from openkspace_amd.server import Server

    ctx = Context(0)
    descs = serve_descs()
    pool = synth.serve_pool(a.pool)
    dev = {}
    for key, cases in pool.items():
        lst = []
        for case in cases:
            nc, _, ny, nx = case["kspace"].shape
            k = DeviceArray.from_host(ctx, case["kspace"])
            n = DeviceArray.from_host(ctx, case["noise"])
            m = np.ascontiguousarray(case["mask"], np.uint8)
            ks = abi.KSpace(k.addr, m.ctypes.data, nc, 1, ny, nx, n.addr, case["noise"].shape[1],
                            abi.Header(**case["hdr"]))
            lst.append(dict(ks=ks, bufs=(k, n), m=m, case=case))
        dev[key] = lst
    seq = synth.serve_requests(a.requests)
    seen = {}
    spec = []
    for key in seq:
        j = seen.get(key, 0)
        seen[key] = j + 1
        spec.append((key, j % a.pool))
    max_out = max(ny * nx for ny, nx, _, _ in synth.SERVE_SHAPES)
    outs = [DeviceArray(ctx, (max_out,), np.float32) for _ in spec]
    N = len(spec)

    def submit_all(srv, arrivals, idx):
        for n, i in enumerate(idx):
            key, j = spec[i]
            d = dev[key][j]
            srv.submit(Server.request(descs[key[1]], d["ks"], outs[i].addr, None, i, arrivals[n]))

    def one_pass(srv, arrivals, idx=None):
        idx = list(range(N)) if idx is None else idx
        submit_all(srv, arrivals, idx)
        got = srv.collect(len(idx))
        bad = [c for c in got.values() if c.status != 0]
        if bad:
            kinds = sorted({(spec[int(c.tag)][0][1], synth.SERVE_SHAPES[spec[int(c.tag)][0][0]][:3], int(c.batch))
                            for c in bad})
            raise RuntimeError(f"{len(bad)} requests failed, status {bad[0].status}: {srv.last_error()}; "
                               f"(strategy, shape, batch) of failures: {kinds[:12]}")
        return got

    stage_id, bound, unit, peak, scale = STAGES[a.stage]
    srv = Server(slots=a.slots, max_batch=a.max_batch, time_stage=stage_id, graphs=not a.no_graphs)
    for _ in range(a.warmup):
        one_pass(srv, [srv.now() + 0.005] * N)
    st0 = srv.stats()
    walls, batch_sizes, sat_lat = [], [], []
    for _ in range(a.steps):
        t0 = srv.now() + 0.005
        got = one_pass(srv, [t0] * N)
        walls.append(max(c.t_done for c in got.values()) - t0)
        batch_sizes += [c.batch for c in got.values()]
        sat_lat += [(c.t_done - c.t_arrival) * 1e3 for c in got.values()]
    st1 = srv.stats()
    value = N * a.steps / sum(walls)
    # one request per (shape, strategy) for the checks: its served image now
    check_idx = {}
    for i, (key, j) in enumerate(spec):
        check_idx.setdefault(key, i)
    served_img = {key: (outs[i].download(), got[i]) for key, i in check_idx.items()}
    # latency under Poisson arrivals at fractions of the served rate
    sweep = []
    for ld in [float(x) for x in a.loads.split(",") if x]:
        rate = ld * value
        t0 = srv.now() + 0.01
        arr = t0 + synth.poisson_arrivals(N, rate)
        got = one_pass(srv, list(arr))
        p_lat = [(c.t_done - c.t_arrival) * 1e3 for c in got.values()]
        sweep.append({"load": ld, "rate_per_s": rate, "p50": _pct(p_lat, 50), "p99": _pct(p_lat, 99),
                      "mean": float(np.mean(p_lat)),
                      "achieved_slices_per_s": N / (max(c.t_done for c in got.values()) - t0)})
    # served passes of one strategy's requests at a time (for the per-strategy ratios)
    by_strat = {}
    for i, (key, _) in enumerate(spec):
        by_strat.setdefault(key[1], []).append(i)
    served_by = {}
    for name, idx in by_strat.items():
        t0 = srv.now() + 0.005
        g = one_pass(srv, [t0] * len(idx), idx)
        served_by[name] = len(idx) / (max(c.t_done for c in g.values()) - t0)
    stats = srv.stats()
    srv.close()
    # the same stream with the resident calibration cache on: every request
    # of a pass repeats one of the pool's slices, so after the warm-up pass
    # GRAPPA's weights come from HBM instead of a fresh calibration (a
    # series whose reference scan is shared); reported beside, not as, value
    csrv = Server(slots=a.slots, max_batch=a.max_batch, graphs=not a.no_graphs, calib_cache_bytes=4 << 30)
    one_pass(csrv, [csrv.now() + 0.005] * N)
    c0 = csrv.stats()
    t0 = csrv.now() + 0.005
    gc = one_pass(csrv, [t0] * N)
    c1 = csrv.stats()
    cache_value = N / (max(c.t_done for c in gc.values()) - t0)
    csrv.close()
    # the plain GPU path: one request per launch chain, one at a time
    plain = Server(slots=1, max_batch=1, graphs=not a.no_graphs)
    one_pass(plain, [plain.now() + 0.005] * N)  # warm-up pass: plans, workspaces, graphs
    t0 = plain.now() + 0.005
    gotp = one_pass(plain, [t0] * N)
    plain_wall = max(c.t_done for c in gotp.values()) - t0
    plain_value = N / plain_wall
    plain_service = [(c.t_done - c.t_start) * 1e3 for c in gotp.values()]
    plain_by = {name: len(idx) / sum(gotp[i].t_done - gotp[i].t_start for i in idx)
                for name, idx in by_strat.items()}
    plain.close()
    plain_img = {key: outs[i].download() for key, i in check_idx.items()}
    # the plain host-buffer path: the blocking oks_recon_* entry points, k-space
    # copied in and the image out per call, one request at a time (what a
    # ReconStrategy impl calling the library per slice does), on a sample
    host_fn = {"ifft_rss": "oks_recon_ifft_rss", "grappa": "oks_recon_grappa", "cs": "oks_recon_cs"}
    host_n = min(N, a.host_requests)
    hbufs = {}
    for key, j in spec[:host_n]:
        case = dev[key][j]["case"]
        hbufs.setdefault((key, j), (np.ascontiguousarray(case["kspace"], np.complex64),
                                    np.ascontiguousarray(case["noise"], np.complex64),
                                    dev[key][j]["m"], np.zeros(max_out, np.float32)))

    def host_call(key, j):
        k, nz_, m, o = hbufs[(key, j)]
        ks = abi.make_kspace_struct(k, m, nz_, dev[key][j]["case"]["hdr"])
        img = abi.Image(o.ctypes.data, None, 0, 0, 0, 0)
        check(ctx.ptr, getattr(lib(), host_fn[key[1]])(ctx.ptr, C.byref(descs[key[1]]), C.byref(ks),
                                                        C.byref(img)))

    for key, j in spec[:min(host_n, 32)]:
        host_call(key, j)  # warm-up: plans, workspaces, graphs
    t0h = time.perf_counter()
    for key, j in spec[:host_n]:
        host_call(key, j)
    host_value = host_n / (time.perf_counter() - t0h)

    d_ms = st1.stage_ms - st0.stage_ms
    d_units = st1.stage_units - st0.stage_units
    d_n = st1.stage_launches - st0.stage_launches
    avg_s = d_ms * 1e-3 / max(d_n, 1)
    achieved = (d_units / max(d_n, 1)) / avg_s / scale if d_n else None
    # algorithmic HBM bytes of one pass of the stream: CS 41 B per
    # pixel-iteration (read z, x, y and the mask, write x and z), ifft-rss and
    # GRAPPA their k-space read once and their image written once
    cs_iters = descs["cs"].iters
    step_cs_bytes, step_bytes = 0.0, 0.0
    for key, j in spec:
        ny, nx, nc, os_x = synth.SERVE_SHAPES[key[0]]
        img_b = ny * (nx // os_x) * 4
        if key[1] == "cs":
            b = CS_BYTES_PER_PIXEL_ITER * ny * nx * nc * cs_iters + img_b
            step_cs_bytes += b
        else:
            b = nc * ny * nx * 8 + img_b
        step_bytes += b
    step_s = sum(walls) / a.steps
    shapes = [f"{ny}x{nx}x{nc}" + (" (2x readout oversampling)" if os_x > 1 else "")
              for ny, nx, nc, os_x in synth.SERVE_SHAPES]
    out = {
        "metric": METRIC, "value": value, "unit": "slices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": sum(walls) * 1e3 / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64/f32 (GRAPPA split-bf16 MFMA, fp64 solves)",
        "data": "synthetic",
        "config": {"workload": "served_mixed_stream", "requests_per_step": N, "shapes": shapes,
                   "mix": {"ifft_rss": 0.5, "grappa_r4": 0.3, "cs_60it": 0.2}, "slots": a.slots,
                   "max_batch": a.max_batch, "distinct_slices_per_shape_and_strategy": a.pool,
                   "remove_oversampling": False, "prewhiten": True, "arrivals": "all at once per step",
                   "cs_iterations_as_hip_graphs": not a.no_graphs},
        "latency_ms": dict(next((w for w in sweep if w["load"] == 0.5), sweep[0] if sweep else {}),
                           arrivals="Poisson"),
        "latency_sweep_ms": sweep,
        "saturation_latency_ms": {"p50": _pct(sat_lat, 50), "p99": _pct(sat_lat, 99)},
        "plain_gpu": {"value": plain_value, "unit": "slices/s",
                      "service_latency_ms": {"p50": _pct(plain_service, 50), "p99": _pct(plain_service, 99)},
                      "path": "same requests one at a time (batch of one, one slot), k-space resident in HBM"},
        "served_vs_plain": value / plain_value,
        "served_with_calib_cache": {"value": cache_value, "unit": "slices/s",
                                    "calib_hits": int(c1.calib_hits - c0.calib_hits),
                                    "calib_misses": int(c1.calib_misses - c0.calib_misses),
                                    "path": "the same stream with GRAPPA weights kept resident across passes "
                                            "(4 GiB cache keyed by a hash of the whitened ACS rows); not the "
                                            "headline, which calibrates every request as the reference does"},
        "plain_host_api": {"value": host_value, "unit": "slices/s", "requests": host_n,
                           "path": "oks_recon_* with host buffers, one request at a time (k-space H2D and image "
                                   "D2H inside every call)"},
        "served_vs_plain_host_api": value / host_value,
        "per_strategy_slices_per_s": {name: {"served": served_by[name], "plain": plain_by[name],
                                             "served_vs_plain": served_by[name] / plain_by[name]}
                                      for name in by_strat},
        "batches": {"mean_requests": float(np.mean(batch_sizes)), "count": int(st1.batches - st0.batches),
                    "peak_inflight_bytes": int(stats.peak_inflight_bytes), "hbm_budget": int(stats.hbm_budget)},
    }
    units_per_launch = d_units / max(d_n, 1)
    tr = TRAFFIC_PER_ALGO_BYTE.get("served")
    cs_gbs = step_cs_bytes / step_s / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": cs_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": cs_gbs / HBM_PEAK_GBS, "traffic": step_cs_bytes * tr if tr else None,
                       "kernel": "CS iteration kernels (k_cs_cols_reg + k_cs_update), whole step",
                       "basis": "the CS requests' algorithmic bytes per pass of the stream (41 B per "
                                "pixel-iteration) over ms_per_step; the other strategies' work shares the step, "
                                "so this is a lower bound on the CS kernels' own rate",
                       "algorithmic_bytes_per_step": step_cs_bytes, "ms_per_step": step_s * 1e3}
    out["step_hbm"] = {"achieved": step_bytes / step_s / 1e9, "unit": "GB/s",
                       "frac": step_bytes / step_s / 1e9 / HBM_PEAK_GBS,
                       "basis": "every request's algorithmic bytes per pass (CS 41 B per pixel-iteration, "
                                "ifft-rss and GRAPPA k-space read once + image written once) over ms_per_step"}
    if d_n:
        out["stage_events"] = {"stage": a.stage, "bound": bound, "unit": unit,
                               "achieved": achieved, "frac": achieved / peak, "launches": int(d_n),
                               "avg_launch_ms": avg_s * 1e3, "algorithmic_units_per_launch": units_per_launch,
                               "basis": "HIP events around each launch of the stage on the slot streams; "
                                        "launches overlap on 4 slots, so the durations include contention"}
    rl = {}
    for key, i in check_idx.items():
        c = served_img[key][1]
        n = int(c.nz * c.ny * c.nx)
        s_img = served_img[key][0][:n]
        p_img = plain_img[key][:n]
        rl[f"{key[1]} {synth.SERVE_SHAPES[key[0]][0]}x{synth.SERVE_SHAPES[key[0]][1]}"] = {
            "served_vs_alone": float(np.linalg.norm(s_img - p_img) / np.linalg.norm(p_img)),
            "batch": int(c.batch)}
    if not a.no_check:
        import oracle as O
        O.set_threads(min(16, os.cpu_count() or 1))
        for key, i in check_idx.items():
            name = f"{key[1]} {synth.SERVE_SHAPES[key[0]][0]}x{synth.SERVE_SHAPES[key[0]][1]}"
            case = dev[key][spec[i][1]]["case"]
            c = served_img[key][1]
            n = int(c.nz * c.ny * c.nx)
            s_img = served_img[key][0][:n]
            _, r32, _, _ = O.recon(key[1], descs[key[1]], case, "f32")
            _, r64, _, _ = O.recon(key[1], descs[key[1]], case, "f64")
            r32, r64 = r32.reshape(-1), r64.reshape(-1)
            rl[name].update({"vs_fp32_oracle": float(np.linalg.norm(s_img - r32) / np.linalg.norm(r32)),
                             "vs_fp64_oracle": float(np.linalg.norm(s_img - r64) / np.linalg.norm(r64)),
                             "fp32_oracle_vs_fp64": float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))})
    out["rel_l2"] = rl
    if not a.no_check:
        out["rel_l2_bars"] = parity_summary(rl)
    if not a.no_cpu:
        host = cpu_host()
        jobs = [(key, dev[key][j]["case"]) for key, j in spec]
        out["cpu_baseline"] = cpu_stream_rate(jobs, descs, 1, host)
        out["cpu_baseline_all_cores"] = cpu_stream_rate(jobs, descs, a.cpu_threads, host)
    for lst in dev.values():
        for d in lst:
            for b in d["bufs"]:
                b.free()
    for o in outs:
        o.free()
    ctx.close()
    print(json.dumps(out))


def parity_summary(rl):
    """The served rows against the strategies' bars, stated where the numbers
    are: 1e-5 for the FP32 paths (ifft-rss, CS) and 1e-3 for the split-BF16
    GRAPPA, against the FP64 oracle, and against the FP32 restatement where
    the restatement itself is within the bar of FP64."""
    bars = {"ifft_rss": 1e-5, "cs": 1e-5, "grappa": 1e-3}
    rows, miss32 = {}, []
    for name, r in rl.items():
        if "vs_fp64_oracle" not in r:
            continue
        bar = bars[name.split()[0]]
        ok64 = r["vs_fp64_oracle"] <= bar
        ok32 = r["vs_fp32_oracle"] <= bar
        rows[name] = {"bar": bar, "meets_bar_vs_fp64": ok64, "meets_bar_vs_fp32_restatement": ok32}
        if not ok32:
            miss32.append(name)
    return {
        "bars": bars, "rows": rows,
        "rows_meeting_bar_vs_fp64": sum(v["meets_bar_vs_fp64"] for v in rows.values()),
        "rows_checked": len(rows),
        "rows_missing_bar_vs_fp32_restatement": miss32,
        "note": "Where a row misses the bar against the FP32 restatement, the restatement's own distance to "
                "FP64 (fp32_oracle_vs_fp64 in rel_l2) is above the bar: its FP32 noise whitener (ill-conditioned "
                "covariance) and, for CS, 60 FISTA iterations of FP32 rounding. The backend computes the small "
                "dense solves in FP64, so its distance to FP64 is the figure that meets the bar.",
    }


def cpu_host():
    """The host CPU the baselines ran on: model name, logical CPUs, usable CPUs."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_stream_rate(spec, descs, threads, host):
    """Slices/s of the FP32 oracle on the served stream's mix, on `threads` host
    threads (the oracle splits each request over coils and rows): one request
    of every (shape, strategy) pair is timed, CS at 2 and 4 iterations with the
    per-iteration cost extrapolated to the served count, and the per-pair times
    are weighted by how often the stream holds each pair."""
    import oracle as O
    from openkspace_amd import abi
    n = max(1, min(threads, host["usable_cpus"] or 1))
    O.set_threads(n)
    counts = {}
    for key, case in spec:
        counts.setdefault(key, [0, case])[0] += 1
    total, t_all = 0.0, time.perf_counter()
    for key, (cnt, case) in counts.items():
        name, d = key[1], descs[key[1]]
        if name == "cs":
            tt = []
            for it in (2, 4):
                dd = abi.CsDesc.default(remove_oversampling=d.remove_oversampling, iters=it, lambda_=d.lambda_)
                t0 = time.perf_counter()
                assert O.recon(name, dd, case, "f32")[0] == 0
                tt.append(time.perf_counter() - t0)
            slope = (tt[1] - tt[0]) / 2.0
            t = max(tt[0] - 2 * slope, 0.0) + d.iters * slope
        else:
            t0 = time.perf_counter()
            assert O.recon(name, d, case, "f32")[0] == 0
            t = time.perf_counter() - t0
        total += cnt * t
    O.set_threads(1)
    dt = time.perf_counter() - t_all
    n_req = sum(c for c, _ in counts.values())
    return {"value": n_req / total, "unit": "slices/s", "cores": n, "kind": "port",
            "sample": f"one request of each of the {len(counts)} (shape, strategy) pairs through the FP32 oracle "
                      f"on {n} thread(s), CS timed at 2 and 4 iterations and extrapolated to its "
                      f"{descs['cs'].iters}, weighted by the pairs' counts in the {n_req}-request stream; "
                      f"{dt:.1f} s of CPU work", "host": host}


def cpu_baseline(wl, seconds, threads=1, host=None):
    """The FP32 oracle on the workload's slices, on `threads` host threads (the
    oracle splits each call over slices, coils and rows), over a bounded
    sample: slices/s."""
    import oracle as O
    host = host or cpu_host()
    n = max(1, min(threads, host["usable_cpus"] or 1))
    O.set_threads(n)
    how = "single thread (the reference's execution model)" if n == 1 else f"{n} threads"
    case = wl["case"]
    try:
        if wl["cpu_unit"] is None:
            # one 3-D volume cannot be cut into slices: time the oracle on the
            # same volume with kz reduced to 8 partitions, per slice
            from openkspace_amd import synth
            sub = synth.make_config(wl["name"], nz=8)
            t0 = time.perf_counter()
            st, _, _, _ = O.recon(wl["oracle"], wl["desc"], sub, "f32")
            dt = time.perf_counter() - t0
            assert st == 0
            return {"value": 8 / dt, "unit": "slices/s", "cores": n, "kind": "port",
                    "sample": f"the same acquisition with 8 kz partitions through the FP32 oracle, {how}, "
                              f"{dt:.1f} s", "host": host}
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
        return {"value": done / dt, "unit": "slices/s", "cores": n, "kind": "port",
                "sample": f"{done} of {nz} slices of {wl['name']} through the FP32 oracle, {how}, {dt:.1f} s",
                "host": host}
    finally:
        O.set_threads(1)


def main():
    a = _args()
    if a.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only (the path has no multi-GPU form); "
                                    "--gpus must be 1"}))
        sys.exit(2)
    if a.workload == "hbm_copy":
        run_hbm_copy(a)
        return
    if a.workload == "c4u":
        run_c4u(a)
        return
    if a.workload == "c1p":
        run_c1p(a)
        return
    if a.workload == "render":
        run_render(a)
        return
    if a.workload == "served":
        if a.steps == 50 and "--steps" not in sys.argv:
            a.steps = 5
        if a.warmup == 5 and "--warmup" not in sys.argv:
            a.warmup = 1
        run_served(a)
        return
    from openkspace_amd._lib import Context, Event
    from openkspace_amd.device import DevCase

    wl = WORKLOADS[a.workload]()
    ctx = Context(0)
    if a.no_graphs:
        ctx.set_option("graphs", 0)
        wl["config"]["iteration_loop"] = "eager launches"
    for opt in a.ctx_option:
        name, val = opt.split("=")
        ctx.set_option(name, int(val))
        wl["config"].setdefault("ctx_options", {})[name] = int(val)
    dc = DevCase(ctx, wl["case"])
    ctx.synchronize()
    print(f"{wl['name']}: inputs resident", file=sys.stderr, flush=True)

    def step():
        dc.run(wl["fn"], wl["desc"])

    for i in range(a.warmup):
        step()
        ctx.synchronize()
        print(f"{wl['name']}: warm-up step {i + 1} of {a.warmup} done", file=sys.stderr, flush=True)
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
        tr = TRAFFIC_PER_ALGO_BYTE.get(a.workload)
        whole = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": achieved / HBM_PEAK_GBS, "traffic": wl["algo"] * tr if tr else None,
                 "kernel": wl["kernel"], "algorithmic_bytes": wl["algo"], "avg_launch_ms": avg_s * 1e3}
    else:
        achieved = wl["algo"] / avg_s / 1e12
        whole = {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None, "kernel": wl["kernel"],
                 "algorithmic_flops": wl["algo"], "avg_launch_ms": avg_s * 1e3}
    out["roofline"] = whole
    stage = WORKLOAD_STAGE.get(a.workload)
    if stage:
        # the dominant stage's own launches, in as many steps again after the
        # timed ones: HIP events on the context stream around each of them,
        # charged their algorithmic bytes or flops
        ctx.set_option("time_stage", STAGES[stage][0])
        for _ in range(a.steps):
            step()
        ms, units, n = ctx.stage_time()
        ctx.set_option("time_stage", 0)
        sid, bound, unit, peak, scale = STAGES[stage]
        if n:
            ach = units / (ms * 1e-3) / scale
            tr = TRAFFIC_PER_ALGO_BYTE.get(a.workload)
            out["roofline"] = {"bound": bound, "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
                               "traffic": (units / n) * tr if (tr and bound == "hbm") else None,
                               "kernel": STAGE_KERNELS[stage], "launches": n, "avg_launch_ms": ms / n,
                               "algorithmic_units_per_launch": units / n}
            out["roofline_whole_call"] = whole
    if not a.no_check and wl["oracle"]:
        import oracle as O
        O.set_threads(min(16, os.cpu_count() or 1))
        got, _ = dc.result()
        _, r32, _, _ = O.recon(wl["oracle"], wl["desc"], wl["case"], "f32")
        _, r64, _, _ = O.recon(wl["oracle"], wl["desc"], wl["case"], "f64")
        out["rel_l2"] = {"vs_fp32_oracle": float(np.linalg.norm(got - r32) / np.linalg.norm(r32)),
                         "vs_fp64_oracle": float(np.linalg.norm(got - r64) / np.linalg.norm(r64)),
                         "fp32_oracle_vs_fp64": float(np.linalg.norm(r32 - r64) / np.linalg.norm(r64))}
    if not a.no_cpu and wl["oracle"]:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds)
        out["cpu_baseline_all_cores"] = cpu_baseline(wl, a.cpu_seconds, a.cpu_threads)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
