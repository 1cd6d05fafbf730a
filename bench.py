#!/usr/bin/env python3
"""S-matrix throughput on one MI355X: slice solves/s at 441 harmonics (N = 882), complex128.

One step = one batched library call (``vrc_batch_stack_rcwa``) over the 512-instance
rectangle-stack sweep (tests/golden/metric_geometry.json): 8 seeded geometries x 64
wavelengths in 0.8-1.6 um; each instance is 10 slabs (1-4 rectangles, eps in {4, 12.25},
0.6 um cell, thickness 0.05-0.2 um) between air half-spaces, Fourier orders +-10 in x
and y. A slice solve = Toeplitz assembly + P/Q + eig(PQ) + slab S-matrix + star product
into the cascade; every instance also solves its two end media and the end projection
(counted as overhead, not as slices). The library splits the batch into launches of at
most 256 eigensolves.

Also reported: the plain per-instance path (RCWAIntegrator over one instance's stack,
one slice after another), live stage times from HIP events on the library stream, the
roofline of the dominant kernel, and the CPU oracle timed on one whole instance.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "S-matrix solves/s @441 harmonics c128, batched+adjoint vs plain; FP64 MFMA %"
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
STAGES = ["fourier", "pq", "hessenberg", "qr", "eigvec", "eigpost", "layer_s", "star", "project", "adjoint"]


def flops_slice(N):
    """Algorithmic real flops of one slice solve (F(N) = 305 N^3)."""
    return 305.0 * N ** 3


def flops_qr(N):
    """Algorithmic real flops of the Schur (QR-iteration) phase of one eigensolve:
    the eig total (316/3) N^3 minus Hessenberg (40/3), Q formation (16/3) and
    eigenvectors + back-transform (28/3)."""
    return (316.0 - 40.0 - 16.0 - 28.0) / 3.0 * N ** 3


TRAFFIC_SUMMARY = ROOT / "profiles" / "pmc_traffic__hqr_kernel_rect_stack_N882_25inst.md"


def hqr_traffic_per_matrix():
    """HBM bytes per matrix of hqr_kernel (FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc,
    each scaled by the counter's reading for a plain copy of known size), from the
    committed profile summary of the 25-instance launch; None when it is absent."""
    import re
    try:
        text = TRAFFIC_SUMMARY.read_text()
    except OSError:
        return None
    m = re.search(r"hbm_bytes_per_matrix:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def load_instances(nh: int, count: int):
    from varrcwa_amd.api import Layer
    d = json.loads((ROOT / "tests" / "golden" / "metric_geometry.json").read_text())
    cell = d["cell"]
    lams = d["wavelengths"]
    geoms = d["geometries"]
    air = Layer([0.0, cell], [0.0, cell], [[1.0]])
    insts = []
    for i in range(count):
        g = geoms[i % len(geoms)]
        lam = lams[(i // len(geoms)) % len(lams)]
        layers = [Layer(l["x"], l["y"], np.asarray(l["eps"], dtype=complex)) for l in g["layers"]]
        insts.append(dict(lam=lam, layers=layers, thickness=g["thickness"], bottom=air, top=air))
    return insts


def read_stages(ctx):
    ms = (C.c_double * len(STAGES))()
    cnt = (C.c_int * len(STAGES))()
    ctx.check(ctx.lib.vrc_timing_read(ctx.h, ms, cnt, len(STAGES)), "timing_read")
    return {s: (ms[i], cnt[i]) for i, s in enumerate(STAGES)}


def cpu_baseline(nh: int):
    """Oracle (restated reference CPU path, numpy/LAPACK) on instance 0 of the sweep:
    RCWAIntegrator over its 10 slabs (end media, slices, star products, projection)."""
    from oracle import rcwa_oracle as O
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # pragma: no cover
        threadpool_limits = None
    d = json.loads((ROOT / "tests" / "golden" / "metric_geometry.json").read_text())
    g = d["geometries"][0]
    cell = d["cell"]
    lam = d["wavelengths"][0]
    air = (np.array([0.0, cell]), np.array([0.0, cell]), np.array([[1.0 + 0j]]))
    layers = [(np.array(l["x"]), np.array(l["y"]), np.array(l["eps"], dtype=complex)) for l in g["layers"]]
    ze = np.concatenate([[0.0], np.cumsum(g["thickness"])])
    s = O.StackLayerSampler(lam, nh, nh, layers, ze, air, air)
    cores = min(16, os.cpu_count() or 1)
    ctxm = threadpool_limits(limits=cores) if threadpool_limits else None
    try:
        t0 = time.perf_counter()
        O.rcwa_integrate(s, ze[0], ze[-1], len(layers), z_edges=ze)
        t = time.perf_counter() - t0
    finally:
        if ctxm is not None:
            ctxm.__exit__(None, None, None)
    nsl = len(layers)
    return dict(value=nsl / t, unit="slice solves/s", cores=cores, kind="port",
                sample=f"instance 0 of the sweep ({nsl} slabs + 2 end media + projection, N={2 * nh * nh}; "
                       f"numpy/OpenBLAS LAPACK, {cores} threads, {t:.1f} s)")


def adjoint_bench(ctx, nh: int, insts, log):
    """Forward S-matrices with the adjoint tape + the reverse pass for the objective
    sum_i zeroth-order transmission (TE+TM), timed against the plain batched forward
    on the same instances (one launch chunk)."""
    import torch
    from varrcwa_amd import api
    from varrcwa_amd.autograd import StackProblem, stack_smatrix, zeroth_order_transmission
    plain = [dict(i, layers=[(l.x, l.y, l.eps) for l in i["layers"]]) for i in insts]
    prob = StackProblem(ctx, nh, nh, plain)
    eps = prob.eps_vector().real.clone().to("cuda").requires_grad_(True)

    def fwd_bwd():
        S, W0, W1 = stack_smatrix(eps, prob)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J = zeroth_order_transmission(S, W0, W1, nh, nh).sum()
        J.backward()
        torch.cuda.synchronize()
        return t1, time.perf_counter()

    fwd_bwd()  # warm
    eps.grad = None
    torch.cuda.synchronize()
    read_stages(ctx)  # reset
    t0 = time.perf_counter()
    t1, t2 = fwd_bwd()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    st = read_stages(ctx)
    api.batch_stack_rcwa(ctx, nh, nh, plain)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    fwd_tape, bwd, fwd = t1 - t0, t2 - t1, t4 - t3
    log(f"adjoint: {len(insts)} instances, forward {fwd * 1e3:.0f} ms, forward+tape {fwd_tape * 1e3:.0f} ms, "
        f"reverse {bwd * 1e3:.0f} ms")
    return {"instances": len(insts), "objective": "sum of 0th-order transmitted power (TE+TM) per instance",
            "gradient_wrt": f"{prob.cells} eps cells", "forward_ms": fwd * 1e3, "forward_with_tape_ms": fwd_tape * 1e3,
            "reverse_ms": bwd * 1e3, "reverse_over_forward": bwd / fwd, "gradient_over_forward": (fwd_tape + bwd) / fwd,
            "stage_ms_forward_and_reverse": {k: v[0] for k, v in st.items()}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=512, help="instances per step (8 geometries x 64 wavelengths)")
    ap.add_argument("--nh", type=int, default=21, help="harmonics per axis (2*order+1)")
    ap.add_argument("--plain", type=int, default=1, help="instances timed through the plain path (0 = skip)")
    ap.add_argument("--cpu", type=int, default=1, help="time the CPU oracle on one instance (0 = skip)")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--adjoint", type=int, default=21, help="instances in the adjoint timing (0 = skip)")
    ap.add_argument("--tune", action="append", default=[], help="kernel switch key=value (vrc_tune_set)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="first copy this many bytes device to device (a plain copy kernel of known size "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(instances would shard across GPUs with no exchange)"}))
        sys.exit(2)

    import torch
    from varrcwa_amd import api
    from varrcwa_amd._lib import Context

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    ctx = Context(0)
    if args.copy_calibration > 0:
        nel = args.copy_calibration // 8
        src = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {nel * 8} bytes read, {nel * 8} bytes written")
        del src, dst
    for kv in args.tune:
        k, v = kv.split("=")
        if ctx.lib.vrc_tune_set(k.encode(), int(v)) != 0:
            raise SystemExit(f"unknown tuning key {k}")
    nh = args.nh
    n = nh * nh
    N = 2 * n
    insts = load_instances(nh, args.instances)
    L = len(insts[0]["layers"])
    ctx.check(ctx.lib.vrc_timing_enable(ctx.h, 1), "timing")

    for w in range(args.warmup):
        api.batch_stack_rcwa(ctx, nh, nh, insts, chunk=args.chunk)
        torch.cuda.synchronize()
        log(f"warmup {w + 1}/{args.warmup} done")
    read_stages(ctx)  # reset
    qs = (C.c_longlong * 6)()
    ctx.check(ctx.lib.vrc_qr_stats(ctx.h, qs), "qr_stats")  # reset
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = api.batch_stack_rcwa(ctx, nh, nh, insts, chunk=args.chunk)
        del out
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = read_stages(ctx)
    ctx.check(ctx.lib.vrc_qr_stats(ctx.h, qs), "qr_stats")  # the timed steps only
    free_b, total_b = torch.cuda.mem_get_info()
    dt = t1 - t0
    slices = args.steps * len(insts) * L
    value = slices / dt
    log(f"batched: {slices} slices in {dt:.2f} s -> {value:.2f} slices/s")

    # dominant stage (stage 1 includes stage 0; report exclusive times)
    excl = {k: v[0] for k, v in st.items()}
    excl["pq"] = max(0.0, excl["pq"] - excl["fourier"])
    dom = max(excl, key=excl.get)
    qr_ms, qr_cnt = st["qr"]
    roof = None
    if qr_cnt:
        # dense eigensolves of the timed steps, counted by the QR kernel itself (the
        # homogeneous end media take the closed form and are not among them)
        mats_per_launch = qs[0] / qr_cnt
        avg_ms = qr_ms / qr_cnt
        per_launch = flops_qr(N) * mats_per_launch
        ach = per_launch / (avg_ms * 1e-3) / 1e12
        tpm = hqr_traffic_per_matrix()
        roof = {"bound": "mfma", "kernel": "hqr_kernel (multishift QR, Schur phase)", "achieved": ach,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                "traffic": tpm * mats_per_launch if tpm else None,
                "traffic_source": TRAFFIC_SUMMARY.relative_to(ROOT).as_posix() if tpm else None,
                "avg_launch_ms": avg_ms, "flops_per_launch": per_launch, "matrices_per_launch": mats_per_launch}

    plain = None
    if args.plain > 0:
        inst = insts[0]
        ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
        s = api.StackLayerSampler(inst["lam"], nh, nh, inst["layers"], ze, inst["bottom"], inst["top"], ctx=ctx)
        it = api.RCWAIntegrator(ze[0], ze[-1], s)
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        for _ in range(args.plain):
            it.compute_edges(ze)
        torch.cuda.synchronize()
        tp = time.perf_counter() - tp0
        plain = args.plain * L / tp
        read_stages(ctx)
        log(f"plain: {args.plain * L} slices in {tp:.2f} s -> {plain:.3f} slices/s")
        del it, s

    adj = None
    if args.adjoint > 0:
        adj = adjoint_bench(ctx, nh, insts[:args.adjoint], log)
        read_stages(ctx)

    cpu = None
    if args.cpu > 0:
        cpu = cpu_baseline(nh)
        log(f"cpu: {cpu}")

    frac = flops_slice(N) * value / (FP64_PEAK_TFLOPS * 1e12)
    line = {
        "metric": METRIC, "value": value, "unit": "slice solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"rectangle stacks, {nh}x{nh} harmonics (N={N}), {L} slabs/instance, "
                               f"{len(insts)} instances/step (8 seeded geometries x 64 wavelengths in 0.8-1.6 um)",
                   "instances_per_step": len(insts), "slices_per_step": len(insts) * L, "N": N,
                   "parallelism": "single GPU", "tuning": args.tune},
        "fp64_matrix_frac": frac,
        "path_tflops": flops_slice(N) * value / 1e12,
        "plain_slices_per_s": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "batched_over_cpu": (value / cpu["value"]) if cpu else None,
        "adjoint": adj,
        "stage_ms_per_step": {k: v / args.steps for k, v in excl.items()},
        "device_memory_in_use_gb": (total_b - free_b) / 1e9,
        "dominant_stage": dom,
        "qr_stats": {"matrices": qs[0], "sweeps_per_matrix": qs[1] / max(qs[0], 1),
                     "windows_per_matrix": qs[2] / max(qs[0], 1),
                     "chase_ms_per_matrix": qs[3] / max(qs[0], 1) / 1e5,
                     "apply_ms_per_matrix": qs[4] / max(qs[0], 1) / 1e5,
                     "small_block_ms_per_matrix": qs[5] / max(qs[0], 1) / 1e5},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
