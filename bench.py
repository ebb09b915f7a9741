"""Benchmark: ray.surface intersections per second, FP64, one MI355X.

Workload (one step): the 12-surface echelle spectrograph (rt_amd/schemes.py,
echelle()) at 4883 rays x 2048 wavelengths = 10,000,384 rays: beam generated
in-kernel (Philox, seed per step), traced through every surface, integer
hit-count frame and QE-weighted FP64 flux frame accumulated on the 4096^2
CCD.  All buffers are resident in HBM before the timed region.

value = sum over non-auxiliary surfaces of the rays alive on entering that
surface's intersection stage (read from the kernel's own counters), divided
by the timed wall time of the K steps.  Each step is also bracketed by HIP
events on its stream: the roofline uses the median step time (K >= 50).
The CPU baseline is the oracle's fused trace (built for speed: -O3, AVX2,
strict IEEE) on the same rays on the host cores, the median of 5 timed passes
after a warm-up pass.  The default line also carries, in detail, the
adjoint/forward time ratios of the differentiable configurations (one scheme
at 1e6 rays; 1024 tolerance instances), (loss forward + adjoint) / plain
forward on the same rays.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "fused_trace_echelle_hbm_traffic.txt")
HOST_CPU_SHARE = 16         # host threads one GPU's job may use


def measured_traffic(block, waves):
    """HBM bytes per launch of the fused trace on this workload under the launch
    (block, waves per SIMD), from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (copy-calibrated; lines 'traffic_bytes_per_launch[B,W] = T'), or
    None when that launch was not profiled."""
    key = f"traffic_bytes_per_launch[{block},{waves}]"
    try:
        for line in open(TRAFFIC_SUMMARY):
            if line.startswith(key):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-per-lambda", type=int, default=4883)
    ap.add_argument("--nlambda", type=int, default=2048)
    ap.add_argument("--cpu-sample", type=int, default=0, help="rays traced by the CPU baseline (0: the whole workload)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="leave the adjoint ratios out of the default line")
    ap.add_argument("--adjoint-instances", type=int, default=1024,
                    help="tolerance instances of the batched adjoint ratio in the default line (0: none)")
    ap.add_argument("--launch", default="", metavar="BLOCK,WAVES,GRID_MULT",
                    help="launch configuration of the generated kernel instead of the autotuner's choice")
    ap.add_argument("--batch", type=int, default=0, metavar="N_INST",
                    help="measure the batched workload instead: N_INST Monte-Carlo tolerance instances of "
                         "the echelle, ray counts log-uniform in [1e4, 1e6], one persistent launch per step")
    ap.add_argument("--batch-adjoint", type=int, default=0, metavar="N_INST",
                    help="batched adjoint/forward ratio over N_INST tolerance instances of the echelle")
    ap.add_argument("--max-rays", type=int, default=0,
                    help="largest instance of the batched modes (log-uniform from 1e4); default 1e6, "
                         "5e5 for --design")
    ap.add_argument("--design", type=int, default=0, metavar="N_INST",
                    help="batched design loop: gradient descent on the compensators of N_INST instances")
    ap.add_argument("--iters", type=int, default=50, help="iterations of the design loop")
    ap.add_argument("--adjoint", action="store_true",
                    help="measure the adjoint instead: gradient of the spot loss over the echelle's 48 design "
                         "parameters at 1e6 rays over 64 wavelengths, as a time ratio to the plain forward trace")
    ap.add_argument("--scheme-run", action="store_true",
                    help="measure the drop-in Scheme path instead: the fixed-order echelle written as the "
                         "reference's XML and run by rt_scheme_run_file (files and PCIe included)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(args.copy_calibration * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine is measured on one GPU; --gpus > 1 is not supported "
                                   "(rays would shard by ray range with one reduction of the CCD frame)"}))
        sys.exit(2)

    import torch
    from rt_amd import flops, lib, schemes
    from rt_amd import scheme as S

    lib.init(0)
    if args.batch:
        return bench_batch(args)
    if args.adjoint:
        return bench_adjoint(args)
    if args.batch_adjoint:
        return bench_batch_adjoint(args)
    if args.design:
        return bench_design(args)
    if args.scheme_run:
        return bench_scheme_run(args)
    sch = schemes.echelle(args.n_per_lambda, args.nlambda)
    n = sch.beam.n_rays
    ds = lib.DeviceScheme(sch)
    det = sch.detector()
    npix = det.ccd["ydim"] * det.ccd["xdim"]
    rays = torch.empty((7, n), dtype=torch.float64, device="cuda")
    flag = torch.empty(n, dtype=torch.int8, device="cuda")
    hist = torch.zeros(npix, dtype=torch.int32, device="cuda")
    flux = torch.zeros(npix, dtype=torch.float64, device="cuda")
    stats = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    flags = S.RT_TRACE_GENERATE | S.RT_TRACE_STORE | S.RT_TRACE_HIST | S.RT_TRACE_FLUX
    stream = torch.cuda.current_stream()

    # launch configuration of the generated kernel chosen by the autotuner on the
    # same workload with another seed (setup, outside the timed region)
    if args.launch:
        ds.set_launch(*(int(v) for v in args.launch.split(",")))
        tuned_ms = None
    else:
        tuned_ms = ds.autotune(n_sample=n, flags=flags, seed=999)
    block, waves, grid_mult = ds.get_launch()

    def step(k):
        ds.trace(rays, flag, seed=k, flags=flags, hist=hist, flux=flux, stats=stats,
                 stream=stream.cuda_stream)

    for k in range(args.warmup):
        step(1000 + k)
    torch.cuda.synchronize()
    generated = lib.last_kernel() == "generated"
    table = generated and os.environ.get("RT_HIP_LTAB", "1") != "0"
    stats.zero_()
    hist.zero_()
    flux.zero_()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for k in range(args.steps):
        step(k)
        evs[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    step_ms = np.array([evs[k].elapsed_time(evs[k + 1]) for k in range(args.steps)])
    st = stats.cpu().numpy().view(np.uint64)
    ray_surf = int(sum(int(st[S.ST_SURF + 4 * s]) for s in range(sch.n_surf)))
    ms_per_step = wall * 1e3 / args.steps
    value = ray_surf / wall
    kernel_ms = float(np.median(step_ms))
    fl = flops.trace_flops(sch, st, generated_kernel=generated, table=table) / args.steps
    fl_splines = flops.trace_flops(sch, st, generated_kernel=generated, table=False) / args.steps
    achieved_tflops = fl / (kernel_ms * 1e-3) / 1e12

    out = {
        "metric": "ray-surface hits/s FP64 on 1 MI355X (12-surface echelle, 1e7 rays)",
        "value": value,
        "unit": "ray*surface/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox-generated beam, seed = step index)",
        "config": {"workload": "echelle-12surf", "rays": n, "n_per_lambda": args.n_per_lambda,
                   "nlambda": args.nlambda, "surfaces": sch.n_surf, "ccd": "4096x4096",
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tflops / PEAK_FP64_TFLOPS,
                     "traffic": measured_traffic(block, waves)},
        "detail": {"kernel_ms_median": kernel_ms, "kernel_ms_mean": float(step_ms.mean()),
                   "kernel_ms_min": float(step_ms.min()), "timed_launches": args.steps,
                   "launch": {"block": block, "waves_per_simd": waves, "grid_mult": grid_mult,
                              "surface_unroll": "full" if ds.get_unroll() else "loop",
                              "kernel": "generated" if generated else "generic",
                              "per_wavelength_table": table, "autotune_sample_ms": tuned_ms},
                   "ray_surface_per_step": ray_surf // args.steps,
                   "ccd_hits_per_step": int(st[S.ST_CCD_HITS]) // args.steps,
                   "ccd_edge_per_step": int(st[S.ST_CCD_EDGE]) // args.steps,
                   "flop_per_step": fl,
                   "flop_per_step_splines_counted": fl_splines,
                   "frac_splines_counted": fl_splines / (kernel_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
    }
    out["detail"]["stages_separately"] = stages_separately(sch, det, rays, flag, stream, kernel_ms)
    ds.close()
    del rays, flag, hist, flux

    if not args.no_adjoint:
        out["detail"]["adjoint_single"] = adjoint_ratio(args.warmup, max(args.steps, 20))
        if args.adjoint_instances:
            out["detail"]["adjoint_batched"] = batch_adjoint_ratio(args.adjoint_instances, args.warmup, 5,
                                                                   args.max_rays or 1_000_000)

    if not args.no_cpu:
        from oracle import oracle as O
        m = min(args.cpu_sample, n) if args.cpu_sample > 0 else n
        threads = min(HOST_CPU_SHARE, len(os.sched_getaffinity(0)))
        O.trace(sch, seed=0, n=m, nthreads=threads, fast=True)             # warm-up pass
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            r = O.trace(sch, seed=0, n=m, nthreads=threads, fast=True)
            times.append(time.perf_counter() - t0)
        cpu_s = float(np.median(times))
        cst = r.stats
        cpu_rs = int(sum(int(cst[S.ST_SURF + 4 * s]) for s in range(sch.n_surf)))
        out["cpu_baseline"] = {"value": cpu_rs / cpu_s, "unit": "ray*surface/s", "cores": threads,
                               "kind": "port",
                               "sample": (f"the whole workload ({m} rays)" if m == n
                                          else f"first {m} rays of the same beam") +
                                         ", oracle fused trace built -O3 -march=x86-64-v3 (strict IEEE), "
                                         f"median of 5 passes after a warm-up on {threads} threads of "
                                         f"{cpu_model()}",
                               "pass_seconds": times}
    print(json.dumps(out))


def stages_separately(sch, det, rays, flag, stream, fused_ms, reps=20):
    """Beam generation and the detector stage timed on their own, with the
    per-stage entry points, beside the fused trace that does both with the
    surfaces in between: rt_hip_beam_positions, _directions and _wavelengths
    write the 7 FP64 rows of every ray of the workload (beam_random,
    beam_gauss_cosins, uniform_range_beam); the detector stage runs on the
    rays the last step delivered to the detector, as ccd_frame reads them
    from a saved beam: the hit counts are rt_hip_ccd_coordinates +
    rt_hip_ccd_accumulate (ccd_coordinates, compute_ccd_image), the flux
    image rt_hip_ccd_flux (pixel, QE spline, weight and atomic add in one
    kernel).  Bytes per ray are the algorithmic ones (rows read and written,
    one atomic word); the rate is reported against the HBM peak, and
    the beam's FP64 generation work (Philox, logarithm, sine and cosine,
    square roots) is beside it.  Median of `reps` HIP-event timings on the
    launch stream."""
    import torch
    from rt_amd import lib
    from rt_amd import scheme as S
    L = lib.lib()
    n = rays.shape[1]
    st = stream.cuda_stream
    desc = np.ascontiguousarray(sch.beam.pack())
    gen = torch.empty((7, n), dtype=torch.float64, device="cuda")
    P = [gen[k].data_ptr() for k in range(7)]
    conic = sch.beam.type == S.RT_BEAM_CONIC
    c = det.ccd
    xdim, ydim = int(c["xdim"]), int(c["ydim"])
    xsize, ysize = c.get("xsize", xdim * c["xpix"]), c.get("ysize", ydim * c["ypix"])
    px = torch.empty(n, dtype=torch.int32, device="cuda")
    py = torch.empty(n, dtype=torch.int32, device="cuda")
    hist = torch.zeros(xdim * ydim, dtype=torch.int32, device="cuda")
    image = torch.zeros(xdim * ydim, dtype=torch.float64, device="cuda")
    counts = torch.zeros(3, dtype=torch.int64, device="cuda")
    off = torch.zeros(1, dtype=torch.int64, device="cuda")
    qe = det.qe
    qe_rec = None
    if isinstance(qe, S.Spline):
        qe_rec = np.concatenate([[len(qe.x)], qe.x, qe.y, qe.y2]).astype(np.float64)
    alive = flag != 0
    hits = [rays[k][alive].contiguous() for k in (0, 1, 6)]
    nh = hits[0].numel()
    X, Y, lam = (h.data_ptr() for h in hits)

    def beam():
        lib.check(L.rt_hip_beam_positions(lib._dptr(desc), 7, 0, n, P[0], P[1], P[2], st), "rt_hip_beam_positions")
        lib.check(L.rt_hip_beam_directions(lib._dptr(desc), 7, 0, n, P[0] if conic else None,
                                           P[1] if conic else None, P[3], P[4], P[5], st), "rt_hip_beam_directions")
        lib.check(L.rt_hip_beam_wavelengths(lib._dptr(desc), 7, 0, n, P[6], st), "rt_hip_beam_wavelengths")

    def ccd_hist():
        lib.check(L.rt_hip_ccd_coordinates(nh, X, Y, xsize, ysize, c["xpix"], c["ypix"], px.data_ptr(),
                                           py.data_ptr(), st), "rt_hip_ccd_coordinates")
        lib.check(L.rt_hip_ccd_accumulate(nh, px.data_ptr(), py.data_ptr(), None, xdim, ydim, None,
                                          hist.data_ptr(), off.data_ptr(), st), "rt_hip_ccd_accumulate")

    def ccd_flux():
        lib.check(L.rt_hip_ccd_flux(nh, X, Y, lam, lib._dptr(qe_rec) if qe_rec is not None else None,
                                    0.0 if qe_rec is not None else float(qe), None, 1.0, xsize, ysize,
                                    c["xpix"], c["ypix"], xdim, ydim, image.data_ptr(), counts.data_ptr(), st),
                  "rt_hip_ccd_flux")

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.median(ms))

    res = {}
    for name, fn, bytes_per_ray, m, what in (
            ("beam_generation", beam, 56, n, "7 FP64 rows written"),
            ("ccd_hit_counts", ccd_hist, 36, nh, "X, Y read; 2 u32 pixel rows written and read; one u32 atomic"),
            ("ccd_flux", ccd_flux, 32, nh, "X, Y, lambda read; one FP64 atomic")):
        ms = timed(fn)
        gbs = bytes_per_ray * m / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "rays": m, "algorithmic_bytes_per_ray": bytes_per_ray, "bytes": what,
                     "achieved_GBs": gbs, "frac_of_hbm_peak": gbs / PEAK_HBM_GBS}
    res["fused_trace_ms"] = fused_ms
    res["ccd_rays_on_frame"] = int(counts[0].item()) // (reps + 1)
    del gen, px, py, hist, image, hits
    return res


def bench_scheme_run(args):
    """The drop-in Scheme path end to end, as ray-tracing/main.cpp runs it on
    this library: the echelle in one fixed order (4,883 rays x 2,048
    wavelengths by default) written as the reference's XML scheme file and run
    by rt_scheme_run_file: Scheme::load_from_file, the input beam generated
    and saved, one recording trace of the range (loss codes, logged extents),
    the Rearrange order and the renormalised lost rays formed on the device,
    the result file written and the log.  Wall time per call, with the two
    .dat files (57 B per ray each) and the PCIe copies inside it; the same
    scheme through the reference's own per-stage host loop on the plugin takes
    seconds (oracle/_ref/ref_pipeline_hip).  At most 5 timed calls after one
    warm-up."""
    import re
    import shutil
    import tempfile
    from rt_amd import lib, scheme_xml, schemes
    from rt_amd import scheme as S
    sc = schemes.echelle_fixed_order(args.n_per_lambda, args.nlambda)
    steps, warmup = max(1, min(args.steps, 5)), min(args.warmup, 1)
    d = tempfile.mkdtemp(prefix="rt_scheme_run_")
    try:
        path = scheme_xml.write_scheme(sc, d)
        lst, log = os.path.join(d, "list.txt"), os.path.join(d, "log.txt")
        times = []
        for k in range(warmup + steps):
            t0 = time.perf_counter()
            rc = lib.scheme_run_file(path, lst, log)
            dt = time.perf_counter() - t0
            if rc != 0:
                raise RuntimeError("rt_scheme_run_file returned %d: %s" % (rc, open(log).read()[-2000:]))
            if k >= warmup:
                times.append(dt)
            print(f"[scheme run] call {k}: {dt:.3f} s", file=sys.stderr, flush=True)
        stages = 3 * sum(1 for s in sc.surfaces if s.type != S.RT_AUXTYPE)
        good = [int(v) for v in re.findall(r"number of good rays: (\d+)\)", open(log).read())][-stages:]
        n = sc.beam.n_rays
        # rays entering each surface's intersection: all rays, then the good
        # rays after the previous surface's action
        entering = [n] + good[2:-1:3]
        ray_surf = sum(entering)
        files = [l.strip() for l in open(lst) if l.strip()]
        file_bytes = sum(os.path.getsize(f) for f in files if os.path.exists(f))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    device = scheme_path_device_ms(sc)
    sec = float(np.median(times))
    print(json.dumps({
        "metric": "drop-in Scheme::run end to end (fixed-order echelle, files and PCIe included)",
        "value": ray_surf / sec, "unit": "ray*surface/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic (Philox beam)",
        "config": {"workload": "echelle-fixed-order-scheme-run", "rays": n, "n_per_lambda": args.n_per_lambda,
                   "nlambda": args.nlambda, "surfaces": len(sc.surfaces)},
        "detail": {"seconds_per_call": times, "median_s": sec, "ray_surface_per_call": ray_surf,
                   "good_rays_at_end": good[-1] if good else None, "result_file_bytes": file_bytes,
                   "device_ms": device}}))


def scheme_path_device_ms(sc, reps=5):
    """Device time of one range of Scheme::run through the C ABI, median of
    `reps` HIP-event timings each: the recording trace (beam generated, every
    stage, loss codes, logged extents), the Rearrange order
    (rt_hip_rearrange, counts to the host) and the file-order arrays with the
    lost rays renormalised (rt_hip_beam_order, with the later-Normalize counts
    Scheme::run forms: 1 + 2 per later surface, 1 per later AUX surface)."""
    import torch
    from rt_amd import lib
    from rt_amd import scheme as S
    n = sc.beam.n_rays
    ds = lib.DeviceScheme(sc)
    rays = torch.empty((7, n), dtype=torch.float64, device="cuda")
    flag = torch.empty(n, dtype=torch.int8, device="cuda")
    loss = torch.empty(n, dtype=torch.uint8, device="cuda")
    ext = torch.empty(sc.n_surf * lib.EXT_PER_SURF, dtype=torch.float64, device="cuda")
    renorm = np.zeros(256, dtype=np.uint8)
    for s in range(sc.n_surf):
        later = sum(1 if t.type == S.RT_AUXTYPE else 2 for t in sc.surfaces[s + 1:])
        for k in range(3):
            renorm[3 * s + k + 1] = min(255, 1 + later)
    stream = torch.cuda.current_stream()
    state = {}

    def record():
        ds.trace_record(rays, flag, loss, seed=1, extents=ext)

    def rearrange():
        state["perm"], state["counts"] = lib.rearrange(loss)

    def order():
        state["out"] = lib.beam_order(state["perm"], loss, renorm, rays)

    out = {}
    for name, fn in (("recording_trace", record), ("rearrange", rearrange), ("beam_order", order)):
        fn()
        ms = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            ms.append(a.elapsed_time(b))
        out[name] = float(np.median(ms))
    out["good_rays"] = int(state["counts"][0])
    ds.close()
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def bench_adjoint(args):
    r = adjoint_ratio(args.warmup, args.steps)
    print(json.dumps({
        "metric": "adjoint/forward time ratio (echelle, 1e6 rays, 48 parameters)",
        "value": r["ratio"], "unit": "x", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False, "dtype": "f64",
        "data": "synthetic (Philox beam)",
        "config": {"workload": "echelle-adjoint", "rays": r["rays"], "wavelengths": 64,
                   "parameters": r["parameters"]},
        "detail": r}))


def adjoint_ratio(warmup, steps):
    """Adjoint/forward time ratio on the differentiable configuration: the
    echelle at 15625 rays x 64 wavelengths (1e6 rays), loss = sum over the
    wavelengths of the RMS spot radius, gradient over 48 design parameters.
    Times: the plain forward trace (beam generated, nothing stored), the loss
    forward (the trace forming the group spots itself, rt_hip_trace_spots;
    loss and adjoint seeds reduced on the device), the adjoint launch (kernel
    + fixed-order gradient reduction);
    ratio = (loss forward + adjoint) / plain forward, on the same rays and
    surfaces."""
    import torch
    from rt_amd import diff, lib, schemes
    from rt_amd import scheme as S
    base = schemes.echelle(n_per_lambda=15625, nlambda=64)
    spec = diff.echelle_design_params(base)
    v0 = spec.values(base)
    n = base.beam.n_rays
    ds = lib.DeviceScheme(base)
    spots = torch.zeros((64, lib.SPOT_STRIDE), dtype=torch.float64, device="cuda")
    ps = np.array([p.surface for p in spec.params], dtype=np.int32)
    pf = np.array([p.field for p in spec.params], dtype=np.int32)
    seeds = torch.zeros((64, 3), dtype=torch.float64, device="cuda")

    def loss_forward():
        sp = ds.trace_spots(64, seed=5, out=spots)
        loss, _ = lib.spot_loss(sp, seeds=seeds)
        return loss

    def timed(fn, k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(k):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    plain = lambda: ds.trace(None, None, seed=5, n=n, flags=S.RT_TRACE_GENERATE)
    for _ in range(warmup):
        plain(); loss_forward(); ds.adjoint(seeds, ps, pf, seed=5)
    t_plain = timed(plain, steps)
    t_loss = timed(loss_forward, steps)
    t_adj = timed(lambda: ds.adjoint(seeds, ps, pf, seed=5), steps)
    ds.close()
    return {"ratio": (t_loss + t_adj) / t_plain, "rays": n, "parameters": len(spec), "forward_ms": t_plain,
            "loss_forward_ms": t_loss, "adjoint_ms": t_adj, "adjoint_kernel_ratio": t_adj / t_plain,
            "definition": "(loss forward + adjoint) / plain forward, same rays and surfaces"}


def bench_batch(args):
    """Batched forward: tolerance instances of the echelle (decentres N(0, 20 um),
    tilts N(0, 30 arcsec), curvatures 1 + N(0, 1e-4)), uneven ray counts, spots
    reduced per instance; one rt_hip_batch_trace per step."""
    import torch
    from rt_amd import lib, schemes
    from rt_amd import scheme as S
    base = schemes.echelle()
    inst, counts = schemes.tolerance_instances(base, args.batch, seed=1, min_rays=10_000, max_rays=1_000_000)
    db = lib.DeviceBatch(inst, counts)
    launch = tuple(int(v) for v in args.launch.split(",")) if args.launch else (256, 7, 16)
    db.set_launch(*launch)
    spots = torch.zeros((len(inst), lib.SPOT_STRIDE), dtype=torch.float64, device="cuda")
    stats = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    for k in range(args.warmup):
        db.trace(seed=1000 + k, spots=spots, stats=stats)
    torch.cuda.synchronize()
    stats.zero_()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for k in range(args.steps):
        db.trace(seed=k * 4096, spots=spots, stats=stats)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = stats.cpu().numpy().view(np.uint64)
    ray_surf = int(sum(int(st[S.ST_SURF + 4 * s]) for s in range(base.n_surf)))
    sp = spots.cpu().numpy()
    print(json.dumps({
        "metric": f"ray-surface hits/s FP64 on 1 MI355X, batched ({len(inst)} echelle tolerance instances)",
        "value": ray_surf / wall, "unit": "ray*surface/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (Philox beams, seed + instance)",
        "config": {"workload": "echelle-tolerance-batch", "instances": len(inst), "rays": int(sum(counts)),
                   "min_rays": int(min(counts)), "max_rays": int(max(counts))},
        "detail": {"kernel_ms": ev0.elapsed_time(ev1) / args.steps,
                   "launch": {"block": launch[0], "waves_per_simd": launch[1], "grid_mult": launch[2]},
                   "ray_surface_per_step": ray_surf // args.steps,
                   "median_rms_spot_um": float(np.median(sp[:, 5]) * 1e3),
                   "rays_per_s": int(sum(counts)) * args.steps / wall}}))


def bench_batch_adjoint(args):
    r = batch_adjoint_ratio(args.batch_adjoint, args.warmup, args.steps, args.max_rays or 1_000_000)
    print(json.dumps({
        "metric": f"batched adjoint/forward time ratio ({r['instances']} echelle tolerance instances, "
                  "48 parameters)",
        "value": r["ratio"], "unit": "x", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False, "dtype": "f64",
        "data": "synthetic (Philox beams, seed + instance)",
        "config": {"workload": "echelle-tolerance-batch-adjoint", "instances": r["instances"], "rays": r["rays"],
                   "wavelengths": 64, "parameters": r["parameters"]},
        "detail": r}))


def batch_adjoint_ratio(n_inst, warmup, steps, max_rays):
    """Batched adjoint/forward ratio: N_INST tolerance instances of the echelle
    (64 wavelengths, ray counts log-uniform in [1e4, --max-rays]), the spot
    loss per instance (sum over wavelengths of the RMS spot radius), gradient of
    every instance over the 48 design parameters.  Times per step: the plain
    batched forward (beams generated, nothing stored), the loss forward (the
    batched trace forming the spots per instance and wavelength itself,
    rt_hip_batch_trace_group_spots; loss and adjoint seeds on the device), the
    batched adjoint (kernel + per-instance reduction).  Ratio = (loss forward
    + adjoint) / plain forward on the same rays and surfaces; the
    per ray*surface ratio against the single-scheme forward rate of the
    default bench is in detail."""
    import torch
    from rt_amd import diff, lib, schemes
    from rt_amd import scheme as S
    base = schemes.echelle(n_per_lambda=4883, nlambda=64)
    spec = diff.echelle_design_params(base)
    inst, counts = schemes.tolerance_instances(base, n_inst, seed=1, min_rays=10_000, max_rays=max_rays)
    db = lib.DeviceBatch(inst, counts)
    total = int(sum(counts))
    ps = np.array([p.surface for p in spec.params], dtype=np.int32)
    pf = np.array([p.field for p in spec.params], dtype=np.int32)
    seeds = torch.zeros((len(inst), 64, 3), dtype=torch.float64, device="cuda")
    grad = torch.zeros((len(inst), len(ps)), dtype=torch.float64, device="cuda")
    stats = torch.zeros(S.STATS_LEN, dtype=torch.int64, device="cuda")
    spots = torch.zeros((len(inst), 64, lib.SPOT_STRIDE), dtype=torch.float64, device="cuda")

    def plain():
        db.trace(seed=5, flags=S.RT_TRACE_GENERATE, stats=stats)

    def loss_forward():
        sp = db.trace_group_spots(64, seed=5, out=spots)
        loss, _ = lib.spot_loss(sp.view(len(inst) * 64, lib.SPOT_STRIDE), seeds=seeds.view(len(inst) * 64, 3))
        return loss

    def adjoint():
        db.adjoint(seeds, ps, pf, seed=5, grad=grad)

    def timed(fn, k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(k):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    for _ in range(warmup):
        plain(); loss_forward(); adjoint()
    torch.cuda.synchronize()
    stats.zero_()
    t_plain = timed(plain, steps)
    st = stats.cpu().numpy().view(np.uint64)
    ray_surf = int(sum(int(st[S.ST_SURF + 4 * s]) for s in range(base.n_surf))) // steps
    t_loss = timed(loss_forward, steps)
    t_adj = timed(adjoint, steps)
    g = grad.cpu().numpy()
    db.close()
    return {"ratio": (t_loss + t_adj) / t_plain, "instances": len(inst), "rays": total,
            "min_rays": int(min(counts)), "max_rays": int(max(counts)), "parameters": len(ps),
            "plain_forward_ms": t_plain, "loss_forward_ms": t_loss, "adjoint_ms": t_adj,
            "adjoint_kernel_ratio": t_adj / t_plain, "ray_surface_per_step": ray_surf,
            "forward_ray_surface_per_s": ray_surf / (t_plain * 1e-3),
            "finite_gradients": bool(np.all(np.isfinite(g))),
            "definition": "(loss forward + adjoint) / plain forward, same rays and surfaces"}


def bench_design(args):
    """Batched and differentiable design loop: N_INST tolerance instances of the
    echelle (64 wavelengths, ray counts log-uniform in [1e4, --max-rays]); one
    iteration = batched trace with stored rays, spots per instance and
    wavelength, the spot loss and its seeds, the batched adjoint over the
    48 design parameters plus the compensators, and a gradient-descent step on
    the compensators (CCD focus and tilts about Y and X) on the device tables;
    the iteration is captured once in a hipGraph and replayed --iters times.
    Reported: wall time per iteration against 256 (N_INST) sequential plain
    single-scheme forwards of the same instances, and the loss before and
    after."""
    import torch
    from rt_amd import diff, lib, schemes
    from rt_amd import scheme as S
    G = 64
    base = schemes.echelle(n_per_lambda=4883, nlambda=G)
    inst, counts = schemes.tolerance_instances(base, args.design, seed=2, min_rays=10_000,
                                               max_rays=args.max_rays or 500_000)
    spec = diff.echelle_design_params(base)
    ps = [p.surface for p in spec.params]
    pf = [p.field for p in spec.params]
    ccd = base.n_surf - 1
    for f in (S.S_DIST + 2, S.S_ANG + 0, S.S_ANG + 1):
        if (ccd, f) not in set(zip(ps, pf)):
            ps.append(ccd)
            pf.append(f)
    ps, pf = np.array(ps, dtype=np.int32), np.array(pf, dtype=np.int32)
    comp = [j for j, (s, f) in enumerate(zip(ps, pf)) if s == ccd and f in (S.S_DIST + 2, S.S_ANG, S.S_ANG + 1)]
    n, total = len(inst), int(sum(counts))
    db = lib.DeviceBatch(inst, counts)
    b = dict(spots=torch.zeros((n, G, lib.SPOT_STRIDE), dtype=torch.float64, device="cuda"),
             seeds=torch.zeros((n * G, 3), dtype=torch.float64, device="cuda"),
             loss=torch.zeros(1, dtype=torch.float64, device="cuda"),
             grad=torch.zeros((n, len(ps)), dtype=torch.float64, device="cuda"))
    lr = torch.zeros(len(ps), dtype=torch.float64, device="cuda")

    def step():
        db.trace_group_spots(G, seed=9, out=b["spots"])
        lib.spot_loss(b["spots"].view(n * G, lib.SPOT_STRIDE), seeds=b["seeds"], loss=b["loss"])
        db.adjoint(b["seeds"].view(n, G, 3), ps, pf, seed=9, grad=b["grad"])
        db.descend(ps, pf, b["grad"], lr)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()                                  # lr = 0: compiles, allocates, measures the gradient scale
    torch.cuda.synchronize()
    loss0 = float(b["loss"].item())
    # step sizes: 2 um of focus and 2 arcsec of tilt for the largest gradient of the batch
    g = b["grad"].abs().max(dim=0).values.cpu().numpy()
    lr_h = np.zeros(len(ps))
    for j in comp:
        step_size = 2e-3 if pf[j] == S.S_DIST + 2 else 2.0 / 3600.0 * S.DEG2RAD
        lr_h[j] = step_size / max(g[j], 1e-300)
    lr.copy_(torch.tensor(lr_h))
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s, capture_error_mode="thread_local"):
        step()
    torch.cuda.synchronize()
    losses = []
    t0 = time.perf_counter()
    for _ in range(args.iters):
        graph.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # loss after the loop (one more replay computes it on the final tables)
    graph.replay()
    torch.cuda.synchronize()
    loss_end = float(b["loss"].item())
    # 256 (n) sequential plain single-scheme forwards of the same instances
    dss = [lib.DeviceScheme(sch) for sch in inst]
    stream = torch.cuda.current_stream()
    for ds, c in zip(dss, counts):
        ds.trace(None, None, seed=9, n=int(c), flags=S.RT_TRACE_GENERATE)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for ds, c in zip(dss, counts):
        ds.trace(None, None, seed=9, n=int(c), flags=S.RT_TRACE_GENERATE, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    seq = time.perf_counter() - t1
    per_iter = wall / args.iters
    print(json.dumps({
        "metric": f"design-loop iteration time / {n} sequential plain forwards (batched forward + adjoint + step)",
        "value": per_iter / seq, "unit": "x", "n_gpus": 1, "steps": args.iters, "warmup": 1,
        "higher_is_better": False, "dtype": "f64", "data": "synthetic (Philox beams, seed + instance)",
        "config": {"workload": "echelle-tolerance-design-loop", "instances": n, "rays": total,
                   "min_rays": int(min(counts)), "max_rays": int(max(counts)), "wavelengths": G,
                   "parameters": len(ps), "compensators": len(comp), "iterations": args.iters},
        "detail": {"iteration_ms": per_iter * 1e3, "sequential_forwards_ms": seq * 1e3,
                   "loss_first": loss0, "loss_last": loss_end}}))


if __name__ == "__main__":
    main()
