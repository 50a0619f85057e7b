"""Benchmark: batched EM on the 256^3 x 10 mm grid with 1e7 muons (paths of
80-600 voxels).  FP32C (FP32 entry terms, FP64 and int64 fixed-point sums,
FP64 2x2 algebra) is timed; the FP64 GPU paths (reference order, and one
wavefront per track) run beside it for the same number of iterations, and
the image error of FP32C against FP64 is reported.

A step is one EM iteration over every track: forward covariance sums,
per-track 2x2 algebra, backprojection and the mean voxel update
(src/MutoMLSD.cpp:179-213, 442-517).  value = track-voxel updates per second.

Other workloads (--workload): scans64 (many-scan batch), map128 (MAP
objective and gradient), siddon200 (plain PoCA + Siddon).

Prints one JSON line.  One GPU only: --gpus above 1 exits.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
# LDS int64 atomic adds: 5.9 lane-ops per cycle per CU (a microbenchmark on
# this GPU), 256 CUs at 2.4 GHz
LDS_ATOMIC_PEAK = 5.9 * 256 * 2.4e9
BYTES_PER_UPDATE_FP32C = 8.0  # u32 voxel index + f32 length per entry
BYTES_PER_UPDATE_FP64 = 12.0  # u32 voxel index + f64 length per entry


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(workload, kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (profiles/<kernel>_<workload>_pmc.txt: 'traffic_bytes: N'), or None."""
    p = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), as lscpu reports it."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(g, rays, cfg_kw, n_sample, iters):
    """The CPU FP64 oracle (reference operation order; serial track loop,
    OpenMP voxel loop) on a seeded subset of the same muons at the full
    grid: updates/s."""
    import oracle
    from muontomography_amd._abi import mlsd_cfg
    cfg = mlsd_cfg(g, **cfg_kw)
    sub = rays[:n_sample]
    t0 = time.perf_counter()
    em = oracle.Em(cfg, sub)
    t_prep = time.perf_counter() - t0
    lam = np.full(g.n_vox, 1e-7)
    t0 = time.perf_counter()
    for _ in range(iters):
        lam, _, _ = em.iterate(lam)
    t = time.perf_counter() - t0
    nnz = em.nnz
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": nnz * iters / t, "unit": "track-voxel updates/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{n_sample} of the same muons (first {n_sample}), full 256^3 grid, "
                      f"{nnz} entries, {iters} FP64 mean-update iterations in {t:.2f} s "
                      f"(set-up {t_prep:.2f} s not counted); serial track loop, OpenMP voxel "
                      f"loop over {threads} threads"}


N_COMPARE = 50  # iterations of the FP32C-vs-FP64 image comparison (the configuration's count)


def _rel(a, ref, mj):
    m = (mj > 0) & (ref > 0)
    r = np.abs(a[m] - ref[m]) / ref[m]
    return (float(r.max()) if r.size else None, float(np.sqrt(np.mean(r * r))) if r.size else None,
            float(np.abs(a - ref).max() / np.abs(ref).max()))


def bench_fp64(args, plan, rays, cfg_kw, init, nnz, wall32, cost32, brick32):
    """The FP64 paths beside FP32C, each for N_COMPARE iterations from the
    same image: the reference-order kernel (one lane per track), the
    wavefront-per-track kernel, and the FP64 brick stream; the FP32C and
    FP64-stream images against the reference-order one.  Then the
    reference's default update rule, the median (src/MutoMLSD.cpp:505-515):
    reference-order FP64 for 3 iterations beside the FP64 stream and FP32C
    (c_ij from FP32 entry terms in each voxel's slot, the median selected
    among the FP32 values), and FP32C against the FP64 stream for
    N_COMPARE iterations."""
    import muontomography_amd as M
    n = N_COMPARE
    res = {}
    mj = plan.voxel_counts()
    # reference order
    plan.set_image(init)
    t0 = time.perf_counter()
    cost64 = plan.run(n, M.MUTO_FP64)
    w64 = (time.perf_counter() - t0) / n
    img64 = plan.image()
    t64 = plan.timing()
    # wavefront per track
    plan.set_option(M.MUTO_OPT_FP64_KERNEL, 1)
    plan.set_image(init)
    plan.run(1, M.MUTO_FP64)
    plan.set_image(init)
    t0 = time.perf_counter()
    plan.run(n, M.MUTO_FP64)
    w64w = (time.perf_counter() - t0) / n
    img64w = plan.image()
    # FP32C after the same iterations
    plan.set_image(init)
    cost32n = plan.run(n, M.MUTO_FP32C)
    img32 = plan.image()
    e32 = _rel(img32, img64, mj)
    # FP32C with FP32 two-sum sums (MUTO_OPT_ACC_MODE 0) for the same iterations
    plan.set_option(M.MUTO_OPT_ACC_MODE, 0)
    plan.set_image(init)
    plan.run(1, M.MUTO_FP32C)  # graph capture
    plan.set_image(init)
    t0 = time.perf_counter()
    plan.run(n, M.MUTO_FP32C)
    w2s = (time.perf_counter() - t0) / n
    e2s = _rel(plan.image(), img64, mj)
    plan.set_option(M.MUTO_OPT_TIMING, 1)
    plan.set_image(init)
    plan.run(args.timing_steps, M.MUTO_FP32C)
    plan.set_option(M.MUTO_OPT_TIMING, 0)
    t2s = plan.timing()
    plan.set_option(M.MUTO_OPT_ACC_MODE, 1)
    log(f"FP32C two-sum sums: {w2s * 1e3:.2f} ms/it ({t2s}), err max {e2s[0]}")
    # FP64 brick stream: its own brick shape (FP64 bricks and tables in LDS),
    # the fastest of a few over 2 iterations
    plan.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
    tune = {}
    for shape in ((16, 16, 16, 256), (16, 8, 32, 256), (8, 16, 32, 512), (16, 16, 32, 512)):
        plan.set_option(M.MUTO_OPT_BRICK, M.brick_code(shape))
        plan.set_image(init)
        plan.run(1, M.MUTO_FP64)
        plan.set_option(M.MUTO_OPT_TIMING, 1)
        plan.run(2, M.MUTO_FP64)
        plan.set_option(M.MUTO_OPT_TIMING, 0)
        tune[shape] = plan.timing()["iteration"]
    b64 = min(tune, key=tune.get)
    plan.set_option(M.MUTO_OPT_BRICK, M.brick_code(b64))
    plan.set_image(init)
    plan.run(1, M.MUTO_FP64)
    plan.set_image(init)
    t0 = time.perf_counter()
    cost64s = plan.run(n, M.MUTO_FP64)
    w64s = (time.perf_counter() - t0) / n
    img64s = plan.image()
    plan.set_option(M.MUTO_OPT_TIMING, 1)
    plan.set_image(init)
    plan.run(args.timing_steps, M.MUTO_FP64)
    plan.set_option(M.MUTO_OPT_TIMING, 0)
    t64s = plan.timing()
    e64s = _rel(img64s, img64, mj)
    log(f"FP64 {n} it: reference order {w64 * 1e3:.2f} ms/it, wavefront {w64w * 1e3:.2f}, "
        f"stream {w64s * 1e3:.2f} (brick {b64}; {tune}); FP32C err max {e32[0]}, "
        f"FP64 stream err max {e64s[0]}")
    # median, the reference's default rule
    kw = dict(cfg_kw)
    kw["update_with_mean"] = 0
    med = plan.ctx.em_plan(M.mlsd_cfg(plan.cfg.grid, **kw), rays=rays)
    nm = 3
    med.set_image(init)
    t0 = time.perf_counter()
    med.run(nm, M.MUTO_FP64)
    wm_ref = (time.perf_counter() - t0) / nm
    imgm_ref = med.image()
    med.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
    med.set_option(M.MUTO_OPT_BRICK, M.brick_code(b64))
    med.set_image(init)
    med.run(1, M.MUTO_FP64)  # stream, slots, graph
    med.set_image(init)
    med.run(nm, M.MUTO_FP64)
    imgm = med.image()
    em_ = _rel(imgm, imgm_ref, mj)
    med.set_image(init)
    t0 = time.perf_counter()
    med.run(args.steps, M.MUTO_FP64)
    wm = (time.perf_counter() - t0) / args.steps
    miss64 = med.counters()["n_median_miss"]
    med.set_option(M.MUTO_OPT_TIMING, 1)
    med.set_image(init)
    med.run(args.timing_steps, M.MUTO_FP64)
    tm = med.timing()
    med.set_option(M.MUTO_OPT_TIMING, 0)
    # the same median with every c_ij written to its slot (no brackets)
    med.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, 0)
    med.set_image(init)
    med.run(1, M.MUTO_FP64)
    t0 = time.perf_counter()
    med.run(args.steps, M.MUTO_FP64)
    wm_slots = (time.perf_counter() - t0) / args.steps
    med.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, 1)
    med.set_image(init)
    med.run(n, M.MUTO_FP64)
    imgm_s = med.image()
    log(f"median: reference order {wm_ref * 1e3:.1f} ms/it, FP64 stream {wm * 1e3:.2f} ms/it by brackets "
        f"({miss64} misses in the last iteration; {wm_slots * 1e3:.2f} with every value in its slot) "
        f"(phases {tm}); err after {nm} it {em_}")
    # FP32C median on the FP32C brick shape
    med.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick32))
    med.set_image(init)
    med.run(nm, M.MUTO_FP32C)  # stream, slots, graph
    e32m = _rel(med.image(), imgm_ref, mj)
    med.set_image(init)
    med.run(n, M.MUTO_FP32C)
    e32ms = _rel(med.image(), imgm_s, mj)
    med.set_image(init)
    t0 = time.perf_counter()
    med.run(args.steps, M.MUTO_FP32C)
    wm32 = (time.perf_counter() - t0) / args.steps
    miss32 = med.counters()["n_median_miss"]
    med.set_option(M.MUTO_OPT_TIMING, 1)
    med.set_image(init)
    med.run(args.timing_steps, M.MUTO_FP32C)
    tm32 = med.timing()
    med.set_option(M.MUTO_OPT_TIMING, 0)
    med.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, 0)
    med.set_image(init)
    med.run(1, M.MUTO_FP32C)
    t0 = time.perf_counter()
    med.run(args.steps, M.MUTO_FP32C)
    wm32_slots = (time.perf_counter() - t0) / args.steps
    med.set_option(M.MUTO_OPT_MEDIAN_BRACKETS, 1)
    med.close()
    log(f"median FP32C: {wm32 * 1e3:.2f} ms/it by brackets ({miss32} misses in the last iteration; "
        f"{wm32_slots * 1e3:.2f} with every value in its slot) (phases {tm32}); err after {nm} it vs "
        f"reference order {e32m}, after {n} it vs the FP64 stream {e32ms}")
    ms32 = wall32 / args.steps * 1e3
    # the FP64 stream's iteration against its HBM roofline (12 B per update)
    bw64s = BYTES_PER_UPDATE_FP64 * nnz / (t64s["iteration"] * 1e-3)
    return {
        "iterations_compared": n,
        "fp64_ms_per_step": w64 * 1e3,
        "fp64_value": nnz / w64,
        "fp32c_speedup_vs_fp64": w64 * 1e3 / ms32,
        "image_err_max_rel": e32[0], "image_err_rms_rel": e32[1], "image_err_norm_inf": e32[2],
        "cost_rel_diff_last": float(abs(cost32n[-1] - cost64[-1]) / abs(cost64[-1])),
        "fp64_kernel_ms": {k: round(v, 4) for k, v in t64.items()},
        "fp64_wave_ms_per_step": w64w * 1e3,
        "fp64_wave_value": nnz / w64w,
        "fp32c_speedup_vs_fp64_wave": w64w * 1e3 / ms32,
        "fp64_wave_vs_fp64_err_max_rel": _rel(img64w, img64, mj)[0],
        "fp32c_two_sum": {"ms_per_step": w2s * 1e3, "value": nnz / w2s,
                          "kernel_ms": {k: round(v, 4) for k, v in t2s.items()},
                          "err_vs_fp64_max_rel": e2s[0], "err_vs_fp64_rms_rel": e2s[1]},
        "fp64_stream": {"ms_per_step": w64s * 1e3, "value": nnz / w64s, "brick": list(b64),
                        "kernel_ms": {k: round(v, 4) for k, v in t64s.items()},
                        "brick_ms": {",".join(map(str, k)): round(v, 3) for k, v in tune.items()},
                        "speedup_vs_fp64": w64 / w64s,
                        "err_vs_fp64_max_rel": e64s[0], "err_vs_fp64_rms_rel": e64s[1],
                        "cost_rel_diff_last": float(abs(cost64s[-1] - cost64[-1]) / abs(cost64[-1])),
                        "iteration_GBps": bw64s / 1e9, "iteration_frac": bw64s / HBM_PEAK},
        "median": {"fp64_ms_per_step": wm_ref * 1e3, "fp64_stream_ms_per_step": wm * 1e3,
                   "roofline": {"bound": "hbm", "kernel": "whole FP64 stream median iteration",
                                "achieved": BYTES_PER_UPDATE_FP64 * nnz / wm / 1e9,
                                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                "frac": BYTES_PER_UPDATE_FP64 * nnz / wm / HBM_PEAK},
                   "fp64_stream_value": nnz / wm, "speedup": wm_ref / wm,
                   "fp64_stream_slots_ms_per_step": wm_slots * 1e3, "fp64_stream_misses_last": miss64,
                   "fp64_stream_kernel_ms": {k: round(v, 4) for k, v in tm.items()},
                   "iterations_compared": nm, "err_max_rel": em_[0], "err_rms_rel": em_[1],
                   "fp32c_ms_per_step": wm32 * 1e3, "fp32c_value": nnz / wm32,
                   "fp32c_slots_ms_per_step": wm32_slots * 1e3, "fp32c_misses_last": miss32,
                   "fp32c_kernel_ms": {k: round(v, 4) for k, v in tm32.items()},
                   "fp32c_speedup_vs_fp64": wm_ref / wm32, "fp32c_speedup_vs_fp64_stream": wm / wm32,
                   "fp32c_err_vs_fp64_max_rel": e32m[0], "fp32c_err_vs_fp64_rms_rel": e32m[1],
                   "fp32c_vs_fp64_stream_iterations": n, "fp32c_err_vs_fp64_stream_max_rel": e32ms[0],
                   "fp32c_err_vs_fp64_stream_rms_rel": e32ms[1]},
    }



def bench_fp64_stream_only(args, ctx, plan, rays, cfg_kw, init, nnz):
    """The FP64 brick stream alone at one brick shape: --steps mean-update
    iterations, then --steps median iterations on a second plan."""
    import muontomography_amd as M
    brick = tuple(int(v) for v in (args.brick or "16,16,16,256").split(","))
    out = {"workload": "FP64 stream, 256^3 grid of 10 mm voxels, 1e7 muons", "brick": list(brick),
           "entries": int(nnz)}
    kw = dict(cfg_kw)
    plans = [("mean", plan)]
    kw["update_with_mean"] = 0
    plans.append(("median", ctx.em_plan(M.mlsd_cfg(plan.cfg.grid, **kw), rays=rays)))
    for name, p in plans:
        p.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
        p.set_option(M.MUTO_OPT_FP64_KERNEL, 2)
        p.set_image(init)
        p.run(1, M.MUTO_FP64)
        p.set_image(init)
        t0 = time.perf_counter()
        p.run(args.steps, M.MUTO_FP64)
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        out[name + "_ms_per_step"] = ms
        log(f"FP64 stream {name}: {ms:.2f} ms/it")
        p.close()
    ctx.close()
    print(json.dumps(out))


def bench_median_fp32c_only(args, ctx, plan, rays, cfg_kw, init, nnz):
    """The FP32C median alone at one brick shape (--brick, default the
    benchmark's 16 x 16 x 32 with 512-thread blocks): --steps iterations
    after a warm-up one (stream, slots, graph)."""
    import muontomography_amd as M
    brick = tuple(int(v) for v in (args.brick or "16,16,32,512").split(","))
    kw = dict(cfg_kw)
    kw["update_with_mean"] = 0
    med = ctx.em_plan(M.mlsd_cfg(plan.cfg.grid, **kw), rays=rays)
    plan.close()
    med.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
    med.set_image(init)
    med.run(1, M.MUTO_FP32C)
    med.set_image(init)
    t0 = time.perf_counter()
    med.run(args.steps, M.MUTO_FP32C)
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    log(f"FP32C median: {ms:.2f} ms/it")
    med.close()
    ctx.close()
    print(json.dumps({"workload": "FP32C median, 256^3 grid of 10 mm voxels, 1e7 muons", "brick": list(brick),
                      "entries": int(nnz), "ms_per_step": ms}))


def bench_scans(args):
    """64 independent scans (grids 32^3-128^3, 10-30 mm voxels, 1e4-1e6
    tracks each) run as one batch: FP32C mean update, one launch per pass
    and one graph per iteration for all scans.  Every scan finishes with
    the batch; the scans' latencies run one at a time are reported beside."""
    import muontomography_amd as M
    from muontomography_amd import workloads as W
    from muontomography_amd._abi import mlsd_cfg
    ctx = M.Context(0)
    t0 = time.perf_counter()
    specs = W.scan_batch(n_scans=args.n_scans, seed=77)
    plans = []
    for g, gcfg, nt in specs:
        rays, _ = M.generate(gcfg, nt)
        cfg = mlsd_cfg(g, update_with_mean=1, mean_percentile=100, error_xy=0.5,
                       error_dz_outer=500.0, error_dz_inner=500.0, initialize_method=2)
        p = ctx.em_plan(cfg, rays=rays)
        p.set_image(np.full(g.n_vox, 1e-7))
        plans.append(p)
    nnz = [p.counters()["nnz"] for p in plans]
    log(f"{len(plans)} scans, {sum(nnz)} entries, set-up {time.perf_counter() - t0:.1f} s")
    # brick shape of the batch (taken from the first plan): the fastest of a
    # few candidates over 3 iterations, each batch built and timed in turn
    tune = {}
    for shape in ((16, 16, 16, 256), (16, 16, 32, 512), (16, 8, 32, 256)):
        plans[0].set_option(M.MUTO_OPT_BRICK, M.brick_code(shape))
        b = ctx.batch(plans)
        b.run(1)
        t1 = time.perf_counter()
        b.run(3)
        tune[shape] = (time.perf_counter() - t1) / 3 * 1e3
        b.close()
    brick = min(tune, key=tune.get)
    plans[0].set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
    log("batch brick shapes (ms per iteration): " + ", ".join(f"{k}: {v:.2f}" for k, v in tune.items()))
    for p in plans:
        p.set_image(np.full(p.cfg.grid.n_vox, 1e-7))
    batch = ctx.batch(plans)
    batch.run(args.warmup)
    for p in plans:
        p.set_image(np.full(p.cfg.grid.n_vox, 1e-7))
    t0 = time.perf_counter()
    batch.run(args.steps)
    wall = time.perf_counter() - t0
    alone = []
    for p in plans:
        p.set_image(np.full(p.cfg.grid.n_vox, 1e-7))
        p.run(1, M.MUTO_FP32C)  # brick build, graph capture
        p.set_image(np.full(p.cfg.grid.n_vox, 1e-7))
        t1 = time.perf_counter()
        p.run(args.steps, M.MUTO_FP32C)
        alone.append(time.perf_counter() - t1)
    alone = np.array(alone) * 1e3
    out = {"metric": "many-scan batch: track-voxel updates/s (FP32C mean update)",
           "value": sum(nnz) * args.steps / wall, "unit": "track-voxel updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
           "higher_is_better": True, "dtype": "f32 entry terms, f64/int64 sums, f64 2x2 algebra",
           "data": "synthetic (seeded generator, one block layout per scan)",
           "config": {"workload": f"{len(plans)} scans, grids 32^3-128^3 of 10-30 mm voxels, "
                                  "1e4-1e6 tracks each, 30 iterations", "entries": int(sum(nnz)),
                      "brick": list(brick)},
           "brick_ms": {",".join(map(str, k)): v for k, v in tune.items()},
           "batch_latency_ms": wall * 1e3,
           "alone_latency_ms_p50": float(np.percentile(alone, 50)),
           "alone_latency_ms_p99": float(np.percentile(alone, 99)),
           "alone_total_ms": float(alone.sum()),
           "alone_value": sum(nnz) * args.steps / (alone.sum() * 1e-3)}
    batch.close()
    for p in plans:
        p.close()
    ctx.close()
    print(json.dumps(out))


def bench_siddon(args):
    """Plain PoCA + Siddon at 200^3 x 10 mm with 1e7 muons: PoCA points and
    the Siddon walk of both legs (pin -> PoCA point -> pout) of every muon,
    kernel times from HIP events (host copies excluded), beside the CPU
    oracle on a sample."""
    import muontomography_amd as M
    import oracle
    from muontomography_amd import workloads as W
    from muontomography_amd._abi import poca_cfg
    ctx = M.Context(0)
    g, rays, _ = W.muons("siddon200", n_muons=args.n_muons)
    ctx.poca_points(g, rays[:1000])
    t = {}
    P, v, a, c = ctx.poca_points(g, rays)
    t["poca_points"] = ctx.kernel_ms()
    entries = 0
    t["siddon_walk"] = 0.0
    for p1, p2 in ((rays["pin"], P), (P, rays["pout"])):
        cnt, _ = ctx.siddon_digest(g, p1, p2)
        t["siddon_walk"] += ctx.kernel_ms()
        entries += int(cnt.sum(dtype=np.uint64))
    cfg = poca_cfg(g, length_weighted=1, straight_line=0)
    # the images' first calls load their kernels and size the context's
    # workspaces (kept for later calls); the second calls are the timings
    cold = {}
    ctx.poca_reconstruct(cfg, rays)
    cold["poca_reconstruct"] = ctx.kernel_ms()
    ctx.poca_reconstruct(cfg, rays)
    t["poca_reconstruct"] = ctx.kernel_ms()
    # the reference's other Siddon consumers on the same muons
    ctx.pocut_reconstruct(cfg, rays, 0.9)
    cold["pocut_reconstruct"] = ctx.kernel_ms()
    ctx.pocut_reconstruct(cfg, rays, 0.9)
    t["pocut_reconstruct"] = ctx.kernel_ms()
    ctx.slr_reconstruct(cfg, rays)
    cold["slr_reconstruct"] = ctx.kernel_ms()
    ctx.slr_reconstruct(cfg, rays)
    t["slr_reconstruct"] = ctx.kernel_ms()
    # muon records to rays (MutoFile::getMuonData's per-record work):
    # float32 records of 6 planes, hits smeared by 0.5 mm
    zi, zo = float(rays["pin"][0, 2]), float(rays["pout"][0, 2])
    zl = [zi + 1000.0, zi + 500.0, zi, zo, zo - 500.0, zo - 1000.0]
    rec = W.records_from_rays(rays, zl, sigma=0.5, seed=1)
    ctx.rays_from_records(rec[:1000], zl, 3)
    ctx.rays_from_records(rec, zl, 3)
    t["rays_from_records"] = ctx.kernel_ms()
    # list output (12 B written per entry as u32 index + f64 length, plus
    # the y, z indices of the CSR API) on a sample that fits host memory
    ns = 2_000_000
    row, *_ = ctx.siddon_csr(g, rays["pin"][:ns], P[:ns])
    t_fill = ctx.kernel_ms()
    fill_entries = int(row[-1])
    n_seg = 2 * len(rays)
    walk_rate = entries / (t["siddon_walk"] * 1e-3)
    cpu = None
    cpu_fit = None
    if not args.no_cpu:
        nf = min(1_000_000, len(rays))
        t0 = time.perf_counter()
        oracle.rays_from_records(rec[:nf], zl, 3)
        cpu_fit = {"records_per_s": nf / (time.perf_counter() - t0), "cores": 1,
                   "sample": f"first {nf} records, one thread"}
        nc = min(2_000_000, len(rays))
        t0 = time.perf_counter()
        cc, _ = oracle.siddon_digest(g, rays["pin"][:nc], P[:nc])
        cc2, _ = oracle.siddon_digest(g, P[:nc], rays["pout"][:nc])
        tc = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": float(cc.sum(dtype=np.uint64) + cc2.sum(dtype=np.uint64)) / tc,
               "unit": "Siddon entries/s", "cores": threads, "kind": "port",
               "sample": f"both legs of the first {nc} muons, {tc:.2f} s, OpenMP over segments"}
    out = {"metric": "Siddon voxel-walk entries/s (both PoCA legs), FP64 bit-exact",
           "value": walk_rate, "unit": "Siddon entries/s", "n_gpus": 1, "steps": 1,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded generator)",
           "config": {"workload": "plain PoCA + Siddon, 200^3 grid of 10 mm voxels, 1e7 muons",
                      "muons": int(len(rays)), "segments": n_seg, "entries": entries,
                      "poca_in_grid": c["n_scat_in_grid"]},
           "kernel_ms": t,
           "first_call_ms": cold,
           "segments_per_s": n_seg / (t["siddon_walk"] * 1e-3),
           "poca_points_per_s": len(rays) / (t["poca_points"] * 1e-3),
           "records_per_s": len(rays) / (t["rays_from_records"] * 1e-3),
           "cpu_records": cpu_fit,
           "csr_fill_sample": {"segments": ns, "entries": fill_entries, "ms": t_fill,
                               "write_GBps": fill_entries * 20 / (t_fill * 1e-3) / 1e9},
           "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out))


def bench_map_grad(args):
    """MAP objective + gradient on 128^3 x 15 mm with 1e6 muons.  The line's
    value is the configuration's gradient run, FP32C with the paths walked
    again (a plan that keeps no lists, its passes per tile); beside it FP32C
    from the brick stream, FP64 from stored lists, FP64 walked again, and
    both walked-again variants with one lane per track and one atomic per
    entry; gradient errors against FP64."""
    import muontomography_amd as M
    from muontomography_amd import workloads as W
    from muontomography_amd._abi import mlsd_cfg
    ctx = M.Context(0)
    g, rays, _ = W.muons("map128", n_muons=args.n_muons)
    cfg = mlsd_cfg(g, update_with_mean=1, mean_percentile=100, error_xy=0.5, error_dz_outer=500.0,
                   error_dz_inner=500.0, beta=1.0, p=1.2, delta=0.1, all_neighbours=1)
    plan = ctx.em_plan(cfg, rays=rays)
    nnz = plan.counters()["nnz"]
    # plans without lists: a budget below the lists' 20 bytes per entry that
    # holds the backward's tile records, and one of 0 (one atomic per entry)
    lean = ctx.em_plan(cfg, rays=rays, list_budget=20 * nnz - 1)
    lean0 = ctx.em_plan(cfg, rays=rays, list_budget=0)
    lam = 1e-7 * (0.5 + np.random.default_rng(1).random(g.n_vox))
    res = {}
    first = {}
    for name, pl, prec, rc in (("fp32c", plan, M.MUTO_FP32C, False), ("fp64", plan, M.MUTO_FP64, False),
                               ("fp64_stream", plan, M.MUTO_FP64_STREAM, False),
                               ("fp64_recompute", lean, M.MUTO_FP64, True),
                               ("fp32c_recompute", lean, M.MUTO_FP32C, True),
                               ("fp64_recompute_per_entry", lean0, M.MUTO_FP64, True),
                               ("fp32c_recompute_per_entry", lean0, M.MUTO_FP32C, True)):
        t0 = time.perf_counter()
        sv = pl.map_forward(lam, prec, recompute=rc)
        gr = sv.backward(1.0)
        sv.close()
        first[name] = (time.perf_counter() - t0) * 1e3  # (the tile records are made here)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sv = pl.map_forward(lam, prec, recompute=rc)
            gr = sv.backward(1.0)
            sv.close()
        t = (time.perf_counter() - t0) / args.steps
        res[name] = (t, gr, sv.phi)
        log(f"{name}: {t * 1e3:.2f} ms per value+gradient")
    # the same with lambda and the gradient in device memory (no host
    # copies: MUTO_MAP_DEVICE_BUFFERS), and the FP64 call that allocates
    # nothing (muto_map_value_grad_noalloc)
    from muontomography_amd.devmem import DeviceArray
    lam_d = DeviceArray.from_numpy(lam)
    grad_d = DeviceArray(g.n_vox)
    phi_d = DeviceArray(1)
    dev = {}
    for name, pl, prec, rc in (("fp32c", plan, M.MUTO_FP32C, False), ("fp64", plan, M.MUTO_FP64, False),
                               ("fp64_stream", plan, M.MUTO_FP64_STREAM, False),
                               ("fp64_recompute", lean, M.MUTO_FP64, True)):
        for k in range(args.steps + 1):
            if k == 1:
                t0 = time.perf_counter()
            sv = pl.map_forward(lam_d, prec, recompute=rc)
            sv.backward(1.0, out=grad_d)
            sv.close()
        dev[name] = (time.perf_counter() - t0) / args.steps
        log(f"{name} on device buffers: {dev[name] * 1e3:.2f} ms per value+gradient")
    for name, pl, rc in (("fp64_noalloc", plan, False), ("fp64_recompute_noalloc", lean, True)):
        for k in range(args.steps + 1):
            if k == 1:
                t0 = time.perf_counter()
            pl.map_value_grad_noalloc(lam_d, grad_d, phi_d, recompute=rc)
        dev[name] = (time.perf_counter() - t0) / args.steps
        log(f"{name}: {dev[name] * 1e3:.2f} ms per value+gradient")
    g_dev = grad_d.numpy()
    for a in (lam_d, grad_d, phi_d):
        a.free()
    g64 = res["fp64"][1]
    # the configuration's gradient run: FP32 entry terms with the Siddon
    # paths walked again instead of stored (a plan without lists)
    out = {"metric": "MAP objective value+gradient evaluations/s",
           "value": 1.0 / res["fp32c_recompute"][0], "unit": "evaluations/s", "n_gpus": 1,
           "steps": args.steps, "higher_is_better": True,
           "dtype": "fp32c (f32 entry terms, fp64 walk); fp64 beside", "data": "synthetic",
           "config": {"workload": "MAP gradient, 128^3 grid of 15 mm voxels, 1e6 muons, Siddon paths "
                                  "walked again (no lists kept)",
                      "entries": int(nnz)},
           "value_fp32c_brick_stream": 1.0 / res["fp32c"][0],
           "ms": {k: v[0] * 1e3 for k, v in res.items()},
           "first_call_ms": first,
           "map_tile_records": int(lean.counters()["n_map_tile_records"]),
           "ms_device_buffers": {k: v * 1e3 for k, v in dev.items()},
           "device_value": 1.0 / dev["fp32c"],
           "grad_err_noalloc_recompute_vs_fp64_norm_inf": float(np.abs(g_dev - g64).max() / np.abs(g64).max()),
           "entries_per_s": {k: 2 * nnz / v[0] for k, v in res.items()},
           "grad_err_fp32c_vs_fp64_norm_inf": float(np.abs(res["fp32c"][1] - g64).max() / np.abs(g64).max()),
           "grad_err_recompute_vs_fp64_norm_inf": float(np.abs(res["fp64_recompute"][1] - g64).max() /
                                                         np.abs(g64).max()),
           "grad_err_fp64_stream_vs_fp64_norm_inf": float(np.abs(res["fp64_stream"][1] - g64).max() /
                                                          np.abs(g64).max()),
           "grad_err_fp32c_recompute_vs_fp64_norm_inf": float(
               np.abs(res["fp32c_recompute"][1] - g64).max() / np.abs(g64).max()),
           "grad_err_recompute_tiled_vs_per_entry_norm_inf": float(
               np.abs(res["fp64_recompute"][1] - res["fp64_recompute_per_entry"][1]).max() /
               np.abs(g64).max())}
    plan.close()
    lean.close()
    lean0.close()
    ctx.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed iterations (default 45; scans64: 30, its configuration's count)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="em256",
                    help="em256 (the headline); scans64 (many-scan batch); map128 (MAP gradient); "
                         "siddon200 (plain PoCA + Siddon)")
    ap.add_argument("--n-muons", type=int, default=None)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=5)
    ap.add_argument("--tune-iters", type=int, default=3,
                    help="iterations timed per brick shape by the autotuner (0: 16^3, no tuning)")
    ap.add_argument("--brick", default=None, metavar="BX,BY,BZ[,BLOCK]",
                    help="em256: use this brick shape (and block size), no tuning (profiling runs)")
    ap.add_argument("--n-scans", type=int, default=64, help="scans64: number of scans")
    ap.add_argument("--median-fp32c-only", action="store_true",
                    help="time only the FP32C median (the command of its profile)")
    ap.add_argument("--fp64-stream-only", action="store_true",
                    help="em256: only the FP64 brick stream, mean then median, --steps iterations "
                         "each at --brick (default 16,16,16,256) (profiling runs)")
    ap.add_argument("--cpu-sample", type=int, default=300_000)
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (HBM "
                         "counter calibration for rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"scans64": 30, "map128": 20}.get(args.workload, 45)
    if args.copy_calib:
        import torch
        n = args.copy_calib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        # a read-only kernel (a reduction: reads the array, writes a few
        # bytes) and a write-only one (a fill), same size, to check that the
        # copy's counter factors hold for reads and writes alone
        for _ in range(3):
            s = a.sum()
        for _ in range(3):
            b.fill_(2.0)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3,
                          "sum_bytes_read": 4 * n, "fill_bytes_written": 4 * n,
                          "sum": float(s)}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "scans64":
        return bench_scans(args)
    if args.workload == "map128":
        return bench_map_grad(args)
    if args.workload == "siddon200":
        return bench_siddon(args)

    import muontomography_amd as M
    from muontomography_amd import workloads as W
    from muontomography_amd._abi import mlsd_cfg

    t0 = time.perf_counter()
    g, rays, info = W.muons(args.workload, n_muons=args.n_muons)
    log(f"generated {len(rays)} muons in {time.perf_counter() - t0:.1f} s "
        f"(acceptance {info['acceptance']:.3f})")
    cfg_kw = dict(update_with_mean=1, mean_percentile=100, error_xy=0.5, error_dz_outer=500.0,
                  error_dz_inner=500.0, initialize_method=2, initialize_value=1e-7)
    cfg = mlsd_cfg(g, **cfg_kw)
    init = np.full(g.n_vox, 1e-7)
    ctx = M.Context(0)
    t0 = time.perf_counter()
    plan = ctx.em_plan(cfg, rays=rays)
    c = plan.counters()
    nnz = c["nnz"]
    log(f"plan: {c['n_scat_in_grid']} selected tracks, {nnz} entries "
        f"({nnz / max(c['n_scat_in_grid'], 1):.0f}/track) in {time.perf_counter() - t0:.1f} s")

    if args.fp64_stream_only:
        return bench_fp64_stream_only(args, ctx, plan, rays, cfg_kw, init, nnz)
    if args.median_fp32c_only:
        return bench_median_fp32c_only(args, ctx, plan, rays, cfg_kw, init, nnz)

    # FP32C: brick (LDS tile) shape chosen by the autotuner, warm-up (the
    # stream build and graph capture), then K steps
    plan.set_image(init)
    t0 = time.perf_counter()
    if args.brick:
        brick = M.brick_shape(M.brick_code(tuple(int(v) for v in args.brick.split(","))))
        plan.set_option(M.MUTO_OPT_BRICK, M.brick_code(brick))
        tune = {}
    elif args.tune_iters > 0:
        brick, tune = plan.autotune(iters=args.tune_iters)
    else:
        brick, tune = (16, 16, 16, 256), {}
    log(f"autotune {time.perf_counter() - t0:.1f} s: brick {brick}; " +
        ", ".join(f"{k}: {v:.2f} ms" for k, v in tune.items()))
    plan.set_image(init)
    t0 = time.perf_counter()
    plan.run(args.warmup, M.MUTO_FP32C)
    log(f"FP32C warm-up {args.warmup} it (incl. brick build) {time.perf_counter() - t0:.1f} s; "
        f"segments {plan.counters()['n_segments']}")
    # timed region: K iterations replaying the captured iteration graph
    # (inputs resident in HBM; one synchronisation at the end)
    t0 = time.perf_counter()
    cost32 = plan.run(args.steps, M.MUTO_FP32C)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = nnz * args.steps / wall
    # per-kernel times (HIP events on the plan's stream), outside the timed
    # region, for the roofline
    plan.set_option(M.MUTO_OPT_TIMING, 1)
    plan.run(args.timing_steps, M.MUTO_FP32C)
    plan.set_option(M.MUTO_OPT_TIMING, 0)
    tk = plan.timing()
    n_seg = plan.counters()["n_segments"]
    log(f"FP32C {args.steps} it: {ms_per_step:.3f} ms/it wall; kernels (events, "
        f"{args.timing_steps} it) {tk}")

    # roofline of the dominant kernel
    kern = max(("forward", "backward"), key=lambda k: tk[k])
    kname = {"forward": "k_fwd32", "backward": "k_bwd32"}[kern]
    t_k = tk[kern] * 1e-3
    achieved = BYTES_PER_UPDATE_FP32C * nnz / t_k
    traffic = read_traffic(args.workload, kname)
    # the whole iteration against the same roofline, with the per-voxel and
    # per-track terms of the algorithmic figure: B_iter = nnz 8 + n_vox (4 +
    # 16 + 4) + n_trk 40 bytes (lambda read, accumulator, M_j; Sx, Sy, Dx,
    # Dy, p_r), over the timed region's ms per iteration
    b_iter = BYTES_PER_UPDATE_FP32C * nnz + 24.0 * g.n_vox + 40.0 * c["n_scat_in_grid"]
    roof = {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
            "timing": f"HIP events on the plan's stream, mean of {args.timing_steps} iterations "
                      f"({kname} {tk[kern]:.4f} ms)",
            "iteration_bytes": b_iter,
            "iteration_achieved": b_iter / (ms_per_step * 1e-3) / 1e9,
            "iteration_frac": b_iter / (ms_per_step * 1e-3) / HBM_PEAK}

    extra = {}
    if not args.no_fp64:
        extra = bench_fp64(args, plan, rays, cfg_kw, init, nnz, wall, cost32, brick)
    plan.close()
    ctx.close()

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(g, rays, cfg_kw, min(args.cpu_sample, len(rays)), args.cpu_iters)
        log(f"cpu baseline {cpu['value']:.3e} upd/s in {time.perf_counter() - t0:.1f} s")

    out = {
        "metric": "Siddon+EM track-voxel updates/s, FP32c vs FP64 at <=1e-4 rel err; % HBM peak",
        "value": value,
        "unit": "track-voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 entry terms, f64/int64 sums, f64 2x2 algebra",
        "data": "synthetic (seeded cosmic-muon generator, inputs resident in HBM)",
        "config": {"workload": "batched EM, 256^3 grid of 10 mm voxels, 1e7 muons with "
                               "80-600-voxel paths, mean update, FP32C beside FP64",
                   "grid": [g.nx, g.ny, g.nz], "voxel_mm": g.dx, "muons": int(len(rays)),
                   "selected_tracks": c["n_scat_in_grid"], "entries": nnz,
                   "segments": n_seg, "brick": list(brick),
                   "parallelism": "1 GPU"},
        "kernel_ms": {k: round(v, 4) for k, v in tk.items()},
        "roofline": roof,
        # the backprojection's scatter-adds (one int64 fixed-point LDS atomic
        # per entry) against the LDS atomic rate a microbenchmark measured
        # on this GPU (ds_add_u64: 5.9 lane-ops per cycle per CU)
        "scatter": {"ops_per_s": nnz / (tk["backward"] * 1e-3), "kind": "ds_add_u64 (LDS)",
                    "peak_ops_per_s": LDS_ATOMIC_PEAK,
                    "frac": nnz / (tk["backward"] * 1e-3) / LDS_ATOMIC_PEAK},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
