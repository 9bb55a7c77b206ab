"""Benchmark: RK-stage cell-updates/s on the 200,000-cell 3-D dome snapshot (FP64).

One step = one full RK stage of the mechanical step on device-resident input:
binning, sort, neighbor list (segment-segment tests), Hertz + JKR force and
torque, surface forces and the velocity solve (bfm_dev_stage). value =
cells / stage time over the timed steps.

roofline: the dominant kernel of the stage (per-kernel HIP events on the
launch stream, measured in this run); cpu_baseline: the CPU oracle (the
reference algorithms restated: serial O(N^2) neighbor search, serial pair
loop, OpenMP at the reference's per-cell sites) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "biofilm-abms_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64, vector and matrix alike
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector
# Operations per unit of work of the candidate kernel (adds, muls, divides
# and square roots each 1): the FP64 segment test of every pair that passed
# both FP32 screens (the reference's prefilter, closest points, |d|), and the
# FP32 centre-distance screen of every screened pair (|r_ij|^2 and the bound).
TEST_FLOPS = 80        # FP64, per tested pair
SCREEN_FLOPS = 12      # FP32, per screened pair


# HBM traffic per launch (FETCH_SIZE and WRITE_SIZE, copy-calibrated) of the
# stage kernels, from the committed rocprofv3 summary of the same workload.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "stage_200k_dome_kernels.md")
KERNEL_SYMBOL = {"candidates": "k_candidates_lane", "velocities": "k_velocities_slots",
                 "pair_forces": "k_slot_order", "scatter": "k_scatter", "keys": "k_keys"}


def profiled_traffic(kernel, cells):
    """Read + write HBM bytes per launch of `kernel` measured by rocprofv3 on
    the 200k-cell dome; None for other workloads or kernels."""
    sym = KERNEL_SYMBOL.get(kernel)
    if cells != 200000 or sym is None or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cols = [c.strip() for c in line.split("|")]
        if len(cols) > 7 and sym in cols[1]:
            try:
                return (float(cols[6]) + float(cols[7])) * 1e6
            except ValueError:
                return None
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=200000)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--jkr", default="exact", choices=["exact", "table", "surrogate", "aniso", "none"],
                    help="JKR radius: exact FP64 solve, 1-D table, learned radius, anisotropic tables "
                         "(adhesion_mode 2, synthetic tables), or no adhesion (Hertz only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-JKR force error and stage time on the snapshot")
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU work for the sampled O(N^2) neighbor search")
    ap.add_argument("--bin-scale", type=float, default=0.0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--workload", default="stage", choices=["stage", "growth", "ensemble", "layer2d", "grow2d"],
                    help="stage: the RK-stage bench (default); growth: the founder-to-N growth run; "
                         "ensemble: batched replicates")
    ap.add_argument("--ens-reps", type=int, default=256)
    ap.add_argument("--ens-lo", type=int, default=500)
    ap.add_argument("--ens-hi", type=int, default=20000)
    ap.add_argument("--ens-minutes", type=float, default=10.0, help="simulated minutes per replicate")
    ap.add_argument("--ens-max-wall", type=float, default=240.0)
    ap.add_argument("--ens-tableau", default="bogacki_shampine")
    ap.add_argument("--ens-jkr", default="surrogate", choices=["exact", "surrogate"])
    ap.add_argument("--growth-cells", type=int, default=100000)
    ap.add_argument("--growth-tableau", default="bogacki_shampine")
    ap.add_argument("--growth-max-wall", type=float, default=900.0)
    ap.add_argument("--growth-seed", type=int, default=43,
                    help="division stream of the founder run; with 42 two cells come to a segment "
                         "distance of 0 near 130 cells and the step is not finite (the oracle's step on "
                         "that state is not finite either)")
    ap.add_argument("--grow2d-cells", type=int, default=2000)
    ap.add_argument("--grow2d-seed", type=int, default=43,
                    help="division stream of the 2-D run; with 42 a length-ratio draw clipped to 0 leaves a "
                         "daughter of length 0 at 810 cells, whose drag l eta vanishes (the reference's "
                         "velocities of that cell are infinite there too)")
    ap.add_argument("--grow2d-cpu-cells", type=int, default=150,
                    help="the oracle's 2-D loop runs the same founder to this many cells, a bounded "
                         "sample of about 20 s (0: skip)")
    ap.add_argument("--growth-cpu-cells", type=int, default=64,
                    help="target of the oracle growth run timed beside the GPU one (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy (calibrates FETCH/WRITE_SIZE)")
    return ap.parse_args()


def learned_jkr_check(ctx, d_cells, cells, n, thr, R, Ldiv, exact_ms, steps):
    """The learned JKR radius on the snapshot's own contacts: the stage's
    list (exported from the device), overlap capped at the shell thickness as
    the force uses it (adhesion.hpp:195-196), gamma = min(gamma_i, gamma_j);
    force error |F - F_exact| / (|term1| + |term2|) of the network alone and
    after the one FP64 Newton step the product applies, against the converged
    FP64 Newton root (40 steps); contacts outside the training envelope take
    the exact solve. Also the stage time with the learned radius."""
    import numpy as np
    import torch

    import bfm
    from bfm import surrogate as sur

    m = ctx.stats()["m"]
    out = torch.zeros((7, max(m, 1)), dtype=torch.float64, device="cuda")
    ctx.dev_export_neighbors(out.data_ptr(), max(m, 1))
    ctx.sync()
    torch.cuda.synchronize()
    nb = out.cpu().numpy().T[:m]
    i, j = nb[:, 0].astype(np.int64), nb[:, 1].astype(np.int64)
    dist = np.sqrt((nb[:, 2] ** 2 + nb[:, 3] ** 2) + nb[:, 4] ** 2)
    delta = 2 * R - dist
    contact = delta > 0
    Rcell, E0 = 0.5, 3900.0
    Req = R / 2
    delta_c = np.minimum(delta[contact], 2 * (R - Rcell))
    gamma = np.minimum(cells[i[contact], 21], cells[j[contact], 21])
    env = (delta_c >= -0.2 * R) & (delta_c <= 0.3 * R) & (gamma >= 10.0) & (gamma <= 1000.0)
    p_s = bfm.core.make_params(jkr_mode=bfm.core.JKR_SURROGATE)
    ctx_s = bfm.Context(n, max_pairs=16 * n, params=p_s)
    sur.install(ctx_s, sur.load())
    dd, gg = delta_c[env], gamma[env]
    _, r0 = sur.evaluate(ctx_s, dd, gg, polish=0)
    _, r1 = sur.evaluate(ctx_s, dd, gg, polish=1)
    _, rx = sur.evaluate(ctx_s, dd, gg, polish=40)

    def force(a):
        a3 = a ** 3
        t1 = 4 * E0 * a3 / (3 * Req)
        t2 = 4 * np.sqrt(np.pi * a3 * gg * E0)
        return t1 - t2, np.abs(t1) + np.abs(t2)

    fx, scale = force(rx)
    e0 = np.abs(force(r0)[0] - fx) / scale
    e1 = np.abs(force(r1)[0] - fx) / scale
    stream = torch.cuda.Stream()
    ctx_s.set_stream(stream.cuda_stream)
    d_vel = torch.zeros((6, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        ctx_s.dev_stage(d_cells.data_ptr(), n, n, thr, R, Ldiv, d_vel.data_ptr())
    ctx_s.sync()
    e_a, e_b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e_a.record(stream)
    for _ in range(steps):
        ctx_s.dev_stage(d_cells.data_ptr(), n, n, thr, R, Ldiv, d_vel.data_ptr())
    e_b.record(stream)
    torch.cuda.synchronize()
    ms_s = e_a.elapsed_time(e_b) / steps
    return {"contacts": int(contact.sum()), "in_envelope": int(env.sum()),
            "force_err_max_polished": float(e1.max()) if e1.size else None,
            "force_err_rms_polished": float(np.sqrt(np.mean(e1 ** 2))) if e1.size else None,
            "force_err_max_network_only": float(e0.max()) if e0.size else None,
            "meets_1e-4": bool(e1.size and e1.max() <= 1e-4),
            "stage_ms_learned": ms_s, "stage_ms_exact": exact_ms,
            "learned_vs_exact_speed": exact_ms / ms_s,
            "reference_radius": "FP64 Newton on the JKR quartic from the network's guess, 40 steps"}


def host_cores():
    """(threads the baseline uses, cores available to this process, host
    logical CPUs, CPU model): every core the process may run on, unless
    OMP_NUM_THREADS limits OpenMP."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or avail
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, avail, os.cpu_count(), model


def cpu_baseline(cells, thr, R, Ldiv, params, budget, aniso_tables=None):
    """Times the oracle on the same snapshot: C1 = getVelocities on a prebuilt
    list (all N cells, OpenMP at the reference sites), C2 = the serial O(N^2)
    getCellNeighbors, sampled on a block of rows centred on N/2 (row i costs
    O(i), so the block has the mean row cost) and scaled to N rows."""
    import numpy as np
    from oracle import oracle as orc

    if aniso_tables is not None:
        orc.set_aniso_tables(aniso_tables)

    n = cells.shape[0]
    threads, avail, nproc, model = host_cores()
    nb = orc.get_cell_neighbors(cells, thr, R, Ldiv, binned=True, nthreads=threads)
    a2d = (cells[:, 6] < params.nz_threshold).astype(np.int32)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        orc.get_velocities(cells, nb, a2d, None, params)
        times.append(time.perf_counter() - t0)
    c1 = float(np.median(times))
    rows = 200
    t0 = time.perf_counter()
    orc.count_neighbor_rows(cells, thr, R, Ldiv, n // 2 - rows // 2, n // 2 + rows // 2)
    t_probe = time.perf_counter() - t0
    rows = int(min(n, max(200, rows * budget / max(t_probe, 1e-6))))
    t0 = time.perf_counter()
    orc.count_neighbor_rows(cells, thr, R, Ldiv, n // 2 - rows // 2, n // 2 - rows // 2 + rows)
    t_rows = time.perf_counter() - t0
    c2 = t_rows * n / rows
    return dict(value=n / (c1 + c2), unit="cell-updates/s", cores=threads, kind="port",
                host={"cores_available": avail, "nproc": nproc, "cpu_model": model},
                sample=(f"oracle on the same {n}-cell snapshot: getVelocities on a prebuilt list "
                        f"(all cells, {threads} OpenMP threads, median of 3: {c1*1e3:.1f} ms) + "
                        f"serial O(N^2) getCellNeighbors timed on {rows} rows centred on N/2 "
                        f"({t_rows:.1f} s) and scaled to N rows ({c2:.1f} s)"),
                stage_only_value=n / c1, velocities_s=c1, neighbors_s=c2)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path "
              "would shard by spatial slabs or by replicates, one process per GPU", file=sys.stderr)
        sys.exit(2)
    if args.workload == "growth":
        return growth(args)
    if args.workload == "ensemble":
        return ensemble(args)
    if args.workload == "grow2d":
        return grow2d(args)
    if args.workload == "layer2d":
        return layer2d(args)
    import numpy as np
    import torch

    import bfm

    if args.copy_calibration:
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8,
                          "launches": 3}))
        return
    n = args.cells
    R, Rcell, Ldiv = 0.8, 0.5, 2.4
    thr = 0.5 * (2 * R + Ldiv)
    cells, dome_r = bfm.make_dome(n, seed=args.seed)
    jkr_mode = {"exact": bfm.core.JKR_EXACT, "table": bfm.core.JKR_TABLE_1D,
                "surrogate": bfm.core.JKR_SURROGATE, "none": bfm.core.JKR_EXACT,
                "aniso": bfm.core.JKR_EXACT}[args.jkr]
    params = bfm.core.make_params(jkr_mode=jkr_mode, adhesion=args.jkr != "none",
                                  anisotropic=args.jkr == "aniso")
    ctx = bfm.Context(n, max_pairs=16 * n, params=params)
    aniso = None
    if args.jkr == "aniso":
        # tables of the reference's shapes: 100 x 50 x 100 contact-point radii,
        # 40 x 40 x 20 x 50 (force, radius) entries (25.6 MB)
        from bfm.jkr import synthetic_aniso_tables
        aniso = synthetic_aniso_tables(n_theta=100, n_half_l=50, n_coords=100, n_rx=40, n_phi=20, n_delta=50)
        ctx.set_aniso_jkr_tables(aniso)
    if args.bin_scale or args.block or args.variant >= 0:
        ctx.set_tuning(args.bin_scale, args.block, args.variant)
    if args.jkr == "table":
        from bfm.jkr import reference_table
        ov, rad = reference_table(R, Rcell, params.E0, 100.0)
        ctx.set_jkr_table(ov, rad)
    if args.jkr == "surrogate":
        from bfm import surrogate as sur
        sur.install(ctx, sur.load())
    stream = torch.cuda.Stream()
    ctx.set_stream(stream.cuda_stream)
    with torch.cuda.stream(stream):
        d_cells = torch.from_numpy(np.ascontiguousarray(cells.T)).to("cuda")
        d_vel = torch.zeros((6, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def stage():
        ctx.dev_stage(d_cells.data_ptr(), n, n, thr, R, Ldiv, d_vel.data_ptr())

    for _ in range(args.warmup):
        stage()
    ctx.sync()
    torch.cuda.synchronize()
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t_start.record(stream)
    for _ in range(args.steps):
        stage()
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ctx.sync()
    ms = t_start.elapsed_time(t_end) / args.steps
    value = n / (ms * 1e-3)

    # per-kernel times (HIP events around each launch on the launch stream)
    ctx.enable_timing(True)
    ctx.reset_timing()
    ctx.stats()  # resets the candidate counters
    for _ in range(args.steps):
        stage()
    kt = ctx.kernel_times()
    st = ctx.stats()
    ctx.enable_timing(False)
    # the pair count from the neighbor list proper (the stage itself keeps
    # only per-cell lists), built after the timed work
    m = ctx.dev_build_neighbors(d_cells.data_ptr(), n, n, thr, R, Ldiv)
    ctx.stats()
    per_kernel = {k: v[0] / v[1] for k, v in kt.items()}
    dom = max(kt, key=lambda k: kt[k][0])
    tests_per_launch = st["segment_tests"] / args.steps
    screened_per_launch = st["screened"] / args.steps
    if dom == "candidates":
        # FP64 work only: the segment tests of the pairs that passed the two
        # FP32 screens; the centre-distance screen is FP32 and reported beside it
        t_s = per_kernel[dom] * 1e-3
        flops64 = tests_per_launch * TEST_FLOPS
        achieved = flops64 / t_s / 1e12
        screen32 = screened_per_launch * SCREEN_FLOPS / t_s / 1e12
        roof = dict(kernel=dom, bound="mfma", achieved=achieved, peak=FP64_PEAK_TFLOPS,
                    unit="TFLOP/s", frac=achieved / FP64_PEAK_TFLOPS,
                    traffic=profiled_traffic(dom, n),
                    engine="FP64 vector ALU: MI355X's FP64 vector and matrix peaks are both 78.6 "
                           "TFLOP/s; the kernel is bound by the returning device-scope atomic per kept pair "
                           "(the partner's per-cell slot), the VALU issue of its FP32 screen loop and the "
                           "latency of its screen-test chain, not by either peak",
                    avg_us=per_kernel[dom] * 1e3, fp64_flops_per_launch=flops64,
                    tests_per_launch=tests_per_launch,
                    fp32_screen={"screened_per_launch": screened_per_launch,
                                 "flops_per_launch": screened_per_launch * SCREEN_FLOPS,
                                 "achieved_tflops": screen32, "peak_tflops": FP32_PEAK_TFLOPS,
                                 "frac": screen32 / FP32_PEAK_TFLOPS},
                    hbm={"algorithmic_bytes": n * 104 + m * 60,
                         "achieved_gbs": (n * 104 + m * 60) / t_s / 1e9, "peak_gbs": HBM_PEAK_GBS})
    else:
        # HBM-bound per-cell / per-pair kernels: algorithmic bytes 176 per cell + 48 per pair
        bytes_ = n * 176 + m * 48
        achieved = bytes_ / (per_kernel[dom] * 1e-3) / 1e9
        roof = dict(kernel=dom, bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=profiled_traffic(dom, n),
                    avg_us=per_kernel[dom] * 1e3,
                    algorithmic_bytes=bytes_)
    stage_bytes = n * 328 + m * 96
    learned = None
    if args.jkr == "exact" and not args.no_learned:
        learned = learned_jkr_check(ctx, d_cells, cells, n, thr, R, Ldiv, ms, args.steps)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cells, thr, R, Ldiv, params, args.cpu_budget, aniso)
    out = {
        "metric": "RK-stage cell-updates/s, 200k-cell 3D FP64; surrogate vs exact JKR @1e-4 err",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "static 200,000-cell 3-D dome snapshot: one RK stage (binning, sort, "
                               "neighbor list, Hertz+JKR force/torque, velocities)",
                   "cells": n, "pairs": m, "pairs_per_cell": m / n, "dome_radius_um": dome_r,
                   "seed": args.seed, "jkr": args.jkr, "neighbor_threshold_um": thr,
                   "bin_edge_um": st["bin_edge"], "nbins": st["nbins"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "learned_jkr": learned,
        "stage_hbm_frac": (stage_bytes / (ms * 1e-3) / 1e9) / HBM_PEAK_GBS,
        "jkr_path_us": round(1e3 * (per_kernel.get("pair_forces", 0) + per_kernel.get("sur_stage", 0)
                                    + per_kernel.get("sur_network", 0)), 2),
        "kernel_us": {k: round(v * 1e3, 2) for k, v in sorted(per_kernel.items(),
                                                              key=lambda x: -x[1])},
        "wall_s_timed": wall,
    }
    print(json.dumps(out))


def growth(args):
    """Founder cell grown to --growth-cells on the GPU (bfm_simulate): wall time,
    iterations, RK step evaluations and the distribution of accepted dt; the
    oracle's loop timed on the host to a small target with the same stream."""
    import numpy as np

    import bfm
    from bfm import sim
    from bfm.core import make_params

    target = args.growth_cells
    cap = int(target * 1.1) + 1024
    cfg = sim.sim_config(n_cells=target, tableau=args.growth_tableau, seed=args.growth_seed,
                         report_seconds=20.0, max_wall_seconds=args.growth_max_wall)
    params = make_params()
    ctx = bfm.Context(cap, max_pairs=16 * cap, params=params)
    c0 = sim.founder_cell()
    cells, parents, trace, st = sim.simulate(ctx, cfg, c0, cap=cap, trace_cap=50_000_000)
    n = cells.shape[0]
    reached = n >= target
    lt = np.log10(trace)
    hist, edges = np.histogram(lt, bins=np.arange(-9.5, -2.4, 0.5))
    gpu = dict(n_final=n, reached_target=bool(reached), wall_s=st["wall_seconds"],
               iterations=st["iterations"], rk_steps=st["rk_steps"], divisions=st["divisions"],
               neighbor_builds=st["neighbor_builds"], t_end_h=st["t_end"],
               host_division_s=st["host_division_seconds"],
               us_per_iteration=1e6 * st["wall_seconds"] / max(st["iterations"], 1),
               dt_percentiles_h={str(q): float(np.percentile(trace, q)) for q in (1, 10, 50, 90, 99)},
               dt_log10_histogram={f"{edges[i]:.1f}": int(hist[i]) for i in range(len(hist))})
    cpu = None
    if args.growth_cpu_cells > 0 and not args.no_cpu:
        from oracle import oracle as orc
        cfg_c = sim.sim_config(n_cells=args.growth_cpu_cells, tableau=args.growth_tableau,
                               seed=args.growth_seed)
        cap_c = 2 * args.growth_cpu_cells + 64
        t0 = time.perf_counter()
        cc, tr_c, it_c, steps_c = orc.simulate(cfg_c, params, c0, cap=cap_c)
        t_cpu = time.perf_counter() - t0
        ctx_s = bfm.Context(cap_c, max_pairs=32 * cap_c + 1024, params=params)
        cg, _, tr_g, st_g = sim.simulate(ctx_s, cfg_c, c0, cap=cap_c)
        cpu = dict(cells=args.growth_cpu_cells, oracle_wall_s=t_cpu, gpu_wall_s=st_g["wall_seconds"],
                   iterations=it_c, gpu_iterations=st_g["iterations"],
                   oracle_us_per_iteration=1e6 * t_cpu / max(it_c, 1),
                   gpu_us_per_iteration=1e6 * st_g["wall_seconds"] / max(st_g["iterations"], 1),
                   same_iterations=bool(it_c == st_g["iterations"]),
                   max_abs_state_diff=float(np.max(np.abs(cc[:, 1:7] - cg[:, 1:7])))
                   if cc.shape == cg.shape else None, kind="port",
                   cores=host_cores()[0])
    out = {"metric": f"founder-to-{target} growth wall time (3D, FP64, tol 1e-8)",
           "value": st["wall_seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"growth from one founder cell to {target} cells",
                      "tableau": args.growth_tableau, "seed": args.growth_seed, "tol": 1e-8,
                      "max_stepsize_h": cfg.max_stepsize, "min_stepsize_h": cfg.min_stepsize,
                      "iter_update_neighbors": cfg.iter_update_neighbors,
                      "daughter_angle_xy_bound": cfg.daughter_angle_xy_bound,
                      "daughter_angle_z_bound": cfg.daughter_angle_z_bound},
           "gpu": gpu, "cpu_baseline": cpu}
    print(json.dumps(out))


def grow2d(args):
    """The 2-D tree's growth run (runSimulationAdaptiveLagrangian of ABM-2d) from
    the 2-D driver's founder cell to --grow2d-cells cells on the GPU
    (bfm2d_simulate): R 0.8, L0 0.4 (division at l > 2.4), growth 1.5 +- 0.15
    per h, daughters N(0.5, 0.125) of l - 2R, E0 3.9 kPa / Ecell 3.9 MPa,
    isotropic JKR adhesion at 0.1 mJ/m^2 between group-1 cells, drag 0.7 Pa h
    (surface 720), Bogacki-Shampine at tolerance 1e-8, seed 42. The oracle's
    loop runs the same founder with the same division stream on one host core
    (--grow2d-cpu-cells), and the two runs are compared."""
    import numpy as np

    from bfm.model2d import (Model2D, founder_cell_2d, make_params, sim_config_2d, simulate_2d,
                             ADHESION2D_JKR_ISOTROPIC)

    target = args.grow2d_cells
    cap = int(target * 1.1) + 64
    p = make_params(adhesion_mode=ADHESION2D_JKR_ISOTROPIC, gamma=100.0)
    amap = np.array([[0, 0], [0, 1]], dtype=np.int32)
    cfg = sim_config_2d(n_cells=target, seed=args.grow2d_seed, report_seconds=20.0,
                        max_wall_seconds=args.growth_max_wall)
    m2 = Model2D(cap, p, max_pairs=16 * cap + 1024)
    c0 = founder_cell_2d()
    cells, parents, trace, st = simulate_2d(m2, cfg, c0, amap, cap=cap, trace_cap=5_000_000)
    n = cells.shape[0]
    # stage evaluations: every RK attempt evaluates the s = 4 stages for all cells
    gpu = dict(n_final=n, reached_target=bool(n >= target), wall_s=st["wall_seconds"],
               iterations=st["iterations"], rk_attempts=st["rk_steps"], divisions=st["divisions"],
               neighbor_builds=st["neighbor_builds"], t_end_h=st["t_end"],
               host_division_s=st["host_division_seconds"],
               us_per_iteration=1e6 * st["wall_seconds"] / max(st["iterations"], 1),
               dt_percentiles_h={str(q): float(np.percentile(trace, q)) for q in (1, 10, 50, 90, 99)})
    cpu = None
    if args.grow2d_cpu_cells > 0 and not args.no_cpu:
        from oracle import oracle as orc
        tgt_c = min(args.grow2d_cpu_cells, target)
        cfg_c = sim_config_2d(n_cells=tgt_c, seed=args.grow2d_seed)
        cap_c = int(tgt_c * 1.1) + 64
        t0 = time.perf_counter()
        cc, par_c, tr_c, it_c, steps_c = orc.simulate_2d(cfg_c, p, amap, c0, cap=cap_c, trace_cap=5_000_000)
        t_cpu = time.perf_counter() - t0
        if tgt_c == target:
            cg, par_g, tr_g, st_g = cells, parents, trace, st
        else:
            m2c = Model2D(cap_c, p, max_pairs=16 * cap_c + 1024)
            cg, par_g, tr_g, st_g = simulate_2d(m2c, cfg_c, c0, amap, cap=cap_c, trace_cap=5_000_000)
        same = cc.shape == cg.shape and it_c == st_g["iterations"]
        cpu = {"value": t_cpu, "unit": "s", "cores": 1, "kind": "port",
               "sample": f"the oracle's 2-D growth loop from the same founder to {tgt_c} cells, one thread",
               "iterations": it_c, "rk_attempts": steps_c, "gpu_wall_s_same_run": st_g["wall_seconds"],
               "same_events": bool(same and np.array_equal(par_c, par_g)),
               "max_abs_state_diff": float(np.max(np.abs(cc[:, 1:5] - cg[:, 1:5]))) if same else None,
               "max_rel_dt_diff": float(np.max(np.abs(tr_c - tr_g) / tr_c)) if same else None,
               "cpu_model": host_cores()[3], "host_logical_cpus": host_cores()[2]}
    out = {"metric": f"founder-to-{target} 2-D growth wall time (FP64, tol 1e-8)",
           "value": st["wall_seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"2-D growth from one founder cell to {target} cells (isotropic JKR, "
                                  f"gamma 100 Pa um, Bogacki-Shampine)",
                      "seed": args.grow2d_seed, "tol": 1e-8, "max_stepsize_h": cfg.max_stepsize,
                      "min_stepsize_h": cfg.min_stepsize,
                      "iter_update_neighbors": cfg.iter_update_neighbors,
                      "daughter_angle_bound": cfg.daughter_angle_xy_bound},
           "gpu": gpu, "cpu_baseline": cpu}
    print(json.dumps(out))


def layer2d(args):
    """The 2-D model (ABM-2d) on a layer of --cells cells in rows (JKR
    adhesion for same-group pairs): the device neighbor search with the
    driver's to_adhere rule, then Bogacki-Shampine steps on device arrays
    (bfm2d_dev_step_rk), timed with HIP events on the context's stream; a
    second pass with per-kernel events gives the breakdown."""
    import numpy as np
    import torch

    import bfm
    from bfm.model2d import Model2D, make_params, random_layer, dev_build_neighbors, dev_step_rk

    n = args.cells
    c = random_layer(n, args.seed)
    p = make_params(adhesion_mode=1, strength=40.0, mindist=1.2)
    m2 = Model2D(n, p, max_pairs=8 * n)
    stream = torch.cuda.Stream()
    m2.ctx.set_stream(stream.cuda_stream)
    with torch.cuda.stream(stream):
        a = torch.from_numpy(np.ascontiguousarray(c.T)).to("cuda")
        b_ = torch.empty_like(a)
    torch.cuda.synchronize()
    amap = np.array([[0, 0], [0, 1]], dtype=np.int32)
    rows = dev_build_neighbors(m2, a.data_ptr(), n, n, 2.0, 0.8, 2.4, amap)
    A, b, bs, _ = bfm.tableau("bogacki_shampine")

    def step():
        dev_step_rk(m2, A, b, bs, 1e-5, a.data_ptr(), n, n, b_.data_ptr(), n)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    step_us = 1e3 * e0.elapsed_time(e1) / args.steps
    m2.ctx.enable_timing(True)
    m2.ctx.reset_timing()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    kt = m2.ctx.kernel_times()
    per = {k: 1e3 * ms / max(cnt, 1) for k, (ms, cnt) in kt.items()}
    # k2d_stage, per launch: each row's 64-byte record and 4-byte CSR entry
    # once per role, seven cell columns, the earlier stages' velocities (on
    # average 1.5 of 4 x 8 B for the 4 stages) and its own 4 x 8 B
    stage_bytes = rows * 2 * (64 + 4) + n * (7 * 8 + 48 + 32)
    st_us = per.get("stage_2d", 0.0)
    roof = {"kernel": "stage_2d", "bound": "hbm", "achieved": stage_bytes / (st_us * 1e-6) / 1e9 if st_us else None,
            "peak": 8000.0, "unit": "GB/s", "traffic": None, "algorithmic_bytes": stage_bytes, "avg_us": st_us}
    roof["frac"] = roof["achieved"] / roof["peak"] if roof["achieved"] else None
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        from bfm.model2d import to_adhere_rule
        # the device list (bit-identical to the oracle's O(N^2) search, which
        # takes about a minute here) feeds the oracle's step
        exp = torch.zeros((7, max(rows, 1)), dtype=torch.float64, device="cuda")
        m2.ctx._check(m2.ctx.lib.bfm_dev_export_neighbors(m2.ctx.h, exp.data_ptr(), max(rows, 1)))
        torch.cuda.synchronize()
        nb = np.asfortranarray(exp.cpu().numpy().T[:rows][:, [0, 1, 2, 3, 5, 6]])
        ta = to_adhere_rule(c, nb, 0.8, 0.5, amap)
        t0 = time.perf_counter()
        orc.step_rk_2d(A, b, bs, 1e-5, c, nb, p, ta)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": n / t_cpu, "unit": "cell-updates/s", "cores": 1, "kind": "port",
               "sample": f"oracle stepRungeKuttaAdaptive of the 2-D tree (Bogacki-Shampine) on the same "
                         f"{n}-cell layer and rows, one step, one thread: {t_cpu:.3f} s"}
    out = {"metric": "2-D RK step cell-updates/s (FP64, device arrays)", "value": n / (step_us * 1e-6),
           "unit": "cell-updates/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{n}-cell 2-D layer in rows, JKR adhesion, Bogacki-Shampine step",
                      "cells": n, "rows": int(rows), "seed": args.seed},
           "us_per_step": step_us, "roofline": roof, "cpu_baseline": cpu,
           "kernel_us": {k: round(v, 2) for k, v in sorted(per.items(), key=lambda kv: -kv[1])}}
    print(json.dumps(out))


def ensemble(args):
    """Batched replicates: --ens-reps domes with log-uniform cell counts in
    [--ens-lo, --ens-hi] (seeds 0..R-1), two phenotypes (gamma 500 / 50 Pa um,
    groups 1 / 2) assigned 50/50 per cell by a seeded draw, advanced --ens-minutes simulated minutes
    (or until --ens-max-wall seconds) by bfm_ensemble_simulate with the
    autotuner and the learned JKR radius. value = RK stage cell-updates per
    second over the whole run (every attempt of every replicate)."""
    import numpy as np

    import bfm
    from bfm import sim
    from bfm.core import make_params, JKR_SURROGATE, JKR_EXACT

    sizes = bfm.ensemble_sizes(args.ens_reps, args.ens_lo, args.ens_hi, seed=0)
    # half-lengths below Ldiv / 2 - R: no cell starts past the division length
    reps, groups = bfm.make_ensemble(sizes, seeds=range(args.ens_reps), hl_hi=1.2)
    group_all = np.concatenate(groups)
    total = int(sizes.sum())
    cap = int(total * 1.3) + 1024
    params = make_params(jkr_mode=JKR_SURROGATE if args.ens_jkr == "surrogate" else JKR_EXACT)
    ctx = bfm.Context(cap, max_pairs=16 * cap, params=params)
    if args.ens_jkr == "surrogate":
        from bfm import surrogate as sur
        sur.install(ctx, sur.load())
    cfg = sim.sim_config(max_time=args.ens_minutes / 60.0, tableau=args.ens_tableau, seed=0,
                         report_seconds=20.0, max_wall_seconds=args.ens_max_wall)
    out, t_out, it_out, st = sim.ensemble_simulate(ctx, cfg, reps, list(range(args.ens_reps)),
                                                   autotune=True, cap=cap)
    value = st["stage_cell_updates"] / st["wall_seconds"]
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        A, b, bs, _ = bfm.tableau(args.ens_tableau)
        p_exact = make_params()
        t_cpu, upd, used = 0.0, 0, []
        for r in np.argsort(sizes)[len(sizes) // 2:len(sizes) // 2 + 4]:
            cells = reps[r]
            nb = orc.get_cell_neighbors(cells, cfg.neighbor_threshold, cfg.R, cfg.Ldiv, binned=True)
            t0 = time.perf_counter()
            orc.step_rk(A, b, bs, 1e-6, cells, nb, p_exact)
            t_cpu += time.perf_counter() - t0
            upd += cells.shape[0] * len(b)
            used.append(int(cells.shape[0]))
        threads = host_cores()[0]
        cpu = dict(value=upd / t_cpu, unit="stage cell-updates/s", cores=threads, kind="port",
                   sample=f"oracle stepRungeKuttaAdaptive ({args.ens_tableau}, exact JKR, prebuilt lists) "
                          f"on 4 median-size replicates {used}, one step each")
    out_j = {"metric": "batched-ensemble RK stage cell-updates/s (3D, FP64)",
             "value": value, "unit": "cell-updates/s", "n_gpus": 1, "higher_is_better": True,
             "vs_baseline": None, "dtype": "f64", "data": "synthetic",
             "config": {"workload": f"{args.ens_reps} replicate domes, {args.ens_lo}-{args.ens_hi} cells "
                                    f"log-uniform, gamma 500/50 Pa um per cell 50/50, {args.ens_minutes} simulated min",
                        "replicates": args.ens_reps, "cells_start": total, "cells_end": st["n_final"],
                        "tableau": args.ens_tableau, "jkr": args.ens_jkr,
                        "phenotype_cells": {"group 1 (gamma 500)": int((group_all == 1).sum()),
                                            "group 2 (gamma 50)": int((group_all == 2).sum())}},
             "ensemble": {"wall_s": st["wall_seconds"], "iterations": st["iterations"],
                          "attempts": st["attempts"], "divisions": st["divisions"],
                          "neighbor_builds": st["neighbor_builds"],
                          "host_division_s": st["host_division_seconds"],
                          "autotune_ms_per_build": st["autotune_ms"],
                          "tuned_bin_scale": st["tuned_bin_scale"], "tuned_variant": st["tuned_variant"],
                          "failed_replicates": st["failed_replicates"],
                          "t_reached_min_h": float(t_out.min()), "t_reached_median_h": float(np.median(t_out)),
                          "t_target_h": args.ens_minutes / 60.0,
                          "replicates_finished": int(np.sum(t_out >= args.ens_minutes / 60.0))},
             "cpu_baseline": cpu}
    print(json.dumps(out_j))


if __name__ == "__main__":
    main()
