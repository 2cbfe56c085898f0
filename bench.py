"""Benchmark of the bem2d hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-extras]
    python bench.py --config large|kite|batched|serving

Headline workload (one step): the large multi-right-hand-side configuration -- sound-soft
scattering by an open circular-arc cavity (radius 1, aperture 0.3 rad centred on +x) at k = 500,
combined-field operator (eta = k), N = 32000 piecewise-constant elements at uniform parameter
spacing, 16-point Gauss with the singular rules at L = 3, sigma = 0.15, and 360 plane-wave
right-hand sides at equispaced angles: Galerkin assembly A = M + 2K (FP64), the right-hand sides
b = 2<f, phi>, the FP32 MFMA LU and FP64 refinement to backward error 1e-12 (the mixed path).  The
plain FP64 path (FP64 assembly + rocSOLVER zgetrf/zgetrs) is timed beside it.  Geometry tables are
resident in HBM before the timed region starts.  value = right-hand sides solved per second.

The same line carries, unless --no-extras, the kite configuration (N = 4000, k = 100, one right-hand
side) on both paths, the 64-problem batch, the full 2000-request served trace (Poisson arrivals;
p50 / p99 latency) and eight kite problems served together.  Progress lines go to stderr; stdout
carries the one JSON line.

--config large / kite / batched / serving prints the line of that configuration alone.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/sec & Gentries/sec assembly, p50/p99 latency; FP32+refine vs FP64 at 1e-10"
FLOP_PER_PAIRPOINT = 128  # convention for the combined-field kernel evaluation
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_assemble_regular_kite4000.md")
FP64_PROFILE = os.path.join(ROOT, "profiles", "fp64_assemble_regular_kite4000.md")
LU_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "traffic_trailing_update_arc32000.md")
ASM_C5_PROFILE = os.path.join(ROOT, "profiles", "fp64_assemble_regular_arc32000.md")


def progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def lu_update_flops(n, pair_tiles, quad_tiles=1 << 30, oct_tiles=1 << 30):
    """Algorithmic FLOP of the FP32 LU's trailing-update launches for an n x n matrix: 8 real FLOP
    per complex multiply-add, in the launch schedule of bem2d_internal_lu_advance (64-wide panels on
    the matrix bordered to whole 64-blocks; while the trailing matrix has >= oct_tiles tiles eight
    panels, >= quad_tiles four, >= pair_tiles two share one trailing update of depth 512, 256, 128,
    each step of the group but the last updating only the cross -- the next block row and column --
    with the group's panels so far).  The total does not depend on the grouping.  Returns (FLOP,
    launches)."""
    nblk = (n + 63) // 64
    total, pend, group = 0, 0, 1
    for b in range(nblk - 1):
        tiles = nblk - b - 1
        if pend == 0:
            group = 1
            if b + 8 <= nblk - 1 and tiles >= oct_tiles:
                group = 8
            elif b + 4 <= nblk - 1 and tiles >= quad_tiles:
                group = 4
            elif b + 2 <= nblk - 1 and tiles >= pair_tiles:
                group = 2
        depth = 64 * (pend + 1)
        if pend + 1 < group:
            total += 8 * depth * 64 * 64 * (2 * tiles - 1)
            pend += 1
        else:
            total += 8 * depth * (64 * tiles) ** 2
            pend = 0
    return total, nblk - 1


def profile_value(path, key):
    """A 'key: value' line of a committed rocprofv3 summary under profiles/, or None."""
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic():
    """HBM bytes per launch of the regular assembly kernel on the default workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    return profile_value(TRAFFIC_PROFILE, "traffic_bytes_per_launch")


def regular_split(ctx, n):
    """Wavefronts per tile of the regular assembly of an n-element mesh under the context's
    launch-shape table (the first bucket whose bound exceeds the tile count; 1 beyond them)."""
    p = (n + 7) // 8
    tiles = p * (p + 1) // 2
    for upper, split in ctx.tune_assembly():
        if tiles < upper:
            return split
    return 1


def pair_points(n, nspecial, nq=16, L=3):
    """Kernel evaluations of one assembly: 256 per regular entry, 2*N^2 L^2 per coincident entry,
    3 N^2 L per adjacent entry."""
    n_ident = n
    n_adj = nspecial - n
    n_reg = n * n - nspecial
    return n_reg * nq * nq, n_ident * 2 * nq * nq * L * L + n_adj * 3 * nq * nq * L, n_reg


def run_batched(ctx, count=64, seed=1):
    """C3: count seeded problems submitted at once; every stage runs as batched launches."""
    import time
    from bem2d_amd import Engine, capi, serving
    out = {"workload": "%d problems (seed %d): ellipses / polygons, k log-uniform in [5, 200], 10 "
                       "elements per wavelength, 1 plane wave each, far field at 360 angles"
                       % (count, seed)}
    far = {}
    for prec, name in ((capi.PRECISION_MIXED, "mixed"), (capi.PRECISION_FP64, "plain_fp64")):
        reqs = serving.batched_problems(count, seed=seed, precision=prec)
        eng = Engine(ctx, 96 << 30, max_active=count)
        warm = [eng.submit(q) for q in reqs]  # one untimed pass of the same batch
        while eng.step() > 0:
            pass
        for t in warm:
            eng.result(t)
        t0 = time.perf_counter()
        tickets = [eng.submit(q) for q in reqs]
        while eng.step() > 0:
            pass
        dt = time.perf_counter() - t0
        res = [eng.result(t, raise_on_fail=False) for t in tickets]
        st = eng.stats()
        eng.close()
        far[name] = [r[1] for r in res]
        out[name] = {"solves_per_s": count / dt, "makespan_s": dt,
                     "failed": sum(r[2].status != 0 for r in res),
                     "lu_step_launches": st["lu_step_launches"],
                     "assembly_launches": st["assembly_launches"]}
        out["n_range"] = [min(q["geom"].n for q in reqs), max(q["geom"].n for q in reqs)]
    errs = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(far["mixed"], far["plain_fp64"])]
    out["farfield_rel_err_mixed_vs_fp64_max"] = float(max(errs))
    out["mixed_over_plain"] = out["mixed"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    return out


def run_kite_served(ctx, copies=8, n=4000, k=100.0, seed=3):
    """The default workload served: `copies` kite problems (N=4000, k=100, one plane wave each at
    seeded random angles) submitted together to the engine, so that the stages of different
    problems overlap on the device (assembly on the FP64 ALUs while another problem's LU runs on
    the matrix cores)."""
    import time
    from bem2d_amd import Engine, capi
    from bem2d_amd import geometry as G
    rng = np.random.default_rng(seed)
    geom = G.kite(n)
    out = {"workload": "%d kite problems N=%d k=%g, 1 plane wave each, submitted together"
                       % (copies, n, k)}
    for prec, name in ((capi.PRECISION_MIXED, "mixed"), (capi.PRECISION_FP64, "plain_fp64")):
        reqs = []
        for _ in range(copies):
            th = 2 * np.pi * rng.uniform()
            reqs.append({"geom": geom, "k": k, "eta": k, "dirs": [[np.cos(th), np.sin(th)]],
                         "theta": np.zeros(0), "precision": prec, "tol": 1e-12})
        eng = Engine(ctx, 32 << 30, max_active=copies)
        warm = eng.submit(reqs[0])
        while eng.step() > 0:
            pass
        eng.result(warm)
        t0 = time.perf_counter()
        tickets = [eng.submit(q) for q in reqs]
        while eng.step() > 0:
            pass
        dt = time.perf_counter() - t0
        res = [eng.result(t, raise_on_fail=False) for t in tickets]
        eng.close()
        out[name] = {"solves_per_s": copies / dt, "makespan_s": dt,
                     "failed": sum(r[2].status != 0 for r in res),
                     "max_backward_error": max(r[2].backward_error for r in res)}
    out["mixed_over_plain"] = out["mixed"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    return out


def run_serving(ctx, count, rate, seed=2):
    """C4: Poisson arrivals at `rate`; latency = completion - arrival; solves = (request, rhs).
    Then each path's capacity: the same requests all present at once."""
    import time
    from bem2d_amd import Engine, capi, serving
    out = {"workload": "%d requests (seed %d), Poisson arrivals at %.1f/s: ellipse / kite / polygon, "
                       "k uniform in [10, 400], 10 elements per wavelength, 1-32 plane waves, far "
                       "field at 36 angles" % (count, seed, rate), "rate_per_s": rate}
    for prec, name in ((capi.PRECISION_MIXED, "mixed"), (capi.PRECISION_FP64, "plain_fp64")):
        reqs, arr = serving.serving_trace(count, rate=rate, seed=seed, precision=prec)
        eng = Engine(ctx, 160 << 30, max_active=64)
        t0 = time.perf_counter()
        tickets = eng.run_trace(reqs, arr)
        dt = time.perf_counter() - t0
        res = [eng.result(t, raise_on_fail=False) for t in tickets]
        eng.close()
        infos = [r[2] for r in res]
        nrhs = sum(q["dirs"].shape[0] for q in reqs)
        lat = serving.latency_summary(infos)
        out[name] = {"p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"], "mean_ms": lat["mean_ms"],
                     "solves_per_s": nrhs / dt, "requests_per_s": count / dt, "wall_s": dt,
                     "failed": sum(i.status != 0 for i in infos),
                     "fallback_fp64": sum(i.fallback_fp64 for i in infos),
                     "max_backward_error": max(i.backward_error for i in infos)}
        progress("C4 trace, %s path: %d requests in %.1f s" % (name, count, dt))
    out["mixed_over_plain_solves"] = out["mixed"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    # capacity: the same requests all present at once, so each path runs as fast as it can
    for prec, name in ((capi.PRECISION_MIXED, "mixed"), (capi.PRECISION_FP64, "plain_fp64")):
        reqs, arr = serving.serving_trace(count, rate=rate, seed=seed, precision=prec)
        eng = Engine(ctx, 160 << 30, max_active=64)
        t0 = time.perf_counter()
        tickets = eng.run_trace(reqs, np.zeros_like(arr))
        dt = time.perf_counter() - t0
        res = [eng.result(t, raise_on_fail=False) for t in tickets]
        eng.close()
        nrhs = sum(q["dirs"].shape[0] for q in reqs)
        out[name]["capacity_solves_per_s"] = nrhs / dt
        out[name]["capacity_requests_per_s"] = count / dt
        out[name]["capacity_failed"] = sum(r[2].status != 0 for r in res)
        progress("C4 capacity, %s path: %.1f s" % (name, dt))
    out["mixed_over_plain_capacity"] = (out["mixed"]["capacity_solves_per_s"] /
                                        out["plain_fp64"]["capacity_solves_per_s"])
    return out


def cpu_baseline_large(geom, k, nrhs, ncols=16, n_lu=8000, seed=7):
    """The oracle (C restatement, OpenMP over columns) timed on a seeded sample of ncols columns of
    the N = 32000 matrix and extrapolated to all columns, plus LAPACK zgesv (scipy's OpenBLAS) of an
    n_lu x n_lu system with the same right-hand sides, extrapolated by (N / n_lu)^3."""
    from oracle import oracle as O
    import scipy.linalg.lapack as lapack
    n = geom.n
    rng = np.random.default_rng(seed)
    cols = np.sort(rng.choice(n, size=ncols, replace=False))
    O.assemble(geom, 16, 3, 0.15, O.K_COMBINED, k, k, 1.0, 2.0, cols=cols[:1])  # warm
    t0 = time.perf_counter()
    blk = O.assemble(geom, 16, 3, 0.15, O.K_COMBINED, k, k, 1.0, 2.0, cols=cols)
    t_cols = time.perf_counter() - t0
    M = (rng.standard_normal((n_lu, n_lu)) + 1j * rng.standard_normal((n_lu, n_lu))) / np.sqrt(n_lu)
    M += 4 * np.eye(n_lu)
    R = rng.standard_normal((n_lu, nrhs)) + 1j * rng.standard_normal((n_lu, nrhs))
    M = np.asfortranarray(M)
    t0 = time.perf_counter()
    _, _, _, info = lapack.zgesv(M, R, overwrite_a=True, overwrite_b=True)
    t_lu = time.perf_counter() - t0
    return {"cols": cols, "block": blk, "t_cols": t_cols, "t_asm": t_cols * n / ncols,
            "t_lu_small": t_lu, "t_lu": t_lu * (n / n_lu) ** 3, "n_lu": n_lu}


def run_large(ctx, steps, warmup, n=32000, k=500.0, nrhs=360, plain_steps=3, cpu=True):
    """C5, the headline: open arc (aperture 0.3 rad), N = 32000, k = 500, 360 plane waves.  One
    step = FP64 assembly + right-hand sides + FP32 MFMA LU + FP64 refinement to 1e-12; the plain
    FP64 path (assembly + rocSOLVER zgetrf/zgetrs) timed over plain_steps steps beside it."""
    import torch
    from bem2d_amd import capi
    from bem2d_amd import geometry as G
    geom = G.open_arc(n, 0.3)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    th = 2 * np.pi * np.arange(nrhs) / nrhs
    dirs = np.stack([np.cos(th), np.sin(th)], 1)
    A = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()
    B = torch.empty((nrhs, n), dtype=torch.complex128, device="cuda").t()
    X = torch.empty_like(B)
    lu = ctx.lu32(n)

    def step_mixed():
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
        lu.factor(A)
        return lu.refine(A, B, X, tol=1e-12, raise_on_fail=False)

    for _ in range(warmup):
        step_mixed()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    reports = [step_mixed() for _ in range(steps)]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    reg_ms, reg_n = ctx.profile_read(capi.PROF_ASSEMBLE_REGULAR)
    sp_ms, sp_n = ctx.profile_read(capi.PROF_ASSEMBLE_SPECIAL)
    lu_ms, lu_n = ctx.profile_read(capi.PROF_LU)
    up_ms, up_n = ctx.profile_read(capi.PROF_LU_UPDATE)
    ctx.profile(False)
    progress("C5 mixed path: %d steps in %.2f s" % (steps, dt))

    # the plain FP64 path: assembly, right-hand sides, zgetrf + zgetrs (A overwritten)
    X64 = torch.empty_like(B)

    def step_plain():
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=X64)
        ctx.solve_fp64(A, X64, preserve_A=False)

    step_plain()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        step_plain()
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t0
    lu64_ms, lu64_n = ctx.profile_read(capi.PROF_LU)
    ctx.profile(False)
    progress("C5 plain FP64 path: %d steps in %.2f s" % (plain_steps, dt64))

    # checks outside the timed regions: both solutions of one assembly, far fields at 360 angles
    ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
    ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
    lu.factor(A)
    rep = lu.refine(A, B, X, tol=1e-12, raise_on_fail=False)
    X64.copy_(B)
    ctx.solve_fp64(A, X64, preserve_A=True)
    ang = torch.tensor(2 * np.pi * np.arange(360) / 360, device="cuda")
    fm = ctx.farfield(mesh, k, X, ang)
    f64 = ctx.farfield(mesh, k, X64, ang)
    e_ff = (torch.linalg.norm(fm - f64) / torch.linalg.norm(f64)).item()
    e_x = (torch.linalg.norm(X - X64) / torch.linalg.norm(X64)).item()
    P_reg, P_sing, n_reg = pair_points(n, mesh.special_count)
    cpu_res = None
    if cpu:
        cpu_res = cpu_baseline_large(geom, k, nrhs)
        got = A[:, torch.tensor(cpu_res["cols"], device="cuda")].cpu().numpy()
        cpu_res["gpu_vs_oracle_cols"] = float(np.linalg.norm(got - cpu_res["block"]) /
                                              np.linalg.norm(cpu_res["block"]))
        progress("C5 CPU baseline sample done")
    lu.close()
    mesh.close()
    del A, B, X, X64
    ctx.release_scratch()
    torch.cuda.empty_cache()

    pair_tiles = ctx.tune_lu_pair_tiles()
    quad_tiles = ctx.tune_lu_quad_tiles()
    oct_tiles = ctx.tune_lu_oct_tiles()
    upd_flop, upd_launches = lu_update_flops(n, pair_tiles, quad_tiles, oct_tiles)
    flu = 8.0 / 3.0 * n ** 3
    reg_avg = reg_ms / max(reg_n, 1)
    up_avg = up_ms / max(up_n, 1)
    lu_avg = lu_ms / max(lu_n, 1)
    achieved_upd = upd_flop * steps / (up_ms * 1e-3) / 1e12 if up_ms > 0 else None
    asm_achieved = FLOP_PER_PAIRPOINT * P_reg / (reg_avg * 1e-3) / 1e12
    out = {
        "workload": "open arc (radius 1, aperture 0.3 rad) N=%d k=%g combined-field (eta=k), P0, "
                    "16-pt Gauss, L=3 sigma=0.15, %d plane waves at equispaced angles; one step = "
                    "FP64 assembly A=M+2K + right-hand sides + FP32 MFMA LU + FP64 refinement to "
                    "backward error 1e-12" % (n, k, nrhs),
        "n": n, "k": k, "nrhs": nrhs,
        "steps": steps, "warmup": warmup, "time_s": dt,
        "solves_per_s_mixed": nrhs * steps / dt,
        "ms_per_step_mixed": 1e3 * dt / steps,
        "solves_per_s_plain_fp64": nrhs * plain_steps / dt64,
        "ms_per_step_plain_fp64": 1e3 * dt64 / plain_steps,
        "plain_steps": plain_steps,
        "mixed_over_plain": (nrhs * steps / dt) / (nrhs * plain_steps / dt64),
        "converged_all_steps": all(r.converged for r in reports),
        "fallback_any_step": any(r.fallback_fp64 for r in reports),
        "backward_error_max": max(r.backward_error for r in reports),
        "refinement_steps_max": max(r.iterations for r in reports),
        "growth": rep.growth,
        "x_rel_err_mixed_vs_fp64": e_x,
        "farfield_rel_err_mixed_vs_fp64": e_ff,
        "assembly_regular_ms": reg_avg,
        "assembly_special_ms": sp_ms / max(sp_n, 1),
        "gentries_per_s": n * n / (reg_avg * 1e-3) / 1e9,
        "lu32_ms": lu_avg,
        "lu32_tflops": flu / (lu_avg * 1e-3) / 1e12,
        "lu32_frac_of_fp32_mfma_peak": flu / (lu_avg * 1e-3) / 1e12 / PEAK_FP32_MFMA_TFLOPS,
        "lu_fp64_ms": lu64_ms / max(lu64_n, 1),
        "lu_fp64_tflops": flu / (lu64_ms / max(lu64_n, 1) * 1e-3) / 1e12,
        "roofline": {
            "kernel": "trailing_update_kernel (FP32 LU, complex rank-64..512 update on "
                      "v_mfma_f32_16x16x4f32, three real products per complex product)",
            "bound": "mfma",
            "achieved": achieved_upd,
            "peak": PEAK_FP32_MFMA_TFLOPS,
            "unit": "TFLOP/s",
            "frac": None if achieved_upd is None else achieved_upd / PEAK_FP32_MFMA_TFLOPS,
            "traffic": profile_value(LU_TRAFFIC_PROFILE, "traffic_bytes_per_launch"),
            "work_per_launch": "%.6e FLOP per factorisation over %d launches = %.6e per launch "
                               "(8 real FLOP per complex multiply-add; panels in twos from %d "
                               "tiles, in fours from %d, in eights from %d)" % (
                                   upd_flop, upd_launches, upd_flop / upd_launches, pair_tiles,
                                   quad_tiles, oct_tiles),
            "executed_frac": None if achieved_upd is None else
            0.75 * achieved_upd / PEAK_FP32_MFMA_TFLOPS,
            "executed_note": "the kernel forms each complex product from three real MFMA products "
                             "(6 real FLOP per complex multiply-add executed): executed_frac is "
                             "the MFMA FLOP the hardware performs over the peak",
            "launches": up_n, "avg_launch_ms": up_avg,
            "measured": "HIP events around each trailing_update_kernel launch on the LU stream "
                        "over the timed steps (eager launches)",
        },
        "roofline_assembly": {
            "kernel": "assemble_regular_kernel<COMBINED,16,%d>" % regular_split(ctx, n),
            "bound": "fp64-valu",
            "achieved": asm_achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": asm_achieved / PEAK_FP64_TFLOPS,
            "convention": "128 FLOP per pair-point of the reference algorithm (256 per regular "
                          "entry); the kernel evaluates each Hankel pair once for the entries "
                          "(i,j) and (j,i), so it executes about 0.6 of the convention's FLOP",
            "work_per_launch": "%d FLOP = 128 x 256 pair-points x %d regular entries" %
                               (FLOP_PER_PAIRPOINT * P_reg, n_reg),
            "hardware_fp64_flop_per_launch": profile_value(ASM_C5_PROFILE,
                                                           "hardware_fp64_flop_per_launch"),
            "valu_issue_fraction": profile_value(ASM_C5_PROFILE, "valu_issue_fraction"),
            "avg_launch_ms": reg_avg,
        },
    }
    hw = out["roofline_assembly"]["hardware_fp64_flop_per_launch"]
    if hw:
        out["roofline_assembly"]["hardware_fp64_frac"] = hw / (reg_avg * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
    if cpu_res is not None:
        t_cpu = cpu_res["t_asm"] + cpu_res["t_lu"]
        cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = {
            "value": nrhs / t_cpu,
            "unit": "solves/s",
            "cores": cores,
            "kind": "port",
            "sample": "oracle assembly of %d seeded random columns of the N=%d matrix (%.1f s, "
                      "extrapolated x%d to %.0f s; the GPU matrix agrees with it on those columns "
                      "to %.1e) + scipy LAPACK zgesv of an N=%d system with the %d right-hand "
                      "sides (%.2f s, extrapolated x%d to %.0f s)" %
                      (len(cpu_res["cols"]), n, cpu_res["t_cols"], n // len(cpu_res["cols"]),
                       cpu_res["t_asm"], cpu_res["gpu_vs_oracle_cols"], cpu_res["n_lu"], nrhs,
                       cpu_res["t_lu_small"], (n // cpu_res["n_lu"]) ** 3, cpu_res["t_lu"]),
        }
    return out


def run_kite(ctx, steps, warmup, n=4000, k=100.0, cpu_cols=0):
    """C2: kite N=4000, k=100, one plane wave at a seeded random angle, both paths; the solutions'
    distance from each other (their distance from the oracle is a GPU test of its own)."""
    import torch
    from bem2d_amd import capi
    from bem2d_amd import geometry as G
    rng = np.random.default_rng(0)
    theta0 = 2 * math.pi * rng.uniform()
    dirs = [[math.cos(theta0), math.sin(theta0)]]
    geom = G.kite(n)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    P_reg, P_sing, n_reg = pair_points(n, mesh.special_count)
    A = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()
    B = torch.empty((1, n), dtype=torch.complex128, device="cuda").t()

    def step():
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
        ctx.solve_fp64(A, B, preserve_A=False)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    lu_ms, lu_n = ctx.profile_read(capi.PROF_LU)
    ctx.profile(False)

    lu = ctx.lu32(n)
    X = torch.empty_like(B)

    def step_mixed():
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
        lu.factor(A)
        return lu.refine(A, B, X, tol=1e-12, raise_on_fail=False)

    for _ in range(warmup):
        step_mixed()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    reports = [step_mixed() for _ in range(steps)]
    torch.cuda.synchronize()
    dt_mixed = time.perf_counter() - t0
    lu32_ms, lu32_n = ctx.profile_read(capi.PROF_LU)
    reg_ms, reg_n = ctx.profile_read(capi.PROF_ASSEMBLE_REGULAR)
    sp_ms, sp_n = ctx.profile_read(capi.PROF_ASSEMBLE_SPECIAL)
    ctx.profile(False)

    ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
    ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
    X64 = B.clone()
    ctx.solve_fp64(A, X64, preserve_A=True)
    lu.factor(A)
    lu.refine(A, B, X, tol=1e-12, raise_on_fail=False)
    torch.cuda.synchronize()
    bw = (torch.linalg.norm(B - A @ X64) / (torch.linalg.norm(A) * torch.linalg.norm(X64) +
                                            torch.linalg.norm(B))).item()
    mixed_vs_fp64 = (torch.linalg.norm(X - X64) / torch.linalg.norm(X64)).item()
    lu.close()
    mesh.close()
    del A, B, X, X64
    ctx.release_scratch()
    torch.cuda.empty_cache()
    reg_avg = reg_ms / max(reg_n, 1)
    achieved = FLOP_PER_PAIRPOINT * P_reg / (reg_avg * 1e-3) / 1e12
    hw_flop = profile_value(FP64_PROFILE, "hardware_fp64_flop_per_launch")
    return {
        "workload": "kite N=%d k=%g combined-field (eta=k), P0, 16-pt Gauss, L=3 sigma=0.15, 1 "
                    "plane-wave right-hand side at a seeded random angle; one step = FP64 "
                    "assembly + right-hand side + solve" % (n, k),
        "steps": steps,
        "solves_per_s_mixed": steps / dt_mixed,
        "solves_per_s_plain_fp64": steps / dt,
        "mixed_over_plain": dt / dt_mixed,
        "ms_per_step_mixed": 1e3 * dt_mixed / steps,
        "ms_per_step_plain_fp64": 1e3 * dt / steps,
        "mixed_converged": all(r.converged for r in reports),
        "mixed_backward_error": max(r.backward_error for r in reports),
        "mixed_refinement_steps": max(r.iterations for r in reports),
        "mixed_vs_fp64_solution_rel_err": mixed_vs_fp64,
        "backward_error_fp64": bw,
        "lu32_ms": lu32_ms / max(lu32_n, 1),
        "lu_fp64_ms": lu_ms / max(lu_n, 1),
        "assembly_regular_ms": reg_avg,
        "assembly_special_ms": sp_ms / max(sp_n, 1),
        "gentries_per_s": n * n / (reg_avg * 1e-3) / 1e9,
        "gpairpts_per_s": (P_reg + P_sing) / (reg_avg * 1e-3) / 1e9,
        "pair_points": P_reg + P_sing,
        "roofline_assembly": {
            "kernel": "assemble_regular_kernel<COMBINED,16,%d>" % regular_split(ctx, n),
            "bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": measured_traffic(),
            "hardware_fp64_frac": None if hw_flop is None else
            hw_flop / (reg_avg * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
            "valu_issue_fraction": profile_value(FP64_PROFILE, "valu_issue_fraction"),
            "avg_launch_ms": reg_avg,
        },
    }


def run_disk(ctx, steps, warmup, k=1.0, ppw=2000, L=7, N=30, cpu=True):
    """The reference's examples/disk.cpp at its largest setting (k = 1, ppw = 2000): the unit circle
    as AnalyticCurve(n, Circle) with n = int(Length k ppw / 2 pi) pieces of equal arc length,
    LinBasis (two per element), SoundSoftScattering with SetQuadOption(7, 30, 0.15); one step =
    assembly of A = M + 2 K_combined (eta = k) with the bases + right-hand side + FP64 solve +
    EvalSol at (-1.5, 0).  Accuracy: EvalSol against the analytic series of the sound-soft disk."""
    import torch
    import scipy.special as sp
    from bem2d_amd import capi
    from bem2d_amd import geometry as G
    n = int(capi.curve_length(capi.ELEM_CIRCLE, [0, 0, 0, 0]) * k * ppw / 2.0 / math.pi)
    geom = G.analytic_curve("circle", n)
    bases = capi.lin_bases()
    nn = n * len(bases)
    mesh = ctx.mesh(geom, N, L, 0.15)
    A = torch.empty((nn, nn), dtype=torch.complex128, device="cuda").t()
    B = torch.empty((1, nn), dtype=torch.complex128, device="cuda").t()
    pt = torch.tensor([[-1.5, 0.0]], dtype=torch.float64, device="cuda")
    dirs = [[1.0, 0.0]]

    def step():
        ctx.assemble_basis(mesh, bases, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        ctx.rhs_basis(mesh, bases, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
        ctx.solve_fp64(A, B, preserve_A=False)
        return ctx.eval_single_layer_basis(mesh, bases, k, B[:, 0].contiguous(), pt)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        s_u = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    reg_ms, reg_n = ctx.profile_read(capi.PROF_ASSEMBLE_REGULAR)
    sp_ms, sp_n = ctx.profile_read(capi.PROF_ASSEMBLE_SPECIAL)
    lu_ms, lu_n = ctx.profile_read(capi.PROF_LU)
    ctx.profile(False)
    u = complex(np.exp(1j * k * -1.5) - s_u[0].item())
    # sound-soft disk: u = u_inc - sum_m i^m J_m(k) / H_m(k) H_m(k r) e^{i m theta}, r = 1.5,
    # theta = pi
    ms = np.arange(-60, 61)
    u_ser = complex(np.exp(1j * k * -1.5) - np.sum(
        (1j ** ms) * sp.jv(ms, k) / sp.hankel1(ms, k) * sp.hankel1(ms, 1.5 * k) *
        np.exp(1j * ms * math.pi)))
    nspecial = mesh.special_count
    P_reg = (n * n - nspecial) * N * N
    reg_avg = reg_ms / max(reg_n, 1)
    res = {
        "workload": "examples/disk.cpp at k=%g ppw=%d: unit circle AnalyticCurve n=%d (arc "
                    "length), LinBasis (%d unknowns), combined field eta=k, N=%d, L=%d, sigma=0.15; "
                    "step = basis assembly + right-hand side + FP64 solve + EvalSol at (-1.5, 0)"
                    % (k, ppw, n, nn, N, L),
        "n_elements": n, "unknowns": nn, "steps": steps,
        "examples_per_s": steps / dt,
        "ms_per_step": 1e3 * dt / steps,
        "assembly_regular_ms": reg_avg,
        "assembly_special_ms": sp_ms / max(sp_n, 1),
        "lu_fp64_ms": lu_ms / max(lu_n, 1),
        "gentries_per_s": nn * nn / (reg_avg * 1e-3) / 1e9,
        "gpairpts_per_s": P_reg / (reg_avg * 1e-3) / 1e9,
        "evalsol": [u.real, u.imag],
        "evalsol_vs_series_rel_err": abs(u - u_ser) / abs(u_ser),
    }
    if cpu:
        # the oracle (OpenMP over columns) on a bounded column sample, scaled to the whole matrix
        from oracle import oracle as O
        cols = np.linspace(0, nn - 1, 8).astype(np.int64)
        t0 = time.perf_counter()
        O.assemble_basis(geom, bases, N, L, 0.15, O.K_COMBINED, k, k, 1.0, 2.0, cols=cols)
        tc = time.perf_counter() - t0
        res["cpu_assembly_estimate_s"] = tc * nn / cols.size
        res["cpu_sample"] = "oracle assemble_basis of 8 of the %d columns in %.2f s on %d threads, " \
            "scaled to the whole matrix" % (nn, tc, os.cpu_count() or 1)
    mesh.close()
    del A, B
    ctx.release_scratch()
    torch.cuda.empty_cache()
    return res


TRAPPING = [(0.0, 0.0), (-1.0, 0.0), (-1.0, -0.69), (0.69, -0.69), (0.69, 1.31), (-1.0, 1.31),
            (-1.0, 0.62), (0.0, 0.62)]  # resonances/trapresonance.cpp:21-35, a = 0.31, c = 1


def resonance_grid():
    """The reference's resonance grid: kr in [4, 8), ki in [-2, 0), 10 x 10 points."""
    return [complex(4 + i * 0.4, -2 + j * 0.2) for i in range(10) for j in range(10)]


def run_resonance(ctx, npts=4, nasm=10, cpu=True):
    """resonances/trapresonance.cpp:44-80: the trapping polygon as Polygon(points, ppw=10,
    kmax=8, L=10, sigma=0.15) (1360 elements), PolBasis(2) (4080 unknowns), N=5, L=3; per complex
    wavenumber k of the 10 x 10 grid: Id + 2 CDL - 2 i |k| SL assembled in one pass (complex-k
    kernels with bases), ChangeBasis with the mass matrix, SingularValues, the smallest singular
    value.  Times: the whole pipeline on npts grid points, the assembly alone on nasm."""
    import torch
    from bem2d_amd import capi
    from bem2d_amd import geometry as G
    geom = G.polygon_ppw(TRAPPING, 10, 8.0, 10, 0.15)
    bases = capi.pol_bases(2)
    nn = geom.n * len(bases)
    mesh = ctx.mesh(geom, 5, 3, 0.15)
    grid = resonance_grid()
    M = ctx.assemble_basis(mesh, bases, capi.KERNEL_NONE, 1.0, 0.0, 1.0, 0.0)
    A = torch.empty((nn, nn), dtype=torch.complex128, device="cuda").t()
    ctx.profile(True)
    ctx.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in grid[:nasm]:
        ctx.assemble_basis(mesh, bases, capi.KERNEL_COMBINED, k, abs(k), 1.0, 2.0, out=A)
    torch.cuda.synchronize()
    t_asm = (time.perf_counter() - t0) / nasm
    progress("resonance: %d assemblies of %d unknowns, %.1f ms each" % (nasm, nn, 1e3 * t_asm))
    reg_ms, reg_n = ctx.profile_read(capi.PROF_ASSEMBLE_REGULAR)
    sp_ms, sp_n = ctx.profile_read(capi.PROF_ASSEMBLE_SPECIAL)
    ctx.profile(False)
    smin = []
    t0 = time.perf_counter()
    t_cb = t_svd = 0.0
    for k in grid[:npts]:
        ctx.assemble_basis(mesh, bases, capi.KERNEL_COMBINED, k, abs(k), 1.0, 2.0, out=A)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        ctx.change_basis(A, M)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        progress("resonance: k = %s change of basis %.1f ms; singular values..." %
                 (k, 1e3 * (tb - ta)))
        s = ctx.singular_values(A)
        smin.append(float(s[-1].item()))
        tc = time.perf_counter()
        progress("resonance: singular values %.1f s" % (tc - tb))
        t_cb += tb - ta
        t_svd += tc - tb
    t_all = (time.perf_counter() - t0) / npts
    res = {
        "workload": "resonances/trapresonance.cpp: trapping polygon, Polygon(ppw=10, kmax=8, L=10, "
                    "sigma=0.15) = %d elements, PolBasis(2) = %d unknowns, N=5, L=3; complex k on "
                    "the 10 x 10 grid kr in [4, 8), ki in [-2, 0); per k: Id + 2 CDL - 2i|k| SL, "
                    "ChangeBasis, SingularValues" % (geom.n, nn),
        "unknowns": nn,
        "assembly_ms": 1e3 * t_asm,
        "assembly_regular_ms": reg_ms / max(reg_n, 1),
        "assembly_special_ms": sp_ms / max(sp_n, 1),
        "assembly_matrices_per_s": 1.0 / t_asm,
        "change_basis_ms": 1e3 * t_cb / npts,
        "svd_ms": 1e3 * t_svd / npts,
        "pipeline_points_per_s": 1.0 / t_all,
        "grid_points_timed": npts,
        "smallest_singular_values": smin,
        "k": [[k.real, k.imag] for k in grid[:npts]],
    }
    if cpu:
        from oracle import oracle as O
        cols = np.linspace(0, nn - 1, 4).astype(np.int64)
        t0 = time.perf_counter()
        O.assemble_basis(geom, bases, 5, 3, 0.15, O.K_COMBINED, grid[0], abs(grid[0]), 1.0, 2.0,
                         cols=cols)
        tc = time.perf_counter() - t0
        res["cpu_assembly_estimate_s"] = tc * nn / cols.size
        res["cpu_sample"] = "oracle assemble_basis (complex k) of 4 of the %d columns in %.2f s on " \
            "%d threads, scaled to the whole matrix" % (nn, tc, os.cpu_count() or 1)
    mesh.close()
    del A, M
    ctx.release_scratch()
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="timed steps of the plain FP64 path on the large configuration")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--config", choices=["all", "large", "kite", "batched", "serving", "disk",
                                         "resonance"],
                    default="all")
    ap.add_argument("--no-extras", action="store_true",
                    help="the large configuration only (no kite, batched and served legs)")
    ap.add_argument("--serving-count", type=int, default=2000)
    ap.add_argument("--serving-rate", type=float, default=27.0,
                    help="arrivals per second (about 0.8 x the plain path's measured capacity)")
    ap.add_argument("--probe-traffic", action="store_true",
                    help="one kite assembly and one 1 GiB device copy, for rocprofv3 --pmc runs")
    ap.add_argument("--probe-lu", action="store_true",
                    help="one N=32000 open-arc assembly and FP32 factorisation and one 1 GiB "
                         "device copy, for rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has a single MI355X, and the path "
              "would shard across GPUs by problem (no collective); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    # per-kernel HIP events need the factorisation's launches made eagerly (a replayed hipGraph of
    # the launch sequence records none; it measured no faster, the launches being queued ahead)
    os.environ.setdefault("BEM2D_LU_GRAPH", "0")

    import torch
    from bem2d_amd import Context

    if args.probe_traffic:
        probe_traffic(args)
        return
    if args.probe_lu:
        probe_lu()
        return
    ctx = Context(0)
    cpu = not args.no_cpu_baseline
    if args.config in ("kite", "batched", "serving", "disk", "resonance"):
        if args.config == "kite":
            res = run_kite(ctx, args.steps, args.warmup)
            value, t = res["solves_per_s_mixed"], res["ms_per_step_mixed"]
        elif args.config == "disk":
            res = run_disk(ctx, min(args.steps, 5), min(args.warmup, 1), cpu=cpu)
            value, t = res["examples_per_s"], res["ms_per_step"]
        elif args.config == "resonance":
            res = run_resonance(ctx, npts=1, cpu=cpu)
            value, t = res["pipeline_points_per_s"], None
        elif args.config == "batched":
            res = run_batched(ctx)
            value, t = res["mixed"]["solves_per_s"], None
        else:
            res = run_serving(ctx, args.serving_count, args.serving_rate)
            value, t = res["mixed"]["solves_per_s"], None
        print(json.dumps({"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": t,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "f64-assembly+fp32-mfma-lu+f64-refinement",
                          "data": "synthetic",
                          "config": {"workload": res["workload"], "name": args.config},
                          "result": res}), flush=True)
        ctx.close()
        return

    large = run_large(ctx, args.steps, args.warmup, plain_steps=args.plain_steps, cpu=cpu)
    out = {
        "metric": METRIC,
        "value": large["solves_per_s_mixed"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": large["ms_per_step_mixed"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64-assembly+fp32-mfma-lu+f64-refinement",
        "data": "synthetic",
        "config": {"workload": large["workload"], "name": "large", "n": large["n"],
                   "k": large["k"], "nrhs": large["nrhs"], "path": "mixed"},
        "roofline": large.pop("roofline"),
    }
    if "cpu_baseline" in large:
        out["cpu_baseline"] = large.pop("cpu_baseline")
    out["large"] = large
    if not args.no_extras and args.config == "all":
        out["kite"] = run_kite(ctx, args.steps, args.warmup)
        progress("C2 kite done")
        out["kite_served"] = run_kite_served(ctx)
        progress("kite served done")
        out["batched"] = run_batched(ctx)
        progress("C3 batch done")
        out["disk_lin"] = run_disk(ctx, 3, 1, cpu=cpu)
        progress("disk.cpp (LinBasis) leg done")
        out["resonance"] = run_resonance(ctx, npts=1, nasm=4, cpu=cpu)
        progress("resonance (complex k, PolBasis(2)) leg done")
        sv = run_serving(ctx, args.serving_count, args.serving_rate)
        out["serving"] = sv
        out["p50_ms"] = sv["mixed"]["p50_ms"]
        out["p99_ms"] = sv["mixed"]["p99_ms"]
    print(json.dumps(out), flush=True)
    ctx.close()


def probe_traffic(args):
    """One regular-assembly launch of the kite workload (N=4000, k=100) and one 1 GiB
    device-to-device copy (the calibration kernel of known size) -- the program rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE runs profile."""
    import torch
    from bem2d_amd import Context, capi
    from bem2d_amd import geometry as G
    ctx = Context(0)
    mesh = ctx.mesh(G.kite(4000), 16, 3, 0.15)
    A = torch.empty((4000, 4000), dtype=torch.complex128, device="cuda").t()
    ctx.assemble(mesh, capi.KERNEL_COMBINED, 100.0, 100.0, 1.0, 2.0, out=A)
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print("probe done: assembly N=4000 + copy of %d bytes" % (src.numel() * 8), flush=True)
    mesh.close()
    ctx.close()


def probe_lu():
    """One assembly and one FP32 factorisation of the N=32000 open arc, then one 1 GiB
    device-to-device copy -- the program the rocprofv3 --pmc runs of the large configuration
    profile."""
    import torch
    from bem2d_amd import Context, capi
    from bem2d_amd import geometry as G
    n, k = 32000, 500.0
    ctx = Context(0)
    mesh = ctx.mesh(G.open_arc(n, 0.3), 16, 3, 0.15)
    A = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()
    ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
    lu = ctx.lu32(n)
    lu.factor(A)
    torch.cuda.synchronize()
    print("probe: assembly and FP32 factorisation N=%d done" % n, flush=True)
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print("probe done: copy of %d bytes" % (src.numel() * 8), flush=True)
    lu.close()
    mesh.close()
    ctx.close()


if __name__ == "__main__":
    main()
