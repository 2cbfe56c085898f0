"""Benchmark of the bem2d hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (one step): sound-soft scattering by the kite x = cos t + 0.65 cos 2t - 0.65, y = 1.5 sin t
at k = 100 with the combined-field operator (eta = k), N = 4000 piecewise-constant elements at
uniform parameter spacing, 16-point Gauss with the singular rules at L = 3, sigma = 0.15, and one
plane-wave right-hand side at a seeded random angle: Galerkin assembly A = M + 2K, the right-hand
side b = 2<f, phi>, and the dense solve.  Inputs (geometry tables) are resident in HBM before the
timed region starts.  Prints one JSON line.
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
PEAK_HBM_GBS = 8000.0


def pair_points(n, nspecial, nq=16, L=3):
    """Kernel evaluations of one assembly: 256 per regular entry, 2*N^2 L^2 per coincident entry,
    3 N^2 L per adjacent entry."""
    n_ident = n
    n_adj = nspecial - n
    n_reg = n * n - nspecial
    return n_reg * nq * nq, n_ident * 2 * nq * nq * L * L + n_adj * 3 * nq * nq * L, n_reg


def cpu_baseline(geom, A_host, k, b_host, ncols, seed=7):
    """The oracle (C restatement, OpenMP over columns) timed on a seeded sample of columns and
    extrapolated to the full matrix, plus LAPACK zgesv (scipy's OpenBLAS) on the full matrix."""
    from oracle import oracle as O
    import scipy.linalg.lapack as lapack
    n = geom.n
    rng = np.random.default_rng(seed)
    cols = np.sort(rng.choice(n, size=ncols, replace=False))
    O.assemble(geom, 16, 3, 0.15, O.K_COMBINED, k, k, 1.0, 2.0, cols=cols[:2])  # warm
    t0 = time.perf_counter()
    blk = O.assemble(geom, 16, 3, 0.15, O.K_COMBINED, k, k, 1.0, 2.0, cols=cols)
    t_cols = time.perf_counter() - t0
    t_asm = t_cols * n / ncols
    err = np.linalg.norm(A_host[:, cols] - blk) / np.linalg.norm(blk)
    t0 = time.perf_counter()
    _, _, x, info = lapack.zgesv(np.asfortranarray(A_host), b_host.copy())
    t_solve = time.perf_counter() - t0
    return t_asm, t_solve, t_cols, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=4000)
    ap.add_argument("--k", type=float, default=100.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cols", type=int, default=64)
    ap.add_argument("--path", choices=["both", "plain", "mixed"], default="both",
                    help="profiling aid: time only one solve path (the JSON then reports it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has a single MI355X, and the path "
              "would shard across GPUs by problem (no collective); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from bem2d_amd import Context, capi
    from bem2d_amd import geometry as G

    n, k = args.n, args.k
    rng = np.random.default_rng(0)
    theta0 = 2 * math.pi * rng.uniform()
    dirs = [[math.cos(theta0), math.sin(theta0)]]
    geom = G.kite(n)
    ctx = Context(0)
    mesh = ctx.mesh(geom, 16, 3, 0.15)
    P_reg, P_sing, n_reg = pair_points(n, mesh.special_count)
    P = P_reg + P_sing

    A = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()
    B = torch.empty((1, n), dtype=torch.complex128, device="cuda").t()

    def step():
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
        ctx.solve_fp64(A, B, preserve_A=False)

    dt = None
    lu_avg = None
    if args.path in ("both", "plain"):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        lu_ms, lu_n = ctx.profile_read(capi.PROF_LU)
        lu_avg = lu_ms / max(lu_n, 1)
        ctx.profile(False)
    # assembly kernels timed on their own (HIP events on the launch stream)
    ctx.profile(True)
    for _ in range(args.steps):
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
    reg_ms, reg_n = ctx.profile_read(capi.PROF_ASSEMBLE_REGULAR)
    sp_ms, sp_n = ctx.profile_read(capi.PROF_ASSEMBLE_SPECIAL)
    ctx.profile(False)
    reg_avg = reg_ms / max(reg_n, 1)
    sp_avg = sp_ms / max(sp_n, 1)
    asm_ms = reg_avg + sp_avg

    # ---- the chosen form: same assembly + FP32 MFMA LU with FP64 refinement -------------------
    lu = ctx.lu32(n)
    X = torch.empty_like(B)

    def step_mixed():
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
        lu.factor(A)
        return lu.refine(A, B, X, tol=1e-12, raise_on_fail=False)

    dt_mixed = None
    mixed_ok, mixed_bw, mixed_it, lu32_ms, lu32_n = False, None, None, 0.0, 0
    if args.path in ("both", "mixed"):
        for _ in range(args.warmup):
            step_mixed()
        torch.cuda.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        reports = [step_mixed() for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt_mixed = time.perf_counter() - t0
        lu32_ms, lu32_n = ctx.profile_read(capi.PROF_LU)
        ctx.profile(False)
        mixed_ok = all(r.converged for r in reports)
        mixed_bw = max(r.backward_error for r in reports)
        mixed_it = max(r.iterations for r in reports)

    # one more assembly for the checks and the CPU baseline (outside the timed region)
    ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
    ctx.rhs(mesh, capi.FIELD_COMBINED_PLANE, k, k, dirs, 2.0, out=B)
    A_host = A.cpu().numpy()
    b_host = B.cpu().numpy()
    X = B.clone()
    ctx.solve_fp64(A.clone(), X, preserve_A=False)
    torch.cuda.synchronize()
    bw = (torch.linalg.norm(B - A @ X) / (torch.linalg.norm(A) * torch.linalg.norm(X) +
                                          torch.linalg.norm(B))).item()
    Xm = torch.empty_like(B)
    lu.factor(A)
    lu.refine(A, B, Xm, tol=1e-12, raise_on_fail=False)
    mixed_vs_fp64 = (torch.linalg.norm(Xm - X) / torch.linalg.norm(X)).item()
    lu.close()

    achieved = FLOP_PER_PAIRPOINT * P_reg / (reg_avg * 1e-3) / 1e12
    plain_sps = args.steps / dt if dt else None
    mixed_sps = args.steps / dt_mixed if dt_mixed else None
    use_mixed = mixed_ok and (plain_sps is None or mixed_sps >= plain_sps)
    out = {
        "metric": METRIC,
        "value": mixed_sps if use_mixed else plain_sps,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * (dt_mixed if use_mixed else dt) / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "kite N=%d k=%g combined-field (eta=k), P0, 16-pt Gauss, L=3 sigma=0.15, "
                        "1 plane-wave RHS at a seeded random angle; one step = FP64 assembly "
                        "A=M+2K + RHS + solve; value = the faster of the mixed path (FP32 MFMA "
                        "LU + FP64 refinement to backward error 1e-12) and the plain FP64 path "
                        "(rocSOLVER zgetrf/zgetrs)" % (n, k),
            "n": n, "k": k, "nrhs": 1, "path": "mixed" if use_mixed else "plain-fp64",
        },
        "solves_per_s_mixed": mixed_sps,
        "solves_per_s_plain_fp64": plain_sps,
        "mixed_over_plain": mixed_sps / plain_sps if (mixed_sps and plain_sps) else None,
        "mixed_converged": mixed_ok,
        "mixed_backward_error": mixed_bw,
        "mixed_refinement_steps": mixed_it,
        "mixed_vs_fp64_solution_rel_err": mixed_vs_fp64,
        "lu32_ms": lu32_ms / max(lu32_n, 1),
        "gentries_per_s": n * n / (asm_ms * 1e-3) / 1e9,
        "gpairpts_per_s": P / (asm_ms * 1e-3) / 1e9,
        "assembly_ms": asm_ms,
        "assembly_regular_ms": reg_avg,
        "assembly_special_ms": sp_avg,
        "lu_fp64_ms": lu_avg,
        "pair_points": P,
        "backward_error_fp64": bw,
        "roofline": {
            "kernel": "assemble_regular_kernel<COMBINED,16>",
            "bound": "fp64-valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": None,
            "work_per_launch": "%d FLOP = 128 FLOP x 256 pair-points x %d regular entries" %
                               (FLOP_PER_PAIRPOINT * P_reg, n_reg),
        },
    }
    if not args.no_cpu_baseline:
        t_asm, t_solve, t_cols, err = cpu_baseline(geom, A_host, k, b_host, args.cpu_cols)
        cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        out["cpu_baseline"] = {
            "value": 1.0 / (t_asm + t_solve),
            "unit": "solves/s",
            "cores": cores,
            "kind": "port",
            "sample": "oracle assembly of %d seeded random columns of the N=%d matrix (%.1f s, "
                      "extrapolated x%d to %.1f s) + scipy LAPACK zgesv on the full matrix "
                      "(%.2f s); sampled columns agree with the GPU to %.1e" %
                      (args.cpu_cols, n, t_cols, n // args.cpu_cols, t_asm, t_solve, err),
        }
    print(json.dumps(out), flush=True)
    mesh.close()
    ctx.close()


if __name__ == "__main__":
    main()
