"""Autotuner of the launch shapes, per size bucket, on the GPU it runs on.

    python -m bem2d_amd.autotune [--out bem2d_amd/tuned/gfx950.tune] [--quick]

Two choices are measured and written as the text table that bem2d_ctx_create loads
(tuned/gfx950.tune beside the library, or $BEM2D_TUNE_FILE):

* Regular-pair assembly (assemble_regular_kernel): wavefronts per 8 x 8-element tile, 1, 2 or 4.
  More wavefronts per tile shorten each lane's chain of node pairs, which pays while a launch has
  too few tiles to fill the device.  For kite meshes of growing N (10 elements per wavelength) the
  assembly is timed with every choice forced for all sizes; each size takes its fastest choice,
  neighbouring sizes with the same choice merge into one bucket, and bucket bounds (in tiles) fall
  at the geometric mean of the neighbouring tile counts.
* FP32 LU: the trailing-matrix size (in 64 x 64 tiles) from which two panels share one depth-128
  update.  The factorisation of random diagonally weighted matrices is timed at a few N for each
  candidate threshold; the threshold with the least total time relative to each size's best wins.

Timings come from the library's own event pairs around each launch (bem2d_profile_*), as medians of
repeated runs.  The measurements are printed as one JSON line and kept as comments in the table.
"""
import argparse
import datetime
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from bem2d_amd import Context, capi  # noqa: E402
from bem2d_amd import geometry as G  # noqa: E402

ASM_SIZES = [64, 128, 256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144]
ASM_SPLITS = [1, 2, 4]
LU_SIZES = [4000, 5440, 8000]
LU_CANDIDATES = [48, 64, 80, 100, 1 << 30]
KITE_PERIMETER = 9.79  # x = cos t + 0.65 cos 2t - 0.65, y = 1.5 sin t


def tiles(n):
    p = (n + 7) // 8
    return p * (p + 1) // 2


def time_assembly(ctx, n, split, reps):
    import torch
    k = 2 * np.pi * n / (10 * KITE_PERIMETER)
    mesh = ctx.mesh(G.kite(n), 16, 3, 0.15)
    A = torch.empty((n, n), dtype=torch.complex128, device="cuda").t()
    ctx.tune_assembly([(1 << 62, split)])
    ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)  # warm-up
    ms = []
    for _ in range(reps):
        ctx.profile(True)
        ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0, out=A)
        torch.cuda.synchronize()
        t, cnt = ctx.profile_read(capi.PROF_ASSEMBLE_REGULAR)
        ms.append(t / max(cnt, 1))
    ctx.profile(False)
    mesh.close()
    return float(np.median(ms))


def time_lu(ctx, n, pair_tiles, reps):
    import torch
    g = torch.Generator(device="cuda").manual_seed(n)
    A = torch.randn((n, n), dtype=torch.complex128, device="cuda", generator=g) / np.sqrt(n)
    A.diagonal().add_(4.0)
    A = A.t()  # column-major view
    ctx.tune_lu_pair_tiles(pair_tiles)
    lu = ctx.lu32(n)
    ms = []
    for r in range(reps + 1):
        ctx.profile(True)
        lu.factor(A)
        torch.cuda.synchronize()
        t, cnt = ctx.profile_read(capi.PROF_LU)
        if r > 0:  # the first factorisation of a buffer is launched eagerly, the rest replay
            ms.append(t / max(cnt, 1))
    ctx.profile(False)
    lu.close()
    return float(np.median(ms))


def buckets_from(best):
    """best: [(tiles, split)] in increasing tiles -> [(upper_tiles, split)] (last choice applies
    beyond the last size only when it is 1)."""
    out = []
    for i, (t, s) in enumerate(best):
        if i + 1 < len(best) and best[i + 1][1] == s:
            continue
        upper = int(np.sqrt(t * best[i + 1][0])) if i + 1 < len(best) else 1 << 62
        out.append((upper, s))
    # one wavefront per tile is the built-in choice beyond the last bucket
    while out and out[-1][1] == 1:
        out.pop()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuned", "gfx950.tune"))
    ap.add_argument("--quick", action="store_true", help="fewer repetitions")
    args = ap.parse_args()
    reps = 3 if args.quick else 7
    ctx = Context(0)
    meas = {"assembly_ms": {}, "lu_ms": {}}
    best = []
    for n in ASM_SIZES:
        row = {s: time_assembly(ctx, n, s, reps) for s in ASM_SPLITS}
        meas["assembly_ms"][n] = row
        choice = min(row, key=row.get)
        best.append((tiles(n), choice))
        print("assembly N=%5d tiles=%6d  %s  -> %d" % (
            n, tiles(n), "  ".join("split %d: %.3f ms" % (s, row[s]) for s in ASM_SPLITS), choice),
            flush=True)
    asm = buckets_from(best)
    lu_tot = {}
    for n in LU_SIZES:
        row = {c: time_lu(ctx, n, c, reps) for c in LU_CANDIDATES}
        meas["lu_ms"][n] = row
        lo = min(row.values())
        for c, t in row.items():
            lu_tot[c] = lu_tot.get(c, 0.0) + t / lo
        print("LU N=%5d  %s" % (n, "  ".join("pair>=%s: %.3f ms" % (
            c if c < 1 << 30 else "never", t) for c, t in row.items())), flush=True)
    pair = min(lu_tot, key=lu_tot.get)
    quad = ctx.tune_lu_quad_tiles()
    props = __import__("torch").cuda.get_device_properties(0)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("# launch-shape table measured by bem2d_amd/autotune.py on %s (%s), %s\n" % (
            props.name, getattr(props, "gcnArchName", "?"), datetime.date.today().isoformat()))
        f.write("# regular assembly, wavefronts per tile by launch size (tiles); 1 beyond the last\n")
        for n in ASM_SIZES:
            row = meas["assembly_ms"][n]
            f.write("#   N=%d tiles=%d: %s\n" % (n, tiles(n), ", ".join(
                "%d -> %.3f ms" % (s, row[s]) for s in ASM_SPLITS)))
        for upper, s in asm:
            f.write("asm_split %d %d\n" % (upper, s))
        f.write("# FP32 LU, relative total over N=%s per pairing threshold: %s\n" % (
            LU_SIZES, ", ".join("%s: %.3f" % (c if c < 1 << 30 else "never", v)
                                for c, v in lu_tot.items())))
        f.write("lu_pair_tiles %d\n" % pair)
        f.write("# FP32 LU, four panels per depth-256 update from this many tiles (carried over: the\n"
                "# sizes above stay below it; measured on the open arc N=32000)\n")
        f.write("lu_quad_tiles %d\n" % quad)
        f.write("# FP32 LU, eight panels per depth-512 update from this many tiles (carried over)\n")
        f.write("lu_oct_tiles %d\n" % ctx.tune_lu_oct_tiles())
        f.write("# FP32 LU, diagonal block factorised apart from the update from this many trailing\n"
                "# tiles (carried over; measured on the open arc N=32000 and the kite N=4000)\n")
        f.write("lu_split_tiles %d\n" % ctx.tune_lu_split_tiles())
    out = {"table": args.out, "assembly_buckets": asm, "lu_pair_tiles": pair,
           "measurements": meas}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
