"""Serving engine on the GPU: requests of uneven size admitted by HBM footprint, batched stages,
results equal to the single-problem paths, each result the same alone and in the mix, and the
far field of the circle against the analytic series."""
import math

import numpy as np
import pytest
import torch

from bem2d_amd import Engine, capi, serving
from bem2d_amd import geometry as G
from conftest import analytic_farfield, rel_fro

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _req(geom, k, nrhs=1, seed=0, nang=36, precision=capi.PRECISION_MIXED, tol=1e-12):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, size=nrhs)
    return {"geom": geom, "k": k, "eta": k, "dirs": np.stack([np.cos(th), np.sin(th)], 1),
            "theta": 2 * np.pi * np.arange(nang) / nang, "precision": precision, "tol": tol}


def _mix():
    return [_req(G.unit_circle(200), 10.0, 1, seed=1),
            _req(G.kite(900), 40.0, 3, seed=2),
            _req(serving.polygon_sides([(1, 0), (-0.5, 0.8), (-0.4, -0.9)], [90, 110, 100]), 30.0,
                 2, seed=3),
            _req(serving.ellipse(1.5, 0.7, 333), 25.0, 9, seed=4),
            _req(G.open_arc(257), 20.0, 1, seed=5)]


def _run(engine, reqs, arrivals=None):
    if arrivals is None:
        tickets = [engine.submit(q) for q in reqs]
        while engine.step() > 0:
            pass
    else:
        tickets = engine.run_trace(reqs, arrivals)
    return [engine.result(t) for t in tickets]


def _single(ctx, q):
    """The same request through the single-problem entry points."""
    m = ctx.mesh(q["geom"], 16, 3, 0.15)
    n = q["geom"].n
    A = ctx.assemble(m, capi.KERNEL_COMBINED, q["k"], q["eta"], 1.0, 2.0)
    B = ctx.rhs(m, capi.FIELD_COMBINED_PLANE, q["k"], q["eta"], q["dirs"], 2.0)
    if q["precision"] == capi.PRECISION_MIXED:
        lu = ctx.lu32(n)
        lu.factor(A)
        X = torch.empty_like(B)
        lu.refine(A, B, X, tol=q["tol"])
        lu.close()
    else:
        X = ctx.solve_fp64(A, B.clone())
    th = torch.tensor(q["theta"], device="cuda")
    ff = ctx.farfield(m, q["k"], X, th)
    torch.cuda.synchronize()
    m.close()
    return X.cpu().numpy(), ff.cpu().numpy()


@pytest.mark.parametrize("precision", [capi.PRECISION_MIXED, capi.PRECISION_FP64])
def test_engine_matches_single_problem_path(ctx, precision):
    print("case engine vs single path precision=%d" % precision, flush=True)
    reqs = _mix()
    for q in reqs:
        q["precision"] = precision
    eng = Engine(ctx, 4 * GB)
    res = _run(eng, reqs)
    st = eng.stats()
    eng.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0, st
    for q, (x, ff, info) in zip(reqs, res):
        assert info.status == capi.OK
        xs, ffs = _single(ctx, q)
        # identical kernels on identical data: the batch composition does not change the arithmetic
        assert rel_fro(x, xs) <= 1e-13, (q["geom"].n, rel_fro(x, xs))
        assert rel_fro(ff, ffs) <= 1e-13
        if precision == capi.PRECISION_MIXED:
            assert info.backward_error <= 1e-12 and info.iterations >= 1


def test_each_result_same_alone_and_in_the_mix(ctx):
    print("case isolation", flush=True)
    reqs = _mix()
    eng = Engine(ctx, 4 * GB)
    alone = [_run(eng, [q])[0] for q in reqs]
    # all together, arriving in reverse order a few milliseconds apart
    order = list(range(len(reqs)))[::-1]
    mixed = _run(eng, [reqs[i] for i in order], arrivals=[0.002 * p for p in range(len(reqs))])
    eng.close()
    for p, i in enumerate(order):
        xa, fa, _ = alone[i]
        xm, fm, info = mixed[p]
        assert info.status == capi.OK
        assert rel_fro(xm, xa) <= 1e-14, (i, rel_fro(xm, xa))
        assert rel_fro(fm, fa) <= 1e-14


def test_grouped_panels_in_batched_launches(ctx):
    """With panels grouped in twos from 2 trailing tiles, fours from 4 and eights from 6, the
    problems of one batched LU launch sit at different points of their groups (different update
    depths in one launch): results still match the single-problem path, which groups the same way,
    and are the same alone and in the mix."""
    print("case engine grouped panels", flush=True)
    saved = ctx.tune_lu_pair_tiles(), ctx.tune_lu_quad_tiles(), ctx.tune_lu_oct_tiles()
    ctx.tune_lu_pair_tiles(2)
    ctx.tune_lu_quad_tiles(4)
    ctx.tune_lu_oct_tiles(6)
    try:
        reqs = _mix()
        eng = Engine(ctx, 4 * GB)
        res = _run(eng, reqs)
        alone = [_run(eng, [q])[0] for q in reqs[:2]]
        st = eng.stats()
        eng.close()
        assert st["failed"] == 0, st
        for q, (x, ff, info) in zip(reqs, res):
            assert info.status == capi.OK and info.backward_error <= 1e-12
            xs, ffs = _single(ctx, q)
            assert rel_fro(x, xs) <= 1e-13, (q["geom"].n, rel_fro(x, xs))
            assert rel_fro(ff, ffs) <= 1e-13
        for (xa, fa, _), (xm, fm, _) in zip(alone, res[:2]):
            assert rel_fro(xm, xa) <= 1e-14 and rel_fro(fm, fa) <= 1e-14
    finally:
        ctx.tune_lu_pair_tiles(saved[0])
        ctx.tune_lu_quad_tiles(saved[1])
        ctx.tune_lu_oct_tiles(saved[2])


def test_circle_far_field_against_series(ctx):
    print("case circle far field", flush=True)
    q = _req(G.unit_circle(200), 10.0, 1, nang=360)
    q["dirs"] = np.array([[1.0, 0.0]])
    eng = Engine(ctx, GB)
    (x, ff, info), = _run(eng, [q])
    eng.close()
    ref = analytic_farfield(10.0, q["theta"])
    # discretisation error of N = 200 P0 elements (the oracle's own value is ~1e-3)
    assert rel_fro(ff[:, 0], ref) < 2.5e-3


def test_admission_waits_for_memory(ctx):
    print("case admission by footprint", flush=True)
    reqs = [_req(G.kite(800), 30.0, 2, seed=s) for s in range(5)]
    one = Engine.footprint(800, 2, 36)
    eng = Engine(ctx, int(1.5 * one), max_active=8)   # room for one request at a time
    tickets = [eng.submit(q) for q in reqs]
    peak_active = 0
    while eng.step() > 0:
        st = eng.stats()
        peak_active = max(peak_active, st["active"])
        assert st["arena_in_use"] <= st["arena_bytes"]
    res = [eng.result(t) for t in tickets]
    eng.close()
    assert peak_active == 1
    for q, (x, ff, info) in zip(reqs, res):
        assert info.status == capi.OK and info.backward_error <= 1e-12
        assert info.t_admit >= info.t_arrival and info.t_done >= info.t_admit


def test_request_larger_than_arena_fails_cleanly(ctx):
    print("case oversized request", flush=True)
    big = _req(G.kite(3000), 50.0)
    small = _req(G.kite(300), 20.0)
    eng = Engine(ctx, Engine.footprint(1000, 1, 36) + (1 << 20))
    tb = eng.submit(big)
    ts = eng.submit(small)
    while eng.step() > 0:
        pass
    _, _, ib = eng.result(tb, raise_on_fail=False)
    _, _, is_ = eng.result(ts)
    eng.close()
    assert ib.status == capi.ERR_NOMEM
    assert is_.status == capi.OK


def test_poisson_trace_latency(ctx):
    print("case short Poisson trace", flush=True)
    reqs, arr = serving.serving_trace(24, rate=200.0, seed=7, k_range=(10.0, 60.0), max_rhs=4)
    eng = Engine(ctx, 8 * GB)
    res = _run(eng, reqs, arrivals=arr)
    st = eng.stats()
    eng.close()
    infos = [r[2] for r in res]
    assert all(i.status == capi.OK for i in infos)
    lat = serving.latency_summary(infos)
    assert 0 < lat["p50_ms"] <= lat["p99_ms"]
    # continuous batching: fewer LU launches than requests x panel steps
    steps = sum(math.ceil(q["geom"].n / 64) for q in reqs)
    assert st["lu_step_launches"] < steps


def test_mixed_and_fp64_requests_share_one_engine(ctx):
    """MIXED, then a larger FP64 request (it grows the rocSOLVER workspace), then MIXED again in
    one engine: every result equals its single-problem path and the engine closes cleanly."""
    print("case mixed and FP64 in one engine", flush=True)
    reqs = [_req(G.kite(400), 30.0, 2, seed=21),
            _req(G.kite(1500), 60.0, 3, seed=22, precision=capi.PRECISION_FP64),
            _req(G.unit_circle(300), 20.0, 9, seed=23),
            _req(G.kite(2000), 70.0, 1, seed=24, precision=capi.PRECISION_FP64),
            _req(G.open_arc(500), 40.0, 4, seed=25)]
    eng = Engine(ctx, 4 * GB)
    res = []
    for q in reqs:  # one after another, so each FP64 request finds the other's state behind it
        res += _run(eng, [q])
    res_mix = _run(eng, reqs)  # and all at once
    eng.close()
    for q, (x, ff, info), (xm, fm, im) in zip(reqs, res, res_mix):
        assert info.status == capi.OK and im.status == capi.OK
        xs, ffs = _single(ctx, q)
        assert rel_fro(x, xs) <= 1e-13 and rel_fro(xm, xs) <= 1e-13, q["geom"].n
        assert rel_fro(ff, ffs) <= 1e-13 and rel_fro(fm, ffs) <= 1e-13


def test_arena_of_one_footprint(ctx):
    """An arena of exactly one request's footprint serves several such requests one after
    another (the refinement's read-back slots live outside the arena)."""
    print("case arena of one footprint", flush=True)
    reqs = [_req(G.kite(600), 30.0, 3, seed=30 + s) for s in range(3)]
    eng = Engine(ctx, Engine.footprint(600, 3, 36))
    res = _run(eng, reqs)
    eng.close()
    for x, ff, info in res:
        assert info.status == capi.OK and info.backward_error <= 1e-12


def test_unreachable_tolerance_falls_back_to_fp64(ctx):
    """A mixed request whose tolerance plain refinement cannot reach (1e-30) stalls, and the
    engine solves it by the pivoted FP64 LU: status OK, fallback flagged, the FP64 result."""
    print("case engine FP64 fallback", flush=True)
    q = _req(G.kite(700), 40.0, 2, seed=41, tol=1e-30)
    ok = _req(G.kite(500), 30.0, 1, seed=42)
    eng = Engine(ctx, 2 * GB)
    (x, ff, info), (x2, ff2, info2) = _run(eng, [q, ok])
    eng.close()
    assert info.status == capi.OK and info.fallback_fp64 == 1
    assert info2.status == capi.OK and info2.fallback_fp64 == 0
    q64 = dict(q, precision=capi.PRECISION_FP64)
    xs, ffs = _single(ctx, q64)
    assert rel_fro(x, xs) <= 1e-13 and rel_fro(ff, ffs) <= 1e-13
