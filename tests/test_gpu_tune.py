"""Launch-shape table (bem2d_tune_*): every assembly shape the autotuner can choose gives the
oracle's matrix, every FP32 LU panel-pairing choice the FP64 solution, and the table loader
accepts the autotuner's format and rejects malformed tables."""
import numpy as np
import pytest
import torch

from bem2d_amd import capi
from bem2d_amd import geometry as G
from conftest import rel_fro

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,geom,k", [("circle200", G.unit_circle(200), 10.0),
                                         ("kite333", G.kite(333), 40.0)])
def test_every_assembly_split_matches_oracle(ctx, oracle, name, geom, k):
    saved = ctx.tune_assembly()
    try:
        mesh = ctx.mesh(geom, 16, 3, 0.15)
        R = oracle.assemble(geom, 16, 3, 0.15, oracle.K_COMBINED, k, k, 1.0, 2.0)
        mats = {}
        for split in (1, 2, 4):
            ctx.tune_assembly([(1 << 62, split)])
            mats[split] = ctx.assemble(mesh, capi.KERNEL_COMBINED, k, k, 1.0, 2.0).cpu().numpy()
            err = rel_fro(mats[split], R)
            print(name, "split", split, "%.2e" % err, flush=True)
            assert err < 1e-12
        assert rel_fro(mats[2], mats[1]) < 1e-14 and rel_fro(mats[4], mats[1]) < 1e-14
        mesh.close()
    finally:
        ctx.tune_assembly(saved)


def test_lu_pairing_choices_solve(ctx):
    n = 900
    rng = np.random.default_rng(9)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A += 4 * np.eye(n)
    b = rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))
    x_ref = np.linalg.solve(A, b)
    At = torch.tensor(np.ascontiguousarray(A.T), device="cuda").t()
    saved = ctx.tune_lu_pair_tiles(), ctx.tune_lu_quad_tiles(), ctx.tune_lu_oct_tiles()
    never = 1 << 30
    try:
        # (twos from, fours from, eights from): pairs whenever possible, from 10 tiles, never; fours
        # whenever possible; fours from 8 tiles with pairs below; eights whenever possible; eights
        # from 12 tiles, fours from 8, pairs below
        for pair, quad, oct_ in ((1, never, never), (10, never, never), (never, never, never),
                                 (never, 1, never), (1, 8, never), (never, never, 1), (1, 8, 12)):
            ctx.tune_lu_pair_tiles(pair)
            ctx.tune_lu_quad_tiles(quad)
            ctx.tune_lu_oct_tiles(oct_)
            B = torch.tensor(np.ascontiguousarray(b.T), device="cuda").t().clone()
            rep = ctx.solve_mixed(At, B, tol=1e-12)
            err = rel_fro(B.cpu().numpy(), x_ref)
            print("pair tiles", pair, "quad tiles", quad, "oct tiles", oct_,
                  "backward %.2e" % rep.backward_error, "err %.2e" % err, flush=True)
            assert rep.fallback_fp64 == 0 and rep.backward_error <= 1e-12 and err < 1e-10
    finally:
        ctx.tune_lu_pair_tiles(saved[0])
        ctx.tune_lu_quad_tiles(saved[1])
        ctx.tune_lu_oct_tiles(saved[2])


def test_lu_split_diagonal_matches_fused(ctx):
    """The FP32 LU with the diagonal block factorised apart from the update (two workgroups per
    compute unit) gives the fused path's solution bit for bit (the same products, summed in the
    same order), with panels single, in twos, fours and eights, and on a replayed graph."""
    n = 1500  # 24 block columns: split on every step with lu_split_tiles = 1
    rng = np.random.default_rng(11)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    A += 4 * np.eye(n)
    b = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    x_ref = np.linalg.solve(A, b)
    At = torch.tensor(np.ascontiguousarray(A.T), device="cuda").t()
    saved_s, saved_p = ctx.tune_lu_split_tiles(), ctx.tune_lu_pair_tiles()
    saved_q, saved_o = ctx.tune_lu_quad_tiles(), ctx.tune_lu_oct_tiles()
    try:
        for pair, quad, oct_ in ((1 << 30, 1 << 30, 1 << 30), (4, 1 << 30, 1 << 30),
                                 (4, 8, 1 << 30), (4, 8, 12)):
            ctx.tune_lu_pair_tiles(pair)
            ctx.tune_lu_quad_tiles(quad)
            ctx.tune_lu_oct_tiles(oct_)
            facs = []
            for split in (1 << 30, 1):
                ctx.tune_lu_split_tiles(split)
                lu = ctx.lu32(n)
                for _ in range(3):  # the third factorisation replays a captured graph
                    lu.factor(At)
                B = torch.tensor(np.ascontiguousarray(b.T), device="cuda").t().clone()
                X = torch.empty_like(B)
                rep = lu.refine(At, B, X, tol=1e-12)
                assert rep.converged and rel_fro(X.cpu().numpy(), x_ref) < 1e-10
                facs.append(X.cpu().numpy())
                print("pair", pair, "quad", quad, "oct", oct_, "split", split, "replays",
                      lu.graph_replays(), flush=True)
                lu.close()
            assert np.array_equal(facs[0], facs[1])
    finally:
        ctx.tune_lu_split_tiles(saved_s)
        ctx.tune_lu_pair_tiles(saved_p)
        ctx.tune_lu_quad_tiles(saved_q)
        ctx.tune_lu_oct_tiles(saved_o)


def test_table_load(ctx, tmp_path):
    saved_a, saved_p = ctx.tune_assembly(), ctx.tune_lu_pair_tiles()
    saved_q, saved_o = ctx.tune_lu_quad_tiles(), ctx.tune_lu_oct_tiles()
    try:
        good = tmp_path / "good.tune"
        good.write_text("# measured\nasm_split 500 4\nasm_split 9000 2  # comment\n"
                        "lu_pair_tiles 77\nlu_quad_tiles 150\nlu_oct_tiles 260\n")
        ctx.tune_load(str(good))
        assert ctx.tune_assembly() == [(500, 4), (9000, 2)]
        assert ctx.tune_lu_pair_tiles() == 77 and ctx.tune_lu_quad_tiles() == 150
        assert ctx.tune_lu_oct_tiles() == 260
        for text in ("asm_split 500 3\n", "asm_split 900 4\nasm_split 500 2\n", "bogus 1\n",
                     "lu_pair_tiles 0\n", "lu_quad_tiles 0\n", "lu_oct_tiles 0\n"):
            bad = tmp_path / "bad.tune"
            bad.write_text(text)
            with pytest.raises(capi.Bem2dError) as e:
                ctx.tune_load(str(bad))
            assert e.value.code == capi.ERR_ARG
        with pytest.raises(capi.Bem2dError):
            ctx.tune_load(str(tmp_path / "missing.tune"))
    finally:
        ctx.tune_assembly(saved_a)
        ctx.tune_lu_pair_tiles(saved_p)
        ctx.tune_lu_quad_tiles(saved_q)
        ctx.tune_lu_oct_tiles(saved_o)
