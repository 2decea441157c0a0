"""GPU checks of problem sets (cmi_batch_*): many grids of uneven size in one
graph-captured iteration.  Problem 0 of a set draws the same packets as a
single context with the same seed, so its result must match that context's
(bar: 1e-9 relative on x_n, the sums differ only in order); every problem of a
set of Stromgren spheres must find its own analytic radius."""
import numpy as np
import pytest

import parallelcmacionize_amd as cmi
from parallelcmacionize_amd import workloads as W

pytestmark = pytest.mark.gpu
PC = W.PC


def test_batch_problem0_matches_single_context():
    box = W.box_10pc()
    n64 = W.lognormal_density(64, 1.0, seed=7)
    pos, L = W.random_sources(4, box, seed=8)
    N = 300000
    # single context, 16^3 subgrids: its device layout is the set's layout
    with cmi.Context(box, [64] * 3, [4, 4, 4], cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES)
        ctx.set_fields(W.global_to_subgrid_major(n64, [4, 4, 4]),
                       np.full(64 ** 3, 1e-4))
        ctx.set_sources(pos, L)
        s1 = ctx.transport(N, 11, 2)
        l1_single = ctx.update_ionization(L.sum(), N, W.ALPHA_B, form=1)
        x_single = W.subgrid_major_to_global(ctx.get_fields()[1], [64] * 3,
                                             [4, 4, 4])
    boxes = [box, [0, 0, 0, 2 * PC, 2 * PC, 2 * PC],
             [PC, PC, PC, 3 * PC, 3 * PC, 3 * PC]]
    with cmi.Batch(boxes, [[64] * 3, [32] * 3, [48] * 3]) as b:
        b.set_fields(0, n64, np.full(64 ** 3, 1e-4))
        b.set_sources(0, pos, L)
        b.set_sources(1, [[PC, PC, PC]], [1e47])
        b.set_sources(2, [[2.5 * PC, 2.5 * PC, 2.5 * PC]], [3e47])
        r = b.iteration([N, 50000, 70000], 11, 2, W.ALPHA_B, form=1)
        x0 = b.get_fields(0)[1]
    assert r["errors"] == 0
    assert r["absorbed"] + r["escaped"] == N + 50000 + 70000
    assert int(r["problem_crossings"].sum()) == r["crossings"]
    assert int(r["problem_crossings"][0]) == s1["crossings"]
    assert np.allclose(x0, x_single, rtol=1e-9, atol=0)
    assert abs(r["l1"][0] - l1_single) < 1e-9 * l1_single
    assert (r["problem_seconds"] > 0).all()
    assert (r["problem_seconds"] <= r["seconds"] * 1.01).all()


def test_batch_stromgren_spheres():
    """Uniform Stromgren spheres of different sizes in one set: each finds
    R = (3 Q / (4 pi n^2 alpha))^(1/3) of its own source."""
    sizes = [32, 48, 64, 48]
    sides = [4.0, 6.0, 8.0, 5.0]
    boxes, ncells, Qs = [], [], []
    for nc, side in zip(sizes, sides):
        boxes.append([-side / 2 * PC] * 3 + [side * PC] * 3)
        ncells.append([nc] * 3)
        Rs = 0.3 * side * PC
        Qs.append(4 / 3 * np.pi * Rs ** 3 * W.N_H ** 2 * W.ALPHA_B)
    N = [600000, 800000, 1000000, 800000]
    with cmi.Batch(boxes, ncells) as b:
        for p in range(len(sizes)):
            b.set_sources(p, [[0.0, 0.0, 0.0]], [Qs[p]])
        for it in range(14):
            r = b.iteration(N, 5, it, W.ALPHA_B, form=1)
            assert r["errors"] == 0
        print("iteration seconds", r["seconds"], "per problem",
              r["problem_seconds"])
        for p in range(len(sizes)):
            x = b.get_fields(p)[1]
            R = W.ionized_radius(x.reshape(-1), np.array(boxes[p]), ncells[p])
            Rs = W.stromgren_radius(Qs[p], W.N_H, W.ALPHA_B)
            print("problem", p, "R/Rs", R / Rs)
            assert abs(R / Rs - 1) < 0.03
