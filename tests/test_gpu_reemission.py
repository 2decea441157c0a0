"""Reemission (do_reemission, testDensitySubGrid_omp.cpp:675-702): with
probability p an absorbed packet leaves its absorption point again in a new
isotropic direction.  Each source photon is then absorbed 1/(1-p) times on
average, so the Stromgren radius grows to
R = (3 Q / ((1 - p) 4 pi n^2 alpha))^(1/3)
(scripts/plot_neutral_fraction_profile.py:68)."""
import numpy as np
import pytest

import parallelcmacionize_amd as cmi
from parallelcmacionize_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p", [0.0, 0.5])
def test_reemission_stromgren_radius(p):
    box = W.box_10pc()
    Rs0 = 3.0 * W.PC
    Q = 4 / 3 * np.pi * Rs0 ** 3 * W.N_H ** 2 * W.ALPHA_B
    N = 1_000_000
    with cmi.Context(box, [64] * 3, [4, 4, 4], cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES)
        ctx.set_fields(np.full(64 ** 3, W.N_H), np.full(64 ** 3, W.XN0))
        ctx.set_sources([[0.0, 0.0, 0.0]], [Q])
        ctx.set_reemission(p)
        for it in range(14):
            s = ctx.transport(N, 3, it)
            assert s["errors"] == 0
            assert s["absorbed"] + s["escaped"] == N
            ctx.update_ionization(Q, N, W.ALPHA_B, form=1)
        x = ctx.get_fields()[1]
    R = W.ionized_radius(x, box, [64] * 3)
    Rp = Rs0 / (1.0 - p) ** (1.0 / 3.0)
    print("p", p, "R/Rp", R / Rp, "rounds", s["rounds"] if "rounds" in s else None)
    assert abs(R / Rp - 1) < 0.02


def test_reemission_needs_tiles_engine():
    box = W.box_10pc()
    with cmi.Context(box, [32] * 3, [2, 2, 2], cmi.PREC_FP64) as ctx:
        ctx.set_sources([[0.0, 0.0, 0.0]], [1e49])
        ctx.set_reemission(0.3)
        with pytest.raises(cmi.CMIError):
            ctx.transport(1000, 1, 0)
        with pytest.raises(cmi.CMIError):
            ctx.set_reemission(1.0)
