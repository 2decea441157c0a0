"""GPU checks of the learned warm start: the device-side features and
x_n write-back against host computations on the same context, and the gain
of the trained network over the plain iteration on the 256^3 clumpy
workload (wall time to mean |dx_n| < 1e-3 from x_n = 1e-6, same packets per
iteration, equal converged field and ionized radius)."""
import os

import numpy as np
import pytest

import parallelcmacionize_amd as cmi
from parallelcmacionize_amd import warmstart as WS
from parallelcmacionize_amd import workloads as W

pytestmark = pytest.mark.gpu
NET = os.path.join(os.path.dirname(cmi.__file__), "data",
                   "warmstart_clumpy256.npz")


@pytest.fixture(scope="module")
def small():
    box, n, pos, L = WS.clumpy_problem(64, 2.5, 2, field_seed=41, src_seed=42,
                                       lum_seed=43, lmin=1e46, lmax=1e47)
    prob = WS.Problem(box, n, pos, L)
    net = WS.new_net(5)
    yield prob, net
    net.close()
    prob.close()


def host_features(prob, npackets):
    J, _ = prob.ctx.get_fields()
    J = W.subgrid_major_to_global(J, prob.ncell, prob.nsub)
    cell = prob.ctx.box[3] / prob.ncell[0]
    jH = prob.L / npackets * J / cell ** 3
    n = W.subgrid_major_to_global(prob.n_ref, prob.ncell, prob.nsub)
    return np.stack([np.log10(n), np.log10(jH + 1e-12)], axis=-1)


def test_context_features_match_host(small):
    prob, net = small
    prob.reset()
    N = 200000
    prob.first_pass(seed=3, npackets=N)
    f = net.context_features(prob.ctx, prob.L, N)
    h = host_features(prob, N)
    assert f.shape == h.shape
    assert np.abs(f - h).max() < 2e-5
    assert (h[..., 1] > -11).mean() > 0.5   # most cells saw packets


def test_warm_start_writes_network_prediction(small):
    prob, net = small
    prob.reset()
    N = 200000
    prob.first_pass(seed=4, npackets=N)
    f = net.context_features(prob.ctx, prob.L, N)
    z = net.predict(f).astype(np.float64)
    x_expect = np.clip(1.0 / (1.0 + np.exp(-z)), 1e-7, 1.0)
    net.warm_start(prob.ctx, prob.L, N)
    J, x = prob.ctx.get_fields()
    x = W.subgrid_major_to_global(x, prob.ncell, prob.nsub)
    assert np.all(J == 0)
    assert np.allclose(x, x_expect, rtol=1e-5, atol=1e-12)
    # the opacity follows the new x_n: a transport now sees it
    s = prob.ctx.transport(N, 5, 1)
    assert s["errors"] == 0
    # the start rule: cells predicted more neutral than the cut start at the
    # fallback value
    prob.reset()
    prob.first_pass(seed=4, npackets=N)
    cut = float(np.median(z))
    net.set_start_rule(cut, 1e-6)
    net.warm_start(prob.ctx, prob.L, N)
    net.set_start_rule(3.0e38, 1e-6)
    x = W.subgrid_major_to_global(prob.ctx.get_fields()[1], prob.ncell, prob.nsub)
    expect = np.where(z * 1.0 > cut, 1e-6, x_expect)
    mism = np.abs(x - expect) > 1e-5 * np.maximum(expect, 1e-12)
    # logits within float rounding of the cut may fall either side
    assert mism.sum() <= 8


def test_coarse_start_restrict_and_prolong():
    """The coarse start's device steps against numpy on a 64^3 grid and its
    16^3 coarsening: the coarse n_H is the mean of each 4^3 block (read back
    through the network's mode-1 features, log10 n_H and the logit of x_n =
    1e-6), and the fine x_n is the coarse logit (clamped) interpolated
    trilinearly at the fine cell centres, the edge cells extended."""
    box = W.box_10pc()
    rng = np.random.default_rng(3)
    nf = 1e8 * np.exp(rng.normal(0, 1, (64, 64, 64)))
    xc = 10 ** rng.uniform(-7.5, 0, (16, 16, 16))
    net = WS.new_net(5)
    net.set_input_mode(1)
    with cmi.Context(box, [64] * 3, [4] * 3, cmi.PREC_FP32) as fine, \
            cmi.Context(box, [16] * 3, [1] * 3, cmi.PREC_FP32) as coarse:
        fine.set_fields(W.global_to_subgrid_major(nf, [4] * 3), np.full(64 ** 3, 1e-6))
        fine.restrict_density_to(coarse)
        f = net.context_features(coarse, 1.0, 1).reshape(16, 16, 16, 2)
        coarse.set_fields(xn=W.global_to_subgrid_major(xc, [1] * 3))
        fine.prolong_neutral_fraction_from(coarse)
        J, xf = fine.get_fields()
        xf = W.subgrid_major_to_global(xf, [64] * 3, [4] * 3)
    net.close()
    n_blk = nf.reshape(16, 4, 16, 4, 16, 4).mean(axis=(1, 3, 5))
    assert np.allclose(f[..., 0], np.log10(n_blk), rtol=0, atol=2e-6)
    assert np.allclose(f[..., 1], np.log(1e-6 / (1 - 1e-6)), rtol=1e-6)
    # trilinear = three 1-D linear interpolations
    u = (np.arange(64) + 0.5) / 4 - 0.5
    i0 = np.floor(u).astype(int)
    w = u - i0
    c0, c1 = np.clip(i0, 0, 15), np.clip(i0 + 1, 0, 15)
    lg = WS.logit(xc)
    lg = lg[c0] * (1 - w)[:, None, None] + lg[c1] * w[:, None, None]
    lg = lg[:, c0] * (1 - w)[None, :, None] + lg[:, c1] * w[None, :, None]
    lg = lg[:, :, c0] * (1 - w)[None, None, :] + lg[:, :, c1] * w[None, None, :]
    assert np.allclose(xf, 1 / (1 + np.exp(-lg)), rtol=1e-12, atol=0)
    assert not J.any()


def test_coarse_start_fp64_and_refusals():
    """The coarse start on FP64 contexts (no opacity array) gives the same
    fields as on FP32 ones, and grids that are not an integer coarsening of
    the same box are refused with CMI_ERR_ARG, leaving the fields alone."""
    box = W.box_10pc()
    rng = np.random.default_rng(4)
    nf = 1e8 * np.exp(rng.normal(0, 1, (32, 32, 32)))
    xc = 10 ** rng.uniform(-7, 0, (8, 8, 8))
    out = {}
    for prec in (cmi.PREC_FP32, cmi.PREC_FP64):
        with cmi.Context(box, [32] * 3, [2] * 3, prec) as fine, \
                cmi.Context(box, [8] * 3, [1] * 3, prec) as coarse:
            fine.set_fields(W.global_to_subgrid_major(nf, [2] * 3), np.full(32 ** 3, 1e-6))
            fine.restrict_density_to(coarse)
            coarse.set_fields(xn=W.global_to_subgrid_major(xc, [1] * 3))
            fine.prolong_neutral_fraction_from(coarse)
            out[prec] = fine.get_fields()[1]
    assert np.array_equal(out[cmi.PREC_FP32], out[cmi.PREC_FP64])
    other = np.array(box, dtype=float)
    other[3:] *= 2.0                       # same cells, another box
    with cmi.Context(box, [32] * 3, [2] * 3, cmi.PREC_FP32) as fine, \
            cmi.Context(box, [12] * 3, [1] * 3, cmi.PREC_FP32) as c12, \
            cmi.Context(other, [8] * 3, [1] * 3, cmi.PREC_FP32) as c_box, \
            cmi.Context(box, [32] * 3, [2] * 3, cmi.PREC_FP32) as same:
        fine.set_fields(W.global_to_subgrid_major(nf, [2] * 3), np.full(32 ** 3, 0.5))
        for c in (c12, c_box, same):
            with pytest.raises(cmi.CMIError):
                fine.restrict_density_to(c)
            with pytest.raises(cmi.CMIError):
                fine.prolong_neutral_fraction_from(c)
        assert np.all(fine.get_fields()[1] == 0.5)


@pytest.mark.skipif(not os.path.exists(NET), reason="no trained network file")
def test_trained_network_gain_on_256_clumpy():
    """The committed network on the 256^3 clumpy workload (never trained on):
    fewer plain iterations and less wall time to mean |dx_n| < 1e-3, at the
    plain iteration's accuracy: distance to the tightly converged fixed point
    of the same packet set within 1.25x the plain run's, ionized radius
    within 0.5 %."""
    box, n, pos, L = WS.clumpy_problem()          # the 256^3 clumpy workload
    prob = WS.Problem(box, n, pos, L)
    net = WS.load_net(NET)
    N = 100_000_000
    prob.first_pass(seed=99)                      # builds the engine buffers
    WS.converge_plain(prob, N, seed=21, tol=1e-5, max_it=60)
    x_ref = prob.xn_lattice()
    plain = WS.converge_plain(prob, N, seed=21)
    x_plain = prob.xn_lattice()
    warm = WS.converge_warm(prob, net, N, seed=21)
    x_warm = prob.xn_lattice()
    net.close()
    prob.close()
    ratio = plain["total_seconds"] / warm["total_seconds"]
    e_p = float(np.mean(np.abs(x_plain - x_ref)))
    e_w = float(np.mean(np.abs(x_warm - x_ref)))
    r_p = W.ionized_radius(x_plain, box, [256] * 3)
    r_w = W.ionized_radius(x_warm, box, [256] * 3)
    print("plain", plain["iterations"], "%.3f s" % plain["total_seconds"], "warm",
          warm["iterations"], "%.3f s" % warm["total_seconds"], "ratio %.2f" % ratio,
          "errors %.2e %.2e" % (e_p, e_w), "R %.5f" % (r_w / r_p))
    assert plain["converged"] and warm["converged"]
    assert warm["iterations"] < plain["iterations"]
    assert ratio >= 1.2
    assert e_w <= 1.25 * e_p
    assert abs(r_w / r_p - 1) < 0.005
