"""The learned warm start's kernels against a CPU oracle: PyTorch CPU conv3d
and autograd in FP32 on the same weights, with the network's BF16 roundings
emulated (inputs, weights and every stored activation rounded to BF16).

Bars: forward logits within 2e-2 of the logit scale; loss within 1e-2
relative; per-layer weight and bias gradients with cosine similarity >= 0.995
and relative norm difference <= 5e-2 (the kernels store activation gradients
in BF16, the oracle keeps them FP32)."""
import numpy as np
import pytest

import parallelcmacionize_amd as cmi

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LAYERS = [(2, 32, 32), (32, 32, 32), (32, 32, 32), (32, 1, 16)]  # cin, cout, NOUT


def split_params(p):
    """master vector -> list of (W torch (cout, cin, 3, 3, 3), b (cout,))"""
    out, off = [], 0
    for cin, cout, nout in LAYERS:
        w = p[off:off + 27 * nout * 32].reshape(27, nout, 32)
        off += 27 * nout * 32
        b = p[off:off + nout]
        off += nout
        wt = w[:, :cout, :cin].reshape(3, 3, 3, cout, cin).transpose(3, 4, 0, 1, 2)
        out.append((np.ascontiguousarray(wt), np.array(b[:cout])))
    assert off == p.size
    return out


def join_grads(p_like, grads):
    """list of (dW torch layout, db) -> master-vector layout"""
    g = np.zeros_like(p_like)
    off = 0
    for (cin, cout, nout), (dw, db) in zip(LAYERS, grads):
        w = np.zeros((27, nout, 32), np.float32)
        w[:, :cout, :cin] = dw.transpose(2, 3, 4, 0, 1).reshape(27, cout, cin)
        g[off:off + 27 * nout * 32] = w.reshape(-1)
        off += 27 * nout * 32
        g[off:off + cout] = db
        off += nout
    return g


def bf16(x):
    return x.to(torch.bfloat16).to(torch.float32)


def torch_forward(params, feat, dil=(1, 1, 1, 1)):
    """feat (B, D, H, W, 2) normalised features -> logits (B, D, H, W)"""
    x = bf16(torch.tensor(feat).permute(0, 4, 1, 2, 3))
    ws = []
    for i, (w, b) in enumerate(params):
        wt = bf16(torch.tensor(w)).requires_grad_(True)
        bt = torch.tensor(b).requires_grad_(True)
        ws.append((wt, bt))
        x = torch.nn.functional.conv3d(x, wt, bt, padding=dil[i], dilation=dil[i])
        if i < 3:
            x = bf16(torch.relu(x))
    return x[:, 0], ws


@pytest.fixture(scope="module")
def net():
    n = cmi.WarmStartNet(seed=3)
    n.set_normalisation([0.0, 0.0], [1.0, 1.0], 1.0)
    yield n
    n.close()


def test_forward_matches_torch(net):
    rng = np.random.default_rng(0)
    feat = rng.normal(size=(16, 16, 32, 2)).astype(np.float32)
    z = net.predict(feat)
    zt, _ = torch_forward(split_params(net.params()), feat[None])
    zt = zt[0].detach().numpy()
    scale = np.abs(zt).max()
    assert np.abs(z - zt).max() < 2e-2 * scale, (np.abs(z - zt).max(), scale)


@pytest.mark.parametrize("shape,dil", [((32, 64, 128), (1, 2, 4, 8)),
                                       ((64, 64, 64), (1, 2, 4, 8)),
                                       ((16, 16, 32), (1, 1, 1, 1))])
def test_tiled_forward_identical_to_per_run(shape, dil):
    """The LDS-tiled forward convolution (blocks of a dilation sub-lattice)
    against the per-run kernel, same network and inputs: identical logits,
    bit for bit, on a volume where every layer tiles and on one where the
    dilation-8 layer does not (64 cells along the last axis)."""
    n = cmi.WarmStartNet(seed=7)
    n.set_normalisation([0.0, 0.0], [1.0, 1.0], 1.0)
    n.set_dilation(dil)
    feat = np.random.default_rng(2).normal(size=shape + (2,)).astype(np.float32)
    n.set_tiled(True)
    z1 = n.predict(feat)
    n.set_tiled(False)
    z0 = n.predict(feat)
    n.close()
    assert np.isfinite(z0).all() and np.abs(z0).max() > 0
    assert np.array_equal(z0, z1)


@pytest.mark.parametrize("dil", [(1, 1, 1, 1), (1, 2, 4, 8), (2, 4, 8, 16)])
def test_gradients_match_torch(net, dil):
    rng = np.random.default_rng(1)
    B = 2
    feat = rng.normal(size=(B, 16, 16, 32, 2)).astype(np.float32)
    tgt = rng.normal(size=(B, 16, 16, 32)).astype(np.float32)
    net.set_dilation(dil)
    loss, g = net.loss_and_grad(feat, tgt)
    net.set_dilation((1, 1, 1, 1))
    zt, ws = torch_forward(split_params(net.params()), feat, dil)
    lt = ((zt - torch.tensor(tgt)) ** 2).mean()
    lt.backward()
    assert abs(loss / lt.item() - 1) < 1e-2
    gt = join_grads(net.params(), [(w.grad.numpy(), b.grad.numpy()) for w, b in ws])
    off = 0
    for li, (cin, cout, nout) in enumerate(LAYERS):
        n_w = 27 * nout * 32
        for name, sl in (("W", slice(off, off + n_w)),
                         ("b", slice(off + n_w, off + n_w + cout))):
            a, b = g[sl], gt[sl]
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
            rel = float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30))
            print("layer", li, name, "cos", cos, "rel", rel)
            assert cos >= 0.995 and rel <= 5e-2, (li, name, cos, rel)
        off += n_w + nout
    # entries of the master vector beyond the real channels get no gradient
    real = join_grads(net.params(), [(np.ones_like(w.grad.numpy()), np.ones_like(b.grad.numpy()))
                                     for w, b in ws]) != 0
    assert np.all(g[~real] == 0)


def test_training_reduces_loss():
    """A synthetic relation the network can represent (a smoothed nonlinear
    function of the two features): 300 Adam steps cut the loss >= 5x."""
    rng = np.random.default_rng(2)
    S, D = 6, 32
    f = rng.normal(size=(S, D, D, D, 2)).astype(np.float32)
    k = np.ones(3) / 3
    for ax in (1, 2, 3):
        f = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, f)
    f = (f / f.std()).astype(np.float32)
    t = (2.0 * np.tanh(f[..., 0]) - 1.5 * f[..., 1] ** 2).astype(np.float32)
    n = cmi.WarmStartNet(seed=5)
    n.set_normalisation([0.0, 0.0], [1.0, 1.0], 1.0)
    n.add_samples(f, t)
    losses = n.train(300, 2, [32, 32, 32], 2e-3, seed=1)
    n.close()
    first, last = losses[:10].mean(), losses[-10:].mean()
    print("loss", first, "->", last)
    assert np.isfinite(losses).all()
    assert last < first / 5
