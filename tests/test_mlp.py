"""The learned proposal's network: the CPU restatement (oracle/mlp_ref.py)
pinned by finite differences, and the HIP training kernels checked against
it on an MI355X."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_ref as M  # noqa: E402


def _params(rng, n_ident, scale=1.0):
    L = M.layout(n_ident)
    p = np.zeros(L["total"])
    fan = {"w1": M.SITES, "emb": 1, "w1e": M.EMBED, "w2": M.HIDDEN, "w3": M.HIDDEN,
           "w4": M.HIDDEN}
    ends = {"w1": "emb", "emb": "w1e", "w1e": "b1", "w2": "b2", "w3": "b3", "w4": "b4"}
    for name, f in fan.items():
        n = L[ends[name]] - L[name]
        p[L[name]:L[name] + n] = rng.uniform(-1, 1, n) * np.sqrt(6.0 / f) * scale
    for b in ("b1", "b2", "b3"):
        p[L[b]:L[b] + M.HIDDEN] = rng.uniform(-0.1, 0.1, M.HIDDEN)
    p[L["b4"]:L["b4"] + M.OUT] = rng.uniform(-0.1, 0.1, M.OUT)
    return p


def test_reference_gradient_matches_finite_differences():
    """Central differences of the float64 loss (no bf16 rounding) agree with
    the analytic gradient on entries of every tensor, including the scattered
    first-layer rows and the embedding rows the batch uses."""
    rng = np.random.default_rng(5)
    n_ident = 5
    p = _params(rng, n_ident)
    ch, ident, legal, target = M.random_batch(rng, 6, n_ident)
    _, g = M.loss_grad(p, n_ident, ch, ident, legal, target, round_bf16=False)
    L = M.layout(n_ident)
    feats = np.arange(M.SITES) * M.CHANNELS + ch[0].astype(int)
    picks = []
    for name in ("w2", "w3", "w4", "b1", "b2", "b3", "b4", "w1e"):
        n = {"w2": M.HIDDEN ** 2, "w3": M.HIDDEN ** 2, "w4": M.HIDDEN * M.OUT, "b1": M.HIDDEN,
             "b2": M.HIDDEN, "b3": M.HIDDEN, "b4": M.OUT, "w1e": M.EMBED * M.HIDDEN}[name]
        picks += [L[name] + int(i) for i in rng.integers(0, n, 6)]
    picks += [L["w1"] + int(f) * M.HIDDEN + int(j) for f, j in zip(feats[:6], rng.integers(0, 256, 6))]
    picks += [L["emb"] + int(ident[0]) * M.EMBED + k for k in range(4)]
    h = 1e-6
    for i in picks:
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        la, _ = M.loss_grad(a, n_ident, ch, ident, legal, target, round_bf16=False)
        lb, _ = M.loss_grad(b, n_ident, ch, ident, legal, target, round_bf16=False)
        fd = (la - lb) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9), i


def test_reference_log_probs_are_normalised_and_masked():
    rng = np.random.default_rng(6)
    p = _params(rng, 3)
    ch, ident, legal, _ = M.random_batch(rng, 8, 3)
    f = M.forward(p, 3, ch, ident, legal)
    lp = f["logp"]
    assert np.all(np.isneginf(lp[legal == 0]))
    assert np.allclose(np.exp(lp).sum(axis=1), 1.0, rtol=1e-12)


@pytest.mark.gpu
def test_gpu_loss_and_gradient_match_reference():
    """ldo_mlp_loss_grad (bf16 MFMA GEMMs, FP32 accumulation) against the
    restatement with the same bf16 operand rounding: the loss to 1e-5
    relative, every tensor's gradient to 1e-3 of its largest entry."""
    from latticednaorigami_amd.mlp import Mlp
    print("case: MLP loss/grad, 64 random samples, 7 identities", flush=True)
    rng = np.random.default_rng(7)
    n_ident = 7
    p = _params(rng, n_ident).astype(np.float32)
    net = Mlp(n_ident, max_batch=64, params=p)
    ch, ident, legal, target = M.random_batch(rng, 64, n_ident)
    loss, g = net.loss_grad(ch, ident, legal, target, want_grads=True)
    ref_loss, ref_g = M.loss_grad(p.astype(np.float64), n_ident, ch, ident, legal, target)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    L = M.layout(n_ident)
    names = ["w1", "emb", "w1e", "b1", "w2", "b2", "w3", "b3", "w4", "b4", "total"]
    for a, b in zip(names, names[1:]):
        ga, gr = g[L[a]:L[b]], ref_g[L[a]:L[b]]
        tol = 1e-3 * np.abs(gr).max()
        assert np.abs(ga - gr).max() <= tol, (a, np.abs(ga - gr).max(), tol)
    lp = net.log_probs(ch, ident, legal)
    ref_lp = M.forward(p.astype(np.float64), n_ident, ch, ident, legal)["logp"]
    fin = legal == 1
    assert np.allclose(lp[fin], ref_lp[fin], atol=1e-4)
    assert np.all(np.isneginf(lp[~fin]))


@pytest.mark.gpu
def test_gpu_training_learns_a_planted_rule():
    """Adam on ldo_mlp_loss_grad learns a rule the input decides (the target
    is the first legal output at or after the channel of site 171 times 2):
    the mean loss falls from the uniform value to under a tenth of it."""
    from latticednaorigami_amd.mlp import Mlp
    print("case: MLP training on a planted rule, 300 Adam steps of 128", flush=True)
    rng = np.random.default_rng(8)
    net = Mlp(4, max_batch=128, seed=3)

    def batch():
        ch, ident, legal, _ = M.random_batch(rng, 128, 4)
        legal[:] = 1
        target = (ch[:, 171].astype(np.int32) * 2) % M.OUT
        return ch, ident, legal, target
    first = net.loss_grad(*batch())
    for _ in range(300):
        net.loss_grad(*batch())
        net.adam_step(lr=1e-3)
    last = np.mean([net.loss_grad(*batch()) for _ in range(5)])
    assert first > 0.8 * np.log(M.OUT)
    assert last < 0.1 * first, (first, last)
