"""CPU restatement of the learned proposal's network (include/ldo_mlp.h).

TEST INFRASTRUCTURE ONLY: the checker of the HIP training kernels
(latticednaorigami_amd/csrc/mlp.hip); nothing in the product imports it.
The reference has no learned proposal, so this follows the network as the
header defines it.  Arithmetic is float64; with bf16=True the operands of the
three hidden GEMMs and of their backward products are first rounded to bf16
(round to nearest even), as the MFMA tiles round them, so the device result
differs only by FP32 accumulation.  It is pinned by central finite
differences of its own loss (bf16=False), which a wrong transpose, a dropped
ReLU mask or a missing bias gradient fails.
"""
import numpy as np

SITES, CHANNELS, FEATURES, EMBED, HIDDEN, OUT = 343, 13, 4459, 16, 256, 36


def layout(n_ident):
    """Element offsets of the flat parameter vector (ldo_mlp_layout)."""
    sizes = [("w1", FEATURES * HIDDEN), ("emb", n_ident * EMBED), ("w1e", EMBED * HIDDEN),
             ("b1", HIDDEN), ("w2", HIDDEN * HIDDEN), ("b2", HIDDEN), ("w3", HIDDEN * HIDDEN),
             ("b3", HIDDEN), ("w4", HIDDEN * OUT), ("b4", OUT)]
    off, o = {}, 0
    for name, n in sizes:
        off[name] = o
        o += n
    off["total"] = o
    return off


def unpack(params, n_ident):
    L = layout(n_ident)
    p = np.asarray(params, dtype=np.float64)

    def t(name, shape):
        return p[L[name]:L[name] + int(np.prod(shape))].reshape(shape)
    return {"w1": t("w1", (FEATURES, HIDDEN)), "emb": t("emb", (n_ident, EMBED)),
            "w1e": t("w1e", (EMBED, HIDDEN)), "b1": t("b1", (HIDDEN,)),
            "w2": t("w2", (HIDDEN, HIDDEN)), "b2": t("b2", (HIDDEN,)),
            "w3": t("w3", (HIDDEN, HIDDEN)), "b3": t("b3", (HIDDEN,)),
            "w4": t("w4", (HIDDEN, OUT)), "b4": t("b4", (OUT,))}


def pack(tensors, n_ident):
    L = layout(n_ident)
    out = np.zeros(L["total"], dtype=np.float64)
    for name in ("w1", "emb", "w1e", "b1", "w2", "b2", "w3", "b3", "w4", "b4"):
        a = np.asarray(tensors[name], dtype=np.float64).ravel()
        out[L[name]:L[name] + a.size] = a
    return out


def bf16(x):
    """Round to bf16 (nearest even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def _mm(a, b, round_bf16):
    if round_bf16:
        return bf16(a) @ bf16(b)
    return a @ b


def forward(params, n_ident, channels, ident, legal, round_bf16=True):
    """Returns the intermediates and the masked log-probabilities."""
    P = unpack(params, n_ident)
    ch = np.asarray(channels, dtype=np.int64).reshape(-1, SITES)
    feats = np.arange(SITES)[None, :] * CHANNELS + ch
    z1 = P["b1"][None, :] + P["w1"][feats].sum(axis=1) + P["emb"][ident] @ P["w1e"]
    h1 = np.maximum(z1, 0)
    z2 = _mm(h1, P["w2"], round_bf16) + P["b2"]
    h2 = np.maximum(z2, 0)
    z3 = _mm(h2, P["w3"], round_bf16) + P["b3"]
    h3 = np.maximum(z3, 0)
    z4 = _mm(h3, P["w4"], round_bf16) + P["b4"]
    lg = np.asarray(legal, dtype=bool).reshape(-1, OUT)
    zm = np.where(lg, z4, -np.inf)
    m = zm.max(axis=1, keepdims=True)
    lse = m + np.log(np.exp(zm - m).sum(axis=1, keepdims=True))
    logp = zm - lse
    return {"feats": feats, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "z3": z3, "h3": h3, "z4": z4,
            "logp": logp, "P": P}


def loss_grad(params, n_ident, channels, ident, legal, target, round_bf16=True):
    """Mean cross-entropy and its gradient (flat, parameter layout)."""
    f = forward(params, n_ident, channels, ident, legal, round_bf16)
    P = f["P"]
    B = f["z4"].shape[0]
    tgt = np.asarray(target, dtype=np.int64)
    loss = -f["logp"][np.arange(B), tgt].mean()
    p = np.exp(f["logp"])
    dz4 = p.copy()
    dz4[np.arange(B), tgt] -= 1.0
    dz4 /= B
    g = {}
    g["w4"] = _mm(f["h3"].T, dz4, round_bf16)
    g["b4"] = dz4.sum(axis=0)
    dz3 = _mm(dz4, P["w4"].T, round_bf16) * (f["z3"] > 0)
    g["w3"] = _mm(f["h2"].T, dz3, round_bf16)
    g["b3"] = dz3.sum(axis=0)
    dz2 = _mm(dz3, P["w3"].T, round_bf16) * (f["z2"] > 0)
    g["w2"] = _mm(f["h1"].T, dz2, round_bf16)
    g["b2"] = dz2.sum(axis=0)
    dz1 = _mm(dz2, P["w2"].T, round_bf16) * (f["z1"] > 0)
    g["b1"] = dz1.sum(axis=0)
    gw1 = np.zeros_like(P["w1"])
    np.add.at(gw1, f["feats"].ravel(), np.repeat(dz1, SITES, axis=0))
    g["w1"] = gw1
    emb = P["emb"][ident]
    g["w1e"] = emb.T @ dz1
    gemb = np.zeros_like(P["emb"])
    np.add.at(gemb, np.asarray(ident), dz1 @ P["w1e"].T)
    g["emb"] = gemb
    return loss, pack(g, n_ident)


def random_batch(rng, batch, n_ident):
    """Random channels, identities, legal masks and legal targets."""
    ch = rng.integers(0, CHANNELS, size=(batch, SITES)).astype(np.uint8)
    ident = rng.integers(0, n_ident, size=batch).astype(np.int32)
    legal = (rng.random((batch, OUT)) < 0.6).astype(np.uint8)
    legal[np.arange(batch), rng.integers(0, OUT, size=batch)] = 1
    target = np.array([rng.choice(np.flatnonzero(row)) for row in legal], dtype=np.int32)
    return ch, ident, legal, target
