"""Training and evaluation of the learned proposal (no reference
counterpart): samples come from the plain sampler's own configurations, the
network trains on the device (ldo_mlp_*), and the trained parameters drive
the learned domain-regrowth move (ldo_learned_*).

Effective sample size uses the integrated autocorrelation time with Sokal's
automatic window (the smallest M with M >= c * tau(M), c = 5), computed on
the host from order-parameter series."""
import sys
import time

import numpy as np

from . import abi
from .mlp import Mlp
from .sampler import Sampler


def collect_samples(system, temps, rounds, steps_per_round, lib=None, seed_steps=5000):
    """Run the plain sampler and extract learned-move training samples after
    every round of `steps_per_round` steps.  Returns stacked arrays
    (channels, ident, legal, target) and the number of configurations read."""
    smp = Sampler(system, temps=temps, lib=lib)
    smp.sweep(seed_steps)
    cap = 2 * system.max_total_staples * len(temps)
    parts = []
    for _ in range(rounds):
        smp.sweep(steps_per_round)
        parts.append(smp.learned_samples(cap))
    smp.close()
    ch = np.concatenate([p[0] for p in parts])
    idn = np.concatenate([p[1] for p in parts])
    lg = np.concatenate([p[2] for p in parts])
    tg = np.concatenate([p[3] for p in parts])
    return (ch, idn, lg, tg), rounds * len(temps)


def train(data, n_ident, steps, batch=256, lr=1e-3, seed=1, lib=None, log_every=0):
    """Adam on the masked cross-entropy of the observed placements.  Returns
    (parameters, loss history)."""
    ch, idn, lg, tg = data
    n = len(idn)
    net = Mlp(n_ident, max_batch=batch, seed=seed, lib=lib)
    rng = np.random.default_rng(seed)
    hist = []
    for it in range(steps):
        pick = rng.integers(0, n, size=min(batch, n))
        loss = net.loss_grad(ch[pick], idn[pick], lg[pick], tg[pick])
        net.adam_step(lr=lr)
        hist.append(loss)
        if log_every and (it + 1) % log_every == 0:
            print(f"train step {it + 1}: loss {np.mean(hist[-log_every:]):.4f}", file=sys.stderr,
                  flush=True)
    params = net.params()
    net.close()
    return params, hist


def mean_loss(params, n_ident, data, lib=None, batch=1024):
    """Mean cross-entropy of `params` on `data` (held-out evaluation)."""
    ch, idn, lg, tg = data
    net = Mlp(n_ident, max_batch=batch, params=params, lib=lib)
    tot, cnt = 0.0, 0
    for s in range(0, len(idn), batch):
        e = min(len(idn), s + batch)
        tot += net.loss_grad(ch[s:e], idn[s:e], lg[s:e], tg[s:e]) * (e - s)
        cnt += e - s
    net.close()
    return tot / max(1, cnt)


def integrated_autocorr_time(x, c=5.0):
    """Integrated autocorrelation time of series x with Sokal's window."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    x = x - x.mean()
    var = x.var()
    if n < 4 or var == 0:
        return 0.5
    f = np.fft.rfft(x, n=2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n] / (var * n)
    tau = 0.5
    for m in range(1, n):
        tau += acf[m]
        if m >= c * tau:
            break
    return max(tau, 0.5)


def ess(series, c=5.0):
    """Effective sample size of each row of `series` summed over rows."""
    return float(sum(len(s) / (2 * integrated_autocorr_time(s, c)) for s in series))


def run_series(smp, rounds, plain_steps, learned_attempts, keys, progress=None):
    """Alternate `learned_attempts` learned moves and `plain_steps` plain
    steps per replica; record the observables `keys` of every replica after
    each round.  Returns ({key: [replica][round]}, seconds spent in the
    sampler's moves, reads excluded).  With `progress` (a label) a line goes
    to standard error every 100 rounds."""
    R = smp.n
    out = {k: [[] for _ in range(R)] for k in keys}
    wall = 0.0
    for i in range(rounds):
        if progress and i % 100 == 0:
            print(f"{progress}: round {i} of {rounds}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        if learned_attempts:
            smp.learned_sweep(learned_attempts)
        if plain_steps:
            smp.sweep(plain_steps)
        wall += time.perf_counter() - t0
        for r in range(R):
            ob = smp.observables(r)
            for k in keys:
                out[k][r].append(ob[k])
    return out, wall


def free_energy_gap(a, b, floor=1e-3):
    """max over order-parameter values with P >= floor in both series of
    |F_a - F_b| with F = -ln P shifted to a common minimum."""
    va = np.concatenate([np.asarray(s) for s in a])
    vb = np.concatenate([np.asarray(s) for s in b])
    top = int(max(va.max(), vb.max())) + 1
    pa = np.bincount(va.astype(int), minlength=top) / len(va)
    pb = np.bincount(vb.astype(int), minlength=top) / len(vb)
    ok = (pa >= floor) & (pb >= floor)
    if not ok.any():
        return None
    fa, fb = -np.log(pa[ok]), -np.log(pb[ok])
    fa -= fa.min()
    fb -= fb.min()
    return float(np.abs(fa - fb).max())
