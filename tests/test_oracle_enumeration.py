"""Exact enumeration of the four-domain system and the oracle's Monte Carlo.

The enumerator states the grand-canonical ensemble with the reference
enumerator's weights (enumerate.cpp:257-275, 480-520, 639-664):
prod_staples [staple_M / 6^(2L-1)] * exp(-E) * orientation multiplicity.
It is pinned here by closed forms (the scaffold-only ensemble), by an
internal consistency check (the factor-6 orientation shortcut against an
explicit enumeration of every orientation), and it then pins the oracle's
move set: the order-parameter distribution sampled by the oracle's Monte
Carlo (staple exchange, CB staple regrowth, CTCB scaffold regrowth,
orientation rotation) must agree with it.  The Monte Carlo runs are seeded,
so every number in this file is reproducible.
"""
import math

import pytest

from latticednaorigami_amd import systems


def _system(max_staples, seed=20261202):
    s = systems.four_domain_system(seed=seed, max_total_staples=max_staples)
    s.max_type_staples = max(1, max_staples)
    return s


def test_scaffold_only_closed_form(oracle_mod):
    # 216 three-step walks from the origin; a reversal puts domains i and i+2
    # on one site as a misbound pair (identities -1/-3 or -2/-4).  Each
    # misbound pair has E = misbinding_h/T - (misbinding_s + ln 6) = -ln 6 and
    # replaces 6*6 orientations by 6, so every walk weighs 6^4.  Walks with
    # k = 0, 1, 2 reversals: 6*5*5, 2*6*5, 6.
    o = oracle_mod.Oracle(_system(0), temps=[330.0])
    probs, log_z, n_configs = o.enumerate()
    assert probs[(0, 0, 0, 0)] == pytest.approx(150 / 216, rel=1e-12)
    assert probs[(0, 0, 1, 0)] == pytest.approx(60 / 216, rel=1e-12)
    assert probs[(0, 0, 2, 0)] == pytest.approx(6 / 216, rel=1e-12)
    assert n_configs == 150 * 6**4 + 60 * 6**3 + 6 * 6**2
    assert log_z == pytest.approx(math.log(216 * 6**4), rel=1e-12)


def test_orientation_shortcut_is_exact(oracle_mod):
    # unbound domains and misbound pairs carry no orientation-dependent
    # energy; counting them as a factor 6 must equal enumerating them
    o = oracle_mod.Oracle(_system(1), temps=[330.0])
    p1, lz1, n1 = o.enumerate()
    p2, lz2, n2 = o.enumerate(all_orientations=True)
    assert set(p1) == set(p2)
    assert n1 == n2
    assert lz1 == pytest.approx(lz2, rel=1e-13)
    for k in p1:
        assert p1[k] == pytest.approx(p2[k], rel=1e-9, abs=1e-15)


def test_probabilities_are_normalised(oracle_mod):
    o = oracle_mod.Oracle(_system(2), temps=[345.0])
    probs, _, _ = o.enumerate()
    assert sum(probs.values()) == pytest.approx(1.0, rel=1e-12)
    by_ns = {}
    for k, v in probs.items():
        by_ns[k[0]] = by_ns.get(k[0], 0) + v
    # at 345 K zero, one and two staples are all probable
    assert by_ns[0] > 0.3 and by_ns[1] > 0.3 and by_ns[2] > 0.05


def _sample(oracle_mod, temp, n_rep, n_samples, stride, seed):
    s = _system(2, seed=seed)
    o = oracle_mod.Oracle(s, temps=[temp] * n_rep)
    exact, _, _ = o.enumerate()
    o.sweep(20000)
    counts = [dict() for _ in range(n_rep)]
    for _ in range(n_samples):
        o.sweep(stride)
        for r in range(n_rep):
            ob = o.observables(r)
            k = (ob["num_staples"], ob["num_fully_bound_domain_pairs"],
                 ob["num_misbound_domain_pairs"], ob["num_stacked_domain_pairs"])
            counts[r][k] = counts[r].get(k, 0) + 1
    return exact, counts


@pytest.mark.parametrize("temp,seed", [(330.0, 7), (345.0, 11)])
def test_monte_carlo_matches_enumeration(oracle_mod, temp, seed):
    """The oracle's sampled distribution agrees with the exact one.

    Errors come from the spread over independent replicas; a state passes if
    it lies within 4 standard errors (plus 0.005 absolute for the
    autocorrelated short run).  The staple-number marginal must agree too,
    which is what the staple-exchange bookkeeping (fugacity, insertion sites,
    bound-domain counts) decides."""
    n_rep, n_samples = 8, 2500
    exact, counts = _sample(oracle_mod, temp, n_rep, n_samples, 100, seed)

    def check(key_fn, p_exact):
        for key, p in p_exact.items():
            if p < 0.01:
                continue
            fr = [sum(v for k, v in c.items() if key_fn(k) == key) / n_samples for c in counts]
            m = sum(fr) / n_rep
            se = math.sqrt(sum((f - m) ** 2 for f in fr) / (n_rep - 1) / n_rep)
            assert abs(m - p) <= 4 * se + 0.005, (temp, key, m, p, se)

    check(lambda k: k, exact)
    marg = {}
    for k, v in exact.items():
        marg[k[0]] = marg.get(k[0], 0) + v
    check(lambda k: k[0], marg)


def _window(op, lo, hi, well=0.0, min_bias=0.0, slope=0.0, off_grid=0.0, mult=1.0):
    from latticednaorigami_amd import abi
    return abi.Window(op, lo, hi, 0, well, min_bias, slope, off_grid, mult)


def _biased_run(oracle_mod, op, win, grid, misbinding=None, moves=None, n_samples=2500):
    """Seeded biased Monte Carlo of the two-staple four-domain system at 345 K
    against the exact distribution reweighted by exp(-B(op)) state by state.
    Returns (target, per-replica counts, n_samples)."""
    from latticednaorigami_amd import abi
    n_rep, temp = 8, 345.0
    s = _system(2, seed=13)
    if misbinding is not None:
        s.misbinding_pot = misbinding
    if moves is not None:
        kept = [m for m in s.moves if m.kind in moves]
        tot = sum(m.freq for m in kept)
        s.moves = [systems.Move(m.kind, m.freq / tot, m.max_regrowth) for m in kept]
    o = oracle_mod.Oracle(s, temps=[temp] * n_rep)
    exact, _, _ = o.enumerate()
    lo, hi, well, min_bias, slope = win
    w = _window(op, lo, hi, well, min_bias, slope)
    glen = len(grid)
    o.set_windows([w] * n_rep, grid * n_rep, glen)

    def bias(key):
        p = key[0] if op == abi.OP_NUMSTAPLES else key[1]
        r = well if lo <= p <= hi else (slope * ((lo - p - 1) if p < lo else (p - hi - 1)) + min_bias)
        g = grid[p] if 0 <= p < glen else 0.0
        return r + g

    weights = {k: v * math.exp(-bias(k)) for k, v in exact.items()}
    z = sum(weights.values())
    target = {k: v / z for k, v in weights.items()}
    assert max(abs(target[k] - exact[k]) for k in exact) > 0.1  # the bias matters
    o.sweep(20000)
    counts = [dict() for _ in range(n_rep)]
    for _ in range(n_samples):
        o.sweep(100)
        for r in range(n_rep):
            ob = o.observables(r)
            k = (ob["num_staples"], ob["num_fully_bound_domain_pairs"],
                 ob["num_misbound_domain_pairs"], ob["num_stacked_domain_pairs"])
            counts[r][k] = counts[r].get(k, 0) + 1
            assert ob["bias"] == pytest.approx(bias(k), abs=1e-12)
    return target, counts, n_samples


def _mean_se(counts, key, n_samples):
    fr = [c.get(key, 0) / n_samples for c in counts]
    m = sum(fr) / len(fr)
    se = math.sqrt(sum((f - m) ** 2 for f in fr) / (len(fr) - 1) / len(fr))
    return m, se


_MET = 1
_ALL_MOVES = None


@pytest.mark.parametrize("op,win,grid,misbinding,moves", [
    # numstaples: grid biases 0 / +1.5 / -2.0 kT on 0 / 1 / 2 staples
    (1, (0, 2, 0.0, 0.0, 0.0), [0.0, 1.5, -2.0], None, _ALL_MOVES),
    # numfulldomains: a LinearStepWell that favours 2-3 fully bound pairs
    (0, (2, 3, -1.0, 1.0, 1.5), [], None, _ALL_MOVES),
    # numfulldomains: a deep well on exactly one fully bound pair, so that
    # half-bound staples dominate and CB/CTCB regrowth moves the order
    # parameter (their exp(-dB) factor decides acceptance: a sign error in
    # it gives p(1,2,0,1) = 0.99 against 0.16).  Misbinding is disallowed so
    # that every bound staple domain is fully bound (see the next test).
    (0, (1, 1, -6.0, 3.0, 3.0), [], "disallowed", _ALL_MOVES),
    # the same well with Opposing misbinding (staples with one bound and one
    # misbound domain are then common) without CB staple regrowth
    (0, (1, 1, -6.0, 3.0, 3.0), [], None, "no_cb"),
])
def test_biased_monte_carlo_matches_reweighted_enumeration(oracle_mod, op, win, grid,
                                                           misbinding, moves):
    """Move-level biases (add_external_bias in every move type, re-sync after
    rejection) sample exp(-B(op)) times the unbiased ensemble: the exact
    distribution reweighted state by state."""
    from latticednaorigami_amd import abi
    mis = abi.MISBIND_DISALLOWED if misbinding == "disallowed" else None
    kinds = None
    if moves == "no_cb":
        kinds = {abi.MOVE_ORIENTATION_ROTATION, abi.MOVE_MET_STAPLE_EXCHANGE,
                 abi.MOVE_CTCB_SCAFFOLD_REGROWTH}
    target, counts, n = _biased_run(oracle_mod, op, win, grid, mis, kinds)
    for key, p in target.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, n)
        assert abs(m - p) <= 4 * se + 0.005, (key, m, p, se)


def test_cb_staple_regrowth_redraws_the_old_growth_point(oracle_mod):
    """A behaviour of the reference kept on purpose: CB staple regrowth draws
    the growth point of the old-conformation regrowth afresh
    (cb_movetypes.cpp:335 -> movetypes.cpp:385-393) instead of reusing the
    forward one.  For a staple with one fully bound and one misbound domain
    the two Rosenbluth weights differ by ~1e5, so detailed balance fails for
    the move (1,1,1,0) -> (1,2,0,1).  Under a deep well on one fully bound
    pair this over-samples the fully bound staple several fold, which the
    oracle (and the engine, step for step equal to it) must reproduce."""
    from latticednaorigami_amd import abi
    target, counts, n = _biased_run(oracle_mod, 0, (1, 1, -6.0, 3.0, 3.0), [], None,
                                    {abi.MOVE_MET_STAPLE_EXCHANGE, abi.MOVE_CB_STAPLE_REGROWTH,
                                     abi.MOVE_CTCB_SCAFFOLD_REGROWTH}, n_samples=800)
    m, se = _mean_se(counts, (1, 2, 0, 1), n)
    assert target[(1, 2, 0, 1)] < 0.15
    assert m > 3 * target[(1, 2, 0, 1)], (m, se)


def _mc_counts(o, n_rep, n_samples, stride=100, burn=20000):
    o.sweep(burn)
    counts = [dict() for _ in range(n_rep)]
    for _ in range(n_samples):
        o.sweep(stride)
        for r in range(n_rep):
            ob = o.observables(r)
            k = (ob["num_staples"], ob["num_fully_bound_domain_pairs"],
                 ob["num_misbound_domain_pairs"], ob["num_stacked_domain_pairs"])
            counts[r][k] = counts[r].get(k, 0) + 1
    return counts


def test_three_quarter_turn_monte_carlo_matches_enumeration(oracle_mod):
    """ThreeQuarterTurn twist, kink and junction rules (domain.cpp:69-118):
    the four-domain system with 3/4-turn domains at 345 K, where unbound,
    misbound, one- and two-staple states are all populated."""
    from latticednaorigami_amd import abi
    s = _system(2, seed=17)
    s.domain_type = abi.THREEQUARTERTURN
    o = oracle_mod.Oracle(s, temps=[345.0] * 8)
    exact, _, _ = o.enumerate()
    assert sum(1 for p in exact.values() if p > 0.02) >= 5
    counts = _mc_counts(o, 8, 2500)
    for key, p in exact.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, 2500)
        assert abs(m - p) <= 4 * se + 0.005, (key, m, p, se)


@pytest.mark.parametrize("domain_type", [0, 1])
def test_three_domain_staples_match_enumeration_with_scaffold_endpoints(oracle_mod, domain_type):
    """Staples of three domains (Met insertion/deletion with sites-before /
    sites-minus-L counting, CB regrowth from either end or the middle, CTCB
    through a staple network bound across the segment boundary) sample the
    enumerated ensemble at 349 K, where the bare scaffold, a misbound scaffold
    and the fully bound staple are all probable, once CTCB constrains the
    in-segment scaffold domains of an externally bound network
    (LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS)."""
    from latticednaorigami_amd import abi
    s = systems.three_domain_system(domain_type=domain_type,
                                    flags=abi.FLAG_CTCB_SCAFFOLD_ENDPOINTS)
    o = oracle_mod.Oracle(s, temps=[349.0] * 8)
    exact, _, _ = o.enumerate()
    assert exact[(1, 3, 0, 2)] == pytest.approx(0.567, abs=0.01)
    counts = _mc_counts(o, 8, 2500)
    for key, p in exact.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, 2500)
        assert abs(m - p) <= 4 * se + 0.005, (key, m, p, se)


def test_reference_ctcb_rule_unbinds_three_domain_staples(oracle_mod):
    """The reference keys the endpoint of an in-segment (staple domain,
    scaffold domain) pair of an externally bound network by the staple's
    chain (top_constraint_points.cpp:553-557), so CTCB regrowth of a segment
    that ends inside a fully bound three-domain staple moves the scaffold
    domain off the staple with both Rosenbluth weights blind to the lost
    pair: the fully bound state (exact probability 0.567) is almost never
    seen, and the scaffold spends its time unbound."""
    s = systems.three_domain_system()
    o = oracle_mod.Oracle(s, temps=[349.0] * 8)
    exact, _, _ = o.enumerate()
    counts = _mc_counts(o, 8, 1000)
    m_full, _ = _mean_se(counts, (1, 3, 0, 2), 1000)
    m_bare, _ = _mean_se(counts, (0, 0, 0, 0), 1000)
    assert exact[(1, 3, 0, 2)] > 0.5 and m_full < 0.05
    assert exact[(0, 0, 0, 0)] < 0.4 and m_bare > 0.7


def _scaffold_only(n_domains, misbinding_h=0.0):
    """A bare HalfTurn scaffold of n_domains on a line (no staples allowed):
    its conformations are the self-avoiding-or-misbound walks the
    enumerator counts, so only the scaffold moves shape the distribution."""
    s = systems.four_domain_system(seed=7, max_total_staples=0)
    s.identities = [[-(i + 1) for i in range(n_domains)], [1, 2]]
    s.chains = [systems.Chain(0, 0, [(i, 0, 0) for i in range(n_domains)],
                              [(0, 0, 1)] * n_domains)]
    s.max_type_staples = 0
    s.misbinding_h = misbinding_h
    return s


@pytest.mark.parametrize("misbinding_h,attempts,max_regrowth,n_samples", [
    # the reference's limits (one recoil, all 36 trial configurations);
    # inverting the acceptance ratio's weights (a planted mistake) gives
    # p(0,0,2,0) = 0.110 against 0.132, seven standard errors out
    (0.0, 36, 6, 1500),
    # a misbound pair costs +1.24 kT, so open probabilities fall below 1;
    # dropping them from the weights (planted) gives 0.979 / 0.021 for no /
    # one misbound pair against 0.962 / 0.037, twelve standard errors out
    (1000.0, 6, 6, 1500),
    # two trial configurations per domain: dead ends are common and the
    # feeler growth (test_config_avail) decides which trials count.  Counting
    # every open trial instead (planted) moves p(0,0,0,0) by 0.010, about
    # three standard errors: this case exercises the feelers but pins their
    # count only loosely.
    (0.0, 2, 4, 8000),
])
def test_ctrg_scaffold_regrowth_matches_enumeration(oracle_mod, misbinding_h, attempts,
                                                     max_regrowth, n_samples):
    """CTRG recoil regrowth (rg_movetypes.cpp) with orientation rotations
    samples the enumerated conformations of a bare six-domain scaffold at
    330 K: the Rosenbluth-like weights (attempts left, open probabilities,
    feelers) and the Metropolis ratio of new to old weight."""
    from latticednaorigami_amd import abi
    s = _scaffold_only(6, misbinding_h)
    s.moves = [systems.Move(abi.MOVE_ORIENTATION_ROTATION, 0.5),
               systems.Move(abi.MOVE_CTRG_SCAFFOLD_REGROWTH, 0.5, max_regrowth, 1, attempts)]
    o = oracle_mod.Oracle(s, temps=[330.0] * 8)
    exact, _, _ = o.enumerate()
    counts = _mc_counts(o, 8, n_samples, stride=20, burn=5000)
    assert o.counters().accepts[1] > 0.3 * o.counters().attempts[1]
    for key, p in exact.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, n_samples)
        assert abs(m - p) <= 4 * se + 0.002, (key, m, p, se)


@pytest.mark.parametrize("kind", ["ctrg_jump", "ctcb_jump"])
def test_jump_matches_enumeration_with_hairpin_staples(oracle_mod, kind):
    """Non-contiguous CTRG and CTCB (segment pairs grown from the stems that
    the crossover staples of a four-domain hairpin offer) with CB staple
    regrowth and orientation rotations sample the enumerated ensemble
    restricted to two bound staples at 345 K.  Without staple exchange the
    staple count stays 2, so the target is the exact distribution
    conditioned on it."""
    from latticednaorigami_amd import abi
    move = {"ctrg_jump": abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH,
            "ctcb_jump": abi.MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH}[kind]
    s = systems.hairpin_system(4, seed=7)
    s.moves = [systems.Move(abi.MOVE_ORIENTATION_ROTATION, 0.2),
               systems.Move(abi.MOVE_MET_STAPLE_EXCHANGE, 0.6),
               systems.Move(abi.MOVE_CB_STAPLE_REGROWTH, 0.2)]
    grow = oracle_mod.Oracle(s, temps=[345.0])
    while grow.observables(0)["num_staples"] != 2:
        grow.sweep(1000)
    s.chains = grow.config(0)
    s.moves = [systems.Move(abi.MOVE_ORIENTATION_ROTATION, 0.3),
               systems.Move(abi.MOVE_CB_STAPLE_REGROWTH, 0.2),
               systems.Move(move, 0.5, 4, 1, 36, 2)]
    o = oracle_mod.Oracle(s, temps=[345.0] * 8)
    exact, _, _ = o.enumerate()
    two = {k: v for k, v in exact.items() if k[0] == 2}
    z = sum(two.values())
    counts = _mc_counts(o, 8, 1000, stride=20, burn=2000)
    assert o.counters().accepts[2] > 0.3 * o.counters().attempts[2]
    for key, p in two.items():
        if p / z < 0.01:
            continue
        m, se = _mean_se(counts, key, 1000)
        assert abs(m - p / z) <= 4 * se + 0.01, (key, m, p / z, se)


@pytest.mark.parametrize("kind", ["ctcb_jump", "ctrg_jump"])
def test_jump_scaffold_regrowth_matches_enumeration(oracle_mod, kind):
    """The non-contiguous selections (select_noncontig_segs) grown by CTCB
    and by CTRG sample the enumerated conformations of the bare six-domain
    scaffold at 330 K (no staples, so every selection is one segment)."""
    from latticednaorigami_amd import abi
    k = {"ctcb_jump": abi.MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH,
         "ctrg_jump": abi.MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH}[kind]
    s = _scaffold_only(6)
    s.moves = [systems.Move(abi.MOVE_ORIENTATION_ROTATION, 0.5),
               systems.Move(k, 0.5, 6, 1, 36, 2)]
    o = oracle_mod.Oracle(s, temps=[330.0] * 8)
    exact, _, _ = o.enumerate()
    counts = _mc_counts(o, 8, 3000, stride=20, burn=5000)
    assert o.counters().accepts[1] > 0.3 * o.counters().attempts[1]
    for key, p in exact.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, 3000)
        assert abs(m - p) <= 4 * se + 0.002, (key, m, p, se)


def _learned_system():
    """A staple [1, 5] whose domain 5 has no complement (misbinding
    disallowed): bound through domain 1, its end domain 5 always dangles, so
    its orientation carries no energy and is uniform in equilibrium."""
    from latticednaorigami_amd import abi
    s = systems.four_domain_system(seed=13, max_total_staples=2)
    s.max_type_staples = 2
    s.identities = [[-1, -2, -3, -4], [1, 5], [3, 4]]
    s.misbinding_pot = abi.MISBIND_DISALLOWED
    return s


def _learned_system_three_domain():
    """The same with a three-domain staple [1, 2, 5]: bound through domains
    1 and 2, its end domain 5 dangles from the middle domain."""
    from latticednaorigami_amd import abi
    s = _learned_system()
    s.identities = [[-1, -2, -3, -4], [1, 2, 5], [3, 4]]
    s.max_staple_size = 3
    s.flags = abi.FLAG_CTCB_SCAFFOLD_ENDPOINTS
    return s


def _dangling_orientations(oracle_mod, network, n_samples=500, system=None):
    """Per-replica frequencies of the six orientations of unbound staple
    domains, sampled between rounds of 20 learned attempts and 5 plain steps
    at 300 K."""
    import numpy as np
    import parity_cases as pc
    o = oracle_mod.Oracle(system or _learned_system(), temps=[300.0] * 8)
    o.learned_set_network(5, pc.random_network(5, network), 0.0)
    o.sweep(5000)
    units = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    hist = np.zeros((8, 6))
    for _ in range(n_samples):
        o.learned_sweep(20)
        o.sweep(5)
        for r in range(8):
            cfg = o.config(r)
            occ = {}
            for ch in cfg:
                for p in ch.positions:
                    occ[p] = occ.get(p, 0) + 1
            for ch in cfg[1:]:
                for p, q in zip(ch.positions, ch.orientations):
                    if occ[p] == 1:
                        hist[r, units.index(q)] += 1
    return hist / hist.sum(axis=1, keepdims=True), o.counters()


def test_learned_move_keeps_unbound_orientations_uniform(oracle_mod):
    """The learned proposal (a seeded random network whose output layer is
    scaled so that it prefers some orientations several-fold) moves the
    dangling domain far more often than the plain moves do.  Its
    Metropolis-Hastings factor q(old)/q(new) must cancel that preference: the
    sampled orientation of unbound staple domains stays uniform (1/6 each,
    within 4 standard errors over 8 replicas).  Without the factor the
    frequencies are 0.12-0.24, 2.8 times the tolerance."""
    import numpy as np
    fr, c = _dangling_orientations(oracle_mod, 5)
    assert c.accepts[7] > 1000
    m = fr.mean(axis=0)
    se = fr.std(axis=0, ddof=1) / np.sqrt(len(fr))
    assert np.all(np.abs(m - 1 / 6) <= 4 * se + 0.005), (m, se)


def test_learned_move_keeps_three_domain_dangling_orientations_uniform(oracle_mod):
    """The uniform-orientation check for the end domain of a three-domain
    staple, which the learned move places next to the staple's middle
    domain.  Without the q(old)/q(new) factor (a planted mistake) the
    frequencies are 0.10-0.29, well outside the tolerance; the enumeration
    test of the three-domain system does not see that mistake."""
    import numpy as np
    fr, c = _dangling_orientations(oracle_mod, 5, system=_learned_system_three_domain())
    assert c.accepts[7] > 1000
    m = fr.mean(axis=0)
    se = fr.std(axis=0, ddof=1) / np.sqrt(len(fr))
    assert np.all(np.abs(m - 1 / 6) <= 4 * se + 0.005), (m, se)


def test_learned_move_matches_enumeration(oracle_mod):
    """Learned attempts (with the CB fallback) interleaved with the plain
    moves sample the enumerated distribution of the four-domain system."""
    import parity_cases as pc
    s = _system(2, seed=13)
    o = oracle_mod.Oracle(s, temps=[345.0] * 8)
    exact, _, _ = o.enumerate()
    o.learned_set_network(4, pc.random_network(4, 5), 0.1)
    o.sweep(20000)
    counts = [dict() for _ in range(8)]
    n = 600
    for _ in range(n):
        o.learned_sweep(10)
        o.sweep(100)
        for r in range(8):
            ob = o.observables(r)
            k = (ob["num_staples"], ob["num_fully_bound_domain_pairs"],
                 ob["num_misbound_domain_pairs"], ob["num_stacked_domain_pairs"])
            counts[r][k] = counts[r].get(k, 0) + 1
    for key, p in exact.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, n)
        assert abs(m - p) <= 4 * se + 0.005, (key, m, p, se)


def test_learned_move_matches_enumeration_three_domain_staples(oracle_mod):
    """Learned end-domain regrowth of a three-domain staple (the anchor is
    its middle domain; the staple stays connected through the other end)
    with the CB fallback, interleaved with the plain moves, samples the
    enumerated ensemble at 349 K (the learned exchange rejects staple types
    of other than two domains in both directions, which keeps balance)."""
    from latticednaorigami_amd import abi
    import parity_cases as pc
    s = systems.three_domain_system(flags=abi.FLAG_CTCB_SCAFFOLD_ENDPOINTS)
    n_ident = len(s.identities[0])
    o = oracle_mod.Oracle(s, temps=[349.0] * 8)
    exact, _, _ = o.enumerate()
    o.learned_set_network(n_ident, pc.random_network(n_ident, 5), 0.1)
    o.sweep(20000)
    counts = [dict() for _ in range(8)]
    n = 1500
    for _ in range(n):
        o.learned_sweep(10)
        o.sweep(100)
        for r in range(8):
            ob = o.observables(r)
            k = (ob["num_staples"], ob["num_fully_bound_domain_pairs"],
                 ob["num_misbound_domain_pairs"], ob["num_stacked_domain_pairs"])
            counts[r][k] = counts[r].get(k, 0) + 1
    acc = o.counters().accepts
    att = o.counters().attempts
    slot = 7  # the learned regrowth's counters (the last of LDO_MAX_MOVES = 8)
    assert att[slot] > 1000 and acc[slot] > 100, (att, acc)
    for key, p in exact.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, n)
        assert abs(m - p) <= 4 * se + 0.005, (key, m, p, se)


def test_learned_exchange_matches_enumeration(oracle_mod):
    """The learned staple exchange alone changes the staple number (Met is
    taken out of the move set): a staple [1, 5] whose domain 5 has no
    complement (misbinding disallowed), at most one staple, 340 K and
    0.1 mM staples, where the exact split is 0.74 bare / 0.26 with a staple
    and the exchange is accepted about a quarter of the time.  The sampled
    split matches enumeration; halving the insertion's growth-path factor
    (a planted mistake) gives 0.15 (6 times the tolerance)."""
    from latticednaorigami_amd import abi
    import parity_cases as pc
    s = systems.four_domain_system(seed=13, max_total_staples=1)
    s.max_type_staples = 1
    s.identities = [[-1, -2, -3, -4], [1, 5]]
    s.misbinding_pot = abi.MISBIND_DISALLOWED
    s.staple_M = 1e-4
    kept = [m for m in s.moves if m.kind != abi.MOVE_MET_STAPLE_EXCHANGE]
    tot = sum(m.freq for m in kept)
    s.moves = [systems.Move(m.kind, m.freq / tot, m.max_regrowth) for m in kept]
    o = oracle_mod.Oracle(s, temps=[340.0] * 8)
    exact, _, _ = o.enumerate()
    o.learned_set_network(5, pc.random_network(5, 5, 0.1), 0.1)
    for _ in range(50):
        o.learned_sweep(10)
        o.sweep(10)
    n = 1000
    counts = [dict() for _ in range(8)]
    for _ in range(n):
        o.learned_sweep(10)
        o.sweep(10)
        for r in range(8):
            ob = o.observables(r)
            k = (ob["num_staples"], ob["num_fully_bound_domain_pairs"],
                 ob["num_misbound_domain_pairs"], ob["num_stacked_domain_pairs"])
            counts[r][k] = counts[r].get(k, 0) + 1
    assert 0.2 < exact[(1, 1, 0, 0)] < 0.35
    assert o.counters().accepts[5] > 5000
    for key, p in exact.items():
        if p < 0.01:
            continue
        m, se = _mean_se(counts, key, n)
        assert abs(m - p) <= 4 * se + 0.005, (key, m, p, se)


def test_four_domain_enumeration_fixture_is_current(oracle_mod):
    """tests/golden/four_domain_enumeration.json (the exact reference of the
    learned proposal's accuracy protocol in bench.py) equals the oracle's
    enumeration."""
    import json
    import os
    import sys
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, golden)
    import make_four_domain_enumeration as mk
    with open(os.path.join(golden, "four_domain_enumeration.json")) as f:
        table = json.load(f)["p_by_temperature"]
    for t in mk.TEMPS:
        assert table[f"{t:g}"] == pytest.approx(mk.exact_numfulldomains(t), rel=1e-12, abs=1e-15)
