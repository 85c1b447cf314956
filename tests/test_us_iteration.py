"""The adaptive umbrella-sampling iteration (us_simulation.cpp:154-452):
the oracle's per-step histogram and the CPU restatement of the weight update
(oracle/us_ref.py), pinned by properties the update must have."""
import math
import os
import sys

import numpy as np

from latticednaorigami_amd import systems

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import us_ref  # noqa: E402

import parity_cases as pc  # noqa: E402


def test_histogram_counts_every_step(oracle_mod):
    """The oracle's histogram equals the order parameter read after each of
    3000 single steps (update_internal runs once per step)."""
    s = systems.raster_system(replicas=2)
    wins, grid, _ = pc._raster_windows(2)
    glen = 3
    grid = grid[:glen] + grid[41:41 + glen]
    ref = oracle_mod.Oracle(s, temps=[330.0, 330.0])
    ref.set_windows(wins, grid, glen)
    ref.set_histogram(True)
    want = np.zeros((2, glen + 1), dtype=np.int64)
    for _ in range(3000):
        ref.sweep(1)
        for r in range(2):
            p = ref.observables(r)["num_staples"]
            want[r, p if p < glen else glen] += 1
    assert np.array_equal(ref.histogram(), want)
    assert want[:, glen].sum() > 0 or want[:, 1:glen].sum() > 0  # not all at zero staples


def _draw(pi, E, n):
    """Visits n * (pi e^-E) / Z: the biased distribution, exactly."""
    w = [p * math.exp(-e) for p, e in zip(pi, E)]
    z = sum(w)
    return [n * x / z for x in w]


def test_update_recovers_unbiased_log_probabilities():
    """With visits drawn from the biased distribution pi e^-E, the new
    biases are log pi (the reference's bias is an energy in kT, so the next
    iteration samples pi e^-log(pi) = flat)."""
    pi = [0.05, 0.3, 0.4, 0.2, 0.05]
    E = [0.0, 1.5, -0.7, 2.0, -1.2]
    u = us_ref.WindowState({p: e for p, e in enumerate(E)})
    f = _draw(pi, E, 10000)
    us_ref.process_iteration(u, f, u.grid(5), 10001, 100.0)
    for p in range(5):
        assert abs(u.E_w[p] - math.log(pi[p])) < 1e-12, p
        assert abs(u.p_i[p] - pi[p]) < 1e-15
    flat = _draw(pi, [u.E_w[p] for p in range(5)], 1.0)
    assert max(flat) - min(flat) < 1e-12


def test_update_clamps_changes_to_max_D_bias():
    pi = [0.01, 0.99]
    u = us_ref.WindowState({0: 0.0, 1: -3.0})
    us_ref.process_iteration(u, _draw(pi, [0.0, -3.0], 1000), u.grid(2), 1001, 0.5)
    assert u.E_w[0] == -0.5  # log 0.01 = -4.6: lowered by 0.5 only
    assert u.E_w[1] == -2.5  # log 0.99 = -0.01: raised by 0.5 only
    # a point visited before and not now sinks by max_D_bias
    us_ref.process_iteration(u, [0, 7], u.grid(2), 8, 0.5)
    assert u.E_w[0] == -1.0 and u.p_i[0] == 0 and u.w_i[0] == 0
    assert u.E_w[1] == -2.0  # p = 1: log 1 = 0, raised by 0.5
    assert u.w_i[1] == 7 / 8
    assert sorted(u.S_n) == [0, 1]


def test_visit_fractions_and_output_formats():
    """w = f / (steps + 1) as the reference divides by simulate()'s return;
    the summary is written through an ostream with precision 3 and width 10
    and the weights with std::to_string."""
    u = us_ref.WindowState()
    us_ref.process_iteration(u, [0, 1, 3], [0.0, 0.0, 0.0], 5, 10.0)
    assert sum(u.w_i.values()) == 4 / 5
    assert us_ref.summary(u, 0) == ("Iteration: 0\n\nGridpoint w, P, E:\n"
                                    "1 :        0.2      0.25     -1.39\n"
                                    "2 :        0.6      0.75    -0.288\n\n")
    assert us_ref.weights_json(u) == (
        '{\n    "biases": [\n        {\n            "point": [\n                1\n'
        '            ],\n            "bias": -1.386294\n        },\n        {\n'
        '            "point": [\n                2\n            ],\n'
        '            "bias": -0.287682\n        }\n    ]\n}\n')
    assert us_ref.weights_json(us_ref.WindowState()) == '{\n    "biases": [\n    ]\n}\n'
