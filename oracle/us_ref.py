"""CPU restatement of the adaptive umbrella-sampling weight iteration.

TEST INFRASTRUCTURE ONLY: the checker of the host driver's MWUS loop
(latticednaorigami_amd/csrc/driver/us_loop.hpp); nothing in the product
imports it.  It follows USGCMCSimulation::process_iteration
(us_simulation.cpp:154-170): fill_grid_sets (341-375),
estimate_current_weights (286-304), SimpleUSGCMCSimulation::update_grids
(418-432) and update_bias (385-416), the summary of output_summary (434-452)
and the weights file of output_weights (206-248), on one-dimensional grid
points.  math.exp / math.log are the C library's, as in the reference.
Pinned in tests/test_us_iteration.py by what the update must satisfy: a
histogram drawn exactly from the biased distribution yields log of the
unbiased probabilities (and a flat biased distribution next), the max_D_bias
clamp in both directions, and points left unvisited sinking by max_D_bias.
"""
import math


class WindowState:
    def __init__(self, E_w=None):
        self.S_n = set()
        self.E_w = dict(E_w or {})
        self.p_i = {}
        self.w_i = {}

    def grid(self, grid_len):
        """Dense grid: E_w where set, the off-grid bias 0 elsewhere."""
        g = [0.0] * grid_len
        for p, v in self.E_w.items():
            g[p] = v
        return g


def process_iteration(u, f, grid, m_steps, max_D_bias):
    """f: visits per point this iteration, grid: the biases it ran with,
    m_steps: simulate()'s return (one past the last step)."""
    s_i = [p for p in range(len(f)) if f[p] > 0]
    old_only = sorted(u.S_n - set(s_i))
    u.S_n |= set(s_i)
    ave = 0.0
    for p in s_i:
        ave += f[p] * math.exp(grid[p])
    for p in s_i:
        u.p_i[p] = f[p] * math.exp(grid[p]) / ave
    for p in old_only:
        u.p_i[p] = 0.0
    u.w_i = {}
    for p in s_i:
        u.w_i[p] = f[p] / m_steps
    for p in old_only:
        u.w_i[p] = 0.0
    for p in sorted(u.S_n):
        old = u.E_w.get(p, 0.0)
        pk = u.p_i[p]
        if pk == 0:
            d = -max_D_bias
            new = old + d
        else:
            new = math.log(pk)
            d = new - old
        upd = new
        if abs(d) > max_D_bias:
            upd = old + max_D_bias if d > 0 else old - max_D_bias
        u.E_w[p] = upd


def _g3(x):
    """A double through an ostream with precision 3 and width 10."""
    return "%10s" % ("%.3g" % x)


def summary(u, n):
    out = f"Iteration: {n}\n\nGridpoint w, P, E:\n"
    for p in sorted(u.S_n):
        out += f"{p} : {_g3(u.w_i[p])}{_g3(u.p_i[p])}{_g3(u.E_w[p])}\n"
    return out + "\n"


def weights_json(u):
    out = '{\n    "biases": ['
    first = True
    for p in sorted(u.E_w):
        out += "\n" if first else ",\n"
        first = False
        out += ("        {\n            \"point\": [\n                %d\n            ],\n"
                "            \"bias\": %f\n        }") % (p, u.E_w[p])
    return out + "\n    ]\n}\n"
