// us_loop.hpp - the adaptive umbrella-sampling weight iteration of one
// window, host code between device runs.
//
// USGCMCSimulation::process_iteration (us_simulation.cpp:154-170) with
// fill_grid_sets (341-375), estimate_current_weights (286-304) and
// SimpleUSGCMCSimulation::update_grids / update_bias (385-432), on a
// one-dimensional grid: a point is a value of the window's order parameter.
// The visit counts f_i come from the sampler's per-step histogram
// (ldo_get_histogram), the grid biases go back through ldo_set_window_grid.
#ifndef LDO_US_LOOP_HPP
#define LDO_US_LOOP_HPP

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iomanip>
#include <map>
#include <ostream>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace ldo_us {

struct WindowState {
    std::set<int> S_n;          // every point sampled by an iteration so far
    std::map<int, double> E_w;  // bias estimate of each point (the grid's values)
    std::map<int, double> p_i;  // probability estimate of each point of S_n
    std::map<int, double> w_i;  // last iteration's visit fraction of each point of S_n
};

// The grid the window samples with: E_w where it holds a value, the grid
// bias's off-grid value (0, bias_functions.hpp m_off_grid_bias) elsewhere.
inline std::vector<double> dense_grid(const WindowState& u, int grid_len) {
    std::vector<double> g(grid_len, 0.0);
    for (const auto& kv: u.E_w) {
        if (kv.first < 0 || kv.first >= grid_len) {
            throw std::out_of_range("grid point " + std::to_string(kv.first) + " off the grid");
        }
        g[kv.first] = kv.second;
    }
    return g;
}

// One iteration's update.  f: visits of each point during the iteration
// (grid_len values), grid: the biases the iteration ran with, m_steps: the
// step count the reference divides by (simulate()'s return value, one more
// than the steps run), max_D_bias: the largest change of a point's bias.
inline void process_iteration(WindowState& u, const int64_t* f, int grid_len,
                              const std::vector<double>& grid, long long m_steps,
                              double max_D_bias) {
    // fill_grid_sets: the points visited now (s_i) and those visited only
    // by earlier iterations (old_only)
    std::vector<int> s_i, old_only;
    for (int p = 0; p < grid_len; p++) {
        if (f[p] > 0) {
            s_i.push_back(p);
        }
    }
    std::set<int> now(s_i.begin(), s_i.end());
    for (int p: u.S_n) {
        if (!now.count(p)) {
            old_only.push_back(p);
        }
    }
    u.S_n.insert(s_i.begin(), s_i.end());
    // estimate_current_weights: unbias the visits with the grid they ran under
    double ave_bias_weight = 0;
    for (int p: s_i) {
        ave_bias_weight += f[p] * std::exp(grid[p]);
    }
    for (int p: s_i) {
        u.p_i[p] = f[p] * std::exp(grid[p]) / ave_bias_weight;
    }
    for (int p: old_only) {
        u.p_i[p] = 0;
    }
    // update_grids
    u.w_i.clear();
    for (int p: s_i) {
        u.w_i[p] = static_cast<double>(f[p]) / m_steps;
    }
    for (int p: old_only) {
        u.w_i[p] = 0;
    }
    double total_w = 0;
    for (const auto& kv: u.w_i) {
        total_w += kv.second;
    }
    if (total_w - 1.0 > 1e-6) {
        throw std::logic_error("visit fractions sum to more than one");
    }
    // update_bias: move each point's bias to log p, by at most max_D_bias
    for (int p: u.S_n) {
        double old_bias = u.E_w[p];
        double p_k_n = u.p_i[p];
        double D_bias, new_bias;
        if (p_k_n == 0) {
            D_bias = -max_D_bias;
            new_bias = old_bias + D_bias;
        }
        else {
            new_bias = std::log(p_k_n);
            D_bias = new_bias - old_bias;
        }
        double updated_bias = new_bias;
        if (std::abs(D_bias) > max_D_bias) {
            updated_bias = D_bias > 0 ? old_bias + max_D_bias : old_bias - max_D_bias;
        }
        u.E_w[p] = updated_bias;
    }
}

// SimpleUSGCMCSimulation::output_summary (us_simulation.cpp:434-452)
inline void write_summary(std::ostream& out, const WindowState& u, int n) {
    out << "Iteration: " << n << "\n";
    out << "\n";
    out << "Gridpoint w, P, E:\n";
    for (int p: u.S_n) {
        out << p << " ";
        out << std::setprecision(3);
        out << ": ";
        out << std::setw(10) << u.w_i.at(p);
        out << std::setw(10) << u.p_i.at(p);
        out << std::setw(10) << u.E_w.at(p);
        out << "\n";
    }
    out << "\n";
}

// std::to_string of a double: fixed notation, six decimals
inline std::string fixed6(double v) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

// USGCMCSimulation::output_weights (us_simulation.cpp:206-248): the grid's
// points and biases as JSON.  Points are written in increasing order (the
// reference writes them in its hash map's order).
inline void write_weights(std::ostream& out, const WindowState& u) {
    out << "{\n";
    out << "    \"biases\": [";
    bool first = true;
    for (const auto& kv: u.E_w) {
        out << (first ? "\n" : ",\n");
        first = false;
        out << "        {\n";
        out << "            \"point\": [\n";
        out << "                " << kv.first << "\n";
        out << "            ],\n";
        out << "            \"bias\": " << fixed6(kv.second) << "\n";
        out << "        }";
    }
    out << "\n    ]\n";
    out << "}\n";
}

} // namespace ldo_us

#endif
