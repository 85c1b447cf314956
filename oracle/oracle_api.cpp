// oracle_api.cpp - extern "C" entry points of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../include/ldo_mlp.h"
#include "ldo_oracle.h"
#include "nn_thermo.hpp"
#include "oracle_core.hpp"

namespace orc {

struct Oracle {
    ldo_system_desc sys {};
    std::vector<std::vector<int>> idents;
    std::vector<std::vector<std::string>> seqs;
    std::vector<std::unique_ptr<Replica>> reps;
    std::vector<double> ladder_t, ladder_u, ladder_b, ladder_s;
    int dim2 {1}; // columns (stacking multipliers) of a 2D grid
    // each replica's m_replica_control_qs: temperature, staple chemical
    // potential multiplier, bias multiplier, stacking multiplier
    std::vector<std::array<double, 4>> ctrl;
    std::vector<int64_t> pair_att, pair_acc; // [2R]: direction d, lower slot q
    std::vector<int> q_to_rep;
    int rng_kind {0};
    uint64_t seed {0};
    std::vector<float> net; // learned proposal parameters
    Rng exchange_rng;
    int64_t swap_counter {0}; // exchange rounds so far (the exchange stream's counter)
    int64_t swap_base {0};    // swap_counter when the pairing index last restarted
    ldo_counters counters {};
    Limits lim {0, 0};
    int exchange_mode {0};
    // umbrella windows, one per ladder slot, and their stored bias values
    std::vector<WindowBias> windows;
    std::vector<WindowState> win_state;
    int grid_len {0};
    // per-window visit counts [n_windows * (grid_len + 1)], when counted
    bool hist_on {false};
    std::vector<int64_t> hist;
    std::string err;
};

// Points every replica at the window of its current ladder slot.
static void assign_windows(Oracle* O) {
    if (O->windows.empty()) {
        return;
    }
    for (size_t q = 0; q < O->q_to_rep.size(); q++) {
        Replica& rep = *O->reps[O->q_to_rep[q]];
        rep.win = &O->windows[q];
        rep.win_state = &O->win_state[q];
        rep.hist = O->hist_on ? O->hist.data() + q * (O->grid_len + 1) : nullptr;
    }
}

static std::vector<ChainSpec> specs_from(const ldo_config* cfg,
                                         const std::vector<std::vector<int>>& idents) {
    std::vector<ChainSpec> out;
    int k = 0;
    for (int c = 0; c < cfg->n_chains; c++) {
        ChainSpec s;
        s.index = cfg->chain_index[c];
        s.identity = cfg->chain_identity[c];
        int len = (int)idents.at(s.identity).size();
        for (int d = 0; d < len; d++, k++) {
            s.pos.push_back({cfg->positions[3 * k], cfg->positions[3 * k + 1],
                             cfg->positions[3 * k + 2]});
            s.ore.push_back({cfg->orientations[3 * k], cfg->orientations[3 * k + 1],
                             cfg->orientations[3 * k + 2]});
        }
        out.push_back(s);
    }
    return out;
}

} // namespace orc

using namespace orc;

static std::string g_create_error;

extern "C" {

void* oracle_create(const ldo_system_desc* sd, const ldo_run_params* run,
                    const ldo_config* init, int32_t n_replicas, const double* temps,
                    int32_t rng_kind) {
    try {
        auto o = new Oracle;
        o->sys = *sd;
        int k = 0;
        for (int c = 0; c < sd->n_chain_types; c++) {
            std::vector<int> ids;
            std::vector<std::string> sq;
            for (int d = 0; d < sd->chain_lengths[c]; d++, k++) {
                ids.push_back(sd->identities[k]);
                sq.push_back(sd->sequences ? std::string(sd->sequences[k]) : std::string());
            }
            o->idents.push_back(ids);
            o->seqs.push_back(sq);
        }
        std::vector<MovesetEntry> moves;
        for (int m = 0; m < run->n_movetypes; m++) {
            MovesetEntry e {run->movetype_kind[m], run->movetype_freq[m],
                            run->movetype_max_regrowth ? run->movetype_max_regrowth[m] : 0};
            if (e.kind == LDO_MOVE_CTRG_SCAFFOLD_REGROWTH ||
                e.kind == LDO_MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH) {
                // simulation.cpp:493-495; the trial configurations of a
                // domain are the 36 pairs of utility::vectors
                if (!run->movetype_max_num_recoils || !run->movetype_max_c_attempts) {
                    throw OrigamiError(LDO_E_ARG, "CTRG needs max_num_recoils and max_c_attempts");
                }
                e.max_num_recoils = run->movetype_max_num_recoils[m];
                e.max_c_attempts = run->movetype_max_c_attempts[m];
                if (e.max_regrowth < 2 || e.max_num_recoils < 0 || e.max_c_attempts < 1 ||
                    e.max_c_attempts > 36) {
                    throw OrigamiError(LDO_E_ARG, "CTRG parameters out of range");
                }
                if (e.kind == LDO_MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH) {
                    if (!run->movetype_max_seg_regrowth || run->movetype_max_seg_regrowth[m] < 1) {
                        throw OrigamiError(LDO_E_ARG, "CTRG jump needs max_seg_regrowth >= 1");
                    }
                    e.max_seg_regrowth = run->movetype_max_seg_regrowth[m];
                }
            }
            if (e.kind == LDO_MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH) {
                // setup_scaffold_regrowth_movetype (simulation.cpp:403-438)
                if (!run->movetype_max_seg_regrowth || run->movetype_max_seg_regrowth[m] < 1 ||
                    e.max_regrowth < 2) {
                    throw OrigamiError(LDO_E_ARG, "CTCB jump parameters out of range");
                }
                e.max_seg_regrowth = run->movetype_max_seg_regrowth[m];
            }
            moves.push_back(e);
        }
        if (run->flags & ~(LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS | LDO_FLAG_WINDOW_TEMPERATURE_SWAPS)) {
            throw OrigamiError(LDO_E_ARG, "unknown ldo_run_params.flags bits");
        }
        if ((run->flags & LDO_FLAG_WINDOW_TEMPERATURE_SWAPS) &&
            run->exchange_mode != LDO_EXCHANGE_WINDOWS) {
            throw OrigamiError(LDO_E_ARG, "window temperature swaps need LDO_EXCHANGE_WINDOWS");
        }
        Limits lim {run->max_total_staples, run->max_type_staples, run->flags};
        o->lim = lim;
        o->exchange_mode = run->exchange_mode;
        std::vector<ChainSpec> cfg = specs_from(init, o->idents);
        o->rng_kind = rng_kind;
        o->seed = run->seed;
        for (int r = 0; r < n_replicas; r++) {
            o->reps.emplace_back(new Replica(*sd, o->idents, o->seqs, cfg, temps[r], moves, lim,
                                             run->centering_freq, run->centering_domain,
                                             run->constraint_check_freq, rng_kind, run->seed,
                                             (uint32_t)r));
            o->ladder_t.push_back(temps[r]);
            o->ladder_u.push_back(1.0);
            o->ladder_b.push_back(1.0);
            o->ladder_s.push_back(1.0);
            o->q_to_rep.push_back(r);
            o->ctrl.push_back({temps[r], 1.0, 0.0, 1.0});
        }
        o->pair_att.assign(2 * (size_t)n_replicas, 0);
        o->pair_acc.assign(2 * (size_t)n_replicas, 0);
        o->exchange_rng.init(0, run->seed, 0xFFFFFFFFu);
        return o;
    } catch (std::exception& e) {
        g_create_error = e.what();
        return nullptr;
    }
}

void oracle_destroy(void* o) { delete static_cast<Oracle*>(o); }

const char* oracle_last_error(void* o) {
    if (o == nullptr) {
        return g_create_error.c_str();
    }
    return static_cast<Oracle*>(o)->err.c_str();
}

#define ORACLE_TRY(o, ...)                                                                         \
    Oracle* O = static_cast<Oracle*>(o);                                                           \
    try {                                                                                          \
        __VA_ARGS__;                                                                               \
    } catch (OrigamiError & e) {                                                                   \
        O->err = e.what();                                                                         \
        return e.code;                                                                             \
    } catch (std::exception & e) {                                                                 \
        O->err = e.what();                                                                         \
        return LDO_E_MISUSE;                                                                       \
    }                                                                                              \
    return LDO_OK;

// OneDPTGCMCSimulation::initialize_control_qs (ptmc_simulation.cpp:337-358)
// as a replica at slot q sees it: the bias multiplier lands in the
// chemical-potential entry (m_bias_i == m_staple_u_mult_i,
// ptmc_simulation.hpp:74, 80) and the bias-multiplier entry stays 0.
// TwoDPTGCMCSimulation::initialize_control_qs (486-492) does the same with
// its bias multipliers of 1.
static void init_control_qs(Oracle* O) {
    for (size_t q = 0; q < O->q_to_rep.size(); q++) {
        int r = O->q_to_rep[q];
        O->ctrl[r] = {O->ladder_t[q], O->ladder_b[r], 0.0, O->ladder_s[q]};
    }
}

// control quantity i of ladder slot q (m_control_qs[i][q])
static double control_q(const Oracle* O, int i, int q) {
    switch (i) {
    case 0: return O->ladder_t[q];
    case 1: return O->ladder_u[q];
    case 2: return O->ladder_b[q];
    default: return O->ladder_s[q];
    }
}

// m_exchange_q_is of each driver (ptmc_simulation.cpp:446-449, 602-650)
static std::vector<int> exchange_q_is(int mode) {
    switch (mode) {
    case LDO_EXCHANGE_T: return {0};
    case LDO_EXCHANGE_ST: return {0, 3};
    case LDO_EXCHANGE_HUT: return {0, 1, 2};
    case LDO_EXCHANGE_2D: return {0, 1, 3};
    default: return {0, 1}; // UT
    }
}

int oracle_set_ladder(void* o, int32_t n, const double* t, const double* u, const double* b,
                      const double* s) {
    ORACLE_TRY(o, {
        if (n != (int)O->reps.size()) {
            throw OrigamiError(LDO_E_ARG, "ladder size must equal the replica count");
        }
        O->ladder_t.assign(t, t + n);
        O->ladder_u.assign(u, u + n);
        O->ladder_b.assign(b, b + n);
        O->ladder_s.assign(s, s + n);
        O->dim2 = 1;
        init_control_qs(O);
    })
}

// TwoDPTGCMCSimulation::initialize_control_qs (ptmc_simulation.cpp:452-493):
// slot i * n_s + j holds temps[i] and stacking_mults[j]
int oracle_set_ladder_2d(void* o, int32_t n_t, const double* temps, int32_t n_s,
                         const double* stacking_mults, double staple_u_mult) {
    ORACLE_TRY(o, {
        if (O->exchange_mode != LDO_EXCHANGE_2D) {
            throw OrigamiError(LDO_E_ARG, "a 2D ladder needs exchange_mode LDO_EXCHANGE_2D");
        }
        if (!temps || !stacking_mults || n_t < 1 || n_s < 1 ||
            (int64_t)n_t * n_s != (int64_t)O->reps.size()) {
            throw OrigamiError(LDO_E_ARG, "the 2D grid must hold one slot per replica");
        }
        O->ladder_t.clear();
        O->ladder_u.clear();
        O->ladder_b.clear();
        O->ladder_s.clear();
        for (int v1 = 0; v1 < n_t; v1++) {
            for (int v2 = 0; v2 < n_s; v2++) {
                O->ladder_t.push_back(temps[v1]);
                O->ladder_u.push_back(staple_u_mult);
                O->ladder_b.push_back(1.0);
                O->ladder_s.push_back(stacking_mults[v2]);
            }
        }
        O->dim2 = n_s;
        init_control_qs(O);
    })
}

int oracle_set_step(void* o, int64_t step) {
    ORACLE_TRY(o, {
        if (step < 0) {
            throw OrigamiError(LDO_E_ARG, "negative step");
        }
        for (auto& r: O->reps) {
            r->step = step;
        }
    })
}

int oracle_get_exchange_counts(void* o, int32_t n, int64_t* att, int64_t* acc) {
    ORACLE_TRY(o, {
        if (n != 2 * (int)O->reps.size() || !att || !acc) {
            throw OrigamiError(LDO_E_ARG, "exchange counts hold 2 values per replica");
        }
        std::copy(O->pair_att.begin(), O->pair_att.end(), att);
        std::copy(O->pair_acc.begin(), O->pair_acc.end(), acc);
    })
}

// element offsets of the flat parameter vector, in the order ldo_mlp.h
// lists the tensors
static ldo_mlp_layout mlp_layout_of(int n_ident) {
    const int64_t H = LDO_MLP_HIDDEN;
    const int64_t sizes[10] = {(int64_t)LDO_MLP_FEATURES * H, (int64_t)n_ident * LDO_MLP_EMBED,
                               (int64_t)LDO_MLP_EMBED * H, H, H * H, H, H * H, H,
                               H * LDO_MLP_OUT, LDO_MLP_OUT};
    int64_t* f[10];
    ldo_mlp_layout L;
    f[0] = &L.w1;
    f[1] = &L.emb;
    f[2] = &L.w1e;
    f[3] = &L.b1;
    f[4] = &L.w2;
    f[5] = &L.b2;
    f[6] = &L.w3;
    f[7] = &L.b3;
    f[8] = &L.w4;
    f[9] = &L.b4;
    int64_t o = 0;
    for (int k = 0; k < 10; k++) {
        *f[k] = o;
        o += sizes[k];
    }
    L.total = o;
    return L;
}

int oracle_learned_set_network(void* o, int32_t n_ident, const float* params, double fallback) {
    ORACLE_TRY(o, {
        if (!params || n_ident < 1 || !(fallback >= 0.0 && fallback <= 1.0) ||
            O->rng_kind == 1) {
            throw OrigamiError(LDO_E_ARG, "bad learned-move arguments");
        }
        ldo_mlp_layout L = mlp_layout_of(n_ident);
        O->net.assign(params, params + L.total);
        int64_t offs[10] = {L.w1, L.emb, L.w1e, L.b1, L.w2, L.b2, L.w3, L.b3, L.w4, L.b4};
        for (auto& r: O->reps) {
            if ((int)r->sys.chains[0].size() > n_ident) {
                throw OrigamiError(LDO_E_ARG, "n_ident must cover every scaffold domain");
            }
            r->net = O->net.data();
            for (int k = 0; k < 10; k++) {
                r->net_off[k] = offs[k];
            }
            r->net_n_ident = n_ident;
            r->learned_fallback = fallback;
            r->attempts.resize(LDO_MAX_MOVETYPES, 0);
            r->accepts.resize(LDO_MAX_MOVETYPES, 0);
        }
    })
}

int oracle_learned_sweep(void* o, int64_t attempts) {
    ORACLE_TRY(o, {
        for (auto& r: O->reps) {
            if (!r->net) {
                throw OrigamiError(LDO_E_MISUSE, "no network");
            }
            for (int64_t i = 0; i < attempts; i++) {
                r->learned_attempt();
            }
        }
    })
}

int oracle_learned_samples(void* o, int32_t cap, uint8_t* channels, int32_t* ident,
                           uint8_t* legal, int32_t* target, int32_t* n_out) {
    ORACLE_TRY(o, {
        int32_t k = 0;
        for (auto& r: O->reps) {
            for (int w = 1; w <= r->sys.num_staples; w++) {
                const std::vector<Dom*>& ch = r->sys.chains[w];
                int len = (int)ch.size();
                for (int end = 0; end < 2 && k < cap; end++) {
                    Dom* d = end == 0 ? ch[0] : ch[len - 1];
                    Dom* a = end == 0 ? ch[1] : ch[len - 2];
                    uint8_t lg[6];
                    target[k] = r->learned_context(d, a, channels + (size_t)k * 343, lg);
                    for (int q = 0; q < 36; q++) {
                        legal[(size_t)k * 36 + q] = lg[q / 6];
                    }
                    ident[k] = std::abs(d->d_ident) - 1;
                    k++;
                }
            }
        }
        *n_out = k;
    })
}

int oracle_set_permutation(void* o, const int32_t* q_to_rep) {
    ORACLE_TRY(o, {
        int R = (int)O->reps.size();
        std::vector<int> seen(R, 0);
        for (int q = 0; q < R; q++) {
            int r = q_to_rep[q];
            if (r < 0 || r >= R || seen[r]++) {
                throw OrigamiError(LDO_E_ARG, "not a permutation of the replicas");
            }
        }
        O->q_to_rep.assign(q_to_rep, q_to_rep + R);
        // the replicas run at their new slots' temperature and stacking
        // multiplier; their chemical-potential multipliers change at the
        // next exchange
        for (int q = 0; q < R; q++) {
            O->ctrl[q_to_rep[q]][0] = O->ladder_t[q];
            O->ctrl[q_to_rep[q]][3] = O->ladder_s[q];
        }
    })
}

int oracle_sweep(void* o, int64_t steps) {
    ORACLE_TRY(o, {
        for (auto& r: O->reps) {
            r->simulate(steps);
        }
    })
}

// PTGCMCSimulation::run / OneDPTGCMCSimulation::attempt_exchange for a
// ut_parallel_tempering run (ptmc_simulation.cpp:106-161, 255-313, 360-412,
// 658-663), with every rank's replica held in this process.
// PTMWUSGCMCSimulation::run_swaps + attempt_exchange
// (us_simulation.cpp:662-706, 770-864) with every window on this process:
// adjacent windows swap configurations when each window's order parameter
// lies inside the other window, with the grid biases of both windows.  A
// window keeps its stored bias values across a swap, as each rank of the
// reference keeps its own SystemBiases.
// The exchange step of PTMWUSGCMCSimulation::run_swaps (us_simulation.cpp:
// 672-697, 770-864): window swaps within each temperature's ladder, then,
// with LDO_FLAG_WINDOW_TEMPERATURE_SWAPS, each window between neighbouring
// ladders.  The replicas' dependent quantities must be current.
static void windows_exchange(Oracle* O, int64_t swap_i, int64_t rng_ctr) {
    int W = (int)O->reps.size();
    bool t_swaps = (O->lim.flags & LDO_FLAG_WINDOW_TEMPERATURE_SWAPS) != 0;
    Rng& xr = (O->rng_kind == 1) ? O->reps[0]->rng : O->exchange_rng;
    if (O->rng_kind != 1) {
        xr.begin_step(rng_ctr);
    }
    // each run of consecutive slots at one temperature is its own window
    // ladder (one reference PTMWUS run), paired from its first slot
    std::vector<int> pairs;
    for (int g0 = 0, g1; g0 < W; g0 = g1) {
        for (g1 = g0 + 1; g1 < W && O->ladder_t[g1] == O->ladder_t[g0]; g1++) {
        }
        for (int i = g0 + (int)(swap_i % 2); i < g1 - 1; i += 2) {
            pairs.push_back(i);
        }
    }
    for (int i: pairs) {
        O->counters.exchange_attempts++;
        O->pair_att[i]++;
        const WindowBias& w1 = O->windows[i];
        const WindowBias& w2 = O->windows[i + 1];
        int r1 = O->q_to_rep[i], r2 = O->q_to_rep[i + 1];
        int p1 = O->reps[r1]->window_op(0);
        int p2 = O->reps[r2]->window_op(0);
        if (p2 < w1.min_op || p2 > w1.max_op || p1 < w2.min_op || p1 > w2.max_op) {
            continue;
        }
        bool accepted = true;
        if (p1 != p2) {
            // GridFloats lookups with operator[]: points without a grid
            // value read 0
            auto g = [](const WindowBias& w, int p) {
                return (p >= 0 && p < (int)w.grid.size()) ? w.grid[p] : 0.0;
            };
            double d1 = g(w1, p1) - g(w2, p1);
            double d2 = g(w2, p2) - g(w1, p2);
            double p = std::min(1.0, std::exp(d1 + d2));
            accepted = (p == 1) ? true : (p > xr.uniform_real());
        }
        if (accepted) {
            O->counters.exchange_accepts++;
            O->pair_acc[i]++;
            O->q_to_rep[i] = r2;
            O->q_to_rep[i + 1] = r1;
        }
    }
    if (t_swaps) {
        // window w of ladders g and g + 1 (g = swap_i mod 2, + 2, ...):
        // the PTMC acceptance of the two slots plus the windows' grid
        // biases at both order parameters
        int L = 1;
        while (L < W && O->ladder_t[L] == O->ladder_t[0]) {
            L++;
        }
        for (int g = (int)(swap_i % 2); (g + 2) * L <= W; g += 2) {
            for (int w = 0; w < L; w++) {
                int a = g * L + w, b = (g + 1) * L + w;
                O->counters.exchange_attempts++;
                O->pair_att[(size_t)W + a]++;
                int r1 = O->q_to_rep[a], r2 = O->q_to_rep[b];
                System& s1 = O->reps[r1]->sys;
                System& s2 = O->reps[r2]->sys;
                double t1 = O->ladder_t[a], t2 = O->ladder_t[b];
                double sm1 = O->ladder_s[a], sm2 = O->ladder_s[b];
                std::vector<int> n1 = s1.staple_counts(), n2 = s2.staple_counts();
                double dbu_dn = 0;
                for (size_t k = 0; k < n1.size(); k++) {
                    dbu_dn += (s2.staple_us[k] / t2 - s1.staple_us[k] / t1) *
                              ((double)n2[k] - (double)n1[k]);
                }
                double DB = 1 / t2 - 1 / t1;
                double DH = s2.hyb_enthalpy * t2 - s1.hyb_enthalpy * t1;
                double Dst = s2.stacking_energy * t2 - s1.stacking_energy * t1;
                double DBM = sm2 / t2 - sm1 / t1;
                const WindowBias& wa = O->windows[a];
                const WindowBias& wb = O->windows[b];
                // each replica's order parameter as its new slot's window reads it
                auto op_of = [](const System& sy, const WindowBias& x) {
                    return x.op == LDO_OP_NUMSTAPLES ? sy.num_staples
                                                     : sy.num_fully_bound_pairs;
                };
                int p1 = op_of(s1, wa);
                int p2 = op_of(s2, wb);
                auto g0 = [](const WindowBias& x, int p) {
                    return (p >= 0 && p < (int)x.grid.size()) ? x.grid[p] : 0.0;
                };
                double dg = (g0(wa, p1) - g0(wb, p1)) + (g0(wb, p2) - g0(wa, p2));
                double p = std::min(1.0, std::exp(DB * DH + DBM * Dst - dbu_dn + wa.mult * dg));
                bool accepted = (p == 1) ? true : (p > xr.uniform_real());
                if (accepted) {
                    O->counters.exchange_accepts++;
                    O->pair_acc[(size_t)W + a]++;
                    O->q_to_rep[a] = r2;
                    O->q_to_rep[b] = r1;
                }
            }
        }
    }
    assign_windows(O);
}

// one round of PTMWUSGCMCSimulation::run_swaps: every replica's interval
// (a replica moved to another temperature by the previous round first
// switches to it, update_temp recomputing its energy), then the exchange
static void windows_round(Oracle* O, int64_t interval, int64_t swap_i, int64_t rng_ctr) {
    int W = (int)O->reps.size();
    bool t_swaps = (O->lim.flags & LDO_FLAG_WINDOW_TEMPERATURE_SWAPS) != 0;
    for (int q = 0; q < W; q++) {
        Replica& rep = *O->reps[O->q_to_rep[q]];
        if (rep.sys.temp != O->ladder_t[q]) {
            rep.sys.update_temp(O->ladder_t[q], O->ladder_s[q]);
            rep.sys.update_staple_us(O->ladder_t[q], 1.0);
        }
    }
    for (auto& rep: O->reps) {
        rep->simulate(interval);
        if (t_swaps) {
            rep->sys.update_enthalpy_and_entropy();
        }
    }
    windows_exchange(O, swap_i, rng_ctr);
}

static void run_windows(Oracle* O, int64_t n_swaps, int64_t interval) {
    for (int64_t s = 0; s < n_swaps; s++) {
        int64_t rng_ctr = ++O->swap_counter;
        windows_round(O, interval, rng_ctr - O->swap_base, rng_ctr);
    }
}

int oracle_set_windows(void* o, int32_t n, const ldo_window* windows, const double* grid,
                       int32_t grid_len) {
    ORACLE_TRY(o, {
        if (n != (int)O->reps.size() || grid_len < 0) {
            throw OrigamiError(LDO_E_ARG, "one window per replica is required");
        }
        if (O->lim.flags & LDO_FLAG_WINDOW_TEMPERATURE_SWAPS) {
            // ladders of L consecutive slots at one temperature, neighbours at
            // different ones, each holding the first ladder's windows
            int L = 1;
            while (L < n && O->ladder_t[L] == O->ladder_t[0]) {
                L++;
            }
            bool ok = n % L == 0;
            for (int q = 0; ok && q < n; q++) {
                const ldo_window& a = windows[q % L];
                const ldo_window& b = windows[q];
                ok = O->ladder_t[q] == O->ladder_t[q - q % L] &&
                     !(q % L == 0 && q > 0 && O->ladder_t[q] == O->ladder_t[q - 1]) &&
                     a.op == b.op && a.min_op == b.min_op && a.max_op == b.max_op &&
                     a.well_bias == b.well_bias && a.min_bias == b.min_bias &&
                     a.slope == b.slope && a.off_grid_bias == b.off_grid_bias &&
                     a.bias_mult == b.bias_mult;
            }
            if (!ok) {
                throw OrigamiError(LDO_E_ARG, "window temperature swaps need equal window ladders");
            }
        }
        O->windows.assign(n, WindowBias {});
        O->win_state.assign(n, WindowState {});
        for (int w = 0; w < n; w++) {
            WindowBias& b = O->windows[w];
            b.op = windows[w].op;
            b.min_op = windows[w].min_op;
            b.max_op = windows[w].max_op;
            b.well_bias = windows[w].well_bias;
            b.min_bias = windows[w].min_bias;
            b.slope = windows[w].slope;
            b.off_grid = windows[w].off_grid_bias;
            b.mult = windows[w].bias_mult;
            b.grid.assign(grid + (size_t)w * grid_len, grid + (size_t)(w + 1) * grid_len);
        }
        O->grid_len = grid_len;
        if (O->hist_on) {
            O->hist.assign((size_t)n * (grid_len + 1), 0);
        }
        assign_windows(O);
        // LinearStepWellBiasFunction stores its value when constructed
        // (bias_functions.cpp:107-122); a GridBiasFunction starts from 0
        // (222-224)
        for (int w = 0; w < n; w++) {
            Replica& rep = *O->reps[O->q_to_rep[w]];
            O->win_state[w].restraint = O->windows[w].restraint(rep.window_op(0));
            O->win_state[w].grid = 0;
        }
    })
}

// GridBiasFunction::replace_biases (bias_functions.cpp:236-239): the grid
// changes, the stored values of the windows' bias functions do not.
int oracle_set_window_grid(void* o, const double* grid) {
    ORACLE_TRY(o, {
        if (O->windows.empty()) {
            throw OrigamiError(LDO_E_ARG, "no umbrella windows set");
        }
        for (size_t w = 0; w < O->windows.size(); w++) {
            O->windows[w].grid.assign(grid + w * O->grid_len, grid + (w + 1) * O->grid_len);
        }
    })
}

int oracle_set_histogram(void* o, int32_t on) {
    ORACLE_TRY(o, {
        if (on && O->windows.empty()) {
            throw OrigamiError(LDO_E_ARG, "set windows first");
        }
        O->hist_on = on != 0;
        O->hist.assign(on ? O->windows.size() * (O->grid_len + 1) : 0, 0);
        assign_windows(O);
        if (!on) {
            for (auto& r: O->reps) {
                r->hist = nullptr;
            }
        }
    })
}

int oracle_get_histogram(void* o, int64_t* counts, int32_t reset) {
    ORACLE_TRY(o, {
        if (!O->hist_on) {
            throw OrigamiError(LDO_E_ARG, "histogram not enabled");
        }
        std::copy(O->hist.begin(), O->hist.end(), counts);
        if (reset) {
            std::fill(O->hist.begin(), O->hist.end(), 0);
        }
    })
}

// PTGCMCSimulation::run (ptmc_simulation.cpp:106-150) for the PT drivers:
// update_control_qs, simulate(exchange_interval), update_dependent_qs, then
// the exchange of OneDPTGCMCSimulation (360-412) or TwoDPTGCMCSimulation
// (495-560) and master_send (212-226).
// The exchange step of the PT drivers: OneDPTGCMCSimulation (ptmc_simulation
// .cpp:360-412) or TwoDPTGCMCSimulation (495-560) attempt_exchange, then
// master_send (212-226).  The replicas' dependent quantities must be current.
static void pt_exchange(Oracle* O, int64_t swap_i, int64_t rng_ctr) {
    int R = (int)O->reps.size();
    int mode = O->exchange_mode;
    std::vector<int> q_is = exchange_q_is(mode);
    Rng& xr = (O->rng_kind == 1) ? O->reps[0]->rng : O->exchange_rng;
    if (O->rng_kind != 1) {
        xr.begin_step(rng_ctr);
    }
    // calc_acceptance_p (275-313) and test_acceptance (255-271) for
    // slots (q1, q2); d: direction of the pair for the counts
    auto attempt = [&](int q1, int q2, int d) {
        O->counters.exchange_attempts++;
        O->pair_att[(size_t)d * R + q1]++;
        int r1 = O->q_to_rep[q1];
        int r2 = O->q_to_rep[q2];
        double t1 = control_q(O, 0, q1), t2 = control_q(O, 0, q2);
        double sm1 = control_q(O, 3, q1), sm2 = control_q(O, 3, q2);
        std::vector<int> n1 = O->reps[r1]->sys.staple_counts();
        std::vector<int> n2 = O->reps[r2]->sys.staple_counts();
        double dbu_dn = 0;
        for (size_t k = 0; k < n1.size(); k++) {
            double u1 = O->reps[r1]->sys.staple_us[k];
            double u2 = O->reps[r2]->sys.staple_us[k];
            dbu_dn += (u2 / t2 - u1 / t1) * ((double)n2[k] - (double)n1[k]);
        }
        double DB = 1 / t2 - 1 / t1;
        double DH = O->reps[r2]->sys.hyb_enthalpy * t2 - O->reps[r1]->sys.hyb_enthalpy * t1;
        double Dst = O->reps[r2]->sys.stacking_energy * t2 - O->reps[r1]->sys.stacking_energy * t1;
        double DBM = sm2 / t2 - sm1 / t1;
        double DBias = 0.0 * t2 - 0.0 * t1;
        double p = std::min(1.0, std::exp(DB * (DH + DBias) + DBM * Dst - dbu_dn));
        bool accept = (p == 1) ? true : (p > xr.uniform_real());
        if (accept) {
            O->counters.exchange_accepts++;
            O->pair_acc[(size_t)d * R + q1]++;
            O->q_to_rep[q1] = r2;
            O->q_to_rep[q2] = r1;
        }
    };
    if (mode == LDO_EXCHANGE_2D) {
        // m_i_starts .. m_rep_incrs (ptmc_simulation.hpp:158-164)
        int v1 = R / O->dim2, v2 = O->dim2;
        const int i_starts[4] = {0, 0, 1, 0};
        const int j_starts[4] = {0, 0, 0, 1};
        const int i_incrs[4] = {2, 1, 2, 1};
        const int j_incrs[4] = {1, 2, 1, 2};
        const int i_ends[4] = {v1 - 1, v1, v1 - 1, v1};
        const int j_ends[4] = {v2, v2 - 1, v2, v2 - 1};
        const int rep_incrs[4] = {v2, 1, v2, 1};
        int set = (int)(swap_i % 4);
        int swap_v = (int)(swap_i % 2);
        for (int i = i_starts[set]; i < i_ends[set]; i += i_incrs[set]) {
            for (int j = j_starts[set]; j < j_ends[set]; j += j_incrs[set]) {
                int rep_i = i * v2 + j;
                attempt(rep_i, rep_i + rep_incrs[set], swap_v);
            }
        }
    }
    else {
        for (int i = (int)(swap_i % 2); i < R - 1; i += 2) {
            attempt(i, i + 1, 0);
        }
    }
    // master_send: each replica takes its slot's exchanged quantities
    for (int q = 0; q < R; q++) {
        for (int i: q_is) {
            O->ctrl[O->q_to_rep[q]][i] = control_q(O, i, q);
        }
    }
}

// one round of PTGCMCSimulation::run: update_control_qs (595-600, 651-680:
// T, UT and HUT call update_temp(temp), whose stacking multiplier defaults to
// 1, origami_system.hpp:163; HUT's update_bias_mult does nothing to the
// system, origami_system.hpp:165), simulate, update_dependent_qs, exchange
static void pt_round(Oracle* O, int64_t interval, int64_t swap_i, int64_t rng_ctr) {
    int mode = O->exchange_mode;
    bool stack = mode == LDO_EXCHANGE_ST || mode == LDO_EXCHANGE_2D;
    for (size_t r = 0; r < O->reps.size(); r++) {
        Replica& rep = *O->reps[r];
        const std::array<double, 4>& c = O->ctrl[r];
        rep.sys.update_temp(c[0], stack ? c[3] : 1.0);
        rep.sys.update_staple_us(c[0], c[1]);
        rep.simulate(interval);
        rep.sys.update_enthalpy_and_entropy();
    }
    pt_exchange(O, swap_i, rng_ctr);
}

// PTGCMCSimulation::run (ptmc_simulation.cpp:106-150) for the PT drivers, or
// PTMWUSGCMCSimulation::run_swaps for umbrella windows
int oracle_run(void* o, int64_t n_swaps, int64_t interval) {
    ORACLE_TRY(o, {
        for (int64_t s = 0; s < n_swaps; s++) {
            int64_t rng_ctr = ++O->swap_counter;
            int64_t swap_i = rng_ctr - O->swap_base;
            if (O->exchange_mode == LDO_EXCHANGE_WINDOWS) {
                windows_round(O, interval, swap_i, rng_ctr);
            }
            else {
                pt_round(O, interval, swap_i, rng_ctr);
            }
        }
    })
}

// one exchange step on the current configurations (ldo_exchange)
int oracle_exchange(void* o, int64_t swap_index) {
    ORACLE_TRY(o, {
        if (swap_index < 0) {
            throw OrigamiError(LDO_E_ARG, "negative swap index");
        }
        // update_control_qs and update_dependent_qs: an interval of 0 steps
        int64_t rng_ctr = ++O->swap_counter;
        if (O->exchange_mode == LDO_EXCHANGE_WINDOWS) {
            windows_round(O, 0, swap_index, rng_ctr);
        }
        else {
            pt_round(O, 0, swap_index, rng_ctr);
        }
    })
}

// site evaluation (ldo_site_eval) on a system rebuilt from the replica's
// configuration at its temperature, so the replica is untouched: the domain
// is taken off its site if placed, then checked at the query's site
int oracle_site_eval(void* o, const ldo_site_query* q, int32_t n, ldo_site_result* res) {
    ORACLE_TRY(o, {
        for (int32_t k = 0; k < n; k++) {
            if (q[k].replica < 0 || q[k].replica >= (int)O->reps.size()) {
                throw OrigamiError(LDO_E_ARG, "replica out of range");
            }
            Replica& rep = *O->reps[q[k].replica];
            System tmp(O->sys, O->idents, O->seqs, rep.sys.snapshot(), rep.sys.temp);
            Dom* d = nullptr;
            for (size_t w = 0; w < tmp.chains.size(); w++) {
                if (tmp.chain_indices[w] == q[k].chain_index && q[k].domain >= 0 &&
                    q[k].domain < (int)tmp.chains[w].size()) {
                    d = tmp.chains[w][q[k].domain];
                }
            }
            if (d == nullptr) {
                throw OrigamiError(LDO_E_ARG, "no such chain or domain in the replica");
            }
            ldo_site_result r {};
            if (d->st != Occ::unassigned) {
                r.unassign_delta = tmp.unassign_domain(*d);
            }
            tmp.violated = false;
            r.delta = tmp.check_domain_constraints(*d, V3 {q[k].pos[0], q[k].pos[1], q[k].pos[2]},
                                                   V3 {q[k].ore[0], q[k].ore[1], q[k].ore[2]});
            r.violated = tmp.violated ? 1 : 0;
            res[k] = r;
        }
    })
}

int oracle_get_observables(void* o, int32_t r, ldo_obs* out) {
    ORACLE_TRY(o, {
        Replica& rep = *O->reps.at(r);
        rep.sys.update_enthalpy_and_entropy();
        out->energy = rep.sys.energy;
        out->enthalpy = rep.sys.hyb_enthalpy;
        out->entropy = rep.sys.hyb_entropy;
        out->stacking_energy = rep.sys.stacking_energy;
        out->temperature = rep.sys.temp;
        out->step = rep.step;
        out->num_staples = rep.sys.num_staples;
        out->num_domains = rep.sys.num_domains;
        out->num_bound_domain_pairs = rep.sys.num_bound_pairs;
        out->num_fully_bound_domain_pairs = rep.sys.num_fully_bound_pairs;
        out->num_misbound_domain_pairs = rep.sys.num_bound_pairs - rep.sys.num_fully_bound_pairs;
        out->num_stacked_domain_pairs = rep.sys.num_stacked;
        out->num_unassigned_domains = rep.sys.num_unassigned;
        out->status = rep.status;
        out->bias = rep.win ? (rep.win_state->restraint + rep.win_state->grid) * rep.win->mult : 0;
        out->last_move = rep.last_move;
        out->last_accepted = rep.last_accepted;
    })
}

int oracle_get_move_tracking(void* o, int32_t r, int64_t* out, int64_t cap, int64_t* n_out) {
    ORACLE_TRY(o, {
        Replica& rep = *O->reps.at(r);
        int64_t T = (int64_t)rep.sys.identities.size();
        int64_t ns = (int64_t)rep.sys.identities[0].size();
        int64_t n = 6 * T + 4 * (ns + 1) + 4 * ((int64_t)O->lim.max_total_staples + 1) + 2;
        if (cap < n) {
            throw OrigamiError(LDO_E_ARG, "tracking capacity too small");
        }
        for (int64_t k = 0; k < n; k++) {
            out[k] = rep.track.empty() ? 0 : rep.track[k];
        }
        *n_out = n;
    })
}

int oracle_get_counters(void* o, ldo_counters* out) {
    ORACLE_TRY(o, {
        ldo_counters c = O->counters;
        for (auto& r: O->reps) {
            for (size_t m = 0; m < r->attempts.size() && m < LDO_MAX_MOVETYPES; m++) {
                c.attempts[m] += r->attempts[m];
                c.accepts[m] += r->accepts[m];
            }
            c.site_empty += r->sys.ev_site_empty;
            c.site_bind += r->sys.ev_site_bind;
            c.commits += r->sys.ev_commits;
        }
        *out = c;
    })
}

int oracle_get_permutation(void* o, int32_t* q_to_rep) {
    ORACLE_TRY(o, {
        for (size_t i = 0; i < O->q_to_rep.size(); i++) {
            q_to_rep[i] = O->q_to_rep[i];
        }
    })
}

int oracle_get_config(void* o, int32_t r, ldo_config* cfg, int32_t cap_chains,
                      int32_t cap_domains) {
    ORACLE_TRY(o, {
        auto snap = O->reps.at(r)->sys.snapshot();
        if ((int)snap.size() > cap_chains) {
            throw OrigamiError(LDO_E_ARG, "chain capacity too small");
        }
        int k = 0;
        for (size_t c = 0; c < snap.size(); c++) {
            cfg->chain_index[c] = snap[c].index;
            cfg->chain_identity[c] = snap[c].identity;
            cfg->chain_length[c] = (int)snap[c].pos.size();
            for (size_t d = 0; d < snap[c].pos.size(); d++, k++) {
                if (k >= cap_domains) {
                    throw OrigamiError(LDO_E_ARG, "domain capacity too small");
                }
                cfg->positions[3 * k] = snap[c].pos[d].x;
                cfg->positions[3 * k + 1] = snap[c].pos[d].y;
                cfg->positions[3 * k + 2] = snap[c].pos[d].z;
                cfg->orientations[3 * k] = snap[c].ore[d].x;
                cfg->orientations[3 * k + 1] = snap[c].ore[d].y;
                cfg->orientations[3 * k + 2] = snap[c].ore[d].z;
            }
        }
        cfg->n_chains = (int)snap.size();
    })
}

int oracle_set_config(void* o, int32_t r, const ldo_config* cfg) {
    ORACLE_TRY(o, { O->reps.at(r)->sys.set_config(specs_from(cfg, O->idents)); })
}

int oracle_record_tape(void* o, int32_t r, int32_t on) {
    ORACLE_TRY(o, { O->reps.at(r)->rng.set_record(on != 0); })
}

int oracle_restart_swaps(void* o) {
    ORACLE_TRY(o, { O->swap_base = O->swap_counter; })
}

int oracle_set_tape(void* o, int32_t r, const uint64_t* values, int64_t n) {
    ORACLE_TRY(o, { O->reps.at(r)->rng.set_replay(values, n); })
}

int64_t oracle_tape_length(void* o, int32_t r) {
    return (int64_t)static_cast<Oracle*>(o)->reps.at(r)->rng.tape.size();
}

int oracle_get_tape(void* o, int32_t r, uint64_t* out, int64_t cap) {
    ORACLE_TRY(o, {
        auto& t = O->reps.at(r)->rng.tape;
        if ((int64_t)t.size() > cap) {
            throw OrigamiError(LDO_E_ARG, "tape capacity too small");
        }
        std::memcpy(out, t.data(), t.size() * sizeof(uint64_t));
    })
}

int oracle_energy_tables(void* o, double temp, double stacking_mult, int32_t ii, int32_t ij,
                         double* out5) {
    ORACLE_TRY(o, {
        Potential& p = O->reps.at(0)->sys.pot;
        Tables t = p.build(temp, stacking_mult);
        size_t k = (size_t)(ii - p.ident_lo) * p.n_ident + (ij - p.ident_lo);
        out5[0] = t.hyb_e[k];
        out5[1] = t.hyb_h[k];
        out5[2] = t.hyb_s[k];
        out5[3] = t.stack[k];
        out5[4] = t.init_e;
    })
}

int oracle_recompute_energy(void* o, int32_t r, double* out_energy) {
    ORACLE_TRY(o, {
        System& s = O->reps.at(r)->sys;
        s.update_energy();
        *out_energy = s.energy;
    })
}

int oracle_energy_in_order(void* o, int32_t r, const int32_t* order, int32_t n,
                           double* out_energy, int64_t* out_stacked) {
    ORACLE_TRY(o, {
        System& s = O->reps.at(r)->sys;
        std::vector<Dom*> all;
        for (auto& ch: s.chains) {
            for (Dom* d: ch) {
                all.push_back(d);
            }
        }
        std::vector<char> seen(all.size(), 0);
        if (n != (int32_t)all.size()) {
            throw OrigamiError(LDO_E_ARG, "the order must list every domain once");
        }
        for (int32_t k = 0; k < n; k++) {
            if (order[k] < 0 || order[k] >= n || seen[order[k]]) {
                throw OrigamiError(LDO_E_ARG, "the order must list every domain once");
            }
            seen[order[k]] = 1;
        }
        for (Dom* d: all) {
            s.unassign_domain(*d);
        }
        s.energy = (s.chains.size() - 1) * s.pot.init_e();
        bool violated = false;
        for (int32_t k = 0; k < n && !violated; k++) {
            Dom* d = all[order[k]];
            s.set_domain_config(*d, d->pos, d->ore);
            violated = s.violated;
        }
        *out_energy = s.energy;
        *out_stacked = s.num_stacked;
        // leave the replica as the chain-order rebuild leaves it
        for (Dom* d: all) {
            if (d->st != Occ::unassigned) {
                s.unassign_domain(*d);
            }
        }
        s.violated = false;
        s.energy = (s.chains.size() - 1) * s.pot.init_e();
        s.set_all_domains();
        if (violated) {
            throw OrigamiError(LDO_E_CONSTRAINT, "a partial configuration violated a constraint");
        }
    })
}

int oracle_nn_hybridization_thermo(const char* seq, double temp, double cation_M,
                                   double* out_h, double* out_s) {
    try {
        orc_nn::HS hs = orc_nn::reduced_duplex(seq, temp, cation_M);
        *out_h = hs.h;
        *out_s = hs.s;
    } catch (std::exception&) {
        return LDO_E_ARG;
    }
    return LDO_OK;
}

double oracle_nn_reduced_energy(const char* seq, double temp, double cation_M) {
    return orc_nn::reduced_duplex_energy(seq, temp, cation_M);
}

int oracle_nn_longest_contig_complement(const char* a, const char* b, char* out, int32_t cap) {
    std::vector<std::string> v = orc_nn::longest_contig_complements(a, b);
    std::string joined;
    for (size_t i = 0; i < v.size(); i++) {
        if (i) {
            joined += ' ';
        }
        joined += v[i];
    }
    if ((int)joined.size() + 1 > cap) {
        return LDO_E_ARG;
    }
    std::memcpy(out, joined.c_str(), joined.size() + 1);
    return LDO_OK;
}

// IdealRandomWalks::num_walks (ideal_random_walk.cpp:14-73): the number of
// `steps`-step walks on Z^3 between points a displacement (dx,dy,dz) apart,
// summed as multinomials over the counts of negative x and y steps.
double oracle_num_walks(int32_t dx, int32_t dy, int32_t dz, int32_t steps) {
    int d[3] = {std::abs(dx), std::abs(dy), std::abs(dz)};
    std::sort(d, d + 3);
    int DX = d[0], DY = d[1], DZ = d[2];
    int sum = DX + DY + DZ;
    if (sum > steps || (steps - sum) % 2 != 0) {
        return 0.0;
    }
    int n_minus = (steps - DX - DY - DZ) / 2;
    int n_plus = (steps - DX - DY + DZ) / 2;
    auto fact = [](int n) {
        long double f = 1;
        for (int i = 2; i <= n; i++) {
            f *= i;
        }
        return f;
    };
    long double walks = 0;
    for (int yb = 0; yb <= n_minus; yb++) {
        for (int xb = 0; xb <= n_minus - yb; xb++) {
            if (n_plus - xb - yb < 0) {
                continue;
            }
            walks += fact(steps) / (fact(xb) * fact(xb + DX) * fact(yb) * fact(yb + DY) *
                                    fact(n_minus - xb - yb) * fact(n_plus - xb - yb));
        }
    }
    return (double)walks;
}

int oracle_rotate_half(const int32_t* v, const int32_t* axis, int32_t* out) {
    V3 r = rotate_half({v[0], v[1], v[2]}, {axis[0], axis[1], axis[2]});
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    return LDO_OK;
}

int oracle_rotate(const int32_t* v, const int32_t* axis, int32_t turns, int32_t* out) {
    V3 r = rotate_turns({v[0], v[1], v[2]}, {axis[0], axis[1], axis[2]}, turns);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    return LDO_OK;
}

void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox4x32_10(ctr, key, out);
}

int oracle_enumerate(void* o, int32_t flags, double* out_rows, int32_t cap_rows, int32_t* n_rows,
                     double* out_log_z, double* out_n_configs) {
    ORACLE_TRY(o, {
        EnumResult r = enumerate_states(O->sys, O->idents, O->seqs, O->lim, O->ladder_t.at(0),
                                        (flags & 1) != 0);
        if ((int64_t)r.weights.size() > cap_rows) {
            throw OrigamiError(LDO_E_ARG, "row capacity too small for the enumerated states");
        }
        int k = 0;
        for (const auto& kv: r.weights) {
            for (int q = 0; q < 4; q++) {
                out_rows[5 * k + q] = kv.first[q];
            }
            out_rows[5 * k + 4] = (double)(kv.second / r.z);
            k++;
        }
        *n_rows = k;
        *out_log_z = (double)std::log(r.z);
        *out_n_configs = (double)r.n_configs;
    })
}

} // extern "C"
