// engine_host.cpp - host build of the HIP sampler's engine, for tests.
//
// TEST INFRASTRUCTURE ONLY.  The kernels of libldo_hip.so run the
// __host__ __device__ functions of latticednaorigami_amd/csrc/replica_ops.h
// one replica per lane or wavefront.  This library compiles the same
// functions for the host (hipcc --cuda-host-only) and calls them replica by
// replica over host memory, with the run preparation of plan.cpp, so the CPU
// test suite can check the engine's own source against the oracle step for
// step.  It exports eh_* functions with the signatures of the ldo_* ABI; it
// is not a fallback of the product, which never loads it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ldo.h"
#include "../../latticednaorigami_amd/csrc/plan.h"
#include "../../latticednaorigami_amd/csrc/replica_ops.h"

namespace {

std::string g_create_error;

struct HostCtx {
    ldo::Plan plan;
    Params P {};
    int R = 0;
    int64_t swap_counter = 0;
    int64_t swap_base = 0;
    std::string err;
    char* blocks = nullptr;
    std::vector<double> tab, init, xdbl;
    std::vector<int32_t> xint; // q_to_rep | rep_slot
    std::vector<int32_t> keys; // table key class of each slot
    int n_classes = 1;
    std::vector<int64_t> xcnt;
    std::vector<uint64_t> tapes_all;
    std::vector<int64_t> tape_meta;
    std::vector<std::vector<uint64_t>> tapes;
    std::vector<ldo_window> windows;
    std::vector<double> grid, win_stored;
    std::vector<int64_t> hist;
    std::vector<float> net;
    ~HostCtx() { std::free(blocks); }
};

// the learned kernel's per-replica loop with the 64 lanes run in turn
// between the phases the kernel separates with barriers
void learned_replica(const Params& P, int r, int64_t attempts) {
    Eng e;
    LearnedMove m;
    e.attach(&P, P.blocks + (size_t)r * P.L.bytes, r);
    e.select_tables();
    for (int64_t it = 0; it < attempts; it++) {
        if (e.S->status != ST_OK) {
            break;
        }
        int g = e.learned_begin(m);
        while (g == 1) {
            for (int ph = 0; ph < LDO_LEARNED_PHASES; ph++) {
                for (int lane = 0; lane < 64; lane++) {
                    learned_phase_lane(e, P, m, ph, lane, 64);
                }
            }
            g = e.learned_continue(m);
        }
    }
}

void set_tables(HostCtx& c) {
    int n = c.R;
    size_t n2 = (size_t)c.plan.spec.n_ident * c.plan.spec.n_ident;
    ldo::slot_tables(c.plan, c.tab, c.init);
    c.P.hyb_e = c.tab.data();
    c.P.hyb_h = c.tab.data() + (size_t)n * n2;
    c.P.hyb_s = c.tab.data() + 2 * (size_t)n * n2;
    c.P.stack = c.tab.data() + 3 * (size_t)n * n2;
    c.P.init_e = c.init.data();
    c.P.init_h = c.init.data() + n;
    c.P.init_s = c.init.data() + 2 * n;
    c.xdbl = ldo::ladder_values(c.plan);
    c.n_classes = ldo::slot_classes(c.plan, c.keys);
    c.P.slot_key = c.keys.data();
    c.P.pt_dim2 = c.plan.dim2;
}

void set_tapes(HostCtx& c) {
    c.tape_meta.assign(2 * c.R, 0);
    c.tapes_all.clear();
    for (int r = 0; r < c.R; r++) {
        c.tape_meta[r] = (int64_t)c.tapes_all.size();
        c.tape_meta[c.R + r] = (int64_t)c.tapes[r].size();
        c.tapes_all.insert(c.tapes_all.end(), c.tapes[r].begin(), c.tapes[r].end());
    }
    if (c.tapes_all.empty()) {
        c.tapes_all.push_back(0);
    }
    c.P.tapes = c.tapes_all.data();
    c.P.tape_off = c.tape_meta.data();
    c.P.tape_len = c.tape_meta.data() + c.R;
}

void check_statuses(const HostCtx& c) {
    for (int r = 0; r < c.R; r++) {
        ldo::check_status(replica_status(c.P, r), r);
    }
}

void init_from(HostCtx& c, const ldo_config* cfg, int only_rep, int mode) {
    std::vector<int32_t> h;
    int nd = 0;
    ldo::pack_config(c.plan, cfg, h, nd);
    int n = cfg->n_chains;
    for (int r = 0; r < c.R; r++) {
        if (only_rep >= 0 && r != only_rep) {
            continue;
        }
        init_replica(c.P, r, h.data(), h.data() + n, h.data() + 2 * n, h.data() + 2 * n + 3 * nd,
                     n, mode, nullptr);
    }
    check_statuses(c);
}

template <typename F>
int guarded(HostCtx* c, F f) {
    std::string& err = c ? c->err : g_create_error;
    try {
        f();
    } catch (ldo::AbiError& e) {
        err = e.msg;
        return e.code;
    } catch (std::exception& e) {
        err = e.what();
        return LDO_E_CONSTRAINT;
    }
    return LDO_OK;
}

} // namespace

extern "C" {

int eh_create(const ldo_system_desc* sys, const ldo_run_params* run, const ldo_config* init,
              int32_t n_replicas, const double* temps, void** out_ctx) {
    HostCtx* c = new HostCtx;
    int rc = guarded(nullptr, [&]() {
        if (!init || !out_ctx) {
            throw ldo::AbiError {LDO_E_ARG, "null argument"};
        }
        ldo::make_plan(sys, run, n_replicas, temps, c->plan);
        c->P = c->plan.P;
        c->R = n_replicas;
        c->P.type_len = c->plan.type_len.data();
        c->P.type_ident = c->plan.type_ident.data();
        set_tables(*c);
        c->xint.assign(2 * (size_t)n_replicas, 0);
        for (int r = 0; r < n_replicas; r++) {
            c->xint[r] = r;
            c->xint[n_replicas + r] = r;
        }
        c->xcnt.assign(2 + 4 * (size_t)n_replicas, 0);
        c->tapes.assign(n_replicas, {});
        set_tapes(*c);
        size_t total = (size_t)c->P.L.bytes * n_replicas;
        c->blocks = (char*)std::aligned_alloc(256, total);
        if (!c->blocks) {
            throw std::runtime_error("out of host memory");
        }
        std::memset(c->blocks, 0, total);
        c->P.blocks = c->blocks;
        init_from(*c, init, -1, 0);
    });
    if (rc != LDO_OK) {
        delete c;
        return rc;
    }
    *out_ctx = c;
    return LDO_OK;
}

void eh_destroy(void* ctx) { delete static_cast<HostCtx*>(ctx); }

const char* eh_last_error(void* ctx) {
    return ctx ? static_cast<HostCtx*>(ctx)->err.c_str() : g_create_error.c_str();
}

int eh_set_config(void* ctx, int32_t replica, const ldo_config* cfg) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R) {
            throw ldo::AbiError {LDO_E_ARG, "replica out of range"};
        }
        init_from(*c, cfg, replica, 1);
    });
}

int eh_get_config(void* ctx, int32_t replica, ldo_config* cfg, int32_t cap_chains,
                  int32_t cap_domains) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R) {
            throw ldo::AbiError {LDO_E_ARG, "replica out of range"};
        }
        ldo::read_config(c->plan, c->blocks + (size_t)replica * c->P.L.bytes, cfg, cap_chains,
                         cap_domains);
    });
}

int eh_set_ladder(void* ctx, int32_t n, const double* temps, const double* u, const double* b,
                  const double* s) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (n != c->R) {
            throw ldo::AbiError {LDO_E_ARG, "ladder size must equal the replica count"};
        }
        c->plan.ladder_t.assign(temps, temps + n);
        c->plan.ladder_u.assign(u, u + n);
        c->plan.ladder_b.assign(b, b + n);
        c->plan.ladder_s.assign(s, s + n);
        c->plan.dim2 = 1;
        set_tables(*c);
    });
}

int eh_set_ladder_2d(void* ctx, int32_t n_temps, const double* temps, int32_t n_stacking,
                     const double* stacking_mults, double staple_u_mult) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        ldo::set_ladder_2d(c->plan, n_temps, temps, n_stacking, stacking_mults, staple_u_mult);
        set_tables(*c);
    });
}

int eh_set_step(void* ctx, int64_t step) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (step < 0) {
            throw ldo::AbiError {LDO_E_ARG, "negative step"};
        }
        for (int r = 0; r < c->R; r++) {
            ((RepScalars*)(c->blocks + (size_t)r * c->P.L.bytes))->step = step;
        }
    });
}

int eh_get_exchange_counts(void* ctx, int32_t n, int64_t* attempts, int64_t* accepts) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (n != 2 * c->R || !attempts || !accepts) {
            throw ldo::AbiError {LDO_E_ARG, "exchange counts hold 2 values per replica"};
        }
        std::memcpy(attempts, c->xcnt.data() + 2, (size_t)n * sizeof(int64_t));
        std::memcpy(accepts, c->xcnt.data() + 2 + n, (size_t)n * sizeof(int64_t));
    });
}

int eh_set_windows(void* ctx, int32_t n, const ldo_window* windows, const double* grid,
                   int32_t grid_len) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        ldo::check_windows(c->plan, n, windows, grid_len);
        std::vector<int32_t> q2r(c->xint.begin(), c->xint.begin() + c->R);
        std::vector<RepScalars> scalars(c->R);
        for (int r = 0; r < c->R; r++) {
            scalars[r] = *(const RepScalars*)(c->blocks + (size_t)r * c->P.L.bytes);
        }
        c->win_stored = ldo::initial_window_state(windows, n, q2r, scalars);
        c->windows.assign(windows, windows + n);
        c->grid.assign(grid, grid + (size_t)n * grid_len);
        if (c->grid.empty()) {
            c->grid.push_back(0);
        }
        c->P.windows = c->windows.data();
        c->P.win_grid = c->grid.data();
        c->P.grid_len = grid_len;
        c->P.win_stored = c->win_stored.data();
        if (c->P.hist) {
            c->hist.assign((size_t)c->R * (grid_len + 1), 0);
            c->P.hist = c->hist.data();
        }
    });
}

int eh_set_window_grid(void* ctx, const double* grid) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (!c->P.windows) {
            throw ldo::AbiError {LDO_E_ARG, "no umbrella windows set"};
        }
        std::copy(grid, grid + (size_t)c->R * c->P.grid_len, c->grid.begin());
    });
}

int eh_set_histogram(void* ctx, int32_t on) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (!on) {
            c->P.hist = nullptr;
            c->hist.clear();
            return;
        }
        if (!c->P.windows) {
            throw ldo::AbiError {LDO_E_ARG, "set windows first"};
        }
        c->hist.assign((size_t)c->R * (c->P.grid_len + 1), 0);
        c->P.hist = c->hist.data();
    });
}

int eh_get_histogram(void* ctx, int64_t* counts, int32_t reset) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (!c->P.hist) {
            throw ldo::AbiError {LDO_E_ARG, "histogram not enabled"};
        }
        std::copy(c->hist.begin(), c->hist.end(), counts);
        if (reset) {
            std::fill(c->hist.begin(), c->hist.end(), 0);
        }
    });
}

int eh_set_tape(void* ctx, int32_t replica, const uint64_t* tape, int64_t n) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R || n < 0) {
            throw ldo::AbiError {LDO_E_ARG, "replica out of range"};
        }
        c->tapes[replica].assign(tape, tape + n);
        set_tapes(*c);
        ((RepScalars*)(c->blocks + (size_t)replica * c->P.L.bytes))->tape_pos = 0;
    });
}

int eh_sweep(void* ctx, int64_t steps) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (steps < 0) {
            throw ldo::AbiError {LDO_E_ARG, "negative step count"};
        }
        for (int r = 0; r < c->R; r++) {
            sweep_replica(c->P, r, steps, nullptr);
        }
        check_statuses(*c);
    });
}

int eh_learned_set_network(void* ctx, int32_t n_ident, const float* params, double fallback) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (!params || n_ident < c->P.L.ns || c->P.n_moves > LDO_MAX_MOVES - 3 ||
            !(fallback >= 0.0 && fallback <= 1.0) || c->P.rng_mode != 0) {
            throw ldo::AbiError {LDO_E_ARG, "bad learned-move arguments"};
        }
        ldo::check_learned_types(c->plan);
        ldo_mlp_layout L;
        ldo::mlp_layout(n_ident, L);
        c->net.assign(params, params + L.total);
        c->P.net = c->net.data();
        int64_t offs[10] = {L.w1, L.emb, L.w1e, L.b1, L.w2, L.b2, L.w3, L.b3, L.w4, L.b4};
        for (int k = 0; k < 10; k++) {
            c->P.net_off[k] = offs[k];
        }
        c->P.net_n_ident = n_ident;
        c->P.learned_slot = LDO_MAX_MOVES - 1;
        c->P.learned_fallback = fallback;
    });
}

int eh_learned_sweep(void* ctx, int64_t attempts) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (attempts < 0 || !c->P.net) {
            throw ldo::AbiError {LDO_E_ARG, "bad learned sweep"};
        }
        for (int r = 0; r < c->R; r++) {
            learned_replica(c->P, r, attempts);
        }
        check_statuses(*c);
    });
}

int eh_learned_samples(void* ctx, int32_t cap, uint8_t* channels, int32_t* ident,
                       uint8_t* legal, int32_t* target, int32_t* n_out) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        ldo::check_learned_types(c->plan);
        int32_t k = 0;
        for (int r = 0; r < c->R; r++) {
            Eng e;
            LearnedMove m;
            e.attach(&c->P, c->P.blocks + (size_t)r * c->P.L.bytes, r);
            for (int w = 1; w <= e.S->num_staples; w++) {
                for (int end = 0; end < 2 && k < cap; end++) {
                    int cs = e.order[w];
                    int len = e.ch[cs].len;
                    int d = end == 0 ? e.slot(cs, 0) : e.slot(cs, len - 1);
                    int a = end == 0 ? e.slot(cs, 1) : e.slot(cs, len - 2);
                    int idv = e.ident_of(d);
                    m.d = d;
                    m.a = a;
                    m.id = (idv < 0 ? -idv : idv) - 1;
                    I3 ap = e.pos(a);
                    m.ax = ap.x;
                    m.ay = ap.y;
                    m.az = ap.z;
                    m.k_old = 6 * unit_index(sub3(e.pos(d), ap)) + e.dom[d].ore;
                    for (int lane = 0; lane < 64; lane++) {
                        learned_context_lane(e, m, lane, 64);
                    }
                    std::memcpy(channels + (size_t)k * 343, m.ch, 343);
                    for (int q = 0; q < 36; q++) {
                        legal[(size_t)k * 36 + q] = m.legal[q / 6];
                    }
                    ident[k] = m.id;
                    target[k] = m.k_old;
                    k++;
                }
            }
        }
        *n_out = k;
    });
}

int eh_run(void* ctx, int64_t n_swaps, int64_t interval) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (n_swaps < 0 || interval < 0) {
            throw ldo::AbiError {LDO_E_ARG, "negative swap count or interval"};
        }
        if (c->n_classes > 64) {
            throw ldo::AbiError {LDO_E_ARG,
                                 "a ladder holds at most 64 (temperature, stacking) pairs"};
        }
        XState X;
        X.q_to_rep = c->xint.data();
        X.rep_slot = c->xint.data() + c->R;
        X.ladder_t = c->xdbl.data();
        X.ladder_u = c->xdbl.data() + c->R;
        X.ladder_s = c->xdbl.data() + 2 * c->R;
        X.reduced_u = c->xdbl.data() + 3 * c->R;
        X.rep_u = c->xdbl.data() + 3 * c->R + c->plan.reduced_u.size();
        X.counts = c->xcnt.data();
        for (int64_t s = 0; s < n_swaps; s++) {
            int64_t rng_ctr = ++c->swap_counter;
            int64_t swap_i = rng_ctr - c->swap_base;
            for (int r = 0; r < c->R; r++) {
                sweep_replica(c->P, r, interval, X.rep_slot);
            }
            if (c->P.exchange_mode == LDO_EXCHANGE_WINDOWS) {
                exchange_windows_round(c->P, X, swap_i, rng_ctr);
            }
            else {
                exchange_round(c->P, X, swap_i, rng_ctr);
            }
        }
        check_statuses(*c);
    });
}

int eh_exchange(void* ctx, int64_t swap_index) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (swap_index < 0) {
            throw ldo::AbiError {LDO_E_ARG, "negative swap index"};
        }
        if (c->n_classes > 64) {
            throw ldo::AbiError {LDO_E_ARG,
                                 "a ladder holds at most 64 (temperature, stacking) pairs"};
        }
        XState X;
        X.q_to_rep = c->xint.data();
        X.rep_slot = c->xint.data() + c->R;
        X.ladder_t = c->xdbl.data();
        X.ladder_u = c->xdbl.data() + c->R;
        X.ladder_s = c->xdbl.data() + 2 * c->R;
        X.reduced_u = c->xdbl.data() + 3 * c->R;
        X.rep_u = c->xdbl.data() + 3 * c->R + c->plan.reduced_u.size();
        X.counts = c->xcnt.data();
        int64_t rng_ctr = ++c->swap_counter;
        // update_control_qs and update_dependent_qs: an interval of 0 steps
        for (int r = 0; r < c->R; r++) {
            sweep_replica(c->P, r, 0, X.rep_slot);
        }
        if (c->P.exchange_mode == LDO_EXCHANGE_WINDOWS) {
            exchange_windows_round(c->P, X, swap_index, rng_ctr);
        }
        else {
            exchange_round(c->P, X, swap_index, rng_ctr);
        }
        check_statuses(*c);
    });
}

int eh_site_eval(void* ctx, const ldo_site_query* q, int32_t n, ldo_site_result* res) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        std::vector<char> copy(c->P.L.bytes);
        for (int32_t k = 0; k < n; k++) {
            int a = std::abs(q[k].ore[0]) + std::abs(q[k].ore[1]) + std::abs(q[k].ore[2]);
            if (q[k].replica < 0 || q[k].replica >= c->R || a != 1) {
                throw ldo::AbiError {LDO_E_ARG, "replica out of range or ore not a unit vector"};
            }
            std::memcpy(copy.data(), c->blocks + (size_t)q[k].replica * c->P.L.bytes,
                        c->P.L.bytes);
            if (!site_eval_on(c->P, copy.data(), q[k], res[k])) {
                throw ldo::AbiError {LDO_E_ARG, "no such chain or domain in the replica"};
            }
        }
    });
}

int eh_restart_swaps(void* ctx) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() { c->swap_base = c->swap_counter; });
}

int eh_get_observables(void* ctx, int32_t replica, ldo_obs* out) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        if (replica < 0 || replica >= c->R) {
            throw ldo::AbiError {LDO_E_ARG, "replica out of range"};
        }
        for (int r = 0; r < c->R; r++) {
            dependent_replica(c->P, r);
        }
        const RepScalars& S = *(const RepScalars*)(c->blocks + (size_t)replica * c->P.L.bytes);
        ldo::fill_observables(c->plan, S, out);
        out->bias = slot_bias(c->P, c->xint[c->R + replica]);
    });
}

int eh_get_move_tracking(void* ctx, int32_t replica, int64_t* out, int64_t cap, int64_t* n_out) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        int64_t n = c->P.L.track_len;
        if (replica < 0 || replica >= c->R || cap < n) {
            throw ldo::AbiError {LDO_E_ARG, "bad tracking arguments"};
        }
        std::memcpy(out, c->blocks + (size_t)replica * c->P.L.bytes + c->P.L.off_track,
                    (size_t)n * sizeof(int64_t));
        *n_out = n;
    });
}

int eh_get_counters(void* ctx, ldo_counters* out) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        std::memset(out, 0, sizeof *out);
        for (int r = 0; r < c->R; r++) {
            ldo::add_counters(*(const RepScalars*)(c->blocks + (size_t)r * c->P.L.bytes), out);
        }
        out->exchange_attempts = c->xcnt[0];
        out->exchange_accepts = c->xcnt[1];
    });
}

int eh_get_permutation(void* ctx, int32_t* q_to_rep) {
    HostCtx* c = static_cast<HostCtx*>(ctx);
    return guarded(c, [&]() {
        std::memcpy(q_to_rep, c->xint.data(), c->R * sizeof(int32_t));
    });
}

} // extern "C"
