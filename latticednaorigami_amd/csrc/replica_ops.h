// replica_ops.h - the per-replica work of the sampler's kernels.
//
// Each function is the body a kernel runs for one replica (or, for the
// exchange, for the whole ladder).  They compile for the device, where the
// kernels in sampler.hip call them one replica per lane or per wavefront,
// and for the host, where a test build drives the same code replica by
// replica against the CPU oracle.
#ifndef LDO_REPLICA_OPS_H
#define LDO_REPLICA_OPS_H

#include "engine.h"
#include "engine_fourbody.h"
#include "engine_moves.h"
#include "engine_system.h"
#include "layout.h"

// Build (mode 0: OrigamiSystem constructor) or replace (mode 1: set_config)
// replica r's configuration.  c_idx/c_ident hold n_chains unique indices and
// identities, c_pos/c_ore three ints per domain in chain order.
DEV void init_replica(const Params& P, int r, const int32_t* c_idx, const int32_t* c_ident,
                      const int32_t* c_pos, const int32_t* c_ore, int n_chains, int mode,
                      const int32_t* rep_slot) {
    Eng e;
    char* blk = P.blocks + (size_t)r * P.L.bytes;
    RepScalars* S = (RepScalars*)blk;
    if (mode == 0) {
        S->temp_slot = rep_slot ? rep_slot[r] : r;
        S->init_slot = S->temp_slot;
        S->visited = 0;
        S->status = ST_OK;
        e.attach(&P, blk, r);
        for (int i = 0; i < P.L.hcap; i++) {
            e.hkey[i] = LDO_HASH_EMPTY;
        }
        for (int c = 0; c < P.L.nchain; c++) {
            e.ch[c].ident = -1;
            e.ch[c].len = 0;
        }
        for (int t = 0; t < P.L.ntypes; t++) {
            e.tcount[t] = 0;
        }
        S->n_free = P.L.smax;
        for (int i = 0; i < P.L.smax; i++) {
            e.freel[i] = (int16_t)(P.L.smax - i);
        }
        // initialize_scaffold (origami_system.cpp:680-693)
        e.add_chain_with(0, c_idx[0]);
        S->num_staples--;
    }
    else {
        e.attach(&P, blk, r);
        // set_config (origami_system.cpp:327-341)
        for (int w = 0; w < S->n_chains; w++) {
            int cs = e.order[w];
            for (int d = 0; d < e.ch[cs].len; d++) {
                e.unassign_domain(e.slot(cs, d));
            }
        }
        for (int w = S->n_chains - 1; w != 0; w--) {
            e.delete_chain(e.order[w]);
        }
    }
    // initialize_staples (origami_system.cpp:659-678)
    int maxidx = c_idx[0];
    for (int c = 1; c < n_chains; c++) {
        e.add_chain_with(c_ident[c], c_idx[c]);
        maxidx = c_idx[c] > maxidx ? c_idx[c] : maxidx;
    }
    S->energy += (double)(S->n_chains - 1) * e.init_e;
    S->current_c_i = maxidx;
    // set_all_domains(config) (origami_system.cpp:588-616)
    int k = 0;
    for (int w = 0; w < S->n_chains; w++) {
        int cs = e.order[w];
        for (int d = 0; d < e.ch[cs].len; d++, k++) {
            int s = e.slot(cs, d);
            e.dom[s].x = c_pos[3 * k];
            e.dom[s].y = c_pos[3 * k + 1];
            e.dom[s].z = c_pos[3 * k + 2];
            I3 o = i3(c_ore[3 * k], c_ore[3 * k + 1], c_ore[3 * k + 2]);
            e.dom[s].ore = (int8_t)unit_index(o);
        }
    }
    e.set_all_domains();
}

// `steps` steps of GCMCSimulation::simulate (simulation.cpp:568-653); with a
// ladder (rep_slot) the replica first takes its slot's control variables
// (update_control_qs: a temperature switch recomputes the energy; windows
// swap only within a ladder of one temperature and change only the bias) and
// ends with its
// dependent quantities.
// One sweep of replica r on a caller-provided engine object and replica block
// (the block in device memory, or a copy of it held closer to the lane).
// par = 1: the 64 lanes of the calling wavefront all make this call (see
// Eng::par).
DEV void sweep_replica_on(Eng& e, const Params& P, int r, char* blk, int64_t steps,
                          const int32_t* rep_slot, int par = 0) {
    e.attach(&P, blk, r, par);
    if (e.S->status != ST_OK) {
        return;
    }
    if (rep_slot != nullptr) {
        // a window swap keeps the temperature; a temperature change
        // recomputes the energy (update_temp)
        int slot = rep_slot[r];
        bool moved = P.slot_key != nullptr && P.slot_key[slot] != P.slot_key[e.S->temp_slot];
        e.begin_interval(slot, P.exchange_mode != LDO_EXCHANGE_WINDOWS || moved);
    }
    e.simulate(steps);
    if (rep_slot != nullptr) {
        e.update_enthalpy_and_entropy();
    }
}

DEV void sweep_replica(const Params& P, int r, int64_t steps, const int32_t* rep_slot,
                       int par = 0) {
    Eng e;
    sweep_replica_on(e, P, r, P.blocks + (size_t)r * P.L.bytes, steps, rep_slot, par);
}

// ldo_site_eval on a replica block (the caller's copy): the domain is taken
// off its site if placed (unassign_domain, origami_system.cpp:375-385) and
// checked at the query's site (check_domain_constraints, :343-355).  Returns
// false when the chain or the domain does not exist.
DEV bool site_eval_on(const Params& P, char* blk, const ldo_site_query& q, ldo_site_result& res) {
    Eng e;
    e.attach(&P, blk, q.replica);
    e.select_tables();
    res.unassign_delta = 0;
    res.delta = 0;
    res.violated = 0;
    res.pad = 0;
    int cs = -1;
    for (int w = 0; w < e.S->n_chains; w++) {
        if (e.ch[e.order[w]].uid == q.chain_index) {
            cs = e.order[w];
        }
    }
    if (cs < 0 || q.domain < 0 || q.domain >= e.ch[cs].len) {
        return false;
    }
    int s = e.slot(cs, q.domain);
    if (e.dom[s].state != OCC_UNASSIGNED) {
        res.unassign_delta = e.unassign_domain(s);
    }
    e.S->violated = 0;
    res.delta = e.check_domain_constraints(s, i3(q.pos[0], q.pos[1], q.pos[2]),
                                           unit_index(i3(q.ore[0], q.ore[1], q.ore[2])));
    res.violated = e.S->violated ? 1 : 0;
    return true;
}

// OrigamiSystem::update_enthalpy_and_entropy (origami_system.cpp:204-246)
DEV void dependent_replica(const Params& P, int r) {
    Eng e;
    e.attach(&P, P.blocks + (size_t)r * P.L.bytes, r);
    e.update_enthalpy_and_entropy();
}

DEV int32_t replica_status(const Params& P, int r) {
    return ((const RepScalars*)(P.blocks + (size_t)r * P.L.bytes))->status;
}

// Exchange state on the device
struct XState {
    int32_t* q_to_rep;
    int32_t* rep_slot;
    const double* ladder_t;
    const double* ladder_u;
    const double* ladder_s;
    const double* reduced_u; // [n_replicas * (ntypes-1)]
    double* rep_u;           // [n_replicas] each replica's chemical-potential multiplier
    // attempts, accepts, then per slot pair attempts[2R] and accepts[2R]
    // (element d * R + q: the pair of slot q with its neighbour in direction d)
    int64_t* counts;
};

// The exchange stream: Philox4x32-10 keyed by (seed, 0xFFFFFFFF), counter
// (draw, 0, exchange round of the context); or the recorded draws of replica
// 0's tape.
DEV double exchange_uniform(const Params& P, int64_t rng_ctr, uint32_t& idx) {
    if (P.rng_mode == 1) {
        RepScalars* S0 = (RepScalars*)P.blocks;
        int64_t pos = S0->tape_pos;
        if (pos >= P.tape_len[0]) {
            S0->status = ST_TAPE;
            return 0.0;
        }
        S0->tape_pos = pos + 1;
        return bits_to_double(P.tapes[P.tape_off[0] + pos]);
    }
    uint32_t c0 = idx++, c1 = 0u, c2 = (uint32_t)(uint64_t)rng_ctr,
             c3 = (uint32_t)((uint64_t)rng_ctr >> 32);
    uint32_t k0 = (uint32_t)P.seed, k1 = 0xFFFFFFFFu;
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    uint64_t x = (uint64_t)c0 | ((uint64_t)c1 << 32);
    return (double)(x >> 11) * 0x1.0p-53;
}

// calc_acceptance_p + test_acceptance (ptmc_simulation.cpp:255-313) for the
// replicas in slots a and b, swapping them when accepted.  The per-staple
// chemical potential of a replica is reduced_u * T * (its own multiplier),
// as update_staple_us leaves m_staple_us (origami_system.cpp:624-628).
// d: direction of the pair (per-pair counters).
DEV void exchange_pair(const Params& P, const XState& X, int a, int b, int d, int64_t rng_ctr,
                       uint32_t& idx) {
    int R = P.n_replicas;
    int nst = P.L.ntypes - 1;
    X.counts[0]++;
    X.counts[2 + (size_t)d * R + a]++;
    int r1 = X.q_to_rep[a], r2 = X.q_to_rep[b];
    RepScalars* S1 = (RepScalars*)(P.blocks + (size_t)r1 * P.L.bytes);
    RepScalars* S2 = (RepScalars*)(P.blocks + (size_t)r2 * P.L.bytes);
    const int16_t* tc1 = (const int16_t*)((char*)S1 + P.L.off_tcount);
    const int16_t* tc2 = (const int16_t*)((char*)S2 + P.L.off_tcount);
    double t1 = X.ladder_t[a], t2 = X.ladder_t[b];
    double sm1 = X.ladder_s[a], sm2 = X.ladder_s[b];
    double dbu_dn = 0;
    for (int k = 0; k < nst; k++) {
        double u1 = X.reduced_u[(size_t)r1 * nst + k] * t1 * X.rep_u[r1];
        double u2 = X.reduced_u[(size_t)r2 * nst + k] * t2 * X.rep_u[r2];
        dbu_dn += (u2 / t2 - u1 / t1) * ((double)tc2[k + 1] - (double)tc1[k + 1]);
    }
    double DB = 1 / t2 - 1 / t1;
    double DH = S2->hyb_h * t2 - S1->hyb_h * t1;
    double Dst = S2->stack_e * t2 - S1->stack_e * t1;
    double DBM = sm2 / t2 - sm1 / t1;
    double DBias = 0.0 * t2 - 0.0 * t1;
    double p = fmin(1.0, exp(DB * (DH + DBias) + DBM * Dst - dbu_dn));
    bool accept = (p == 1) ? true : (p > exchange_uniform(P, rng_ctr, idx));
    if (accept) {
        X.counts[1]++;
        X.counts[2 + 2 * (size_t)R + (size_t)d * R + a]++;
        X.q_to_rep[a] = r2;
        X.q_to_rep[b] = r1;
        X.rep_slot[r2] = a;
        X.rep_slot[r1] = b;
    }
}

// One exchange round of the PT drivers: OneDPTGCMCSimulation::
// attempt_exchange (ptmc_simulation.cpp:360-412; swap_i picks the even or
// odd neighbour pairs) or TwoDPTGCMCSimulation::attempt_exchange (495-560;
// swap_i % 4 picks even or odd rows of temperature neighbours, or even or
// odd columns of stacking neighbours, in the grid of pt_dim2 columns), then
// master_send (212-226): with UT, HUT and 2D every replica takes its new
// slot's chemical-potential multiplier.  rng_ctr is the round's
// exchange-stream counter.
DEV void exchange_round(const Params& P, const XState& X, int64_t swap_i, int64_t rng_ctr) {
    int R = P.n_replicas;
    uint32_t idx = 0;
    if (P.exchange_mode == LDO_EXCHANGE_2D) {
        int v2 = P.pt_dim2, v1 = R / v2;
        int set = (int)(swap_i % 4);
        bool temps = set % 2 == 0; // sets 0 and 2 pair temperatures
        int i0 = set == 2 ? 1 : 0, j0 = set == 3 ? 1 : 0;
        int di = temps ? 2 : 1, dj = temps ? 1 : 2;
        int i1 = temps ? v1 - 1 : v1, j1 = temps ? v2 : v2 - 1;
        int step = temps ? v2 : 1;
        for (int i = i0; i < i1; i += di) {
            for (int j = j0; j < j1; j += dj) {
                int a = i * v2 + j;
                exchange_pair(P, X, a, a + step, temps ? 0 : 1, rng_ctr, idx);
            }
        }
    }
    else {
        for (int i = (int)(swap_i % 2); i < R - 1; i += 2) {
            exchange_pair(P, X, i, i + 1, 0, rng_ctr, idx);
        }
    }
    int m = P.exchange_mode;
    if (m == LDO_EXCHANGE_TEMPERATURE || m == LDO_EXCHANGE_HUT || m == LDO_EXCHANGE_2D) {
        for (int q = 0; q < R; q++) {
            X.rep_u[X.q_to_rep[q]] = X.ladder_u[q];
        }
    }
}

// A replica's value of a window's order parameter (GridBiasFunction::get_point)
DEV int replica_window_op(const Params& P, int r, const ldo_window& w) {
    const RepScalars* S = (const RepScalars*)(P.blocks + (size_t)r * P.L.bytes);
    return w.op == LDO_OP_NUMSTAPLES ? S->num_staples : S->num_fully;
}

// grid bias of window q at point p as the exchange reads it (a GridFloats
// lookup: points without a value read 0)
DEV double exchange_grid(const Params& P, int q, int p) {
    return (p >= 0 && p < P.grid_len) ? P.win_grid[(size_t)q * P.grid_len + p] : 0.0;
}

// PTMWUSGCMCSimulation::attempt_exchange (us_simulation.cpp:770-864):
// adjacent windows swap configurations when each window's order parameter
// lies inside the other window.  Windows are ladder slots here, so a swap
// exchanges the two replicas' slots; each window keeps its stored biases.
DEV void exchange_window_group(const Params& P, const XState& X, int64_t swap_i,
                               int64_t rng_ctr, int g0, int g1, uint32_t& idx) {
    for (int i = g0 + (int)(swap_i % 2); i < g1 - 1; i += 2) {
        X.counts[0]++;
        X.counts[2 + i]++;
        const ldo_window& w1 = P.windows[i];
        const ldo_window& w2 = P.windows[i + 1];
        int r1 = X.q_to_rep[i], r2 = X.q_to_rep[i + 1];
        int p1 = replica_window_op(P, r1, w1);
        int p2 = replica_window_op(P, r2, w2);
        if (p2 < w1.min_op || p2 > w1.max_op || p1 < w2.min_op || p1 > w2.max_op) {
            continue;
        }
        bool accept = true;
        if (p1 != p2) {
            double d1 = exchange_grid(P, i, p1) - exchange_grid(P, i + 1, p1);
            double d2 = exchange_grid(P, i + 1, p2) - exchange_grid(P, i, p2);
            double p = fmin(1.0, exp(d1 + d2));
            accept = (p == 1) ? true : (p > exchange_uniform(P, rng_ctr, idx));
        }
        if (accept) {
            X.counts[1]++;
            X.counts[2 + 2 * (size_t)P.n_replicas + i]++;
            X.q_to_rep[i] = r2;
            X.q_to_rep[i + 1] = r1;
            X.rep_slot[r2] = i;
            X.rep_slot[r1] = i + 1;
        }
    }
}

// Swap of the replicas in window w of two ladders at neighbouring
// temperatures (slots a < b, LDO_FLAG_WINDOW_TEMPERATURE_SWAPS): the PTMC log
// ratio of the two slots plus the grid biases of both windows at both
// replicas' order parameters; the windows' restraints are equal and cancel.
DEV void exchange_window_temperatures(const Params& P, const XState& X, int a, int b,
                                      int64_t rng_ctr, uint32_t& idx) {
    int R = P.n_replicas;
    int nst = P.L.ntypes - 1;
    X.counts[0]++;
    X.counts[2 + (size_t)R + a]++;
    int r1 = X.q_to_rep[a], r2 = X.q_to_rep[b];
    const RepScalars* S1 = (const RepScalars*)(P.blocks + (size_t)r1 * P.L.bytes);
    const RepScalars* S2 = (const RepScalars*)(P.blocks + (size_t)r2 * P.L.bytes);
    const int16_t* tc1 = (const int16_t*)((const char*)S1 + P.L.off_tcount);
    const int16_t* tc2 = (const int16_t*)((const char*)S2 + P.L.off_tcount);
    double t1 = X.ladder_t[a], t2 = X.ladder_t[b];
    double sm1 = X.ladder_s[a], sm2 = X.ladder_s[b];
    double dbu_dn = 0;
    for (int k = 0; k < nst; k++) {
        double u1 = X.reduced_u[(size_t)r1 * nst + k] * t1 * X.rep_u[r1];
        double u2 = X.reduced_u[(size_t)r2 * nst + k] * t2 * X.rep_u[r2];
        dbu_dn += (u2 / t2 - u1 / t1) * ((double)tc2[k + 1] - (double)tc1[k + 1]);
    }
    double DB = 1 / t2 - 1 / t1;
    double DH = S2->hyb_h * t2 - S1->hyb_h * t1;
    double Dst = S2->stack_e * t2 - S1->stack_e * t1;
    double DBM = sm2 / t2 - sm1 / t1;
    int p1 = replica_window_op(P, r1, P.windows[a]);
    int p2 = replica_window_op(P, r2, P.windows[b]);
    double dg = (exchange_grid(P, a, p1) - exchange_grid(P, b, p1)) +
                (exchange_grid(P, b, p2) - exchange_grid(P, a, p2));
    double p = fmin(1.0, exp(DB * DH + DBM * Dst - dbu_dn + P.windows[a].bias_mult * dg));
    bool accept = (p == 1) ? true : (p > exchange_uniform(P, rng_ctr, idx));
    if (accept) {
        X.counts[1]++;
        X.counts[2 + 3 * (size_t)R + a]++;
        X.q_to_rep[a] = r2;
        X.q_to_rep[b] = r1;
        X.rep_slot[r2] = a;
        X.rep_slot[r1] = b;
    }
}

// One exchange round over every window ladder.
DEV void exchange_windows_round(const Params& P, const XState& X, int64_t swap_i,
                                int64_t rng_ctr) {
    int R = P.n_replicas;
    uint32_t idx = 0;
    // each run of consecutive slots at one temperature is a window ladder of
    // its own (one PTMWUS run of the reference), paired from its first slot
    int W = R;
    for (int g0 = 0, g1; g0 < R; g0 = g1) {
        for (g1 = g0 + 1; g1 < R && X.ladder_t[g1] == X.ladder_t[g0]; g1++) {
        }
        if (g0 == 0) {
            W = g1;
        }
        exchange_window_group(P, X, swap_i, rng_ctr, g0, g1, idx);
    }
    if (P.flags & LDO_FLAG_WINDOW_TEMPERATURE_SWAPS) {
        for (int g = (int)(swap_i % 2); (g + 2) * W <= R; g += 2) {
            for (int w = 0; w < W; w++) {
                exchange_window_temperatures(P, X, g * W + w, (g + 1) * W + w, rng_ctr, idx);
            }
        }
    }
}

// the stored move-level bias of the window in ladder slot q (0 without
// windows).  The replica in slot q is rep_slot's inverse, not the block's
// temp_slot, which follows a swap only when the next interval begins.
DEV double slot_bias(const Params& P, int q) {
    if (P.windows == nullptr) {
        return 0;
    }
    return (P.win_stored[2 * q] + P.win_stored[2 * q + 1]) * P.windows[q].bias_mult;
}

// ------------------------------------------------------ learned proposal
// The lane parts of a learned-move attempt (Eng::learned_begin/finish): lane
// `lane` of `nl` handles sites lane, lane + nl, ... of the input box and
// hidden units lane, lane + nl, ...; every value is computed by one lane in a
// fixed order, so any lane count gives the same bits.

// the one-hot channels of the 7x7x7 box around the anchor (site index
// (dz+3)*49 + (dy+3)*7 + (dx+3)) and the six neighbour sites that can take d
// (at most one other domain there)
DEV void learned_context_lane(const Eng& e, LearnedMove& m, int lane, int nl) {
    for (int s = lane; s < 343; s += nl) {
        I3 p = i3(m.ax + s % 7 - 3, m.ay + (s / 7) % 7 - 3, m.az + s / 49 - 3);
        m.ch[s] = (uint8_t)e.site_channel(p, m.d);
    }
    for (int n = lane; n < 6; n += nl) {
        I3 p = add3(i3(m.ax, m.ay, m.az), unit3(n));
        m.legal[n] = e.site_count(p, m.d) <= 1 ? 1 : 0;
    }
}

// h[0] = relu(b1 + sum over sites of W1[site*13 + channel] + emb[id] W1e)
DEV void net_layer1_lane(const Params& P, LearnedMove& m, int lane, int nl) {
    const float* W = P.net;
    const float* e = W + P.net_off[1] + (int64_t)m.id * 16;
    for (int j = lane; j < 256; j += nl) {
        float s = W[P.net_off[3] + j];
        for (int site = 0; site < 343; site++) {
            s += W[P.net_off[0] + (int64_t)(site * 13 + m.ch[site]) * 256 + j];
        }
        for (int k = 0; k < 16; k++) {
            s = fmaf(e[k], W[P.net_off[2] + k * 256 + j], s);
        }
        m.h[0][j] = fmaxf(s, 0.f);
    }
}

// out[j] = b[j] + sum_i in[i] W[i][j] (fused multiply-adds in i order),
// ReLU when `relu`
DEV void net_dense_lane(const float* W, const float* b, const float* in, float* out, int n_in,
                        int n_out, bool relu, int lane, int nl) {
    for (int j = lane; j < n_out; j += nl) {
        float s = b[j];
        for (int i = 0; i < n_in; i++) {
            s = fmaf(in[i], W[(int64_t)i * n_out + j], s);
        }
        out[j] = relu ? fmaxf(s, 0.f) : s;
    }
}

// network phase `ph` (0: context, 1: layer 1, 2-3: hidden, 4: logits); the
// kernel separates phases with barriers, the host build runs them in order
DEV void learned_phase_lane(const Eng& e, const Params& P, LearnedMove& m, int ph, int lane,
                            int nl) {
    const float* W = P.net;
    switch (ph) {
    case 0:
        learned_context_lane(e, m, lane, nl);
        break;
    case 1:
        net_layer1_lane(P, m, lane, nl);
        break;
    case 2:
        net_dense_lane(W + P.net_off[4], W + P.net_off[5], m.h[0], m.h[1], 256, 256, true, lane, nl);
        break;
    case 3:
        net_dense_lane(W + P.net_off[6], W + P.net_off[7], m.h[1], m.h[0], 256, 256, true, lane, nl);
        break;
    case 4:
        net_dense_lane(W + P.net_off[8], W + P.net_off[9], m.h[0], m.logits, 256, 36, false, lane,
                       nl);
        break;
    }
}

#define LDO_LEARNED_PHASES 5

#endif
