// plan.cpp - host-side preparation of a sampler run (see plan.h).
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ldo {

namespace {

int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

} // namespace

void mlp_layout(int n_ident, ldo_mlp_layout& L) {
    const int64_t H = LDO_MLP_HIDDEN;
    int64_t o = 0;
    auto take = [&](int64_t& field, int64_t n) {
        field = o;
        o += n;
    };
    take(L.w1, (int64_t)LDO_MLP_FEATURES * H);
    take(L.emb, (int64_t)n_ident * LDO_MLP_EMBED);
    take(L.w1e, (int64_t)LDO_MLP_EMBED * H);
    take(L.b1, H);
    take(L.w2, H * H);
    take(L.b2, H);
    take(L.w3, H * H);
    take(L.b3, H);
    take(L.w4, H * LDO_MLP_OUT);
    take(L.b4, LDO_MLP_OUT);
    L.total = o;
}

Layout make_layout(int ns, int lmax, int smax, int ntypes, int tcap, bool rg, bool jump) {
    Layout L;
    std::memset(&L, 0, sizeof L);
    L.ns = ns;
    L.lmax = lmax;
    // staple chain k owns the slots ns + (k-1)*2^lshift + [0, lmax): a
    // power-of-two stride turns the slot -> (chain, domain) maps into a
    // shift and a mask
    L.lshift = 0;
    while ((1 << L.lshift) < lmax) {
        L.lshift++;
    }
    L.smax = smax;
    L.nslots = ns + (smax << L.lshift);
    L.nchain = 1 + smax;
    L.ntypes = ntypes;
    L.tcap = tcap;
    // at most nslots occupied sites: load factor <= 1/2
    int hcap = 64, hbits = 6;
    while (hcap < 2 * L.nslots) {
        hcap *= 2;
        hbits++;
    }
    L.hcap = hcap;
    L.hbits = hbits;
    // one CTCB pass adds at most one active endpoint per domain of the
    // scanned networks (find_growthpoints_endpoints) and one per inactive
    // endpoint reached (update_endpoints), plus the segment ends
    L.ecap = 2 * L.nslots + 16;
    int64_t off = round_up(sizeof(RepScalars), 16);
    auto take = [&](int64_t& field, int64_t bytes) {
        field = off;
        off = round_up(off + bytes, 16);
    };
    int64_t S = L.nslots, C = L.nchain;
    take(L.off_dom, S * (int64_t)sizeof(DomRec));
    take(L.off_dident, S * 2);
    take(L.off_chain, C * (int64_t)sizeof(ChainRec));
    take(L.off_order, C * 2);
    take(L.off_free, C * 2);
    take(L.off_tlist, (int64_t)ntypes * tcap * 2);
    take(L.off_tcount, (int64_t)ntypes * 2);
    take(L.off_hkey, (int64_t)hcap * 8);
    take(L.off_hval, (int64_t)hcap * 4);
    take(L.off_modified, 2 * S * 2);
    take(L.off_assigned, 2 * S * 2);
    take(L.off_prev, S * (int64_t)sizeof(PosOre));
    take(L.off_old, S * (int64_t)sizeof(PosOre));
    take(L.off_seg, S);
    take(L.off_gp, S * 2);
    take(L.off_stem, S * 2);
    take(L.off_inact, S * 2);
    take(L.off_netseg, S);
    take(L.off_netds_flag, S);
    take(L.off_checked, C);
    take(L.off_netcs, C);
    take(L.off_regrow, C);
    take(L.off_dir0, C);
    take(L.off_dir1, C);
    take(L.off_active, (int64_t)L.ecap * (int64_t)sizeof(Endpoint));
    take(L.off_active0, (int64_t)L.ecap * (int64_t)sizeof(Endpoint));
    take(L.off_gps, 2 * S * 2);
    take(L.off_iaes, 2 * S * 2);
    take(L.off_netds, S * 2);
    take(L.off_frames, C * (int64_t)sizeof(NetFrame));
    take(L.off_part, C);
    take(L.off_sframes, C * (int64_t)sizeof(NetFrame));
    take(L.off_gstack, LDO_GROW_DEPTH * (int64_t)sizeof(GrowFrame));
    take(L.off_gpl, 2 * LDO_MAX_STAPLE_LEN * 2);
    // CTRG recoil growth, only when the moveset holds it: the regrowth
    // stack, per level the attempts made and configurations left (growth and
    // weight copies) and the open probability, and the erased endpoints of
    // the assigned levels (a stack of positions and one count per level)
    int64_t R = rg ? S : 0;
    take(L.off_rgs, R * 2);
    take(L.off_rgatt, R);
    take(L.off_rgwatt, R);
    take(L.off_rgavail, R * 8);
    take(L.off_rgwavail, R * 8);
    take(L.off_rgopen, R * 8);
    take(L.off_erstk, (rg ? (int64_t)L.ecap : 0) * (int64_t)sizeof(Endpoint));
    take(L.off_ercnt, R * 2);
    int64_t J = jump ? 1 : 0;
    take(L.off_jdom, J * ns);
    take(L.off_jps, J * LDO_JUMP_QUEUE * 2);
    take(L.off_jseg, J * LDO_JUMP_MAX_SEGS * (int64_t)sizeof(JumpSeg));
    take(L.off_jstem, J * LDO_JUMP_MAX_SEGS);
    take(L.off_stemseg, J * ns);
    take(L.off_sdir, J * LDO_JUMP_MAX_SEGS);
    L.track_len = 6 * ntypes + 4 * (ns + 1) + 4 * (smax + 1) + 2;
    take(L.off_track, (int64_t)L.track_len * 8);
    L.bytes = round_up(off, 256);
    return L;
}

void make_plan(const ldo_system_desc* sys, const ldo_run_params* run, int32_t n_replicas,
               const double* temps, Plan& out) {
    if (!sys || !run || !temps || n_replicas < 1) {
        throw AbiError {LDO_E_ARG, "null argument or no replicas"};
    }
    if (sys->apply_mean_field_cor) {
        throw AbiError {LDO_E_NOTIMPL, "apply_mean_field_cor is not supported"};
    }
    if (sys->n_chain_types < 2) {
        throw AbiError {LDO_E_ARG, "a system needs a scaffold and at least one staple type"};
    }
    if (run->n_movetypes < 1 || run->n_movetypes > LDO_MAX_MOVES) {
        throw AbiError {LDO_E_ARG, "moveset must hold 1..8 movetypes"};
    }
    if (run->exchange_mode < LDO_EXCHANGE_NONE || run->exchange_mode > LDO_EXCHANGE_2D) {
        throw AbiError {LDO_E_ARG, "unknown exchange_mode"};
    }
    if (run->exchange_mode != LDO_EXCHANGE_NONE && run->exchange_mode != LDO_EXCHANGE_WINDOWS &&
        n_replicas > 64) {
        throw AbiError {LDO_E_ARG, "a temperature ladder holds at most 64 slots"};
    }
    out.R = n_replicas;
    out.ntypes = sys->n_chain_types;
    SystemSpec& sp = out.spec;
    int k = 0, lo = 0, hi = 0, lmax = 0;
    bool first = true;
    for (int t = 0; t < sys->n_chain_types; t++) {
        std::vector<int> ids;
        std::vector<std::string> sq;
        for (int d = 0; d < sys->chain_lengths[t]; d++, k++) {
            int v = sys->identities[k];
            ids.push_back(v);
            if (first) {
                lo = hi = v;
                first = false;
            }
            lo = std::min(lo, v);
            hi = std::max(hi, v);
            if (sys->hybridization_pot == LDO_HYB_NEARESTNEIGHBOUR ||
                sys->stacking_pot == LDO_STACK_SEQUENCESPECIFIC) {
                if (!sys->sequences) {
                    throw AbiError {LDO_E_ARG, "sequences are required by this potential"};
                }
                sq.push_back(sys->sequences[k]);
            }
            else {
                sq.push_back("");
            }
        }
        if (t > 0) {
            lmax = std::max(lmax, (int)ids.size());
        }
        sp.idents.push_back(ids);
        sp.seqs.push_back(sq);
    }
    if (lmax > 8 || lmax > run->max_staple_size) {
        throw AbiError {LDO_E_ARG, "staples longer than max_staple_size (at most 8)"};
    }
    sp.ident_lo = lo;
    sp.n_ident = hi - lo + 1;
    sp.hybridization_pot = sys->hybridization_pot;
    sp.stacking_pot = sys->stacking_pot;
    sp.binding_h = sys->binding_h;
    sp.binding_s = sys->binding_s;
    sp.misbinding_h = sys->misbinding_h;
    sp.misbinding_s = sys->misbinding_s;
    sp.stacking_ene = sys->stacking_ene;
    sp.cation_M = sys->cation_M;
    for (int i = 0; i < sys->n_specified; i++) {
        sp.spec_h.push_back(sys->enthalpies[i]);
        sp.spec_s.push_back(sys->entropies[i]);
    }
    int ns = (int)sp.idents[0].size();
    int smax = run->max_total_staples;
    if (smax < 1 || smax > 4000) {
        throw AbiError {LDO_E_ARG, "max_total_staples must be in 1..4000"};
    }
    int lstride = 1;
    while (lstride < lmax) {
        lstride *= 2;
    }
    if (ns + (int64_t)smax * lstride > 30000) {
        throw AbiError {LDO_E_ARG, "too many domain slots for 16-bit slot indices"};
    }
    int tcap = std::min(std::max(run->max_type_staples, 1), smax);
    Params& P = out.P;
    std::memset(&P, 0, sizeof P);
    bool rg = false, jump = false;
    for (int m = 0; m < run->n_movetypes; m++) {
        int k = run->movetype_kind[m];
        bool j = k == LDO_MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH || k == LDO_MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH;
        jump = jump || j;
        // the jump moves keep their regrowth stack in the CTRG sections
        rg = rg || j || k == LDO_MOVE_CTRG_SCAFFOLD_REGROWTH;
    }
    P.L = make_layout(ns, lmax, smax, sys->n_chain_types, tcap, rg, jump);
    P.n_replicas = n_replicas;
    P.domain_type = sys->domain_type;
    P.misbinding_pot = sys->misbinding_pot;
    P.cyclic = sys->cyclic;
    P.n_moves = run->n_movetypes;
    double cum = 0;
    for (int m = 0; m < run->n_movetypes; m++) {
        int kind = run->movetype_kind[m];
        if (kind < 0 || kind > 6) {
            throw AbiError {LDO_E_NOTIMPL, "movetype not supported"};
        }
        P.move_kind[m] = kind;
        cum += run->movetype_freq[m];
        P.move_cum[m] = cum;
        P.move_maxreg[m] = run->movetype_max_regrowth ? run->movetype_max_regrowth[m] : 0;
        P.move_recoils[m] = 0;
        P.move_cattempts[m] = 0;
        P.move_segreg[m] = 0;
        if (kind >= 3 && P.move_maxreg[m] < 2) {
            throw AbiError {LDO_E_ARG, "scaffold regrowth max_regrowth must be at least 2"};
        }
        if (kind == 5 || kind == 6) {
            // the first segment's length is drawn from [2, max_seg_regrowth + 1]
            // (movetypes.cpp:522-523)
            if (!run->movetype_max_seg_regrowth || run->movetype_max_seg_regrowth[m] < 1) {
                throw AbiError {LDO_E_ARG, "a jump move needs max_seg_regrowth >= 1"};
            }
            if (P.move_maxreg[m] > LDO_JUMP_MAX_REGROWTH) {
                throw AbiError {LDO_E_ARG, "a jump move's max_regrowth must be at most 60"};
            }
            P.move_segreg[m] = run->movetype_max_seg_regrowth[m];
        }
        if (kind == 4 || kind == 5) {
            if (!run->movetype_max_num_recoils || !run->movetype_max_c_attempts) {
                throw AbiError {LDO_E_ARG, "CTRG needs max_num_recoils and max_c_attempts"};
            }
            P.move_recoils[m] = run->movetype_max_num_recoils[m];
            P.move_cattempts[m] = run->movetype_max_c_attempts[m];
            // the trial configurations of a domain are the 36 (site,
            // orientation) pairs, drawn without replacement
            if (P.move_recoils[m] < 0 || P.move_cattempts[m] < 1 || P.move_cattempts[m] > 36) {
                throw AbiError {LDO_E_ARG,
                                "CTRG needs max_num_recoils >= 0 and 1 <= max_c_attempts <= 36"};
            }
        }
    }
    P.max_total_staples = smax;
    P.max_type_staples = run->max_type_staples;
    P.centering_freq = run->centering_freq;
    P.centering_domain = run->centering_domain;
    if (P.centering_domain < 0 || P.centering_domain >= ns) {
        throw AbiError {LDO_E_ARG, "centering_domain out of range"};
    }
    P.check_freq = run->constraint_check_freq;
    P.rng_mode = run->rng_mode;
    P.seed = run->seed;
    P.ident_lo = lo;
    P.n_ident = sp.n_ident;
    P.fugacity = sys->staple_M;
    int lpr = run->lanes_per_replica;
    if (lpr == 0) {
        lpr = n_replicas >= 2048 ? 1 : 64;
    }
    if (lpr != 1 && lpr != 64) {
        throw AbiError {LDO_E_ARG, "lanes_per_replica must be 0, 1 or 64"};
    }
    P.lanes_per_replica = lpr;
    if (run->flags & ~(LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS | LDO_FLAG_WINDOW_TEMPERATURE_SWAPS)) {
        throw AbiError {LDO_E_ARG, "unknown ldo_run_params.flags bits"};
    }
    if ((run->flags & LDO_FLAG_WINDOW_TEMPERATURE_SWAPS) &&
        run->exchange_mode != LDO_EXCHANGE_WINDOWS) {
        throw AbiError {LDO_E_ARG, "window temperature swaps need LDO_EXCHANGE_WINDOWS"};
    }
    P.flags = run->flags;
    P.n_slots = n_replicas;
    P.exchange_mode = run->exchange_mode;
    // chain lengths by identity, and the table index of every domain of
    // every chain type
    out.type_len.assign(sys->n_chain_types, 0);
    int stride = std::max(lmax, ns);
    out.type_ident.assign((size_t)sys->n_chain_types * stride, 0);
    for (int t = 0; t < sys->n_chain_types; t++) {
        out.type_len[t] = (int)sp.idents[t].size();
        for (size_t d = 0; d < sp.idents[t].size(); d++) {
            out.type_ident[(size_t)t * stride + d] = (int16_t)(sp.idents[t][d] - lo);
        }
    }
    P.type_ident_stride = stride;
    // ladder and per-replica chemical potentials (molarity_to_chempots,
    // chempots_to_reduced_chempots: origami_system.cpp:965-989)
    int nst = sys->n_chain_types - 1;
    out.reduced_u.assign((size_t)n_replicas * nst, 0.0);
    for (int r = 0; r < n_replicas; r++) {
        out.ladder_t.push_back(temps[r]);
        out.ladder_u.push_back(1.0);
        out.ladder_b.push_back(1.0);
        out.ladder_s.push_back(1.0);
        for (int t = 0; t < nst; t++) {
            double extra = (2 * sp.idents[t + 1].size() - 1) * std::log(6);
            double u = temps[r] * (std::log(sys->staple_M) - extra);
            out.reduced_u[(size_t)r * nst + t] = u / temps[r];
        }
    }
}

void slot_tables(const Plan& p, std::vector<double>& tab, std::vector<double>& init) {
    int n = p.R;
    size_t n2 = (size_t)p.spec.n_ident * p.spec.n_ident;
    tab.assign(4 * n2 * n, 0.0);
    init.assign(3 * (size_t)n, 0.0);
    for (int q = 0; q < n; q++) {
        TableSet t = build_tables(p.spec, p.ladder_t[q], slot_stacking(p, q));
        std::copy(t.hyb_e.begin(), t.hyb_e.end(), tab.begin() + (0 * (size_t)n + q) * n2);
        std::copy(t.hyb_h.begin(), t.hyb_h.end(), tab.begin() + (1 * (size_t)n + q) * n2);
        std::copy(t.hyb_s.begin(), t.hyb_s.end(), tab.begin() + (2 * (size_t)n + q) * n2);
        std::copy(t.stack.begin(), t.stack.end(), tab.begin() + (3 * (size_t)n + q) * n2);
        init[q] = t.init_e;
        init[n + q] = t.init_h;
        init[2 * n + q] = t.init_s;
    }
}

std::vector<double> ladder_values(const Plan& p) {
    int n = p.R;
    std::vector<double> xd(4 * (size_t)n + p.reduced_u.size());
    for (int q = 0; q < n; q++) {
        xd[q] = p.ladder_t[q];
        xd[n + q] = p.ladder_u[q];
        xd[2 * n + q] = p.ladder_s[q];
    }
    std::copy(p.reduced_u.begin(), p.reduced_u.end(), xd.begin() + 3 * n);
    size_t o = 3 * (size_t)n + p.reduced_u.size();
    for (int r = 0; r < n; r++) {
        xd[o + r] = p.ladder_b[r];
    }
    return xd;
}

double slot_stacking(const Plan& p, int q) {
    int m = p.P.exchange_mode;
    if (m == LDO_EXCHANGE_TEMPERATURE || m == LDO_EXCHANGE_T || m == LDO_EXCHANGE_HUT) {
        return 1.0;
    }
    return p.ladder_s[q];
}

int slot_classes(const Plan& p, std::vector<int32_t>& key) {
    key.assign(p.R, 0);
    std::vector<std::pair<double, double>> seen;
    for (int q = 0; q < p.R; q++) {
        std::pair<double, double> k {p.ladder_t[q], slot_stacking(p, q)};
        size_t c = 0;
        while (c < seen.size() && seen[c] != k) {
            c++;
        }
        if (c == seen.size()) {
            seen.push_back(k);
        }
        key[q] = (int32_t)c;
    }
    return (int)seen.size();
}

int window_ladder_size(const Plan& p) {
    int W = 1;
    while (W < p.R && p.ladder_t[W] == p.ladder_t[0]) {
        W++;
    }
    if (p.R % W != 0) {
        return -1;
    }
    for (int q = 0; q < p.R; q++) {
        if (p.ladder_t[q] != p.ladder_t[q - q % W]) {
            return -1;
        }
        if (q % W == 0 && q > 0 && p.ladder_t[q] == p.ladder_t[q - 1]) {
            return -1;
        }
    }
    return W;
}

bool u_follows_slot(int m) {
    return m == LDO_EXCHANGE_TEMPERATURE || m == LDO_EXCHANGE_HUT || m == LDO_EXCHANGE_2D;
}

void set_ladder_2d(Plan& p, int32_t n_temps, const double* temps, int32_t n_stacking,
                   const double* stacking_mults, double staple_u_mult) {
    if (p.P.exchange_mode != LDO_EXCHANGE_2D) {
        throw AbiError {LDO_E_ARG, "a 2D ladder needs exchange_mode LDO_EXCHANGE_2D"};
    }
    if (!temps || !stacking_mults || n_temps < 1 || n_stacking < 1 ||
        (int64_t)n_temps * n_stacking != p.R) {
        throw AbiError {LDO_E_ARG, "the 2D grid must hold one slot per replica"};
    }
    p.ladder_t.clear();
    p.ladder_u.clear();
    p.ladder_b.clear();
    p.ladder_s.clear();
    for (int i = 0; i < n_temps; i++) {
        for (int j = 0; j < n_stacking; j++) {
            p.ladder_t.push_back(temps[i]);
            p.ladder_u.push_back(staple_u_mult);
            p.ladder_b.push_back(1.0);
            p.ladder_s.push_back(stacking_mults[j]);
        }
    }
    p.dim2 = n_stacking;
}

void check_learned_types(const Plan& p) {
    for (int t = 1; t < p.ntypes; t++) {
        if (p.type_len[t] < 2) {
            throw AbiError {LDO_E_NOTIMPL, "the learned moves need staples of two or more domains"};
        }
    }
}

int config_domains(const Plan& p, const ldo_config* cfg) {
    if (!cfg || cfg->n_chains < 1 || cfg->chain_identity[0] != 0) {
        throw AbiError {LDO_E_ARG, "configuration must start with the scaffold chain"};
    }
    if (cfg->n_chains - 1 > p.P.max_total_staples) {
        throw AbiError {LDO_E_ARG, "configuration holds more staples than max_total_staples"};
    }
    int n_domains = 0;
    for (int i = 0; i < cfg->n_chains; i++) {
        int id = cfg->chain_identity[i];
        if (id < 0 || id >= p.ntypes) {
            throw AbiError {LDO_E_ARG, "chain identity out of range"};
        }
        n_domains += (int)p.spec.idents[id].size();
    }
    for (int k = 0; k < 3 * n_domains; k++) {
        int v = cfg->orientations[k];
        if (v < -1 || v > 1) {
            throw AbiError {LDO_E_ARG, "orientation is not a unit vector"};
        }
    }
    return n_domains;
}

void pack_config(const Plan& p, const ldo_config* cfg, std::vector<int32_t>& packed,
                 int& n_domains) {
    n_domains = config_domains(p, cfg);
    size_t need = (size_t)2 * cfg->n_chains + (size_t)6 * n_domains;
    packed.assign(need, 0);
    std::copy(cfg->chain_index, cfg->chain_index + cfg->n_chains, packed.begin());
    std::copy(cfg->chain_identity, cfg->chain_identity + cfg->n_chains,
              packed.begin() + cfg->n_chains);
    std::copy(cfg->positions, cfg->positions + 3 * n_domains, packed.begin() + 2 * cfg->n_chains);
    std::copy(cfg->orientations, cfg->orientations + 3 * n_domains,
              packed.begin() + 2 * cfg->n_chains + 3 * n_domains);
}

void read_config(const Plan& p, const char* block, ldo_config* cfg, int32_t cap_chains,
                 int32_t cap_domains) {
    const Layout& L = p.P.L;
    const RepScalars* S = (const RepScalars*)block;
    const DomRec* dom = (const DomRec*)(block + L.off_dom);
    const ChainRec* ch = (const ChainRec*)(block + L.off_chain);
    const int16_t* order = (const int16_t*)(block + L.off_order);
    if (S->n_chains > cap_chains) {
        throw AbiError {LDO_E_ARG, "chain capacity too small"};
    }
    static const int U[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                                {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    int k = 0;
    for (int w = 0; w < S->n_chains; w++) {
        int cs = order[w];
        cfg->chain_index[w] = ch[cs].uid;
        cfg->chain_identity[w] = ch[cs].ident;
        cfg->chain_length[w] = ch[cs].len;
        for (int d = 0; d < ch[cs].len; d++, k++) {
            if (k >= cap_domains) {
                throw AbiError {LDO_E_ARG, "domain capacity too small"};
            }
            int s = cs == 0 ? d : L.ns + ((cs - 1) << L.lshift) + d;
            cfg->positions[3 * k] = dom[s].x;
            cfg->positions[3 * k + 1] = dom[s].y;
            cfg->positions[3 * k + 2] = dom[s].z;
            for (int q = 0; q < 3; q++) {
                cfg->orientations[3 * k + q] = U[dom[s].ore][q];
            }
        }
    }
    cfg->n_chains = S->n_chains;
}

void fill_observables(const Plan& p, const RepScalars& S, ldo_obs* out) {
    out->energy = S.energy;
    out->enthalpy = S.hyb_h;
    out->entropy = S.hyb_s;
    out->stacking_energy = S.stack_e;
    out->temperature = p.ladder_t[S.temp_slot];
    out->step = S.step;
    out->num_staples = S.num_staples;
    out->num_domains = S.num_domains;
    out->num_bound_domain_pairs = S.num_bound;
    out->num_fully_bound_domain_pairs = S.num_fully;
    out->num_misbound_domain_pairs = S.num_bound - S.num_fully;
    out->num_stacked_domain_pairs = S.num_stacked;
    out->num_unassigned_domains = S.num_unassigned;
    out->status = S.status;
    out->last_move = S.last_move;
    out->last_accepted = S.last_acc;
}

void add_counters(const RepScalars& S, ldo_counters* out) {
    for (int m = 0; m < LDO_MAX_MOVES; m++) {
        out->attempts[m] += S.attempts[m];
        out->accepts[m] += S.accepts[m];
    }
    out->site_empty += S.site_empty;
    out->site_bind += S.site_bind;
    out->commits += S.commits;
}

void check_status(int32_t status, int replica) {
    if (status == ST_OK) {
        return;
    }
    const char* why = "unknown";
    int code = LDO_E_MISUSE;
    switch (status) {
    case ST_CONSTRAINT:
        why = "constraint violation (OrigamiMisuse)";
        code = LDO_E_CONSTRAINT;
        break;
    case ST_CAPACITY:
        why = "replica capacity exceeded";
        break;
    case ST_TAPE:
        why = "recorded draw sequence exhausted";
        break;
    case ST_COORD:
        why = "lattice coordinate outside the packed-key range";
        break;
    case ST_MISUSE:
        why = "sampler misuse";
        break;
    }
    throw AbiError {code, "replica " + std::to_string(replica) + ": " + why};
}

void check_windows(const Plan& p, int32_t n, const ldo_window* windows, int32_t grid_len) {
    if (n != p.R || !windows || grid_len < 0) {
        throw AbiError {LDO_E_ARG, "one window per replica (ladder slot) is required"};
    }
    if (p.P.exchange_mode != LDO_EXCHANGE_NONE && p.P.exchange_mode != LDO_EXCHANGE_WINDOWS) {
        throw AbiError {LDO_E_NOTIMPL, "umbrella windows with a temperature ladder"};
    }
    for (int w = 0; w < n; w++) {
        if (windows[w].op != LDO_OP_NUMFULLDOMAINS && windows[w].op != LDO_OP_NUMSTAPLES) {
            throw AbiError {LDO_E_ARG, "unknown window order parameter"};
        }
    }
    if (p.P.flags & LDO_FLAG_WINDOW_TEMPERATURE_SWAPS) {
        // equal ladders of consecutive slots at one temperature, each holding
        // the first ladder's windows in its order
        int W = window_ladder_size(p);
        if (W <= 0) {
            throw AbiError {LDO_E_ARG, "window temperature swaps need window ladders of one size"};
        }
        for (int q = W; q < n; q++) {
            const ldo_window& a = windows[q % W];
            const ldo_window& b = windows[q];
            if (a.op != b.op || a.min_op != b.min_op || a.max_op != b.max_op ||
                a.well_bias != b.well_bias || a.min_bias != b.min_bias || a.slope != b.slope ||
                a.off_grid_bias != b.off_grid_bias || a.bias_mult != b.bias_mult) {
                throw AbiError {LDO_E_ARG, "window temperature swaps need every ladder to hold "
                                           "the first ladder's windows"};
            }
        }
    }
}

double window_restraint(const ldo_window& w, int p) {
    if (p < w.min_op) {
        return w.slope * (w.min_op - p - 1) + w.min_bias;
    }
    if (p > w.max_op) {
        return w.slope * (p - w.max_op - 1) + w.min_bias;
    }
    return w.well_bias;
}

int window_op_value(const ldo_window& w, const RepScalars& S) {
    return w.op == LDO_OP_NUMSTAPLES ? S.num_staples : S.num_fully;
}

std::vector<double> initial_window_state(const ldo_window* windows, int32_t n,
                                         const std::vector<int32_t>& q_to_rep,
                                         const std::vector<RepScalars>& scalars) {
    std::vector<double> st(2 * (size_t)n, 0.0);
    for (int w = 0; w < n; w++) {
        const RepScalars& S = scalars[q_to_rep[w]];
        st[2 * w] = window_restraint(windows[w], window_op_value(windows[w], S));
    }
    return st;
}

} // namespace ldo
