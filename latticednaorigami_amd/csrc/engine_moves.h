// engine_moves.h - the BASELINE move types and the simulate loop.
//
// MCMovetype and subclasses (movetypes.cpp, met_movetypes.cpp,
// cb_movetypes.cpp, orientation_movetype.cpp) and the constraint-point
// machinery of conserved-topology regrowth (top_constraint_points.cpp).
// Recursions of the reference (staple-network scans, connector scans,
// nested staple growth) run on explicit stacks in the replica block.
#ifndef LDO_ENGINE_MOVES_H
#define LDO_ENGINE_MOVES_H

#include "engine.h"

// MCMovetype::select_random_domain (movetypes.cpp:98-113)
DEVIN int Eng::select_random_domain() {
    int idx = uniform_int(0, S->num_domains - 1);
    int n = S->n_chains;
    if (!par) {
        for (int w = 0; w < n; w++) {
            int cs = order[w];
            int len = ch[cs].len;
            if (idx < len) {
                return slot(cs, idx);
            }
            idx -= len;
        }
        return LDO_NONE;
    }
    // lanes: 64 chains of the working order at a time; the chain holding
    // domain idx is the first whose running length total exceeds idx
    for (int base = 0; base < n; base += 64) {
        int w = base + plane();
        int cs = 0, len = 0;
        if (w < n) {
            cs = order[w];
            len = ch[cs].len;
        }
        int incl = pscan(len);
        uint64_t hit = pballot(w < n && idx < incl && idx >= incl - len);
        if (hit) {
            int src = __builtin_ctzll(hit);
            return slot(pshfl(cs, src), idx - (pshfl(incl, src) - pshfl(len, src)));
        }
        idx -= pshfl(incl, 63);
    }
    return LDO_NONE;
}

// MCMovetype::test_acceptance (movetypes.cpp:142-158)
DEVIN bool Eng::test_acceptance(double p_ratio) {
    double p_accept = fmin(1.0, p_ratio) * S->modifier;
    if (p_accept == 1) {
        return true;
    }
    return p_accept > uniform_real();
}

// MCMovetype::scan_for_scaffold_domain (movetypes.cpp:191-232) as an
// explicit depth-first search; part[] holds the participating chains
DEVNI bool Eng::scan_for_scaffold(int dom0, int excl_cs) {
    for (int c = plane(); c < nchain; c += pwidth()) {
        part[c] = 0;
    }
    if (excl_cs >= 0) {
        part[excl_cs] = 1;
    }
    int sp = 0;
    sframes[sp].growth = (int16_t)dom0;
    sframes[sp].idx = 0;
    sp++;
    part[cslot_of(dom0)] = 1;
    while (sp > 0) {
        NetFrame& f = sframes[sp - 1];
        int c = cslot_of(f.growth);
        if (f.idx >= ch[c].len) {
            sp--;
            continue;
        }
        int cur = slot(c, f.idx);
        f.idx++;
        if (cur == f.growth) {
            continue;
        }
        int b = dom[cur].bound;
        if (b < 0) {
            continue;
        }
        int bc = cslot_of(b);
        if (bc == c) {
            continue;
        }
        if (bc == 0) {
            return true;
        }
        if (part[bc]) {
            continue;
        }
        part[bc] = 1;
        if (sp >= nchain) {
            fail(ST_CAPACITY);
            return true;
        }
        sframes[sp].growth = (int16_t)b;
        sframes[sp].idx = 0;
        sp++;
    }
    return false;
}

// MCMovetype::staple_is_connector (movetypes.cpp:160-177)
DEVNI bool Eng::staple_is_connector(int cs) {
    for (int d = 0; d < ch[cs].len; d++) {
        int s = slot(cs, d);
        if (dom[s].state != OCC_UNBOUND) {
            int b = dom[s].bound;
            if (cslot_of(b) == 0) {
                continue;
            }
            if (!scan_for_scaffold(b, cs)) {
                return true;
            }
        }
    }
    return false;
}

// RegrowthMCMovetype::num_bound_staple_domains (movetypes.cpp:395-405)
DEVIN int Eng::num_bound_staple_domains(int cs) {
    int n = 0;
    for (int d = 0; d < ch[cs].len; d++) {
        int stt = dom[slot(cs, d)].state;
        if (stt == OCC_BOUND || stt == OCC_MISBOUND) {
            n++;
        }
    }
    return n;
}

// MCMovetype::reset_origami (movetypes.cpp:53-86)
DEVNI void Eng::reset_origami() {
    for (int k = 0; k < S->n_assigned; k++) {
        unassign_domain(assigned[k]);
    }
    for (int k = 0; k < S->n_added; k++) {
        delete_chain(S->added[k]);
        S->current_c_i -= 1;
    }
    for (int k = 0; k < S->n_modified; k++) {
        int s = modified[k];
        set_checked_domain_config(s, i3(prev[s].x, prev[s].y, prev[s].z), prev[s].ore);
    }
    S->violated = 0;
}

// RegrowthMCMovetype::set_growth_point (movetypes.cpp:322-341)
DEVIN double Eng::set_growth_point(int gnew, int gold) {
    double de = set_domain_config(gnew, pos(gold), dom[gold].ore ^ 1);
    if (S->violated) {
        S->rejected = 1;
    }
    else {
        push_assigned(gnew);
    }
    return de;
}

// RegrowthMCMovetype::grow_staple (movetypes.cpp:343-370)
DEVNI void Eng::grow_staple(int cs, int d_i) {
    int len = ch[cs].len;
    int first = slot(cs, d_i);
    if (len - d_i > 1) {
        grow_chain_seq(first, len - d_i, 1);
    }
    if (S->rejected) {
        return;
    }
    if (d_i + 1 > 1) {
        grow_chain_seq(first, d_i + 1, -1);
    }
}

// grow_chain of the Metropolis (met_movetypes.cpp:50-67) and CB staple
// (cb_movetypes.cpp:354-361) moves over the domains first, first+dir, ...
DEVNI void Eng::grow_chain_seq(int first, int count, int dir) {
    int prev_d = first;
    for (int i = 1; i < count; i++) {
        int d = walk(prev_d, dir);
        if (S->kind == LDO_MOVE_KIND_MET) {
            I3 p = add3(pos(prev_d), unit3(uniform_int(0, 5)));
            int o = uniform_int(0, 5);
            S->delta_e += set_domain_config(d, p, o);
            if (S->violated) {
                S->rejected = 1;
                break;
            }
            push_assigned(d);
        }
        else {
            cb_select_and_set(d, prev_d);
            if (S->rejected) {
                break;
            }
        }
        prev_d = d;
    }
}

// OrientationRotationMCMovetype::internal_attempt_move (orientation_movetype.cpp:30-65)
DEVNI bool Eng::orientation_move() {
    bool accepted = false;
    int d = select_random_domain();
    if (d < 0) {
        fail(ST_MISUSE);
        return false;
    }
    int o_new = uniform_int(0, 5);
    int stt = dom[d].state;
    if (stt == OCC_BOUND || stt == OCC_MISBOUND) {
        double de = 0;
        int o_old = dom[d].ore;
        int b = dom[d].bound;
        de += unassign_domain(b);
        set_domain_orientation(d, o_new);
        I3 p = pos(d);
        de += set_domain_config(b, p, o_new ^ 1);
        if (!S->violated) {
            accepted = test_acceptance(exp(-de));
        }
        if (!accepted) {
            unassign_domain(b);
            set_domain_orientation(d, o_old);
            set_checked_domain_config(b, p, o_old ^ 1);
        }
    }
    else {
        set_domain_orientation(d, o_new);
        accepted = true;
    }
    return accepted;
}

// MetStapleExchangeMCMovetype::insert_staple + staple_insertion_accepted
// (met_movetypes.cpp:303-359, 212-253); exchange multipliers are 1
DEVNI bool Eng::insert_staple() {
    int ntypes = P->L.ntypes;
    int ident = uniform_int(1, ntypes - 1);
    S->trk_type = ident;
    if (S->num_staples == P->max_total_staples) {
        return false;
    }
    if (tcount[ident] == P->max_type_staples) {
        return false;
    }
    int cs = add_chain(ident);
    if (cs < 0) {
        return false;
    }
    S->delta_e += init_e;
    S->added[S->n_added++] = cs;
    int len = ch[cs].len;
    int gi = uniform_int(0, len - 1);
    int gnew = slot(cs, gi);
    int gold = select_random_domain();
    int guard = 0;
    while (gold < 0 || cslot_of(gold) == cs) {
        gold = select_random_domain();
        if (++guard > (1 << 20) || S->status != ST_OK) {
            fail(ST_MISUSE);
            return false;
        }
    }
    S->delta_e += set_growth_point(gnew, gold);
    if (S->rejected) {
        return false;
    }
    grow_staple(cs, gi);
    if (S->rejected) {
        return false;
    }
    S->delta_e += calc_move_bias(0); // add_external_bias (met_movetypes.cpp:69-75, 216)
    double boltz = exp(-S->delta_e);
    int n_new = tcount[ident];
    double pratio = (double)len / 6.0 / n_new * boltz;
    pratio *= S->insertion_sites * P->fugacity;
    pratio /= num_bound_staple_domains(cs);
    return test_acceptance(pratio);
}

// MetStapleExchangeMCMovetype::delete_staple + staple_deletion_accepted
// (met_movetypes.cpp:361-403, 255-301)
DEVNI bool Eng::delete_staple() {
    int ntypes = P->L.ntypes;
    int ident = uniform_int(1, ntypes - 1);
    S->trk_type = ident;
    int n = tcount[ident];
    if (n == 0) {
        S->rejected = 1;
        return false;
    }
    int cs = tlist[ident * P->L.tcap + uniform_int(0, n - 1)];
    if (staple_is_connector(cs)) {
        return false;
    }
    int n_bd = num_bound_staple_domains(cs);
    int len = ch[cs].len;
    for (int d = 0; d < len; d++) {
        int s = slot(cs, d);
        prev[s].x = dom[s].x;
        prev[s].y = dom[s].y;
        prev[s].z = dom[s].z;
        prev[s].ore = dom[s].ore;
        push_modified(s);
        S->delta_e += unassign_domain(s);
    }
    S->delta_e -= init_e;
    // add_external_bias with the staple count reduced by one
    // (met_movetypes.cpp:259-261)
    S->delta_e += calc_move_bias(-1);
    double boltz = exp(-S->delta_e);
    int ni = tcount[ident];
    double pratio = ni * 6.0 / (double)len * boltz;
    pratio /= P->fugacity * (double)(S->insertion_sites - len);
    pratio *= n_bd;
    bool accepted = test_acceptance(pratio);
    if (accepted) {
        delete_chain(cs);
    }
    return accepted;
}

// CBMCMovetype::calc_biases + calc_bias + select_and_set_config
// (cb_movetypes.cpp:58-166, 363-387, 486-539)
DEVNI void Eng::cb_select_and_set(int d, int prev_d) {
    I3 pp = pos(prev_d);
    // Per trial site v (the six unit vectors in order): whether it is a
    // candidate, its weight, the orientation it forces (-1: drawn after the
    // selection) and its unbound occupant (LDO_NONE: an empty site).  The
    // loops over v are unrolled, so these stay in registers; visiting the
    // candidates in v order sums and selects in the reference's order.
    bool cand[6];
    int core[6];
    int cub[6];
    double w[6];
    // lanes: lane v < 6 looks up trial site v.  The checks below place d
    // and take it off again, which leaves every site's occupant as it was,
    // so the six lookups can all be made first (d is unassigned here; an
    // assigned d is looked up site by site).
    bool looked = par && dom[d].state == OCC_UNASSIGNED;
    int occ_l = OCC_UNASSIGNED, ub_l = -1;
    if (looked && plane() < 6) {
        int hi = hfind(pkey(add3(pp, unit3(plane()))));
        if (hi >= 0) {
            occ_l = (int)((hval[hi] >> 16) & 3u);
            ub_l = (int)(int16_t)(hval[hi] & 0xFFFFu);
        }
    }
#pragma unroll
    for (int v = 0; v < 6; v++) {
        I3 pn = add3(pp, unit3(v));
        cand[v] = false;
        core[v] = -1;
        cub[v] = LDO_NONE;
        w[v] = 0;
        int occ = looked ? pshfl(occ_l, v) : occupancy(pn);
        if (occ == OCC_BOUND || occ == OCC_MISBOUND) {
            continue;
        }
        if (occ == OCC_UNBOUND) {
            int ub;
            if (looked) {
                // unbound_at's check of the occupant slot
                ub = pshfl(ub_l, v);
                if (ub < 0 || ub >= nslots) {
                    fail(ST_CONSTRAINT);
                    ub = 0;
                }
            }
            else {
                ub = unbound_at(pn);
            }
            int on = dom[ub].ore ^ 1;
            double de = check_domain_constraints(d, pn, on);
            if (!S->violated) {
                cand[v] = true;
                core[v] = on;
                cub[v] = ub;
                w[v] = exp(-de);
            }
            else {
                S->violated = 0;
            }
        }
        else {
            cand[v] = true;
            if (dom[d].state == OCC_UNASSIGNED) {
                // check_domain_constraints at an empty site, with the
                // occupancy already known: the domain is left unassigned at
                // pn and weighs 6 exp(-0) = 6
                dom[d].x = pn.x;
                dom[d].y = pn.y;
                dom[d].z = pn.z;
                dom[d].ore = 0;
                S->site_empty++;
                w[v] = 6.0;
            }
            else {
                double de = check_domain_constraints(d, pn, 0);
                w[v] = 6 * exp(-de);
            }
        }
    }
    double sum = 0;
    if (S->kind == LDO_MOVE_KIND_CTCB) {
#pragma unroll
        for (int v = 0; v < 6; v++) {
            if (!cand[v]) {
                continue;
            }
            I3 pn = add3(pp, unit3(v));
            if (cub[v] >= 0) {
                bool same_chain = cslot_of(cub[v]) == cslot_of(d);
                bool endpoint = cp_endpoint_reached(d, pn);
                if (!(same_chain || endpoint)) {
                    w[v] = 0;
                }
            }
            if (!cp_walks_remain(d, pn)) {
                w[v] = 0;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 6; v++) {
        if (cand[v]) {
            sum += w[v];
        }
    }
    if (sum == 0) {
        S->rejected = 1;
        return;
    }
    S->bias *= sum;
    if (!S->regrow_old) {
        double u = uniform_real();
        double cum = 0;
        int sel = -1;
        int o = -1;
#pragma unroll
        for (int v = 0; v < 6; v++) {
            if (sel < 0 && cand[v]) {
                cum += w[v] / sum;
                if (u < cum) {
                    sel = v;
                    o = core[v];
                }
            }
        }
        I3 p = sel >= 0 ? add3(pp, unit3(sel)) : i3(0, 0, 0);
        if (o < 0) {
            o = uniform_int(0, 5);
        }
        set_checked_domain_config(d, p, o);
    }
    else {
        set_checked_domain_config(d, i3(old[d].x, old[d].y, old[d].z), old[d].ore);
    }
    push_assigned(d);
}

// CBMCMovetype::unassign_domains (cb_movetypes.cpp:200-213) over the domains
// of chain cs with d in [from, to)
DEVIN void Eng::cb_unassign_chain(int cs, int from, int to) {
    for (int d = from; d < to; d++) {
        int s = slot(cs, d);
        prev[s].x = dom[s].x;
        prev[s].y = dom[s].y;
        prev[s].z = dom[s].z;
        prev[s].ore = dom[s].ore;
        push_modified(s);
        unassign_domain(s);
    }
}

// CBMCMovetype::setup_for_regrow_old (cb_movetypes.cpp:215-225)
DEVIN void Eng::cb_setup_regrow_old() {
    S->regrow_old = 1;
    S->new_bias = S->bias;
    S->new_modifier = S->modifier;
    S->bias = 1;
    for (int k = 0; k < S->n_modified; k++) {
        int s = modified[k];
        old[s] = prev[s];
    }
    S->n_modified = 0;
    S->n_assigned = 0;
}

// CBMCMovetype::test_cb_acceptance (cb_movetypes.cpp:182-198)
DEVNI bool Eng::cb_test_acceptance() {
    double ratio = S->new_bias / S->bias;
    if (test_acceptance(ratio)) {
        reset_origami();
        calc_move_bias(0);
        return true;
    }
    S->n_modified = 0;
    S->n_assigned = 0;
    return false;
}

// CBMCMovetype::set_old_growth_point (cb_movetypes.cpp:168-180)
DEVIN void Eng::cb_set_old_growth_point(int gnew, int gold) {
    double de = set_checked_domain_config(gnew, pos(gold), old[gnew].ore);
    S->bias *= exp(-de);
    push_assigned(gnew);
}

// CBStapleRegrowthMCMovetype::internal_attempt_move (cb_movetypes.cpp:292-348)
DEVNI bool Eng::cb_staple_move() {
    if (S->num_staples == 0) {
        S->trk_nost = 1;
        return false;
    }
    S->trk_nost = 0;
    int w = uniform_int(1, S->num_staples);
    int cs = order[w];
    S->trk_type = ch[cs].ident;
    if (staple_is_connector(cs)) {
        return false;
    }
    int len = ch[cs].len;
    // (staple domain, partner) pairs of the old configuration, in the
    // replica block: a draw picks one by index
    int16_t* gnew_l = gpl;
    int16_t* gold_l = gpl + LDO_MAX_STAPLE_LEN;
    int nb = 0;
    for (int d = 0; d < len; d++) {
        int s = slot(cs, d);
        int b = dom[s].bound;
        if (b >= 0 && cslot_of(b) != cs) {
            gnew_l[nb] = (int16_t)s;
            gold_l[nb] = (int16_t)b;
            nb++;
        }
    }
    if (nb == 0) {
        fail(ST_CONSTRAINT);
        return false;
    }
    S->bias *= nb;
    int k = uniform_int(0, nb - 1);
    cb_unassign_chain(cs, 0, len);
    double de = set_growth_point(gnew_l[k], gold_l[k]);
    S->bias *= exp(-de);
    if (!S->rejected) {
        grow_staple(cs, d_of(gnew_l[k]));
    }
    if (S->rejected) {
        return false;
    }
    S->bias /= num_bound_staple_domains(cs);
    S->bias *= exp(-calc_move_bias(0)); // add_external_bias (cb_movetypes.cpp:52-56)
    cb_setup_regrow_old();
    k = uniform_int(0, nb - 1);
    cb_unassign_chain(cs, 0, len);
    cb_set_old_growth_point(gnew_l[k], gold_l[k]);
    if (!S->rejected) {
        grow_staple(cs, d_of(gnew_l[k]));
    }
    S->modifier = S->new_modifier;
    return cb_test_acceptance();
}

// ------------------------------------------------- constraint points
// Constraintpoints::reset_internal (top_constraint_points.cpp:191-205)
DEVNI void Eng::cp_reset() {
    // -1 in every int8 and int16 entry is the byte pattern 0xff
    fill_section(seg, nslots, 0xffffffffu);
    fill_section(gp, 2 * (int64_t)nslots, 0xffffffffu);
    fill_section(stem, 2 * (int64_t)nslots, 0xffffffffu);
    fill_section(inact, 2 * (int64_t)nslots, 0xffffffffu);
    fill_section(netseg, nslots, 0xffffffffu);
    fill_section(netds_flag, nslots, 0u);
    fill_section(checked, nchain, 0u);
    fill_section(netcs, nchain, 0u);
    fill_section(regrow, nchain, 0u);
    fill_section(dir0, nchain, 0u);
    fill_section(dir1, nchain, 0u);
    S->n_active = 0;
    S->n_active0 = 0;
    S->n_rgs = 0;
    S->n_jsegs = 0;
    S->n_jstems = 0;
    n_netds = 0;
}

// Constraintpoints::add_active_endpoint (:264-278)
DEVIN void Eng::cp_add_active(int s, I3 p, int sg) {
    if (S->n_active >= P->L.ecap) {
        fail(ST_CAPACITY);
        return;
    }
    Endpoint& e = active[S->n_active++];
    e.cslot = (int16_t)cslot_of(s);
    e.seg = sg;
    e.d = (int16_t)d_of(s);
    e.x = p.x;
    e.y = p.y;
    e.z = p.z;
}

// Constraintpoints::remove_active_endpoint (:299-320).  With rec, the
// removed endpoints (get_erased_endpoints) go on top of the erased-endpoint
// stack as one level.
DEVIN void Eng::cp_remove_active(int s, bool rec) {
    int sg = seg[s];
    int cs = cslot_of(s), dd = d_of(s);
    int n = S->n_active;
    int k = 0, r = 0;
    int top = rec ? S->er_top : 0;
    // lanes: order-preserving compaction, 64 entries at a time (an entry
    // moves to an index at or below its own, after every lane has read)
    for (int base = 0; base < n; base += pwidth()) {
        int i = base + plane();
        bool keep = false, drop = false;
        Endpoint e;
        if (i < n) {
            e = active[i];
            keep = !(e.cslot == cs && e.seg == sg && e.d == dd);
            drop = !keep;
        }
        uint64_t below = (1ull << plane()) - 1;
        uint64_t m = pballot(keep);
        uint64_t md = pballot(drop);
        if (keep) {
            active[k + __builtin_popcountll(m & below)] = e;
        }
        if (rec && drop) {
            int j = top + r + __builtin_popcountll(md & below);
            if (j < P->L.ecap) {
                erstk[j] = e;
            }
        }
        k += __builtin_popcountll(m);
        r += __builtin_popcountll(md);
    }
    S->n_active = k;
    if (rec) {
        if (top + r > P->L.ecap || S->er_depth >= nslots) {
            fail(ST_CAPACITY);
            return;
        }
        ercnt[S->er_depth++] = (int16_t)r;
        S->er_top = top + r;
    }
}

// Constraintpoints::remove_activated_endpoint (:322-338): the first active
// endpoint of the inactive-endpoint domain paired with s
DEVIN void Eng::cp_remove_activated(int s) {
    int e = inact[s];
    if (e < 0) {
        return;
    }
    int sg = seg[e];
    int cs = cslot_of(e), dd = d_of(e);
    int n = S->n_active;
    int j = -1;
    for (int i = 0; i < n; i++) {
        const Endpoint& a = active[i];
        if (a.cslot == cs && a.seg == sg && a.d == dd) {
            j = i;
            break;
        }
    }
    if (j < 0) {
        return;
    }
    for (int i = j; i + 1 < n; i++) {
        active[i] = active[i + 1];
    }
    S->n_active = n - 1;
}

// Constraintpoints::update_endpoints (:340-349)
DEVIN void Eng::cp_update_endpoints(int s) {
    cp_remove_active(s);
    int e = inact[s];
    if (e >= 0) {
        cp_add_active(e, pos(s), seg[e]);
    }
}

// Constraintpoints::endpoint_reached (:359-372)
DEVIN bool Eng::cp_endpoint_reached(int s, I3 p) {
    int sg = seg[s];
    int cs = cslot_of(s), dd = d_of(s);
    int n = S->n_active;
    bool hit = false;
    for (int i = plane(); i < n && !(hit && !par); i += pwidth()) {
        const Endpoint& e = active[i];
        if (e.cslot == cs && e.seg == sg && e.d == dd && e.x == p.x && e.y == p.y &&
            e.z == p.z) {
            hit = true;
        }
    }
    return pany(hit);
}

// m_domain_to_dir of (chain, segment): the scaffold segments of a
// non-contiguous selection have their own (0 for an empty one, which the
// reference never sets)
DEVIN int Eng::cp_seg_dir(int cs, int sg) {
    if (cs == 0 && S->n_jsegs > 0) {
        return sdir[sg];
    }
    return sg == 0 ? dir0[cs] : dir1[cs];
}

// Constraintpoints::walks_remain (:397-442, 571-601).  The ideal-walk count
// is only tested against zero, which holds exactly when the Manhattan
// distance exceeds the remaining steps or differs from them in parity
// (ideal_random_walk.cpp:35-41).  Stem domains look up their segment pair in
// m_stemd_to_segs, which only holds the scaffold stems of a non-contiguous
// selection; other stems have no walk constraint.
DEVIN bool Eng::cp_walks_remain(int s, I3 p) {
    if (stem[s] >= 0) {
        if (S->n_jsegs == 0 || cslot_of(s) != 0 || stemseg[s] < 0) {
            return true;
        }
        return cp_walks_seg(s, p, stemseg[s]) && cp_walks_seg(s, p, stemseg[s] + 1);
    }
    return cp_walks_seg(s, p, seg[s]);
}

DEVIN bool Eng::cp_walks_seg(int s, I3 p, int sg) {
    int cs = cslot_of(s), dd = d_of(s);
    int dir = cp_seg_dir(cs, sg);
    int n = S->n_active;
    bool bad = false;
    for (int i = plane(); i < n && !(bad && !par); i += pwidth()) {
        const Endpoint& e = active[i];
        if (e.cslot != cs || e.seg != sg) {
            continue;
        }
        int steps;
        int ed = e.d;
        if (cyc && cs == 0) {
            if (dir > 0 && ed < dd) {
                steps = ns + ed - dd;
            }
            else if (dir < 0 && ed > dd) {
                steps = dd + ns - ed;
            }
            else if (dir == 0 || ed == dd) {
                steps = 0;
            }
            else {
                steps = ed > dd ? ed - dd : dd - ed;
            }
        }
        else {
            steps = ed > dd ? ed - dd : dd - ed;
        }
        int ax = e.x - p.x, ay = e.y - p.y, az = e.z - p.z;
        int l1 = (ax < 0 ? -ax : ax) + (ay < 0 ? -ay : ay) + (az < 0 ? -az : az);
        if (l1 > steps || ((steps - l1) & 1)) {
            bad = true;
        }
    }
    return !pany(bad);
}

// StapleNetwork::scan_network + scan_staple_topology + make_staple_stack
// (top_constraint_points.cpp:36-44, 78-170).  The recursion into bound
// staples runs on frames[]; gps/iaes hold (domain, domain) pairs.
DEVNI void Eng::cp_scan_network(int d0) {
    for (int i = 0; i < n_netds; i++) {
        int s = netds[i];
        netseg[s] = -1;
        netds_flag[s] = 0;
        netcs[cslot_of(s)] = 0;
    }
    net_external = false;
    n_gps = 0;
    n_iaes = 0;
    n_netds = 0;
    netcs[0] = 1;
    int sp = 0;
    int growth = d0;
    // enter a recursion level for `growth`
    {
        int ci = cslot_of(growth);
        netcs[ci] = 1;
        netds[n_netds++] = (int16_t)growth;
        netds_flag[growth] = 1;
        int gd = d_of(growth), len = ch[ci].len;
        netseg[growth] = 0;
        for (int k = gd + 1; k < len; k++) {
            netseg[slot(ci, k)] = 0;
        }
        for (int k = gd - 1; k >= 0; k--) {
            netseg[slot(ci, k)] = 1;
        }
        frames[sp].growth = (int16_t)growth;
        frames[sp].idx = 0;
        sp++;
    }
    while (sp > 0) {
        NetFrame& f = frames[sp - 1];
        int ci = cslot_of(f.growth);
        int len = ch[ci].len;
        int gd = d_of(f.growth);
        if (f.idx >= len - 1) {
            sp--;
            continue;
        }
        int k = f.idx++;
        int dd = (k < len - 1 - gd) ? slot(ci, gd + 1 + k) : slot(ci, gd - 1 - (k - (len - 1 - gd)));
        if (n_netds >= nslots || n_gps >= nslots || n_iaes >= nslots) {
            fail(ST_CAPACITY);
            return;
        }
        netds[n_netds++] = (int16_t)dd;
        netds_flag[dd] = 1;
        int b = dom[dd].bound;
        if (b < 0 || cslot_of(b) == ci) {
            continue;
        }
        int bci = cslot_of(b);
        if (netcs[bci]) {
            bool external = false;
            if (bci == 0) {
                external = !cp_selected(b);
            }
            if (!external) {
                if (netds_flag[b]) {
                    iaes[2 * n_iaes] = (int16_t)b;
                    iaes[2 * n_iaes + 1] = (int16_t)dd;
                }
                else {
                    iaes[2 * n_iaes] = (int16_t)dd;
                    iaes[2 * n_iaes + 1] = (int16_t)b;
                }
                n_iaes++;
            }
            if (!net_external && external) {
                net_external = true;
            }
        }
        else {
            gps[2 * n_gps] = (int16_t)dd;
            gps[2 * n_gps + 1] = (int16_t)b;
            n_gps++;
            if (sp >= nchain) {
                fail(ST_CAPACITY);
                return;
            }
            // recurse into the bound staple
            int cj = bci;
            netcs[cj] = 1;
            netds[n_netds++] = (int16_t)b;
            netds_flag[b] = 1;
            int bd0 = d_of(b), blen = ch[cj].len;
            netseg[b] = 0;
            for (int q = bd0 + 1; q < blen; q++) {
                netseg[slot(cj, q)] = 0;
            }
            for (int q = bd0 - 1; q >= 0; q--) {
                netseg[slot(cj, q)] = 1;
            }
            frames[sp].growth = (int16_t)b;
            frames[sp].idx = 0;
            sp++;
        }
    }
}

// whether scaffold domain b is among the selected (m_scaffold_domains)
DEVIN bool Eng::cp_selected(int b) {
    if (S->n_jsegs > 0) {
        return jdom[b] != 0;
    }
    int rel = (d_of(b) - d_of(seg_first)) * seg_dir;
    if (P->cyclic) {
        rel = ((rel % ns) + ns) % ns;
    }
    return rel >= 0 && rel < seg_count;
}

// Constraintpoints::find_growthpoints_endpoints (top_constraint_points.cpp:
// 454-569) for scaffold segment sg: first, first+dir, ... (count domains)
DEVNI void Eng::cp_find(int first, int count, int dir, int sg) {
    // CTRG grows Constraintpoints::m_d_stack (domains_to_be_regrown): each
    // segment domain, then the domain stack of the network it roots
    bool rec = S->kind == LDO_MOVE_KIND_CTRG || S->kind == LDO_MOVE_KIND_CTRG_JUMP ||
               S->n_jsegs > 0;
    if (S->n_jsegs == 0) {
        dir0[0] = (int8_t)dir;
    }
    int d = first;
    for (int i = 0; i < count; i++) {
        if (i > 0) {
            d = walk(d, dir);
        }
        if (rec) {
            if (S->n_rgs >= nslots) {
                fail(ST_CAPACITY);
                return;
            }
            rgs[S->n_rgs++] = (int16_t)d;
        }
        seg[d] = (int8_t)sg;
        int b = dom[d].bound;
        if (b < 0 || cslot_of(b) == 0 || checked[cslot_of(b)]) {
            continue;
        }
        cp_scan_network(b);
        gps[2 * n_gps] = (int16_t)d;
        gps[2 * n_gps + 1] = (int16_t)b;
        n_gps++;
        if (net_external) {
            for (int g = 0; g < n_gps; g++) {
                int g1 = gps[2 * g];
                if (cslot_of(g1) == 0) {
                    cp_add_active(g1, pos(g1), sg);
                }
            }
            // the reference keys these endpoints by the staple domain (a1),
            // where no scaffold growth looks for them
            bool on_scaffold = (P->flags & LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS) != 0;
            for (int a = 0; a < n_iaes; a++) {
                int a1 = iaes[2 * a], a2 = iaes[2 * a + 1];
                if (cslot_of(a2) == 0) {
                    cp_add_active(on_scaffold ? a2 : a1, pos(a2), sg);
                }
            }
        }
        else {
            for (int g = 0; g < n_gps; g++) {
                gp[gps[2 * g]] = gps[2 * g + 1];
                stem[gps[2 * g + 1]] = gps[2 * g];
            }
            for (int a = 0; a < n_iaes; a++) {
                inact[iaes[2 * a]] = iaes[2 * a + 1];
            }
            if (rec) {
                if (S->n_rgs + n_netds > nslots) {
                    fail(ST_CAPACITY);
                    return;
                }
                for (int q = 0; q < n_netds; q++) {
                    rgs[S->n_rgs++] = netds[q];
                }
            }
            for (int q = 0; q < n_netds; q++) {
                int s = netds[q];
                int c = cslot_of(s);
                if (c != 0) {
                    regrow[c] = 1;
                    dir0[c] = 1;
                    dir1[c] = -1;
                }
                if (seg[s] < 0) {
                    seg[s] = netseg[s];
                }
            }
        }
        for (int q = 0; q < n_netds; q++) {
            checked[cslot_of(netds[q])] = 1;
        }
        checked[0] = 1;
    }
}

// the end of calculate_constraintpoints (top_constraint_points.cpp:220,
// 247): m_initial_active_endpoints = m_active_endpoints
DEVNI void Eng::cp_save_initial() {
    for (int i = plane(), n = S->n_active; i < n; i += pwidth()) {
        active0[i] = active[i];
    }
    S->n_active0 = S->n_active;
}

// CTCBRegrowthMCMovetype::grow_chain (cb_movetypes.cpp:463-484) and
// grow_staple_and_update_endpoints (:541-562), interleaved through an
// explicit stack.  Growing starts with the scaffold segment and descends
// into staples at growthpoints; any rejection unwinds the whole growth.
DEVNI void Eng::ctcb_grow_chain_scaffold(int first, int count, int dir, bool first_gp) {
    GrowFrame* stk = gstack;
    int sp = 0;
    if (count > 1) {
        GrowFrame f;
        f.first = (int16_t)first;
        f.count = (int16_t)count;
        f.i = 1;
        f.prev = (int16_t)first;
        f.dir = (int8_t)dir;
        f.kind = 0;
        f.phase = 0;
        stk[sp++] = f;
    }
    if (first_gp && gp[first] >= 0) {
        GrowFrame f;
        f.gold = (int16_t)first;
        f.kind = 1;
        f.phase = 0;
        stk[sp++] = f;
    }
    while (sp > 0 && !S->rejected && S->status == ST_OK) {
        GrowFrame& F = stk[sp - 1];
        if (F.kind == 0) {
            if (F.i >= F.count) {
                sp--;
                continue;
            }
            int d = walk(F.prev, F.dir);
            if (seg[d] < 0) {
                seg[d] = 0; // get_dir's map insertion (top_constraint_points.cpp:179-185)
            }
            cb_select_and_set(d, F.prev);
            if (S->rejected) {
                break;
            }
            cp_update_endpoints(d);
            F.prev = (int16_t)d;
            F.i++;
            if (gp[d] >= 0) {
                if (sp >= LDO_GROW_DEPTH) {
                    fail(ST_CAPACITY);
                    break;
                }
                GrowFrame g;
                g.gold = (int16_t)d;
                g.kind = 1;
                g.phase = 0;
                stk[sp++] = g;
            }
            continue;
        }
        // staple growth from growthpoint F.gold
        if (F.phase == 0) {
            int gnew = gp[F.gold];
            int cs = cslot_of(gnew);
            if (cs == 0) {
                sp--;
                continue;
            }
            if (S->regrow_old) {
                cb_set_old_growth_point(gnew, F.gold);
            }
            else {
                double de = set_growth_point(gnew, F.gold);
                S->bias *= exp(-de);
            }
            if (S->rejected) {
                break;
            }
            cp_update_endpoints(gnew);
            F.gnew = (int16_t)gnew;
            F.phase = 1;
            int len = ch[cs].len, di = d_of(gnew);
            if (len - di > 1) {
                if (sp >= LDO_GROW_DEPTH) {
                    fail(ST_CAPACITY);
                    break;
                }
                GrowFrame l;
                l.first = (int16_t)gnew;
                l.count = (int16_t)(len - di);
                l.i = 1;
                l.prev = (int16_t)gnew;
                l.dir = 1;
                l.kind = 0;
                l.phase = 0;
                stk[sp++] = l;
            }
            continue;
        }
        if (F.phase == 1) {
            F.phase = 2;
            int di = d_of(F.gnew);
            if (di + 1 > 1) {
                if (sp >= LDO_GROW_DEPTH) {
                    fail(ST_CAPACITY);
                    break;
                }
                GrowFrame l;
                l.first = F.gnew;
                l.count = (int16_t)(di + 1);
                l.i = 1;
                l.prev = F.gnew;
                l.dir = -1;
                l.kind = 0;
                l.phase = 0;
                stk[sp++] = l;
            }
            continue;
        }
        sp--;
    }
}

// CTRegrowthMCMovetype::select_indices (movetypes.cpp:469-510): the
// segment first, first + dir, ... (count domains), and the active endpoint
// of the domain past its end
DEVIN bool Eng::ct_select_indices(int max_regrowth, int& first, int& count, int& dir) {
    int guard = 0;
    while (true) {
        int max_len = ns < max_regrowth ? ns : max_regrowth;
        int sel = uniform_int(2, max_len);
        int start = uniform_int(0, ns - 1);
        dir = uniform_int(0, 1);
        if (dir == 0) {
            dir = -1;
        }
        int start_s = slot(0, start);
        int c = start_s;
        int nf = 0;
        while (c >= 0 && nf != sel) {
            nf++;
            c = walk(c, dir);
        }
        int cur_end = c;
        first = start_s;
        int nb = 0;
        int b = walk(start_s, -dir);
        while (b >= 0 && nf + nb != sel) {
            first = b;
            nb++;
            b = walk(b, -dir);
        }
        count = nf + nb;
        if (count < 2) {
            if (++guard > 1000 || S->status != ST_OK) {
                fail(ST_MISUSE);
                return false;
            }
            continue;
        }
        if (cur_end >= 0) {
            cp_add_active(cur_end, pos(cur_end), 0);
        }
        return true;
    }
}

// CTCBScaffoldRegrowthMCMovetype::internal_attempt_move (cb_movetypes.cpp:680-747)
DEVNI bool Eng::ctcb_scaffold_move(int max_regrowth) {
    S->regrow_old = 0;
    cp_reset();
    int first, count, dir;
    if (!ct_select_indices(max_regrowth, first, count, dir)) {
        return false;
    }
    seg_first = first;
    seg_count = count;
    seg_dir = dir;
    S->trk_len = count;
    cp_find(first, count, dir, 0);
    cp_save_initial();
    bool whole_ring = P->cyclic && count == ns;
    if (!whole_ring) {
        cp_remove_active(first);
    }
    // staples to be regrown, in ascending unique-index order (std::set<int>);
    // the list borrows the connector-scan frames, which are idle here
    int16_t* sl = (int16_t*)sframes;
    int nst = 0;
    for (int c2 = 1; c2 < nchain; c2++) {
        if (regrow[c2] && ch[c2].ident >= 0) {
            int k = nst++;
            while (k > 0 && ch[sl[k - 1]].uid > ch[c2].uid) {
                sl[k] = sl[k - 1];
                k--;
            }
            sl[k] = (int16_t)c2;
        }
    }
    S->trk_nst = nst;
    for (int phase = 0; phase < 2; phase++) {
        if (phase == 1) {
            cb_setup_regrow_old();
            for (int i = plane(), n = S->n_active0; i < n; i += pwidth()) {
                active[i] = active0[i];
            }
            S->n_active = S->n_active0;
            if (!whole_ring) {
                cp_remove_active(first);
            }
        }
        int d = first;
        for (int i = 1; i < count; i++) {
            d = walk(d, dir);
            prev[d].x = dom[d].x;
            prev[d].y = dom[d].y;
            prev[d].z = dom[d].z;
            prev[d].ore = dom[d].ore;
            push_modified(d);
            unassign_domain(d);
        }
        for (int k = 0; k < nst; k++) {
            cb_unassign_chain(sl[k], 0, ch[sl[k]].len);
        }
        ctcb_grow_chain_scaffold(first, count, dir);
        if (phase == 0 && S->rejected) {
            return false;
        }
        if (phase == 0) {
            S->bias *= exp(-calc_move_bias(0)); // add_external_bias (cb_movetypes.cpp:719)
        }
    }
    S->modifier = 1;
    return cb_test_acceptance();
}

// ------------------------------------------------- CTRG scaffold regrowth
// CTRGRegrowthMCMovetype (rg_movetypes.cpp) with
// CTRGScaffoldRegrowthMCMovetype::internal_attempt_move (:629-689).  The
// moveset's excluded staples are always none (simulation.cpp:450-452), so
// nothing is drawn for them and the excluded-staple test passes.  A trial
// configuration is one of the 36 (site vector, orientation vector) pairs of
// m_all_configs (:58, utility.hpp:168-177), index 6 * site + orientation;
// the configurations a level has left are a bit mask, whose set bits in
// ascending order are the reference's list after its erasures.

// Constraintpoints::get_dir (top_constraint_points.cpp:179-185)
DEVIN int Eng::rg_get_dir(int s) {
    int cs = cslot_of(s);
    int sg = seg[s];
    if (cs == 0 && S->n_jsegs > 0) {
        return sdir[sg < 0 ? 0 : sg];
    }
    return sg == 1 ? dir1[cs] : dir0[cs];
}

// unassign_and_save_domains (:87-104) with save and modify,
// unassign_domains (:122-129) with neither, and the third pass of the move
// (whose saved positions are the ones the old configuration keeps in `old`)
// with modify only
DEVNI double Eng::rg_unassign_and_save(bool save) {
    double de = 0;
    int n = S->n_rgs;
    for (int i = 0; i < n; i++) {
        int d = rgs[i];
        if (save) {
            prev[d].x = dom[d].x;
            prev[d].y = dom[d].y;
            prev[d].z = dom[d].z;
            prev[d].ore = dom[d].ore;
        }
        if (i > 0) {
            push_modified(d);
            de += unassign_domain(d);
        }
    }
    S->er_top = 0;
    S->er_depth = 0;
    return de;
}

// set_config (:233-241) and, with rec = RG_NONE, the
// set_checked_domain_config + update_endpoints of calc_weights (:388-392,
// 412-418).  RG_LIST (recoil growth) records the assignment for
// reset_origami; RG_DUP (feelers, calc_old_c_opens) only counts it: those
// records repeat domains the recoil growth recorded first, and their second
// unassignment in an accepting reset_origami changes nothing but the
// unassignment count.
DEVNI double Eng::rg_set_config(int d, I3 p, int o, int rec) {
    double de = set_checked_domain_config(d, p, o);
    if (rec == RG_LIST) {
        push_assigned(d);
    }
    else if (rec == RG_DUP) {
        rg_dup++;
    }
    cp_remove_active(d, true);
    int e = inact[d];
    if (e >= 0) {
        cp_add_active(e, pos(d), seg[e]);
    }
    return de;
}

// prepare_for_growth (:243-260)
DEVIN void Eng::rg_prepare_for_growth(RgCur& c) {
    c.di++;
    c.d = rgs[c.di];
    c.stemd = stem[c.d] >= 0;
    c.dir = rg_get_dir(c.d);
    c.att = 0;
    if (c.stemd) {
        c.maxatt = 1;
        c.avail = 0;
        c.ref = stem[c.d];
    }
    else {
        c.maxatt = rg_max_att;
        c.avail = LDO_RG_ALL;
        c.ref = walk(c.d, -c.dir);
    }
}

// prepare_for_regrowth (:262-286)
DEVNI double Eng::rg_prepare_for_regrowth(RgCur& c, int rec) {
    c.di--;
    c.d = rgs[c.di];
    c.dir = rg_get_dir(c.d);
    double de = unassign_domain(c.d);
    if (rec == RG_LIST) {
        S->n_assigned--;
    }
    else if (rec == RG_DUP) {
        rg_dup--;
    }
    rg_restore_endpoints(c.d);
    c.stemd = stem[c.d] >= 0;
    if (c.stemd) {
        c.maxatt = 1;
        c.att = 1;
        c.avail = 0;
        c.ref = stem[c.d];
    }
    else {
        c.maxatt = rg_max_att;
        c.att = rgatt[c.di];
        c.avail = rgavail[c.di];
        c.ref = walk(c.d, -c.dir);
    }
    return de;
}

// restore_endpoints (:288-296)
DEVNI void Eng::rg_restore_endpoints(int d) {
    cp_remove_activated(d);
    if (S->er_depth <= 0) {
        fail(ST_MISUSE);
        return;
    }
    int r = ercnt[--S->er_depth];
    int top = S->er_top - r;
    for (int k = 0; k < r; k++) {
        const Endpoint& e = erstk[top + k];
        cp_add_active(d, i3(e.x, e.y, e.z), seg[d]);
    }
    S->er_top = top;
}

// select_trial_config (:298-313)
DEVIN void Eng::rg_select(RgCur& c, I3& p, int& o) {
    if (c.stemd) {
        p = pos(c.ref);
        o = dom[c.ref].ore ^ 1;
        return;
    }
    int k = uniform_int(0, __builtin_popcountll(c.avail) - 1);
    int i = 0;
    for (uint64_t m = c.avail;; m &= m - 1) {
        if (k-- == 0) {
            i = __builtin_ctzll(m);
            break;
        }
    }
    c.avail &= ~(1ull << i);
    p = add3(pos(c.ref), unit3(i / 6));
    o = i % 6;
}

// calc_p_config_open (:315-343)
DEVNI double Eng::rg_p_open(const RgCur& c, I3 p, int o) {
    double de = check_domain_constraints(c.d, p, o);
    if (S->violated) {
        S->violated = 0;
        return 0;
    }
    if (!cp_walks_remain(c.d, p)) {
        return 0;
    }
    if (occupancy(p) == OCC_UNBOUND) {
        int ub = unbound_at(p);
        bool same_chain = cslot_of(ub) == cslot_of(c.d);
        bool endpoint = cp_endpoint_reached(c.d, p);
        if (!(same_chain || endpoint || c.stemd)) {
            return 0;
        }
    }
    return fmin(1.0, exp(-de));
}

// test_config_open (:345-361)
DEVIN bool Eng::rg_test_open(double p) {
    p = fmin(1.0, p);
    if (p == 1) {
        return true;
    }
    return p > uniform_real();
}

// recoil_regrow (:177-231)
DEVNI double Eng::rg_recoil_regrow() {
    double de = 0;
    int n = S->n_rgs;
    RgCur c;
    c.di = 0;
    c.d = rgs[0];
    c.dir = rg_get_dir(c.d);
    rgopen[0] = 1;
    rg_prepare_for_growth(c);
    int recoils = 0;
    while (S->status == ST_OK) {
        I3 p = i3(0, 0, 0);
        int o = 0;
        double pc = 0;
        bool open = false;
        while (!open && c.att != c.maxatt) {
            c.att++;
            rg_select(c, p, o);
            pc = rg_p_open(c, p, o);
            open = rg_test_open(pc);
        }
        if (open) {
            if (recoils != 0) {
                recoils--;
            }
            de += rg_set_config(c.d, p, o, RG_LIST);
            rgatt[c.di] = (int8_t)c.att;
            rgavail[c.di] = c.avail;
            rgopen[c.di] = pc;
            if (c.di == n - 1) {
                break;
            }
            rg_prepare_for_growth(c);
        }
        else {
            if (recoils == rg_max_recoils || c.di == 1) {
                S->rejected = 1;
                break;
            }
            recoils++;
            de += rg_prepare_for_regrowth(c, RG_LIST);
        }
    }
    return de;
}

// test_config_avail (:426-480): feeler growth past level c.di
DEVNI bool Eng::rg_test_avail(RgCur& c) {
    int n = S->n_rgs;
    int feels = 0;
    if (feels == rg_max_recoils || c.di == n - 1) {
        return true;
    }
    rg_prepare_for_growth(c);
    bool avail = false;
    while (S->status == ST_OK) {
        I3 p = i3(0, 0, 0);
        int o = 0;
        bool open = false;
        while (!open && c.att != c.maxatt) {
            c.att++;
            rg_select(c, p, o);
            double pc = rg_p_open(c, p, o);
            open = rg_test_open(pc);
        }
        if (open) {
            feels++;
            if (feels == rg_max_recoils || c.di == n - 1) {
                feels--;
                avail = true;
                break;
            }
            rg_set_config(c.d, p, o, RG_DUP);
            rgatt[c.di] = (int8_t)c.att;
            rgavail[c.di] = c.avail;
            rg_prepare_for_growth(c);
        }
        else {
            if (feels == 0) {
                avail = false;
                break;
            }
            feels--;
            rg_prepare_for_regrowth(c, RG_DUP);
        }
    }
    while (feels != 0) {
        c.di--;
        feels--;
        c.d = rgs[c.di];
        unassign_domain(c.d);
        rg_restore_endpoints(c.d);
    }
    c.di--;
    c.d = rgs[c.di];
    return avail;
}

// calc_weights (:363-421): the Rosenbluth-like weight of the configuration
// in `actual`, built level by level from the attempts and configurations
// the growth left (rgwatt, rgwavail) and the open probabilities (rgopen)
DEVNI double Eng::rg_calc_weights(const PosOre* actual) {
    int n = S->n_rgs;
    double weight = 1;
    RgCur c;
    c.di = 0;
    c.d = rgs[0];
    while (c.di != n - 1 && S->status == ST_OK) {
        c.di++;
        c.d = rgs[c.di];
        c.stemd = stem[c.d] >= 0;
        c.dir = rg_get_dir(c.d);
        int avail_cs = 1;
        if (!c.stemd) {
            c.ref = walk(c.d, -c.dir);
            int att = rgwatt[c.di];
            c.avail = rgwavail[c.di];
            while (att != rg_max_att) {
                att++;
                I3 p;
                int o;
                rg_select(c, p, o);
                double pc = rg_p_open(c, p, o);
                if (rg_test_open(pc)) {
                    rg_set_config(c.d, p, o, RG_NONE);
                    RgCur keep = c;
                    avail_cs += rg_test_avail(c) ? 1 : 0;
                    c.avail = keep.avail;
                    c.stemd = keep.stemd;
                    c.ref = keep.ref;
                    c.dir = keep.dir;
                    unassign_domain(c.d);
                    rg_restore_endpoints(c.d);
                }
            }
        }
        weight *= avail_cs / rgopen[c.di - 1];
        const PosOre& a = actual[c.d];
        rg_set_config(c.d, i3(a.x, a.y, a.z), a.ore, RG_NONE);
    }
    weight /= rgopen[c.di];
    return weight;
}

// calc_old_c_opens (:482-512): open probabilities of the old configuration
// (in `old`) and the configurations its growth would have left
DEVNI void Eng::rg_calc_old_c_opens() {
    int n = S->n_rgs;
    RgCur c;
    c.di = 0;
    rgopen[0] = 1;
    while (c.di != n - 1 && S->status == ST_OK) {
        c.di++;
        c.d = rgs[c.di];
        rgatt[c.di] = 1;
        I3 p = i3(old[c.d].x, old[c.d].y, old[c.d].z);
        int o = old[c.d].ore;
        c.stemd = stem[c.d] >= 0;
        rgopen[c.di] = rg_p_open(c, p, o);
        rg_set_config(c.d, p, o, RG_DUP);
        if (c.stemd) {
            rgavail[c.di] = 0;
        }
        else {
            c.dir = rg_get_dir(c.d);
            c.ref = walk(c.d, -c.dir);
            I3 v = sub3(p, i3(old[c.ref].x, old[c.ref].y, old[c.ref].z));
            int l1 = (v.x < 0 ? -v.x : v.x) + (v.y < 0 ? -v.y : v.y) + (v.z < 0 ? -v.z : v.z);
            if (l1 != 1) {
                fail(ST_CONSTRAINT);
                return;
            }
            rgavail[c.di] = LDO_RG_ALL & ~(1ull << (6 * unit_index(v) + o));
        }
    }
}

// Constraintpoints::reset_active_endpoints (:295-297), then
// remove_active_endpoint of the segment's first domain when asked
DEVIN void Eng::rg_reset_active(bool remove_first) {
    for (int i = plane(), n = S->n_active0; i < n; i += pwidth()) {
        active[i] = active0[i];
    }
    S->n_active = S->n_active0;
    if (remove_first) {
        cp_remove_active(rgs[0]);
    }
}

// The common part of CTRGScaffoldRegrowthMCMovetype::internal_attempt_move
// (:640-686) and CTRGJumpScaffoldRegrowthMCMovetype's (:811-850) once the
// regrowth stack is set up: recoil growth, the weights of the new and the
// old configuration, and test_rg_acceptance (:514-533), with
// setup_for_calc_new_weights (:144-150) and setup_for_calc_old_weights
// (:152-160).  rm_new, rm_opens and rm_old say whether the stack's first
// domain loses its active endpoints before the new weights, the old open
// probabilities and the old weights.
DEVNI bool Eng::rg_regrow_and_accept(bool rm_new, bool rm_opens, bool rm_old) {
    int n = S->n_rgs;
    rg_dup = 0;
    S->delta_e += rg_unassign_and_save(true);
    S->delta_e += rg_recoil_regrow();
    if (S->rejected || S->status != ST_OK) {
        return false;
    }
    S->delta_e += calc_move_bias(0); // add_external_bias (:82-85)

    // weights of the new configuration (kept in prev)
    for (int i = 0; i < n; i++) {
        int d = rgs[i];
        rgwatt[i] = rgatt[i];
        rgwavail[i] = rgavail[i];
        old[d] = prev[d];
    }
    S->n_modified = 0;
    rg_unassign_and_save(true);
    rg_reset_active(rm_new);
    double w_new = rg_calc_weights(prev);

    // weights of the old configuration (kept in old)
    for (int i = 1; i < n; i++) {
        unassign_domain(rgs[i]);
    }
    S->er_top = 0;
    S->er_depth = 0;
    rg_reset_active(rm_opens);
    rg_calc_old_c_opens();
    for (int i = 0; i < n; i++) {
        rgwatt[i] = rgatt[i];
        rgwavail[i] = rgavail[i];
    }
    S->n_modified = 0;
    rg_unassign_and_save(false);
    rg_reset_active(rm_old);
    double w_old = rg_calc_weights(old);
    if (S->status != ST_OK) {
        return false;
    }
    double ratio = w_new / w_old * exp(-S->delta_e);
    if (test_acceptance(ratio)) {
        // reset_origami with the new configuration as the one to restore
        reset_origami();
        S->commits += rg_dup;
        return true;
    }
    S->n_modified = 0;
    S->n_assigned = 0;
    return false;
}

// CTRGScaffoldRegrowthMCMovetype::internal_attempt_move (:629-689) with
// setup_constraints (:131-142)
DEVNI bool Eng::ctrg_scaffold_move(int max_regrowth) {
    cp_reset();
    int first, count, dir;
    if (!ct_select_indices(max_regrowth, first, count, dir)) {
        return false;
    }
    seg_first = first;
    seg_count = count;
    seg_dir = dir;
    S->trk_len = count;
    cp_find(first, count, dir, 0);
    cp_save_initial();
    if (S->status != ST_OK) {
        return false;
    }
    if (!(P->cyclic && count == ns)) {
        cp_remove_active(first);
    }
    int n = S->n_rgs;
    bool whole = P->cyclic && n == ns;
    // the reference's test before calc_old_c_opens (:665-668) differs from
    // the other two
    return rg_regrow_and_accept(!whole, !P->cyclic && n == ns, !whole);
}

// ------------------------------------ non-contiguous scaffold selection
// CTRegrowthMCMovetype::check_for_stemds (movetypes.cpp:704-720): queue the
// scaffold domains bound to the chain neighbours of the staple domain bound
// to cur
DEVIN void Eng::jump_check_stems(int cur, int& tail) {
    if (dom[cur].state != OCC_BOUND) {
        return;
    }
    int b = dom[cur].bound;
    for (int sd = -1; sd <= 1; sd += 2) {
        int nd = walk(b, sd);
        if (nd >= 0 && dom[nd].state == OCC_BOUND && cslot_of(dom[nd].bound) == 0) {
            if (tail >= LDO_JUMP_QUEUE) {
                fail(ST_CAPACITY);
                return;
            }
            jps[tail++] = dom[nd].bound;
        }
    }
}

// CTRegrowthMCMovetype::fill_seg (movetypes.cpp:666-702) for a segment that
// already holds `have` domains; last is the last domain added (unchanged
// when none is)
DEVNI bool Eng::jump_fill(int start, int have, int max_length, int seg_max, int dir, int& tail,
                          int& last) {
    int cur = start;
    jump_check_stems(cur, tail);
    int next = start;
    int size = have;
    while (size != seg_max && next >= 0) {
        next = walk(cur, dir);
        if (next < 0) {
            break;
        }
        int nn = walk(next, dir);
        if (nn >= 0 && jdom[nn]) {
            break;
        }
        size++;
        jdom[next] = 1;
        S->n_jdom++;
        last = next;
        if (S->n_jdom == max_length) {
            return true;
        }
        cur = next;
        jump_check_stems(cur, tail);
    }
    return false;
}

// CTRegrowthMCMovetype::select_noncontig_segs (movetypes.cpp:512-648): the
// segments (jseg), their directions (sdir), the scaffold stems (jstem) with
// their growthpoints and segment pairs, and the active endpoints past each
// segment's end
DEVNI bool Eng::jump_select(int max_regrowth) {
    fill_section(jdom, ns, 0u);
    fill_section(stemseg, ns, 0xffffffffu);
    fill_section(sdir, LDO_JUMP_MAX_SEGS, 0u);
    S->n_jdom = 0;
    int max_length = uniform_int(2, max_regrowth);
    int seg_max = uniform_int(2, rg_max_segreg + 1);
    int start = slot(0, uniform_int(0, ns - 1));
    int dir = uniform_int(0, 1);
    if (dir == 0) {
        dir = -1;
    }
    if (walk(start, dir) < 0) {
        dir = -dir;
    }
    jdom[start] = 1;
    S->n_jdom = 1;
    int head = 0, tail = 0;
    int last = start;
    bool reached = jump_fill(start, 1, max_length, seg_max, dir, tail, last);
    JumpSeg s0;
    s0.first = (int16_t)start;
    s0.count = (int16_t)S->n_jdom; // start and the domains filled after it
    s0.last = (int16_t)last;
    s0.dir = (int8_t)dir;
    s0.pad = 0;
    jseg[0] = s0;
    S->n_jsegs = 1;
    S->n_jstems = 0;
    while (!reached && head < tail && S->status == ST_OK) {
        int stemd = jps[head++];
        if (jdom[stemd]) {
            continue;
        }
        int a = walk(stemd, -1), b = walk(stemd, 1);
        if ((a >= 0 && jdom[a]) || (b >= 0 && jdom[b])) {
            continue;
        }
        jdom[stemd] = 1;
        S->n_jdom++;
        if (S->n_jstems >= LDO_JUMP_MAX_SEGS / 2 || S->n_jsegs + 2 > LDO_JUMP_MAX_SEGS) {
            fail(ST_CAPACITY);
            return false;
        }
        jstem[S->n_jstems++] = (int16_t)stemd;
        int k = S->n_jsegs;
        S->n_jsegs += 2;
        JumpSeg p0, p1;
        p0.first = (int16_t)stemd;
        p0.count = 1;
        p0.last = (int16_t)stemd;
        p0.pad = 0;
        p1.first = LDO_NONE;
        p1.count = 0;
        p1.last = (int16_t)stemd;
        p1.pad = 0;
        if (S->n_jdom == max_length) {
            reached = true;
            p0.dir = 1;
            p1.dir = -1;
            jseg[k] = p0;
            jseg[k + 1] = p1;
            break;
        }
        int dir1 = uniform_int(0, 1);
        if (dir1 == 0) {
            dir1 = -1;
        }
        p0.dir = (int8_t)dir1;
        p1.dir = (int8_t)-dir1;
        for (int i = 0; i < 2; i++) {
            int d = i == 0 ? dir1 : -dir1;
            int sm = uniform_int(0, rg_max_segreg);
            if (i == 0) {
                max_length++;
            }
            int lst = stemd;
            int before = S->n_jdom;
            reached = jump_fill(stemd, 0, max_length, sm, d, tail, lst);
            int added = S->n_jdom - before;
            if (i == 0) {
                p0.count = (int16_t)(1 + added);
                p0.last = (int16_t)lst;
            }
            else {
                p1.first = added > 0 ? (int16_t)walk(stemd, d) : LDO_NONE;
                p1.count = (int16_t)added;
                p1.last = (int16_t)lst;
            }
            if (reached) {
                break;
            }
        }
        jseg[k] = p0;
        jseg[k + 1] = p1;
    }
    if (S->status != ST_OK) {
        return false;
    }
    int nx = walk(jseg[0].last, jseg[0].dir);
    if (nx >= 0) {
        cp_add_active(nx, pos(nx), 0);
    }
    for (int i = 0; i < S->n_jsegs; i++) {
        sdir[i] = jseg[i].count > 0 ? jseg[i].dir : 0;
    }
    for (int t = 0; t < S->n_jstems; t++) {
        int stemd = jstem[t];
        int k = 1 + 2 * t;
        gp[dom[stemd].bound] = (int16_t)stemd;
        stem[stemd] = dom[stemd].bound;
        stemseg[stemd] = (int8_t)k;
        for (int j = 0; j < 2; j++) {
            int x = walk(jseg[k + j].last, jseg[k + j].dir);
            if (x >= 0) {
                cp_add_active(x, pos(x), k + j);
            }
        }
    }
    return true;
}

// CTRGJumpScaffoldRegrowthMCMovetype::internal_attempt_move (:791-853)
DEVNI bool Eng::ctrg_jump_move(int max_regrowth) {
    cp_reset();
    if (!jump_select(max_regrowth)) {
        return false;
    }
    S->trk_len = 0;
    for (int i = 0; i < S->n_jsegs; i++) {
        if (jseg[i].count > 0) {
            cp_find(jseg[i].first, jseg[i].count, jseg[i].dir, i);
        }
    }
    cp_save_initial();
    if (S->status != ST_OK) {
        return false;
    }
    int first = jseg[0].first;
    cp_remove_active(first);
    for (int t = 0; t < S->n_jstems; t++) {
        int stemd = jstem[t];
        gp[dom[stemd].bound] = (int16_t)stemd;
        stem[stemd] = dom[stemd].bound;
    }
    return rg_regrow_and_accept(true, true, true);
}

// CTCBJumpScaffoldRegrowthMCMovetype::set_first_seg_domain
// (cb_movetypes.cpp:973-983): a stem binds its growthpoint as in growth
// from a staple, in the new and the old growth alike
DEVIN void Eng::ctcb_set_first_seg_domain(int stemd) {
    int d_old = stem[stemd];
    if (!cp_walks_remain(stemd, pos(d_old))) {
        S->rejected = 1;
        return;
    }
    double de = set_growth_point(stemd, d_old);
    S->bias *= exp(-de);
}

// the growth of both passes of CTCBJumpScaffoldRegrowthMCMovetype
// (cb_movetypes.cpp:896-921, 935-960): the first segment with the staple
// grown from its first domain, then per stem the stem and each of its two
// segments grown out of it
DEVNI void Eng::ctcb_jump_grow() {
    ctcb_grow_chain_scaffold(jseg[0].first, jseg[0].count, jseg[0].dir);
    for (int t = 0; t < S->n_jstems && !S->rejected && S->status == ST_OK; t++) {
        int stemd = jstem[t];
        int k = 1 + 2 * t;
        ctcb_set_first_seg_domain(stemd);
        for (int j = 0; j < 2 && !S->rejected && S->status == ST_OK; j++) {
            int count = j == 0 ? jseg[k].count : 1 + jseg[k + 1].count;
            ctcb_grow_chain_scaffold(stemd, count, jseg[k + j].dir, false);
        }
    }
}

// CTCBJumpScaffoldRegrowthMCMovetype::internal_attempt_move
// (cb_movetypes.cpp:872-967); unlike the contiguous move, a dead end in the
// growth of the old configuration rejects too
DEVNI bool Eng::ctcb_jump_move(int max_regrowth) {
    S->regrow_old = 0;
    cp_reset();
    if (!jump_select(max_regrowth)) {
        return false;
    }
    S->trk_len = 0;
    for (int i = 0; i < S->n_jsegs; i++) {
        if (jseg[i].count > 0) {
            cp_find(jseg[i].first, jseg[i].count, jseg[i].dir, i);
        }
    }
    cp_save_initial();
    if (S->status != ST_OK) {
        return false;
    }
    int first = jseg[0].first;
    cp_remove_active(first);
    int nst = 0;
    for (int c = 1; c < nchain; c++) {
        if (regrow[c] && ch[c].ident >= 0) {
            nst++;
        }
    }
    S->trk_nst = nst;
    int n = S->n_rgs;
    for (int phase = 0; phase < 2; phase++) {
        if (phase == 1) {
            cb_setup_regrow_old();
            for (int i = plane(), m = S->n_active0; i < m; i += pwidth()) {
                active[i] = active0[i];
            }
            S->n_active = S->n_active0;
            cp_remove_active(first);
        }
        // CBMCMovetype::unassign_domains (cb_movetypes.cpp:200-213) over the
        // regrowth stack past its first domain
        for (int i = 1; i < n; i++) {
            int d = rgs[i];
            prev[d].x = dom[d].x;
            prev[d].y = dom[d].y;
            prev[d].z = dom[d].z;
            prev[d].ore = dom[d].ore;
            push_modified(d);
            unassign_domain(d);
        }
        ctcb_jump_grow();
        if (S->rejected || S->status != ST_OK) {
            return false;
        }
        if (phase == 0) {
            S->bias *= exp(-calc_move_bias(0)); // add_external_bias (cb_movetypes.cpp:928)
        }
    }
    S->modifier = 1;
    return cb_test_acceptance();
}

// MCMovetype::attempt_move and the reset_internal chain (movetypes.cpp:41-51,
// 242-250; met_movetypes.cpp:45-48, 139-143; cb_movetypes.cpp:44-50, 457-461)
DEVIN bool Eng::attempt(int m) {
    return attempt_kind(P->move_kind[m], P->move_maxreg[m], m);
}

DEVNI bool Eng::attempt_kind(int kind, int max_regrowth, int m) {
    S->n_modified = 0;
    S->n_assigned = 0;
    S->n_added = 0;
    S->rejected = 0;
    S->modifier = 1;
    S->delta_e = 0;
    S->bias = 1;
    S->new_bias = 1;
    S->new_modifier = 1;
    S->regrow_old = 0;
    S->insertion_sites = S->num_domains;
    switch (kind) {
    case 0:
        S->kind = LDO_MOVE_KIND_ORIENT;
        return orientation_move();
    case 1:
        S->kind = LDO_MOVE_KIND_MET;
        S->trk_ins = uniform_real() < 0.5 ? 1 : 0;
        return S->trk_ins ? insert_staple() : delete_staple();
    case 2:
        S->kind = LDO_MOVE_KIND_CB;
        return cb_staple_move();
    case 3:
        S->kind = LDO_MOVE_KIND_CTCB;
        return ctcb_scaffold_move(max_regrowth);
    case 4:
        S->kind = LDO_MOVE_KIND_CTRG;
        rg_max_recoils = m >= 0 ? P->move_recoils[m] : 0;
        rg_max_att = m >= 0 ? P->move_cattempts[m] : 36;
        return ctrg_scaffold_move(max_regrowth);
    case 5:
        S->kind = LDO_MOVE_KIND_CTRG_JUMP;
        rg_max_recoils = m >= 0 ? P->move_recoils[m] : 0;
        rg_max_att = m >= 0 ? P->move_cattempts[m] : 36;
        rg_max_segreg = m >= 0 ? P->move_segreg[m] : 0;
        return ctrg_jump_move(max_regrowth);
    case 6:
        // a CTCB move throughout (grow_chain, calc_bias), on a jump selection
        S->kind = LDO_MOVE_KIND_CTCB;
        rg_max_segreg = m >= 0 ? P->move_segreg[m] : 0;
        return ctcb_jump_move(max_regrowth);
    }
    fail(ST_MISUSE);
    return false;
}

// MCMovetype::add_tracker (movetypes.cpp:41-51) for the move kinds whose
// summaries list tracked counts: MetStapleExchange by staple type and
// direction (met_movetypes.cpp:145-190), CBStapleRegrowth by staple type
// unless there were no staples (cb_movetypes.cpp:264-290),
// CTCBScaffoldRegrowth by segment length and staples regrown
// (cb_movetypes.cpp:626-678)
DEVIN void Eng::track_move(int m, bool acc) {
    S->last_move = m;
    S->last_acc = acc ? 1 : 0;
    int T = P->L.ntypes;
    int a = acc ? 1 : 0;
    int kind = P->move_kind[m];
    if (kind == 1) {
        int b = S->trk_ins ? 0 : 2 * T;
        track[b + S->trk_type] += 1;
        track[b + T + S->trk_type] += a;
    }
    else if (kind == 2) {
        if (!S->trk_nost) {
            track[4 * T + S->trk_type] += 1;
            track[5 * T + S->trk_type] += a;
        }
    }
    else if (kind == 3) {
        int o = 6 * T, n1 = ns + 1, s1 = P->L.smax + 1;
        track[o + S->trk_len] += 1;
        track[o + n1 + S->trk_len] += a;
        track[o + 2 * n1 + S->trk_nst] += 1;
        track[o + 2 * n1 + s1 + S->trk_nst] += a;
    }
    else if (kind == 4) {
        int o = 6 * T + 2 * (ns + 1) + 2 * (P->L.smax + 1);
        track[o + S->trk_len] += 1;
        track[o + ns + 1 + S->trk_len] += a;
    }
    else if (kind == 5) {
        int o = 6 * T + 4 * (ns + 1) + 2 * (P->L.smax + 1);
        track[o] += 1;
        track[o + 1] += a;
    }
    else if (kind == 6) {
        int s1 = P->L.smax + 1;
        int o = 6 * T + 4 * (ns + 1) + 2 * s1 + 2;
        track[o + S->trk_nst] += 1;
        track[o + s1 + S->trk_nst] += a;
    }
}

// GCMCSimulation::simulate (simulation.cpp:568-653) and select_movetype
// (:655-665); output writers are host-side and not part of the step
DEVNI void Eng::simulate(int64_t steps) {
    for (int64_t n = 0; n < steps; n++) {
        if (S->status != ST_OK) {
            break;
        }
        S->step += 1;
        S->rng_idx = 0;
        double prob = uniform_real();
        int m = 0;
        while (m < P->n_moves && !(prob < P->move_cum[m])) {
            m++;
        }
        if (m == P->n_moves) {
            m = P->n_moves - 1;
        }
        S->attempts[m]++;
        bool acc = attempt(m);
        S->accepts[m] += acc ? 1 : 0;
        track_move(m, acc);
        if (!acc) {
            reset_origami();
            calc_move_bias(0);
        }
        if (P->centering_freq != 0 && S->step % P->centering_freq == 0) {
            center();
        }
        if (P->check_freq != 0 && S->step % P->check_freq == 0) {
            check_all_constraints();
        }
        if (P->hist != nullptr && S->status == ST_OK) {
            count_window_op();
        }
    }
}

// USGCMCSimulation::update_internal (us_simulation.cpp:262-266): one visit
// of the window's order parameter after each step.  A slot belongs to one
// replica for a whole sweep, so its bins have a single writer.
DEVNI void Eng::count_window_op() {
    int p = window_op(0);
    int b = (p >= 0 && p < P->grid_len) ? p : P->grid_len;
    P->hist[(size_t)S->temp_slot * (P->grid_len + 1) + b] += 1;
}

// update_control_qs of the PT drivers (ptmc_simulation.cpp:595-680): switch
// to the tables of ladder slot `slot` and, with `recompute` (update_temp),
// recompute the energy.  A (temperature, stacking multiplier) key that
// update_temp meets for the first time rebuilds its tables in the reference,
// which also refreshes the initiation terms (origami_potential.cpp:1019-1042,
// 1057-1063); on later visits of the key the initiation terms keep their
// last values.  slot_key numbers the keys.  Without `recompute` (a window
// swap at one temperature) only the tables are selected.
DEVNI void Eng::begin_interval(int slot_, bool recompute) {
    S->temp_slot = slot_;
    if (recompute) {
        uint64_t bit = 1ull << (P->slot_key != nullptr ? P->slot_key[slot_] : slot_);
        if (!(S->visited & bit)) {
            S->visited |= bit;
            S->init_slot = slot_;
        }
    }
    select_tables();
    if (recompute) {
        update_energy();
    }
}

// the window's order parameter: fully bound pairs (numfulldomains) or
// staples (numstaples), the latter offset while a deletion is tested
DEVNI int Eng::window_op(int staple_offset) {
    const ldo_window& w = P->windows[S->temp_slot];
    if (w.op == LDO_OP_NUMSTAPLES) {
        return S->num_staples + staple_offset;
    }
    return S->num_fully;
}

// SystemOrderParams::update_move_params + SystemBiases::calc_move
// (order_params.cpp:595-603; bias_functions.cpp:484-496) for the replica's
// umbrella window: a LinearStepWell restraint (bias_functions.cpp:139-151)
// then a Grid bias (242-284).  The stored values belong to the window.
DEVNI double Eng::calc_move_bias(int staple_offset) {
    if (P->windows == nullptr) {
        return 0;
    }
    int q = S->temp_slot;
    const ldo_window& w = P->windows[q];
    int p = window_op(staple_offset);
    double r;
    if (p < w.min_op) {
        r = w.slope * (w.min_op - p - 1) + w.min_bias;
    }
    else if (p > w.max_op) {
        r = w.slope * (p - w.max_op - 1) + w.min_bias;
    }
    else {
        r = w.well_bias;
    }
    double g = (p >= 0 && p < P->grid_len) ? P->win_grid[(size_t)q * P->grid_len + p]
                                           : w.off_grid_bias;
    double* st = P->win_stored + 2 * q;
    double diff = r - st[0];
    diff += g - st[1];
    st[0] = r;
    st[1] = g;
    return diff * w.bias_mult;
}

// ------------------------------------------------------ learned proposal
// Learned domain regrowth (no reference counterpart): a staple end domain d
// moves next to its chain neighbour a (the anchor) to one of the 6 sites x 6
// orientations, drawn from the network's masked softmax over a context that
// leaves d out, so forward and reverse moves share one distribution.
// Acceptance is Metropolis-Hastings: min(1, exp(-dE - dB) q(old) / q(new)).
// With probability learned_fallback the attempt is a plain CB staple
// regrowth instead.  Returns 1 when the lanes must build the context and
// evaluate the network, 0 when the attempt is complete.
// whether staple chain cs reaches the scaffold through bound pairs
DEVNI bool Eng::chain_connected(int cs) {
    for (int k = 0; k < ch[cs].len; k++) {
        int b = dom[slot(cs, k)].bound;
        if (b < 0 || cslot_of(b) == cs) {
            continue;
        }
        if (cslot_of(b) == 0 || scan_for_scaffold(b, cs)) {
            return true;
        }
    }
    return false;
}

DEVNI int Eng::learned_begin(LearnedMove& m) {
    S->rng_sub = 1;
    S->learned_step += 1;
    S->rng_idx = 0;
    m.mode = 0;
    int ls = P->learned_slot;
    if (uniform_real() < P->learned_fallback) {
        S->attempts[ls - 1]++;
        bool acc = attempt_kind(2, 0);
        S->accepts[ls - 1] += acc ? 1 : 0;
        if (!acc) {
            reset_origami();
            calc_move_bias(0);
        }
        S->rng_sub = 0;
        return 0;
    }
    if (uniform_real() < 0.5) {
        S->attempts[ls - 2]++;
        return learned_exchange_begin(m);
    }
    S->attempts[ls]++;
    m.kind = 1;
    if (S->num_staples == 0) {
        S->rng_sub = 0;
        return 0;
    }
    int w = uniform_int(1, S->num_staples);
    int cs = order[w];
    int len = ch[cs].len;
    bool first = uniform_real() < 0.5;
    int d = first ? slot(cs, 0) : slot(cs, len - 1);
    int a = first ? slot(cs, 1) : slot(cs, len - 2);
    int idv = ident_of(d);
    int id = (idv < 0 ? -idv : idv) - 1;
    if (len < 2 || id < 0 || id >= P->net_n_ident) {
        fail(ST_MISUSE);
        S->rng_sub = 0;
        return 0;
    }
    I3 ap = pos(a);
    int n_old = unit_index(sub3(pos(d), ap));
    m.d = d;
    m.a = a;
    m.id = id;
    m.k_old = 6 * n_old + dom[d].ore;
    m.ax = ap.x;
    m.ay = ap.y;
    m.az = ap.z;
    m.mode = 1;
    return 1;
}

// unnormalised masked softmax of the network's logits and its sum
DEVNI void Eng::learned_softmax(const LearnedMove& m, double* w, double& z) {
    float mx = -INFINITY;
    for (int k = 0; k < 36; k++) {
        if (m.legal[k / 6]) {
            mx = fmaxf(mx, m.logits[k]);
        }
    }
    z = 0;
    for (int k = 0; k < 36; k++) {
        w[k] = m.legal[k / 6] ? exp((double)m.logits[k] - (double)mx) : 0.0;
        z += w[k];
    }
}

DEVNI int Eng::learned_continue(LearnedMove& m) {
    if (m.kind == 1) {
        learned_finish(m);
        return 0;
    }
    return learned_exchange_continue(m);
}

DEVNI void Eng::learned_finish(LearnedMove& m) {
    int ls = P->learned_slot;
    double w[36];
    double z;
    learned_softmax(m, w, z);
    double u = uniform_real() * z;
    double cum = 0;
    int kn = -1;
    for (int k = 0; k < 36; k++) {
        if (w[k] == 0) {
            continue;
        }
        kn = k;
        cum += w[k];
        if (u < cum) {
            break;
        }
    }
    bool accepted = false;
    if (kn < 0) {
        // unreachable (d's own site is always legal); kept as a bound on kn
        S->accepts[ls] += 0;
        S->rng_sub = 0;
        return;
    }
    if (kn == m.k_old) {
        accepted = true;
    }
    else {
        S->n_modified = 0;
        S->n_assigned = 0;
        S->n_added = 0;
        S->rejected = 0;
        S->violated = 0;
        S->modifier = 1;
        int d = m.d;
        int cs = cslot_of(d);
        int partner = dom[d].bound;
        prev[d].x = dom[d].x;
        prev[d].y = dom[d].y;
        prev[d].z = dom[d].z;
        prev[d].ore = dom[d].ore;
        push_modified(d);
        double de = unassign_domain(d);
        I3 pn = add3(i3(m.ax, m.ay, m.az), unit3(kn / 6));
        de += set_domain_config(d, pn, kn % 6);
        push_assigned(d);
        if (!S->violated) {
            // every staple must stay connected to the scaffold (the ensemble
            // of the reference's moves): d's own, and a staple d left
            bool ok = chain_connected(cs);
            if (ok && partner >= 0 && cslot_of(partner) != 0 && cslot_of(partner) != cs) {
                ok = chain_connected(cslot_of(partner));
            }
            if (ok) {
                de += calc_move_bias(0);
                accepted = test_acceptance(exp(-de) * (w[m.k_old] / w[kn]));
            }
        }
        if (!accepted) {
            reset_origami();
            calc_move_bias(0);
        }
    }
    S->accepts[ls] += accepted ? 1 : 0;
    S->rng_sub = 0;
}

// Learned staple exchange (no reference counterpart), for staples of two
// domains: an insertion draws a staple type uniformly and one of its two
// domains (the growth domain), binds it to its complementary scaffold domain
// (which must be unbound), and places the other domain from the network's
// distribution around the growth domain.  A deletion draws one of the N
// staples uniformly.  With P_ins the summed probability of the growth paths
// that produce the staple (a path per domain bound to its complement:
// (1 / types) (1 / 2) q), the stationary weight of the reference's ensemble
// (staple_M / 6^3 per staple, exp(-E) with the initiation term) gives the
// acceptance ratio exp(-dE - dB) (staple_M / 216) (1 / N) / P_ins for an
// insertion and its inverse for a deletion (N counted with the staple).
// the scaffold domain a staple domain grows onto: scaffold domain v - 1 for
// identity v when it carries identity -v, else none
DEVNI int Eng::complement_of(int s) {
    int v = ident_of(s);
    if (v < 1 || v > ns) {
        return -1;
    }
    int x = slot(0, v - 1);
    return ident_of(x) == -v ? x : -1;
}

DEVNI int Eng::learned_end(LearnedMove& m, bool changed, bool accepted) {
    if (changed && !accepted) {
        reset_origami();
        calc_move_bias(0);
    }
    S->accepts[P->learned_slot - 2] += accepted ? 1 : 0;
    S->rng_sub = 0;
    m.mode = 0;
    return 0;
}

// sets up the scoring of growth path i of staple m.cs (domain i grows, the
// other is placed by the network) for a deletion
DEVNI int Eng::learned_path(LearnedMove& m, int i) {
    int gi = slot(m.cs, i), dj = slot(m.cs, 1 - i);
    int idv = ident_of(dj);
    m.a = gi;
    m.d = dj;
    m.id = (idv < 0 ? -idv : idv) - 1;
    I3 ap = pos(gi);
    m.ax = ap.x;
    m.ay = ap.y;
    m.az = ap.z;
    m.k_old = 6 * unit_index(sub3(pos(dj), ap)) + dom[dj].ore;
    m.stage = i;
    m.mode = 1;
    return 1;
}

DEVNI int Eng::learned_exchange_begin(LearnedMove& m) {
    S->n_modified = 0;
    S->n_assigned = 0;
    S->n_added = 0;
    S->rejected = 0;
    S->violated = 0;
    S->modifier = 1;
    S->delta_e = 0;
    m.qsum = 0;
    int types = P->L.ntypes - 1;
    if (uniform_real() < 0.5) {
        m.kind = 2;
        if (S->num_staples == P->max_total_staples) {
            return learned_end(m, false, false);
        }
        int ident = uniform_int(1, types);
        if (tcount[ident] == P->max_type_staples || P->type_len[ident] != 2) {
            return learned_end(m, false, false);
        }
        int i = uniform_int(0, 1);
        int cs = add_chain(ident);
        if (cs < 0) {
            return learned_end(m, false, false);
        }
        S->delta_e += init_e;
        S->added[S->n_added++] = cs;
        int gi = slot(cs, i), dj = slot(cs, 1 - i);
        int x = complement_of(gi);
        if (x < 0 || dom[x].state != OCC_UNBOUND) {
            return learned_end(m, true, false);
        }
        S->delta_e += set_domain_config(gi, pos(x), dom[x].ore ^ 1);
        push_assigned(gi);
        if (S->violated) {
            return learned_end(m, true, false);
        }
        int idv = ident_of(dj);
        m.cs = cs;
        m.a = gi;
        m.d = dj;
        m.id = (idv < 0 ? -idv : idv) - 1;
        I3 ap = pos(gi);
        m.ax = ap.x;
        m.ay = ap.y;
        m.az = ap.z;
        m.k_old = -1;
        m.stage = 0;
        m.mode = 1;
        return 1;
    }
    m.kind = 3;
    if (S->num_staples == 0) {
        return learned_end(m, false, false);
    }
    int w = uniform_int(1, S->num_staples);
    int cs = order[w];
    if (ch[cs].len != 2 || staple_is_connector(cs)) {
        return learned_end(m, false, false);
    }
    m.cs = cs;
    for (int k = 0; k < 2; k++) {
        int s = slot(cs, k);
        m.valid[k] = (dom[s].state == OCC_BOUND && dom[s].bound == complement_of(s)) ? 1 : 0;
    }
    if (!m.valid[0] && !m.valid[1]) {
        return learned_end(m, false, false);
    }
    return learned_path(m, m.valid[0] ? 0 : 1);
}

DEVNI int Eng::learned_exchange_continue(LearnedMove& m) {
    double w[36];
    double z;
    learned_softmax(m, w, z);
    double types = (double)(P->L.ntypes - 1);
    if (m.kind == 2 && m.k_old < 0) {
        // place the second domain from the network's distribution
        double u = uniform_real() * z;
        double cum = 0;
        int kn = -1;
        for (int k = 0; k < 36; k++) {
            if (w[k] == 0) {
                continue;
            }
            kn = k;
            cum += w[k];
            if (u < cum) {
                break;
            }
        }
        if (kn < 0) {
            // every site next to the growth domain is taken: no placement
            // can be proposed, so the insertion is rejected
            return learned_end(m, true, false);
        }
        m.qsum = w[kn] / z;
        int dj = m.d;
        S->delta_e += set_domain_config(dj, add3(pos(m.a), unit3(kn / 6)), kn % 6);
        push_assigned(dj);
        if (S->violated) {
            return learned_end(m, true, false);
        }
        if (dom[dj].state == OCC_BOUND && dom[dj].bound == complement_of(dj)) {
            // the staple also grows from the other domain: score that path
            int gi = m.a;
            int idv = ident_of(gi);
            m.a = dj;
            m.d = gi;
            m.id = (idv < 0 ? -idv : idv) - 1;
            I3 ap = pos(dj);
            m.ax = ap.x;
            m.ay = ap.y;
            m.az = ap.z;
            m.k_old = 6 * unit_index(sub3(pos(gi), ap)) + dom[gi].ore;
            m.stage = 1;
            return 1;
        }
    }
    else {
        m.qsum += w[m.k_old] / z;
        if (m.kind == 3 && m.stage == 0 && m.valid[1]) {
            return learned_path(m, 1);
        }
    }
    double p_ins = 0.5 * m.qsum / types;
    if (m.kind == 2) {
        S->delta_e += calc_move_bias(0);
        double ratio = exp(-S->delta_e) * (P->fugacity / 216.0) / (double)S->num_staples / p_ins;
        return learned_end(m, true, test_acceptance(ratio));
    }
    // deletion
    int cs = m.cs;
    double n_before = (double)S->num_staples;
    for (int k = 0; k < 2; k++) {
        int s = slot(cs, k);
        prev[s].x = dom[s].x;
        prev[s].y = dom[s].y;
        prev[s].z = dom[s].z;
        prev[s].ore = dom[s].ore;
        push_modified(s);
        S->delta_e += unassign_domain(s);
    }
    S->delta_e -= init_e;
    S->delta_e += calc_move_bias(-1);
    double ratio = exp(-S->delta_e) * p_ins / ((P->fugacity / 216.0) / n_before);
    bool acc = test_acceptance(ratio);
    if (acc) {
        delete_chain(cs);
    }
    return learned_end(m, true, acc);
}

#endif
