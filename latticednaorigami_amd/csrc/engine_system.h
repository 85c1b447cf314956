// engine_system.h - lattice state updates of one replica.
//
// OrigamiSystem (origami_system.cpp:193-871): occupancy bookkeeping on the
// site hash, domain set/unset with FP64 energy accumulation in the
// reference's order, chain add/delete with the working chain order, and the
// periodic centring / full constraint check / energy recomputation.
#ifndef LDO_ENGINE_SYSTEM_H
#define LDO_ENGINE_SYSTEM_H

#include "engine.h"

// OrigamiSystem::update_occupancies (origami_system.cpp:770-806)
DEVIN void Eng::update_occupancies(int s, I3 p) {
    int occ = occupancy(p);
    if (occ == OCC_UNBOUND) {
        int dj = unbound_at(p);
        S->num_bound += 1;
        int nst;
        if (ident_of(s) == -ident_of(dj)) {
            nst = OCC_BOUND;
            S->num_fully += 1;
        }
        else {
            if (cslot_of(s) == cslot_of(dj)) {
                S->num_self += 1;
            }
            nst = OCC_MISBOUND;
        }
        // a bound site keeps the domain placed last (the learned proposal
        // reads both domains of a site through it)
        hput(p, nst, s);
        dom[s].state = (int8_t)nst;
        dom[dj].state = (int8_t)nst;
        dom[dj].bound = (int16_t)s;
        dom[s].bound = (int16_t)dj;
    }
    else if (occ == OCC_UNASSIGNED) {
        dom[s].state = OCC_UNBOUND;
        hput(p, OCC_UNBOUND, s);
    }
    else {
        fail(ST_CONSTRAINT);
    }
}

// OrigamiSystem::internal_check_domain_constraints (:828-871)
DEVIN Delta Eng::internal_check(int s, I3 p, int o) {
    Delta dc;
    dc.e = 0;
    dc.st = 0;
    int occ = occupancy(p);
    if (occ == OCC_BOUND || occ == OCC_MISBOUND) {
        S->violated = 1;
        S->num_unassigned++;
    }
    else if (occ == OCC_UNBOUND) {
        dom[s].x = p.x;
        dom[s].y = p.y;
        dom[s].z = p.z;
        dom[s].ore = (int8_t)o;
        update_occupancies(s, p);
        dc = bind_domain(s);
        S->violated = pot_viol ? 1 : 0;
    }
    else {
        dom[s].x = p.x;
        dom[s].y = p.y;
        dom[s].z = p.z;
        dom[s].ore = (int8_t)o;
        update_occupancies(s, p);
        S->site_empty++;
    }
    return dc;
}

// OrigamiSystem::internal_unassign_domain (:695-760).  need_delta=false is
// used where the caller discards the returned change (a peek, or undoing a
// violating placement); the stacking walk is then skipped.
DEVIN Delta Eng::internal_unassign(int s, bool need_delta) {
    Delta dc;
    dc.e = 0;
    dc.st = 0;
    int occ = dom[s].state;
    int dj = dom[s].bound;
    if (occ == OCC_BOUND) {
        S->num_fully -= 1;
        S->num_bound -= 1;
        if (need_delta) {
            dc_e = 0;
            dc_st = 0;
            fb_viol = false;
            S->site_bind++;
            fb_terms(s, dj);
            dc.e = dc_e * -1;
            dc.st = dc_st * -1;
        }
    }
    else if (occ == OCC_MISBOUND) {
        S->num_bound -= 1;
        if (cslot_of(dj) == cslot_of(s)) {
            S->num_self -= 1;
        }
    }
    if (occ == OCC_BOUND || occ == OCC_MISBOUND) {
        // unassign_bound_domain (:741-753)
        double de = -hyb(s, dj);
        dom[s].bound = LDO_NONE;
        dom[dj].bound = LDO_NONE;
        dom[s].state = OCC_UNASSIGNED;
        hput(pos(s), OCC_UNBOUND, dj);
        dom[dj].state = OCC_UNBOUND;
        dc.e += de;
    }
    else if (occ == OCC_UNBOUND) {
        // unassign_unbound_domain (:755-760)
        herase(pos(s));
        dom[s].state = OCC_UNASSIGNED;
    }
    else {
        S->num_unassigned--;
    }
    return dc;
}

// OrigamiSystem::check_domain_constraints (:343-355)
DEVNI double Eng::check_domain_constraints(int s, I3 p, int o) {
    // An unassigned domain checked at an empty site: placing it and taking it
    // off again leaves the site hash as it was (the key is appended at the
    // end of its probe run, so its removal shifts nothing back) and the
    // domain unassigned at p with orientation o.  Only that and the event
    // count remain, so they are written directly.
    if (dom[s].state == OCC_UNASSIGNED && occupancy(p) == OCC_UNASSIGNED) {
        dom[s].x = p.x;
        dom[s].y = p.y;
        dom[s].z = p.z;
        dom[s].ore = (int8_t)o;
        S->site_empty++;
        return 0;
    }
    Delta dc = internal_check(s, p, o);
    internal_unassign(s, false);
    return dc.e;
}

// OrigamiSystem::unassign_domain (:375-385)
DEVNI double Eng::unassign_domain(int s) {
    Delta dc = internal_unassign(s, true);
    S->energy += dc.e;
    S->num_stacked += dc.st;
    S->num_unassigned++;
    S->commits++;
    return dc.e;
}

// OrigamiSystem::set_checked_domain_config (:478-515)
DEVNI double Eng::set_checked_domain_config(int s, I3 p, int o) {
    dom[s].x = p.x;
    dom[s].y = p.y;
    dom[s].z = p.z;
    dom[s].ore = (int8_t)o;
    update_occupancies(s, p);
    double de = 0;
    if (dom[s].state == OCC_MISBOUND) {
        de += hyb(s, dom[s].bound);
    }
    else if (dom[s].state == OCC_BOUND) {
        int dj = dom[s].bound;
        de += hyb(s, dj);
        dc_e = 0;
        dc_st = 0;
        fb_viol = false;
        S->site_bind++;
        fb_terms(s, dj);
        de += dc_e;
        S->num_stacked += dc_st;
    }
    S->energy += de;
    S->num_unassigned--;
    S->commits++;
    return de;
}

// OrigamiSystem::set_domain_config (:517-541)
DEVNI double Eng::set_domain_config(int s, I3 p, int o) {
    if (dom[s].state != OCC_UNASSIGNED) {
        fail(ST_CONSTRAINT);
        return 0;
    }
    Delta dc = internal_check(s, p, o);
    if (S->violated) {
        internal_unassign(s, false);
    }
    else {
        S->energy += dc.e;
        S->num_stacked += dc.st;
        S->num_unassigned--;
        S->commits++;
    }
    return dc.e;
}

// OrigamiSystem::add_chain(ident, index) (:402-443)
DEVNI int Eng::add_chain_with(int ident, int uid) {
    int cs;
    if (ident == 0) {
        cs = 0;
    }
    else {
        if (S->n_free == 0) {
            fail(ST_CAPACITY);
            return -1;
        }
        cs = freel[--S->n_free];
        int tc = tcount[ident];
        if (tc >= P->L.tcap) {
            fail(ST_CAPACITY);
            return -1;
        }
        tlist[ident * P->L.tcap + tc] = (int16_t)cs;
        tcount[ident] = (int16_t)(tc + 1);
    }
    int len = P->type_len[ident];
    ch[cs].uid = uid;
    ch[cs].ident = (int16_t)ident;
    ch[cs].len = (int16_t)len;
    order[S->n_chains++] = (int16_t)cs;
    for (int d = 0; d < len; d++) {
        int s = slot(cs, d);
        dom[s].x = 0;
        dom[s].y = 0;
        dom[s].z = 0;
        dom[s].ore = 0;
        dom[s].state = OCC_UNASSIGNED;
        dom[s].bound = LDO_NONE;
        dident[s] = P->type_ident[ident * P->type_ident_stride + d];
    }
    S->num_staples++;
    S->num_domains += len;
    S->num_unassigned += len;
    return cs;
}

// OrigamiSystem::add_chain(ident) (:387-400)
DEVNI int Eng::add_chain(int ident) {
    S->current_c_i += 1;
    S->energy += init_e;
    return add_chain_with(ident, S->current_c_i);
}

// OrigamiSystem::delete_chain (:445-472)
DEVNI void Eng::delete_chain(int cs) {
    int w = 0;
    while (w < S->n_chains && order[w] != cs) {
        w++;
    }
    int ident = ch[cs].ident;
    int tc = tcount[ident];
    int k = 0;
    while (k < tc && tlist[ident * P->L.tcap + k] != cs) {
        k++;
    }
    for (; k + 1 < tc; k++) {
        tlist[ident * P->L.tcap + k] = tlist[ident * P->L.tcap + k + 1];
    }
    tcount[ident] = (int16_t)(tc - 1);
    for (; w + 1 < S->n_chains; w++) {
        order[w] = order[w + 1];
    }
    S->n_chains--;
    int len = ch[cs].len;
    S->num_domains -= len;
    S->num_staples--;
    S->num_unassigned -= len;
    ch[cs].ident = -1;
    freel[S->n_free++] = (int16_t)cs;
    S->energy -= init_e;
}

// OrigamiSystem::center (:553-571): translate so that the centring scaffold
// domain sits at the origin and rebuild the site hash
DEVNI void Eng::center() {
    I3 ref = pos(slot(0, P->centering_domain));
    for (int i = 0; i < P->L.hcap; i++) {
        hkey[i] = LDO_HASH_EMPTY;
    }
    for (int w = 0; w < S->n_chains; w++) {
        int cs = order[w];
        for (int d = 0; d < ch[cs].len; d++) {
            int s = slot(cs, d);
            I3 np = sub3(pos(s), ref);
            dom[s].x = np.x;
            dom[s].y = np.y;
            dom[s].z = np.z;
            int stt = dom[s].state;
            if (stt == OCC_UNBOUND) {
                hput(np, OCC_UNBOUND, s);
            }
            else if (stt != OCC_UNASSIGNED) {
                hput(np, stt, s);
            }
        }
    }
}

// OrigamiSystem::set_all_domains() + check_distance_constraints (:573-586, 357-373)
DEVNI void Eng::set_all_domains() {
    for (int w = 0; w < S->n_chains; w++) {
        int cs = order[w];
        for (int d = 0; d < ch[cs].len; d++) {
            int s = slot(cs, d);
            set_domain_config(s, pos(s), dom[s].ore);
            if (S->violated) {
                fail(ST_CONSTRAINT);
                return;
            }
        }
    }
    for (int w = 0; w < S->n_chains; w++) {
        int cs = order[w];
        for (int d = 0; d < ch[cs].len; d++) {
            int s = slot(cs, d);
            int n = walk(s, 1);
            if (n < 0) {
                continue;
            }
            I3 dr = sub3(pos(n), pos(s));
            int l1 = (dr.x < 0 ? -dr.x : dr.x) + (dr.y < 0 ? -dr.y : dr.y) +
                     (dr.z < 0 ? -dr.z : dr.z);
            if (l1 != 1) {
                fail(ST_CONSTRAINT);
                return;
            }
        }
    }
}

// OrigamiSystem::check_all_constraints (:267-325)
DEVNI void Eng::check_all_constraints() {
    if (S->num_unassigned != 0) {
        fail(ST_CONSTRAINT);
        return;
    }
    for (int w = 0; w < S->n_chains; w++) {
        int cs = order[w];
        for (int d = 0; d < ch[cs].len; d++) {
            int s = slot(cs, d);
            if (dom[s].state == OCC_UNASSIGNED) {
                fail(ST_CONSTRAINT);
                return;
            }
            unassign_domain(s);
        }
    }
    S->energy -= (double)(S->n_chains - 1) * init_e;
    if (S->num_stacked != 0) {
        set_all_domains();
        fail(ST_CONSTRAINT);
        return;
    }
    double eps = 0.000001;
    if (S->energy < -eps || S->energy > eps) {
        set_all_domains();
        fail(ST_CONSTRAINT);
        return;
    }
    S->energy = 0;
    set_all_domains();
    S->energy += (double)(S->n_chains - 1) * init_e;
}

// OrigamiSystem::update_energy (:808-826)
DEVNI void Eng::update_energy() {
    for (int w = 0; w < S->n_chains; w++) {
        int cs = order[w];
        for (int d = 0; d < ch[cs].len; d++) {
            unassign_domain(slot(cs, d));
        }
    }
    S->energy = 0;
    S->energy += (double)(S->n_chains - 1) * init_e;
    set_all_domains();
}

// OrigamiSystem::update_enthalpy_and_entropy (:204-246); the per-slot flags
// stand in for the reference's std::set of accounted partners
DEVNI void Eng::update_enthalpy_and_entropy() {
    size_t n2 = (size_t)P->n_ident * P->n_ident;
    const double* th = P->hyb_h + (size_t)S->temp_slot * n2;
    const double* ts = P->hyb_s + (size_t)S->temp_slot * n2;
    for (int s = 0; s < nslots; s++) {
        netds_flag[s] = 0;
    }
    double h = 0, e = 0;
    for (int w = 0; w < S->n_chains; w++) {
        int cs = order[w];
        for (int d = 0; d < ch[cs].len; d++) {
            int s = slot(cs, d);
            int stt = dom[s].state;
            if (stt == OCC_BOUND || stt == OCC_MISBOUND) {
                if (!netds_flag[s]) {
                    int b = dom[s].bound;
                    size_t k = (size_t)dident[s] * P->n_ident + dident[b];
                    h += th[k];
                    e += ts[k];
                    netds_flag[b] = 1;
                }
            }
        }
    }
    h += (double)(S->n_chains - 1) * P->init_h[S->init_slot];
    e += (double)(S->n_chains - 1) * P->init_s[S->init_slot];
    S->hyb_h = h;
    S->hyb_s = e;
    S->stack_e = S->energy - (h - e);
}

#endif
