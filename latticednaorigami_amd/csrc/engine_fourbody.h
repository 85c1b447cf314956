// engine_fourbody.h - FourBody binding potential on slot-indexed domains.
//
// JunctionBindingPotential (origami_potential.cpp:158-934) and
// OrigamiPotential::bind_domain (:1282-1312): binding a complementary pair
// requires opposing orientations and is scored by walking the chain
// neighbours of both partners, their partners and the crossover junctions
// they form.  The walk is the same sequence of checks as the reference, so
// the FP64 energy is accumulated in the same order.
#ifndef LDO_ENGINE_FOURBODY_H
#define LDO_ENGINE_FOURBODY_H

#include "engine.h"

DEVIN void Eng::fb_single(int h1, int h2, int h3) {
    I3 n1 = sub3(pos(h2), pos(h1));
    if (eq3(n1, ore(h1))) {
        return;
    }
    I3 n2 = sub3(pos(h3), pos(h2));
    if (!eq3(n1, n2)) {
        dc_e -= stk(h2, h3);
        dc_st -= 1;
    }
}

DEVIN void Eng::fb_double(int h1, int h2, int h3) {
    I3 n1 = sub3(pos(h2), pos(h1));
    I3 n2 = sub3(pos(h3), pos(h2));
    if (!eq3(n1, n2)) {
        dc_e -= stk(h1, h2) / 2;
        dc_e -= stk(h2, h3) / 2;
        dc_st -= 1;
    }
}

DEVIN void Eng::fb_triply(int h1, int h2, int h3) {
    I3 n1 = sub3(pos(h2), pos(h1));
    I3 n2 = sub3(pos(h3), pos(h2));
    if (!eq3(n1, n2)) {
        fb_viol = true;
    }
}

// check_junction (origami_potential.cpp:419-462) with the stacking penalty
// of check_junction_stacking_penalty (:85-127)
DEVNI void Eng::fb_junction(int j1, int j2, int j3, int j4, int k1, int k2) {
    if (d_of(k1) > d_of(k2)) {
        int t = k1;
        k1 = k2;
        k2 = t;
        t = j1;
        j1 = j4;
        j4 = t;
        t = j2;
        j2 = j3;
        j3 = t;
    }
    if (!(pair_stacked(j1, j2) && pair_stacked(j3, j4))) {
        return;
    }
    if (!junction_ok(j1, j2, k1, k2)) {
        fb_viol = true;
        return;
    }
    int pen = 0;
    I3 nk = sub3(pos(k2), pos(k1));
    if (eq3(nk, ore(k1))) {
        I3 n1 = sub3(pos(j2), pos(j1));
        if (d_of(j1) > d_of(j2)) {
            n1 = neg3(n1);
        }
        I3 n3 = sub3(pos(j4), pos(j3));
        if (d_of(j3) > d_of(j4)) {
            n3 = neg3(n3);
        }
        pen = eq3(n1, n3) ? 2 : (eq3(n1, neg3(n3)) ? 0 : 1);
    }
    if (pen == 1) {
        dc_e -= stk(j1, j2) / 2;
        dc_e -= stk(j3, j4) / 2;
        dc_st -= 1;
    }
    else if (pen == 2) {
        dc_e -= stk(j1, j2);
        dc_e -= stk(j3, j4);
        dc_st -= 2;
    }
}

// check_doubly_contig_junction_pair (:359-417)
DEVIN void Eng::fb_junction_pair(int a, int b, int j) {
    int k1 = a, k2 = b;
    I3 ndr = sub3(pos(k2), pos(k1));
    if (!eq3(ore(k1), ndr)) {
        fb_viol = true;
        return;
    }
    if (j == 1) {
        return;
    }
    int sel_j1[2] = {walk(a, -1), walk(bd(a), 1)};
    int sel_j2[2] = {a, bd(a)};
    for (int q = 0; q < 2; q++) {
        int j1 = sel_j1[q], j2 = sel_j2[q];
        int j3 = k2;
        int j4 = walk(k2, 1);
        if (eb(j4) && eb(j1)) {
            fb_junction(j1, j2, j3, j4, k1, k2);
        }
        j3 = bd(k2);
        j4 = walk(j3, -1);
        if (eb(j4) && eb(j1)) {
            fb_junction(j1, j2, j3, j4, k1, k2);
        }
    }
}

// check_backward_triplet_stacking_combos (:464-519)
DEVIN void Eng::fb_backward_triplets(int a, int b) {
    int h2 = a, h3 = b;
    if (!pair_stacked(h2, h3)) {
        return;
    }
    int h1 = walk(a, -1);
    bool dc23 = doubly_contig(h2, h3);
    if (eb(h1)) {
        bool dc12 = doubly_contig(h1, h2);
        if (pair_stacked(h1, h2)) {
            if (dc12 && dc23) {
                fb_triply(h1, h2, h3);
                if (fb_viol) {
                    return;
                }
            }
            fb_double(h1, h2, h3);
        }
        else {
            fb_single(h1, h2, h3);
        }
    }
    int h2_prev = h1;
    int ba = bd(a);
    h1 = walk(ba, 1);
    if (eb(h1) && bd(h1) != h2_prev && bd(h1) != h3) {
        if (pair_stacked(ba, h1)) {
            fb_double(h1, h2, h3);
        }
    }
    h1 = walk(ba, -1);
    if (eb(h1) && bd(h1) != h2_prev && bd(h1) != h3) {
        if (pair_stacked(h1, ba)) {
            fb_double(h1, h2, h3);
        }
        else {
            fb_single(h1, ba, h3);
        }
    }
}

// check_forward_triplet_stacking_combos (:521-585)
DEVIN void Eng::fb_forward_triplets(int a, int b) {
    int h2 = b, h1 = a;
    bool first = pair_stacked(h1, h2);
    int h3 = walk(b, 1);
    bool dc12 = doubly_contig(h1, h2);
    if (eb(h3)) {
        bool second = pair_stacked(h2, h3);
        if (first && second) {
            bool dc23 = doubly_contig(h2, h3);
            if (dc12 && dc23) {
                fb_triply(h1, h2, h3);
                if (fb_viol) {
                    return;
                }
            }
            fb_double(h1, h2, h3);
        }
        else if (second) {
            fb_single(h1, h2, h3);
        }
    }
    int h2_next = h3;
    int bb = bd(b);
    h3 = walk(bb, 1);
    if (eb(h3) && bd(h3) != h2_next && bd(h3) != h1) {
        if (pair_stacked(bb, h3)) {
            if (first) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, h2, h3);
            }
        }
    }
    h3 = walk(bb, -1);
    if (eb(h3) && bd(h3) != h2_next && bd(h3) != h1) {
        if (pair_stacked(h3, bb)) {
            if (first) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, h2, h3);
            }
        }
        else if (first) {
            fb_single(h3, bb, h1);
        }
    }
}

// check_central_triplet_stacking_combos (:587-652)
DEVIN void Eng::fb_central_triplets(int di, int dj) {
    int h1 = walk(di, -1);
    int h2 = di;
    int h3 = walk(dj, 1);
    int h2_next = walk(di, 1);
    if (eb(h1) && eb(h3) && bd(h3) != h2_next && !doubly_contig(h1, h2)) {
        if (pair_stacked(dj, h3)) {
            if (pair_stacked(h1, h2)) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, h2, h3);
            }
        }
    }
    h3 = walk(dj, -1);
    if (eb(h1) && eb(h3) && bd(h3) != h1 && !doubly_contig(h1, h2)) {
        if (pair_stacked(h1, h2)) {
            if (pair_stacked(h3, dj)) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h3, dj, h1);
            }
        }
        else if (pair_stacked(h3, dj)) {
            fb_single(h1, h2, h3);
        }
    }
    h2 = di;
    h3 = walk(di, 1);
    h1 = walk(dj, 1);
    int h2_prev = walk(di, -1);
    if (eb(h1) && eb(h3) && bd(h1) != h3 && !doubly_contig(h2, h3)) {
        if (pair_stacked(dj, h1) && pair_stacked(h2, h3)) {
            fb_double(h1, h2, h3);
        }
    }
    h1 = walk(dj, -1);
    if (eb(h1) && eb(h3) && bd(h1) != h2_prev && !doubly_contig(h2, h3)) {
        if (pair_stacked(h2, h3)) {
            if (pair_stacked(h1, dj)) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, dj, h3);
            }
        }
    }
}

// The single-junction checks below enumerate up to three outer candidate
// pairs and, for each, up to three inner ones.  The lists are kept in scalar
// variables (LDO_NONE for an entry the reference does not add; eb() is false
// on it), each outer candidate in a function of its own, and no list lives in
// a local array: gfx950 code for a small stack array filled on divergent
// branches was found to store the last entry at the wrong index and leave the
// next one unwritten, which the loop then used as a domain slot.  The order
// of the fb_junction calls, and of every read before them, is the reference's.

// inner candidates of one outer pair (k2, k1) of check_backward_single_junction
DEVIN void Eng::fb_backward_outer(int j3, int j4, int k2, int k1) {
    if (!eb(k1) || pair_stacked(k1, k2)) {
        return;
    }
    int k1_next = walk(k1, d_of(k1) - d_of(k2));
    int k1b = bd(k1);
    int k1b_for = fwd(k1b);
    int k1b_bac = bwd(k1b);
    int e1 = k1b_for, e2 = k1b_bac;
    if (eb(k1_next)) {
        int k1nb = bd(k1_next);
        if (k1b_for == k1nb) {
            e1 = LDO_NONE;
        }
        if (k1b_bac == k1nb) {
            e2 = LDO_NONE;
        }
    }
    if (eb(k1_next)) {
        fb_junction(k1_next, k1, j3, j4, k1, k2);
    }
    if (eb(e1)) {
        fb_junction(e1, k1b, j3, j4, k1, k2);
    }
    if (eb(e2)) {
        fb_junction(e2, k1b, j3, j4, k1, k2);
    }
}

// check_backward_single_junction (:654-748)
DEVIN void Eng::fb_backward_single_junction(int a, int b) {
    int j3 = a, j4 = b;
    int j3_bac = bwd(j3);
    int j3b = bd(j3);
    int j3b_for = fwd(j3b);
    int j3b_bac = bwd(j3b);
    int c1, c2 = LDO_NONE; // outer candidates (j3b, c) after (j3, j3_bac)
    if (j3b_for == bd(j4)) {
        c1 = j3b_bac;
    }
    else if (eb(j3_bac) && eb(j3b_bac) && bd(j3_bac) == j3b_bac) {
        c1 = j3b_for;
    }
    else if (eb(j3_bac) && eb(j3b_for) && bd(j3_bac) == j3b_for) {
        c1 = j3b_bac;
    }
    else {
        c1 = j3b_for;
        c2 = j3b_bac;
    }
    fb_backward_outer(j3, j4, j3, j3_bac);
    fb_backward_outer(j3, j4, j3b, c1);
    fb_backward_outer(j3, j4, j3b, c2);
}

// inner candidates of one outer pair (k1, k2) of check_forward_single_junction
DEVIN void Eng::fb_forward_outer(int j1, int j2, int k1, int k2) {
    if (!eb(k2) || pair_stacked(k1, k2)) {
        return;
    }
    int k2_next = walk(k2, d_of(k2) - d_of(k1));
    int k2b = bd(k2);
    int k2b_for = fwd(k2b);
    int k2b_bac = bwd(k2b);
    int e1 = k2b_for, e2 = k2b_bac;
    if (eb(k2_next)) {
        int k2nb = bd(k2_next);
        if (k2b_for == k2nb) {
            e1 = LDO_NONE;
        }
        if (k2b_bac == k2nb) {
            e2 = LDO_NONE;
        }
    }
    if (eb(k2_next)) {
        fb_junction(j1, j2, k2, k2_next, k1, k2);
    }
    if (eb(e1)) {
        fb_junction(j1, j2, k2b, e1, k1, k2);
    }
    if (eb(e2)) {
        fb_junction(j1, j2, k2b, e2, k1, k2);
    }
}

// check_forward_single_junction (:750-844)
DEVIN void Eng::fb_forward_single_junction(int a, int b) {
    int j1 = a, j2 = b;
    int j2_for = fwd(j2);
    int j1b = bd(j1);
    int j2b = bd(j2);
    int j2b_for = fwd(j2b);
    int j2b_bac = bwd(j2b);
    int c1, c2 = LDO_NONE; // outer candidates (j2b, c) after (j2, j2_for)
    if (walk(j1b, 1) == j2b) {
        c1 = j2b_for;
    }
    else if (eb(j2_for) && eb(j2b_bac) && bd(j2_for) == j2b_bac) {
        c1 = j2b_for;
    }
    else if (eb(j2_for) && eb(j2b_for) && bd(j2_for) == j2b_for) {
        c1 = j2b_bac;
    }
    else {
        c1 = j2b_for;
        c2 = j2b_bac;
    }
    fb_forward_outer(j1, j2, j2, j2_for);
    fb_forward_outer(j1, j2, j2b, c1);
    fb_forward_outer(j1, j2, j2b, c2);
}

// inner candidates of one outer pair (j2, j1) of check_central_single_junction
DEVIN void Eng::fb_central_outer(int j1, int j2, int k1, int k2, int k2b) {
    if (!eb(j1)) {
        return;
    }
    int k2_for = fwd(k2);
    int k2b_for = fwd(k2b);
    int k2b_bac = bwd(k2b);
    int e1 = k2b_for, e2 = k2b_bac;
    if (eb(k2_for)) {
        int k2fb = bd(k2_for);
        if (k2b_for == k2fb) {
            e1 = LDO_NONE;
        }
        if (k2b_bac == k2fb) {
            e2 = LDO_NONE;
        }
    }
    if (eb(k2_for)) {
        fb_junction(j1, j2, k2, k2_for, k1, k2);
    }
    if (eb(e1)) {
        fb_junction(j1, j2, k2b, e1, k1, k2);
    }
    if (eb(e2)) {
        fb_junction(j1, j2, k2b, e2, k1, k2);
    }
}

// check_central_single_junction (:846-934)
DEVIN void Eng::fb_central_single_junction(int a, int b) {
    int k1 = a, k2 = b;
    int k1b = bd(k1), k2b = bd(k2);
    if (cslot_of(k1b) == cslot_of(k2b)) {
        int dd = d_of(k1b) - d_of(k2b);
        if (dd == 1 || dd == -1) {
            return;
        }
    }
    int k1_bac = bwd(k1);
    int k1b_for = fwd(k1b);
    int k1b_bac = bwd(k1b);
    int e1 = k1b_for, e2 = k1b_bac; // outer candidates (k1b, e) after (k1, k1_bac)
    if (eb(k1_bac)) {
        int k1bacb = bd(k1_bac);
        if (k1b_for == k1bacb) {
            e1 = LDO_NONE;
        }
        if (k1b_bac == k1bacb) {
            e2 = LDO_NONE;
        }
    }
    fb_central_outer(k1_bac, k1, k1, k2, k2b);
    fb_central_outer(e1, k1b, k1, k2, k2b);
    fb_central_outer(e2, k1b, k1, k2, k2b);
}

// check_regular_pair_constraints (:288-320)
DEVIN void Eng::fb_regular_pair(int a, int b, int i) {
    I3 ndr = sub3(pos(b), pos(a));
    if (!kink_ok(a, ndr, b)) {
        fb_viol = true;
        return;
    }
    if (pair_stacked(a, b)) {
        dc_e += stk(a, b);
        dc_st += 1;
        if (i == -1) {
            fb_backward_single_junction(a, b);
        }
        else {
            fb_forward_single_junction(a, b);
        }
    }
    else {
        fb_central_single_junction(a, b);
    }
    if (i == -1) {
        fb_backward_triplets(a, b);
    }
    else {
        fb_forward_triplets(a, b);
    }
}

// check_doubly_contig_helix_pair (:322-357)
DEVIN void Eng::fb_helix_pair(int a, int b, int i, int j) {
    if (j == 1) {
        return;
    }
    I3 ndr = sub3(pos(b), pos(a));
    I3 oa = ore(a);
    if (eq3(ndr, oa) || eq3(ndr, neg3(oa))) {
        fb_viol = true;
        return;
    }
    if (twist_ok(a, ndr, b)) {
        dc_e += stk(a, b);
        dc_st += 1;
    }
    else {
        fb_viol = true;
        return;
    }
    if (i == -1) {
        fb_backward_triplets(a, b);
        fb_backward_single_junction(a, b);
    }
    else {
        fb_forward_triplets(a, b);
        fb_forward_single_junction(a, b);
    }
}

// JunctionBindingPotential::check_constraints (:231-286)
DEVIN void Eng::fb_around(int cd, int j) {
    for (int i = -1; i <= 0; i++) {
        int a = walk(cd, i);
        int b = walk(cd, i + 1);
        if (!(eb(a) && eb(b))) {
            continue;
        }
        int ba = bd(a), bb = bd(b);
        if (cslot_of(ba) == cslot_of(bb)) {
            if (d_of(ba) == d_of(bb) - 1) {
                fb_helix_pair(a, b, i, j);
            }
            else if (d_of(ba) == d_of(bb) + 1) {
                fb_junction_pair(a, b, j);
            }
            else {
                fb_regular_pair(a, b, i);
            }
        }
        else {
            fb_regular_pair(a, b, i);
        }
        if (fb_viol) {
            return;
        }
    }
    int pv = walk(cd, -1);
    int nx = walk(cd, 1);
    if (eb(pv) && eb(nx)) {
        if (pair_stacked(cd, nx)) {
            if (pair_stacked(pv, cd)) {
                if (doubly_contig(pv, cd) && doubly_contig(cd, nx)) {
                    fb_triply(pv, cd, nx);
                    if (fb_viol) {
                        return;
                    }
                }
                fb_double(pv, cd, nx);
            }
            else {
                fb_single(pv, cd, nx);
            }
        }
    }
}

// calc_stacking_and_steric_terms (:212-229)
DEVNI void Eng::fb_terms(int di, int dj) {
    fb_around(di, 0);
    if (fb_viol) {
        return;
    }
    fb_around(dj, 1);
    if (fb_viol) {
        return;
    }
    fb_central_triplets(di, dj);
}

// OrigamiPotential::bind_domain (:1282-1312), BindingPotential::bind_domains
// (:131-147) and OpposingMisbindingPotential / Disallowed (:938-950)
DEVIN Delta Eng::bind_domain(int di) {
    Delta r;
    r.e = 0;
    r.st = 0;
    pot_viol = false;
    int dj = bd(di);
    S->site_bind++;
    if (ident_of(di) == -ident_of(dj)) {
        dc_e = 0;
        dc_st = 0;
        fb_viol = false;
        if (dom[di].ore != (dom[dj].ore ^ 1)) {
            fb_viol = true;
        }
        else {
            fb_terms(di, dj);
            if (!fb_viol) {
                dc_e += hyb(di, dj);
                r.e = dc_e;
                r.st = dc_st;
            }
        }
        pot_viol = fb_viol;
    }
    else if (mispot == 0) {
        if (dom[di].ore != (dom[dj].ore ^ 1)) {
            pot_viol = true;
        }
        else {
            r.e += hyb(di, dj);
        }
    }
    else {
        pot_viol = true;
    }
    return r;
}

#endif
