// oracle_core.cpp - CPU restatement of the LatticeDNAOrigami hot path.
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).
#include "oracle_core.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "nn_thermo.hpp"

namespace orc {

const V3 kUnit[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

static V3 absv(V3 v) { return {std::abs(v.x), std::abs(v.y), std::abs(v.z)}; }

// utility.cpp:68-85 - half turn about the absolute axis
V3 rotate_half(V3 v, V3 axis) {
    V3 a = absv(axis);
    V3 r = v;
    if (a == V3 {1, 0, 0}) {
        r.y = -r.y;
        r.z = -r.z;
    }
    else if (a == V3 {0, 1, 0}) {
        r.x = -r.x;
        r.z = -r.z;
    }
    else if (a == V3 {0, 0, 1}) {
        r.x = -r.x;
        r.y = -r.y;
    }
    return r;
}

// utility.cpp:102-142 - quarter turns with the reference's sign rules
V3 rotate_turns(V3 v, V3 axis, int turns) {
    if (turns % 2 == 0) {
        return rotate_half(v, axis);
    }
    bool odd_turns_even = ((turns - 1) / 2 % 2 == 0);
    int dir = (turns >= 0 && odd_turns_even) ? 1 : -1;
    V3 a = absv(axis);
    if (axis != a) {
        dir = -dir;
    }
    V3 r = v;
    if (a == V3 {1, 0, 0}) {
        std::swap(r.y, r.z);
        r.y *= -dir;
        r.z *= dir;
    }
    else if (a == V3 {0, 1, 0}) {
        std::swap(r.x, r.z);
        r.x *= -dir;
        r.z *= dir;
    }
    else if (a == V3 {0, 0, 1}) {
        std::swap(r.x, r.y);
        r.x *= -dir;
        r.y *= dir;
    }
    return r;
}

// domain.cpp:9-31
Dom* walk(Dom* d, int incr) {
    while (d != nullptr && incr > 0) {
        d = d->fwd;
        incr--;
    }
    while (d != nullptr && incr < 0) {
        d = d->bwd;
        incr++;
    }
    return d;
}

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
// 1, 2, 3"): 10 rounds of two 32x32->64 multiplies with Weyl key schedule.
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
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
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

void Rng::init(int kind, uint64_t seed, uint32_t stream) {
    m_kind = kind;
    m_seed = seed;
    m_stream = stream;
    m_step = 0;
    m_idx = 0;
    if (kind == 1) {
        // RandomGens::set_seed(int) (random_gens.cpp:25)
        m_mt.seed((uint64_t)(int64_t)(int)seed);
    }
}

uint64_t Rng::philox64() {
    uint32_t ctr[4] = {m_idx, m_sub, (uint32_t)(uint64_t)m_step,
                       (uint32_t)((uint64_t)m_step >> 32)};
    uint32_t key[2] = {(uint32_t)m_seed, m_stream};
    uint32_t out[4];
    philox4x32_10(ctr, key, out);
    m_idx++;
    return (uint64_t)out[0] | ((uint64_t)out[1] << 32);
}

uint64_t Rng::replay_next() {
    if (m_replay_pos >= m_replay.size()) {
        throw OrigamiError(LDO_E_MISUSE, "replayed tape exhausted");
    }
    return m_replay[m_replay_pos++];
}

double Rng::uniform_real() {
    double u;
    if (m_kind == 2) {
        uint64_t b = replay_next();
        std::memcpy(&u, &b, sizeof u);
    }
    else if (m_kind == 1) {
        std::uniform_real_distribution<double> dist;
        u = dist(m_mt);
    }
    else {
        u = (double)(philox64() >> 11) * 0x1.0p-53;
    }
    if (m_record) {
        uint64_t bits;
        std::memcpy(&bits, &u, sizeof bits);
        tape.push_back(bits);
    }
    return u;
}

int Rng::uniform_int(int lo, int hi) {
    if (hi < lo) {
        throw OrigamiError(LDO_E_MISUSE, "uniform_int with empty range");
    }
    int v;
    if (m_kind == 2) {
        v = (int)(int64_t)replay_next();
    }
    else if (m_kind == 1) {
        std::uniform_int_distribution<int> dist {lo, hi};
        v = dist(m_mt);
    }
    else {
        uint64_t x = philox64();
        uint64_t n = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
        v = lo + (int)((((x >> 32) & 0xFFFFFFFFull) * n) >> 32);
    }
    if (m_record) {
        tape.push_back((uint64_t)(int64_t)v);
    }
    return v;
}

// ----------------------------------------------------------------- energies
static bool eb(const Dom* d) { return d != nullptr && d->st == Occ::bound; }

void Potential::setup(const ldo_system_desc& sys,
                      const std::vector<std::vector<int>>& idents,
                      const std::vector<std::vector<std::string>>& seqs) {
    m_sys = sys;
    m_idents = idents;
    m_seqs = seqs;
    domain_type = sys.domain_type;
    misbinding = sys.misbinding_pot;
    int lo = 0, hi = 0;
    bool first = true;
    for (auto& c: idents) {
        for (int v: c) {
            if (first) {
                lo = hi = v;
                first = false;
            }
            lo = std::min(lo, v);
            hi = std::max(hi, v);
        }
    }
    ident_lo = lo;
    n_ident = hi - lo + 1;
    if (sys.hybridization_pot == LDO_HYB_SPECIFIED) {
        for (int i = 0; i < sys.n_specified; i++) {
            m_spec_h.push_back(sys.enthalpies[i]);
            m_spec_s.push_back(sys.entropies[i]);
        }
    }
}

// OrigamiPotential::calc_energies and helpers (origami_potential.cpp:1057-1221)
Tables Potential::build(double temp, double stacking_mult) const {
    Tables t;
    size_t n2 = (size_t)n_ident * n_ident;
    t.hyb_e.assign(n2, 0.0);
    t.hyb_h.assign(n2, 0.0);
    t.hyb_s.assign(n2, 0.0);
    t.stack.assign(n2, 0.0);
    orc_nn::HS init = orc_nn::reduced_initiation(temp);
    t.init_h = init.h;
    t.init_s = init.s;
    t.init_e = init.h - init.s;
    double stacking_ene = m_sys.stacking_ene * stacking_mult;
    bool mf = m_sys.apply_mean_field_cor != 0;
    for (size_t ci = 0; ci < m_idents.size(); ci++) {
        for (size_t cj = 0; cj < m_idents.size(); cj++) {
            for (size_t di = 0; di < m_idents[ci].size(); di++) {
                for (size_t dj = 0; dj < m_idents[cj].size(); dj++) {
                    int ii = m_idents[ci][di];
                    int ij = m_idents[cj][dj];
                    size_t k = (size_t)(ii - ident_lo) * n_ident + (ij - ident_lo);
                    bool comp = (ii == -ij);
                    double h = 0, s = 0;
                    if (m_sys.hybridization_pot == LDO_HYB_NEARESTNEIGHBOUR) {
                        const std::string& a = m_seqs[ci][di];
                        const std::string& b = m_seqs[cj][dj];
                        auto comps = orc_nn::longest_contig_complements(a, b);
                        if (!comps.empty()) {
                            int n = 0;
                            for (auto& cs: comps) {
                                orc_nn::HS hs = orc_nn::reduced_duplex(cs, temp, m_sys.cation_M);
                                h += hs.h - t.init_h;
                                s += hs.s - t.init_s;
                                n++;
                            }
                            h /= n;
                            s /= n;
                            s += std::log(6);
                            if (comp) {
                                if (mf) {
                                    s += 3 * std::log(6);
                                }
                                if (comps[0].size() != a.size() && a.size() == b.size()) {
                                    throw OrigamiError(LDO_E_CONSTRAINT,
                                                       "Sequences that should be complementary are not");
                                }
                            }
                            else if (comps[0].size() == a.size() && a.size() == b.size()) {
                                throw OrigamiError(LDO_E_CONSTRAINT,
                                                   "Sequences that should not be complementary are");
                            }
                        }
                    }
                    else {
                        if (comp) {
                            if (m_sys.hybridization_pot == LDO_HYB_UNIFORM) {
                                h = m_sys.binding_h / temp;
                                s = m_sys.binding_s;
                            }
                            else {
                                int idx = std::abs(ii) - 1;
                                h = m_spec_h.at(idx) / temp;
                                s = m_spec_s.at(idx);
                            }
                            if (mf) {
                                s += 3 * std::log(6);
                            }
                        }
                        else {
                            h = m_sys.misbinding_h / temp;
                            s = m_sys.misbinding_s;
                        }
                        s += std::log(6);
                    }
                    t.hyb_h[k] = h;
                    t.hyb_s[k] = s;
                    t.hyb_e[k] = h - s;
                    if (m_sys.stacking_pot == LDO_STACK_SEQUENCESPECIFIC) {
                        t.stack[k] = orc_nn::seq_stacking_energy(m_seqs[ci][di], m_seqs[cj][dj]);
                    }
                    else {
                        t.stack[k] = stacking_ene / temp;
                    }
                }
            }
        }
    }
    return t;
}

// OrigamiPotential::update_temp (origami_potential.cpp:1019-1042).  On a
// table-cache hit the initiation terms keep the values of the last table
// build, as in the reference (calc_energies is the only writer of them).
void Potential::update_temp(double temp, double stacking_mult) {
    auto key = std::make_pair(temp, stacking_mult);
    auto it = m_cache.find(key);
    if (it == m_cache.end()) {
        Tables t = build(temp, stacking_mult);
        m_cache[key] = t;
        m_cur = t;
    }
    else {
        double ih = m_cur.init_h, is = m_cur.init_s, ie = m_cur.init_e;
        m_cur = it->second;
        m_cur.init_h = ih;
        m_cur.init_s = is;
        m_cur.init_e = ie;
    }
}

// domain.cpp:33-57, 69-97 (twist and kink rules by domain type)
bool Potential::twist_ok(const Dom& a, V3 ndr, const Dom& b) const {
    V3 rot = (domain_type == LDO_HALFTURN) ? rotate_half(a.ore, ndr)
                                           : rotate_turns(a.ore, ndr, -1);
    return rot == b.ore;
}

bool Potential::kink_ok(const Dom& a, V3 ndr, const Dom& b) const {
    if (ndr == -a.ore) {
        return false;
    }
    if (ndr == a.ore) {
        if (domain_type == LDO_HALFTURN) {
            return !(b.ore == -a.ore);
        }
        return !(a.ore == b.ore || a.ore == -b.ore);
    }
    if (ndr == b.ore || ndr == -b.ore) {
        return false;
    }
    return true;
}

// domain.cpp:59-67, 99-118
bool Potential::junction_ok(const Dom& j1, const Dom& j2, const Dom&, const Dom&,
                            const Dom& k1, const Dom& k2) const {
    if (domain_type == LDO_HALFTURN) {
        return true;
    }
    V3 ndr_k1 = k2.pos - k1.pos;
    if (ndr_k1 == k1.ore) {
        V3 ndr_1 = j2.pos - j1.pos;
        if (j1.d > j2.d) {
            ndr_1 = -ndr_1;
        }
        if (!twist_ok(k1, ndr_1, k2)) {
            return false;
        }
    }
    return true;
}

// doubly_contiguous (origami_potential.cpp:51-68)
static bool doubly_contig(const Dom* a, const Dom* b) {
    if (a->c != b->c || b->d != a->d + 1) {
        return false;
    }
    const Dom* ba = a->bd;
    const Dom* bb = b->bd;
    if (ba->c != bb->c) {
        return false;
    }
    return std::abs(bb->d - ba->d) == 1;
}

// check_pair_stacked (origami_potential.cpp:70-83)
bool Potential::pair_stacked(Dom* a, Dom* b) const {
    if (a->d > b->d) {
        std::swap(a, b);
    }
    V3 ndr = b->pos - a->pos;
    if (ndr != a->ore && ndr != -a->ore) {
        return twist_ok(*a, ndr, *b);
    }
    return false;
}

// check_junction_stacking_penalty (origami_potential.cpp:85-127)
static int junction_penalty(const Dom& j1, const Dom& j2, const Dom& j3, const Dom& j4,
                            const Dom& k1, const Dom& k2) {
    V3 ndr_k = k2.pos - k1.pos;
    if (ndr_k != k1.ore) {
        return 0;
    }
    V3 n1 = j2.pos - j1.pos;
    if (j1.d > j2.d) {
        n1 = -n1;
    }
    V3 n3 = j4.pos - j3.pos;
    if (j3.d > j4.d) {
        n3 = -n3;
    }
    if (n1 == n3) {
        return 2;
    }
    if (n1 == -n3) {
        return 0;
    }
    return 1;
}

// BindingPotential::bind_domains (origami_potential.cpp:131-147)
Delta Potential::fb_bind(Dom& di, Dom& dj) {
    m_dc = {};
    m_fb_violated = false;
    if (di.ore != -dj.ore) {
        m_fb_violated = true;
        return {};
    }
    fb_terms(di, dj);
    if (m_fb_violated) {
        return {};
    }
    m_dc.e += hyb_e(di, dj);
    return m_dc;
}

// BindingPotential::check_stacking (origami_potential.cpp:149-156)
Delta Potential::fb_check(Dom& di, Dom& dj) {
    m_dc = {};
    m_fb_violated = false;
    fb_terms(di, dj);
    return m_dc;
}

// origami_potential.cpp:158-179
void Potential::fb_single(Dom* h1, Dom* h2, Dom* h3) {
    V3 n1 = h2->pos - h1->pos;
    if (n1 == h1->ore) {
        return;
    }
    V3 n2 = h3->pos - h2->pos;
    if (n1 != n2) {
        m_dc.e -= stack(*h2, *h3);
        m_dc.stacked -= 1;
    }
}

// origami_potential.cpp:181-198
void Potential::fb_double(Dom* h1, Dom* h2, Dom* h3) {
    V3 n1 = h2->pos - h1->pos;
    V3 n2 = h3->pos - h2->pos;
    if (n1 != n2) {
        m_dc.e -= stack(*h1, *h2) / 2;
        m_dc.e -= stack(*h2, *h3) / 2;
        m_dc.stacked -= 1;
    }
}

// origami_potential.cpp:200-210
void Potential::fb_triply(Dom* h1, Dom* h2, Dom* h3) {
    V3 n1 = h2->pos - h1->pos;
    V3 n2 = h3->pos - h2->pos;
    if (n1 != n2) {
        m_fb_violated = true;
    }
}

// JunctionBindingPotential::calc_stacking_and_steric_terms (:212-229)
void Potential::fb_terms(Dom& di, Dom& dj) {
    fb_around(&di, 0);
    if (m_fb_violated) {
        return;
    }
    fb_around(&dj, 1);
    if (m_fb_violated) {
        return;
    }
    fb_central_triplets(di, dj);
}

// JunctionBindingPotential::check_constraints (:231-286)
void Potential::fb_around(Dom* cd, int j) {
    for (int i = -1; i <= 0; i++) {
        Dom* a = walk(cd, i);
        Dom* b = walk(cd, i + 1);
        if (!(eb(a) && eb(b))) {
            continue;
        }
        Dom& ba = *a->bd;
        Dom& bb = *b->bd;
        if (ba.c == bb.c) {
            if (ba.d == bb.d - 1) {
                fb_helix_pair(a, b, i, j);
            }
            else if (ba.d == bb.d + 1) {
                fb_junction_pair(a, b, j);
            }
            else {
                fb_regular_pair(a, b, i);
            }
        }
        else {
            fb_regular_pair(a, b, i);
        }
        if (m_fb_violated) {
            return;
        }
    }
    Dom* prev = walk(cd, -1);
    Dom* next = walk(cd, 1);
    if (eb(prev) && eb(next)) {
        if (pair_stacked(cd, next)) {
            if (pair_stacked(prev, cd)) {
                if (doubly_contig(prev, cd) && doubly_contig(cd, next)) {
                    fb_triply(prev, cd, next);
                    if (m_fb_violated) {
                        return;
                    }
                }
                fb_double(prev, cd, next);
            }
            else {
                fb_single(prev, cd, next);
            }
        }
    }
}

// check_regular_pair_constraints (:288-320)
void Potential::fb_regular_pair(Dom* a, Dom* b, int i) {
    V3 ndr = b->pos - a->pos;
    if (!kink_ok(*a, ndr, *b)) {
        m_fb_violated = true;
        return;
    }
    if (pair_stacked(a, b)) {
        m_dc.e += stack(*a, *b);
        m_dc.stacked += 1;
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
void Potential::fb_helix_pair(Dom* a, Dom* b, int i, int j) {
    if (j == 1) {
        return;
    }
    V3 ndr = b->pos - a->pos;
    if (ndr == a->ore || ndr == -a->ore) {
        m_fb_violated = true;
        return;
    }
    if (twist_ok(*a, ndr, *b)) {
        m_dc.e += stack(*a, *b);
        m_dc.stacked += 1;
    }
    else {
        m_fb_violated = true;
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

// check_doubly_contig_junction_pair (:359-417)
void Potential::fb_junction_pair(Dom* a, Dom* b, int j) {
    Dom* k1 = a;
    Dom* k2 = b;
    V3 ndr = k2->pos - k1->pos;
    if (k1->ore != ndr) {
        m_fb_violated = true;
        return;
    }
    if (j == 1) {
        return;
    }
    std::pair<Dom*, Dom*> first_sel[2] = {{walk(a, -1), a}, {walk(a->bd, 1), a->bd}};
    for (auto& sel: first_sel) {
        Dom* j1 = sel.first;
        Dom* j2 = sel.second;
        Dom* j3 = k2;
        Dom* j4 = walk(k2, 1);
        if (eb(j4) && eb(j1)) {
            fb_junction(j1, j2, j3, j4, k1, k2);
        }
        j3 = k2->bd;
        j4 = walk(j3, -1);
        if (eb(j4) && eb(j1)) {
            fb_junction(j1, j2, j3, j4, k1, k2);
        }
    }
}

// check_junction (:419-462)
void Potential::fb_junction(Dom* j1, Dom* j2, Dom* j3, Dom* j4, Dom* k1, Dom* k2) {
    if (k1->d > k2->d) {
        std::swap(k1, k2);
        std::swap(j1, j4);
        std::swap(j2, j3);
    }
    if (!(pair_stacked(j1, j2) && pair_stacked(j3, j4))) {
        return;
    }
    if (!junction_ok(*j1, *j2, *j3, *j4, *k1, *k2)) {
        m_fb_violated = true;
        return;
    }
    int pen = junction_penalty(*j1, *j2, *j3, *j4, *k1, *k2);
    if (pen == 1) {
        m_dc.e -= stack(*j1, *j2) / 2;
        m_dc.e -= stack(*j3, *j4) / 2;
        m_dc.stacked -= 1;
    }
    else if (pen == 2) {
        m_dc.e -= stack(*j1, *j2);
        m_dc.e -= stack(*j3, *j4);
        m_dc.stacked -= 2;
    }
}

// check_backward_triplet_stacking_combos (:464-519)
void Potential::fb_backward_triplets(Dom* a, Dom* b) {
    Dom* h2 = a;
    Dom* h3 = b;
    if (!pair_stacked(h2, h3)) {
        return;
    }
    Dom* h1 = walk(a, -1);
    bool dc23 = doubly_contig(h2, h3);
    if (eb(h1)) {
        bool dc12 = doubly_contig(h1, h2);
        if (pair_stacked(h1, h2)) {
            if (dc12 && dc23) {
                fb_triply(h1, h2, h3);
                if (m_fb_violated) {
                    return;
                }
            }
            fb_double(h1, h2, h3);
        }
        else {
            fb_single(h1, h2, h3);
        }
    }
    Dom* h2_prev = h1;
    h1 = walk(a->bd, 1);
    if (eb(h1) && h1->bd != h2_prev && h1->bd != h3) {
        if (pair_stacked(a->bd, h1)) {
            fb_double(h1, h2, h3);
        }
    }
    h1 = walk(a->bd, -1);
    if (eb(h1) && h1->bd != h2_prev && h1->bd != h3) {
        if (pair_stacked(h1, a->bd)) {
            fb_double(h1, h2, h3);
        }
        else {
            fb_single(h1, a->bd, h3);
        }
    }
}

// check_forward_triplet_stacking_combos (:521-585)
void Potential::fb_forward_triplets(Dom* a, Dom* b) {
    Dom* h2 = b;
    Dom* h1 = a;
    bool first_stacked = pair_stacked(h1, h2);
    Dom* h3 = walk(b, 1);
    bool dc12 = doubly_contig(h1, h2);
    if (eb(h3)) {
        bool second_stacked = pair_stacked(h2, h3);
        if (first_stacked && second_stacked) {
            bool dc23 = doubly_contig(h2, h3);
            if (dc12 && dc23) {
                fb_triply(h1, h2, h3);
                if (m_fb_violated) {
                    return;
                }
            }
            fb_double(h1, h2, h3);
        }
        else if (second_stacked) {
            fb_single(h1, h2, h3);
        }
    }
    Dom* h2_next = h3;
    h3 = walk(b->bd, 1);
    if (eb(h3) && h3->bd != h2_next && h3->bd != h1) {
        if (pair_stacked(b->bd, h3)) {
            if (first_stacked) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, h2, h3);
            }
        }
    }
    h3 = walk(b->bd, -1);
    if (eb(h3) && h3->bd != h2_next && h3->bd != h1) {
        if (pair_stacked(h3, b->bd)) {
            if (first_stacked) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, h2, h3);
            }
        }
        else if (first_stacked) {
            fb_single(h3, b->bd, h1);
        }
    }
}

// check_central_triplet_stacking_combos (:587-652)
void Potential::fb_central_triplets(Dom& di, Dom& dj) {
    Dom* h1 = walk(&di, -1);
    Dom* h2 = &di;
    Dom* h3 = walk(&dj, 1);
    Dom* h2_next = walk(&di, 1);
    if (eb(h1) && eb(h3) && h3->bd != h2_next && !doubly_contig(h1, h2)) {
        if (pair_stacked(&dj, h3)) {
            if (pair_stacked(h1, h2)) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, h2, h3);
            }
        }
    }
    h3 = walk(&dj, -1);
    if (eb(h1) && eb(h3) && h3->bd != h1 && !doubly_contig(h1, h2)) {
        if (pair_stacked(h1, h2)) {
            if (pair_stacked(h3, &dj)) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h3, &dj, h1);
            }
        }
        else if (pair_stacked(h3, &dj)) {
            fb_single(h1, h2, h3);
        }
    }
    h2 = &di;
    h3 = walk(&di, 1);
    h1 = walk(&dj, 1);
    Dom* h2_prev = walk(&di, -1);
    if (eb(h1) && eb(h3) && h1->bd != h3 && !doubly_contig(h2, h3)) {
        if (pair_stacked(&dj, h1) && pair_stacked(h2, h3)) {
            fb_double(h1, h2, h3);
        }
    }
    h1 = walk(&dj, -1);
    if (eb(h1) && eb(h3) && h1->bd != h2_prev && !doubly_contig(h2, h3)) {
        if (pair_stacked(h2, h3)) {
            if (pair_stacked(h1, &dj)) {
                fb_double(h1, h2, h3);
            }
            else {
                fb_single(h1, &dj, h3);
            }
        }
    }
}

// check_backward_single_junction (:654-748)
void Potential::fb_backward_single_junction(Dom* a, Dom* b) {
    Dom* j3 = a;
    Dom* j4 = b;
    std::vector<std::pair<Dom*, Dom*>> first_sel;
    Dom* j3_bac = j3->bwd;
    first_sel.push_back({j3, j3_bac});
    Dom* j3b = j3->bd;
    Dom* j3b_for = j3b->fwd;
    Dom* j3b_bac = j3b->bwd;
    if (j3b_for == j4->bd) {
        first_sel.push_back({j3b, j3b_bac});
    }
    else if (eb(j3_bac) && eb(j3b_bac) && j3_bac->bd == j3b_bac) {
        first_sel.push_back({j3b, j3b_for});
    }
    else if (eb(j3_bac) && eb(j3b_for) && j3_bac->bd == j3b_for) {
        first_sel.push_back({j3b, j3b_bac});
    }
    else {
        first_sel.push_back({j3b, j3b_for});
        first_sel.push_back({j3b, j3b_bac});
    }
    for (auto& s1: first_sel) {
        Dom* k2 = s1.first;
        Dom* k1 = s1.second;
        if (!eb(k1)) {
            continue;
        }
        if (pair_stacked(k1, k2)) {
            continue;
        }
        std::vector<std::pair<Dom*, Dom*>> second_sel;
        int dir = k1->d - k2->d;
        Dom* k1_next = walk(k1, dir);
        second_sel.push_back({k1, k1_next});
        Dom* k1b = k1->bd;
        Dom* k1b_for = k1b->fwd;
        Dom* k1b_bac = k1b->bwd;
        if (eb(k1_next)) {
            Dom* k1nb = k1_next->bd;
            if (k1b_for != k1nb) {
                second_sel.push_back({k1b, k1b_for});
            }
            if (k1b_bac != k1nb) {
                second_sel.push_back({k1b, k1b_bac});
            }
        }
        else {
            second_sel.push_back({k1b, k1b_for});
            second_sel.push_back({k1b, k1b_bac});
        }
        for (auto& s2: second_sel) {
            Dom* j2 = s2.first;
            Dom* j1 = s2.second;
            if (!eb(j1)) {
                continue;
            }
            fb_junction(j1, j2, j3, j4, k1, k2);
        }
    }
}

// check_forward_single_junction (:750-844)
void Potential::fb_forward_single_junction(Dom* a, Dom* b) {
    Dom* j1 = a;
    Dom* j2 = b;
    std::vector<std::pair<Dom*, Dom*>> first_sel;
    Dom* j2_for = j2->fwd;
    first_sel.push_back({j2, j2_for});
    Dom* j1b = j1->bd;
    Dom* j2b = j2->bd;
    Dom* j2b_for = j2b->fwd;
    Dom* j2b_bac = j2b->bwd;
    if (walk(j1b, 1) == j2b) {
        first_sel.push_back({j2b, j2b_for});
    }
    else if (eb(j2_for) && eb(j2b_bac) && j2_for->bd == j2b_bac) {
        first_sel.push_back({j2b, j2b_for});
    }
    else if (eb(j2_for) && eb(j2b_for) && j2_for->bd == j2b_for) {
        first_sel.push_back({j2b, j2b_bac});
    }
    else {
        first_sel.push_back({j2b, j2b_for});
        first_sel.push_back({j2b, j2b_bac});
    }
    for (auto& s1: first_sel) {
        Dom* k1 = s1.first;
        Dom* k2 = s1.second;
        if (!eb(k2)) {
            continue;
        }
        if (pair_stacked(k1, k2)) {
            continue;
        }
        std::vector<std::pair<Dom*, Dom*>> second_sel;
        int dir = k2->d - k1->d;
        Dom* k2_next = walk(k2, dir);
        second_sel.push_back({k2, k2_next});
        Dom* k2b = k2->bd;
        Dom* k2b_for = k2b->fwd;
        Dom* k2b_bac = k2b->bwd;
        if (eb(k2_next)) {
            Dom* k2nb = k2_next->bd;
            if (k2b_for != k2nb) {
                second_sel.push_back({k2b, k2b_for});
            }
            if (k2b_bac != k2nb) {
                second_sel.push_back({k2b, k2b_bac});
            }
        }
        else {
            second_sel.push_back({k2b, k2b_for});
            second_sel.push_back({k2b, k2b_bac});
        }
        for (auto& s2: second_sel) {
            Dom* j3 = s2.first;
            Dom* j4 = s2.second;
            if (!eb(j4)) {
                continue;
            }
            fb_junction(j1, j2, j3, j4, k1, k2);
        }
    }
}

// check_central_single_junction (:846-934)
void Potential::fb_central_single_junction(Dom* a, Dom* b) {
    Dom* k1 = a;
    Dom* k2 = b;
    if (k1->bd->c == k2->bd->c && std::abs(k1->bd->d - k2->bd->d) == 1) {
        return;
    }
    std::vector<std::pair<Dom*, Dom*>> first_sel;
    Dom* k1_bac = k1->bwd;
    first_sel.push_back({k1, k1_bac});
    Dom* k1b = k1->bd;
    Dom* k1b_for = k1b->fwd;
    Dom* k1b_bac = k1b->bwd;
    if (eb(k1_bac)) {
        Dom* k1bacb = k1_bac->bd;
        if (k1b_for != k1bacb) {
            first_sel.push_back({k1b, k1b_for});
        }
        if (k1b_bac != k1bacb) {
            first_sel.push_back({k1b, k1b_bac});
        }
    }
    else {
        first_sel.push_back({k1b, k1b_for});
        first_sel.push_back({k1b, k1b_bac});
    }
    for (auto& s1: first_sel) {
        Dom* j2 = s1.first;
        Dom* j1 = s1.second;
        if (!eb(j1)) {
            continue;
        }
        std::vector<std::pair<Dom*, Dom*>> second_sel;
        Dom* k2_for = k2->fwd;
        second_sel.push_back({k2, k2_for});
        Dom* k2b = k2->bd;
        Dom* k2b_for = k2b->fwd;
        Dom* k2b_bac = k2b->bwd;
        if (eb(k2_for)) {
            Dom* k2fb = k2_for->bd;
            if (k2b_for != k2fb) {
                second_sel.push_back({k2b, k2b_for});
            }
            if (k2b_bac != k2fb) {
                second_sel.push_back({k2b, k2b_bac});
            }
        }
        else {
            second_sel.push_back({k2b, k2b_for});
            second_sel.push_back({k2b, k2b_bac});
        }
        for (auto& s2: second_sel) {
            Dom* j3 = s2.first;
            Dom* j4 = s2.second;
            if (!eb(j4)) {
                continue;
            }
            fb_junction(j1, j2, j3, j4, k1, k2);
        }
    }
}

// OrigamiPotential::bind_domain (:1282-1312) with the misbinding potentials
// (:938-950)
Delta Potential::bind_domain(Dom& di) {
    violated = false;
    Dom& dj = *di.bd;
    Delta dc {};
    if (di.d_ident == -dj.d_ident) {
        dc = fb_bind(di, dj);
        violated = m_fb_violated;
    }
    else if (misbinding == LDO_MISBIND_OPPOSING) {
        if (di.ore != -dj.ore) {
            violated = true;
        }
        else {
            dc.e += hyb_e(di, dj);
        }
    }
    else {
        violated = true;
    }
    return dc;
}

Delta Potential::check_stacking(Dom& di, Dom& dj) { return fb_check(di, dj); }

// ------------------------------------------------------------------- system
// molarity_to_chempots / chempots_to_reduced_chempots (origami_system.cpp:965-989)
System::System(const ldo_system_desc& sd, const std::vector<std::vector<int>>& idents,
               const std::vector<std::vector<std::string>>& seqs,
               const std::vector<ChainSpec>& cfg, double t):
        identities {idents},
        cyclic {sd.cyclic != 0},
        temp {t},
        staple_M {sd.staple_M},
        reduced_fugacity {sd.staple_M} {
    if (sd.apply_mean_field_cor) {
        throw OrigamiError(LDO_E_NOTIMPL, "apply_mean_field_cor is not supported");
    }
    for (size_t i = 1; i < idents.size(); i++) {
        double extra = (2 * idents[i].size() - 1) * std::log(6);
        staple_us.push_back(t * (std::log(sd.staple_M) - extra));
    }
    for (double u: staple_us) {
        reduced_staple_us.push_back(u / t);
    }
    pot.setup(sd, idents, seqs);
    // the constructor's get_energies() does not enter the per-temperature
    // table cache (origami_potential.cpp:1011, 1019-1042)
    pot.set_uncached(t);
    ident_to_index.push_back({});
    for (size_t i = 1; i < idents.size(); i++) {
        ident_to_index.push_back({});
    }
    initialize_scaffold(cfg.at(0));
    initialize_staples(cfg);
    set_all_domains(cfg);
}

System::~System() {
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            delete d;
        }
    }
}

Occ System::occupancy(V3 pos) const {
    auto it = pos_occ.find(pos);
    return it == pos_occ.end() ? Occ::unassigned : it->second;
}

std::vector<Dom*>& System::chain_by_index(int c_i) {
    for (size_t i = 0; i < chain_indices.size(); i++) {
        if (chain_indices[i] == c_i) {
            return chains[i];
        }
    }
    throw OrigamiError(LDO_E_MISUSE, "no such chain index");
}

std::vector<int> System::staple_counts() const {
    std::vector<int> counts(identities.size() - 1, 0);
    for (size_t i = 1; i < chains.size(); i++) {
        counts[chains[i][0]->c_ident - 1]++;
    }
    return counts;
}

// origami_system.cpp:680-693
void System::initialize_scaffold(const ChainSpec& c) {
    add_chain(c.identity, c.index);
    num_staples--;
    if (cyclic) {
        Dom* first = chains[0].front();
        Dom* last = chains[0].back();
        last->fwd = first;
        first->bwd = last;
    }
}

// origami_system.cpp:659-678
void System::initialize_staples(const std::vector<ChainSpec>& cfg) {
    for (size_t i = 1; i < cfg.size(); i++) {
        add_chain(cfg[i].identity, cfg[i].index);
    }
    energy += (chains.size() - 1) * pot.init_e();
    current_c_i = *std::max_element(chain_indices.begin(), chain_indices.end());
}

int System::add_chain(int ident) {
    current_c_i += 1;
    energy += pot.init_e();
    return add_chain(ident, current_c_i);
}

int System::add_chain(int ident, int c_i) {
    ident_to_index[ident].push_back(c_i);
    chain_indices.push_back(c_i);
    chain_identities.push_back(ident);
    int len = (int)identities[ident].size();
    chains.push_back({});
    num_staples++;
    Dom* prev = nullptr;
    for (int d = 0; d < len; d++) {
        Dom* dom = new Dom {c_i, ident, d, identities[ident][d], len};
        if (prev != nullptr) {
            prev->fwd = dom;
        }
        dom->bwd = prev;
        chains.back().push_back(dom);
        prev = dom;
        num_domains++;
        num_unassigned++;
    }
    return c_i;
}

// origami_system.cpp:445-472
void System::delete_chain(int c_i) {
    size_t w = 0;
    while (w < chain_indices.size() && chain_indices[w] != c_i) {
        w++;
    }
    if (w == chain_indices.size()) {
        throw OrigamiError(LDO_E_MISUSE, "delete of unknown chain");
    }
    int ident = chain_identities[w];
    auto& lst = ident_to_index[ident];
    lst.erase(std::find(lst.begin(), lst.end(), c_i));
    chain_indices.erase(chain_indices.begin() + w);
    chain_identities.erase(chain_identities.begin() + w);
    num_domains -= (int)chains[w].size();
    num_staples--;
    for (Dom* d: chains[w]) {
        delete d;
        num_unassigned--;
    }
    chains.erase(chains.begin() + w);
    energy -= pot.init_e();
}

// origami_system.cpp:828-871
Delta System::internal_check(Dom& di, V3 pos, V3 ore) {
    Occ occ = occupancy(pos);
    Delta dc {};
    switch (occ) {
    case Occ::bound:
    case Occ::misbound:
        violated = true;
        num_unassigned++;
        break;
    case Occ::unbound:
        di.pos = pos;
        di.ore = ore;
        update_occupancies(di, pos);
        ev_site_bind++;
        dc = pot.bind_domain(di);
        violated = pot.violated;
        break;
    case Occ::unassigned:
        di.pos = pos;
        di.ore = ore;
        update_occupancies(di, pos);
        ev_site_empty++;
        break;
    }
    return dc;
}

// origami_system.cpp:770-806
void System::update_occupancies(Dom& di, V3 pos) {
    Occ occ = occupancy(pos);
    if (occ == Occ::unbound) {
        Dom* dj = unbound_at(pos);
        num_bound_pairs += 1;
        Occ ns;
        if (di.d_ident == -dj->d_ident) {
            ns = Occ::bound;
            num_fully_bound_pairs += 1;
        }
        else {
            if (di.c == dj->c) {
                num_self_bound_pairs += 1;
            }
            ns = Occ::misbound;
        }
        pos_to_unbound.erase(pos);
        di.st = ns;
        dj->st = ns;
        pos_occ[pos] = ns;
        dj->bd = &di;
        di.bd = dj;
    }
    else if (occ == Occ::unassigned) {
        di.st = Occ::unbound;
        pos_occ[pos] = Occ::unbound;
        pos_to_unbound[pos] = &di;
    }
    else {
        throw OrigamiError(LDO_E_CONSTRAINT, "Trying to bind to an already bound domain");
    }
}

// origami_system.cpp:695-739
Delta System::internal_unassign(Dom& di) {
    Occ occ = di.st;
    Dom* dj = di.bd;
    Delta dc {};
    switch (occ) {
    case Occ::bound:
        num_fully_bound_pairs -= 1;
        num_bound_pairs -= 1;
        dc = pot.check_stacking(di, *dj);
        dc.e *= -1;
        dc.stacked *= -1;
        dc.e += unassign_bound(di);
        break;
    case Occ::misbound:
        num_bound_pairs -= 1;
        if (dj->c == di.c) {
            num_self_bound_pairs -= 1;
        }
        dc.e += unassign_bound(di);
        break;
    case Occ::unbound:
        unassign_unbound(di);
        break;
    case Occ::unassigned:
        num_unassigned--;
        break;
    }
    return dc;
}

// origami_system.cpp:741-760
double System::unassign_bound(Dom& di) {
    Dom& dj = *di.bd;
    double de = -pot.hyb_e(di, dj);
    di.bd = nullptr;
    dj.bd = nullptr;
    di.st = Occ::unassigned;
    pos_to_unbound[di.pos] = &dj;
    pos_occ[di.pos] = Occ::unbound;
    dj.st = Occ::unbound;
    return de;
}

void System::unassign_unbound(Dom& di) {
    pos_to_unbound.erase(di.pos);
    pos_occ.erase(di.pos);
    di.st = Occ::unassigned;
}

double System::check_domain_constraints(Dom& di, V3 pos, V3 ore) {
    Delta dc = internal_check(di, pos, ore);
    internal_unassign(di);
    return dc.e;
}

double System::unassign_domain(Dom& di) {
    if (di.st == Occ::bound) {
        ev_site_bind++; // the stacking walk of a committed unset
    }
    ev_commits++;
    Delta dc = internal_unassign(di);
    energy += dc.e;
    num_stacked += dc.stacked;
    num_unassigned++;
    return dc.e;
}

double System::set_checked_domain_config(Dom& di, V3 pos, V3 ore) {
    di.pos = pos;
    di.ore = ore;
    update_occupancies(di, pos);
    double de = 0;
    if (di.st == Occ::misbound) {
        de += pot.hyb_e(di, *di.bd);
    }
    else if (di.st == Occ::bound) {
        de += pot.hyb_e(di, *di.bd);
        ev_site_bind++;
        Delta dc = pot.check_stacking(di, *di.bd);
        de += dc.e;
        num_stacked += dc.stacked;
    }
    energy += de;
    num_unassigned--;
    ev_commits++;
    return de;
}

double System::set_domain_config(Dom& di, V3 pos, V3 ore) {
    if (di.st != Occ::unassigned) {
        throw OrigamiError(LDO_E_CONSTRAINT, "Trying to set an already assigned domain");
    }
    Delta dc = internal_check(di, pos, ore);
    if (violated) {
        internal_unassign(di);
    }
    else {
        energy += dc.e;
        num_stacked += dc.stacked;
        num_unassigned--;
        ev_commits++;
    }
    return dc.e;
}

void System::set_domain_orientation(Dom& di, V3 ore) {
    if (di.st == Occ::bound) {
        violated = true;
    }
    else {
        di.ore = ore;
    }
}

// origami_system.cpp:553-571
void System::center(int centering_domain) {
    std::unordered_map<V3, Dom*, V3Hash> new_unbound;
    std::unordered_map<V3, Occ, V3Hash> new_occ;
    V3 ref = chains[0][centering_domain]->pos;
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            V3 np = d->pos - ref;
            auto it = pos_to_unbound.find(d->pos);
            new_unbound[np] = (it == pos_to_unbound.end()) ? nullptr : it->second;
            new_occ[np] = occupancy(d->pos);
            d->pos = np;
        }
    }
    // entries for bound sites hold nullptr; unbound_at() is only called on
    // unbound sites, so they are never read
    pos_to_unbound.clear();
    for (auto& kv: new_unbound) {
        if (kv.second != nullptr) {
            pos_to_unbound[kv.first] = kv.second;
        }
    }
    pos_occ = new_occ;
}

void System::check_distance_constraints() {
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            Dom* n = walk(d, 1);
            if (n == nullptr) {
                continue;
            }
            V3 dist = n->pos - d->pos;
            if (std::abs(dist.x) + std::abs(dist.y) + std::abs(dist.z) != 1) {
                throw OrigamiError(LDO_E_CONSTRAINT, "Contiguous domains not on adjacent sites");
            }
        }
    }
}

void System::set_all_domains() {
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            set_domain_config(*d, d->pos, d->ore);
            if (violated) {
                throw OrigamiError(LDO_E_CONSTRAINT, "Constaints in violation after move complete");
            }
        }
    }
    check_distance_constraints();
}

void System::set_all_domains(const std::vector<ChainSpec>& cfg) {
    for (size_t i = 0; i < cfg.size(); i++) {
        for (size_t d = 0; d < chains[i].size(); d++) {
            chains[i][d]->pos = cfg[i].pos[d];
            chains[i][d]->ore = cfg[i].ore[d];
        }
    }
    set_all_domains();
}

void System::set_config(const std::vector<ChainSpec>& cfg) {
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            unassign_domain(*d);
        }
    }
    for (int i = (int)chains.size() - 1; i != 0; i--) {
        delete_chain(chains[i][0]->c);
    }
    initialize_staples(cfg);
    set_all_domains(cfg);
}

// origami_system.cpp:267-325
void System::check_all_constraints() {
    if (num_unassigned != 0) {
        throw OrigamiError(LDO_E_CONSTRAINT, "Number of unassigned domains is no zero on constraint check");
    }
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            if (d->st == Occ::unassigned) {
                throw OrigamiError(LDO_E_CONSTRAINT, "Domain unassigned after move complete");
            }
            unassign_domain(*d);
        }
    }
    energy -= (chains.size() - 1) * pot.init_e();
    if (num_stacked != 0) {
        set_all_domains();
        throw OrigamiError(LDO_E_CONSTRAINT, "Stack count inconsistency");
    }
    double eps = 0.000001;
    if (energy < -eps || energy > eps) {
        set_all_domains();
        throw OrigamiError(LDO_E_CONSTRAINT, "System energy inconsistency");
    }
    energy = 0;
    set_all_domains();
    energy += (chains.size() - 1) * pot.init_e();
}

void System::update_temp(double t, double stacking_mult) {
    temp = t;
    pot.update_temp(t, stacking_mult);
    update_energy();
}

void System::update_staple_us(double t, double mult) {
    for (size_t i = 0; i < staple_us.size(); i++) {
        staple_us[i] = reduced_staple_us[i] * t * mult;
    }
}

void System::update_energy() {
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            unassign_domain(*d);
        }
    }
    energy = 0;
    energy += (chains.size() - 1) * pot.init_e();
    set_all_domains();
}

// origami_system.cpp:204-246
void System::update_enthalpy_and_entropy() {
    hyb_enthalpy = 0;
    hyb_entropy = 0;
    std::set<Dom*> accounted;
    for (auto& ch: chains) {
        for (Dom* d: ch) {
            if (d->st == Occ::bound || d->st == Occ::misbound) {
                if (accounted.find(d) == accounted.end()) {
                    hyb_enthalpy += pot.hyb_h(*d, *d->bd);
                    hyb_entropy += pot.hyb_s(*d, *d->bd);
                    accounted.insert(d->bd);
                }
            }
        }
    }
    hyb_enthalpy += (chains.size() - 1) * pot.init_h();
    hyb_entropy += (chains.size() - 1) * pot.init_s();
    stacking_energy = energy - (hyb_enthalpy - hyb_entropy);
}

std::vector<ChainSpec> System::snapshot() const {
    std::vector<ChainSpec> out;
    for (size_t i = 0; i < chains.size(); i++) {
        ChainSpec c;
        c.index = chain_indices[i];
        c.identity = chain_identities[i];
        for (Dom* d: chains[i]) {
            c.pos.push_back(d->pos);
            c.ore.push_back(d->ore);
        }
        out.push_back(c);
    }
    return out;
}

} // namespace orc
