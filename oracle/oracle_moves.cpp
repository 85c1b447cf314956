// oracle_moves.cpp - CPU restatement of the reference's move types, the
// simulate loop and the constraint-point machinery.
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).
#include <algorithm>
#include <cmath>
#include <deque>
#include <functional>
#include <set>
#include <unordered_map>

#include "oracle_core.hpp"

namespace orc {

static std::pair<int, int> key_of(const Dom* d) { return {d->c, d->d}; }

// ------------------------------------------------------- Constraintpoints
// top_constraint_points.cpp:191-205
void Constraintpoints::reset() {
    m_regrowth_staples.clear();
    m_checked_staples.clear();
    m_d_stack.clear();
    m_segs.clear();
    m_dirs.clear();
    m_growthpoints.clear();
    m_stemdomains.clear();
    m_active.clear();
    m_initial_active.clear();
    m_inactive.clear();
    m_erased.clear();
    m_stemd_to_segs.clear();
    m_net_scaffold_ds.clear();
    m_net_dirs.clear();
}

// top_constraint_points.cpp:179-185 (operator[] inserts 0 for unknown keys)
int Constraintpoints::get_dir(Dom* d) {
    int seg = m_segs[d];
    return m_dirs[{d->c, seg}];
}

// top_constraint_points.cpp:207-221
void Constraintpoints::calculate(const std::vector<Dom*>& sds, int dir) {
    m_net_scaffold_ds = sds;
    m_dirs[{sds[0]->c, 0}] = dir;
    find_growthpoints_endpoints(sds, 0);
    m_initial_active = m_active;
}

// top_constraint_points.cpp:223-248: segment i is segment index i, empty
// segments included (their direction stays unset)
void Constraintpoints::calculate(const std::vector<std::vector<Dom*>>& segs,
                                 const std::vector<int>& dirs) {
    m_net_scaffold_ds.clear();
    for (auto& sg: segs) {
        m_net_scaffold_ds.insert(m_net_scaffold_ds.end(), sg.begin(), sg.end());
    }
    for (size_t i = 0; i < segs.size(); i++) {
        if (!segs[i].empty()) {
            m_dirs[{segs[i][0]->c, (int)i}] = dirs[i];
            find_growthpoints_endpoints(segs[i], (int)i);
        }
    }
    m_initial_active = m_active;
}

void Constraintpoints::add_active_endpoint(Dom* d, V3 pos) {
    int seg = m_segs.at(d);
    m_active[{d->c, seg}].push_back({d->d, pos});
}

void Constraintpoints::add_active_endpoint(Dom* d, V3 pos, int seg) {
    m_active[{d->c, seg}].push_back({d->d, pos});
}

// top_constraint_points.cpp:299-320
void Constraintpoints::remove_active_endpoint(Dom* d) {
    m_erased.clear();
    int seg = m_segs.at(d);
    auto& lst = m_active[{d->c, seg}];
    std::vector<std::pair<int, V3>> kept;
    for (auto& e: lst) {
        if (e.first != d->d) {
            kept.push_back(e);
        }
        else {
            m_erased.push_back(e.second);
        }
    }
    lst = kept;
}

// top_constraint_points.cpp:322-338: the first endpoint of the inactive
// endpoint domain paired with d goes
void Constraintpoints::remove_activated_endpoint(Dom* d) {
    auto it = m_inactive.find(d);
    if (it == m_inactive.end()) {
        return;
    }
    Dom* ed = it->second;
    auto& lst = m_active[{ed->c, m_segs.at(ed)}];
    for (size_t j = 0; j < lst.size(); j++) {
        if (lst[j].first == ed->d) {
            lst.erase(lst.begin() + (long)j);
            break;
        }
    }
}

// top_constraint_points.cpp:340-349
void Constraintpoints::update_endpoints(Dom* d) {
    remove_active_endpoint(d);
    auto it = m_inactive.find(d);
    if (it != m_inactive.end()) {
        add_active_endpoint(it->second, d->pos);
    }
}

// top_constraint_points.cpp:359-372
bool Constraintpoints::endpoint_reached(Dom* d, V3 pos) {
    int seg = m_segs.at(d);
    for (auto& e: m_active[{d->c, seg}]) {
        if (d->d == e.first && pos == e.second) {
            return true;
        }
    }
    return false;
}

// IdealRandomWalks::num_walks(...) == 0 exactly when the Manhattan distance
// exceeds the steps or differs from them in parity (ideal_random_walk.cpp:35-41);
// otherwise the sum (:44-67) has a positive term.
static bool walks_exist(V3 a, V3 b, int steps) {
    V3 dr = b - a;
    int s = std::abs(dr.x) + std::abs(dr.y) + std::abs(dr.z);
    return !(s > steps || (steps - s) % 2 != 0);
}

// top_constraint_points.cpp:397-442, 571-601.  A stem domain looks its
// segments up in m_stemd_to_segs, which only the non-contiguous move fills,
// so for the contiguous move a stem domain has no walk constraint.
bool Constraintpoints::walks_remain(Dom* d, V3 pos) {
    std::vector<int> segs;
    if (is_stemdomain(d)) {
        segs = m_stemd_to_segs[d];
    }
    else {
        segs = {m_segs.at(d)};
    }
    for (int seg: segs) {
        if (!walks_remain_seg(d, pos, seg)) {
            return false;
        }
    }
    return true;
}

bool Constraintpoints::walks_remain_seg(Dom* d, V3 pos, int seg) {
    std::pair<int, int> key {d->c, seg};
    int dir = m_dirs[key];
    for (auto& e: m_active[key]) {
        int steps;
        if (m_sys.cyclic && d->c == 0) {
            int n = (int)m_sys.chains[0].size();
            if (dir > 0 && e.first < d->d) {
                steps = n + e.first - d->d;
            }
            else if (dir < 0 && e.first > d->d) {
                steps = d->d + n - e.first;
            }
            else if (dir == 0 || e.first == d->d) {
                steps = 0;
            }
            else {
                steps = std::abs(e.first - d->d);
            }
        }
        else {
            steps = std::abs(e.first - d->d);
        }
        if (!walks_exist(pos, e.second, steps)) {
            return false;
        }
    }
    return true;
}

// StapleNetwork::scan_network (top_constraint_points.cpp:36-44, 69-76).  The
// per-network direction map is not cleared between scans, as in the
// reference; its entries are always +1/-1 for segments 0/1.
void Constraintpoints::scan_network(Dom* d) {
    m_net_external = false;
    m_net_cs.clear();
    m_net_gps.clear();
    m_net_ds.clear();
    m_net_iaes.clear();
    m_net_segs.clear();
    m_net_cs.insert(0);
    scan_staple_topology(d);
}

// top_constraint_points.cpp:135-161
std::vector<Dom*> Constraintpoints::make_staple_stack(Dom* d, int ci) {
    std::vector<Dom*> ds;
    std::vector<Dom*>& staple = m_sys.chain_by_index(ci);
    m_net_segs[d] = 0;
    m_net_dirs[{ci, 0}] = 1;
    for (size_t di = d->d + 1; di < staple.size(); di++) {
        ds.push_back(staple[di]);
        m_net_segs[staple[di]] = 0;
    }
    m_net_dirs[{ci, 1}] = -1;
    for (int di = d->d - 1; di >= 0; di--) {
        ds.push_back(staple[di]);
        m_net_segs[staple[di]] = 1;
    }
    return ds;
}

// top_constraint_points.cpp:78-133, 163-170 (excluded staples are always
// empty: the moveset's num_excluded_staples is hard-wired to 0,
// simulation.cpp:408-410)
void Constraintpoints::scan_staple_topology(Dom* growth_d) {
    int ci = growth_d->c;
    m_net_cs.insert(ci);
    m_net_ds.push_back(growth_d);
    std::vector<Dom*> ds = make_staple_stack(growth_d, ci);
    for (Dom* d: ds) {
        m_net_ds.push_back(d);
        Dom* bd = d->bd;
        if (bd == nullptr || bd->c == ci) {
            continue;
        }
        int bd_ci = bd->c;
        if (m_net_cs.count(bd_ci) > 0) {
            bool external = false;
            if (bd_ci == 0) {
                external = std::find(m_net_scaffold_ds.begin(), m_net_scaffold_ds.end(), bd) ==
                           m_net_scaffold_ds.end();
            }
            if (!external) {
                if (std::find(m_net_ds.begin(), m_net_ds.end(), bd) != m_net_ds.end()) {
                    m_net_iaes.push_back({bd, d});
                }
                else {
                    m_net_iaes.push_back({d, bd});
                }
            }
            if (!m_net_external && external) {
                m_net_external = true;
            }
        }
        else {
            m_net_gps.push_back({d, bd});
            scan_staple_topology(bd);
        }
    }
}

// top_constraint_points.cpp:454-494 plus the add_* helpers (:500-569)
void Constraintpoints::find_growthpoints_endpoints(const std::vector<Dom*>& sds, int seg) {
    for (Dom* d: sds) {
        m_d_stack.push_back(d);
        m_segs[d] = seg;
        if (d->bd == nullptr || d->bd->c == d->c || m_checked_staples.count(d->bd->c) > 0) {
            continue;
        }
        Dom* bd = d->bd;
        scan_network(bd);
        std::vector<std::pair<Dom*, Dom*>> gps = m_net_gps;
        gps.push_back({d, bd});
        if (m_net_external) {
            for (auto& g: gps) {
                if (g.first->c == 0) {
                    add_active_endpoint(g.first, g.first->pos, seg);
                }
            }
            for (auto& ia: m_net_iaes) {
                if (ia.second->c == 0) {
                    // the reference keys this endpoint by the staple domain
                    // (ia.first), where no scaffold growth looks for it
                    add_active_endpoint(scaffold_endpoints ? ia.second : ia.first,
                                        ia.second->pos, seg);
                }
            }
        }
        else {
            for (auto& g: gps) {
                m_growthpoints[g.first] = g.second;
                m_stemdomains[g.second] = g.first;
            }
            for (auto& ia: m_net_iaes) {
                m_inactive[ia.first] = ia.second;
            }
            for (int ci: m_net_cs) {
                if (ci != 0) {
                    m_regrowth_staples.insert(ci);
                }
            }
            m_d_stack.insert(m_d_stack.end(), m_net_ds.begin(), m_net_ds.end());
            m_segs.insert(m_net_segs.begin(), m_net_segs.end());
            for (auto& kv: m_net_dirs) {
                m_dirs.insert(kv);
            }
        }
        for (int ci: m_net_cs) {
            m_checked_staples.insert(ci);
        }
    }
}

// ----------------------------------------------------------------- Replica
Replica::Replica(const ldo_system_desc& sd, const std::vector<std::vector<int>>& idents,
                 const std::vector<std::vector<std::string>>& seqs,
                 const std::vector<ChainSpec>& cfg, double temp,
                 const std::vector<MovesetEntry>& moves, const Limits& lim,
                 int centering_freq, int centering_domain, int check_freq, int rng_kind,
                 uint64_t seed, uint32_t stream):
        sys {sd, idents, seqs, cfg, temp},
        m_moves {moves},
        m_lim {lim},
        m_centering_freq {centering_freq},
        m_centering_domain {centering_domain},
        m_check_freq {check_freq},
        m_cp {sys} {
    m_cp.scaffold_endpoints = (lim.flags & LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS) != 0;
    rng.init(rng_kind, seed, stream);
    double cum = 0;
    for (auto& m: moves) {
        cum += m.freq;
        m_cum.push_back(cum);
    }
    attempts.assign(moves.size(), 0);
    accepts.assign(moves.size(), 0);
}

// GCMCSimulation::simulate (simulation.cpp:568-653) without file output
// MCMovetype::add_tracker (movetypes.cpp:41-51) summed as the summaries of
// met_movetypes.cpp:145-190, cb_movetypes.cpp:264-290 and 626-678 read them
void Replica::add_tracker(int slot, bool accepted) {
    last_move = slot;
    last_accepted = accepted ? 1 : 0;
    int T = (int)sys.identities.size();
    int ns = (int)sys.identities[0].size();
    int a = accepted ? 1 : 0;
    if (track.empty()) {
        track.assign(6 * T + 4 * (ns + 1) + 4 * (m_lim.max_total_staples + 1) + 2, 0);
    }
    switch (m_moves[slot].kind) {
    case LDO_MOVE_MET_STAPLE_EXCHANGE: {
        int b = tracker.insertion ? 0 : 2 * T;
        track[b + tracker.staple_type] += 1;
        track[b + T + tracker.staple_type] += a;
        break;
    }
    case LDO_MOVE_CB_STAPLE_REGROWTH:
        if (!tracker.no_staples) {
            track[4 * T + tracker.staple_type] += 1;
            track[5 * T + tracker.staple_type] += a;
        }
        break;
    case LDO_MOVE_CTCB_SCAFFOLD_REGROWTH: {
        int o = 6 * T, n1 = ns + 1, s1 = m_lim.max_total_staples + 1;
        track[o + tracker.num_scaffold_domains] += 1;
        track[o + n1 + tracker.num_scaffold_domains] += a;
        track[o + 2 * n1 + tracker.num_staples] += 1;
        track[o + 2 * n1 + s1 + tracker.num_staples] += a;
        break;
    }
    case LDO_MOVE_CTRG_SCAFFOLD_REGROWTH: {
        // rg_movetypes.cpp:583-627 reads num_scaffold_domains only
        int o = 6 * T + 2 * (ns + 1) + 2 * (m_lim.max_total_staples + 1);
        track[o + tracker.num_scaffold_domains] += 1;
        track[o + ns + 1 + tracker.num_scaffold_domains] += a;
        break;
    }
    case LDO_MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH: {
        int o = 6 * T + 4 * (ns + 1) + 2 * (m_lim.max_total_staples + 1);
        track[o] += 1;
        track[o + 1] += a;
        break;
    }
    case LDO_MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH: {
        // cb_movetypes.cpp:816-870: all under 0 scaffold domains, by staples
        int s1 = m_lim.max_total_staples + 1;
        int o = 6 * T + 4 * (ns + 1) + 2 * s1 + 2;
        track[o + tracker.num_staples] += 1;
        track[o + s1 + tracker.num_staples] += a;
        break;
    }
    default:
        break;
    }
}

void Replica::simulate(int64_t steps) {
    int64_t end = step + steps;
    while (step < end) {
        step++;
        rng.begin_step(step);
        // select_movetype (simulation.cpp:655-665)
        double prob = rng.uniform_real();
        size_t i = 0;
        for (; i < m_cum.size(); i++) {
            if (prob < m_cum[i]) {
                break;
            }
        }
        if (i == m_cum.size()) {
            i = m_cum.size() - 1;
        }
        m_kind = m_moves[i].kind;
        attempts[i]++;
        m_max_recoils = m_moves[i].max_num_recoils;
        m_max_c_attempts = m_moves[i].max_c_attempts;
        m_max_seg_regrowth = m_moves[i].max_seg_regrowth;
        bool accepted = attempt(m_moves[i].kind, m_moves[i].max_regrowth);
        accepts[i] += accepted;
        add_tracker((int)i, accepted);
        if (!accepted) {
            reset_origami();
            calc_move_bias(0);
        }
        if (m_centering_freq != 0 && step % m_centering_freq == 0) {
            sys.center(m_centering_domain);
        }
        if (m_check_freq != 0 && step % m_check_freq == 0) {
            try {
                sys.check_all_constraints();
            } catch (OrigamiError& e) {
                status = LDO_E_CONSTRAINT;
                throw;
            }
        }
        // update_internal (us_simulation.cpp:262-266): f_i[point]++
        if (hist != nullptr) {
            int p = window_op(0);
            int glen = (int)win->grid.size();
            hist[(p >= 0 && p < glen) ? p : glen]++;
        }
    }
}

// MCMovetype::attempt_move plus the reset_internal chain of each move type
// (movetypes.cpp:41-51, 242-250; met_movetypes.cpp:45-48, 139-143;
// cb_movetypes.cpp:44-50, 457-461)
bool Replica::attempt(int kind, int max_regrowth) {
    m_modified.clear();
    m_assigned.clear();
    m_added.clear();
    m_prev_pos.clear();
    m_prev_ore.clear();
    m_rejected = false;
    m_modifier = 1;
    m_delta_e = 0;
    m_bias = 1;
    m_new_bias = 1;
    m_new_modifier = 1;
    m_regrow_old = false;
    m_insertion_sites = sys.num_domains;
    switch (kind) {
    case LDO_MOVE_ORIENTATION_ROTATION:
        return orientation_move();
    case LDO_MOVE_MET_STAPLE_EXCHANGE:
        return staple_exchange_move();
    case LDO_MOVE_CB_STAPLE_REGROWTH:
        return cb_staple_move();
    case LDO_MOVE_CTCB_SCAFFOLD_REGROWTH:
        m_cp.reset();
        return ctcb_scaffold_move(max_regrowth);
    case LDO_MOVE_CTRG_SCAFFOLD_REGROWTH:
        // CTRGRegrowthMCMovetype::reset_internal (rg_movetypes.cpp:66-80)
        m_cp.reset();
        m_regrow_ds.clear();
        m_c_attempts_q.clear();
        m_c_attempts_wq.clear();
        m_avail_cis_q.clear();
        m_avail_cis_wq.clear();
        m_c_opens.clear();
        m_erased_q.clear();
        m_weight = 1;
        m_weight_new = 1;
        return ctrg_scaffold_move(max_regrowth);
    case LDO_MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH:
        m_cp.reset();
        m_regrow_ds.clear();
        m_c_attempts_q.clear();
        m_c_attempts_wq.clear();
        m_avail_cis_q.clear();
        m_avail_cis_wq.clear();
        m_c_opens.clear();
        m_erased_q.clear();
        m_weight = 1;
        m_weight_new = 1;
        return ctrg_jump_move(max_regrowth);
    case LDO_MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH:
        m_cp.reset();
        return ctcb_jump_move(max_regrowth);
    }
    throw OrigamiError(LDO_E_NOTIMPL, "movetype not supported");
}

// MCMovetype::reset_origami (movetypes.cpp:53-86)
void Replica::reset_origami() {
    for (auto& k: m_assigned) {
        sys.unassign_domain(*sys.domain(k.first, k.second));
    }
    for (int c: m_added) {
        sys.delete_chain(c);
        sys.current_c_i -= 1;
    }
    for (auto& k: m_modified) {
        V3 pos = m_prev_pos[k];
        V3 ore = m_prev_ore[k];
        sys.set_checked_domain_config(*sys.domain(k.first, k.second), pos, ore);
    }
    sys.violated = false;
}

// movetypes.cpp:98-113
Dom* Replica::select_random_domain() {
    int idx = rng.uniform_int(0, sys.num_domains - 1);
    int counter = 0;
    for (auto& ch: sys.chains) {
        for (Dom* d: ch) {
            if (counter == idx) {
                return d;
            }
            counter++;
        }
    }
    return nullptr;
}

// movetypes.cpp:142-158
bool Replica::test_acceptance(long double p_ratio) {
    double p_accept = (double)(std::fmin((long double)1, p_ratio) * m_modifier);
    if (p_accept == 1) {
        return true;
    }
    return p_accept > rng.uniform_real();
}

// movetypes.cpp:160-177
bool Replica::staple_is_connector(const std::vector<Dom*>& staple) {
    for (Dom* d: staple) {
        if (d->st != Occ::unbound) {
            Dom* b = d->bd;
            if (b->c == 0) {
                continue;
            }
            std::set<int> participating {d->c};
            if (!scan_for_scaffold_domain(b, participating)) {
                return true;
            }
        }
    }
    return false;
}

// movetypes.cpp:191-232
bool Replica::scan_for_scaffold_domain(Dom* dom, std::set<int>& participating) {
    int c_i = dom->c;
    participating.insert(c_i);
    std::vector<Dom*> staple = sys.chain_by_index(c_i);
    for (Dom* cur: staple) {
        if (cur == dom) {
            continue;
        }
        Dom* b = cur->bd;
        if (b != nullptr) {
            if (b->c == dom->c) {
                continue;
            }
            if (b->c == 0) {
                return true;
            }
            if (participating.count(b->c) > 0) {
                continue;
            }
            if (scan_for_scaffold_domain(b, participating)) {
                return true;
            }
        }
    }
    return false;
}

// movetypes.cpp:258-277
std::vector<std::pair<Dom*, Dom*>> Replica::find_bound_domains(const std::vector<Dom*>& ch) {
    std::vector<std::pair<Dom*, Dom*>> out;
    int ci = ch[0]->c;
    for (Dom* d: ch) {
        if (d->bd != nullptr && d->bd->c != ci) {
            out.push_back({d, d->bd});
        }
    }
    if (out.empty()) {
        throw OrigamiError(LDO_E_CONSTRAINT, "System has unbound staple");
    }
    return out;
}

// movetypes.cpp:395-405
int Replica::num_bound_staple_domains(const std::vector<Dom*>& staple) const {
    int n = 0;
    for (Dom* d: staple) {
        if (d->st == Occ::bound || d->st == Occ::misbound) {
            n++;
        }
    }
    return n;
}

// movetypes.cpp:322-341
double Replica::set_growth_point(Dom& gnew, Dom& gold) {
    V3 o_new = -gold.ore;
    double de = sys.set_domain_config(gnew, gold.pos, o_new);
    if (sys.violated) {
        m_rejected = true;
    }
    else {
        m_assigned.push_back(key_of(&gnew));
    }
    return de;
}

// movetypes.cpp:343-370
void Replica::grow_staple(int d_i, const std::vector<Dom*>& chain) {
    std::vector<Dom*> three(chain.begin() + d_i, chain.end());
    if (three.size() > 1) {
        grow_chain(three);
    }
    if (m_rejected) {
        return;
    }
    std::vector<Dom*> five(chain.begin(), chain.begin() + d_i + 1);
    std::reverse(five.begin(), five.end());
    if (five.size() > 1) {
        grow_chain(five);
    }
}

// virtual grow_chain dispatch
void Replica::grow_chain(const std::vector<Dom*>& ds) {
    switch (m_kind) {
    case LDO_MOVE_MET_STAPLE_EXCHANGE:
        met_grow_chain(ds);
        break;
    case LDO_MOVE_CB_STAPLE_REGROWTH:
        for (size_t i = 1; i < ds.size(); i++) {
            cb_select_and_set(i, ds);
            if (m_rejected) {
                break;
            }
        }
        break;
    case LDO_MOVE_CTCB_SCAFFOLD_REGROWTH:
    case LDO_MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH:
        ctcb_grow_chain(ds);
        break;
    }
}

// --------------------------------------------------------- orientation move
// orientation_movetype.cpp:30-65
bool Replica::orientation_move() {
    bool accepted = false;
    Dom* d = select_random_domain();
    V3 o_new = kUnit[rng.uniform_int(0, 5)];
    if (d->st == Occ::bound || d->st == Occ::misbound) {
        double de = 0;
        V3 o_old = d->ore;
        Dom* b = d->bd;
        de += sys.unassign_domain(*b);
        sys.set_domain_orientation(*d, o_new);
        V3 pos = d->pos;
        de += sys.set_domain_config(*b, pos, -o_new);
        if (!sys.violated) {
            accepted = test_acceptance(std::exp(-de));
        }
        if (!accepted) {
            sys.unassign_domain(*b);
            sys.set_domain_orientation(*d, o_old);
            sys.set_checked_domain_config(*b, pos, -o_old);
        }
    }
    else {
        sys.set_domain_orientation(*d, o_new);
        accepted = true;
    }
    return accepted;
}

// ------------------------------------------------------ staple exchange move
// met_movetypes.cpp:192-206
bool Replica::staple_exchange_move() {
    tracker.insertion = rng.uniform_real() < 0.5;
    if (tracker.insertion) {
        return insert_staple();
    }
    return delete_staple();
}

// met_movetypes.cpp:50-67
void Replica::met_grow_chain(const std::vector<Dom*>& ds) {
    for (size_t i = 1; i < ds.size(); i++) {
        Dom* d = ds[i];
        Dom* prev = ds[i - 1];
        V3 p = prev->pos + kUnit[rng.uniform_int(0, 5)];
        V3 o = kUnit[rng.uniform_int(0, 5)];
        m_delta_e += sys.set_domain_config(*d, p, o);
        if (sys.violated) {
            m_rejected = true;
            break;
        }
        m_assigned.push_back(key_of(d));
    }
}

// met_movetypes.cpp:77-85
void Replica::met_unassign_domains(const std::vector<Dom*>& ds) {
    for (Dom* d: ds) {
        auto k = key_of(d);
        m_prev_pos[k] = d->pos;
        m_prev_ore[k] = d->ore;
        m_modified.push_back(k);
        m_delta_e += sys.unassign_domain(*d);
    }
}

// met_movetypes.cpp:303-359, 212-253 (exchange multipliers are all 1 and the
// move-level external bias is 0 without bias functions)
bool Replica::insert_staple() {
    int ident = rng.uniform_int(1, (int)sys.identities.size() - 1);
    tracker.staple_type = ident;
    if (sys.num_staples == m_lim.max_total_staples) {
        return false;
    }
    if ((int)sys.ident_to_index[ident].size() == m_lim.max_type_staples) {
        return false;
    }
    int c_i = sys.add_chain(ident);
    m_delta_e += sys.pot.init_e();
    m_added.push_back(c_i);
    std::vector<Dom*> chain = sys.chains.back();
    // select_new_growthpoint (movetypes.cpp:372-383)
    int gi = rng.uniform_int(0, (int)chain.size() - 1);
    Dom* gnew = chain[gi];
    Dom* gold = select_random_domain();
    while (gold->c == gnew->c) {
        gold = select_random_domain();
    }
    m_delta_e += set_growth_point(*gnew, *gold);
    if (m_rejected) {
        return false;
    }
    grow_staple(gnew->d, chain);
    if (m_rejected) {
        return false;
    }
    std::vector<Dom*>& staple = sys.chain_by_index(c_i);
    m_delta_e += calc_move_bias(0); // add_external_bias (met_movetypes.cpp:69-75, 216)
    double boltz = std::exp(-m_delta_e);
    int n_new = (int)sys.ident_to_index[ident].size();
    size_t len = sys.identities[ident].size();
    double pratio = static_cast<double>(len) / 6.0 / n_new * boltz;
    pratio *= m_insertion_sites * sys.reduced_fugacity;
    pratio /= num_bound_staple_domains(staple);
    return test_acceptance(pratio);
}

// met_movetypes.cpp:361-403, 255-301
bool Replica::delete_staple() {
    int ident = rng.uniform_int(1, (int)sys.identities.size() - 1);
    tracker.staple_type = ident;
    int n = (int)sys.ident_to_index[ident].size();
    if (n == 0) {
        m_rejected = true;
        return false;
    }
    int c_i = sys.ident_to_index[ident][rng.uniform_int(0, n - 1)];
    std::vector<Dom*> staple = sys.chain_by_index(c_i);
    if (staple_is_connector(staple)) {
        return false;
    }
    int n_bd = num_bound_staple_domains(staple);
    met_unassign_domains(staple);
    m_delta_e -= sys.pot.init_e();
    // add_external_bias with the staple count reduced by one
    // (met_movetypes.cpp:259-261)
    m_delta_e += calc_move_bias(-1);
    double boltz = std::exp(-m_delta_e);
    int ni = (int)sys.ident_to_index[ident].size();
    size_t len = sys.identities[ident].size();
    double pratio = ni * 6.0 / static_cast<double>(len) * boltz;
    pratio /= sys.reduced_fugacity * (m_insertion_sites - len);
    pratio *= n_bd;
    bool accepted = test_acceptance(pratio);
    if (accepted) {
        sys.delete_chain(c_i);
    }
    return accepted;
}

// -------------------------------------------------- CB staple regrowth move
// cb_movetypes.cpp:58-102
void Replica::cb_calc_biases(V3 p_prev, Dom& d, std::vector<std::pair<V3, V3>>& cfgs,
                             std::vector<double>& bf) {
    for (int v = 0; v < 6; v++) {
        V3 p_new = p_prev + kUnit[v];
        Occ occ = sys.occupancy(p_new);
        if (occ == Occ::bound || occ == Occ::misbound) {
            continue;
        }
        if (occ == Occ::unbound) {
            Dom* ub = sys.unbound_at(p_new);
            V3 o_new = -ub->ore;
            double de = sys.check_domain_constraints(d, p_new, o_new);
            if (!sys.violated) {
                cfgs.push_back({p_new, o_new});
                bf.push_back(std::exp(-de));
            }
            else {
                sys.violated = false;
            }
        }
        else {
            cfgs.push_back({p_new, V3 {0, 0, 0}});
            double de = sys.check_domain_constraints(d, p_new, V3 {1, 0, 0});
            bf.push_back(6 * std::exp(-de));
        }
    }
}

// CBStapleRegrowthMCMovetype::calc_bias (cb_movetypes.cpp:363-387) and
// CTCBRegrowthMCMovetype::calc_bias (:486-539)
std::vector<double> Replica::cb_calc_bias(const std::vector<double>& bf,
                                          const std::vector<std::pair<V3, V3>>& cfgs, Dom* d) {
    std::vector<double> w;
    if (m_kind == LDO_MOVE_CB_STAPLE_REGROWTH) {
        double ros = 0;
        for (double b: bf) {
            ros += b;
        }
        if (ros == 0) {
            m_rejected = true;
        }
        else {
            m_bias *= ros;
            for (double b: bf) {
                w.push_back(b / ros);
            }
        }
        return w;
    }
    std::vector<long double> wl(bf.begin(), bf.end());
    for (size_t i = 0; i < cfgs.size(); i++) {
        V3 pos = cfgs[i].first;
        if (sys.occupancy(pos) == Occ::unbound) {
            Dom* occ = sys.unbound_at(pos);
            bool same_chain = occ->c == d->c;
            bool endpoint = m_cp.endpoint_reached(d, pos);
            if (!(same_chain || endpoint)) {
                wl[i] = 0;
            }
        }
        if (!m_cp.walks_remain(d, pos)) {
            wl[i] = 0;
        }
    }
    long double sum = 0;
    for (long double x: wl) {
        sum += x;
    }
    if (sum == 0) {
        m_rejected = true;
    }
    else {
        m_bias *= sum;
        for (size_t i = 0; i < wl.size(); i++) {
            w.push_back(static_cast<double>(wl[i] / sum));
        }
    }
    return w;
}

// cb_movetypes.cpp:104-166
void Replica::cb_select_and_set(size_t i, const std::vector<Dom*>& ds) {
    Dom* d = ds[i];
    V3 p_prev = ds[i - 1]->pos;
    std::vector<std::pair<V3, V3>> cfgs;
    std::vector<double> bf;
    cb_calc_biases(p_prev, *d, cfgs, bf);
    std::vector<double> w = cb_calc_bias(bf, cfgs, d);
    if (m_rejected) {
        return;
    }
    if (!m_regrow_old) {
        std::pair<V3, V3> cfg {};
        double cum = 0;
        double u = rng.uniform_real();
        for (size_t k = 0; k < w.size(); k++) {
            cum += w[k];
            if (u < cum) {
                cfg = cfgs[k];
                break;
            }
        }
        V3 p_new = cfg.first;
        V3 o_new = cfg.second;
        if (o_new == V3 {0, 0, 0}) {
            o_new = kUnit[rng.uniform_int(0, 5)];
        }
        sys.set_checked_domain_config(*d, p_new, o_new);
    }
    else {
        auto k = key_of(d);
        sys.set_checked_domain_config(*d, m_old_pos[k], m_old_ore[k]);
    }
    m_assigned.push_back(key_of(d));
}

// cb_movetypes.cpp:200-213
void Replica::cb_unassign_domains(const std::vector<Dom*>& ds) {
    for (Dom* d: ds) {
        auto k = key_of(d);
        m_prev_pos[k] = d->pos;
        m_prev_ore[k] = d->ore;
        m_modified.push_back(k);
        sys.unassign_domain(*d);
    }
}

// cb_movetypes.cpp:215-225
void Replica::cb_setup_regrow_old() {
    m_regrow_old = true;
    m_new_bias = m_bias;
    m_new_modifier = m_modifier;
    m_bias = 1;
    m_modified.clear();
    m_assigned.clear();
    m_old_pos = m_prev_pos;
    m_old_ore = m_prev_ore;
}

// cb_movetypes.cpp:168-180
double Replica::cb_set_old_growth_point(Dom& gnew, Dom& gold) {
    auto k = key_of(&gnew);
    V3 o_old = m_old_ore[k];
    double de = sys.set_checked_domain_config(gnew, gold.pos, o_old);
    m_bias *= std::exp(-de);
    m_assigned.push_back(k);
    return de;
}

// cb_movetypes.cpp:182-198 (ops/bias resynchronisation is a no-op without
// bias functions)
bool Replica::cb_test_acceptance() {
    long double ratio = m_new_bias / m_bias;
    if (test_acceptance(ratio)) {
        reset_origami();
        calc_move_bias(0);
        return true;
    }
    m_modified.clear();
    m_assigned.clear();
    return false;
}

// cb_movetypes.cpp:292-348, 389-404
bool Replica::cb_staple_move() {
    if (sys.num_staples == 0) {
        tracker.no_staples = true;
        return false;
    }
    tracker.no_staples = false;
    int w = rng.uniform_int(1, sys.num_staples);
    std::vector<Dom*> chain = sys.chains[w];
    tracker.staple_type = chain[0]->c_ident;
    if (staple_is_connector(chain)) {
        return false;
    }
    auto bound = find_bound_domains(chain);
    m_bias *= bound.size();
    auto gp = bound[rng.uniform_int(0, (int)bound.size() - 1)];
    cb_unassign_domains(chain);
    {
        double de = set_growth_point(*gp.first, *gp.second);
        m_bias *= std::exp(-de);
        if (!m_rejected) {
            grow_staple(gp.first->d, chain);
        }
    }
    if (m_rejected) {
        return false;
    }
    m_bias /= num_bound_staple_domains(chain);
    cb_add_external_bias();
    cb_setup_regrow_old();
    gp = bound[rng.uniform_int(0, (int)bound.size() - 1)];
    cb_unassign_domains(chain);
    cb_set_old_growth_point(*gp.first, *gp.second);
    if (!m_rejected) {
        grow_staple(gp.first->d, chain);
    }
    m_modifier = m_new_modifier;
    return cb_test_acceptance();
}

// ------------------------------------------ CTCB scaffold regrowth move
// CTRegrowthMCMovetype::select_indices (movetypes.cpp:469-510)
std::vector<Dom*> Replica::select_indices(const std::vector<Dom*>& seg, unsigned min_len,
                                          unsigned max_regrowth) {
    unsigned seg_len = (unsigned)seg.size();
    unsigned max_len = std::min(seg_len, max_regrowth);
    unsigned sel_len = (unsigned)rng.uniform_int((int)min_len, (int)max_len);
    int start = rng.uniform_int(0, (int)seg_len - 1);
    m_dir = rng.uniform_int(0, 1);
    if (m_dir == 0) {
        m_dir = -1;
    }
    Dom* cur = seg[start];
    std::vector<Dom*> ds;
    while (cur != nullptr && ds.size() != sel_len) {
        ds.push_back(cur);
        cur = walk(cur, m_dir);
    }
    Dom* back = walk(seg[start], -m_dir);
    while (back != nullptr && ds.size() != sel_len) {
        ds.insert(ds.begin(), back);
        back = walk(back, -m_dir);
    }
    if (ds.size() < min_len) {
        return select_indices(seg, min_len, max_regrowth);
    }
    if (cur != nullptr) {
        m_cp.add_active_endpoint(cur, cur->pos, 0);
    }
    return ds;
}

// cb_movetypes.cpp:463-484
void Replica::ctcb_grow_chain(const std::vector<Dom*>& ds) {
    if (ds.size() <= 1) {
        return;
    }
    for (size_t i = 1; i < ds.size(); i++) {
        Dom* d = ds[i];
        m_dir = m_cp.get_dir(d);
        cb_select_and_set(i, ds);
        if (m_rejected) {
            break;
        }
        m_cp.update_endpoints(d);
        if (m_cp.is_growthpoint(d)) {
            ctcb_grow_staple_and_update_endpoints(d);
            if (m_rejected) {
                break;
            }
        }
    }
}

// cb_movetypes.cpp:541-562
void Replica::ctcb_grow_staple_and_update_endpoints(Dom* gold) {
    Dom* gnew = m_cp.domain_to_grow(gold);
    int c_i = gnew->c;
    if (c_i == 0) {
        return;
    }
    if (m_regrow_old) {
        cb_set_old_growth_point(*gnew, *gold);
    }
    else {
        double de = set_growth_point(*gnew, *gold);
        m_bias *= std::exp(-de);
    }
    if (!m_rejected) {
        m_cp.update_endpoints(gnew);
        std::vector<Dom*> staple = sys.chain_by_index(c_i);
        grow_staple(gnew->d, staple);
    }
}

// cb_movetypes.cpp:680-747
bool Replica::ctcb_scaffold_move(int max_regrowth) {
    m_regrow_old = false;
    std::vector<Dom*> scaffold = sys.chains[0];
    std::vector<Dom*> sds = select_indices(scaffold, 2, (unsigned)max_regrowth);
    tracker.num_scaffold_domains = (int)sds.size();
    m_cp.calculate(sds, m_dir);
    bool whole_ring = sys.cyclic && sds.size() == scaffold.size();
    if (!whole_ring) {
        m_cp.remove_active_endpoint(sds[0]);
    }
    std::set<int> staples = m_cp.staples_to_be_regrown();
    tracker.num_staples = (int)staples.size();
    std::vector<Dom*> to_unassign(sds.begin() + 1, sds.end());
    cb_unassign_domains(to_unassign);
    for (int c_i: staples) {
        cb_unassign_domains(sys.chain_by_index(c_i));
    }
    if (m_cp.is_growthpoint(sds[0])) {
        ctcb_grow_staple_and_update_endpoints(sds[0]);
        if (m_rejected) {
            return false;
        }
    }
    ctcb_grow_chain(sds);
    if (m_rejected) {
        return false;
    }
    cb_add_external_bias();
    cb_setup_regrow_old();
    m_cp.reset_active_endpoints();
    if (!whole_ring) {
        m_cp.remove_active_endpoint(sds[0]);
    }
    cb_unassign_domains(to_unassign);
    for (int c_i: staples) {
        cb_unassign_domains(sys.chain_by_index(c_i));
    }
    if (m_cp.is_growthpoint(sds[0])) {
        ctcb_grow_staple_and_update_endpoints(sds[0]);
    }
    ctcb_grow_chain(sds);
    m_modifier = 1;
    return cb_test_acceptance();
}

// ------------------------------------------ CTRG scaffold regrowth move
// CTRGRegrowthMCMovetype (rg_movetypes.cpp) with
// CTRGScaffoldRegrowthMCMovetype::internal_attempt_move (:629-689).  Excluded
// staples are always empty (simulation.cpp:450-452), so sel_excluded_staples
// draws nothing and excluded_staples_bound is false.

// m_all_configs = all_pairs(utility::vectors) (:58, utility.hpp:168-177):
// index 6 * (site vector) + (orientation vector)
static std::pair<V3, V3> rg_config(int i) { return {kUnit[i / 6], kUnit[i % 6]}; }

static int rg_vector_index(V3 v) {
    for (int k = 0; k < 6; k++) {
        if (kUnit[k] == v) {
            return k;
        }
    }
    throw OrigamiError(LDO_E_CONSTRAINT, "CTRG: domain not adjacent to its reference");
}

// :87-104
double Replica::rg_unassign_and_save() {
    double de = 0;
    Dom* d = m_regrow_ds[0];
    m_prev_pos[key_of(d)] = d->pos;
    m_prev_ore[key_of(d)] = d->ore;
    for (size_t i = 1; i < m_regrow_ds.size(); i++) {
        d = m_regrow_ds[i];
        m_prev_pos[key_of(d)] = d->pos;
        m_prev_ore[key_of(d)] = d->ore;
        m_modified.push_back(key_of(d));
        de += sys.unassign_domain(*d);
    }
    m_erased_q.clear();
    return de;
}

// :122-129
void Replica::rg_unassign_domains() {
    for (size_t i = 1; i < m_regrow_ds.size(); i++) {
        sys.unassign_domain(*m_regrow_ds[i]);
    }
    m_erased_q.clear();
}

// :177-231
double Replica::rg_recoil_regrow() {
    double de = 0;
    m_di = 0;
    m_d = m_regrow_ds[0];
    m_dir = m_cp.get_dir(m_d);
    size_t n = m_regrow_ds.size();
    m_c_attempts_q.resize(n);
    m_avail_cis_q.resize(n);
    m_c_opens.resize(n);
    m_c_opens[0] = 1;
    rg_prepare_for_growth();
    int recoils = 0;
    while (true) {
        Cfg c;
        double p = 0;
        bool open = false;
        while (!open && m_c_attempts != m_d_max_c_attempts) {
            m_c_attempts++;
            c = rg_select_trial_config();
            p = rg_calc_p_config_open(c);
            open = rg_test_config_open(p);
        }
        if (open) {
            if (recoils != 0) {
                recoils--;
            }
            de += rg_set_config(m_d, c);
            m_c_attempts_q[m_di] = m_c_attempts;
            m_avail_cis_q[m_di] = m_avail_cis;
            m_c_opens[m_di] = p;
            if (m_di == n - 1) {
                break;
            }
            rg_prepare_for_growth();
        }
        else {
            if (recoils == m_max_recoils || m_di == 1) {
                m_rejected = true;
                break;
            }
            recoils++;
            de += rg_prepare_for_regrowth();
        }
    }
    return de;
}

// :233-241
double Replica::rg_set_config(Dom* d, Cfg c) {
    double de = sys.set_checked_domain_config(*d, c.first, c.second);
    m_assigned.push_back(key_of(d));
    m_cp.update_endpoints(m_d);
    m_erased_q.push_back(m_cp.get_erased_endpoints());
    return de;
}

// :243-260
void Replica::rg_prepare_for_growth() {
    m_di++;
    m_d = m_regrow_ds[m_di];
    m_stemd = m_cp.is_stemdomain(m_d);
    m_dir = m_cp.get_dir(m_d);
    m_c_attempts = 0;
    if (m_stemd) {
        m_d_max_c_attempts = 1;
        m_avail_cis.clear();
        m_ref_d = m_cp.get_growthpoint(m_d);
    }
    else {
        m_d_max_c_attempts = m_max_c_attempts;
        m_avail_cis.resize(36);
        for (int i = 0; i < 36; i++) {
            m_avail_cis[i] = i;
        }
        m_ref_d = walk(m_d, -m_dir);
    }
}

// :262-286
double Replica::rg_prepare_for_regrowth() {
    m_di--;
    m_d = m_regrow_ds[m_di];
    m_dir = m_cp.get_dir(m_d);
    double de = sys.unassign_domain(*m_d);
    m_assigned.pop_back();
    rg_restore_endpoints();
    m_stemd = m_cp.is_stemdomain(m_d);
    if (m_stemd) {
        m_d_max_c_attempts = 1;
        m_c_attempts = 1;
        m_avail_cis.clear();
        m_ref_d = m_cp.get_growthpoint(m_d);
    }
    else {
        m_d_max_c_attempts = m_max_c_attempts;
        m_c_attempts = m_c_attempts_q[m_di];
        m_avail_cis = m_avail_cis_q[m_di];
        m_ref_d = walk(m_d, -m_dir);
    }
    return de;
}

// :288-296
void Replica::rg_restore_endpoints() {
    m_cp.remove_activated_endpoint(m_d);
    std::vector<V3> erased = m_erased_q.back();
    m_erased_q.pop_back();
    for (V3 p: erased) {
        m_cp.add_active_endpoint(m_d, p);
    }
}

// :298-313
Replica::Cfg Replica::rg_select_trial_config() {
    if (m_stemd) {
        return {m_ref_d->pos, -m_ref_d->ore};
    }
    int ci = rng.uniform_int(0, (int)m_avail_cis.size() - 1);
    int i = m_avail_cis[ci];
    m_avail_cis.erase(m_avail_cis.begin() + ci);
    Cfg c = rg_config(i);
    c.first = c.first + m_ref_d->pos;
    return c;
}

// :315-343
double Replica::rg_calc_p_config_open(Cfg c) {
    double de = sys.check_domain_constraints(*m_d, c.first, c.second);
    if (sys.violated) {
        sys.violated = false;
        return 0;
    }
    if (!m_cp.walks_remain(m_d, c.first)) {
        return 0;
    }
    if (sys.occupancy(c.first) == Occ::unbound) {
        Dom* occ = sys.unbound_at(c.first);
        bool same_chain = occ->c == m_d->c;
        bool endpoint = m_cp.endpoint_reached(m_d, c.first);
        if (!(same_chain || endpoint || m_stemd)) {
            return 0;
        }
    }
    return std::fmin(1, std::exp(-de));
}

// :345-361
bool Replica::rg_test_config_open(double p) {
    p = std::fmin(1, p);
    if (p == 1) {
        return true;
    }
    return p > rng.uniform_real();
}

// :363-421
void Replica::rg_calc_weights() {
    m_di = 0;
    m_d = m_regrow_ds[0];
    while (m_di != m_regrow_ds.size() - 1) {
        m_di++;
        m_d = m_regrow_ds[m_di];
        m_stemd = m_cp.is_stemdomain(m_d);
        m_dir = m_cp.get_dir(m_d);
        int avail_cs = 1;
        if (!m_stemd) {
            m_ref_d = walk(m_d, -m_dir);
            int c_attempts = m_c_attempts_wq[m_di];
            m_avail_cis = m_avail_cis_wq[m_di];
            while (c_attempts != m_max_c_attempts) {
                c_attempts++;
                Cfg c = rg_select_trial_config();
                double p = rg_calc_p_config_open(c);
                if (rg_test_config_open(p)) {
                    sys.set_checked_domain_config(*m_d, c.first, c.second);
                    m_cp.update_endpoints(m_d);
                    m_erased_q.push_back(m_cp.get_erased_endpoints());
                    int dir = m_dir;
                    Dom* ref_d = m_ref_d;
                    bool stemd = m_stemd;
                    std::vector<int> avail = m_avail_cis;
                    avail_cs += rg_test_config_avail();
                    m_avail_cis = avail;
                    m_stemd = stemd;
                    m_ref_d = ref_d;
                    m_dir = dir;
                    sys.unassign_domain(*m_d);
                    rg_restore_endpoints();
                }
            }
        }
        m_weight *= avail_cs / m_c_opens[m_di - 1];
        auto key = key_of(m_d);
        sys.set_checked_domain_config(*m_d, m_prev_pos[key], m_prev_ore[key]);
        m_cp.update_endpoints(m_d);
        m_erased_q.push_back(m_cp.get_erased_endpoints());
    }
    m_weight /= m_c_opens[m_di];
}

// :426-480
bool Replica::rg_test_config_avail() {
    int feels = 0;
    if (feels == m_max_recoils || m_di == m_regrow_ds.size() - 1) {
        return true;
    }
    rg_prepare_for_growth();
    bool avail;
    while (true) {
        Cfg c;
        bool open = false;
        while (!open && m_c_attempts != m_d_max_c_attempts) {
            m_c_attempts++;
            c = rg_select_trial_config();
            double p = rg_calc_p_config_open(c);
            open = rg_test_config_open(p);
        }
        if (open) {
            feels++;
            if (feels == m_max_recoils || m_di == m_regrow_ds.size() - 1) {
                feels--;
                avail = true;
                break;
            }
            rg_set_config(m_d, c);
            m_c_attempts_q[m_di] = m_c_attempts;
            m_avail_cis_q[m_di] = m_avail_cis;
            rg_prepare_for_growth();
        }
        else {
            if (feels == 0) {
                avail = false;
                break;
            }
            feels--;
            rg_prepare_for_regrowth();
        }
    }
    while (feels != 0) {
        m_di--;
        feels--;
        m_d = m_regrow_ds[m_di];
        sys.unassign_domain(*m_d);
        rg_restore_endpoints();
    }
    m_di--;
    m_d = m_regrow_ds[m_di];
    return avail;
}

// :482-512
void Replica::rg_calc_old_c_opens() {
    m_di = 0;
    m_c_opens[0] = 1;
    while (m_di != m_regrow_ds.size() - 1) {
        m_di++;
        m_d = m_regrow_ds[m_di];
        m_c_attempts_q[m_di] = 1;
        auto key = key_of(m_d);
        V3 p = m_old_pos[key];
        Cfg c {p, m_old_ore[key]};
        m_stemd = m_cp.is_stemdomain(m_d);
        m_c_opens[m_di] = rg_calc_p_config_open(c);
        rg_set_config(m_d, c);
        if (m_stemd) {
            m_avail_cis_q[m_di].clear();
        }
        else {
            std::vector<int> all(36);
            for (int i = 0; i < 36; i++) {
                all[i] = i;
            }
            m_dir = m_cp.get_dir(m_d);
            m_ref_d = walk(m_d, -m_dir);
            V3 pref = m_old_pos[key_of(m_ref_d)];
            int ci = 6 * rg_vector_index(p - pref) + rg_vector_index(c.second);
            all.erase(std::remove(all.begin(), all.end(), ci), all.end());
            m_avail_cis_q[m_di] = all;
        }
    }
}

// :514-533
bool Replica::rg_test_acceptance() {
    long double ratio = m_weight_new / m_weight * std::exp(-m_delta_e);
    if (test_acceptance(ratio)) {
        m_prev_pos = m_new_pos;
        m_prev_ore = m_new_ore;
        reset_origami();
        return true;
    }
    m_modified.clear();
    m_assigned.clear();
    return false;
}

// :629-689 with setup_constraints (:131-142), setup_for_calc_new_weights
// (:144-150) and setup_for_calc_old_weights (:152-160)
bool Replica::ctrg_scaffold_move(int max_regrowth) {
    std::vector<Dom*> scaffold = sys.chains[0];
    std::vector<Dom*> sds = select_indices(scaffold, 2, (unsigned)max_regrowth);
    tracker.num_scaffold_domains = (int)sds.size();
    m_cp.calculate(sds, m_dir);
    if (!(sys.cyclic && sds.size() == scaffold.size())) {
        m_cp.remove_active_endpoint(sds[0]);
    }
    m_regrow_ds = m_cp.domains_to_be_regrown();
    bool whole = sys.cyclic && m_regrow_ds.size() == scaffold.size();
    m_delta_e += rg_unassign_and_save();
    m_delta_e += rg_recoil_regrow();
    if (m_rejected) {
        return false;
    }
    m_delta_e += calc_move_bias(0); // add_external_bias (:82-85)

    // weights of the new configuration
    m_c_attempts_wq = m_c_attempts_q;
    m_avail_cis_wq = m_avail_cis_q;
    m_old_pos = m_prev_pos;
    m_old_ore = m_prev_ore;
    m_modified.clear();
    rg_unassign_and_save();
    m_cp.reset_active_endpoints();
    if (!whole) {
        m_cp.remove_active_endpoint(m_regrow_ds[0]);
    }
    rg_calc_weights();

    // weights of the old configuration; the reference's test before
    // calc_old_c_opens (:665-668) differs from the other three
    rg_unassign_domains();
    m_cp.reset_active_endpoints();
    if (!sys.cyclic && m_regrow_ds.size() == scaffold.size()) {
        m_cp.remove_active_endpoint(m_regrow_ds[0]);
    }
    rg_calc_old_c_opens();
    m_weight_new = m_weight;
    m_weight = 1;
    m_c_attempts_wq = m_c_attempts_q;
    m_avail_cis_wq = m_avail_cis_q;
    m_new_pos = m_prev_pos;
    m_new_ore = m_prev_ore;
    m_modified.clear();
    rg_unassign_and_save();
    m_cp.reset_active_endpoints();
    if (!whole) {
        m_cp.remove_active_endpoint(m_regrow_ds[0]);
    }
    rg_calc_weights();
    return rg_test_acceptance();
}

// ------------------------------------ non-contiguous scaffold selection
// CTRegrowthMCMovetype::check_for_stemds (movetypes.cpp:704-720): the
// scaffold domains bound to the chain neighbours of the staple domain bound
// to cur_d
void Replica::check_for_stemds(Dom* cur_d, std::deque<Dom*>& possible_stems) {
    if (cur_d->st != Occ::bound) {
        return;
    }
    Dom* bound_d = cur_d->bd;
    for (int staple_dir: {-1, 1}) {
        Dom* nd = walk(bound_d, staple_dir);
        if (nd != nullptr && nd->st == Occ::bound && nd->bd->c == 0) {
            possible_stems.push_back(nd->bd);
        }
    }
}

// CTRegrowthMCMovetype::fill_seg (movetypes.cpp:666-702)
bool Replica::fill_seg(Dom* start_d, size_t max_length, size_t seg_max_length, int dir,
                       std::set<Dom*>& domains, std::deque<Dom*>& possible_stems,
                       std::vector<Dom*>& seg) {
    Dom* cur_d = start_d;
    check_for_stemds(cur_d, possible_stems);
    Dom* next_d = start_d;
    while (seg.size() != seg_max_length && next_d != nullptr) {
        next_d = walk(cur_d, dir);
        if (next_d == nullptr) {
            break;
        }
        Dom* next_next_d = walk(next_d, dir);
        if (next_next_d != nullptr && domains.count(next_next_d) != 0) {
            break;
        }
        seg.push_back(next_d);
        domains.insert(next_d);
        if (domains.size() == max_length) {
            return true;
        }
        cur_d = next_d;
        check_for_stemds(cur_d, possible_stems);
    }
    return false;
}

// CTRegrowthMCMovetype::select_noncontig_segs (movetypes.cpp:512-648)
void Replica::select_noncontig_segs(std::vector<std::vector<Dom*>>& segs,
                                    std::vector<std::vector<std::vector<Dom*>>>& paired_segs,
                                    std::vector<Dom*>& seg_stems, std::vector<int>& dirs,
                                    int max_regrowth) {
    std::vector<Dom*>& given = sys.chains[0];
    std::set<Dom*> domains;
    size_t max_length = (size_t)rng.uniform_int(2, max_regrowth);
    size_t seg_max = (size_t)rng.uniform_int(2, m_max_seg_regrowth + 1);
    int start_i = rng.uniform_int(0, (int)given.size() - 1);
    Dom* seg_start_d = given[start_i];
    int dir = rng.uniform_int(0, 1);
    if (dir == 0) {
        dir = -1;
    }
    if (walk(seg_start_d, dir) == nullptr) {
        dir *= -1;
    }
    std::vector<Dom*> cur_seg {seg_start_d};
    domains.insert(seg_start_d);
    std::deque<Dom*> possible_stems;
    dirs.push_back(dir);
    bool reached = fill_seg(seg_start_d, max_length, seg_max, dir, domains, possible_stems,
                            cur_seg);
    segs.push_back(cur_seg);
    while (!reached && !possible_stems.empty()) {
        cur_seg.clear();
        Dom* stemd = possible_stems.front();
        possible_stems.pop_front();
        if (domains.count(stemd) != 0) {
            continue;
        }
        bool adjacent = false;
        for (int test_dir: {-1, 1}) {
            Dom* nd = walk(stemd, test_dir);
            if (nd != nullptr && domains.count(nd) != 0) {
                adjacent = true;
                break;
            }
        }
        if (adjacent) {
            continue;
        }
        cur_seg.push_back(stemd);
        domains.insert(stemd);
        seg_stems.push_back(stemd);
        if (domains.size() == max_length) {
            reached = true;
            segs.push_back(cur_seg);
            segs.push_back({});
            paired_segs.push_back({{}, {}});
            dirs.push_back(1);
            dirs.push_back(-1);
            break;
        }
        int dir1 = rng.uniform_int(0, 1);
        if (dir1 == 0) {
            dir1 = -1;
        }
        int dir_pair[2] = {dir1, -dir1};
        dirs.push_back(dir_pair[0]);
        dirs.push_back(dir_pair[1]);
        std::vector<std::vector<Dom*>> seg_pair {{}, {}};
        segs.push_back({});
        segs.push_back({});
        for (size_t i: {0, 1}) {
            dir = dir_pair[i];
            seg_max = (size_t)rng.uniform_int(0, m_max_seg_regrowth);
            if (i == 0) {
                max_length++;
            }
            std::vector<Dom*> seg;
            reached = fill_seg(stemd, max_length, seg_max, dir, domains, possible_stems, seg);
            seg_pair[i] = seg;
            if (!cur_seg.empty()) {
                seg.insert(seg.begin(), cur_seg[0]);
            }
            segs[segs.size() - 2 + i] = seg;
            if (reached) {
                break;
            }
            cur_seg.clear();
        }
        paired_segs.push_back(seg_pair);
    }

    Dom* last_d = segs[0].back();
    Dom* next_d = walk(last_d, dirs[0]);
    if (next_d != nullptr) {
        m_cp.add_active_endpoint(next_d, next_d->pos, 0);
    }
    int seg_i = 1;
    for (size_t i = 0; i < paired_segs.size(); i++) {
        Dom* stem_d = seg_stems[i];
        m_cp.add_growthpoint(stem_d->bd, stem_d);
        m_cp.add_stem_seg_pair(stem_d, {seg_i, seg_i + 1});
        for (auto& sg: paired_segs[i]) {
            int d = dirs[seg_i];
            last_d = sg.empty() ? stem_d : sg.back();
            Dom* nd = walk(last_d, d);
            if (nd != nullptr) {
                m_cp.add_active_endpoint(nd, nd->pos, seg_i);
            }
            seg_i++;
        }
    }
}

// CTRGJumpScaffoldRegrowthMCMovetype::internal_attempt_move
// (rg_movetypes.cpp:791-853)
bool Replica::ctrg_jump_move(int max_regrowth) {
    std::vector<std::vector<Dom*>> segs;
    std::vector<std::vector<std::vector<Dom*>>> paired_segs;
    std::vector<Dom*> seg_stems;
    std::vector<int> dirs;
    select_noncontig_segs(segs, paired_segs, seg_stems, dirs, max_regrowth);
    tracker.num_scaffold_domains = 0;
    m_cp.calculate(segs, dirs);
    m_cp.remove_active_endpoint(segs[0][0]);
    for (Dom* stem_d: seg_stems) {
        m_cp.add_growthpoint(stem_d->bd, stem_d);
    }
    m_regrow_ds = m_cp.domains_to_be_regrown();
    m_delta_e += rg_unassign_and_save();
    m_delta_e += rg_recoil_regrow();
    if (m_rejected) {
        return false;
    }
    m_delta_e += calc_move_bias(0); // add_external_bias (:82-85)

    m_c_attempts_wq = m_c_attempts_q;
    m_avail_cis_wq = m_avail_cis_q;
    m_old_pos = m_prev_pos;
    m_old_ore = m_prev_ore;
    m_modified.clear();
    rg_unassign_and_save();
    m_cp.reset_active_endpoints();
    m_cp.remove_active_endpoint(segs[0][0]);
    rg_calc_weights();

    rg_unassign_domains();
    m_cp.reset_active_endpoints();
    m_cp.remove_active_endpoint(segs[0][0]);
    rg_calc_old_c_opens();
    m_weight_new = m_weight;
    m_weight = 1;
    m_c_attempts_wq = m_c_attempts_q;
    m_avail_cis_wq = m_avail_cis_q;
    m_new_pos = m_prev_pos;
    m_new_ore = m_prev_ore;
    m_modified.clear();
    rg_unassign_and_save();
    m_cp.reset_active_endpoints();
    m_cp.remove_active_endpoint(segs[0][0]);
    rg_calc_weights();
    return rg_test_acceptance();
}

// CTCBJumpScaffoldRegrowthMCMovetype::set_first_seg_domain
// (cb_movetypes.cpp:973-983): a stem binds its growthpoint as in growth
// from a staple, in the new and the old growth alike
void Replica::ctcb_set_first_seg_domain(Dom* d_new) {
    Dom* d_old = m_cp.get_growthpoint(d_new);
    if (!m_cp.walks_remain(d_new, d_old->pos)) {
        m_rejected = true;
        return;
    }
    m_dir = m_cp.get_dir(d_new);
    double de = set_growth_point(*d_new, *d_old);
    m_bias *= std::exp(-de);
}

// the growth of both passes of CTCBJumpScaffoldRegrowthMCMovetype
// (cb_movetypes.cpp:896-921, 935-960)
void Replica::ctcb_jump_grow(const std::vector<std::vector<Dom*>>& segs,
                             const std::vector<std::vector<std::vector<Dom*>>>& paired_segs,
                             const std::vector<Dom*>& seg_stems) {
    if (m_cp.is_growthpoint(segs[0][0])) {
        ctcb_grow_staple_and_update_endpoints(segs[0][0]);
        if (m_rejected) {
            return;
        }
    }
    ctcb_grow_chain(segs[0]);
    if (m_rejected) {
        return;
    }
    for (size_t i = 0; i < seg_stems.size(); i++) {
        ctcb_set_first_seg_domain(seg_stems[i]);
        if (m_rejected) {
            return;
        }
        for (auto sg: paired_segs[i]) {
            sg.insert(sg.begin(), seg_stems[i]);
            ctcb_grow_chain(sg);
            if (m_rejected) {
                return;
            }
        }
    }
}

// CTCBJumpScaffoldRegrowthMCMovetype::internal_attempt_move
// (cb_movetypes.cpp:872-967)
bool Replica::ctcb_jump_move(int max_regrowth) {
    m_regrow_old = false;
    std::vector<std::vector<Dom*>> segs;
    std::vector<std::vector<std::vector<Dom*>>> paired_segs;
    std::vector<Dom*> seg_stems;
    std::vector<int> dirs;
    select_noncontig_segs(segs, paired_segs, seg_stems, dirs, max_regrowth);
    tracker.num_scaffold_domains = 0;
    m_cp.calculate(segs, dirs);
    m_cp.remove_active_endpoint(segs[0][0]);
    tracker.num_staples = (int)m_cp.staples_to_be_regrown().size();
    std::vector<Dom*> regrow = m_cp.domains_to_be_regrown();
    std::vector<Dom*> to_unassign(regrow.begin() + 1, regrow.end());
    cb_unassign_domains(to_unassign);
    ctcb_jump_grow(segs, paired_segs, seg_stems);
    if (m_rejected) {
        return false;
    }
    cb_add_external_bias();
    cb_setup_regrow_old();
    m_cp.reset_active_endpoints();
    m_cp.remove_active_endpoint(segs[0][0]);
    cb_unassign_domains(to_unassign);
    ctcb_jump_grow(segs, paired_segs, seg_stems);
    if (m_rejected) {
        return false;
    }
    m_modifier = 1;
    return cb_test_acceptance();
}

// ------------------------------------------------------ move-level biases
double WindowBias::restraint(int p) const {
    if (p < min_op) {
        return slope * (min_op - p - 1) + min_bias;
    }
    if (p > max_op) {
        return slope * (p - max_op - 1) + min_bias;
    }
    return well_bias;
}

double WindowBias::grid_bias(int p) const {
    if (p >= 0 && p < (int)grid.size()) {
        return grid[p];
    }
    return off_grid;
}

int Replica::window_op(int staple_offset) const {
    if (win->op == LDO_OP_NUMSTAPLES) {
        return sys.num_staples + staple_offset;
    }
    return sys.num_fully_bound_pairs;
}

// SystemOrderParams::update_move_params + SystemBiases::calc_move
// (order_params.cpp:595-603; bias_functions.cpp:484-496): refresh the
// window's bias functions at the current order parameter and return the
// change since their stored values, times the bias multiplier.
double Replica::calc_move_bias(int staple_offset) {
    if (win == nullptr) {
        return 0;
    }
    int p = window_op(staple_offset);
    double r = win->restraint(p);
    double g = win->grid_bias(p);
    double diff = r - win_state->restraint;
    diff += g - win_state->grid;
    win_state->restraint = r;
    win_state->grid = g;
    return diff * win->mult;
}

// CBMCMovetype::add_external_bias (cb_movetypes.cpp:52-56)
void Replica::cb_add_external_bias() {
    m_bias *= std::exp(-calc_move_bias(0));
}


// ------------------------------------------------------ learned proposal
// Learned domain regrowth, restated from the sampler's definition (ldo.h,
// ldo_mlp.h); the network arithmetic follows the sampler's order (layer 1 as
// sequential float sums over the sites, then fused multiply-adds over the
// embedding, dense layers as fused multiply-adds in input order) so that
// every decision matches.
static int unit_of(V3 v) {
    for (int k = 0; k < 6; k++) {
        if (kUnit[k] == v) {
            return k;
        }
    }
    return -1;
}

int Replica::learned_context(Dom* d, Dom* a, uint8_t ch[343], uint8_t legal[6]) const {
    std::unordered_map<V3, std::vector<Dom*>, V3Hash> at;
    for (auto& c: sys.chains) {
        for (Dom* x: c) {
            if (x->st != Occ::unassigned && x != d) {
                at[x->pos].push_back(x);
            }
        }
    }
    V3 A = a->pos;
    for (int s = 0; s < 343; s++) {
        V3 p {A.x + s % 7 - 3, A.y + (s / 7) % 7 - 3, A.z + s / 49 - 3};
        auto it = at.find(p);
        int scaf = 6, stap = 6;
        if (it != at.end()) {
            for (Dom* x: it->second) {
                int o = unit_of(x->ore);
                if (x->c == 0) {
                    scaf = std::min(scaf, o);
                }
                else {
                    stap = std::min(stap, o);
                }
            }
        }
        ch[s] = (uint8_t)(scaf < 6 ? 1 + scaf : (stap < 6 ? 7 + stap : 0));
    }
    for (int n = 0; n < 6; n++) {
        V3 p = A + kUnit[n];
        auto it = at.find(p);
        legal[n] = (it == at.end() || it->second.size() <= 1) ? 1 : 0;
    }
    return 6 * unit_of(d->pos - A) + unit_of(d->ore);
}

void Replica::learned_logits(const uint8_t ch[343], int id, float logits[36]) const {
    const float* W = net;
    float h0[256], h1[256];
    const float* e = W + net_off[1] + (int64_t)id * 16;
    for (int j = 0; j < 256; j++) {
        float s = W[net_off[3] + j];
        for (int site = 0; site < 343; site++) {
            s += W[net_off[0] + (int64_t)(site * 13 + ch[site]) * 256 + j];
        }
        for (int k = 0; k < 16; k++) {
            s = std::fmaf(e[k], W[net_off[2] + k * 256 + j], s);
        }
        h0[j] = std::fmax(s, 0.f);
    }
    auto dense = [&](int64_t wo, int64_t bo, const float* in, float* out, int n_out, bool relu) {
        for (int j = 0; j < n_out; j++) {
            float s = W[bo + j];
            for (int i = 0; i < 256; i++) {
                s = std::fmaf(in[i], W[wo + (int64_t)i * n_out + j], s);
            }
            out[j] = relu ? std::fmax(s, 0.f) : s;
        }
    };
    dense(net_off[4], net_off[5], h0, h1, 256, true);
    dense(net_off[6], net_off[7], h1, h0, 256, true);
    dense(net_off[8], net_off[9], h0, logits, 36, false);
}

// whether staple chain c reaches the scaffold through bound pairs
static bool chain_connected(const std::vector<Dom*>& chain,
                            const std::function<bool(Dom*, std::set<int>&)>& scan) {
    for (Dom* x: chain) {
        Dom* b = x->bd;
        if (b == nullptr || b->c == x->c) {
            continue;
        }
        if (b->c == 0) {
            return true;
        }
        std::set<int> part {x->c};
        if (scan(b, part)) {
            return true;
        }
    }
    return false;
}

void Replica::learned_attempt() {
    learned_step++;
    rng.begin_sub(1, learned_step);
    const int fb = LDO_MAX_MOVETYPES - 2, ls = LDO_MAX_MOVETYPES - 1;
    if (rng.uniform_real() < learned_fallback) {
        attempts[fb]++;
        m_kind = LDO_MOVE_CB_STAPLE_REGROWTH;
        bool acc = attempt(LDO_MOVE_CB_STAPLE_REGROWTH, 0);
        accepts[fb] += acc;
        if (!acc) {
            reset_origami();
            calc_move_bias(0);
        }
        return;
    }
    if (rng.uniform_real() < 0.5) {
        attempts[ls - 2]++;
        learned_exchange();
        return;
    }
    attempts[ls]++;
    if (sys.num_staples == 0) {
        return;
    }
    int w = rng.uniform_int(1, sys.num_staples);
    const std::vector<Dom*>& chain = sys.chains[w];
    int len = (int)chain.size();
    bool first = rng.uniform_real() < 0.5;
    Dom* d = first ? chain[0] : chain[len - 1];
    Dom* a = first ? chain[1] : chain[len - 2];
    int id = std::abs(d->d_ident) - 1;
    if (len < 2 || id < 0 || id >= net_n_ident) {
        throw OrigamiError(LDO_E_MISUSE, "learned move: identity outside the network");
    }
    uint8_t ch[343], legal[6];
    int k_old = learned_context(d, a, ch, legal);
    float logits[36];
    learned_logits(ch, id, logits);
    double w8[36];
    float mx = -INFINITY;
    for (int k = 0; k < 36; k++) {
        if (legal[k / 6]) {
            mx = std::fmax(mx, logits[k]);
        }
    }
    double z = 0;
    for (int k = 0; k < 36; k++) {
        w8[k] = legal[k / 6] ? std::exp((double)logits[k] - (double)mx) : 0.0;
        z += w8[k];
    }
    double u = rng.uniform_real() * z;
    double cum = 0;
    int kn = -1;
    for (int k = 0; k < 36; k++) {
        if (w8[k] == 0) {
            continue;
        }
        kn = k;
        cum += w8[k];
        if (u < cum) {
            break;
        }
    }
    bool accepted = false;
    if (kn < 0) {
        return;
    }
    if (kn == k_old) {
        accepted = true;
    }
    else {
        m_modified.clear();
        m_assigned.clear();
        m_added.clear();
        m_prev_pos.clear();
        m_prev_ore.clear();
        m_modifier = 1;
        sys.violated = false;
        Dom* partner = d->bd;
        auto key = std::make_pair(d->c, d->d);
        m_prev_pos[key] = d->pos;
        m_prev_ore[key] = d->ore;
        m_modified.push_back(key);
        double de = sys.unassign_domain(*d);
        V3 pn = a->pos + kUnit[kn / 6];
        de += sys.set_domain_config(*d, pn, kUnit[kn % 6]);
        m_assigned.push_back(key);
        if (!sys.violated) {
            auto scan = [this](Dom* x, std::set<int>& part) {
                return scan_for_scaffold_domain(x, part);
            };
            bool ok = chain_connected(chain, scan);
            if (ok && partner != nullptr && partner->c != 0 && partner->c != d->c) {
                ok = chain_connected(sys.chain_by_index(partner->c), scan);
            }
            if (ok) {
                de += calc_move_bias(0);
                accepted = test_acceptance((long double)(std::exp(-de) * (w8[k_old] / w8[kn])));
            }
        }
        if (!accepted) {
            reset_origami();
            calc_move_bias(0);
        }
    }
    accepts[ls] += accepted;
}


// the scaffold domain a staple domain grows onto in the learned exchange:
// scaffold domain v - 1 for identity v when it carries identity -v
Dom* Replica::complement_of(Dom* s) const {
    int v = s->d_ident;
    if (v < 1 || v > (int)sys.chains[0].size()) {
        return nullptr;
    }
    Dom* x = sys.chains[0][v - 1];
    return x->d_ident == -v ? x : nullptr;
}

// unnormalised masked softmax of the network around anchor a with d left
// out, its sum, and the output d occupies
void Replica::learned_weights(Dom* d, Dom* a, int id, double w[36], double& z,
                              int& k_of_d) const {
    uint8_t ch[343], legal[6];
    k_of_d = learned_context(d, a, ch, legal);
    float logits[36];
    learned_logits(ch, id, logits);
    float mx = -INFINITY;
    for (int k = 0; k < 36; k++) {
        if (legal[k / 6]) {
            mx = std::fmax(mx, logits[k]);
        }
    }
    z = 0;
    for (int k = 0; k < 36; k++) {
        w[k] = legal[k / 6] ? std::exp((double)logits[k] - (double)mx) : 0.0;
        z += w[k];
    }
}

// learned staple exchange, restated from the sampler's definition (the
// acceptance ratios are derived there from the Met move's stationary weight)
void Replica::learned_exchange() {
    const int xs = LDO_MAX_MOVETYPES - 3;
    m_modified.clear();
    m_assigned.clear();
    m_added.clear();
    m_prev_pos.clear();
    m_prev_ore.clear();
    m_modifier = 1;
    m_delta_e = 0;
    m_rejected = false;
    sys.violated = false;
    double types = (double)(sys.identities.size() - 1);
    auto end = [&](bool changed, bool acc) {
        if (changed && !acc) {
            reset_origami();
            calc_move_bias(0);
        }
        accepts[xs] += acc;
    };
    auto id_of = [](Dom* s) { return std::abs(s->d_ident) - 1; };
    if (rng.uniform_real() < 0.5) {
        if (sys.num_staples == m_lim.max_total_staples) {
            return end(false, false);
        }
        int ident = rng.uniform_int(1, (int)types);
        if ((int)sys.ident_to_index[ident].size() == m_lim.max_type_staples ||
            sys.identities[ident].size() != 2) {
            return end(false, false);
        }
        int i = rng.uniform_int(0, 1);
        int c_i = sys.add_chain(ident);
        m_delta_e += sys.pot.init_e();
        m_added.push_back(c_i);
        std::vector<Dom*> chain = sys.chains.back();
        Dom* g = chain[i];
        Dom* dj = chain[1 - i];
        Dom* x = complement_of(g);
        if (x == nullptr || x->st != Occ::unbound) {
            return end(true, false);
        }
        m_delta_e += sys.set_domain_config(*g, x->pos, -x->ore);
        m_assigned.push_back({g->c, g->d});
        if (sys.violated) {
            return end(true, false);
        }
        double w[36], z;
        int unused;
        learned_weights(dj, g, id_of(dj), w, z, unused);
        double u = rng.uniform_real() * z;
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
            return end(true, false);
        }
        double q = w[kn] / z;
        m_delta_e += sys.set_domain_config(*dj, g->pos + kUnit[kn / 6], kUnit[kn % 6]);
        m_assigned.push_back({dj->c, dj->d});
        if (sys.violated) {
            return end(true, false);
        }
        if (dj->st == Occ::bound && dj->bd == complement_of(dj)) {
            int k_g;
            learned_weights(g, dj, id_of(g), w, z, k_g);
            q += w[k_g] / z;
        }
        m_delta_e += calc_move_bias(0);
        double ratio = std::exp(-m_delta_e) * (sys.reduced_fugacity / 216.0) /
                       (double)sys.num_staples / (0.5 * q / types);
        return end(true, test_acceptance((long double)ratio));
    }
    if (sys.num_staples == 0) {
        return end(false, false);
    }
    int w8 = rng.uniform_int(1, sys.num_staples);
    std::vector<Dom*> staple = sys.chains[w8];
    if (staple.size() != 2 || staple_is_connector(staple)) {
        return end(false, false);
    }
    bool valid[2];
    for (int k = 0; k < 2; k++) {
        valid[k] = staple[k]->st == Occ::bound && staple[k]->bd == complement_of(staple[k]);
    }
    if (!valid[0] && !valid[1]) {
        return end(false, false);
    }
    double q = 0;
    for (int k = 0; k < 2; k++) {
        if (valid[k]) {
            double w[36], z;
            int k_d;
            learned_weights(staple[1 - k], staple[k], id_of(staple[1 - k]), w, z, k_d);
            q += w[k_d] / z;
        }
    }
    double p_ins = 0.5 * q / types;
    double n_before = (double)sys.num_staples;
    met_unassign_domains(staple);
    m_delta_e -= sys.pot.init_e();
    m_delta_e += calc_move_bias(-1);
    double ratio = std::exp(-m_delta_e) * p_ins / ((sys.reduced_fugacity / 216.0) / n_before);
    bool acc = test_acceptance((long double)ratio);
    if (acc) {
        sys.delete_chain(staple[0]->c);
    }
    end(true, acc);
}

} // namespace orc
