// oracle_enum.cpp - exact enumeration of the grand-canonical ensemble of a
// small system.
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).
//
// The ensemble and its weights follow the reference enumerator
// (enumerate.cpp:257-275 add_staple, 480-520 orientation multiplicities,
// 639-664 calc_and_save_weights):
//   weight = prod_staples [staple_M / 6^(2L-1)] * exp(-E) * multiplicity
// with E the system energy including the per-staple initiation energy, the
// first scaffold domain pinned at the origin, and every domain whose
// orientation cannot enter the energy (an unbound domain, or either member
// of a misbound pair) counted as 6 orientations.
//
// The construction is independent of the reference's growthpoint/overcount
// logic: for every scaffold walk, connected staple sets are built
// breadth-first (each new staple binds with at least one domain at a site
// that holds exactly one domain) and de-duplicated as canonical sets, so each
// configuration is visited exactly once.  Orientations are then enumerated
// depth-first on the oracle's System, which supplies energies, constraint
// violations and the order parameters.
#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <map>
#include <set>
#include <unordered_map>
#include <vector>

#include "oracle_core.hpp"

namespace orc {

namespace {

int64_t pack(V3 p) {
    return ((int64_t)(p.x + 512) << 20) | ((int64_t)(p.y + 512) << 10) | (int64_t)(p.z + 512);
}

V3 unpack(int64_t k) {
    return {(int)((k >> 20) & 1023) - 512, (int)((k >> 10) & 1023) - 512, (int)(k & 1023) - 512};
}

// staple placement: identity, then the positions of its domains in order
using Placement = std::vector<int>;
using StapleSet = std::vector<Placement>; // sorted

struct Enumerator {
    System& sys;
    const std::vector<std::vector<int>>& idents;
    Limits lim;
    bool all_orientations;
    bool misbinding_allowed;
    EnumResult& out;

    std::vector<V3> scaffold_pos;

    // ------------------------------------------------------ position sets
    void occupancy_of(const StapleSet& st, std::unordered_map<int64_t, int>& occ) const {
        occ.clear();
        for (V3 p: scaffold_pos) {
            occ[pack(p)]++;
        }
        for (const Placement& pl: st) {
            for (size_t k = 1; k + 2 < pl.size(); k += 3) {
                occ[pack({pl[k], pl[k + 1], pl[k + 2]})]++;
            }
        }
    }

    // all placements of staple type t with domain a at site s whose sites
    // keep every occupancy at most 2
    void placements(int t, int a, V3 s, const std::unordered_map<int64_t, int>& occ,
                    std::vector<Placement>& outp) const {
        int len = (int)idents[t].size();
        std::vector<V3> pos(len);
        pos[a] = s;
        std::unordered_map<int64_t, int> local;
        local[pack(s)] = 1;
        auto count = [&](V3 p) {
            auto it = occ.find(pack(p));
            int n = it == occ.end() ? 0 : it->second;
            auto jt = local.find(pack(p));
            return n + (jt == local.end() ? 0 : jt->second);
        };
        // domains after a, then domains before a
        std::vector<int> order;
        for (int d = a + 1; d < len; d++) {
            order.push_back(d);
        }
        for (int d = a - 1; d >= 0; d--) {
            order.push_back(d);
        }
        std::function<void(size_t)> rec = [&](size_t k) {
            if (k == order.size()) {
                Placement pl {t};
                for (V3 p: pos) {
                    pl.push_back(p.x);
                    pl.push_back(p.y);
                    pl.push_back(p.z);
                }
                outp.push_back(pl);
                return;
            }
            int d = order[k];
            V3 prev = d > a ? pos[d - 1] : pos[d + 1];
            for (const V3& u: kUnit) {
                V3 p = prev + u;
                if (count(p) >= 2) {
                    continue;
                }
                pos[d] = p;
                local[pack(p)]++;
                rec(k + 1);
                if (--local[pack(p)] == 0) {
                    local.erase(pack(p));
                }
            }
        };
        rec(0);
    }

    std::vector<std::vector<StapleSet>> staple_levels() const {
        std::vector<std::vector<StapleSet>> levels;
        levels.push_back({StapleSet {}});
        std::unordered_map<int64_t, int> occ;
        for (int n = 1; n <= lim.max_total_staples; n++) {
            std::set<StapleSet> next;
            for (const StapleSet& st: levels.back()) {
                occupancy_of(st, occ);
                std::vector<int> per_type(idents.size(), 0);
                for (const Placement& pl: st) {
                    per_type[pl[0]]++;
                }
                for (const auto& kv: occ) {
                    if (kv.second != 1) {
                        continue;
                    }
                    V3 s = unpack(kv.first);
                    for (int t = 1; t < (int)idents.size(); t++) {
                        if (per_type[t] >= lim.max_type_staples) {
                            continue;
                        }
                        for (int a = 0; a < (int)idents[t].size(); a++) {
                            std::vector<Placement> pls;
                            placements(t, a, s, occ, pls);
                            for (Placement& pl: pls) {
                                StapleSet ns = st;
                                ns.push_back(pl);
                                std::sort(ns.begin(), ns.end());
                                next.insert(ns);
                            }
                        }
                    }
                }
            }
            levels.emplace_back(next.begin(), next.end());
            if (levels.back().empty()) {
                break;
            }
        }
        return levels;
    }

    // ---------------------------------------------------- orientations
    struct Item {
        Dom* d;
        V3 pos;
        Dom* first; // the domain already at this site when d is placed
        int role;   // 0 alone, 1 first of a complementary pair, 2 first of a
                    // misbound pair, 3 second at its site
    };
    std::vector<Item> items;
    long double prefix {1};

    void leaf(long double mult) {
        int nf = sys.num_fully_bound_pairs;
        std::array<int, 4> key {sys.num_staples, nf, sys.num_bound_pairs - nf, sys.num_stacked};
        long double w = prefix * std::exp(-(long double)sys.energy) * mult;
        out.weights[key] += w;
        out.z += w;
        out.n_configs += mult;
    }

    void dfs(size_t i, long double mult) {
        if (i == items.size()) {
            leaf(mult);
            return;
        }
        Item& it = items[i];
        if (it.role == 3) {
            sys.set_domain_config(*it.d, it.pos, -it.first->ore);
            if (sys.violated) {
                sys.violated = false;
                return;
            }
            dfs(i + 1, mult);
            sys.unassign_domain(*it.d);
            return;
        }
        if (it.role == 1 || all_orientations) {
            for (const V3& o: kUnit) {
                sys.set_domain_config(*it.d, it.pos, o);
                if (sys.violated) {
                    sys.violated = false;
                    continue;
                }
                dfs(i + 1, mult);
                sys.unassign_domain(*it.d);
            }
            return;
        }
        sys.set_domain_config(*it.d, it.pos, kUnit[0]);
        if (sys.violated) {
            sys.violated = false;
            return;
        }
        dfs(i + 1, mult * 6);
        sys.unassign_domain(*it.d);
    }

    void evaluate(const StapleSet& st) {
        std::vector<int> added;
        for (const Placement& pl: st) {
            added.push_back(sys.add_chain(pl[0]));
        }
        items.clear();
        prefix = 1;
        std::unordered_map<int64_t, Dom*> first_at;
        std::unordered_map<int64_t, int> occ;
        std::vector<std::pair<Dom*, V3>> order;
        for (size_t d = 0; d < scaffold_pos.size(); d++) {
            order.push_back({sys.chains[0][d], scaffold_pos[d]});
        }
        for (size_t s = 0; s < st.size(); s++) {
            const Placement& pl = st[s];
            int len = (int)idents[pl[0]].size();
            prefix *= (long double)sys.staple_M / std::pow((long double)6, 2 * len - 1);
            std::vector<Dom*>& ch = sys.chain_by_index(added[s]);
            for (int d = 0; d < len; d++) {
                order.push_back({ch[d], {pl[1 + 3 * d], pl[2 + 3 * d], pl[3 + 3 * d]}});
            }
        }
        for (auto& od: order) {
            occ[pack(od.second)]++;
        }
        bool ok = true;
        for (auto& od: order) {
            int64_t k = pack(od.second);
            Item it {od.first, od.second, nullptr, 0};
            auto f = first_at.find(k);
            if (f != first_at.end()) {
                it.first = f->second;
                it.role = 3;
            }
            else {
                first_at[k] = od.first;
                if (occ[k] == 2) {
                    // find the partner that will arrive later
                    Dom* partner = nullptr;
                    bool seen = false;
                    for (auto& oe: order) {
                        if (oe.first == od.first) {
                            seen = true;
                            continue;
                        }
                        if (seen && pack(oe.second) == k) {
                            partner = oe.first;
                        }
                    }
                    bool comp = partner && partner->d_ident == -od.first->d_ident;
                    if (!comp && !misbinding_allowed) {
                        ok = false;
                    }
                    it.role = comp ? 1 : 2;
                }
            }
            items.push_back(it);
        }
        if (ok) {
            dfs(0, 1);
        }
        for (auto it = added.rbegin(); it != added.rend(); ++it) {
            sys.delete_chain(*it);
        }
    }

    void run() {
        int ns = (int)sys.chains[0].size();
        for (Dom* d: sys.chains[0]) {
            if (d->st != Occ::unassigned) {
                sys.unassign_domain(*d);
            }
        }
        int n_walks = 1;
        for (int i = 1; i < ns; i++) {
            n_walks *= 6;
        }
        for (int w = 0; w < n_walks; w++) {
            scaffold_pos.assign(1, V3 {0, 0, 0});
            int code = w;
            for (int i = 1; i < ns; i++) {
                scaffold_pos.push_back(scaffold_pos.back() + kUnit[code % 6]);
                code /= 6;
            }
            std::unordered_map<int64_t, int> occ;
            bool ok = true;
            for (V3 p: scaffold_pos) {
                if (++occ[pack(p)] > 2) {
                    ok = false;
                }
            }
            if (!ok) {
                continue;
            }
            if (sys.cyclic) {
                V3 dv = scaffold_pos.front() - scaffold_pos.back();
                if (std::abs(dv.x) + std::abs(dv.y) + std::abs(dv.z) != 1) {
                    continue;
                }
            }
            for (const auto& level: staple_levels()) {
                for (const StapleSet& st: level) {
                    evaluate(st);
                }
            }
        }
    }
};

} // namespace

EnumResult enumerate_states(const ldo_system_desc& desc, const std::vector<std::vector<int>>& idents,
                            const std::vector<std::vector<std::string>>& seqs, const Limits& lim,
                            double temp, bool all_orientations) {
    std::vector<ChainSpec> scaffold(1);
    scaffold[0].index = 0;
    scaffold[0].identity = 0;
    for (size_t d = 0; d < idents[0].size(); d++) {
        scaffold[0].pos.push_back(V3 {(int)d, 0, 0});
        scaffold[0].ore.push_back(V3 {0, 0, 1});
    }
    System sys(desc, idents, seqs, scaffold, temp);
    EnumResult out;
    Enumerator e {sys, idents, lim, all_orientations, desc.misbinding_pot != LDO_MISBIND_DISALLOWED, out};
    e.run();
    return out;
}

} // namespace orc
