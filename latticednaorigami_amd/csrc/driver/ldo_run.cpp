// ldo_run.cpp - the reference's simulation loop on the sampler's C ABI.
//
// Reads the reference's inputs (the .inp parameter file of parser.cpp, the
// system JSON of OrigamiInputFile, the moveset JSON of
// setup_movetypes, files.cpp:140-218, 296-306) and writes the reference's
// output files (files.cpp:493-791: .vsf .trj .vcf .states .ores .counts
// .staples .staplestates .times .ene .ops, and .swp for parallel tempering,
// ptmc_simulation.cpp:91-104, 315-322) while every replica runs on the GPU
// through ldo.h.  Supported: simulation_type constant_temp, annealing,
// t_, ut_, hut_, st_ and 2d_parallel_tempering, umbrella_sampling,
// mw_umbrella_sampling and ptmw_umbrella_sampling
// (with the per-window,
// per-iteration files of us_simulation.cpp:103-248: _iter-<n>.* outputs,
// -inp.biases / .biases weights and the .out iteration summary) with the
// movetypes OrientationRotation,
// MetStapleExchange, CBStapleRegrowth, CTCBScaffoldRegrowth,
// CTRGScaffoldRegrowth, CTRGJumpScaffoldRegrowth (the movetypes of the
// reference's examples/moveset_standard.json) and CTCBJumpScaffoldRegrowth.
//
//   ldo_run <parameter file> [key=value ...]
//
// Later key=value arguments override the file, as the reference's command
// line overrides its config file.  Outputs are written after the MC steps
// whose index is a multiple of the output's frequency, as GCMCSimulation::
// simulate writes them (simulation.cpp:620-650); with parallel tempering the
// frequencies must be multiples of exchange_interval.  ldo_flags sets
// ldo_run_params.flags (no reference key).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/ldo.h"
#include "us_loop.hpp"

namespace {

struct Fail : std::runtime_error {
    int code;
    Fail(int c, const std::string& m): std::runtime_error(m), code(c) {}
};

// ------------------------------------------------------------------ JSON
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json& at(const std::string& k) const {
        auto it = obj.find(k);
        if (kind != Obj || it == obj.end()) {
            throw Fail(LDO_E_IO, "JSON key missing: " + k);
        }
        return it->second;
    }
    bool has(const std::string& k) const { return kind == Obj && obj.count(k) > 0; }
};

struct JsonReader {
    const std::string& s;
    size_t i = 0;
    explicit JsonReader(const std::string& t): s(t) {}

    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) {
            i++;
        }
    }
    [[noreturn]] void bad() { throw Fail(LDO_E_IO, "malformed JSON at byte " + std::to_string(i)); }
    char peek() {
        ws();
        if (i >= s.size()) {
            bad();
        }
        return s[i];
    }
    void expect(char c) {
        if (peek() != c) {
            bad();
        }
        i++;
    }
    Json value() {
        char c = peek();
        Json v;
        if (c == '{') {
            i++;
            v.kind = Json::Obj;
            if (peek() == '}') {
                i++;
                return v;
            }
            while (true) {
                Json k = value();
                if (k.kind != Json::Str) {
                    bad();
                }
                expect(':');
                v.obj[k.str] = value();
                if (peek() == ',') {
                    i++;
                    continue;
                }
                expect('}');
                return v;
            }
        }
        if (c == '[') {
            i++;
            v.kind = Json::Arr;
            if (peek() == ']') {
                i++;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                if (peek() == ',') {
                    i++;
                    continue;
                }
                expect(']');
                return v;
            }
        }
        if (c == '"') {
            i++;
            v.kind = Json::Str;
            while (i < s.size() && s[i] != '"') {
                if (s[i] == '\\' && i + 1 < s.size()) {
                    i++;
                }
                v.str += s[i++];
            }
            if (i >= s.size()) {
                bad();
            }
            i++;
            return v;
        }
        if (s.compare(i, 4, "true") == 0 || s.compare(i, 5, "false") == 0) {
            v.kind = Json::Bool;
            v.b = s[i] == 't';
            i += v.b ? 4 : 5;
            return v;
        }
        if (s.compare(i, 4, "null") == 0) {
            i += 4;
            return v;
        }
        size_t end = i;
        while (end < s.size() && std::string("+-0123456789.eE").find(s[end]) != std::string::npos) {
            end++;
        }
        if (end == i) {
            bad();
        }
        v.kind = Json::Num;
        v.num = std::stod(s.substr(i, end - i));
        i = end;
        return v;
    }
};

std::string slurp(const std::string& path) {
    std::ifstream f(path);
    if (!f.good()) {
        throw Fail(LDO_E_IO, path + ": cannot be read");
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

Json read_json(const std::string& path) {
    std::string text = slurp(path);
    JsonReader r(text);
    return r.value();
}

// --------------------------------------------------------------- params
struct Params {
    std::map<std::string, std::string> kv;

    std::string str(const std::string& k, const std::string& def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : it->second;
    }
    double num(const std::string& k, double def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : std::stod(it->second);
    }
    long long integer(const std::string& k, long long def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : std::stoll(it->second);
    }
    bool flag(const std::string& k, bool def) const {
        auto it = kv.find(k);
        if (it == kv.end()) {
            return def;
        }
        return it->second == "true" || it->second == "1";
    }
    std::vector<double> list(const std::string& k) const {
        std::vector<double> out;
        std::stringstream ss(str(k, ""));
        double v;
        while (ss >> v) {
            out.push_back(v);
        }
        return out;
    }
    std::vector<std::string> words(const std::string& k) const {
        std::vector<std::string> out;
        std::stringstream ss(str(k, ""));
        std::string w;
        while (ss >> w) {
            out.push_back(w);
        }
        return out;
    }
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

void parse_assignment(const std::string& line, Params& p) {
    std::string l = line.substr(0, line.find('#'));
    size_t eq = l.find('=');
    if (eq == std::string::npos) {
        if (!trim(l).empty()) {
            throw Fail(LDO_E_IO, "parameter line without '=': " + line);
        }
        return;
    }
    p.kv[trim(l.substr(0, eq))] = trim(l.substr(eq + 1));
}

// "a/b" fractions or plain numbers (utility.cpp:266-282)
double fraction(const Json& v) {
    if (v.kind == Json::Num) {
        return v.num;
    }
    size_t slash = v.str.find('/');
    if (slash == std::string::npos) {
        return std::stod(v.str);
    }
    return std::stod(v.str.substr(0, slash)) / std::stod(v.str.substr(slash + 1));
}

int enum_of(const std::string& v, const std::vector<std::string>& names, const char* what) {
    for (size_t k = 0; k < names.size(); k++) {
        if (names[k] == v) {
            return (int)k;
        }
    }
    throw Fail(LDO_E_NOTIMPL, std::string(what) + " " + v + " is not supported");
}

// ----------------------------------------------------------------- run
struct Chain {
    int index, identity;
    std::vector<int> pos, ore; // 3 per domain
};

struct Run {
    Params p;
    std::vector<std::vector<int>> identities;
    std::vector<std::string> move_labels; // the moveset's "label" fields
    std::vector<int32_t> move_kinds;
    int max_staple_size = 2, max_total_staples = 999, n_scaffold = 0;
    void* ctx = nullptr;
    int R = 1;
    std::vector<double> temps;

    void check(int rc) {
        if (rc != LDO_OK) {
            throw Fail(rc, ldo_last_error(ctx));
        }
    }

    std::vector<Chain> config(int r) {
        int cap_c = max_total_staples + 1;
        int cap_d = n_scaffold + max_total_staples * max_staple_size;
        std::vector<int32_t> idx(cap_c), idn(cap_c), len(cap_c), pos(3 * cap_d), ore(3 * cap_d);
        ldo_config c {};
        c.chain_index = idx.data();
        c.chain_identity = idn.data();
        c.chain_length = len.data();
        c.positions = pos.data();
        c.orientations = ore.data();
        check(ldo_get_config(ctx, r, &c, cap_c, cap_d));
        std::vector<Chain> out;
        int o = 0;
        for (int k = 0; k < c.n_chains; k++) {
            Chain ch {idx[k], idn[k], {}, {}};
            ch.pos.assign(pos.begin() + 3 * o, pos.begin() + 3 * (o + len[k]));
            ch.ore.assign(ore.begin() + 3 * o, ore.begin() + 3 * (o + len[k]));
            o += len[k];
            out.push_back(ch);
        }
        return out;
    }
};

// occupancy state of every domain (OrigamiStateOutputFile codes: 1 unbound,
// 2 bound, 3 misbound) from the positions: a site holds one domain, or a
// pair that is bound when the identities are complementary
std::vector<std::vector<int>> domain_states(const Run& run, const std::vector<Chain>& chains) {
    std::map<std::vector<int>, std::vector<std::pair<int, int>>> at;
    for (size_t c = 0; c < chains.size(); c++) {
        for (size_t d = 0; d < chains[c].pos.size() / 3; d++) {
            std::vector<int> p(chains[c].pos.begin() + 3 * d, chains[c].pos.begin() + 3 * d + 3);
            at[p].push_back({(int)c, (int)d});
        }
    }
    std::vector<std::vector<int>> st(chains.size());
    for (size_t c = 0; c < chains.size(); c++) {
        st[c].assign(chains[c].pos.size() / 3, 1);
    }
    for (auto& kv: at) {
        if (kv.second.size() == 2) {
            auto a = kv.second[0], b = kv.second[1];
            int ia = run.identities[chains[a.first].identity][a.second];
            int ib = run.identities[chains[b.first].identity][b.second];
            int code = (ia == -ib) ? 2 : 3;
            st[a.first][a.second] = code;
            st[b.first][b.second] = code;
        }
    }
    return st;
}

struct Outputs {
    std::string base;
    long long f_cfg = 0, f_vtf = 0, f_counts = 0, f_times = 0, f_ene = 0, f_ops = 0;
    std::ofstream trj, vcf, states, ores, counts, staples, staplestates, times, ene, ops;
    std::vector<std::string> op_tags;
    int max_domains = 0;

    static void open(std::ofstream& f, const std::string& path) {
        f.open(path);
        if (!f.good()) {
            throw Fail(LDO_E_IO, path + ": output path does not exist");
        }
    }
};

std::unique_ptr<Outputs> setup_outputs(const Run& run, const std::string& base) {
    auto o = std::make_unique<Outputs>();
    const Params& p = run.p;
    o->base = base;
    o->max_domains = run.max_staple_size * run.max_total_staples + run.n_scaffold;
    {
        std::ofstream vsf;
        Outputs::open(vsf, base + ".vsf");
        vsf << "atom 0:" << run.n_scaffold << " radius 0.25 type scaffold\n";
        vsf << "atom " << run.n_scaffold << ":" << o->max_domains - 1;
        vsf << " radius 0.25 type staple";
    }
    o->f_cfg = p.integer("configs_output_freq", 0);
    o->f_vtf = p.integer("vtf_output_freq", 0);
    o->f_counts = p.integer("counts_output_freq", 0);
    o->f_times = p.integer("times_output_freq", 0);
    o->f_ene = p.integer("energies_output_freq", 0);
    o->f_ops = p.integer("order_params_output_freq", 0);
    if (o->f_cfg) {
        Outputs::open(o->trj, base + ".trj");
    }
    if (o->f_vtf) {
        Outputs::open(o->vcf, base + ".vcf");
        Outputs::open(o->states, base + ".states");
        Outputs::open(o->ores, base + ".ores");
    }
    if (o->f_counts) {
        Outputs::open(o->counts, base + ".counts");
        Outputs::open(o->staples, base + ".staples");
        Outputs::open(o->staplestates, base + ".staplestates");
    }
    if (o->f_times) {
        Outputs::open(o->times, base + ".times");
        o->times << "step time\n";
    }
    if (o->f_ene) {
        Outputs::open(o->ene, base + ".ene");
        o->ene << "step tenergy henthalpy hentropy stacking bias\n";
        o->ene.precision(10);
    }
    if (o->f_ops) {
        Outputs::open(o->ops, base + ".ops");
        o->op_tags = p.words("ops_to_output");
        for (auto& t: o->op_tags) {
            if (t != "numfulldomains" && t != "nummisdomains" && t != "numstackedpairs" &&
                t != "numstaples") {
                throw Fail(LDO_E_NOTIMPL, "order parameter " + t + " is not supported");
            }
            o->ops << t << ", ";
        }
        o->ops << "\n";
        o->ops.precision(10);
        o->ops.flush();
    }
    return o;
}

bool due(long long freq, long long step) { return freq != 0 && step % freq == 0; }

// Outputs of replica r at `step`; bias_rep >= 0 takes the bias column from
// that replica instead (the replica that now holds the window whose stored
// bias the record must show, after a window exchange)
// GCMCSimulation::write_log_entry (simulation.cpp:667-704) for replica r
// after its step `step`; the helix-triplet and junction-quadruplet counters
// of the reference are never updated (origami_system.hpp:194-195) and print 0,
// as does the domain-update bias (no domain-level biases here).
void write_log_entry(Run& run, std::ostream& log, int r, long long step) {
    ldo_obs o {};
    run.check(ldo_get_observables(run.ctx, r, &o));
    std::vector<Chain> chains = run.config(r);
    std::vector<int> counts(run.identities.size() - 1, 0);
    for (size_t c = 1; c < chains.size(); c++) {
        counts[chains[c].identity - 1]++;
    }
    int unique = 0;
    for (int n: counts) {
        unique += n > 0 ? 1 : 0;
    }
    log << "Step: " << step << "\n";
    log << "Temperature: " << o.temperature << "\n";
    log << "Bound staples: " << o.num_staples << "\n";
    log << "Unique bound staples: " << unique << "\n";
    log << "Fully bound domain pairs: " << o.num_fully_bound_domain_pairs << "\n";
    log << "Misbound domain pairs: " << o.num_misbound_domain_pairs << "\n";
    log << "Stacked domain pairs: " << o.num_stacked_domain_pairs << "\n";
    log << "Linear helix triplets: " << 0 << "\n";
    log << "Stacked junction quadruplets: " << 0 << "\n";
    log << "Staple counts: ";
    for (int n: counts) {
        log << n << " ";
    }
    log << "\n";
    log << "System energy: " << o.energy << "\n";
    log << "Total external bias: " << o.bias << "\n";
    log << "Domain update external bias: " << 0.0 << "\n";
    log << "Move update external bias: " << o.bias << "\n";
    log << "Movetype: " << run.move_labels.at(o.last_move) << "\n";
    log << "Accepted: " << std::boolalpha << (o.last_accepted != 0) << std::noboolalpha << "\n";
    log << "\n";
}

// GCMCSimulation::write_log_summary (simulation.cpp:706-718) of a one-replica
// run: each movetype's header (movetypes.cpp:88-96) to the log, and the
// movetype summaries (met_movetypes.cpp:145-190, cb_movetypes.cpp:264-290,
// 626-678, 816-870, rg_movetypes.cpp:583-627, 745-789; none for orientation
// rotation) to the .moves file.
void write_log_summary(Run& run, std::ostream& log, const std::string& moves_path) {
    ldo_counters c {};
    run.check(ldo_get_counters(run.ctx, &c));
    int T = (int)run.identities.size();
    int ns = run.n_scaffold;
    int S1 = run.max_total_staples + 1;
    std::vector<int64_t> tr((size_t)(6 * T + 4 * (ns + 1) + 4 * S1 + 2));
    int64_t n_tr = 0;
    run.check(ldo_get_move_tracking(run.ctx, 0, tr.data(), (int64_t)tr.size(), &n_tr));
    auto at = [&](int part, int i) { return tr[(size_t)part * T + i]; };
    std::ofstream moves;
    Outputs::open(moves, moves_path);
    auto header = [&](std::ostream& os, size_t m) {
        int att = (int)c.attempts[m], acc = (int)c.accepts[m];
        os << "Movetype: " << run.move_labels[m] << "\n";
        os << "    Attempts: " << att << "\n";
        os << "    Accepts: " << acc << "\n";
        os << "    Frequency: " << static_cast<double>(acc) / att << "\n";
    };
    log << "Run summary" << "\n\n";
    for (size_t m = 0; m < run.move_kinds.size(); m++) {
        header(log, m);
        int kind = run.move_kinds[m];
        if (kind == LDO_MOVE_MET_STAPLE_EXCHANGE) {
            header(moves, m);
            std::vector<int> types;
            for (int t = 1; t < T; t++) {
                if (at(0, t) + at(2, t) > 0) {
                    types.push_back(t);
                }
            }
            moves << "       Exchange multipliers\n        ";
            for (size_t i = 0; i < types.size(); i++) {
                moves << 1 << ", ";
            }
            moves << "\n";
            for (int t: types) {
                int iats = (int)at(0, t), iacs = (int)at(1, t);
                int dats = (int)at(2, t), dacs = (int)at(3, t);
                moves << "    Staple type: " << t << "\n";
                moves << "        Insertion attempts: " << iats << "\n";
                moves << "        Insertion accepts: " << iacs << "\n";
                moves << "        Insertion frequency: " << static_cast<float>(iacs) / iats
                      << "\n";
                moves << "        Deletion attempts: " << dats << "\n";
                moves << "        Deletion accepts: " << dacs << "\n";
                moves << "        Deletion frequency: " << static_cast<float>(dacs) / dats
                      << "\n";
            }
        }
        else if (kind == LDO_MOVE_CB_STAPLE_REGROWTH) {
            header(moves, m);
            moves << "\n";
            for (int t = 1; t < T; t++) {
                int ats = (int)at(4, t), acs = (int)at(5, t);
                if (ats == 0) {
                    continue;
                }
                moves << "    Staple type: " << t << "\n";
                moves << "        Attempts: " << ats << "\n";
                moves << "        Accepts: " << acs << "\n";
                moves << "        Frequency: " << static_cast<double>(acs) / ats << "\n";
            }
        }
        else if (kind == LDO_MOVE_CTCB_SCAFFOLD_REGROWTH) {
            header(moves, m);
            const int64_t* len_att = tr.data() + 6 * T;
            const int64_t* len_acc = len_att + (ns + 1);
            const int64_t* st_att = len_acc + (ns + 1);
            const int64_t* st_acc = st_att + S1;
            for (int l = 0; l <= ns; l++) {
                if (len_att[l] == 0) {
                    continue;
                }
                moves << "    Number of scaffold domains: " << l << "\n";
                moves << "        Attempts: " << (int)len_att[l] << "\n";
                moves << "        Accepts: " << (int)len_acc[l] << "\n";
                moves << "        Frequency: " << static_cast<double>(len_acc[l]) / len_att[l]
                      << "\n";
            }
            moves << "\n";
            for (int s = 0; s < S1; s++) {
                if (st_att[s] == 0) {
                    continue;
                }
                moves << "    Number of staples: " << s << "\n";
                moves << "        Attempts: " << (int)st_att[s] << "\n";
                moves << "        Accepts: " << (int)st_acc[s] << "\n";
                moves << "        Frequency: " << static_cast<double>(st_acc[s]) / st_att[s]
                      << "\n";
            }
        }
        else if (kind == LDO_MOVE_CTRG_SCAFFOLD_REGROWTH) {
            // CTRGScaffoldRegrowthMCMovetype::write_log_summary (rg_movetypes.cpp:583-627)
            header(moves, m);
            const int64_t* len_att = tr.data() + 6 * T + 2 * (ns + 1) + 2 * S1;
            const int64_t* len_acc = len_att + (ns + 1);
            for (int l = 0; l <= ns; l++) {
                if (len_att[l] == 0) {
                    continue;
                }
                moves << "    Number of scaffold domains: " << l << "\n";
                moves << "        Attempts: " << (int)len_att[l] << "\n";
                moves << "        Accepts: " << (int)len_acc[l] << "\n";
                moves << "        Frequency: " << static_cast<double>(len_acc[l]) / len_att[l]
                      << "\n";
            }
            moves << "\n";
        }
        else if (kind == LDO_MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH) {
            // CTCBJumpScaffoldRegrowthMCMovetype::write_log_summary
            // (cb_movetypes.cpp:816-870): every move under 0 scaffold domains
            header(moves, m);
            const int64_t* st_att = tr.data() + 6 * T + 4 * (ns + 1) + 2 * S1 + 2;
            const int64_t* st_acc = st_att + S1;
            int64_t att = 0, acc = 0;
            for (int s = 0; s < S1; s++) {
                att += st_att[s];
                acc += st_acc[s];
            }
            if (att != 0) {
                moves << "    Number of scaffold domains: " << 0 << "\n";
                moves << "        Attempts: " << (int)att << "\n";
                moves << "        Accepts: " << (int)acc << "\n";
                moves << "        Frequency: " << static_cast<double>(acc) / att << "\n";
            }
            moves << "\n";
            for (int s = 0; s < S1; s++) {
                if (st_att[s] == 0) {
                    continue;
                }
                moves << "    Number of staples: " << s << "\n";
                moves << "        Attempts: " << (int)st_att[s] << "\n";
                moves << "        Accepts: " << (int)st_acc[s] << "\n";
                moves << "        Frequency: " << static_cast<double>(st_acc[s]) / st_att[s]
                      << "\n";
            }
        }
        else if (kind == LDO_MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH) {
            // CTRGJumpScaffoldRegrowthMCMovetype::write_log_summary
            // (rg_movetypes.cpp:745-789): every move is tracked under 0 domains
            header(moves, m);
            const int64_t* jt = tr.data() + 6 * T + 4 * (ns + 1) + 2 * S1;
            if (jt[0] != 0) {
                moves << "    Number of scaffold domains: " << 0 << "\n";
                moves << "        Attempts: " << (int)jt[0] << "\n";
                moves << "        Accepts: " << (int)jt[1] << "\n";
                moves << "        Frequency: " << static_cast<double>(jt[1]) / jt[0] << "\n";
            }
            moves << "\n";
        }
        log << "\n";
    }
}

void write_outputs(Run& run, Outputs& o, int r, long long step, double seconds,
                   int bias_rep = -1) {
    bool need_cfg = due(o.f_cfg, step) || due(o.f_vtf, step) || due(o.f_counts, step);
    std::vector<Chain> chains;
    std::vector<std::vector<int>> st;
    if (need_cfg) {
        chains = run.config(r);
        st = domain_states(run, chains);
    }
    if (due(o.f_cfg, step)) {
        o.trj << step << "\n";
        for (auto& ch: chains) {
            o.trj << ch.index << " " << ch.identity << "\n";
            for (int v: ch.pos) {
                o.trj << v << " ";
            }
            o.trj << "\n";
            for (int v: ch.ore) {
                o.trj << v << " ";
            }
            o.trj << "\n";
        }
        o.trj << "\n";
        o.trj.flush();
    }
    if (due(o.f_vtf, step)) {
        o.vcf << "timestep\n";
        int written = 0;
        for (size_t c = 0; c < chains.size(); c++) {
            int n = 0;
            for (size_t d = 0; d < chains[c].pos.size() / 3; d++, n++) {
                for (int i = 0; i < 3; i++) {
                    o.vcf << chains[c].pos[3 * d + i] << " ";
                    o.ores << chains[c].ore[3 * d + i] << " ";
                }
                o.vcf << "\n";
                o.states << st[c][d] << " ";
            }
            for (; n < run.max_staple_size; n++) {
                o.vcf << "0 0 0 \n";
                o.ores << "0 0 0 ";
                o.states << "-1 ";
            }
            written += n;
        }
        for (int k = written; k < o.max_domains; k++) {
            o.vcf << "0 0 0 \n";
            o.ores << "0 0 0 ";
            o.states << "-1 ";
        }
        o.vcf << "\n";
        o.ores << "\n";
        o.states << "\n";
        o.vcf.flush();
        o.ores.flush();
        o.states.flush();
    }
    ldo_obs ob {};
    if (due(o.f_counts, step) || due(o.f_ene, step) || due(o.f_ops, step)) {
        run.check(ldo_get_observables(run.ctx, r, &ob));
        if (bias_rep >= 0) {
            ldo_obs ob2 {};
            run.check(ldo_get_observables(run.ctx, bias_rep, &ob2));
            ob.bias = ob2.bias;
        }
    }
    if (due(o.f_counts, step)) {
        int n_types = (int)run.identities.size();
        std::vector<int> per(n_types, 0);
        std::vector<int> full(n_types, 0);
        for (size_t c = 1; c < chains.size(); c++) {
            per[chains[c].identity]++;
            bool all = true;
            for (int s: st[c]) {
                all = all && s == 2;
            }
            full[chains[c].identity] |= all ? 1 : 0;
        }
        int unique = 0;
        for (int t = 1; t < n_types; t++) {
            unique += per[t] > 0;
        }
        o.counts << step << " " << ob.num_staples << " " << unique << " "
                 << ob.num_bound_domain_pairs << " " << ob.num_fully_bound_domain_pairs << " "
                 << ob.num_misbound_domain_pairs << " \n";
        o.staples << step << " ";
        o.staplestates << step << " ";
        for (int t = 1; t < n_types; t++) {
            o.staples << per[t] << " ";
            o.staplestates << full[t] << " ";
        }
        o.staples << "\n";
        o.staplestates << "\n";
        o.counts.flush();
        o.staples.flush();
        o.staplestates.flush();
    }
    if (due(o.f_times, step)) {
        o.times << step << " " << seconds << "\n";
        o.times.flush();
    }
    if (due(o.f_ene, step)) {
        o.ene << step << " " << ob.energy << " " << ob.enthalpy << " " << ob.entropy << " "
              << ob.stacking_energy << " " << ob.bias << " \n";
        o.ene.flush();
    }
    if (due(o.f_ops, step)) {
        o.ops << step;
        for (auto& t: o.op_tags) {
            int v = t == "numfulldomains"    ? ob.num_fully_bound_domain_pairs
                    : t == "nummisdomains"   ? ob.num_misbound_domain_pairs
                    : t == "numstackedpairs" ? ob.num_stacked_domain_pairs
                                             : ob.num_staples;
            o.ops << " " << v;
        }
        o.ops << "\n";
        o.ops.flush();
    }
}

long long gcd_of(long long a, long long b) { return b == 0 ? a : gcd_of(b, a % b); }

// OrigamiTrajInputFile::read_config (files.cpp:129-218): skip `frame`
// frames (each ends with an empty line), read the step line, then chains
// (index identity / positions / orientations) until an empty line
std::vector<Chain> read_traj_frame(const std::string& path, int frame) {
    std::ifstream f(path);
    if (!f.good()) {
        throw Fail(LDO_E_IO, "Trajectory input file " + path + " does not exist");
    }
    std::string line;
    for (int i = 0; i < frame; i++) {
        while (true) {
            if (!std::getline(f, line)) {
                throw Fail(LDO_E_IO, "Step " + std::to_string(frame) +
                                         " not found in trajectory input file" + path);
            }
            if (line.empty()) {
                break;
            }
        }
    }
    std::getline(f, line); // the step number
    std::vector<Chain> out;
    auto ints = [](const std::string& s) {
        std::vector<int> v;
        std::stringstream ss(s);
        int x;
        while (ss >> x) {
            v.push_back(x);
        }
        v.resize(v.size() / 3 * 3);
        return v;
    };
    while (std::getline(f, line) && !line.empty()) {
        Chain ch {};
        std::stringstream ss(line);
        ss >> ch.index >> ch.identity;
        std::string pl, ol;
        std::getline(f, pl);
        std::getline(f, ol);
        ch.pos = ints(pl);
        ch.ore = ints(ol);
        if (ch.pos.size() != ch.ore.size()) {
            throw Fail(LDO_E_IO, path + ": positions and orientations differ in length");
        }
        out.push_back(ch);
    }
    return out;
}

void set_config_from(Run& run, int r, const std::vector<Chain>& chains) {
    std::vector<int32_t> idx, idn, len, pos, ore;
    for (auto& ch: chains) {
        idx.push_back(ch.index);
        idn.push_back(ch.identity);
        len.push_back((int32_t)(ch.pos.size() / 3));
        pos.insert(pos.end(), ch.pos.begin(), ch.pos.end());
        ore.insert(ore.end(), ch.ore.begin(), ch.ore.end());
    }
    ldo_config c {};
    c.n_chains = (int32_t)chains.size();
    c.chain_index = idx.data();
    c.chain_identity = idn.data();
    c.chain_length = len.data();
    c.positions = pos.data();
    c.orientations = ore.data();
    run.check(ldo_set_config(run.ctx, r, &c));
}

// ------------------------------------------------ multi-window umbrella sampling
// MWUSGCMCSimulation (us_simulation.cpp:466-600) with every window a replica
// of one context: the windows file, the window restraint and the grid bias of
// the bias-function file (bias_functions.cpp:300-420), the order-parameter
// file's tags (order_params.cpp), then the equilibration run and the
// adaptive iterations of USGCMCSimulation::run (93-170).
struct MwusSetup {
    int op = LDO_OP_NUMFULLDOMAINS;
    int grid_len = 0;
    std::vector<ldo_window> windows;
    std::vector<std::string> postfix; // MWUSGCMCSimulation::setup_window_variables
};

// single: SimpleUSGCMCSimulation (simulation_type umbrella_sampling), one
// window whose restraint, if any, is the bias file's own well; otherwise the
// windows file of MWUSGCMCSimulation
MwusSetup parse_mwus(const Run& run, bool single) {
    const Params& p = run.p;
    MwusSetup m;
    // order-parameter tags and types
    std::map<std::string, std::string> op_type;
    const Json ops = read_json(p.str("order_parameter_file", "")).at("origami").at("order_params");
    for (auto& o: ops.arr) {
        op_type[o.at("tag").str] = o.at("type").str;
    }
    // windows file: bias tags, then "min, max" per window (parse_windows_file, 552-600)
    std::string wpath = p.str("windows_file", "");
    std::ifstream wf;
    std::string line;
    std::vector<std::string> wtags;
    if (!single) {
        wf.open(wpath);
        if (!wf.good()) {
            throw Fail(LDO_E_IO, "Windows file " + wpath + " does not exist");
        }
        std::getline(wf, line);
    }
    if (!single) {
        std::stringstream ss(line);
        std::string t;
        while (ss >> t) {
            wtags.push_back(t);
        }
    }
    if (!single && wtags.size() != 1) {
        throw Fail(LDO_E_NOTIMPL, "windows over one bias function only (the file names " +
                                      std::to_string(wtags.size()) + ")");
    }
    std::vector<std::pair<int, int>> ranges;
    while (!single && std::getline(wf, line)) {
        if (trim(line).empty()) {
            continue;
        }
        size_t comma = line.find(',');
        if (comma == std::string::npos) {
            throw Fail(LDO_E_IO, wpath + ": window line without a comma: " + line);
        }
        std::stringstream a(line.substr(0, comma)), b(line.substr(comma + 1));
        int lo, hi;
        if (!(a >> lo) || !(b >> hi)) {
            throw Fail(LDO_E_IO, wpath + ": bad window line: " + line);
        }
        int extra;
        if (a >> extra || b >> extra) {
            throw Fail(LDO_E_NOTIMPL, "windows over one order parameter only");
        }
        ranges.push_back({lo, hi});
    }
    if (!single && ranges.empty()) {
        throw Fail(LDO_E_IO, wpath + ": no windows");
    }
    // bias functions: the windows' restraint and the grid bias, on one order parameter
    std::string grid_tag = p.str("us_grid_bias_tag", "");
    const Json bfs =
        read_json(p.str("bias_functions_file", "")).at("origami").at("bias_functions");
    const Json* well = nullptr;
    const Json* grid = nullptr;
    for (auto& b: bfs.arr) {
        const std::string& tag = b.at("tag").str;
        if (tag == grid_tag) {
            grid = &b;
        }
        else if ((single && !well) || (!single && tag == wtags[0])) {
            well = &b;
        }
        else {
            throw Fail(LDO_E_NOTIMPL, "bias function " + tag +
                                          ": only the window restraint and the grid bias");
        }
    }
    if ((!single && !well) || !grid) {
        throw Fail(LDO_E_IO, "bias-function file lacks " + (grid ? wtags[0] : grid_tag));
    }
    if (grid->at("type").str != "Grid" || grid->at("ops").arr.size() != 1) {
        throw Fail(LDO_E_NOTIMPL, "us_grid_bias_tag must name a Grid bias over one order parameter");
    }
    const std::string& op_tag = grid->at("ops").arr[0].str;
    if (well && (well->at("ops").arr.size() != 1 || well->at("ops").arr[0].str != op_tag)) {
        throw Fail(LDO_E_NOTIMPL, "the window restraint and the grid bias must share their order parameter");
    }
    auto t = op_type.find(op_tag);
    if (t == op_type.end()) {
        throw Fail(LDO_E_IO, "order parameter " + op_tag + " not in the order-parameter file");
    }
    if (t->second == "NumBoundDomainPairs") {
        m.op = LDO_OP_NUMFULLDOMAINS;
    }
    else if (t->second == "NumStaples") {
        m.op = LDO_OP_NUMSTAPLES;
    }
    else {
        throw Fail(LDO_E_NOTIMPL, "umbrella sampling over " + t->second + " is not supported");
    }
    // every value either parameter can take lies on the grid
    m.grid_len = run.n_scaffold + run.max_total_staples * run.max_staple_size + 1;
    if (single) {
        // the well's own range (bias_functions.cpp:309-430), or no restraint
        ranges.push_back(well ? std::make_pair((int)well->at("min_op").num,
                                               (int)well->at("max_op").num)
                              : std::make_pair(0, m.grid_len - 1));
    }
    const std::string wtype = well ? well->at("type").str : "";
    double mult = p.num("bias_functions_mult", 1);
    for (auto& r: ranges) {
        ldo_window w {};
        w.op = m.op;
        w.min_op = r.first;
        w.max_op = r.second;
        if (!well) {
            w.well_bias = 0;
            w.min_bias = 0;
            w.slope = 0;
        }
        else if (wtype == "LinearStepWell") {
            w.well_bias = well->at("well_bias").num;
            w.min_bias = well->at("min_bias").num;
            w.slope = well->at("slope").num;
        }
        else if (wtype == "SquareWell") {
            // outside_bias everywhere outside: a LinearStepWell of slope 0
            w.well_bias = well->at("well_bias").num;
            w.min_bias = well->at("outside_bias").num;
            w.slope = 0;
        }
        else {
            throw Fail(LDO_E_NOTIMPL, "window restraint of type " + wtype);
        }
        w.off_grid_bias = 0;
        w.bias_mult = mult;
        m.windows.push_back(w);
        m.postfix.push_back(single ? std::string() : "_win-" + std::to_string(r.first) + "--" +
                                                         std::to_string(r.second));
    }
    for (const char* k: {"restart_us_iter"}) {
        if (p.flag(k, false)) {
            throw Fail(LDO_E_NOTIMPL, std::string(k) + " (the iteration archives) is not supported");
        }
    }
    return m;
}

// USGCMCSimulation::read_weights (us_simulation.cpp:190-204)
void read_weights(const std::string& path, ldo_us::WindowState& u) {
    std::ifstream f(path);
    if (!f.good()) {
        throw Fail(LDO_E_IO, "Restart bias file " + path + " does not exist");
    }
    const Json j = read_json(path);
    for (auto& e: j.at("biases").arr) {
        if (e.at("point").arr.size() != 1) {
            throw Fail(LDO_E_IO, path + ": grid points of one order parameter expected");
        }
        u.E_w[(int)e.at("point").arr[0].num] = e.at("bias").num;
    }
}

int run_mwus(Run& run, const ldo_system_desc& sd, ldo_run_params rp, const ldo_config& init,
             bool pt, bool single = false) {
    const Params& p = run.p;
    MwusSetup m = parse_mwus(run, single);
    // PTMWUS (us_simulation.cpp:602-919): windows swap configurations every
    // exchange_interval steps; the swap file records which configuration
    // each window holds whenever the iteration's step count is a multiple
    // of configs_output_freq
    long long interval = p.integer("exchange_interval", 0);
    long long cfg_freq = p.integer("configs_output_freq", 0);
    if (pt) {
        if (interval < 1 || cfg_freq < 1) {
            throw Fail(LDO_E_ARG, "ptmw_umbrella_sampling needs exchange_interval and "
                                  "configs_output_freq");
        }
        for (const char* k: {"configs_output_freq", "vtf_output_freq", "counts_output_freq",
                             "times_output_freq", "energies_output_freq",
                             "order_params_output_freq"}) {
            if (p.integer(k, 0) % interval != 0) {
                throw Fail(LDO_E_NOTIMPL, "output frequencies must be multiples of exchange_interval");
            }
        }
        rp.exchange_mode = LDO_EXCHANGE_WINDOWS;
    }
    int W = (int)m.windows.size();
    std::string base = p.str("output_filebase", "");
    std::vector<ldo_us::WindowState> us(W);
    // each window's summary stream: its .out file (MWUSGCMCSimulation::
    // update_internal, us_simulation.cpp:548-551), or standard output for a
    // single window, whose constructor also reports its inputs there
    // (USGCMCSimulation::USGCMCSimulation, 31-91)
    std::vector<std::unique_ptr<std::ofstream>> summary_files;
    std::vector<std::ostream*> summary;
    for (int w = 0; w < W; w++) {
        if (p.flag("read_biases", false)) {
            read_weights(single ? p.str("biases_file", "")
                                : p.str("biases_filebase", "") + m.postfix[w] + ".biases",
                         us[w]);
        }
        if (single) {
            summary.push_back(&std::cout);
            continue;
        }
        summary_files.push_back(std::make_unique<std::ofstream>(base + m.postfix[w] + ".out"));
        if (!summary_files.back()->good()) {
            throw Fail(LDO_E_IO, base + m.postfix[w] + ".out: output path does not exist");
        }
        summary.push_back(summary_files.back().get());
    }
    if (single) {
        std::cout << (p.flag("read_biases", false) ? "Reading biases from file\n"
                                                   : "No biases read in\n");
        std::cout << (p.flag("restart_from_config", false)
                          ? "Reading starting config from file\n"
                          : "Starting from configuration in system file\n");
        std::cout << "Starting new iteration\n";
    }
    run.R = W;
    run.temps.assign(W, p.num("temp", 300));
    rp.rng_mode = LDO_RNG_PHILOX;
    int rc = ldo_create(&sd, &rp, &init, run.R, run.temps.data(), &run.ctx);
    if (rc != LDO_OK) {
        throw Fail(rc, ldo_last_error(nullptr));
    }
    if (p.flag("restart_from_config", false)) {
        std::string fb = p.str("restart_traj_filebase", "");
        std::string post = p.str("restart_traj_postfix", ".trj");
        // restart_traj_files names each window's file in window order
        // (MWUSGCMCSimulation::setup_window_sims, us_simulation.cpp:542-546);
        // the reference never parses restart_steps into the list it reads
        // there, so here restart_steps, when given, lists each window's step
        // and restart_step serves every window otherwise
        std::vector<std::string> files = p.words("restart_traj_files");
        std::vector<double> steps = p.list("restart_steps");
        if (!single && !files.empty() && (int)files.size() != W) {
            throw Fail(LDO_E_ARG, "restart_traj_files must name one file per window");
        }
        if (!single && !steps.empty() && (int)steps.size() != W) {
            throw Fail(LDO_E_ARG, "restart_steps must list one step per window");
        }
        for (int w = 0; w < W; w++) {
            std::string file = single ? p.str("restart_traj_file", "")
                               : files.empty() ? fb + m.postfix[w] + post
                                               : files[w];
            int step = (!single && !steps.empty()) ? (int)steps[w]
                                                   : (int)p.integer("restart_step", 0);
            set_config_from(run, w, read_traj_frame(file, step));
        }
    }
    std::vector<double> grid((size_t)W * m.grid_len);
    auto upload_grid = [&]() {
        for (int w = 0; w < W; w++) {
            std::vector<double> g = ldo_us::dense_grid(us[w], m.grid_len);
            std::copy(g.begin(), g.end(), grid.begin() + (size_t)w * m.grid_len);
        }
    };
    upload_grid();
    run.check(ldo_set_windows(run.ctx, W, m.windows.data(), grid.data(), m.grid_len));
    run.check(ldo_set_histogram(run.ctx, 1));

    // one simulate() call of every window: outputs under <window>_iter-<tag>,
    // steps counted from 1; a duration limit is checked between launches
    const long long launch_cap = 20000;
    auto phase = [&](const std::string& tag, long long total, double max_dur) {
        std::vector<std::unique_ptr<Outputs>> outs;
        for (int w = 0; w < W; w++) {
            outs.push_back(setup_outputs(run, base + m.postfix[w] + "_iter-" + tag));
        }
        const Outputs& o0 = *outs[0];
        long long chunk = 0;
        for (long long f: {o0.f_cfg, o0.f_vtf, o0.f_counts, o0.f_times, o0.f_ene, o0.f_ops}) {
            if (f > 0) {
                chunk = gcd_of(f, chunk);
            }
        }
        if (chunk == 0) {
            chunk = launch_cap;
        }
        auto start = std::chrono::steady_clock::now();
        auto elapsed = [&]() {
            return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        };
        long long step = 0;
        while (step < total) {
            long long n = std::min({chunk - step % chunk, total - step, launch_cap});
            run.check(ldo_sweep(run.ctx, n));
            step += n;
            double t = elapsed();
            for (int w = 0; w < W; w++) {
                write_outputs(run, *outs[w], w, step, t);
            }
            if (t > max_dur) {
                break;
            }
        }
        return step;
    };
    // one PTMWUS iteration (run_swaps, 658-694): exchange rounds with the
    // outputs of each window taken before the round's swap
    std::vector<int32_t> q_before(W), q_after(W);
    auto swap_phase = [&](const std::string& tag, long long swaps, double max_dur) {
        std::vector<std::unique_ptr<Outputs>> outs;
        for (int w = 0; w < W; w++) {
            outs.push_back(setup_outputs(run, base + m.postfix[w] + "_iter-" + tag));
        }
        std::ofstream swp;
        Outputs::open(swp, base + "_iter-" + tag + ".swp");
        // run_swaps counts swap_i from 1 in every iteration
        // (us_simulation.cpp:662-706)
        run.check(ldo_restart_swaps(run.ctx));
        run.check(ldo_get_permutation(run.ctx, q_before.data()));
        for (int w = 0; w < W; w++) {
            swp << q_before[w] << " ";
        }
        swp << "\n";
        std::vector<int64_t> att0(2 * run.R), acc0(2 * run.R), att1(2 * run.R), acc1(2 * run.R);
        run.check(ldo_get_exchange_counts(run.ctx, 2 * run.R, att0.data(), acc0.data()));
        auto start = std::chrono::steady_clock::now();
        long long m_steps = 0;
        for (long long s = 1; s <= swaps; s++) {
            run.check(ldo_get_permutation(run.ctx, q_before.data()));
            run.check(ldo_run(run.ctx, 1, interval));
            m_steps += interval;
            run.check(ldo_get_permutation(run.ctx, q_after.data()));
            double t =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
            for (int w = 0; w < W; w++) {
                write_outputs(run, *outs[w], q_before[w], m_steps, t, q_after[w]);
            }
            if (m_steps % cfg_freq == 0) {
                for (int w = 0; w < W; w++) {
                    swp << q_before[w] << " ";
                }
                swp << "\n";
            }
            if (t > max_dur) {
                break;
            }
        }
        if (m_steps % cfg_freq == 0) {
            for (int w = 0; w < W; w++) {
                swp << q_after[w] << " ";
            }
            swp << "\n";
        }
        // write_acceptance_freqs (884-898): the iteration's swaps and
        // attempts of each pair of neighbouring windows
        run.check(ldo_get_exchange_counts(run.ctx, 2 * run.R, att1.data(), acc1.data()));
        std::cout << "Iter " << tag << "\n";
        std::cout << "Window 1, Window 2, Swaps, Attempts, Frequency\n";
        for (int i = 0; i + 1 < W; i++) {
            int64_t n_att = att1[i] - att0[i], n_acc = acc1[i] - acc0[i];
            std::cout << i << " " << i + 1 << " " << n_acc << " " << n_att << " "
                      << (static_cast<double>(n_acc) / n_att) << " \n";
        }
        std::cout << "\n";
        std::cout.flush();
        return m_steps;
    };
    // run_equilibration (103-124): its visits are not used; PTMWUS skips it
    // when the biases are read in (PTMWUSGCMCSimulation::run, 646-648)
    if (!(pt && p.flag("read_biases", false))) {
        phase("equil", p.integer("equil_steps", 0), p.num("max_equil_dur", 0));
    }
    std::vector<int64_t> hist((size_t)W * (m.grid_len + 1));
    run.check(ldo_get_histogram(run.ctx, hist.data(), 1));
    long long iters = p.integer("max_num_iters", 0);
    double max_D_bias = p.num("max_D_bias", 0);
    for (long long n = 0; n < iters; n++) {
        std::string tag = std::to_string(n);
        for (int w = 0; w < W; w++) {
            std::ofstream bf(base + m.postfix[w] + "_iter-" + tag + "-inp.biases");
            ldo_us::write_weights(bf, us[w]);
        }
        // the step count the visits are divided by: simulate()'s return
        // value (one past the last step) for MWUS, the summed exchange
        // intervals for PTMWUS
        long long m_steps =
            pt ? swap_phase(tag, p.integer("iter_swaps", 0), p.num("max_iter_dur", 0))
               : phase(tag, p.integer("iter_steps", 0), p.num("max_iter_dur", 0)) + 1;
        run.check(ldo_get_histogram(run.ctx, hist.data(), 1));
        for (int w = 0; w < W; w++) {
            const int64_t* f = hist.data() + (size_t)w * (m.grid_len + 1);
            std::vector<double> g(grid.begin() + (size_t)w * m.grid_len,
                                  grid.begin() + (size_t)(w + 1) * m.grid_len);
            try {
                ldo_us::process_iteration(us[w], f, m.grid_len, g, m_steps, max_D_bias);
            } catch (std::logic_error& e) {
                throw Fail(LDO_E_MISUSE, e.what());
            }
            ldo_us::write_summary(*summary[w], us[w], (int)n);
            summary[w]->flush();
            std::ofstream bf(base + m.postfix[w] + "_iter-" + tag + ".biases");
            ldo_us::write_weights(bf, us[w]);
        }
        upload_grid();
        run.check(ldo_set_window_grid(run.ctx, grid.data()));
    }
    ldo_destroy(run.ctx);
    return 0;
}

int main_run(int argc, char** argv) {
    if (argc < 2 || std::string(argv[1]) == "-h" || std::string(argv[1]) == "--help") {
        std::cout << "usage: ldo_run <parameter file> [key=value ...]\n";
        return argc < 2 ? 2 : 0;
    }
    Run run;
    Params& p = run.p;
    {
        std::stringstream ss(slurp(argv[1]));
        std::string line;
        while (std::getline(ss, line)) {
            parse_assignment(line, p);
        }
        for (int a = 2; a < argc; a++) {
            parse_assignment(argv[a], p);
        }
    }
    // system JSON (OrigamiInputFile, files.cpp:140-218)
    const Json sys = read_json(p.str("origami_input_filename", "")).at("origami");
    for (auto& ch: sys.at("identities").arr) {
        std::vector<int> ids;
        for (auto& v: ch.arr) {
            ids.push_back((int)v.num);
        }
        run.identities.push_back(ids);
    }
    run.n_scaffold = (int)run.identities.at(0).size();
    std::vector<std::string> seq_store;
    std::vector<const char*> seqs;
    if (sys.has("sequences")) {
        for (auto& ch: sys.at("sequences").arr) {
            for (auto& v: ch.arr) {
                seq_store.push_back(v.str);
            }
        }
        for (auto& s: seq_store) {
            seqs.push_back(s.c_str());
        }
    }
    std::vector<int32_t> lens, flat;
    for (auto& ids: run.identities) {
        lens.push_back((int32_t)ids.size());
        flat.insert(flat.end(), ids.begin(), ids.end());
    }
    ldo_system_desc sd {};
    sd.n_chain_types = (int32_t)run.identities.size();
    sd.chain_lengths = lens.data();
    sd.identities = flat.data();
    sd.sequences = seqs.empty() ? nullptr : seqs.data();
    sd.cyclic = sys.has("cyclic") && sys.at("cyclic").b;
    sd.domain_type = enum_of(p.str("domain_type", "Halfturn"), {"HalfTurn", "ThreeQuarterTurn"},
                             "domain_type");
    if (p.str("binding_pot", "FourBody") != "FourBody") {
        throw Fail(LDO_E_NOTIMPL, "binding_pot " + p.str("binding_pot", "") + " is not supported");
    }
    sd.hybridization_pot = enum_of(p.str("hybridization_pot", "NearestNeighbour"),
                                   {"NearestNeighbour", "Uniform", "Specified"},
                                   "hybridization_pot");
    sd.misbinding_pot = enum_of(p.str("misbinding_pot", "Opposing"), {"Opposing", "Disallowed"},
                                "misbinding_pot");
    sd.stacking_pot = enum_of(p.str("stacking_pot", "Constant"),
                              {"Constant", "SequenceSpecific"}, "stacking_pot");
    sd.apply_mean_field_cor = p.flag("apply_mean_field_cor", false);
    sd.binding_h = p.num("binding_h", 0);
    sd.binding_s = p.num("binding_s", 0);
    sd.misbinding_h = p.num("misbinding_h", 0);
    sd.misbinding_s = p.num("misbinding_s", 0);
    sd.stacking_ene = p.num("stacking_ene", 1);
    sd.cation_M = p.num("cation_M", 1);
    sd.staple_M = p.num("staple_M", 1);
    // moveset JSON (setup_movetypes, files.cpp:296-306)
    const Json mv = read_json(p.str("movetype_file", "")).at("origami").at("movetypes");
    std::vector<int32_t> kinds, maxreg, recoils, c_attempts, seg_regrowth;
    std::vector<double> freqs;
    for (auto& m: mv.arr) {
        kinds.push_back(enum_of(m.at("type").str,
                                {"OrientationRotation", "MetStapleExchange", "CBStapleRegrowth",
                                 "CTCBScaffoldRegrowth", "CTRGScaffoldRegrowth",
                                 "CTRGJumpScaffoldRegrowth", "CTCBJumpScaffoldRegrowth"},
                                "movetype"));
        run.move_labels.push_back(m.has("label") ? m.at("label").str : m.at("type").str);
        freqs.push_back(fraction(m.at("freq")));
        maxreg.push_back(m.has("max_regrowth") ? (int32_t)m.at("max_regrowth").num : 0);
        // setup_rg_movetype (simulation.cpp:493-495)
        recoils.push_back(m.has("max_num_recoils") ? (int32_t)m.at("max_num_recoils").num : 0);
        c_attempts.push_back(m.has("max_c_attempts") ? (int32_t)m.at("max_c_attempts").num : 0);
        seg_regrowth.push_back(m.has("max_seg_regrowth") ? (int32_t)m.at("max_seg_regrowth").num
                                                         : 0);
    }
    run.max_total_staples = (int)p.integer("max_total_staples", 999);
    run.max_staple_size = (int)p.integer("max_staple_size", 2);
    ldo_run_params rp {};
    run.move_kinds = kinds;
    rp.n_movetypes = (int32_t)kinds.size();
    rp.movetype_kind = kinds.data();
    rp.movetype_freq = freqs.data();
    rp.movetype_max_regrowth = maxreg.data();
    rp.movetype_max_num_recoils = recoils.data();
    rp.movetype_max_c_attempts = c_attempts.data();
    rp.movetype_max_seg_regrowth = seg_regrowth.data();
    rp.max_total_staples = run.max_total_staples;
    rp.max_type_staples = (int32_t)p.integer("max_type_staples", 999);
    rp.max_staple_size = run.max_staple_size;
    rp.centering_freq = (int32_t)p.integer("centering_freq", 0);
    rp.centering_domain = (int32_t)p.integer("centering_domain", 0);
    rp.constraint_check_freq = (int32_t)p.integer("constraint_check_freq", 0);
    long long seed = p.integer("random_seed", -1);
    if (seed < 0) {
        seed = (long long)std::chrono::steady_clock::now().time_since_epoch().count() & 0x7FFFFFFF;
    }
    rp.seed = (uint64_t)seed;
    rp.flags = (int32_t)p.integer("ldo_flags", 0);
    // initial configuration: the first of the file's configurations
    std::vector<int32_t> c_idx, c_idn, c_len, c_pos, c_ore;
    const Json& cfg0 = sys.at("configurations").arr.at(0);
    for (auto& ch: cfg0.at("chains").arr) {
        c_idx.push_back((int32_t)ch.at("index").num);
        c_idn.push_back((int32_t)ch.at("identity").num);
        c_len.push_back((int32_t)ch.at("positions").arr.size());
        for (auto& v: ch.at("positions").arr) {
            for (auto& x: v.arr) {
                c_pos.push_back((int32_t)x.num);
            }
        }
        for (auto& v: ch.at("orientations").arr) {
            for (auto& x: v.arr) {
                c_ore.push_back((int32_t)x.num);
            }
        }
    }
    ldo_config init {};
    init.n_chains = (int32_t)c_idx.size();
    init.chain_index = c_idx.data();
    init.chain_identity = c_idn.data();
    init.chain_length = c_len.data();
    init.positions = c_pos.data();
    init.orientations = c_ore.data();

    std::string sim = p.str("simulation_type", "constant_temp");
    if (sim == "mw_umbrella_sampling" || sim == "ptmw_umbrella_sampling") {
        return run_mwus(run, sd, rp, init, sim == "ptmw_umbrella_sampling");
    }
    if (sim == "umbrella_sampling") {
        return run_mwus(run, sd, rp, init, false, true);
    }
    // the PT drivers of apps/main.cpp and their exchange modes
    static const std::map<std::string, int> pt_modes = {
        {"t_parallel_tempering", LDO_EXCHANGE_T},
        {"ut_parallel_tempering", LDO_EXCHANGE_TEMPERATURE},
        {"hut_parallel_tempering", LDO_EXCHANGE_HUT},
        {"st_parallel_tempering", LDO_EXCHANGE_ST},
        {"2d_parallel_tempering", LDO_EXCHANGE_2D}};
    auto ptm = pt_modes.find(sim);
    bool pt = ptm != pt_modes.end();
    bool anneal = sim == "annealing";
    if (!pt && !anneal && sim != "constant_temp") {
        throw Fail(LDO_E_NOTIMPL, "simulation_type " + sim + " is not supported");
    }
    // The reference builds SystemBiases from the bias-function file for
    // every simulation type (bias_functions.cpp:304-305) and every move adds
    // them; here biases exist only as the umbrella windows of the MWUS
    // drivers, so a biased constant-T, annealing or PTMC run is refused
    // rather than run unbiased.
    if (!p.str("bias_functions_file", "").empty()) {
        throw Fail(LDO_E_NOTIMPL, "bias_functions_file is supported only with "
                                  "umbrella_sampling, mw_umbrella_sampling and "
                                  "ptmw_umbrella_sampling");
    }
    // control quantities of every ladder slot: temperature, chemical
    // potential, bias and stacking multipliers (m_control_qs)
    std::vector<double> lad_t, lad_u, lad_b, lad_s;
    std::vector<double> v1s, v2s; // the 2D grid's temperatures and stacking multipliers
    int mode = pt ? ptm->second : LDO_EXCHANGE_NONE;
    double max_temp = 0, min_temp = 0, temp_interval = 1;
    long long steps_per_temp = 0;
    if (mode == LDO_EXCHANGE_2D) {
        // TwoDPTGCMCSimulation::initialize_control_qs (ptmc_simulation.cpp:452-493)
        v1s = p.list("temps");
        v2s = p.list("stacking_mults");
        run.R = (int)p.integer("num_reps", (long long)(v1s.size() * v2s.size()));
        if (v1s.empty() || v2s.empty() || (size_t)run.R != v1s.size() * v2s.size()) {
            throw Fail(LDO_E_ARG, "num_reps must equal the temps x stacking_mults grid");
        }
        for (double t: v1s) {
            for (double sm: v2s) {
                lad_t.push_back(t);
                lad_u.push_back(p.num("staple_u_mult", 1));
                lad_b.push_back(1.0);
                lad_s.push_back(sm);
            }
        }
        run.temps = lad_t;
    }
    else if (pt) {
        // OneDPTGCMCSimulation::initialize_control_qs (337-358); a missing
        // multiplier list reads as ones (the reference indexes the empty list)
        run.temps = p.list("temps");
        run.R = (int)p.integer("num_reps", (long long)run.temps.size());
        if ((int)run.temps.size() != run.R) {
            throw Fail(LDO_E_ARG, "temps must list num_reps temperatures");
        }
        lad_t = run.temps;
        std::vector<double>* mults[3] = {&lad_u, &lad_b, &lad_s};
        const char* keys[3] = {"chem_pot_mults", "bias_mults", "stacking_mults"};
        for (int k = 0; k < 3; k++) {
            *mults[k] = p.list(keys[k]);
            if (mults[k]->empty()) {
                mults[k]->assign(run.R, 1.0);
            }
            if ((int)mults[k]->size() != run.R) {
                throw Fail(LDO_E_ARG, std::string(keys[k]) + " must list num_reps values");
            }
        }
    }
    else if (anneal) {
        // AnnealingGCMCSimulation (annealing_simulation.cpp:12-36)
        max_temp = p.num("max_temp", 0);
        min_temp = p.num("min_temp", 0);
        temp_interval = p.num("temp_interval", 0);
        steps_per_temp = p.integer("steps_per_temp", 0);
        if (std::fmod(max_temp - min_temp, temp_interval) != 0) {
            throw Fail(LDO_E_ARG, "Bad temperature interval");
        }
        run.temps = {max_temp};
    }
    else {
        run.temps = {p.num("temp", 300)};
    }
    rp.exchange_mode = mode;
    rp.rng_mode = LDO_RNG_PHILOX;
    int rc = ldo_create(&sd, &rp, &init, run.R, run.temps.data(), &run.ctx);
    if (rc != LDO_OK) {
        throw Fail(rc, ldo_last_error(nullptr));
    }
    if (mode == LDO_EXCHANGE_2D) {
        run.check(ldo_set_ladder_2d(run.ctx, (int32_t)v1s.size(), v1s.data(),
                                    (int32_t)v2s.size(), v2s.data(), p.num("staple_u_mult", 1)));
    }
    else if (pt) {
        run.check(ldo_set_ladder(run.ctx, run.R, lad_t.data(), lad_u.data(), lad_b.data(),
                                 lad_s.data()));
    }

    // restarts of parallel-tempering runs (ptmc_simulation.cpp:38-84): each
    // replica's configuration from frame restart_step of its own trajectory,
    // and the slot permutation from the last line of the swap file
    if (pt && p.flag("restart_from_config", false)) {
        std::string fb = p.str("restart_traj_filebase", "");
        std::string post = p.str("restart_traj_postfix", ".trj");
        for (int r = 0; r < run.R; r++) {
            set_config_from(run, r, read_traj_frame(fb + "-" + std::to_string(r) + post,
                                                    (int)p.integer("restart_step", 0)));
        }
    }
    if (pt && p.flag("restart_from_swap", false)) {
        std::string path = p.str("restart_swap_file", "");
        std::ifstream f(path);
        if (!f.good()) {
            throw Fail(LDO_E_IO, "Restart swap file " + path + " does not exist");
        }
        std::string line, last;
        while (std::getline(f, line)) {
            if (!line.empty()) {
                last = line;
            }
        }
        std::stringstream ss(last);
        std::vector<int32_t> q_to_rep(run.R);
        for (int q = 0; q < run.R; q++) {
            if (!(ss >> q_to_rep[q])) {
                throw Fail(LDO_E_IO, path + ": last line is not a permutation");
            }
        }
        run.check(ldo_set_permutation(run.ctx, q_to_rep.data()));
    }
    std::string base = p.str("output_filebase", "");
    std::vector<std::unique_ptr<Outputs>> outs;
    for (int r = 0; r < run.R; r++) {
        outs.push_back(setup_outputs(run, pt ? base + "-" + std::to_string(r) : base));
    }
    const Outputs& o0 = *outs[0];
    long long log_freq = p.integer("logging_freq", 0);
    long long chunk = 0;
    for (long long f: {o0.f_cfg, o0.f_vtf, o0.f_counts, o0.f_times, o0.f_ene, o0.f_ops, log_freq}) {
        if (f > 0) {
            chunk = gcd_of(f, chunk);
        }
    }
    auto start = std::chrono::steady_clock::now();
    auto elapsed = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    };
    // steps start + 1 .. start + total of one replica run in chunks that end
    // on every step an output or log entry falls on (simulation.cpp:568-653)
    auto simulate = [&](long long start_step, long long total) {
        long long step = start_step;
        while (step < start_step + total) {
            long long n = start_step + total - step;
            if (chunk > 0) {
                n = std::min(n, chunk - step % chunk);
            }
            run.check(ldo_sweep(run.ctx, n));
            step += n;
            // the constant-T and annealing drivers log to standard output
            // (constant_temp_simulation.cpp:19, annealing_simulation.cpp:27)
            if (log_freq != 0 && step % log_freq == 0) {
                write_log_entry(run, std::cout, 0, step);
            }
            write_outputs(run, *outs[0], 0, step, elapsed());
        }
    };
    if (!pt && !anneal) {
        simulate(0, p.integer("ct_steps", 0));
        // simulate(ct_steps) ends with the run summary and the .moves file
        write_log_summary(run, std::cout, base + ".moves");
        std::cout.flush();
    }
    else if (anneal) {
        // AnnealingGCMCSimulation::run (annealing_simulation.cpp:38-50): every
        // temperature from max_temp down to min_temp runs steps_per_temp
        // steps and writes the run summary; simulate returns its last step
        // index + 1, which run adds to its own count, so temperature k starts
        // at step (2^k - 1)(steps_per_temp + 1)
        long long step = 0;
        for (double temp = max_temp; temp >= min_temp; temp -= temp_interval) {
            double one = 1.0;
            run.check(ldo_set_ladder(run.ctx, 1, &temp, &one, &one, &one));
            // update_temp recomputes the energy (origami_system.cpp:618-622)
            run.check(ldo_run(run.ctx, 1, 0));
            run.check(ldo_set_step(run.ctx, step));
            simulate(step, steps_per_temp);
            write_log_summary(run, std::cout, base + ".moves");
            step += step + steps_per_temp + 1;
        }
        std::cout.flush();
    }
    else {
        long long swaps = p.integer("swaps", 0);
        long long interval = p.integer("exchange_interval", 1);
        // each ldo_run interval starts by switching every replica to its
        // ladder slot, so an output may not fall inside an interval
        if (interval < 1 || (chunk > 0 && chunk % interval != 0)) {
            throw Fail(LDO_E_NOTIMPL, "output frequencies must be multiples of exchange_interval");
        }
        // PTGCMCSimulation::initialize_swap_file (ptmc_simulation.cpp:91-104):
        // each slot's exchanged control quantities (m_exchange_q_is)
        std::vector<int> q_is = {0};
        if (mode == LDO_EXCHANGE_TEMPERATURE || mode == LDO_EXCHANGE_HUT ||
            mode == LDO_EXCHANGE_2D) {
            q_is.push_back(1);
        }
        if (mode == LDO_EXCHANGE_HUT) {
            q_is.push_back(2);
        }
        if (mode == LDO_EXCHANGE_ST || mode == LDO_EXCHANGE_2D) {
            q_is.push_back(3);
        }
        const std::vector<double>* qs[4] = {&lad_t, &lad_u, &lad_b, &lad_s};
        std::ofstream swp;
        Outputs::open(swp, base + ".swp");
        for (int r = 0; r < run.R; r++) {
            for (int i: q_is) {
                swp << (*qs[i])[r] << "/";
            }
            swp << " ";
        }
        swp << "\n";
        long long cfg_freq = o0.f_cfg;
        std::vector<int32_t> q_to_rep(run.R);
        auto swap_entry = [&]() {
            for (int q = 0; q < run.R; q++) {
                swp << q_to_rep[q] << " ";
            }
            swp << "\n";
            swp.flush();
        };
        // each rank logs to its own .out file (ptmc_simulation.cpp:48-56);
        // simulate runs without a summary between exchanges
        std::vector<std::unique_ptr<std::ofstream>> logs;
        for (int r = 0; r < run.R; r++) {
            logs.push_back(std::make_unique<std::ofstream>());
            Outputs::open(*logs.back(), base + "-" + std::to_string(r) + ".out");
        }
        long long step = 0;
        for (long long s = 1; s <= swaps; s++) {
            // write_swap_entry runs before the round's exchange
            // (ptmc_simulation.cpp:140-147): it records the permutation the
            // round's interval ran under
            bool entry = cfg_freq != 0 && (step + interval) % cfg_freq == 0;
            if (entry) {
                run.check(ldo_get_permutation(run.ctx, q_to_rep.data()));
            }
            run.check(ldo_run(run.ctx, 1, interval));
            step += interval;
            for (int r = 0; r < run.R; r++) {
                if (log_freq != 0 && step % log_freq == 0) {
                    write_log_entry(run, *logs[r], r, step);
                }
                write_outputs(run, *outs[r], r, step, elapsed());
            }
            if (entry) {
                swap_entry();
            }
        }
        // the closing entry and the acceptance frequencies on standard
        // output (ptmc_simulation.cpp:143-149, 414-426, 562-592)
        if (cfg_freq != 0 && step % cfg_freq == 0) {
            run.check(ldo_get_permutation(run.ctx, q_to_rep.data()));
            swap_entry();
        }
        std::vector<int64_t> att(2 * run.R), acc(2 * run.R);
        run.check(ldo_get_exchange_counts(run.ctx, 2 * run.R, att.data(), acc.data()));
        auto freq_line = [&](double a, double b, double c, bool third, int64_t n_acc,
                             int64_t n_att) {
            std::cout << a << " " << b << " ";
            if (third) {
                std::cout << c << " ";
            }
            std::cout << n_acc << " " << n_att << " " << (static_cast<double>(n_acc) / n_att)
                      << " \n";
        };
        if (mode == LDO_EXCHANGE_2D) {
            int v1 = (int)v1s.size(), v2 = (int)v2s.size();
            for (int i = 0; i + 1 < v1; i++) {
                for (int j = 0; j < v2; j++) {
                    int q = i * v2 + j;
                    freq_line(v1s[i], v1s[i + 1], v2s[j], true, acc[q], att[q]);
                }
            }
            std::cout << "\n";
            for (int j = 0; j + 1 < v2; j++) {
                for (int i = 0; i < v1; i++) {
                    int q = run.R + i * v2 + j;
                    freq_line(v2s[j], v2s[j + 1], v1s[i], true, acc[q], att[q]);
                }
            }
            std::cout << "\n";
        }
        else {
            for (int i = 0; i + 1 < run.R; i++) {
                freq_line(lad_t[i], lad_t[i + 1], 0, false, acc[i], att[i]);
            }
            std::cout << "\n";
        }
        std::cout.flush();
    }
    ldo_destroy(run.ctx);
    return 0;
}

} // namespace

int main(int argc, char** argv) {
    try {
        return main_run(argc, argv);
    } catch (Fail& e) {
        std::cerr << "ldo_run: " << e.what() << " (status " << e.code << ")\n";
        return 1;
    } catch (std::exception& e) {
        std::cerr << "ldo_run: " << e.what() << "\n";
        return 1;
    }
}
