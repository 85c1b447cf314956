// oracle_core.hpp - CPU restatement of the LatticeDNAOrigami hot path.
//
// TEST INFRASTRUCTURE ONLY (see ldo_oracle.h).  Each piece cites the
// reference function it follows.  The structure deliberately mirrors the
// reference (pointer-linked domains, hash maps keyed by lattice site,
// std::set iteration orders) so that it reproduces the reference's
// arithmetic and random-draw order, which the HIP sampler is then checked
// against draw for draw.
//
// Parity unpinned (no golden vector, known answer or statistic fixes them;
// the HIP sampler is checked against this restatement only):
//   - the values of the FourBody junction terms in networks of several
//     crossovers (Potential::fb_*_single_junction, origami_potential.cpp:
//     654-934); which neighbours those walks read is pinned by the energy
//     being a state function of folded networks (order and rotation
//     invariance), single crossovers by hand-built configurations;
//   - CTCB endpoint bookkeeping for networks of several staples bound to
//     each other (cb_movetypes.cpp);
//   - the mt19937_64 draws of the tape mode, which are libstdc++'s;
//   - CTRG's feeler growth (rg_movetypes.cpp:426-480), pinned only loosely
//     by the enumeration of the bare scaffold.
#ifndef ORACLE_CORE_HPP
#define ORACLE_CORE_HPP

#include <cstdint>
#include <array>
#include <deque>
#include <map>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../include/ldo.h"

namespace orc {

// ---------------------------------------------------------------- vectors
// utility::VectorThree (utility.hpp:66-95, utility.cpp:26-161)
struct V3 {
    int x {0}, y {0}, z {0};
    V3() = default;
    V3(int a, int b, int c): x {a}, y {b}, z {c} {}
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator-() const { return {-x, -y, -z}; }
    bool operator==(const V3& o) const { return x == o.x && y == o.y && z == o.z; }
    bool operator!=(const V3& o) const { return !(*this == o); }
    int& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    int get(int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct V3Hash {
    size_t operator()(const V3& v) const {
        uint64_t h = (uint64_t)(uint32_t)v.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)v.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)v.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

// The six unit vectors in the reference's order (utility.hpp:156-162)
extern const V3 kUnit[6];

V3 rotate_half(V3 v, V3 axis);           // utility.cpp:68-85
V3 rotate_turns(V3 v, V3 axis, int turns); // utility.cpp:102-142

// ----------------------------------------------------------------- domains
enum class Occ { unassigned = 0, unbound = 1, bound = 2, misbound = 3 };

// domainContainer::Domain (domain.hpp:14-52)
struct Dom {
    int c;       // unique chain index
    int c_ident; // chain identity
    int d;       // index within its chain
    int d_ident; // domain identity
    int c_len;
    Dom* fwd {nullptr};
    Dom* bwd {nullptr};
    V3 pos {};
    V3 ore {};
    Occ st {Occ::unassigned};
    Dom* bd {nullptr};
};

// Domain::operator+ (domain.cpp:9-31)
Dom* walk(Dom* d, int incr);

// --------------------------------------------------------------------- rng
// RandomGens (random_gens.cpp:12-49) or the Philox stream of the sampler.
class Rng {
  public:
    void init(int kind, uint64_t seed, uint32_t stream);
    void begin_step(int64_t step) {
        m_step = step;
        m_sub = 0;
        m_idx = 0;
    }
    // sub-stream `sub` (1: learned-move attempts) at counter `ctr`
    void begin_sub(uint32_t sub, int64_t ctr) {
        m_step = ctr;
        m_sub = sub;
        m_idx = 0;
    }
    double uniform_real();
    int uniform_int(int lo, int hi);
    void set_record(bool on) { m_record = on; }
    // replay recorded draws instead of generating them (doubles as IEEE-754
    // bit patterns, integers as integers, in draw order)
    void set_replay(const uint64_t* v, int64_t n) {
        m_replay.assign(v, v + n);
        m_replay_pos = 0;
        m_kind = 2;
    }
    std::vector<uint64_t> tape;

  private:
    uint64_t philox64();
    uint64_t replay_next();
    std::vector<uint64_t> m_replay;
    size_t m_replay_pos {0};
    int m_kind {0};
    uint64_t m_seed {0};
    uint32_t m_stream {0};
    int64_t m_step {0};
    uint32_t m_sub {0};
    uint32_t m_idx {0};
    bool m_record {false};
    std::mt19937_64 m_mt;
};

void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

// ------------------------------------------------------------------- energy
struct Delta {
    double e {0};
    int stacked {0};
};

class System;

// Energy tables keyed by domain identity pairs (origami_potential.cpp:1057-1221)
struct Tables {
    std::vector<double> hyb_e, hyb_h, hyb_s, stack;
    double init_h {0}, init_s {0}, init_e {0};
};

class Potential {
  public:
    void setup(const ldo_system_desc& sys, const std::vector<std::vector<int>>& idents,
               const std::vector<std::vector<std::string>>& seqs);
    void update_temp(double temp, double stacking_mult); // :1019-1042
    void set_uncached(double temp) { m_cur = build(temp, 1.0); }
    double hyb_e(const Dom& a, const Dom& b) const { return m_cur.hyb_e[key(a, b)]; }
    double hyb_h(const Dom& a, const Dom& b) const { return m_cur.hyb_h[key(a, b)]; }
    double hyb_s(const Dom& a, const Dom& b) const { return m_cur.hyb_s[key(a, b)]; }
    double stack(const Dom& a, const Dom& b) const { return m_cur.stack[key(a, b)]; }
    double init_e() const { return m_cur.init_e; }
    double init_h() const { return m_cur.init_h; }
    double init_s() const { return m_cur.init_s; }
    Tables build(double temp, double stacking_mult) const; // calc_energies
    int ident_lo {0};
    int n_ident {0};
    int domain_type {0};
    int misbinding {0};
    bool violated {false};

    Delta bind_domain(Dom& di);                 // :1282-1312
    Delta check_stacking(Dom& di, Dom& dj);     // :1314-1317

  private:
    size_t key(const Dom& a, const Dom& b) const {
        return (size_t)(a.d_ident - ident_lo) * n_ident + (b.d_ident - ident_lo);
    }
    ldo_system_desc m_sys {};
    std::vector<std::vector<int>> m_idents;
    std::vector<std::vector<std::string>> m_seqs;
    std::vector<double> m_spec_h, m_spec_s;
    Tables m_cur;
    std::map<std::pair<double, double>, Tables> m_cache;

    // JunctionBindingPotential (origami_potential.cpp:131-934)
    Delta m_dc;
    bool m_fb_violated {false};
    Delta fb_bind(Dom& di, Dom& dj);
    Delta fb_check(Dom& di, Dom& dj);
    void fb_terms(Dom& di, Dom& dj);
    void fb_around(Dom* cd, int j);
    void fb_regular_pair(Dom* a, Dom* b, int i);
    void fb_helix_pair(Dom* a, Dom* b, int i, int j);
    void fb_junction_pair(Dom* a, Dom* b, int j);
    void fb_junction(Dom* j1, Dom* j2, Dom* j3, Dom* j4, Dom* k1, Dom* k2);
    void fb_backward_triplets(Dom* a, Dom* b);
    void fb_forward_triplets(Dom* a, Dom* b);
    void fb_central_triplets(Dom& di, Dom& dj);
    void fb_backward_single_junction(Dom* a, Dom* b);
    void fb_forward_single_junction(Dom* a, Dom* b);
    void fb_central_single_junction(Dom* a, Dom* b);
    void fb_single(Dom* h1, Dom* h2, Dom* h3);
    void fb_double(Dom* h1, Dom* h2, Dom* h3);
    void fb_triply(Dom* h1, Dom* h2, Dom* h3);

  public:
    bool pair_stacked(Dom* a, Dom* b) const;
    bool twist_ok(const Dom& a, V3 ndr, const Dom& b) const;
    bool kink_ok(const Dom& a, V3 ndr, const Dom& b) const;
    bool junction_ok(const Dom& j1, const Dom& j2, const Dom& j3, const Dom& j4,
                     const Dom& k1, const Dom& k2) const;
};

// ------------------------------------------------------------------ system
struct ChainSpec {
    int index;
    int identity;
    std::vector<V3> pos;
    std::vector<V3> ore;
};

class OrigamiError : public std::runtime_error {
  public:
    OrigamiError(int code, const std::string& m): std::runtime_error(m), code {code} {}
    int code;
};

// OrigamiSystem (origami_system.hpp:66-225)
class System {
  public:
    System(const ldo_system_desc& sys, const std::vector<std::vector<int>>& idents,
           const std::vector<std::vector<std::string>>& seqs,
           const std::vector<ChainSpec>& chains, double temp);
    ~System();
    System(const System&) = delete;
    System& operator=(const System&) = delete;

    std::vector<std::vector<int>> identities;
    bool cyclic;
    double temp;
    double staple_M;
    std::vector<double> staple_us;
    std::vector<double> reduced_staple_us;
    double reduced_fugacity;
    Potential pot;

    std::vector<std::vector<Dom*>> chains;     // m_domains (working order)
    std::vector<int> chain_indices;            // working -> unique
    std::vector<int> chain_identities;         // working -> identity
    std::vector<std::vector<int>> ident_to_index;
    std::unordered_map<V3, Dom*, V3Hash> pos_to_unbound;
    std::unordered_map<V3, Occ, V3Hash> pos_occ;
    int num_domains {0};
    int num_staples {0};
    int num_bound_pairs {0};
    int num_fully_bound_pairs {0};
    int num_self_bound_pairs {0};
    int num_unassigned {0};
    int num_stacked {0};
    double energy {0};
    double hyb_enthalpy {0};
    double hyb_entropy {0};
    double stacking_energy {0};
    bool violated {false};
    int current_c_i {0};
    // counted events of the throughput byte model: trial placements onto an
    // empty site, binding/stacking evaluations, committed sets and unsets
    int64_t ev_site_empty {0};
    int64_t ev_site_bind {0};
    int64_t ev_commits {0};

    Occ occupancy(V3 pos) const;
    Dom* unbound_at(V3 pos) const { return pos_to_unbound.at(pos); }
    std::vector<Dom*>& chain_by_index(int c_i);
    Dom* domain(int c_i, int d_i) { return chain_by_index(c_i)[d_i]; }
    std::vector<int> staple_counts() const;

    double check_domain_constraints(Dom& di, V3 pos, V3 ore); // :343-355
    double unassign_domain(Dom& di);                          // :375-385
    int add_chain(int ident);                                 // :387-400
    int add_chain(int ident, int c_i);                        // :402-443
    void delete_chain(int c_i);                               // :445-472
    double set_checked_domain_config(Dom& di, V3 pos, V3 ore);// :478-515
    double set_domain_config(Dom& di, V3 pos, V3 ore);        // :517-541
    void set_domain_orientation(Dom& di, V3 ore);             // :543-551
    void center(int centering_domain);                        // :553-571
    void set_all_domains();                                   // :573-586
    void set_all_domains(const std::vector<ChainSpec>& cfg);  // :588-616
    void set_config(const std::vector<ChainSpec>& cfg);       // :327-341
    void check_all_constraints();                             // :267-325
    void check_distance_constraints();                        // :357-373
    void update_temp(double t, double stacking_mult);         // :618-622
    void update_staple_us(double t, double mult);             // :624-628
    void update_energy();                                     // :808-826
    void update_enthalpy_and_entropy();                       // :204-246
    std::vector<ChainSpec> snapshot() const;                  // :175-191

  private:
    void initialize_scaffold(const ChainSpec& c);
    void initialize_staples(const std::vector<ChainSpec>& c);
    Delta internal_unassign(Dom& di);                    // :695-739
    double unassign_bound(Dom& di);                      // :741-753
    void unassign_unbound(Dom& di);                      // :755-760
    void update_occupancies(Dom& di, V3 pos);            // :770-806
    Delta internal_check(Dom& di, V3 pos, V3 ore);       // :828-871
};

// -------------------------------------------------------------- movetypes
struct MovesetEntry {
    int kind;
    double freq;
    int max_regrowth;
    int max_num_recoils {0}; // CTRG (simulation.cpp:493-495)
    int max_c_attempts {0};
    int max_seg_regrowth {0}; // CTRGJump
};

struct Limits {
    int max_total_staples;
    int max_type_staples;
    int flags {0}; // ldo_run_flags
};

// Move-level biases of one umbrella window: a LinearStepWell restraint
// (bias_functions.cpp:139-151) as the first bias level and a Grid bias
// (242-284) as the second, on one system-wide order parameter, multiplied by
// the window's bias multiplier (SystemBiases::calc_move, 484-496).
struct WindowBias {
    int op {0}; // LDO_OP_*
    int min_op {0}, max_op {0};
    double well_bias {0}, min_bias {0}, slope {0};
    double off_grid {0}, mult {1};
    std::vector<double> grid;
    double restraint(int p) const; // LinearStepWellBiasFunction::calc_bias
    double grid_bias(int p) const; // GridBiasFunction::calc_bias
};

// the stored values of a window's bias functions (BiasFunction::m_bias)
struct WindowState {
    double restraint {0};
    double grid {0};
};

class Replica; // forward

// Constraintpoints + StapleNetwork (top_constraint_points.cpp:14-601)
class Constraintpoints {
  public:
    explicit Constraintpoints(System& s): m_sys {s} {}
    // LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS: endpoints of externally bound
    // networks go on the in-segment scaffold domains (off: the reference)
    bool scaffold_endpoints {false};
    void reset();
    void calculate(const std::vector<Dom*>& scaffold_domains, int dir);
    void calculate(const std::vector<std::vector<Dom*>>& segments, const std::vector<int>& dirs);
    void add_growthpoint(Dom* growthpoint, Dom* stemd) {
        m_growthpoints[growthpoint] = stemd;
        m_stemdomains[stemd] = growthpoint;
    }
    void add_stem_seg_pair(Dom* stemd, std::vector<int> segs) { m_stemd_to_segs[stemd] = segs; }
    std::set<int> staples_to_be_regrown() const { return m_regrowth_staples; }
    bool is_growthpoint(Dom* d) const { return m_growthpoints.count(d) > 0; }
    bool is_stemdomain(Dom* d) const { return m_stemdomains.count(d) > 0; }
    void add_active_endpoint(Dom* d, V3 pos);
    void add_active_endpoint(Dom* d, V3 pos, int seg);
    void reset_active_endpoints() { m_active = m_initial_active; }
    void remove_active_endpoint(Dom* d);
    void remove_activated_endpoint(Dom* d);
    void update_endpoints(Dom* d);
    const std::vector<V3>& get_erased_endpoints() const { return m_erased; }
    const std::vector<Dom*>& domains_to_be_regrown() const { return m_d_stack; }
    Dom* get_growthpoint(Dom* d) { return m_stemdomains[d]; }
    Dom* domain_to_grow(Dom* d) { return m_growthpoints[d]; }
    bool endpoint_reached(Dom* d, V3 pos);
    bool walks_remain(Dom* d, V3 pos);
    bool walks_remain_seg(Dom* d, V3 pos, int seg);
    int get_dir(Dom* d);

  private:
    void find_growthpoints_endpoints(const std::vector<Dom*>& sds, int seg);
    // StapleNetwork state (top_constraint_points.hpp:73-86)
    void scan_network(Dom* d);
    void scan_staple_topology(Dom* growth_d);
    std::vector<Dom*> make_staple_stack(Dom* d, int ci);
    System& m_sys;
    std::vector<Dom*> m_net_scaffold_ds;
    bool m_net_external {false};
    std::set<int> m_net_cs;
    std::vector<std::pair<Dom*, Dom*>> m_net_gps;
    std::vector<Dom*> m_net_ds;
    std::vector<std::pair<Dom*, Dom*>> m_net_iaes;
    std::unordered_map<Dom*, int> m_net_segs;
    std::map<std::pair<int, int>, int> m_net_dirs;

    std::set<int> m_regrowth_staples;
    std::vector<Dom*> m_d_stack;
    std::set<int> m_checked_staples;
    std::unordered_map<Dom*, int> m_segs;
    std::map<std::pair<int, int>, int> m_dirs;
    std::unordered_map<Dom*, Dom*> m_growthpoints;
    std::unordered_map<Dom*, Dom*> m_stemdomains;
    std::map<std::pair<int, int>, std::vector<std::pair<int, V3>>> m_active;
    std::map<std::pair<int, int>, std::vector<std::pair<int, V3>>> m_initial_active;
    std::unordered_map<Dom*, Dom*> m_inactive;
    std::vector<V3> m_erased; // positions the last remove_active_endpoint took
    std::unordered_map<Dom*, std::vector<int>> m_stemd_to_segs;
};

// One replica: system + moves + RNG (the per-rank GCMCSimulation state)
class Replica {
  public:
    Replica(const ldo_system_desc& sys, const std::vector<std::vector<int>>& idents,
            const std::vector<std::vector<std::string>>& seqs,
            const std::vector<ChainSpec>& chains, double temp,
            const std::vector<MovesetEntry>& moves, const Limits& lim,
            int centering_freq, int centering_domain, int constraint_check_freq,
            int rng_kind, uint64_t seed, uint32_t stream);
    System sys;
    Rng rng;
    int64_t step {0};
    int status {0};
    std::vector<int64_t> attempts, accepts;
    // the movetypes' m_tracker fields of the move in flight, the tracked
    // counts in the layout of ldo_get_move_tracking, and the last step's
    // move slot and outcome
    struct Tracker {
        int staple_type {0};
        bool insertion {false};
        bool no_staples {false};
        int num_scaffold_domains {0};
        int num_staples {0};
    } tracker;
    std::vector<int64_t> track;
    int last_move {0};
    int last_accepted {0};
    // the umbrella window this replica samples in, and that window's stored
    // bias values (both null without biases)
    const WindowBias* win {nullptr};
    WindowState* win_state {nullptr};
    // the window's visit counts ([grid size + 1], the last bin off the grid),
    // or null when not counted (USGCMCSimulation::update_internal)
    int64_t* hist {nullptr};

    void simulate(int64_t steps); // simulation.cpp:568-653
    void add_tracker(int slot, bool accepted);
    int window_op(int staple_offset) const;
    double calc_move_bias(int staple_offset);

    // learned domain regrowth (no reference counterpart; the sampler's
    // ldo_learned_*): network parameters in the ldo_mlp.h layout
    const float* net {nullptr};
    int64_t net_off[10] {};
    int net_n_ident {0};
    double learned_fallback {0};
    int64_t learned_step {0};
    void learned_attempt();
    void learned_exchange();
    Dom* complement_of(Dom* s) const;
    void learned_weights(Dom* d, Dom* a, int id, double w[36], double& z, int& k_of_d) const;
    // the network input around the chain neighbour of staple end d (d left
    // out); returns the output d occupies
    int learned_context(Dom* d, Dom* a, uint8_t ch[343], uint8_t legal[6]) const;
    void learned_logits(const uint8_t ch[343], int id, float logits[36]) const;

  private:
    std::vector<MovesetEntry> m_moves;
    std::vector<double> m_cum;
    Limits m_lim;
    int m_centering_freq, m_centering_domain, m_check_freq;

    // shared move state (MCMovetype / RegrowthMCMovetype / CBMCMovetype)
    std::vector<std::pair<int, int>> m_modified, m_assigned;
    std::vector<int> m_added;
    std::map<std::pair<int, int>, V3> m_prev_pos, m_prev_ore, m_old_pos, m_old_ore;
    bool m_rejected {false};
    double m_modifier {1};
    double m_delta_e {0};
    long double m_bias {1}, m_new_bias {1};
    double m_new_modifier {1};
    bool m_regrow_old {false};
    int m_insertion_sites {0};
    int m_kind {0};
    int m_dir {0};
    Constraintpoints m_cp;

    bool attempt(int kind, int max_regrowth);
    void reset_origami();
    Dom* select_random_domain();
    bool test_acceptance(long double p);
    bool staple_is_connector(const std::vector<Dom*>& staple);
    bool scan_for_scaffold_domain(Dom* d, std::set<int>& participating);
    std::vector<std::pair<Dom*, Dom*>> find_bound_domains(const std::vector<Dom*>& ch);
    int num_bound_staple_domains(const std::vector<Dom*>& staple) const;
    double set_growth_point(Dom& gnew, Dom& gold);
    void grow_staple(int d_i, const std::vector<Dom*>& chain);
    void grow_chain(const std::vector<Dom*>& ds);

    bool orientation_move();
    bool staple_exchange_move();
    bool insert_staple();
    bool delete_staple();
    void met_grow_chain(const std::vector<Dom*>& ds);
    void met_unassign_domains(const std::vector<Dom*>& ds);

    bool cb_staple_move();
    void cb_calc_biases(V3 p_prev, Dom& d, std::vector<std::pair<V3, V3>>& cfgs,
                        std::vector<double>& bf);
    std::vector<double> cb_calc_bias(const std::vector<double>& bf,
                                     const std::vector<std::pair<V3, V3>>& cfgs, Dom* d);
    void cb_select_and_set(size_t i, const std::vector<Dom*>& ds);
    void cb_unassign_domains(const std::vector<Dom*>& ds);
    void cb_setup_regrow_old();
    bool cb_test_acceptance();
    double cb_set_old_growth_point(Dom& gnew, Dom& gold);
    void cb_add_external_bias();

    bool ctcb_scaffold_move(int max_regrowth);
    std::vector<Dom*> select_indices(const std::vector<Dom*>& seg, unsigned min_len,
                                     unsigned max_regrowth);
    void ctcb_grow_chain(const std::vector<Dom*>& ds);
    void ctcb_grow_staple_and_update_endpoints(Dom* gold);

    // CTRGRegrowthMCMovetype (rg_movetypes.hpp): the regrowth stack, per
    // level the trial attempts made and configuration indices left (growth
    // and weight copies), the open probability and the erased endpoints, and
    // the level being grown
    using Cfg = std::pair<V3, V3>;
    std::vector<Dom*> m_regrow_ds;
    std::vector<int> m_c_attempts_q, m_c_attempts_wq;
    std::vector<std::vector<int>> m_avail_cis_q, m_avail_cis_wq;
    std::vector<double> m_c_opens;
    std::vector<std::vector<V3>> m_erased_q;
    std::map<std::pair<int, int>, V3> m_new_pos, m_new_ore;
    size_t m_di {0};
    Dom* m_d {nullptr};
    Dom* m_ref_d {nullptr};
    bool m_stemd {false};
    int m_c_attempts {0}, m_d_max_c_attempts {0};
    std::vector<int> m_avail_cis;
    double m_weight {1}, m_weight_new {1};
    int m_max_recoils {0}, m_max_c_attempts {0};
    bool ctrg_scaffold_move(int max_regrowth);
    // the non-contiguous selection (movetypes.cpp:512-648, 666-721)
    int m_max_seg_regrowth {0};
    void select_noncontig_segs(std::vector<std::vector<Dom*>>& segs,
                               std::vector<std::vector<std::vector<Dom*>>>& paired_segs,
                               std::vector<Dom*>& seg_stems, std::vector<int>& dirs,
                               int max_regrowth);
    bool fill_seg(Dom* start_d, size_t max_length, size_t seg_max_length, int dir,
                  std::set<Dom*>& domains, std::deque<Dom*>& possible_stems,
                  std::vector<Dom*>& seg);
    void check_for_stemds(Dom* cur_d, std::deque<Dom*>& possible_stems);
    bool ctrg_jump_move(int max_regrowth);
    void ctcb_set_first_seg_domain(Dom* d_new);
    void ctcb_jump_grow(const std::vector<std::vector<Dom*>>& segs,
                        const std::vector<std::vector<std::vector<Dom*>>>& paired_segs,
                        const std::vector<Dom*>& seg_stems);
    bool ctcb_jump_move(int max_regrowth);
    double rg_unassign_and_save();
    void rg_unassign_domains();
    double rg_recoil_regrow();
    double rg_set_config(Dom* d, Cfg c);
    void rg_prepare_for_growth();
    double rg_prepare_for_regrowth();
    void rg_restore_endpoints();
    Cfg rg_select_trial_config();
    double rg_calc_p_config_open(Cfg c);
    bool rg_test_config_open(double p);
    void rg_calc_weights();
    bool rg_test_config_avail();
    void rg_calc_old_c_opens();
    bool rg_test_acceptance();
};

// ------------------------------------------------------------ enumeration
// Exact enumeration of a small system's grand-canonical ensemble
// (oracle_enum.cpp).  Weights are keyed by (numstaples, numfulldomains,
// nummisdomains, numstackedpairs).
struct EnumResult {
    std::map<std::array<int, 4>, long double> weights;
    long double z {0};
    long double n_configs {0};
};

EnumResult enumerate_states(const ldo_system_desc& desc, const std::vector<std::vector<int>>& idents,
                            const std::vector<std::vector<std::string>>& seqs, const Limits& lim,
                            double temp, bool all_orientations);

} // namespace orc

#endif
