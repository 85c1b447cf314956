// plan.h - host-side preparation of a sampler run.
//
// Everything the library derives from the C-ABI inputs before it touches a
// device: validation, the replica-block layout, the launch parameters, the
// energy tables of every ladder slot, per-slot identity tables, the packed
// initial configuration, and the reading of replica blocks back into ABI
// structs.  Plain C++ (no HIP), shared by sampler.hip and the host test
// build of the engine.
#ifndef LDO_PLAN_H
#define LDO_PLAN_H

#include <string>
#include <vector>

#include "../../include/ldo.h"
#include "../../include/ldo_mlp.h"
#include "layout.h"
#include "tables.h"

namespace ldo {

// an ABI call rejected with a status code and message
struct AbiError {
    int code;
    std::string msg;
};

Layout make_layout(int ns, int lmax, int smax, int ntypes, int tcap, bool rg, bool jump);

// element offsets of the learned proposal's flat parameter vector (ldo_mlp.h)
void mlp_layout(int n_ident, ldo_mlp_layout& L);

struct Plan {
    SystemSpec spec;
    Params P {};  // scalars and layout; the device pointers are the caller's
    int R = 0;
    int ntypes = 0;
    std::vector<int32_t> type_len;   // [ntypes]
    std::vector<int16_t> type_ident; // [ntypes * P.type_ident_stride] table index
    std::vector<double> ladder_t, ladder_u, ladder_b, ladder_s;
    std::vector<double> reduced_u;   // [R * (ntypes - 1)]
    int dim2 = 1;                    // stacking multipliers of a 2D ladder grid
};

// Validates the inputs of ldo_create and fills `out`; throws AbiError.
void make_plan(const ldo_system_desc* sys, const ldo_run_params* run, int32_t n_replicas,
               const double* temps, Plan& out);

// Dense energy tables of every ladder slot: tab = hyb_e | hyb_h | hyb_s |
// stack, each [R][n_ident^2]; init = init_e | init_h | init_s, each [R].
void slot_tables(const Plan& p, std::vector<double>& tab, std::vector<double>& init);

// ladder_t | ladder_u | ladder_s | reduced_u | rep_u, as the exchange reads
// them; rep_u (each replica's own chemical-potential multiplier) starts at
// the replica's bias multiplier, ladder_b[r]
std::vector<double> ladder_values(const Plan& p);

// The stacking multiplier slot q's sweep uses: its own with
// LDO_EXCHANGE_ST and LDO_EXCHANGE_2D, 1 with the T/UT/HUT ladders, whose
// update_control_qs calls update_temp(temp) (ptmc_simulation.cpp:652-670).
// Modes without a ladder keep the slot's value.
double slot_stacking(const Plan& p, int q);

// Table key class of every slot: slots with the same (temperature, sweep
// stacking multiplier) share a class, numbered from 0 in order of first
// appearance.  A table-cache hit keeps the initiation terms of the last
// build (OrigamiPotential::update_temp, origami_potential.cpp:1019-1042), so
// a replica's initiation terms change only on its first visit of a class.
// Returns the class count.
int slot_classes(const Plan& p, std::vector<int32_t>& key);

// The window count W of every ladder when the slots form equal ladders of
// consecutive slots at one temperature (neighbouring ladders at different
// temperatures), else -1.
int window_ladder_size(const Plan& p);

// Whether a replica takes its slot's chemical-potential multiplier after
// every exchange (m_exchange_q_is holds m_staple_u_mult_i: UT, HUT, 2D).
bool u_follows_slot(int exchange_mode);

// Checks the 2D grid (n_temps x n_stacking = R) and lays it out in the
// plan's ladder, slot i * n_stacking + j (ptmc_simulation.cpp:452-493).
void set_ladder_2d(Plan& p, int32_t n_temps, const double* temps, int32_t n_stacking,
                   const double* stacking_mults, double staple_u_mult);

// The learned moves place a staple end next to its chain neighbour: every
// staple type needs at least two domains (LDO_E_NOTIMPL otherwise).
void check_learned_types(const Plan& p);
// Checks a configuration against the plan and returns its domain count.
int config_domains(const Plan& p, const ldo_config* cfg);
// Validates a configuration and packs it as chain_index | chain_identity |
// positions | orientations (three ints per domain).
void pack_config(const Plan& p, const ldo_config* cfg, std::vector<int32_t>& packed,
                 int& n_domains);

// Reads the configuration held in a host copy of a replica block.
void read_config(const Plan& p, const char* block, ldo_config* cfg, int32_t cap_chains,
                 int32_t cap_domains);

void fill_observables(const Plan& p, const RepScalars& S, ldo_obs* out);
void add_counters(const RepScalars& S, ldo_counters* out);

// Throws the ABI error that a replica's status word stands for.
void check_status(int32_t status, int replica);

// Validates umbrella windows for this run (one per ladder slot).
void check_windows(const Plan& p, int32_t n, const ldo_window* windows, int32_t grid_len);

// LinearStepWellBiasFunction::calc_bias (bias_functions.cpp:139-151)
double window_restraint(const ldo_window& w, int op_value);

// the value of a window's order parameter in a replica
int window_op_value(const ldo_window& w, const RepScalars& S);

// Stored bias values of every window for the replicas now in them: the
// restraint at the replica's order parameter and a zero grid value, as a
// freshly built SystemBiases holds them (bias_functions.cpp:107-122, 222-224).
std::vector<double> initial_window_state(const ldo_window* windows, int32_t n,
                                         const std::vector<int32_t>& q_to_rep,
                                         const std::vector<RepScalars>& scalars);

} // namespace ldo

#endif
