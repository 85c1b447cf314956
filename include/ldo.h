/*
 * ldo.h - C ABI of libldo_hip.so, the MI355X (gfx950) lattice Monte Carlo
 * sampler for the LatticeDNAOrigami domain-resolution DNA-origami model.
 *
 * The reference has no FFI; its hot path sits behind the C++ simulation
 * loop contract (include/LatticeDNAOrigami/simulation.hpp:77-81,
 * src/simulation.cpp:568-653) driven one replica per MPI rank by the
 * constant-T, PTMC and PTMWUS drivers.  This ABI replaces that contract:
 * every replica of a run lives on one GPU, and each entry point below names
 * the reference interface it stands in for.
 *
 * Conventions
 *   - every function returns an int status (LDO_OK on success) and leaves a
 *     message for ldo_last_error();
 *   - the caller owns every buffer it passes in; the library copies inputs
 *     and writes outputs into caller memory;
 *   - calls are synchronous (the device stream is synchronised before
 *     return) and one host thread drives one ldo_ctx.
 *   - identities, positions and orientations follow the reference's system
 *     JSON (src/files.cpp:36-104): chain 0 is the scaffold, staple chain
 *     identities are 1-based, domain identities are signed integers and a
 *     staple domain binds the scaffold domain of opposite sign.
 */
#ifndef LDO_H
#define LDO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the reference's C++ exceptions (utility.hpp:30-58). */
#define LDO_OK 0
#define LDO_E_CONSTRAINT 1 /* utility::OrigamiMisuse */
#define LDO_E_MISUSE 2     /* utility::SimulationMisuse */
#define LDO_E_NOTIMPL 3    /* utility::NotImplemented */
#define LDO_E_IO 4         /* utility::FileError */
#define LDO_E_HIP 5        /* HIP runtime failure */
#define LDO_E_ARG 6        /* argument outside the documented range */

/* Option values (the reference's string options, src/parser.cpp:39-57). */
enum ldo_domain_type { LDO_HALFTURN = 0, LDO_THREEQUARTERTURN = 1 };
enum ldo_hybridization_pot {
    LDO_HYB_NEARESTNEIGHBOUR = 0,
    LDO_HYB_UNIFORM = 1,
    LDO_HYB_SPECIFIED = 2
};
enum ldo_misbinding_pot { LDO_MISBIND_OPPOSING = 0, LDO_MISBIND_DISALLOWED = 1 };
enum ldo_stacking_pot { LDO_STACK_CONSTANT = 0, LDO_STACK_SEQUENCESPECIFIC = 1 };

/* Move types of a moveset file (src/simulation.cpp:267-318). */
enum ldo_movetype {
    LDO_MOVE_ORIENTATION_ROTATION = 0,  /* orientation_movetype.cpp:30-65 */
    LDO_MOVE_MET_STAPLE_EXCHANGE = 1,   /* met_movetypes.cpp:192-403 */
    LDO_MOVE_CB_STAPLE_REGROWTH = 2,    /* cb_movetypes.cpp:292-348 */
    LDO_MOVE_CTCB_SCAFFOLD_REGROWTH = 3, /* cb_movetypes.cpp:680-747 */
    LDO_MOVE_CTRG_SCAFFOLD_REGROWTH = 4, /* rg_movetypes.cpp:177-533, 629-689 */
    LDO_MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH = 5, /* rg_movetypes.cpp:791-853,
                                               movetypes.cpp:512-721 */
    LDO_MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH = 6  /* cb_movetypes.cpp:872-983 */
};

/* Random streams.  PHILOX: a counter-based Philox4x32-10 stream keyed by
 * (seed, replica) whose counter is (step, draw) so every MC step has its own
 * sub-stream.  TAPE: each replica consumes a caller-supplied sequence of
 * recorded draws (ldo_set_tape), in the order the reference's RandomGens
 * would have produced them (random_gens.cpp:22-43). */
enum ldo_rng_mode { LDO_RNG_PHILOX = 0, LDO_RNG_TAPE = 1 };

/* Replica exchange (ptmc_simulation.cpp:360-412, 495-560;
 * us_simulation.cpp:770-864).  The parallel-tempering modes share one
 * acceptance (calc_acceptance_p, ptmc_simulation.cpp:275-313) and differ in
 * which control variables a replica takes from its ladder slot after each
 * exchange (m_exchange_q_is, ptmc_simulation.cpp:602-650) and in what its
 * sweep uses (update_control_qs, 595-680):
 *   T   temperature; the sweep's stacking multiplier is 1
 *   UT  temperature and staple chemical-potential multiplier; stacking 1
 *   HUT as UT, plus the bias multiplier (which the system ignores:
 *       OrigamiSystem::update_bias_mult is empty, origami_system.hpp:165)
 *   ST  temperature and stacking multiplier; the sweep's stacking energies
 *       are scaled by the slot's multiplier
 *   2D  a grid of temperatures x stacking multipliers (ldo_set_ladder_2d),
 *       swept as ST; exchanges alternate between the two directions.
 * The chemical-potential multiplier only enters the exchange. */
enum ldo_exchange_mode {
    LDO_EXCHANGE_NONE = 0,        /* independent constant-T chains */
    LDO_EXCHANGE_TEMPERATURE = 1, /* ut_parallel_tempering */
    LDO_EXCHANGE_WINDOWS = 2,     /* ptmwus: umbrella windows at one temperature */
    LDO_EXCHANGE_T = 3,           /* t_parallel_tempering */
    LDO_EXCHANGE_ST = 4,          /* st_parallel_tempering */
    LDO_EXCHANGE_HUT = 5,         /* hut_parallel_tempering */
    LDO_EXCHANGE_2D = 6           /* 2d_parallel_tempering */
};

/* System-wide order parameters a move-level bias can act on
 * (order_params.cpp: NumBoundDomainPairsOrderParam counts fully bound pairs,
 * 288-312; NumStaplesOrderParam, 194-209). */
enum ldo_order_param { LDO_OP_NUMFULLDOMAINS = 0, LDO_OP_NUMSTAPLES = 1 };

/* One umbrella window (one ladder slot): a LinearStepWell restraint on the
 * window's order parameter (bias_functions.cpp:139-151) and a Grid bias over
 * the same parameter (242-284), both multiplied by bias_mult (SystemBiases,
 * 475-496).  Grid values come separately (ldo_set_windows). */
typedef struct ldo_window {
    int32_t op;     /* ldo_order_param */
    int32_t min_op; /* window [min_op, max_op] */
    int32_t max_op;
    int32_t pad0;
    double well_bias;     /* restraint inside the window */
    double min_bias;      /* restraint one step outside */
    double slope;         /* restraint growth per further step */
    double off_grid_bias; /* grid bias where no grid value is given */
    double bias_mult;
} ldo_window;

/* System and energy model: the data setup_origami() reads
 * (origami_system.cpp:991-1030) and the OrigamiPotential options
 * (origami_potential.cpp:952-1012). */
typedef struct ldo_system_desc {
    int32_t n_chain_types;        /* identities.size(): scaffold + staple types */
    const int32_t* chain_lengths; /* [n_chain_types] */
    const int32_t* identities;    /* flattened, chain-type major */
    const char* const* sequences; /* [sum chain_lengths] or NULL */
    const double* enthalpies;     /* "Specified": per staple-domain index, or NULL */
    const double* entropies;
    int32_t n_specified;
    int32_t cyclic;
    int32_t domain_type;       /* ldo_domain_type */
    int32_t hybridization_pot; /* ldo_hybridization_pot */
    int32_t misbinding_pot;    /* ldo_misbinding_pot */
    int32_t stacking_pot;      /* ldo_stacking_pot */
    int32_t apply_mean_field_cor;
    int32_t pad0;
    double binding_h;    /* K   */
    double binding_s;    /* k_B */
    double misbinding_h; /* K   */
    double misbinding_s; /* k_B */
    double stacking_ene; /* K   */
    double cation_M;
    double staple_M;
} ldo_system_desc;

/* Run parameters: the moveset file and the .inp keys the hot loop reads
 * (parser.cpp:108-122, 241-251; simulation.cpp:192-237). */
typedef struct ldo_run_params {
    int32_t n_movetypes;
    const int32_t* movetype_kind;         /* ldo_movetype */
    const double* movetype_freq;          /* parsed "a/b" fractions */
    const int32_t* movetype_max_regrowth; /* CTCB max_regrowth, else ignored */
    int32_t max_total_staples;
    int32_t max_type_staples;
    int32_t max_staple_size;
    int32_t centering_freq;
    int32_t centering_domain;
    int32_t constraint_check_freq;
    uint64_t seed;
    int32_t rng_mode;      /* ldo_rng_mode */
    int32_t exchange_mode; /* ldo_exchange_mode */
    int32_t lanes_per_replica; /* 0: pick automatically; 1 or 64 */
    int32_t flags;             /* ldo_run_flags, 0 = the reference's behaviour */
    /* CTRG recoil growth (simulation.cpp:493-495): max_num_recoils (>= 0) and
     * max_c_attempts (1..36) per movetype; read for CTRG entries only, and
     * both must be given when the moveset holds one */
    const int32_t* movetype_max_num_recoils;
    const int32_t* movetype_max_c_attempts;
    /* max_seg_regrowth of the non-contiguous moves (CTRGJump, CTCBJump;
     * >= 1, with their max_regrowth <= 60), read for those entries only */
    const int32_t* movetype_max_seg_regrowth;
} ldo_run_params;

/* Opt-in deviations from the reference (ldo_run_params.flags).
 *
 * LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS: when a staple network bound to the
 * regrown scaffold segment is also bound outside it, the reference turns
 * each in-segment (staple domain, scaffold domain) pair into an active
 * endpoint keyed by the staple domain's chain
 * (Constraintpoints::add_active_endpoints_on_scaffold,
 * top_constraint_points.cpp:553-557), an endpoint no scaffold growth ever
 * consults.  A staple of three or more domains bound across the segment
 * boundary then leaves its in-segment scaffold domains unconstrained: CTCB
 * regrowth pulls them off the staple while both Rosenbluth weights drop the
 * binding trial, so the move does not satisfy detailed balance.  With this
 * flag the endpoint is placed on the scaffold domain at its own position, as
 * the growthpoint pairs of the same function are.  Two-domain staples that
 * bind only the scaffold never reach this code path.
 *
 * LDO_FLAG_WINDOW_TEMPERATURE_SWAPS (an extension: the reference's PTMWUS
 * exchanges windows at one temperature, us_simulation.cpp:770-864): with
 * LDO_EXCHANGE_WINDOWS and several window ladders (runs of consecutive slots
 * at one temperature, all of W windows, slot g * W + w holding window w of
 * ladder g, every ladder's windows equal to the first's), each exchange
 * round ends with swaps between window w of ladders g and g + 1 for every w
 * and every g = swap_i mod 2, 2 + swap_i mod 2, ...  A swap is accepted with
 * min(1, exp(a + m [G_a(p1) - G_b(p1) + G_b(p2) - G_a(p2)])): a the PTMC
 * log ratio of the two slots (calc_acceptance_p, ptmc_simulation.cpp:
 * 275-313), G the slots' grid biases, p1 and p2 the replicas' order
 * parameters and m the windows' bias multiplier; the restraints, equal in
 * both windows, cancel.  A replica that changes temperature has its energy
 * recomputed, as update_temp does. */
enum ldo_run_flags {
    LDO_FLAG_CTCB_SCAFFOLD_ENDPOINTS = 1,
    LDO_FLAG_WINDOW_TEMPERATURE_SWAPS = 2
};

/* One configuration: the Chain records of origami_system.hpp:46-62. */
typedef struct ldo_config {
    int32_t n_chains;
    int32_t pad0;
    int32_t* chain_index;    /* [n_chains] unique chain indices */
    int32_t* chain_identity; /* [n_chains] */
    int32_t* chain_length;   /* [n_chains] (output; ignored on input) */
    int32_t* positions;      /* [3 * domains] chain by chain */
    int32_t* orientations;   /* [3 * domains] */
} ldo_config;

/* Per-replica observables (origami_system.hpp:99-130, 204-246). */
typedef struct ldo_obs {
    double energy;          /* m_energy, kT */
    double enthalpy;        /* hybridization enthalpy (update_enthalpy_and_entropy) */
    double entropy;
    double stacking_energy; /* energy - (enthalpy - entropy) */
    double temperature;
    int64_t step;
    int32_t num_staples;
    int32_t num_domains;
    int32_t num_bound_domain_pairs;
    int32_t num_fully_bound_domain_pairs;
    int32_t num_misbound_domain_pairs;
    int32_t num_stacked_domain_pairs;
    int32_t num_unassigned_domains;
    int32_t status; /* 0, or LDO_E_CONSTRAINT after a failed constraint check */
    double bias;    /* the replica's move-level bias (kT) under its window */
    int32_t last_move;     /* moveset slot of the replica's last MC step */
    int32_t last_accepted; /* whether that move was accepted */
} ldo_obs;

#define LDO_MAX_MOVETYPES 8

typedef struct ldo_counters {
    int64_t attempts[LDO_MAX_MOVETYPES];
    int64_t accepts[LDO_MAX_MOVETYPES];
    /* trial placements onto empty sites / onto unbound occupants (binding
     * evaluations) and committed domain sets+unsets; the byte model of the
     * throughput roofline is built from these counted events */
    int64_t site_empty;
    int64_t site_bind;
    int64_t commits;
    int64_t exchange_attempts;
    int64_t exchange_accepts;
} ldo_counters;

/* Create a context holding n_replicas replicas of one system.  Every replica
 * starts from `init` at temperature temps[r] (n_replicas values).
 * Replaces setup_origami + GCMCSimulation construction
 * (origami_system.cpp:991-1030; simulation.cpp:182-258). */
int ldo_create(const ldo_system_desc* sys,
               const ldo_run_params* run,
               const ldo_config* init,
               int32_t n_replicas,
               const double* temps,
               void** out_ctx);

void ldo_destroy(void* ctx);

/* Message of the last failed call on ctx (or of the last failed
 * ldo_create when ctx is NULL). */
const char* ldo_last_error(void* ctx);

/* Replace the configuration of one replica (OrigamiSystem::set_config,
 * origami_system.cpp:327-341). */
int ldo_set_config(void* ctx, int32_t replica, const ldo_config* cfg);

/* ldo_set_config without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops).  The configuration is staged in a static
 * device array of LDO_NOALLOC_CONFIG_INTS int32 values per GPU (2 per chain
 * + 6 per domain; 256 KiB, held until the process ends); a larger
 * configuration returns LDO_E_ARG.  Every call on one GPU shares that array:
 * calls that overlap on one GPU race. */
#define LDO_NOALLOC_CONFIG_INTS 65536
int ldo_set_config_noalloc(void* ctx, int32_t replica, const ldo_config* cfg);

/* Read one replica's configuration (OrigamiSystem::chains(),
 * origami_system.cpp:175-191).  cfg arrays must hold cap_chains chains and
 * cap_domains domains. */
int ldo_get_config(void* ctx, int32_t replica, ldo_config* cfg,
                   int32_t cap_chains, int32_t cap_domains);

/* Control variables of a temperature ladder slot (PTGCMCSimulation control
 * q's, ptmc_simulation.cpp:337-358): temperature, staple chemical potential
 * multiplier, bias multiplier and stacking multiplier.  Every replica's own
 * chemical-potential multiplier, the one its first exchange sees, is reset
 * to bias_mults[replica]: OneDPTGCMCSimulation::initialize_control_qs stores
 * the bias multiplier under index m_bias_i, which equals m_staple_u_mult_i
 * (ptmc_simulation.cpp:349-356, ptmc_simulation.hpp:74, 80).  With
 * LDO_EXCHANGE_T and LDO_EXCHANGE_ST a replica keeps that value for the whole
 * run. */
int ldo_set_ladder(void* ctx, int32_t n, const double* temps,
                   const double* staple_u_mults, const double* bias_mults,
                   const double* stacking_mults);

/* The grid of LDO_EXCHANGE_2D (TwoDPTGCMCSimulation::initialize_control_qs,
 * ptmc_simulation.cpp:452-493): slot i * n_stacking + j holds temps[i] and
 * stacking_mults[j], every slot the chemical-potential multiplier
 * staple_u_mult and bias multiplier 1 (so every replica's own multiplier
 * starts at 1, as with ldo_set_ladder); n_temps * n_stacking must equal the
 * replica count.  Exchange round swap_i pairs set swap_i % 4 of
 * attempt_exchange (495-560): 0 and 2 neighbouring temperatures from even
 * and odd rows, 1 and 3 neighbouring stacking multipliers from even and odd
 * columns. */
int ldo_set_ladder_2d(void* ctx, int32_t n_temps, const double* temps, int32_t n_stacking,
                      const double* stacking_mults, double staple_u_mult);

/* Umbrella windows, one per ladder slot (n = the replica count): window
 * parameters and grid biases grid[w * grid_len + p] for order-parameter
 * values 0 <= p < grid_len.  Replaces the bias-function file and the
 * umbrella-sampling grid of USGCMCSimulation (us_simulation.cpp:35-67,
 * 503-516).  Each window's stored bias values restart from its replica's
 * current configuration. */
int ldo_set_windows(void* ctx, int32_t n, const ldo_window* windows, const double* grid,
                    int32_t grid_len);

/* New grid biases for every window (grid[w * grid_len + p], the grid_len of
 * ldo_set_windows), leaving the windows' stored bias values as they are:
 * GridBiasFunction::replace_biases (bias_functions.cpp:236-239) as the
 * adaptive umbrella-sampling loop calls it between iterations
 * (us_simulation.cpp:385-416).  The next move's bias difference is taken
 * against the value stored under the old grid, as in the reference. */
int ldo_set_window_grid(void* ctx, const double* grid);

/* Counting of each window's order parameter after every step of ldo_sweep
 * and ldo_run (USGCMCSimulation::update_internal, us_simulation.cpp:262-266:
 * the f_i histogram of the adaptive loop).  on != 0 starts counting from
 * zero; on == 0 stops and frees the counts.  Requires windows; a later
 * ldo_set_windows zeroes the counts.  Steps that end in a constraint failure
 * are not counted. */
int ldo_set_histogram(void* ctx, int32_t on);

/* counts[w * (grid_len + 1) + p]: steps window w ended with its order
 * parameter at p (0 <= p < grid_len); bin grid_len counts values off the
 * grid.  reset != 0 zeroes the counts after the copy. */
int ldo_get_histogram(void* ctx, int64_t* counts, int32_t reset);

/* Supply a recorded draw sequence for one replica (LDO_RNG_TAPE).  Entry k
 * is the k-th value the reference's RandomGens returned: a uniform_real as
 * its IEEE-754 bit pattern, a uniform_int as the integer. */
int ldo_set_tape(void* ctx, int32_t replica, const uint64_t* tape, int64_t n);

/* Set every replica's step counter: the next MC step is step + 1, as when
 * GCMCSimulation::simulate starts from start_step (simulation.cpp:567-573).
 * Centering, constraint checks and the Philox sub-streams follow this
 * numbering.  AnnealingGCMCSimulation::run starts each temperature's
 * simulate from the previous start plus that call's return value
 * (annealing_simulation.cpp:38-50). */
int ldo_set_step(void* ctx, int64_t step);

/* Run `steps` MC steps on every replica with no exchange
 * (GCMCSimulation::simulate, simulation.cpp:568-653). */
int ldo_sweep(void* ctx, int64_t steps);

/* Learned moves: no counterpart in the reference (the learned proposal of
 * this sampler).  With probability `fallback` an attempt is a plain CB
 * staple regrowth.  Otherwise, with probability 1/2 each:
 *  - domain regrowth: a staple drawn uniformly and one of its two end
 *    domains d; the network of ldo_mlp.h, given the 7x7x7 box around d's
 *    chain neighbour with d left out, proposes d's new site (one of the
 *    neighbour's 6) and orientation; accepted with
 *    min(1, exp(-dE - dB) q(old) / q(new)) if every staple stays connected
 *    to the scaffold;
 *  - staple exchange (two-domain staples; scaffold domain k must carry
 *    identity -(k+1)): an insertion (1/2) binds a uniformly drawn domain of
 *    a uniformly drawn type to its complementary scaffold domain and places
 *    the other domain from the network; a deletion (1/2) removes a uniformly
 *    drawn staple; acceptance targets the reference's grand-canonical
 *    weight, summing the proposal over the staple's growth paths.
 * Counters: attempts/accepts[7] domain regrowth, [6] CB fallbacks, [5]
 * staple exchange (so the move set may use slots 0-4).
 * The parameters are copied to the device; n_ident must be at least the
 * number of scaffold domains (identity index = |identity| - 1). */
int ldo_learned_set_network(void* ctx, int32_t n_ident, const float* params, double fallback);

/* `attempts` learned-move attempts on every replica, each with the tables
 * and umbrella window of its last sweep. */
int ldo_learned_sweep(void* ctx, int64_t attempts);

/* Training samples from the current configurations (no state change): for
 * every staple of every replica and both of its ends, the network's input
 * box around the end's chain neighbour (end left out, channels [343]), the
 * end's identity index, the legal outputs [36] and the output the end
 * occupies (target).  Writes at most `cap` samples, replica by replica, and
 * their number to n_out. */
int ldo_learned_samples(void* ctx, int32_t cap, uint8_t* channels, int32_t* ident,
                        uint8_t* legal, int32_t* target, int32_t* n_out);

/* PTMC: n_swaps rounds of [apply ladder slot -> simulate(exchange_interval)
 * -> dependent q's -> device exchange] (PTGCMCSimulation::run,
 * ptmc_simulation.cpp:106-150).  With LDO_EXCHANGE_WINDOWS the exchange is
 * PTMWUSGCMCSimulation::attempt_exchange (us_simulation.cpp:662-706,
 * 770-864): adjacent windows swap configurations. */
int ldo_run(void* ctx, int64_t n_swaps, int64_t exchange_interval);

/* One exchange round on the replicas' current configurations: a round of
 * PTGCMCSimulation::run (ptmc_simulation.cpp:106-150) or of
 * PTMWUSGCMCSimulation::run_swaps (us_simulation.cpp:672-697) with an
 * interval of 0 steps, i.e. every replica takes its slot's control values
 * (update_control_qs), its dependent quantities are recomputed
 * (update_dependent_qs), then attempt_exchange and master_send run with pair
 * set swap_index (its parity; swap_index % 4 for the 2D grid) instead of the
 * run's own round count.  Draws from the exchange stream at its next
 * counter. */
int ldo_exchange(void* ctx, int64_t swap_index);

/* Site evaluation: domain `domain` of the chain with unique index
 * `chain_index` in `replica` placed at pos with orientation ore, on the
 * replica's current configuration and tables.  The domain is first taken
 * off its site if it is placed (OrigamiSystem::unassign_domain,
 * origami_system.cpp:375-385), then checked at the new site
 * (check_domain_constraints, 343-355), as the CB and CTCB trial loops score
 * a trial (cb_movetypes.cpp:58-102).  Every query sees the configuration as
 * it was before the call; the replica is left unchanged. */
typedef struct ldo_site_query {
    int32_t replica;
    int32_t chain_index; /* ldo_config.chain_index of the chain */
    int32_t domain;      /* domain index within the chain */
    int32_t pos[3];
    int32_t ore[3];      /* a unit lattice vector */
    int32_t pad;
} ldo_site_query;

typedef struct ldo_site_result {
    double unassign_delta; /* energy change of taking the domain off (0 if unplaced) */
    double delta;          /* energy change of the placement (kT) */
    int32_t violated;      /* 1: the placement breaks a constraint (the reference's
                              OrigamiMisuse), delta is then not meaningful */
    int32_t pad;
} ldo_site_result;

int ldo_site_eval(void* ctx, const ldo_site_query* queries, int32_t n, ldo_site_result* results);

/* The next exchange round pairs as the first of a run (swap_i = 1, odd
 * pairs first), as PTMWUSGCMCSimulation::run_swaps restarts swap_i every
 * iteration (us_simulation.cpp:662-706).  The exchange's random stream
 * continues. */
int ldo_restart_swaps(void* ctx);

/* Per-move tracking counts of one replica, as the movetypes'
 * write_log_summary report them in the .moves file (MCMovetype::add_tracker,
 * movetypes.cpp:41-51; met_movetypes.cpp:145-190; cb_movetypes.cpp:264-290,
 * 626-678).  With T = number of chain identities (scaffold included, index =
 * identity), ns scaffold domains and S = max_total_staples, out holds
 *   MetStapleExchange insertion attempts[T], accepts[T],
 *                     deletion attempts[T], accepts[T],
 *   CBStapleRegrowth attempts[T], accepts[T] (moves with staples present),
 *   CTCBScaffoldRegrowth attempts[ns+1], accepts[ns+1] by segment length,
 *                        attempts[S+1], accepts[S+1] by staples regrown,
 *   CTRGScaffoldRegrowth attempts[ns+1], accepts[ns+1] by segment length
 *                        (rg_movetypes.cpp:583-627),
 *   CTRGJumpScaffoldRegrowth attempts, accepts (the reference tracks them
 *                        under 0 scaffold domains, rg_movetypes.cpp:745-789),
 *   CTCBJumpScaffoldRegrowth attempts[S+1], accepts[S+1] by staples regrown
 *                        (all under 0 scaffold domains, cb_movetypes.cpp:816-870);
 * 6T + 4(ns+1) + 4(S+1) + 2 values in all, written to *n_out (cap must hold
 * them).  Counts are per move kind, summed over moveset slots of that kind. */
int ldo_get_move_tracking(void* ctx, int32_t replica, int64_t* out, int64_t cap,
                          int64_t* n_out);

int ldo_get_observables(void* ctx, int32_t replica, ldo_obs* out);
int ldo_get_counters(void* ctx, ldo_counters* out);

/* Exchange attempts and accepts per slot pair, for the acceptance
 * frequencies the PTMC drivers print at the end of a run
 * (write_acceptance_freqs, ptmc_simulation.cpp:414-426, 562-592):
 * element d * R + q (n = 2R) counts the pair of slot q with its neighbour in
 * direction d, where d = 0 is the ladder's own direction (slot q + 1; in the
 * 2D grid the next temperature, slot q + n_stacking) and d = 1 the stacking
 * direction of the 2D grid (slot q + 1).  Window exchanges count under d = 0
 * too.  No reference counterpart as an entry point. */
int ldo_get_exchange_counts(void* ctx, int32_t n, int64_t* attempts, int64_t* accepts);

/* Ladder slot -> replica permutation (m_q_to_repi, written to .swp). */
int ldo_get_permutation(void* ctx, int32_t* q_to_rep);

/* Set the permutation, as a restart from a swap file does (the file's last
 * line, ptmc_simulation.cpp:54-84); every replica takes its new slot's
 * temperature and stacking multiplier when the next ldo_run interval begins
 * (its own chemical-potential multiplier changes at the next exchange, as
 * the exchange mode has it). */
int ldo_set_permutation(void* ctx, const int32_t* q_to_rep);

/* The HIP stream the sampler's kernels are launched on (hipStream_t). */
void* ldo_get_stream(void* ctx);

/* Name of the sweep kernel this context launches (ldo_sweep_kernel, or
 * ldo_sweep_lds_kernel when every replica block is held in LDS), for matching
 * profiler output. */
const char* ldo_kernel_name(void* ctx);

/* Kernel timing with HIP events recorded on the sampler's stream around
 * every sweep-kernel and exchange-kernel launch of ldo_sweep / ldo_run.
 * No reference counterpart: it serves throughput and roofline measurement.
 * ldo_set_profiling(ctx, 1) starts a fresh tally, 0 stops recording.
 * ldo_get_profile returns the summed kernel milliseconds and launch counts
 * since the tally started (reading synchronizes the stream). */
int ldo_set_profiling(void* ctx, int32_t on);
int ldo_get_profile(void* ctx, double* sweep_ms, int64_t* sweep_launches,
                    double* exchange_ms, int64_t* exchange_launches);

/* ------------------------------------------------------------- batching
 * Many independent constant-temperature chains, each a system of its own
 * (scaffold, staple types, energies, moveset, seed), advanced together by one
 * kernel launch per sweep.  The reference runs each such chain as a separate
 * ConstantTGCMCSimulation process (constant_temp_simulation.hpp:31); no
 * single reference interface batches them.  Chain k runs exactly the chain a
 * one-replica context of system k at temps[k] would run (same seed, same
 * random stream), so ldo_create / ldo_sweep on system k alone reproduces it.
 * Philox random streams only.  Errors as for ldo_create. */
int ldo_batch_create(int32_t n_chains, const ldo_system_desc* systems,
                     const ldo_run_params* runs, const ldo_config* init,
                     const double* temps, void** out_batch);
void ldo_batch_destroy(void* batch);
const char* ldo_batch_last_error(void* batch);

/* Sweep kernels a batch can use: one chain per workgroup with its block in
 * LDS, one chain per wavefront with its block in device memory (all 64 lanes
 * run the chain, splitting its independent lookups), one chain per lane (64
 * chains per wavefront), or one chain per wavefront run by its first lane
 * alone (fewer memory-pipeline issues per step when thousands of wavefronts
 * share the chip). */
enum ldo_batch_kernel {
    LDO_BATCH_LDS = 0,
    LDO_BATCH_WAVE = 1,
    LDO_BATCH_LANE = 2,
    LDO_BATCH_WAVE_SOLO = 3
};

/* Times every available sweep kernel on `steps_per_trial` steps of every
 * chain (real steps: the chains advance), keeps the fastest and returns it
 * with the milliseconds of each trial in ms[0 .. LDO_BATCH_WAVE_SOLO]
 * (ms[k] < 0: kernel k unavailable; ms may be NULL). */
int ldo_batch_tune(void* batch, int64_t steps_per_trial, int32_t* chosen, double* ms);
int ldo_batch_set_kernel(void* batch, int32_t kernel);

/* `steps` MC steps on every chain; the kernel time in *kernel_ms. */
int ldo_batch_sweep(void* batch, int64_t steps, double* kernel_ms);

/* Observables of chains [first, first + n), counters summed over all chains,
 * and one chain's configuration. */
int ldo_batch_get_observables(void* batch, int32_t first, int32_t n, ldo_obs* out);
int ldo_batch_get_counters(void* batch, ldo_counters* out);
int ldo_batch_get_config(void* batch, int32_t chain, ldo_config* cfg, int32_t cap_chains,
                         int32_t cap_domains);

#ifdef __cplusplus
}
#endif

#endif /* LDO_H */
