/*
 * ldo_oracle.h - C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a single-threaded FP64
 * restatement of the LatticeDNAOrigami hot path (lattice state, FourBody
 * energy, the four BASELINE move types, the simulate loop, PTMC exchange and
 * the exact enumerator).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the CPU baseline;
 * the product library (libldo_hip.so) never links or calls it.
 *
 * It shares the plain data structs of include/ldo.h so that one Python
 * description of a system drives both the oracle and the HIP sampler.
 */
#ifndef LDO_ORACLE_H
#define LDO_ORACLE_H

#include <stdint.h>

#include "../include/ldo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rng_kind: 0 = Philox stream identical to the HIP sampler's LDO_RNG_PHILOX;
 *           1 = std::mt19937_64 with libstdc++ distributions, i.e. the
 *               reference's RandomGens (random_gens.cpp:12-49). */
void* oracle_create(const ldo_system_desc* sys, const ldo_run_params* run,
                    const ldo_config* init, int32_t n_replicas,
                    const double* temps, int32_t rng_kind);
void oracle_destroy(void* o);
const char* oracle_last_error(void* o);

int oracle_set_ladder(void* o, int32_t n, const double* temps,
                      const double* staple_u_mults, const double* bias_mults,
                      const double* stacking_mults);
/* Umbrella windows (see ldo_set_windows); with LDO_EXCHANGE_WINDOWS in the
 * run parameters oracle_run performs the window exchange. */
int oracle_set_windows(void* o, int32_t n, const ldo_window* windows, const double* grid,
                       int32_t grid_len);
int oracle_sweep(void* o, int64_t steps);
int oracle_set_permutation(void* o, const int32_t* q_to_rep);
int oracle_set_ladder_2d(void* o, int32_t n_temps, const double* temps, int32_t n_stacking,
                         const double* stacking_mults, double staple_u_mult);
int oracle_get_exchange_counts(void* o, int32_t n, int64_t* attempts, int64_t* accepts);
int oracle_set_step(void* o, int64_t step);
int oracle_exchange(void* o, int64_t swap_index);
int oracle_site_eval(void* o, const ldo_site_query* q, int32_t n, ldo_site_result* res);
/* ldo_set_window_grid, ldo_set_histogram, ldo_get_histogram */
int oracle_set_window_grid(void* o, const double* grid);
int oracle_set_histogram(void* o, int32_t on);
int oracle_get_histogram(void* o, int64_t* counts, int32_t reset);
/* learned domain regrowth (the sampler's ldo_learned_* restated) */
int oracle_learned_set_network(void* o, int32_t n_ident, const float* params, double fallback);
int oracle_learned_sweep(void* o, int64_t attempts);
int oracle_learned_samples(void* o, int32_t cap, uint8_t* channels, int32_t* ident,
                           uint8_t* legal, int32_t* target, int32_t* n_out);
int oracle_run(void* o, int64_t n_swaps, int64_t exchange_interval);
int oracle_get_observables(void* o, int32_t replica, ldo_obs* out);
int oracle_get_counters(void* o, ldo_counters* out);
int oracle_get_permutation(void* o, int32_t* q_to_rep);
int oracle_get_config(void* o, int32_t replica, ldo_config* cfg,
                      int32_t cap_chains, int32_t cap_domains);
int oracle_set_config(void* o, int32_t replica, const ldo_config* cfg);

/* Tape recording (mt19937_64 mode): every draw of replica r is appended to
 * its tape; the tape is what ldo_set_tape replays on the GPU. */
// per-move tracking counts (layout of ldo_get_move_tracking)
int oracle_get_move_tracking(void* o, int32_t replica, int64_t* out, int64_t cap, int64_t* n_out);
// the next exchange round pairs as the first of a run (ldo_restart_swaps)
int oracle_restart_swaps(void* o);
// replica `replica` replays these draws (as ldo_set_tape does)
int oracle_set_tape(void* o, int32_t replica, const uint64_t* values, int64_t n);
int oracle_record_tape(void* o, int32_t replica, int32_t on);
int64_t oracle_tape_length(void* o, int32_t replica);
int oracle_get_tape(void* o, int32_t replica, uint64_t* out, int64_t cap);

/* Energy tables at a temperature (kT units; origami_potential.cpp:1057-1221) */
int oracle_energy_tables(void* o, double temp, double stacking_mult,
                         int32_t ident_i, int32_t ident_j, double* out5);

/* Recompute the energy of replica r's configuration from scratch by the
 * reference's unassign-all / set-all sequence (origami_system.cpp:808-826) */
int oracle_recompute_energy(void* o, int32_t replica, double* out_energy);

/* Test hook: unassign every domain of replica r, then assign them again in
 * `order` (indices into the chain-order list of all domains, each once),
 * through the same set_domain_config as the chain-order rebuild. Reports the
 * energy and stacked-pair count so built; the replica is left rebuilt in chain
 * order. The energy is a function of the configuration, so every order must
 * give the chain-order value. */
int oracle_energy_in_order(void* o, int32_t replica, const int32_t* order, int32_t n,
                           double* out_energy, int64_t* out_stacked);

/* Nearest-neighbour thermodynamics (nearest_neighbour.cpp:34-177) */
int oracle_nn_hybridization_thermo(const char* seq, double temp, double cation_M,
                                   double* out_h, double* out_s);
double oracle_nn_reduced_energy(const char* seq, double temp, double cation_M);
/* writes the longest contiguous complements, separated by ' ', into out */
int oracle_nn_longest_contig_complement(const char* seq_i, const char* seq_j,
                                        char* out, int32_t cap);

/* Ideal random walk count (ideal_random_walk.cpp:14-73), long double sum */
double oracle_num_walks(int32_t dx, int32_t dy, int32_t dz, int32_t steps);

/* VectorThree rotations (utility.cpp:68-142): v, axis as 3-int arrays */
int oracle_rotate_half(const int32_t* v, const int32_t* axis, int32_t* out);
int oracle_rotate(const int32_t* v, const int32_t* axis, int32_t turns,
                  int32_t* out);

/* Philox4x32-10: ctr[4], key[2] -> out[4] */
void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out);

/* Exact enumeration of the grand-canonical ensemble at the first ladder
 * temperature under the run's staple limits, with the weights of
 * enumerate.cpp:257-275, 480-520, 639-664.  Writes rows of five doubles
 * (numstaples, numfulldomains, nummisdomains, numstackedpairs, probability)
 * into out_rows (capacity cap_rows rows), the row count into n_rows, ln Z and
 * the number of configurations counted.  flags bit 0: enumerate every
 * orientation explicitly instead of counting the orientations of unbound
 * and misbound domains as a factor of 6. */
int oracle_enumerate(void* o, int32_t flags, double* out_rows, int32_t cap_rows,
                     int32_t* n_rows, double* out_log_z, double* out_n_configs);

#ifdef __cplusplus
}
#endif

#endif /* LDO_ORACLE_H */
