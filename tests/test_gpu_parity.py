"""Parity of the HIP sampler with the CPU oracle on an MI355X (cases in
parity_cases.py)."""
import pytest

import parity_cases as pc

pytestmark = [pytest.mark.gpu]


@pytest.fixture(scope="module")
def sampler_cls():
    from latticednaorigami_amd.sampler import Sampler
    return Sampler


def test_four_domain_constant_t_bit_exact(sampler_cls, oracle_mod):
    pc.case_four_domain_constant_t_bit_exact(sampler_cls, oracle_mod)


def test_lane_mappings_agree(sampler_cls, oracle_mod):
    pc.case_lane_mappings_agree(sampler_cls, oracle_mod)


def test_raster_nearest_neighbour_bit_exact(sampler_cls, oracle_mod):
    pc.case_raster_nearest_neighbour_bit_exact(sampler_cls, oracle_mod)


def test_sweep_kernels_agree(sampler_cls, oracle_mod):
    pc.case_sweep_kernels_agree(sampler_cls, oracle_mod)


def test_ptmc_exchange_matches_oracle(sampler_cls, oracle_mod):
    pc.case_ptmc_exchange_matches_oracle(sampler_cls, oracle_mod)


def test_pt_variants_match_oracle(sampler_cls, oracle_mod):
    pc.case_pt_variants_match_oracle(sampler_cls, oracle_mod)


def test_windows_exchange_matches_oracle(sampler_cls, oracle_mod):
    pc.case_windows_exchange_matches_oracle(sampler_cls, oracle_mod)


def test_window_histogram_and_grid_replacement(sampler_cls, oracle_mod):
    pc.case_window_histogram_and_grid_replacement(sampler_cls, oracle_mod)


def test_mwus3d_window_ladders_by_temperature(sampler_cls, oracle_mod):
    pc.case_mwus3d_window_ladders_by_temperature(sampler_cls, oracle_mod)


def test_three_quarter_turn_and_three_domain_staples(sampler_cls, oracle_mod):
    pc.case_three_quarter_turn_and_three_domain_staples(sampler_cls, oracle_mod)


def test_learned_move_matches_oracle(sampler_cls, oracle_mod):
    pc.case_learned_move_matches_oracle(sampler_cls, oracle_mod)


def test_learned_moves_in_3d_umbrella_windows(sampler_cls, oracle_mod):
    pc.case_learned_moves_in_3d_umbrella_windows(sampler_cls, oracle_mod)


def test_batch_matches_oracle(oracle_mod):
    from latticednaorigami_amd.sampler import BatchSampler
    pc.case_batch_matches_oracle(BatchSampler, oracle_mod)


def test_reference_rng_tape_replay(sampler_cls, oracle_mod):
    pc.case_reference_rng_tape_replay(sampler_cls, oracle_mod)


def test_config_roundtrip_and_errors(sampler_cls, oracle_mod):
    pc.case_config_roundtrip_and_errors(sampler_cls, oracle_mod)


def test_hand_built_fourbody_configurations(sampler_cls, oracle_mod):
    pc.case_hand_built_fourbody_configurations(sampler_cls, oracle_mod)


def test_folded_crossover_networks(sampler_cls, oracle_mod):
    pc.case_folded_crossover_networks(sampler_cls, oracle_mod)


def test_fixed_configuration_exchange_chains(sampler_cls, oracle_mod):
    pc.case_fixed_configuration_exchange_chains(sampler_cls, oracle_mod)


def test_met_known_answer_tapes(sampler_cls, oracle_mod):
    pc.case_met_known_answer_tapes(sampler_cls, oracle_mod)


def test_batch_lengths_8_to_128(oracle_mod):
    from latticednaorigami_amd.sampler import BatchSampler
    pc.case_batch_lengths_8_to_128(BatchSampler, oracle_mod)


def test_batch_full_workload(oracle_mod):
    from latticednaorigami_amd.sampler import BatchSampler
    pc.case_batch_full_workload(BatchSampler, oracle_mod)


def test_learned_rejects_single_domain_staples(sampler_cls, oracle_mod):
    pc.case_learned_rejects_single_domain_staples(sampler_cls, oracle_mod)


def test_ctrg_scaffold_regrowth_bit_exact(sampler_cls, oracle_mod):
    pc.case_ctrg_scaffold_regrowth_bit_exact(sampler_cls, oracle_mod)


def test_ctrg_recoil_and_attempt_limits(sampler_cls, oracle_mod):
    pc.case_ctrg_recoil_and_attempt_limits(sampler_cls, oracle_mod)


def test_ctrg_three_domain_staples(sampler_cls, oracle_mod):
    pc.case_ctrg_three_domain_staples(sampler_cls, oracle_mod)


def test_ctrg_rejects_bad_limits(sampler_cls, oracle_mod):
    pc.case_ctrg_rejects_bad_limits(sampler_cls, oracle_mod)


def test_ctrg_jump_bit_exact(sampler_cls, oracle_mod):
    pc.case_ctrg_jump_bit_exact(sampler_cls, oracle_mod)


def test_ctrg_jump_limits(sampler_cls, oracle_mod):
    pc.case_ctrg_jump_limits(sampler_cls, oracle_mod)


def test_ctcb_jump_bit_exact(sampler_cls, oracle_mod):
    pc.case_ctcb_jump_bit_exact(sampler_cls, oracle_mod)


def test_window_temperature_swaps(sampler_cls, oracle_mod):
    pc.case_window_temperature_swaps(sampler_cls, oracle_mod)


def test_mwus3d_full_size_exchanges(sampler_cls, oracle_mod):
    pc.case_mwus3d_full_size_exchanges(sampler_cls, oracle_mod)


def test_exchange_and_site_eval(sampler_cls, oracle_mod):
    pc.case_exchange_and_site_eval(sampler_cls, oracle_mod)
