"""The engine's own source, built for the host, against the CPU oracle.

The HIP kernels run the __host__ __device__ functions of
latticednaorigami_amd/csrc/replica_ops.h; tests/engine_host builds the same
functions for the CPU.  These are the GPU parity cases (parity_cases.py)
run on that build: they check the engine's logic step for step against the
oracle wherever no GPU is at hand.
"""
import ctypes as C
import os
import subprocess

import pytest

import parity_cases as pc
from latticednaorigami_amd import abi, systems

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_DIR = os.path.join(HERE, "engine_host")
HOST_LIB = os.path.join(HOST_DIR, "_build", "libldo_engine_host.so")
NAMES = ["create", "destroy", "last_error", "set_config", "get_config", "set_ladder", "set_tape",
         "set_windows", "sweep", "run", "get_observables", "get_counters", "get_permutation",
         "set_window_grid", "set_histogram", "get_histogram", "restart_swaps",
         "get_move_tracking", "set_ladder_2d", "get_exchange_counts", "set_step", "exchange",
         "site_eval"]


class HostLib:
    """The host build's eh_* functions under the ldo_* names Sampler calls."""

    def __init__(self, cdll):
        abi.declare_abi(cdll, "eh_")
        cdll.eh_create.restype = C.c_int
        cdll.eh_create.argtypes = [C.POINTER(abi.SystemDesc), C.POINTER(abi.RunParams),
                                   C.POINTER(abi.Config), C.c_int32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_void_p)]
        cdll.eh_set_tape.restype = C.c_int
        cdll.eh_set_tape.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64), C.c_int64]
        cdll.eh_restart_swaps.restype = C.c_int
        cdll.eh_restart_swaps.argtypes = [C.c_void_p]
        abi.declare_learned_abi(cdll, "eh_")
        for n in NAMES + ["learned_set_network", "learned_sweep", "learned_samples"]:
            setattr(self, "ldo_" + n, getattr(cdll, "eh_" + n))


@pytest.fixture(scope="module")
def sampler_cls():
    if not os.path.exists(HOST_LIB):
        subprocess.run(["make", "-C", HOST_DIR], check=True, capture_output=True)
    lib = HostLib(C.CDLL(HOST_LIB))
    from latticednaorigami_amd.sampler import Sampler

    def make(system, **kw):
        return Sampler(system, lib=lib, **kw)
    return make


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


def test_windows_exchange_matches_oracle(sampler_cls, oracle_mod):
    pc.case_windows_exchange_matches_oracle(sampler_cls, oracle_mod)


def test_window_histogram_and_grid_replacement(sampler_cls, oracle_mod):
    pc.case_window_histogram_and_grid_replacement(sampler_cls, oracle_mod)


def test_reference_rng_tape_replay(sampler_cls, oracle_mod):
    pc.case_reference_rng_tape_replay(sampler_cls, oracle_mod)


def test_config_roundtrip_and_errors(sampler_cls, oracle_mod):
    pc.case_config_roundtrip_and_errors(sampler_cls, oracle_mod)


def test_mwus3d_window_ladders_by_temperature(sampler_cls, oracle_mod):
    pc.case_mwus3d_window_ladders_by_temperature(sampler_cls, oracle_mod)


def test_three_quarter_turn_and_three_domain_staples(sampler_cls, oracle_mod):
    pc.case_three_quarter_turn_and_three_domain_staples(sampler_cls, oracle_mod)


def test_learned_move_matches_oracle(sampler_cls, oracle_mod):
    pc.case_learned_move_matches_oracle(sampler_cls, oracle_mod)


def test_learned_moves_in_3d_umbrella_windows(sampler_cls, oracle_mod):
    pc.case_learned_moves_in_3d_umbrella_windows(sampler_cls, oracle_mod)


def test_reference_ctcb_rule_fails_alike_on_3d_system(sampler_cls, oracle_mod):
    """Under the reference's endpoint rule the 3D three-domain-staple system
    corrupts its scaffold (a dangling staple domain blocks the rejected
    CTCB move's old-configuration regrowth, cb_movetypes.cpp:104-130, which
    returns early and leaves scaffold domains unassigned) and a later move
    binds onto a bound site: oracle and engine stop with the same
    constraint error (OrigamiMisuse in the reference, origami_system.cpp:804)."""
    s, wins, grid, glen = systems.mwus3d_system(flags=0)
    s.temps = s.temps[:8]
    for make in (lambda: sampler_cls(s, temps=s.temps, exchange_mode=abi.EXCHANGE_WINDOWS),
                 lambda: oracle_mod.Oracle(s, temps=s.temps, exchange_mode=abi.EXCHANGE_WINDOWS)):
        x = make()
        x.set_windows(wins[:8], grid, glen)
        with pytest.raises(abi.LdoError) as e:
            x.run(3, 500)
        assert e.value.code == abi.LDO_E_CONSTRAINT


def test_hand_built_fourbody_configurations(sampler_cls, oracle_mod):
    pc.case_hand_built_fourbody_configurations(sampler_cls, oracle_mod)


def test_folded_crossover_networks(sampler_cls, oracle_mod):
    pc.case_folded_crossover_networks(sampler_cls, oracle_mod)


def test_fixed_configuration_exchange_chains(sampler_cls, oracle_mod):
    pc.case_fixed_configuration_exchange_chains(sampler_cls, oracle_mod)


def test_met_known_answer_tapes(sampler_cls, oracle_mod):
    pc.case_met_known_answer_tapes(sampler_cls, oracle_mod)


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


def test_pt_variants_match_oracle(sampler_cls, oracle_mod):
    pc.case_pt_variants_match_oracle(sampler_cls, oracle_mod)


def test_window_temperature_swaps(sampler_cls, oracle_mod):
    pc.case_window_temperature_swaps(sampler_cls, oracle_mod)


def test_exchange_and_site_eval(sampler_cls, oracle_mod):
    pc.case_exchange_and_site_eval(sampler_cls, oracle_mod)
