"""ctypes mirror of include/ldo.h and the loader of libldo_hip.so.

The structs here are field-for-field copies of the C declarations; the
library is the product (HIP kernels behind a C ABI).  Loading fails loudly
when the library has not been built: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libldo_hip.so")

LDO_OK = 0
LDO_E_CONSTRAINT = 1
LDO_E_MISUSE = 2
LDO_E_NOTIMPL = 3
LDO_E_IO = 4
LDO_E_HIP = 5
LDO_E_ARG = 6

HALFTURN, THREEQUARTERTURN = 0, 1
HYB_NEARESTNEIGHBOUR, HYB_UNIFORM, HYB_SPECIFIED = 0, 1, 2
MISBIND_OPPOSING, MISBIND_DISALLOWED = 0, 1
STACK_CONSTANT, STACK_SEQUENCESPECIFIC = 0, 1
MOVE_ORIENTATION_ROTATION = 0
MOVE_MET_STAPLE_EXCHANGE = 1
MOVE_CB_STAPLE_REGROWTH = 2
MOVE_CTCB_SCAFFOLD_REGROWTH = 3
MOVE_CTRG_SCAFFOLD_REGROWTH = 4
MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH = 5
MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH = 6
RNG_PHILOX, RNG_TAPE = 0, 1
FLAG_CTCB_SCAFFOLD_ENDPOINTS = 1
# swaps between the same window of neighbouring temperature ladders
FLAG_WINDOW_TEMPERATURE_SWAPS = 2
EXCHANGE_NONE, EXCHANGE_TEMPERATURE, EXCHANGE_WINDOWS = 0, 1, 2
# the other parallel-tempering drivers (ptmc_simulation.cpp:602-680);
# EXCHANGE_TEMPERATURE is ut_parallel_tempering
EXCHANGE_T, EXCHANGE_ST, EXCHANGE_HUT, EXCHANGE_2D = 3, 4, 5, 6
PT_EXCHANGE_MODES = {"t_parallel_tempering": EXCHANGE_T,
                     "ut_parallel_tempering": EXCHANGE_TEMPERATURE,
                     "hut_parallel_tempering": EXCHANGE_HUT,
                     "st_parallel_tempering": EXCHANGE_ST,
                     "2d_parallel_tempering": EXCHANGE_2D}
OP_NUMFULLDOMAINS, OP_NUMSTAPLES = 0, 1
MAX_MOVETYPES = 8

MOVETYPE_NAMES = {
    "OrientationRotation": MOVE_ORIENTATION_ROTATION,
    "MetStapleExchange": MOVE_MET_STAPLE_EXCHANGE,
    "CBStapleRegrowth": MOVE_CB_STAPLE_REGROWTH,
    "CTCBScaffoldRegrowth": MOVE_CTCB_SCAFFOLD_REGROWTH,
    "CTRGScaffoldRegrowth": MOVE_CTRG_SCAFFOLD_REGROWTH,
    "CTRGJumpScaffoldRegrowth": MOVE_CTRG_JUMP_SCAFFOLD_REGROWTH,
    "CTCBJumpScaffoldRegrowth": MOVE_CTCB_JUMP_SCAFFOLD_REGROWTH,
}

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_chain_types", C.c_int32),
        ("chain_lengths", _i32p),
        ("identities", _i32p),
        ("sequences", C.POINTER(C.c_char_p)),
        ("enthalpies", _f64p),
        ("entropies", _f64p),
        ("n_specified", C.c_int32),
        ("cyclic", C.c_int32),
        ("domain_type", C.c_int32),
        ("hybridization_pot", C.c_int32),
        ("misbinding_pot", C.c_int32),
        ("stacking_pot", C.c_int32),
        ("apply_mean_field_cor", C.c_int32),
        ("pad0", C.c_int32),
        ("binding_h", C.c_double),
        ("binding_s", C.c_double),
        ("misbinding_h", C.c_double),
        ("misbinding_s", C.c_double),
        ("stacking_ene", C.c_double),
        ("cation_M", C.c_double),
        ("staple_M", C.c_double),
    ]


class RunParams(C.Structure):
    _fields_ = [
        ("n_movetypes", C.c_int32),
        ("movetype_kind", _i32p),
        ("movetype_freq", _f64p),
        ("movetype_max_regrowth", _i32p),
        ("max_total_staples", C.c_int32),
        ("max_type_staples", C.c_int32),
        ("max_staple_size", C.c_int32),
        ("centering_freq", C.c_int32),
        ("centering_domain", C.c_int32),
        ("constraint_check_freq", C.c_int32),
        ("seed", C.c_uint64),
        ("rng_mode", C.c_int32),
        ("exchange_mode", C.c_int32),
        ("lanes_per_replica", C.c_int32),
        ("flags", C.c_int32),
        ("movetype_max_num_recoils", _i32p),
        ("movetype_max_c_attempts", _i32p),
        ("movetype_max_seg_regrowth", _i32p),
    ]


class Config(C.Structure):
    _fields_ = [
        ("n_chains", C.c_int32),
        ("pad0", C.c_int32),
        ("chain_index", _i32p),
        ("chain_identity", _i32p),
        ("chain_length", _i32p),
        ("positions", _i32p),
        ("orientations", _i32p),
    ]


class Obs(C.Structure):
    _fields_ = [
        ("energy", C.c_double),
        ("enthalpy", C.c_double),
        ("entropy", C.c_double),
        ("stacking_energy", C.c_double),
        ("temperature", C.c_double),
        ("step", C.c_int64),
        ("num_staples", C.c_int32),
        ("num_domains", C.c_int32),
        ("num_bound_domain_pairs", C.c_int32),
        ("num_fully_bound_domain_pairs", C.c_int32),
        ("num_misbound_domain_pairs", C.c_int32),
        ("num_stacked_domain_pairs", C.c_int32),
        ("num_unassigned_domains", C.c_int32),
        ("status", C.c_int32),
        ("bias", C.c_double),
        ("last_move", C.c_int32),
        ("last_accepted", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Window(C.Structure):
    _fields_ = [
        ("op", C.c_int32),
        ("min_op", C.c_int32),
        ("max_op", C.c_int32),
        ("pad0", C.c_int32),
        ("well_bias", C.c_double),
        ("min_bias", C.c_double),
        ("slope", C.c_double),
        ("off_grid_bias", C.c_double),
        ("bias_mult", C.c_double),
    ]


class SiteQuery(C.Structure):
    """ldo_site_query: a trial placement of one domain of one replica."""
    _fields_ = [("replica", C.c_int32), ("chain_index", C.c_int32), ("domain", C.c_int32),
                ("pos", C.c_int32 * 3), ("ore", C.c_int32 * 3), ("pad", C.c_int32)]


class SiteResult(C.Structure):
    _fields_ = [("unassign_delta", C.c_double), ("delta", C.c_double),
                ("violated", C.c_int32), ("pad", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [
        ("attempts", C.c_int64 * MAX_MOVETYPES),
        ("accepts", C.c_int64 * MAX_MOVETYPES),
        ("site_empty", C.c_int64),
        ("site_bind", C.c_int64),
        ("commits", C.c_int64),
        ("exchange_attempts", C.c_int64),
        ("exchange_accepts", C.c_int64),
    ]


class LdoError(RuntimeError):
    """A non-zero status from the sampler (or oracle) C ABI."""

    def __init__(self, code, message):
        super().__init__(f"status {code}: {message}")
        self.code = code


def tracking_len(n_types, n_scaffold, max_staples):
    """Length of the ldo_get_move_tracking array."""
    return 6 * n_types + 4 * (n_scaffold + 1) + 4 * (max_staples + 1) + 2


def split_tracking(values, n_types, n_scaffold, max_staples):
    """The ldo_get_move_tracking array as named per-index lists."""
    T, n1, s1 = n_types, n_scaffold + 1, max_staples + 1
    v = list(values)
    out, o = {}, 0
    for name, n in (("met_insert_attempts", T), ("met_insert_accepts", T),
                    ("met_delete_attempts", T), ("met_delete_accepts", T),
                    ("cb_attempts", T), ("cb_accepts", T),
                    ("ctcb_length_attempts", n1), ("ctcb_length_accepts", n1),
                    ("ctcb_staples_attempts", s1), ("ctcb_staples_accepts", s1),
                    ("ctrg_length_attempts", n1), ("ctrg_length_accepts", n1),
                    ("ctrg_jump_attempts", 1), ("ctrg_jump_accepts", 1),
                    ("ctcb_jump_staples_attempts", s1), ("ctcb_jump_staples_accepts", s1)):
        out[name] = v[o:o + n]
        o += n
    assert o == len(v), (o, len(v))
    return out


def site_queries(queries):
    """(replica, chain_index, domain, pos, ore) tuples as an ldo_site_query array."""
    arr = (SiteQuery * max(1, len(queries)))()
    for k, (r, c, d, pos, ore) in enumerate(queries):
        arr[k].replica, arr[k].chain_index, arr[k].domain = r, c, d
        arr[k].pos[:] = list(pos)
        arr[k].ore[:] = list(ore)
    return arr


def declare_abi(lib, prefix):
    """Attach argtypes/restype for the ldo_* (or oracle_*) entry points."""
    v = C.c_void_p
    sig = {
        "destroy": (None, [v]),
        "last_error": (C.c_char_p, [v]),
        "set_ladder": (C.c_int, [v, C.c_int32, _f64p, _f64p, _f64p, _f64p]),
        "set_ladder_2d": (C.c_int, [v, C.c_int32, _f64p, C.c_int32, _f64p, C.c_double]),
        "set_step": (C.c_int, [v, C.c_int64]),
        "exchange": (C.c_int, [v, C.c_int64]),
        "site_eval": (C.c_int, [v, C.POINTER(SiteQuery), C.c_int32, C.POINTER(SiteResult)]),
        "get_exchange_counts": (C.c_int, [v, C.c_int32, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]),
        "sweep": (C.c_int, [v, C.c_int64]),
        "run": (C.c_int, [v, C.c_int64, C.c_int64]),
        "get_observables": (C.c_int, [v, C.c_int32, C.POINTER(Obs)]),
        "get_counters": (C.c_int, [v, C.POINTER(Counters)]),
        "get_permutation": (C.c_int, [v, _i32p]),
        "get_config": (C.c_int, [v, C.c_int32, C.POINTER(Config), C.c_int32, C.c_int32]),
        "set_config": (C.c_int, [v, C.c_int32, C.POINTER(Config)]),
        "set_windows": (C.c_int, [v, C.c_int32, C.POINTER(Window), _f64p, C.c_int32]),
        "set_window_grid": (C.c_int, [v, _f64p]),
        "set_histogram": (C.c_int, [v, C.c_int32]),
        "get_histogram": (C.c_int, [v, C.POINTER(C.c_int64), C.c_int32]),
        "get_move_tracking": (C.c_int, [v, C.c_int32, C.POINTER(C.c_int64), C.c_int64,
                                        C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args


BATCH_LDS, BATCH_WAVE, BATCH_LANE, BATCH_WAVE_SOLO = 0, 1, 2, 3
# int32 values staged by ldo_set_config_noalloc (2 per chain + 6 per domain)
NOALLOC_CONFIG_INTS = 65536
BATCH_KERNEL_NAMES = {BATCH_LDS: "ldo_batch_sweep_lds_kernel",
                      BATCH_WAVE: "ldo_batch_sweep_kernel (one chain per wavefront)",
                      BATCH_LANE: "ldo_batch_sweep_kernel (one chain per lane)",
                      BATCH_WAVE_SOLO: "ldo_batch_sweep_kernel (one chain per wavefront, "
                                       "first lane only)"}


def declare_batch_abi(lib):
    """Attach argtypes/restype for the ldo_batch_* entry points."""
    v = C.c_void_p
    sig = {
        "create": (C.c_int, [C.c_int32, C.POINTER(SystemDesc), C.POINTER(RunParams),
                             C.POINTER(Config), _f64p, C.POINTER(C.c_void_p)]),
        "destroy": (None, [v]),
        "last_error": (C.c_char_p, [v]),
        "set_kernel": (C.c_int, [v, C.c_int32]),
        "tune": (C.c_int, [v, C.c_int64, _i32p, _f64p]),
        "sweep": (C.c_int, [v, C.c_int64, _f64p]),
        "get_observables": (C.c_int, [v, C.c_int32, C.c_int32, C.POINTER(Obs)]),
        "get_counters": (C.c_int, [v, C.POINTER(Counters)]),
        "get_config": (C.c_int, [v, C.c_int32, C.POINTER(Config), C.c_int32, C.c_int32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, "ldo_batch_" + name)
        fn.restype = res
        fn.argtypes = args


class MlpLayout(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("w1", "emb", "w1e", "b1", "w2", "b2", "w3", "b3", "w4",
                                         "b4", "total")]


MLP_SITES, MLP_CHANNELS, MLP_FEATURES, MLP_EMBED, MLP_HIDDEN, MLP_OUT = 343, 13, 4459, 16, 256, 36


def declare_mlp_abi(lib):
    """Attach argtypes/restype for the ldo_mlp_* entry points (ldo_mlp.h)."""
    v = C.c_void_p
    u8p = C.POINTER(C.c_uint8)
    f32p = C.POINTER(C.c_float)
    sig = {
        "layout_of": (C.c_int, [C.c_int32, C.POINTER(MlpLayout)]),
        "create": (C.c_int, [C.c_int32, C.c_int32, f32p, C.c_uint64, C.POINTER(C.c_void_p)]),
        "destroy": (None, [v]),
        "last_error": (C.c_char_p, [v]),
        "get_params": (C.c_int, [v, f32p]),
        "set_params": (C.c_int, [v, f32p]),
        "loss_grad": (C.c_int, [v, C.c_int32, u8p, _i32p, u8p, _i32p, _f64p, f32p]),
        "adam_step": (C.c_int, [v, C.c_float, C.c_float, C.c_float, C.c_float]),
        "log_probs": (C.c_int, [v, C.c_int32, u8p, _i32p, u8p, f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, "ldo_mlp_" + name)
        fn.restype = res
        fn.argtypes = args


def declare_learned_abi(lib, prefix):
    """argtypes/restype of the learned-move entry points (ldo_learned_*)."""
    u8p = C.POINTER(C.c_uint8)
    sig = {
        "learned_set_network": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_float),
                                          C.c_double]),
        "learned_sweep": (C.c_int, [C.c_void_p, C.c_int64]),
        "learned_samples": (C.c_int, [C.c_void_p, C.c_int32, u8p, _i32p, u8p, _i32p, _i32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args


_LIB = None


def load_library(path=None):
    """Load libldo_hip.so (the product).  Raises if it is not built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(
            f"{p} is missing: build the HIP sampler first (python -c "
            "'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    declare_abi(lib, "ldo_")
    lib.ldo_restart_swaps.restype = C.c_int
    lib.ldo_restart_swaps.argtypes = [C.c_void_p]
    lib.ldo_set_config_noalloc.restype = C.c_int
    lib.ldo_set_config_noalloc.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Config)]
    lib.ldo_create.restype = C.c_int
    lib.ldo_create.argtypes = [C.POINTER(SystemDesc), C.POINTER(RunParams),
                               C.POINTER(Config), C.c_int32, _f64p,
                               C.POINTER(C.c_void_p)]
    lib.ldo_set_tape.restype = C.c_int
    lib.ldo_set_tape.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64), C.c_int64]
    lib.ldo_get_stream.restype = C.c_void_p
    lib.ldo_get_stream.argtypes = [C.c_void_p]
    lib.ldo_set_permutation.restype = C.c_int
    lib.ldo_set_permutation.argtypes = [C.c_void_p, _i32p]
    lib.ldo_kernel_name.restype = C.c_char_p
    lib.ldo_kernel_name.argtypes = [C.c_void_p]
    lib.ldo_set_profiling.restype = C.c_int
    lib.ldo_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    lib.ldo_get_profile.restype = C.c_int
    lib.ldo_get_profile.argtypes = [C.c_void_p, _f64p, C.POINTER(C.c_int64), _f64p,
                                    C.POINTER(C.c_int64)]
    declare_batch_abi(lib)
    declare_mlp_abi(lib)
    declare_learned_abi(lib, "ldo_")
    if path is None:
        _LIB = lib
    return lib
