"""ctypes binding of the CPU oracle (oracle/_build/libldo_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

from latticednaorigami_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libldo_oracle.so")
REF_NN = os.path.join(HERE, "_ref", "libref_nn.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        abi.declare_abi(L, "oracle_")
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(abi.SystemDesc), C.POINTER(abi.RunParams),
                                    C.POINTER(abi.Config), C.c_int32,
                                    C.POINTER(C.c_double), C.c_int32]
        L.oracle_record_tape.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.oracle_tape_length.restype = C.c_int64
        L.oracle_tape_length.argtypes = [C.c_void_p, C.c_int32]
        L.oracle_get_tape.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64), C.c_int64]
        L.oracle_energy_tables.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32,
                                           C.c_int32, C.POINTER(C.c_double)]
        L.oracle_recompute_energy.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        L.oracle_energy_in_order.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                             C.c_int32, C.POINTER(C.c_double),
                                             C.POINTER(C.c_int64)]
        L.oracle_nn_hybridization_thermo.argtypes = [C.c_char_p, C.c_double, C.c_double,
                                                     C.POINTER(C.c_double),
                                                     C.POINTER(C.c_double)]
        L.oracle_nn_reduced_energy.restype = C.c_double
        L.oracle_nn_reduced_energy.argtypes = [C.c_char_p, C.c_double, C.c_double]
        L.oracle_nn_longest_contig_complement.argtypes = [C.c_char_p, C.c_char_p,
                                                          C.c_char_p, C.c_int32]
        L.oracle_num_walks.restype = C.c_double
        L.oracle_num_walks.argtypes = [C.c_int32] * 4
        L.oracle_rotate_half.argtypes = [C.POINTER(C.c_int32)] * 3
        L.oracle_rotate.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                    C.POINTER(C.c_int32)]
        L.oracle_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.oracle_enumerate.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_int32,
                                       C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        u8p = C.POINTER(C.c_uint8)
        i32p = C.POINTER(C.c_int32)
        L.oracle_learned_set_network.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_float),
                                                 C.c_double]
        L.oracle_learned_sweep.argtypes = [C.c_void_p, C.c_int64]
        L.oracle_learned_samples.argtypes = [C.c_void_p, C.c_int32, u8p, i32p, u8p, i32p, i32p]
        _lib = L
    return _lib


def _check(L, h, rc):
    if rc != 0:
        raise abi.LdoError(rc, L.oracle_last_error(h).decode())


class Oracle:
    """The oracle's replicas of one system (rng_kind 0: Philox stream of the
    sampler; 1: the reference's mt19937_64 RandomGens)."""

    def __init__(self, system, temps=None, rng_kind=0, exchange_mode=abi.EXCHANGE_NONE):
        self.L = lib()
        self.system = system
        temps = list(temps if temps is not None else (system.temps or [300.0]))
        self.n = len(temps)
        sd, rp, cfg, self._keep = system.to_abi(exchange_mode=exchange_mode)
        t = (C.c_double * self.n)(*temps)
        h = self.L.oracle_create(C.byref(sd), C.byref(rp), C.byref(cfg), self.n, t, rng_kind)
        if not h:
            raise abi.LdoError(abi.LDO_E_MISUSE, self.L.oracle_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sweep(self, steps):
        _check(self.L, self.h, self.L.oracle_sweep(self.h, steps))

    def set_permutation(self, q_to_rep):
        p = (C.c_int32 * self.n)(*q_to_rep)
        self.L.oracle_set_permutation.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        _check(self.L, self.h, self.L.oracle_set_permutation(self.h, p))

    def learned_set_network(self, n_ident, params, fallback=0.1):
        import numpy as np
        self._net = np.ascontiguousarray(params, dtype=np.float32)
        _check(self.L, self.h, self.L.oracle_learned_set_network(
            self.h, n_ident, self._net.ctypes.data_as(C.POINTER(C.c_float)), fallback))

    def learned_sweep(self, attempts):
        _check(self.L, self.h, self.L.oracle_learned_sweep(self.h, attempts))

    def learned_samples(self, cap):
        """(channels [n, 343], ident [n], legal [n, 36], target [n])"""
        import numpy as np
        ch = np.zeros((max(1, cap), 343), dtype=np.uint8)
        lg = np.zeros((max(1, cap), 36), dtype=np.uint8)
        idn = np.zeros(max(1, cap), dtype=np.int32)
        tg = np.zeros(max(1, cap), dtype=np.int32)
        n = C.c_int32()
        u8 = C.POINTER(C.c_uint8)
        i32 = C.POINTER(C.c_int32)
        _check(self.L, self.h, self.L.oracle_learned_samples(
            self.h, cap, ch.ctypes.data_as(u8), idn.ctypes.data_as(i32), lg.ctypes.data_as(u8),
            tg.ctypes.data_as(i32), C.byref(n)))
        k = n.value
        return ch[:k], idn[:k], lg[:k], tg[:k]

    def run(self, n_swaps, interval):
        _check(self.L, self.h, self.L.oracle_run(self.h, n_swaps, interval))

    def set_windows(self, windows, grid, grid_len):
        n = len(windows)
        arr = (abi.Window * n)(*windows)
        g = (C.c_double * max(1, len(grid)))(*grid)
        _check(self.L, self.h, self.L.oracle_set_windows(self.h, n, arr, g, grid_len))
        self._grid_len = grid_len

    def _fn(self, name):
        return getattr(self.L, "oracle_" + name)

    def _check_call(self, rc):
        _check(self.L, self.h, rc)

    def set_window_grid(self, grid):
        """New grid biases for every window; the stored bias values stay."""
        g = (C.c_double * max(1, len(grid)))(*grid)
        self._check_call(self._fn("set_window_grid")(self.h, g))

    def set_histogram(self, on=True):
        self._check_call(self._fn("set_histogram")(self.h, 1 if on else 0))

    def histogram(self, reset=False):
        """Visits [n_windows, grid_len + 1] of each window's order parameter
        (last column: off the grid) since counting started or was reset."""
        import numpy as np
        out = np.zeros((self.n, self._grid_len + 1), dtype=np.int64)
        self._check_call(self._fn("get_histogram")(
            self.h, out.ctypes.data_as(C.POINTER(C.c_int64)), 1 if reset else 0))
        return out

    def set_ladder(self, temps, u=None, b=None, s=None):
        n = len(temps)
        one = [1.0] * n
        arrs = [(C.c_double * n)(*v) for v in (temps, u or one, b or one, s or one)]
        _check(self.L, self.h, self.L.oracle_set_ladder(self.h, n, *arrs))

    def exchange(self, swap_index):
        """One exchange round with pair set swap_index on the current configurations."""
        _check(self.L, self.h, self.L.oracle_exchange(self.h, swap_index))

    def site_eval(self, queries):
        """Site evaluations of (replica, chain_index, domain, pos, ore) tuples:
        a list of (unassign_delta, delta, violated)."""
        arr = abi.site_queries(queries)
        out = (abi.SiteResult * max(1, len(queries)))()
        _check(self.L, self.h, self.L.oracle_site_eval(self.h, arr, len(queries), out))
        return [(out[k].unassign_delta, out[k].delta, out[k].violated)
                for k in range(len(queries))]

    def set_step(self, step):
        _check(self.L, self.h, self.L.oracle_set_step(self.h, step))

    def set_ladder_2d(self, temps, stacking_mults, staple_u_mult=1.0):
        t = (C.c_double * len(temps))(*temps)
        s = (C.c_double * len(stacking_mults))(*stacking_mults)
        _check(self.L, self.h, self.L.oracle_set_ladder_2d(
            self.h, len(temps), t, len(stacking_mults), s, staple_u_mult))

    def exchange_counts(self):
        import numpy as np
        n = 2 * self.n
        att = np.zeros(n, dtype=np.int64)
        acc = np.zeros(n, dtype=np.int64)
        _check(self.L, self.h, self.L.oracle_get_exchange_counts(
            self.h, n, att.ctypes.data_as(C.POINTER(C.c_int64)),
            acc.ctypes.data_as(C.POINTER(C.c_int64))))
        return att.reshape(2, self.n), acc.reshape(2, self.n)

    def observables(self, r=0):
        o = abi.Obs()
        _check(self.L, self.h, self.L.oracle_get_observables(self.h, r, C.byref(o)))
        return o.as_dict()

    def counters(self):
        c = abi.Counters()
        _check(self.L, self.h, self.L.oracle_get_counters(self.h, C.byref(c)))
        return c

    def move_tracking(self, r=0):
        cap = abi.tracking_len(len(self.system.identities), len(self.system.identities[0]),
                               self.system.max_total_staples)
        buf = (C.c_int64 * cap)()
        n = C.c_int64()
        _check(self.L, self.h, self.L.oracle_get_move_tracking(self.h, r, buf, cap, C.byref(n)))
        return abi.split_tracking(buf[:n.value], len(self.system.identities),
                                  len(self.system.identities[0]), self.system.max_total_staples)

    def permutation(self):
        p = (C.c_int32 * self.n)()
        _check(self.L, self.h, self.L.oracle_get_permutation(self.h, p))
        return list(p)

    def config(self, r=0):
        from latticednaorigami_amd.systems import config_from_abi, empty_config
        cap_c = self.system.max_total_staples + 1
        cap_d = self.system.max_domains()
        cfg, keep = empty_config(cap_c, cap_d)
        _check(self.L, self.h, self.L.oracle_get_config(self.h, r, C.byref(cfg), cap_c, cap_d))
        return config_from_abi(cfg)

    def enumerate(self, all_orientations=False, cap_rows=4096):
        """Exact state probabilities {(numstaples, numfulldomains, nummisdomains,
        numstackedpairs): p} at the first ladder temperature, with ln Z and the
        number of configurations counted."""
        rows = (C.c_double * (5 * cap_rows))()
        n = C.c_int32()
        lz, nc = C.c_double(), C.c_double()
        _check(self.L, self.h, self.L.oracle_enumerate(self.h, int(all_orientations), rows,
                                                       cap_rows, C.byref(n), C.byref(lz),
                                                       C.byref(nc)))
        probs = {tuple(int(rows[5 * k + q]) for q in range(4)): rows[5 * k + 4]
                 for k in range(n.value)}
        return probs, lz.value, nc.value

    def set_config(self, r, chains):
        from latticednaorigami_amd.systems import config_to_abi
        keep = []
        cfg = config_to_abi(chains, keep)
        _check(self.L, self.h, self.L.oracle_set_config(self.h, r, C.byref(cfg)))

    def restart_swaps(self):
        """The next exchange round pairs as the first of a run."""
        self.L.oracle_restart_swaps.argtypes = [C.c_void_p]
        _check(self.L, self.h, self.L.oracle_restart_swaps(self.h))

    def set_tape(self, r, values):
        """Replica r replays these draws (doubles as bit patterns)."""
        buf = (C.c_uint64 * max(1, len(values)))(*values)
        self.L.oracle_set_tape.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64), C.c_int64]
        _check(self.L, self.h, self.L.oracle_set_tape(self.h, r, buf, len(values)))

    def record_tape(self, r, on=True):
        _check(self.L, self.h, self.L.oracle_record_tape(self.h, r, int(on)))

    def tape(self, r):
        n = self.L.oracle_tape_length(self.h, r)
        buf = (C.c_uint64 * max(1, n))()
        _check(self.L, self.h, self.L.oracle_get_tape(self.h, r, buf, n))
        return list(buf)[:n]

    def tables(self, temp, ii, ij, stacking_mult=1.0):
        out = (C.c_double * 5)()
        _check(self.L, self.h, self.L.oracle_energy_tables(self.h, temp, stacking_mult, ii, ij,
                                                           out))
        return dict(zip(("hyb_e", "hyb_h", "hyb_s", "stack", "init_e"), list(out)))

    def recompute_energy(self, r=0):
        e = C.c_double()
        _check(self.L, self.h, self.L.oracle_recompute_energy(self.h, r, C.byref(e)))
        return e.value

    def energy_in_order(self, order, r=0):
        """(energy, stacked pairs) rebuilt by assigning the domains in `order`
        (indices into the chain-order list of all domains)."""
        arr = (C.c_int32 * len(order))(*order)
        e, st = C.c_double(), C.c_int64()
        _check(self.L, self.h, self.L.oracle_energy_in_order(self.h, r, arr, len(order),
                                                             C.byref(e), C.byref(st)))
        return e.value, st.value


def nn_thermo(seq, temp, cation):
    h, s = C.c_double(), C.c_double()
    rc = lib().oracle_nn_hybridization_thermo(seq.encode(), temp, cation, C.byref(h), C.byref(s))
    if rc:
        raise ValueError(seq)
    return h.value, s.value


def nn_reduced_energy(seq, temp, cation):
    return lib().oracle_nn_reduced_energy(seq.encode(), temp, cation)


def longest_contig_complement(a, b):
    buf = C.create_string_buffer(4096)
    lib().oracle_nn_longest_contig_complement(a.encode(), b.encode(), buf, 4096)
    v = buf.value.decode()
    return v.split(" ") if v else []


def num_walks(d, steps):
    return lib().oracle_num_walks(d[0], d[1], d[2], steps)


def rotate_half(v, axis):
    a = (C.c_int32 * 3)(*v)
    b = (C.c_int32 * 3)(*axis)
    o = (C.c_int32 * 3)()
    lib().oracle_rotate_half(a, b, o)
    return tuple(o)


def rotate(v, axis, turns):
    a = (C.c_int32 * 3)(*v)
    b = (C.c_int32 * 3)(*axis)
    o = (C.c_int32 * 3)()
    lib().oracle_rotate(a, b, turns, o)
    return tuple(o)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return tuple(o)


def ref_nn_lib():
    """The reference's own nearest_neighbour.cpp (oracle/_ref), or None."""
    if not os.path.exists(REF_NN):
        return None
    L = C.CDLL(REF_NN)
    L.ref_nn_hybridization_thermo.argtypes = [C.c_char_p, C.c_double, C.c_double,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ref_nn_init_thermo.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ref_nn_longest_contig_complement.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    L.ref_nn_stacking_energy.restype = C.c_double
    L.ref_nn_stacking_energy.argtypes = [C.c_char_p, C.c_char_p]
    return L
