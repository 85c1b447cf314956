"""Python driver over the C ABI of libldo_hip.so.

`Sampler` mirrors the reference's simulation drivers for the hot path: a
constant-T run is `sweep(steps)` (ConstantTGCMCSimulation::run,
constant_temp_simulation.hpp:31) and a temperature-exchange run is
`run(n_swaps, interval)` (PTGCMCSimulation::run, ptmc_simulation.cpp:106-150).
All replicas live on the GPU; every call is synchronous.
"""
import ctypes as C

from . import abi
from .systems import config_from_abi, config_to_abi, empty_config


class Sampler:
    def __init__(self, system, temps=None, rng_mode=abi.RNG_PHILOX,
                 exchange_mode=abi.EXCHANGE_NONE, lanes_per_replica=0, lib=None):
        self.lib = lib if lib is not None else abi.load_library()
        self.system = system
        temps = list(temps if temps is not None else (system.temps or [300.0]))
        self.n = len(temps)
        sd, rp, cfg, self._keep = system.to_abi(rng_mode, exchange_mode, lanes_per_replica)
        t = (C.c_double * self.n)(*temps)
        h = C.c_void_p()
        rc = self.lib.ldo_create(C.byref(sd), C.byref(rp), C.byref(cfg), self.n, t, C.byref(h))
        if rc != 0:
            raise abi.LdoError(rc, self.lib.ldo_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise abi.LdoError(rc, self.lib.ldo_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ldo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sweep(self, steps):
        self._check(self.lib.ldo_sweep(self.h, steps))

    def run(self, n_swaps, interval):
        self._check(self.lib.ldo_run(self.h, n_swaps, interval))

    def restart_swaps(self):
        """The next exchange round pairs as the first of a run (swap_i = 1)."""
        self._check(self.lib.ldo_restart_swaps(self.h))

    def set_ladder(self, temps, u=None, b=None, s=None):
        n = len(temps)
        one = [1.0] * n
        arrs = [(C.c_double * n)(*v) for v in (temps, u or one, b or one, s or one)]
        self._check(self.lib.ldo_set_ladder(self.h, n, *arrs))

    def exchange(self, swap_index):
        """One exchange round with pair set swap_index on the current configurations."""
        self._check(self.lib.ldo_exchange(self.h, swap_index))

    def site_eval(self, queries):
        """Site evaluations of (replica, chain_index, domain, pos, ore) tuples:
        a list of (unassign_delta, delta, violated)."""
        arr = abi.site_queries(queries)
        out = (abi.SiteResult * max(1, len(queries)))()
        self._check(self.lib.ldo_site_eval(self.h, arr, len(queries), out))
        return [(out[k].unassign_delta, out[k].delta, out[k].violated)
                for k in range(len(queries))]

    def set_step(self, step):
        """Every replica's next MC step is step + 1."""
        self._check(self.lib.ldo_set_step(self.h, step))

    def set_ladder_2d(self, temps, stacking_mults, staple_u_mult=1.0):
        """The temperature x stacking-multiplier grid of 2d_parallel_tempering
        (slot i * len(stacking_mults) + j)."""
        t = (C.c_double * len(temps))(*temps)
        s = (C.c_double * len(stacking_mults))(*stacking_mults)
        self._check(self.lib.ldo_set_ladder_2d(self.h, len(temps), t, len(stacking_mults), s,
                                               staple_u_mult))

    def exchange_counts(self):
        """Per slot pair exchange (attempts, accepts), each [2][n_replicas]:
        [d][q] is the pair of slot q with its neighbour in direction d."""
        import numpy as np
        n = 2 * self.n
        att = np.zeros(n, dtype=np.int64)
        acc = np.zeros(n, dtype=np.int64)
        self._check(self.lib.ldo_get_exchange_counts(
            self.h, n, att.ctypes.data_as(C.POINTER(C.c_int64)),
            acc.ctypes.data_as(C.POINTER(C.c_int64))))
        return att.reshape(2, self.n), acc.reshape(2, self.n)

    def set_windows(self, windows, grid, grid_len):
        """Umbrella windows, one per replica slot (abi.Window records) and
        their grid biases grid[w * grid_len + p]."""
        n = len(windows)
        arr = (abi.Window * n)(*windows)
        g = (C.c_double * max(1, len(grid)))(*grid)
        self._check(self.lib.ldo_set_windows(self.h, n, arr, g, grid_len))
        self._grid_len = grid_len

    def _fn(self, name):
        return getattr(self.lib, "ldo_" + name)

    def _check_call(self, rc):
        self._check(rc)

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

    def set_permutation(self, q_to_rep):
        p = (C.c_int32 * self.n)(*q_to_rep)
        self._check(self.lib.ldo_set_permutation(self.h, p))

    def learned_set_network(self, n_ident, params, fallback=0.1):
        """Enable learned domain regrowth with these network parameters
        (ldo_mlp.h layout); `fallback` is the plain CB probability."""
        import numpy as np
        self._net = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.ldo_learned_set_network(
            self.h, n_ident, self._net.ctypes.data_as(C.POINTER(C.c_float)), fallback))

    def learned_sweep(self, attempts):
        self._check(self.lib.ldo_learned_sweep(self.h, attempts))

    def learned_samples(self, cap):
        """Training samples of the current configurations:
        (channels [n, 343], ident [n], legal [n, 36], target [n])."""
        import numpy as np
        ch = np.zeros((max(1, cap), 343), dtype=np.uint8)
        lg = np.zeros((max(1, cap), 36), dtype=np.uint8)
        idn = np.zeros(max(1, cap), dtype=np.int32)
        tg = np.zeros(max(1, cap), dtype=np.int32)
        n = C.c_int32()
        u8 = C.POINTER(C.c_uint8)
        i32 = C.POINTER(C.c_int32)
        self._check(self.lib.ldo_learned_samples(
            self.h, cap, ch.ctypes.data_as(u8), idn.ctypes.data_as(i32), lg.ctypes.data_as(u8),
            tg.ctypes.data_as(i32), C.byref(n)))
        k = n.value
        return ch[:k], idn[:k], lg[:k], tg[:k]

    def set_tape(self, r, values):
        buf = (C.c_uint64 * max(1, len(values)))(*values)
        self._check(self.lib.ldo_set_tape(self.h, r, buf, len(values)))

    def observables(self, r=0):
        o = abi.Obs()
        self._check(self.lib.ldo_get_observables(self.h, r, C.byref(o)))
        return o.as_dict()

    def counters(self):
        c = abi.Counters()
        self._check(self.lib.ldo_get_counters(self.h, C.byref(c)))
        return c

    def move_tracking(self, r=0):
        """Per-move tracking counts of replica r (ldo_get_move_tracking)."""
        cap = abi.tracking_len(len(self.system.identities), len(self.system.identities[0]),
                               self.system.max_total_staples)
        buf = (C.c_int64 * cap)()
        n = C.c_int64()
        self._check(self.lib.ldo_get_move_tracking(self.h, r, buf, cap, C.byref(n)))
        return abi.split_tracking(buf[:n.value], len(self.system.identities),
                                  len(self.system.identities[0]), self.system.max_total_staples)

    def permutation(self):
        p = (C.c_int32 * self.n)()
        self._check(self.lib.ldo_get_permutation(self.h, p))
        return list(p)

    def config(self, r=0):
        cap_c = self.system.max_total_staples + 1
        cap_d = self.system.max_domains()
        cfg, keep = empty_config(cap_c, cap_d)
        self._check(self.lib.ldo_get_config(self.h, r, C.byref(cfg), cap_c, cap_d))
        return config_from_abi(cfg)

    def set_config(self, r, chains):
        keep = []
        cfg = config_to_abi(chains, keep)
        self._check(self.lib.ldo_set_config(self.h, r, C.byref(cfg)))

    def set_config_noalloc(self, r, chains):
        """set_config through ldo_set_config_noalloc (no device allocation)."""
        keep = []
        cfg = config_to_abi(chains, keep)
        self._check(self.lib.ldo_set_config_noalloc(self.h, r, C.byref(cfg)))

    def stream(self):
        return self.lib.ldo_get_stream(self.h)

    def set_profiling(self, on=True):
        """Record HIP events around every sweep and exchange kernel launch."""
        self._check(self.lib.ldo_set_profiling(self.h, int(on)))

    def profile(self):
        """Summed kernel times (ms) and launch counts since profiling started."""
        sm, en = C.c_double(), C.c_double()
        sn, xn = C.c_int64(), C.c_int64()
        self._check(self.lib.ldo_get_profile(self.h, C.byref(sm), C.byref(sn), C.byref(en),
                                             C.byref(xn)))
        return {"sweep_ms": sm.value, "sweep_launches": sn.value,
                "exchange_ms": en.value, "exchange_launches": xn.value}


class BatchSampler:
    """Many independent constant-T chains, each a system of its own, advanced
    by one kernel launch per sweep (the batch C ABI, ldo_batch_*).  Chain k
    follows exactly the chain Sampler(systems[k], temps=[temps[k]]) follows."""

    def __init__(self, systems, temps, lib=None):
        if len(systems) != len(temps) or not systems:
            raise ValueError("one temperature per system, at least one system")
        self.lib = lib if lib is not None else abi.load_library()
        self.systems = list(systems)
        n = self.n = len(systems)
        packed = [s.to_abi() for s in systems]
        self._keep = [p[3] for p in packed]
        sds = (abi.SystemDesc * n)(*[p[0] for p in packed])
        rps = (abi.RunParams * n)(*[p[1] for p in packed])
        cfgs = (abi.Config * n)(*[p[2] for p in packed])
        self._keep += [sds, rps, cfgs]
        t = (C.c_double * n)(*temps)
        h = C.c_void_p()
        rc = self.lib.ldo_batch_create(n, sds, rps, cfgs, t, C.byref(h))
        if rc != 0:
            raise abi.LdoError(rc, self.lib.ldo_batch_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise abi.LdoError(rc, self.lib.ldo_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ldo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, steps_per_trial):
        """Time each available sweep kernel on real steps; keep the fastest.
        Returns (chosen kernel, {kernel: ms or None})."""
        chosen = C.c_int32()
        n = len(abi.BATCH_KERNEL_NAMES)
        ms = (C.c_double * n)()
        self._check(self.lib.ldo_batch_tune(self.h, steps_per_trial, C.byref(chosen), ms))
        return chosen.value, {k: (ms[k] if ms[k] >= 0 else None) for k in range(n)}

    def set_kernel(self, kernel):
        self._check(self.lib.ldo_batch_set_kernel(self.h, kernel))

    def sweep(self, steps):
        """`steps` MC steps on every chain; returns the kernel milliseconds."""
        ms = C.c_double()
        self._check(self.lib.ldo_batch_sweep(self.h, steps, C.byref(ms)))
        return ms.value

    def observables(self, first=0, n=None):
        n = self.n - first if n is None else n
        out = (abi.Obs * max(1, n))()
        self._check(self.lib.ldo_batch_get_observables(self.h, first, n, out))
        return [out[i].as_dict() for i in range(n)]

    def counters(self):
        c = abi.Counters()
        self._check(self.lib.ldo_batch_get_counters(self.h, C.byref(c)))
        return c

    def config(self, k):
        s = self.systems[k]
        cfg, keep = empty_config(s.max_total_staples + 1, s.max_domains())
        self._check(self.lib.ldo_batch_get_config(self.h, k, C.byref(cfg), s.max_total_staples + 1,
                                                  s.max_domains()))
        return config_from_abi(cfg)
