"""The learned proposal's network on the device (the ldo_mlp_* C ABI)."""
import ctypes as C

import numpy as np

from . import abi


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Mlp:
    """A network with `n_ident` domain identities; batches up to `max_batch`.
    `params` (flat float32, ldo_mlp_layout order) or a seeded initialisation."""

    def __init__(self, n_ident, max_batch=256, params=None, seed=1, lib=None):
        self.lib = lib if lib is not None else abi.load_library()
        self.n_ident = n_ident
        lay = abi.MlpLayout()
        rc = self.lib.ldo_mlp_layout_of(n_ident, C.byref(lay))
        if rc != 0:
            raise abi.LdoError(rc, "bad n_ident")
        self.layout = {n: getattr(lay, n) for n, _ in abi.MlpLayout._fields_}
        p = None
        if params is not None:
            self._init = np.ascontiguousarray(params, dtype=np.float32)
            p = _ptr(self._init, C.c_float)
        h = C.c_void_p()
        rc = self.lib.ldo_mlp_create(n_ident, max_batch, p, seed, C.byref(h))
        if rc != 0:
            raise abi.LdoError(rc, self.lib.ldo_mlp_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise abi.LdoError(rc, self.lib.ldo_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ldo_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.layout["total"], dtype=np.float32)
        self._check(self.lib.ldo_mlp_get_params(self.h, _ptr(out, C.c_float)))
        return out

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.ldo_mlp_set_params(self.h, _ptr(a, C.c_float)))

    @staticmethod
    def _batch(channels, ident, legal, target=None):
        ch = np.ascontiguousarray(channels, dtype=np.uint8)
        idn = np.ascontiguousarray(ident, dtype=np.int32)
        lg = np.ascontiguousarray(legal, dtype=np.uint8)
        tg = None if target is None else np.ascontiguousarray(target, dtype=np.int32)
        return ch, idn, lg, tg

    def loss_grad(self, channels, ident, legal, target, want_grads=False):
        ch, idn, lg, tg = self._batch(channels, ident, legal, target)
        loss = C.c_double()
        g = np.zeros(self.layout["total"], dtype=np.float32) if want_grads else None
        self._check(self.lib.ldo_mlp_loss_grad(
            self.h, len(idn), _ptr(ch, C.c_uint8), _ptr(idn, C.c_int32), _ptr(lg, C.c_uint8),
            _ptr(tg, C.c_int32), C.byref(loss), _ptr(g, C.c_float) if want_grads else None))
        return (loss.value, g) if want_grads else loss.value

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self._check(self.lib.ldo_mlp_adam_step(self.h, lr, beta1, beta2, eps))

    def log_probs(self, channels, ident, legal):
        ch, idn, lg, _ = self._batch(channels, ident, legal)
        out = np.zeros((len(idn), abi.MLP_OUT), dtype=np.float32)
        self._check(self.lib.ldo_mlp_log_probs(self.h, len(idn), _ptr(ch, C.c_uint8),
                                               _ptr(idn, C.c_int32), _ptr(lg, C.c_uint8),
                                               _ptr(out, C.c_float)))
        return out
