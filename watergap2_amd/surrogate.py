"""The calibration surrogate: a 16-128-128-128-1 GELU perceptron that maps basin
features and a (gamma, CFA) pair to the log of the long-term mean discharge,
trained and evaluated on the GPU by the library's BF16 MFMA kernels
(``wg_mlp_*`` in include/watergap2_amd.h).

It stands in for the search of calibGammaClass::findNewGamma (calibration.cpp:
266-636): the network proposes each basin's gamma, the plain model verifies the
proposal against the 1 % criterion, and the reference's search runs for the
basins the proposal misses.
"""
import ctypes as C

import numpy as np

from . import _abi

N_FEATURES = 16


class MLP:
    """Handle of one network on the GPU."""

    def __init__(self, seed=0):
        self._lib = _abi.load_library()
        h = C.c_void_p()
        rc = self._lib.wg_mlp_create(int(seed), C.byref(h))
        if rc != 0:
            msg = (self._lib.wg_mlp_last_error(h) or b"").decode() if h else ""
            raise RuntimeError(f"wg_mlp_create failed ({rc}): {msg}")
        self._h = h
        self.n_params = int(self._lib.wg_mlp_n_params())

    def _c(self, rc):
        if rc != 0:
            raise RuntimeError(f"wg_mlp call failed ({rc}): {(self._lib.wg_mlp_last_error(self._h) or b'').decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wg_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        self._c(self._lib.wg_mlp_get_params(self._h, _abi.ptr(p, _abi.P_f)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self._c(self._lib.wg_mlp_set_params(self._h, _abi.ptr(p, _abi.P_f)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.ndim == 2 and x.shape[1] == N_FEATURES
        y = np.empty(len(x), dtype=np.float32)
        self._c(self._lib.wg_mlp_forward(self._h, len(x), _abi.ptr(x, _abi.P_f), _abi.ptr(y, _abi.P_f)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._c(self._lib.wg_mlp_gradient(self._h, len(x), _abi.ptr(x, _abi.P_f), _abi.ptr(t, _abi.P_f),
                                          _abi.ptr(loss, _abi.P_f), _abi.ptr(g, _abi.P_f)))
        return float(loss[0]), g

    def train(self, x, t, epochs=20, batch=1024, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        self._c(self._lib.wg_mlp_train(self._h, len(x), _abi.ptr(x, _abi.P_f), _abi.ptr(t, _abi.P_f), epochs, batch,
                                       lr, int(seed), _abi.ptr(losses, _abi.P_f)))
        return losses
