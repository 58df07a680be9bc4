"""wg_mlp_forward_noalloc (activations in the library's static scratch array)
against wg_mlp_forward, through the C ABI: bit-identical outputs on two calls in
a row, the second with fewer rows, on a caller's stream and on the handle's own,
nothing written past the last row, and sizes outside 1..WG_MLP_NOALLOC_MAX_ROWS
refused without device work.  Device buffers and the stream come from the HIP
runtime the library itself links (libamdhip64), called through ctypes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ROWS = 8192  # WG_MLP_NOALLOC_MAX_ROWS
WG_E_ARG = -1
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.lib.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.lib.hipStreamDestroy.argtypes = [C.c_void_p]
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        self.bufs.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.check(self.lib.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy")
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        self.check(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def test_forward_noalloc_matches_forward(wg):
    lib = wg._abi.load_library()
    hip = Hip()
    h = C.c_void_p()
    assert lib.wg_mlp_create(C.c_uint64(11), C.byref(h)) == 0
    stream = C.c_void_p()
    hip.check(hip.lib.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    try:
        rng = np.random.default_rng(4)
        params = (rng.standard_normal(lib.wg_mlp_n_params()) * 0.08).astype(np.float32)
        assert lib.wg_mlp_set_params(h, params.ctypes.data_as(wg._abi.P_f)) == 0
        for rows, on in ((MAX_ROWS, stream), (3000, stream), (77, None)):  # largest first, then smaller
            print(f"case: wg_mlp_forward_noalloc rows={rows} stream={'caller' if on else 'handle'}", flush=True)
            x = rng.standard_normal((rows, 16)).astype(np.float32)
            y_ref = np.zeros(rows, np.float32)
            assert lib.wg_mlp_forward(h, rows, x.ctypes.data_as(wg._abi.P_f), y_ref.ctypes.data_as(wg._abi.P_f)) == 0
            xd = hip.put(x)
            yd = hip.put(np.full(rows + 64, np.nan, np.float32))
            rc = lib.wg_mlp_forward_noalloc(h, rows, xd, yd, on)
            assert rc == 0, lib.wg_mlp_last_error(h)
            if on is not None:
                hip.check(hip.lib.hipStreamSynchronize(on), "hipStreamSynchronize")
            else:
                out = np.zeros(lib.wg_mlp_n_params(), np.float32)
                assert lib.wg_mlp_get_params(h, out.ctypes.data_as(wg._abi.P_f)) == 0  # waits for the handle's stream
            y = hip.get(yd, rows + 64, np.float32)
            assert np.array_equal(y[:rows].view(np.uint32), y_ref.view(np.uint32))
            assert np.isnan(y[rows:]).all()  # nothing past the last row
            hip.free()
        # sizes beyond the scratch array, or empty, are refused before any launch
        for bad in (MAX_ROWS + 1, 0, -5):
            assert lib.wg_mlp_forward_noalloc(h, bad, C.c_void_p(16), C.c_void_p(16), stream) == WG_E_ARG
        assert b"rows must be 1..8192" in lib.wg_mlp_last_error(h)
    finally:
        hip.free()
        hip.lib.hipStreamDestroy(stream)
        assert lib.wg_mlp_destroy(h) == 0
