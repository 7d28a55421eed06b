import ctypes
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)


class Oracle:
    """ctypes view of oracle/liboracle.so (the checker)."""

    def __init__(self, path):
        L = ctypes.CDLL(str(path))
        L.orc_prime.restype = ctypes.c_uint32
        L.orc_generator.restype = ctypes.c_uint32
        L.orc_root.restype = ctypes.c_uint32
        L.orc_root.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_zeta_table.argtypes = [ctypes.c_int, ctypes.c_int, _u32p, _u32p]
        L.orc_dft_textbook.argtypes = [_u32p, _u32p, ctypes.c_int, ctypes.c_int]
        for f in (L.orc_ntt_ct_layers, L.orc_ntt_gs_layers):
            f.argtypes = [_u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_crt3.argtypes = [_u32p, _u32p]
        L.orc_crt3_set.argtypes = [ctypes.c_int, _u32p, _u32p]
        L.orc_set_for.argtypes = [ctypes.c_int]
        L.orc_shoup_quotient.restype = ctypes.c_uint32
        L.orc_shoup_quotient.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        _pu32 = ctypes.POINTER(ctypes.c_uint32)
        for f in (L.orc_narrow_ct, L.orc_narrow_ct_lazy, L.orc_narrow_gs, L.orc_narrow_gs_lazy):
            f.argtypes = [_pu32, _pu32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_narrow_pointwise.restype = ctypes.c_uint32
        L.orc_narrow_pointwise.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_narrow_pointwise_shared.restype = ctypes.c_uint32
        L.orc_narrow_pointwise_shared.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_narrow_shared_operand.restype = ctypes.c_uint32
        L.orc_narrow_shared_operand.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_narrow_load.restype = ctypes.c_uint32
        L.orc_narrow_load.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        for f in (L.orc_narrow_ct_layers, L.orc_narrow_gs_layers):
            f.argtypes = [_u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_ntt_mul.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p]
        L.orc_gmp_mul.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p]
        L.orc_conv_residues.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_int, _u32p]
        L.orc_mlp_param_count.restype = ctypes.c_size_t
        L.orc_mlp_param_count.argtypes = [ctypes.c_int]
        L.orc_mlp_forward.argtypes = [_f64p, ctypes.c_int, _f64p, ctypes.c_size_t, _f64p]
        L.orc_mlp_loss.restype = ctypes.c_double
        L.orc_mlp_loss.argtypes = [_f64p, ctypes.c_int, _f64p, _i32p, _f64p, ctypes.c_size_t]
        L.orc_mlp_gradient.restype = ctypes.c_double
        L.orc_mlp_gradient.argtypes = [_f64p, ctypes.c_int, _f64p, _i32p, _f64p, ctypes.c_size_t, _f64p]
        L.orc_mlp_adam.argtypes = [_f64p, _f64p, _f64p, _f64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double]
        self.L = L

    @staticmethod
    def _p64(a):
        return a.ctypes.data_as(_u64p)

    @staticmethod
    def _p32(a):
        return a.ctypes.data_as(_u32p)

    def prime(self, i):
        return self.L.orc_prime(i)

    def gmp_mul(self, a, b):
        a = np.ascontiguousarray(a, np.uint64)
        b = np.ascontiguousarray(b, np.uint64)
        out = np.zeros(a.size + b.size, np.uint64)
        self.L.orc_gmp_mul(self._p64(a), a.size, self._p64(b), b.size, self._p64(out))
        return out

    def ntt_mul(self, a, b):
        a = np.ascontiguousarray(a, np.uint64)
        b = np.ascontiguousarray(b, np.uint64)
        out = np.zeros(a.size + b.size, np.uint64)
        rc = self.L.orc_ntt_mul(self._p64(a), a.size, self._p64(b), b.size, self._p64(out))
        assert rc == 0, rc
        return out

    def ct_layers(self, x, log2n, prime, lo, hi):
        x = np.ascontiguousarray(x, np.uint32).copy()
        self.L.orc_ntt_ct_layers(self._p32(x), log2n, prime, lo, hi)
        return x

    def gs_layers(self, x, log2n, prime, lo, hi):
        x = np.ascontiguousarray(x, np.uint32).copy()
        self.L.orc_ntt_gs_layers(self._p32(x), log2n, prime, lo, hi)
        return x

    def dft(self, x, log2n, prime):
        x = np.ascontiguousarray(x, np.uint32)
        X = np.zeros(1 << log2n, np.uint32)
        self.L.orc_dft_textbook(self._p32(x), self._p32(X), log2n, prime)
        return X

    def zeta(self, prime, log2n):
        z = np.zeros(1 << log2n, np.uint32)
        zi = np.zeros(1 << log2n, np.uint32)
        self.L.orc_zeta_table(prime, log2n, self._p32(z), self._p32(zi))
        return z, zi

    # --- dispatch cost model (double precision) ---
    def mlp_param_count(self, k):
        return self.L.orc_mlp_param_count(k)

    @staticmethod
    def _f64(a):
        return np.ascontiguousarray(a, np.float64)

    def mlp_forward(self, w, k, x):
        w, x = self._f64(w), self._f64(x)
        y = np.zeros((x.shape[0], k), np.float64)
        self.L.orc_mlp_forward(w.ctypes.data_as(_f64p), k, x.ctypes.data_as(_f64p), x.shape[0],
                               y.ctypes.data_as(_f64p))
        return y

    def mlp_loss(self, w, k, x, path, t):
        w, x, t = self._f64(w), self._f64(x), self._f64(t)
        path = np.ascontiguousarray(path, np.int32)
        return self.L.orc_mlp_loss(w.ctypes.data_as(_f64p), k, x.ctypes.data_as(_f64p),
                                   path.ctypes.data_as(_i32p), t.ctypes.data_as(_f64p), x.shape[0])

    def mlp_gradient(self, w, k, x, path, t):
        w, x, t = self._f64(w), self._f64(x), self._f64(t)
        path = np.ascontiguousarray(path, np.int32)
        g = np.zeros(w.size, np.float64)
        L = self.L.orc_mlp_gradient(w.ctypes.data_as(_f64p), k, x.ctypes.data_as(_f64p),
                                    path.ctypes.data_as(_i32p), t.ctypes.data_as(_f64p), x.shape[0],
                                    g.ctypes.data_as(_f64p))
        return g, L

    def mlp_adam(self, w, g, m, v, t, lr):
        for a in (w, g, m, v):
            assert a.dtype == np.float64 and a.flags.c_contiguous
        self.L.orc_mlp_adam(w.ctypes.data_as(_f64p), g.ctypes.data_as(_f64p), m.ctypes.data_as(_f64p),
                            v.ctypes.data_as(_f64p), w.size, t, lr)

    def crt3(self, r, prime_set=0):
        r = np.ascontiguousarray(r, np.uint32)
        out = np.zeros(3, np.uint32)
        self.L.orc_crt3_set(prime_set, self._p32(r), self._p32(out))
        return int(out[0]) | (int(out[1]) << 32) | (int(out[2]) << 64)

    # --- lazy butterflies of the narrow prime set ---
    def narrow_ct(self, a, b, w, p):
        x, y = ctypes.c_uint32(a), ctypes.c_uint32(b)
        self.L.orc_narrow_ct(ctypes.byref(x), ctypes.byref(y), w, p)
        return x.value, y.value

    def narrow_ct_lazy(self, a, b, w, p):
        x, y = ctypes.c_uint32(a), ctypes.c_uint32(b)
        self.L.orc_narrow_ct_lazy(ctypes.byref(x), ctypes.byref(y), w, p)
        return x.value, y.value

    def narrow_gs(self, a, b, winv, p):
        x, y = ctypes.c_uint32(a), ctypes.c_uint32(b)
        self.L.orc_narrow_gs(ctypes.byref(x), ctypes.byref(y), winv, p)
        return x.value, y.value

    def narrow_gs_lazy(self, a, b, winv, p):
        x, y = ctypes.c_uint32(a), ctypes.c_uint32(b)
        self.L.orc_narrow_gs_lazy(ctypes.byref(x), ctypes.byref(y), winv, p)
        return x.value, y.value

    def narrow_layers(self, x, log2n, prime, lo, hi, inverse=False):
        x = np.ascontiguousarray(x, np.uint32).copy()
        f = self.L.orc_narrow_gs_layers if inverse else self.L.orc_narrow_ct_layers
        f(self._p32(x), log2n, prime, lo, hi)
        return x


@pytest.fixture(scope="session")
def oracle():
    path = ROOT / "oracle" / "liboracle.so"
    if not path.exists():
        pytest.fail("oracle/liboracle.so is not built (run __graft_entry__.build())")
    return Oracle(path)


@pytest.fixture(scope="session")
def engine():
    import pi_compute_amd as pc

    lib = pc.load_library()
    rc = lib.gpu_ntt_init(0)
    if rc != 0:
        pytest.fail(f"gpu_ntt_init failed: {lib.gpu_ntt_last_error().decode()}")
    return pc
