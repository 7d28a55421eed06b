"""pi_compute_amd - MI355X (gfx950) exact big-integer product engine.

The product path is the HIP library ``libpintt.so`` built from ``csrc/`` and
reached through the C ABI declared in ``include/pintt_gpu.h``.  This module is
the thin Python view of that ABI used by the tests, the smoke check and the
benchmark: it loads the library (and fails loudly if it was not built), binds
the entry points with ctypes and provides the seeded operand generator.

There is no Python or CPU fallback for a product: if the library or the device
is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libpintt.so"

PINTT_OK = 0
STATUS_NAMES = {
    0: "PINTT_OK",
    1: "PINTT_ERR_NO_DEVICE",
    2: "PINTT_ERR_TOO_LARGE",
    3: "PINTT_ERR_NO_MEMORY",
    4: "PINTT_ERR_HIP",
    5: "PINTT_ERR_ARGS",
}

# Every symbol include/*.h declares, with (restype, argtypes).
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
ABI = {
    "gpu_ntt_init": (ctypes.c_int, [ctypes.c_int]),
    "gpu_ntt_shutdown": (None, []),
    "gpu_ntt_mul": (ctypes.c_int, [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p]),
    "gpu_ntt_sqr": (ctypes.c_int, [_u64p, ctypes.c_size_t, _u64p]),
    "gpu_ntt_mul_device": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "gpu_ntt_mul_device_noalloc": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p],
    ),
    "gpu_ntt_transform_log2": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_size_t]),
    "gpu_ntt_max_limbs": (ctypes.c_size_t, []),
    "gpu_ntt_product_plan": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
         ctypes.POINTER(ctypes.c_size_t)],
    ),
    "gpu_ntt_last_error": (ctypes.c_char_p, []),
    "gpu_ntt_set_chunk_cost": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_ntt_chunk_plan": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)]),
    "gpu_ntt_set_timing": (None, [ctypes.c_int]),
    "gpu_ntt_pass_times": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "gpu_ntt_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "gpu_ntt_get_plan": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
         ctypes.POINTER(ctypes.c_float)],
    ),
    "gpu_ntt_set_plan": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "gpu_ntt_list_plans": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "gpu_ntt_set_batch_plan": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "gpu_ntt_get_batch_plan": (
        ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                       ctypes.POINTER(ctypes.c_int)]),
    "gpu_ntt_mul_batch": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_u64p),
         ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_u64p)],
    ),
    "gpu_ntt_mul_batch_device": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_void_p),
         ctypes.c_void_p],
    ),
    "gpu_ntt_set_split": (None, [ctypes.c_int]),
    "gpu_ntt_set_graph_branches": (None, [ctypes.c_int]),
    "gpu_ntt_graph_create": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_void_p),
         ctypes.POINTER(ctypes.c_void_p)],
    ),
    "gpu_ntt_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_ntt_graph_nodes": (ctypes.c_size_t, [ctypes.c_void_p]),
    "gpu_ntt_graph_destroy": (None, [ctypes.c_void_p]),
    "gpu_ntt_synchronize": (ctypes.c_int, []),
    "gpu_ntt_save_plans": (ctypes.c_int, [ctypes.c_char_p]),
    "gpu_ntt_load_plans": (ctypes.c_int, [ctypes.c_char_p]),
    "gpu_ntt_dev_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_u64p)]),
    "gpu_ntt_dev_free": (ctypes.c_int, [_u64p]),
    "gpu_ntt_upload": (
        ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_u64p), ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t)]
    ),
    "gpu_ntt_download": (
        ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_u64p), ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t)]
    ),
    "gpu_ntt_dev_copy": (ctypes.c_int, [_u64p, _u64p, ctypes.c_size_t]),
    "gpu_ntt_add_batch_device": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_u64p),
         ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_u64p)],
    ),
    "gpu_ntt_add_batch_device_async": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_u64p),
         ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_u64p), ctypes.c_void_p],
    ),
    "gpu_ntt_add_batch_device_flags": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_u64p),
         ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_u64p),
         ctypes.POINTER(ctypes.c_int)],
    ),
    "gpu_ntt_bs_leaves": (
        ctypes.c_int,
        [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t, _u64p, ctypes.POINTER(ctypes.c_uint32),
         ctypes.POINTER(ctypes.c_int32)],
    ),
    "gpu_ntt_decimal_leaves_device": (
        ctypes.c_int,
        [ctypes.c_size_t, ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
         ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p, ctypes.c_size_t],
    ),
    # include/pintt_mlp.h
    "gpu_mlp_param_count": (ctypes.c_size_t, [ctypes.c_int]),
    "gpu_mlp_forward": (
        ctypes.c_int,
        [_f32p, ctypes.c_int, _f32p, ctypes.c_size_t, _f32p],
    ),
    "gpu_mlp_gradient": (
        ctypes.c_int,
        [_f32p, ctypes.c_int, _f32p, _i32p, _f32p, ctypes.c_size_t, _f32p, _f32p],
    ),
    "gpu_mlp_train": (
        ctypes.c_int,
        [_f32p, ctypes.c_int, _f32p, _i32p, _f32p, ctypes.c_size_t, ctypes.c_int, ctypes.c_float, _f32p],
    ),
    "gpu_mlp_last_error": (ctypes.c_char_p, []),
}

PASS_NAMES = ("forward_columns", "row_convolution", "inverse_columns", "crt_carry")

_lib = None


class PinttError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def load_library(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load libpintt.so and bind every ABI symbol.  Raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP engine first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in ABI.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(lib, status: int) -> None:
    if status != PINTT_OK:
        raise PinttError(status, lib.gpu_ntt_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """|a| * |b| for little-endian uint64 limb arrays, on the GPU (host buffers)."""
    lib = load_library()
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.empty(a.size + b.size, dtype=np.uint64)
    _check(lib, lib.gpu_ntt_mul(_ptr(a), a.size, _ptr(b), b.size, _ptr(out)))
    return out


def sqr(a: np.ndarray) -> np.ndarray:
    """|a|^2 on the GPU (host buffers)."""
    lib = load_library()
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.empty(2 * a.size, dtype=np.uint64)
    _check(lib, lib.gpu_ntt_sqr(_ptr(a), a.size, _ptr(out)))
    return out


def mul_batch(pairs) -> list:
    """Products of a list of (a, b) limb-array pairs in one batch (host buffers).
    A pair whose two members are the same array object is squared."""
    lib = load_library()
    ops = []
    for a, b in pairs:
        same = a is b
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = a if same else np.ascontiguousarray(b, dtype=np.uint64)
        ops.append((a, b, np.empty(a.size + b.size, dtype=np.uint64)))
    n = len(ops)
    pa = (_u64p * n)(*[_ptr(a) for a, _, _ in ops])
    pb = (_u64p * n)(*[_ptr(b) for _, b, _ in ops])
    po = (_u64p * n)(*[_ptr(o) for _, _, o in ops])
    na = (ctypes.c_size_t * n)(*[a.size for a, _, _ in ops])
    nb = (ctypes.c_size_t * n)(*[b.size for _, b, _ in ops])
    _check(lib, lib.gpu_ntt_mul_batch(n, pa, na, pb, nb, po))
    return [o for _, _, o in ops]


def mul_batch_device(d_a, an, d_b, bn, d_out, stream: int = 0) -> None:
    """Enqueue a batch of products on device pointers (lists of equal length)."""
    lib = load_library()
    n = len(d_a)
    vp = ctypes.c_void_p
    _check(lib, lib.gpu_ntt_mul_batch_device(
        n, (vp * n)(*d_a), (ctypes.c_size_t * n)(*an), (vp * n)(*d_b), (ctypes.c_size_t * n)(*bn),
        (vp * n)(*d_out), stream or None))


class BatchGraph:
    """A batch of device-pointer products captured into a HIP graph."""

    def __init__(self, d_a, an, d_b, bn, d_out):
        lib = load_library()
        n = len(d_a)
        vp = ctypes.c_void_p
        h = vp()
        _check(lib, lib.gpu_ntt_graph_create(
            n, (vp * n)(*d_a), (ctypes.c_size_t * n)(*an), (vp * n)(*d_b), (ctypes.c_size_t * n)(*bn),
            (vp * n)(*d_out), ctypes.byref(h)))
        self.handle = h.value
        self.nodes = lib.gpu_ntt_graph_nodes(self.handle)

    def launch(self, stream: int = 0) -> None:
        lib = load_library()
        _check(lib, lib.gpu_ntt_graph_launch(self.handle, stream or None))

    def close(self) -> None:
        if self.handle:
            load_library().gpu_ntt_graph_destroy(self.handle)
            self.handle = None


def synchronize() -> None:
    lib = load_library()
    _check(lib, lib.gpu_ntt_synchronize())


def set_split(mode: int) -> None:
    """0 never cut operands, 1 cut when cheaper (default), 2..4 force that many chunks."""
    load_library().gpu_ntt_set_split(int(mode))


def set_graph_branches(enable: bool) -> None:
    """Graphs made after this capture each length bucket as its own branch (default) or one chain."""
    load_library().gpu_ntt_set_graph_branches(1 if enable else 0)


def mul_device(d_a: int, an: int, d_b: int, bn: int, d_out: int, stream: int = 0) -> None:
    """Enqueue a product on device pointers (e.g. torch tensors' data_ptr())."""
    lib = load_library()
    _check(lib, lib.gpu_ntt_mul_device(d_a, an, d_b, bn, d_out, stream or None))


def transform_log2(an: int, bn: int) -> int:
    return load_library().gpu_ntt_transform_log2(an, bn)


def list_plans(log2n: int) -> list:
    buf = (ctypes.c_int * 3 * 256)()
    k = load_library().gpu_ntt_list_plans(log2n, ctypes.cast(buf, ctypes.POINTER(ctypes.c_int)), 256)
    return [tuple(buf[i]) for i in range(k)]


def get_plan(log2n: int) -> dict:
    lib = load_library()
    r, w, rows, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
    _check(lib, lib.gpu_ntt_get_plan(log2n, ctypes.byref(r), ctypes.byref(w), ctypes.byref(rows), ctypes.byref(ms)))
    return {"log2n": log2n, "log2r": r.value, "log2c": log2n - r.value, "log2w": w.value,
            "log2rows": rows.value, "tuned_ms": ms.value}


def product_plan(an: int, bn: int) -> dict:
    """Transform length and chunking gpu_ntt_mul would use for an an x bn product."""
    lib = load_library()
    k, m, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    _check(lib, lib.gpu_ntt_product_plan(an, bn, ctypes.byref(k), ctypes.byref(m), ctypes.byref(h)))
    return {"log2n": k.value, "chunks": m.value, "chunk_limbs": h.value}


def set_plan(log2n: int, log2r: int, log2w: int, log2rows: int) -> None:
    lib = load_library()
    _check(lib, lib.gpu_ntt_set_plan(log2n, log2r, log2w, log2rows))


def set_batch_plan(log2n: int, log2r: int, log2w: int, log2rows: int) -> None:
    """Plan for large buckets of length 2^log2n (log2r <= 0 clears it)."""
    lib = load_library()
    _check(lib, lib.gpu_ntt_set_batch_plan(log2n, log2r, log2w, log2rows))


def get_batch_plan(log2n: int):
    """(log2r, log2w, log2rows) of the length's batch plan, or None."""
    lib = load_library()
    r, w, rows = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if lib.gpu_ntt_get_batch_plan(log2n, ctypes.byref(r), ctypes.byref(w), ctypes.byref(rows)) != 0:
        return None
    return (r.value, w.value, rows.value)


def autotune(lo: int, hi: int) -> None:
    lib = load_library()
    _check(lib, lib.gpu_ntt_autotune(lo, hi))


def set_timing(enable: bool) -> None:
    load_library().gpu_ntt_set_timing(1 if enable else 0)


def pass_times() -> dict:
    lib = load_library()
    buf = (ctypes.c_double * 8)()
    k = lib.gpu_ntt_pass_times(buf, 8)
    return {PASS_NAMES[i]: buf[i] for i in range(k)}


# ---------------------------------------------------------------------------
# Seeded operands: splitmix64 stream (Steele, Lea, Flood 2014), limbs drawn
# least significant first, top limb masked to the exact bit length with the top
# bit set; operand a first, then b, from one stream.

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start .. start+count-1 of the splitmix64 stream seeded with `seed`."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def limbs_for_bits(bits: int) -> int:
    return (bits + 63) // 64


def draw_operand(seed: int, start: int, bits: int) -> np.ndarray:
    """`bits`-bit operand (top bit set) from stream outputs start.. ."""
    n = limbs_for_bits(bits)
    x = splitmix64(seed, start, n)
    top = bits - 64 * (n - 1)
    if top < 64:
        x[-1] &= np.uint64((1 << top) - 1)
    x[-1] |= np.uint64(1 << (top - 1))
    return x


def operands(seed: int, bits_a: int, bits_b: int):
    """Operands a (bits_a bits) then b (bits_b bits) from one splitmix64 stream."""
    a = draw_operand(seed, 0, bits_a)
    b = draw_operand(seed, a.size, bits_b)
    return a, b


def to_int(limbs: np.ndarray) -> int:
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little")


def from_int(x: int, nlimbs: int | None = None) -> np.ndarray:
    n = nlimbs if nlimbs is not None else max(1, (x.bit_length() + 63) // 64)
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64)


# ---- device-resident limb arrays (binary splitting's upper levels) ----

def upload(pairs) -> None:
    """Host limbs into device memory in one call: [(device address, numpy uint64 array)]."""
    lib = load_library()
    srcs = [np.ascontiguousarray(x, np.uint64) for _, x in pairs]
    n = len(pairs)
    dst = (_u64p * n)(*[ctypes.cast(d, _u64p) for d, _ in pairs])
    src = (_u64p * n)(*[_ptr(x) for x in srcs])
    _check(lib, lib.gpu_ntt_upload(n, dst, src, (ctypes.c_size_t * n)(*[x.size for x in srcs])))


def download(pairs) -> None:
    """Device memory into host limbs in one call: [(numpy uint64 array, written in place; device address)]."""
    lib = load_library()
    for x, _ in pairs:
        if x.dtype != np.uint64 or not x.flags.c_contiguous:
            raise ValueError("download needs contiguous uint64 arrays")
    n = len(pairs)
    dst = (_u64p * n)(*[_ptr(x) for x, _ in pairs])
    src = (_u64p * n)(*[ctypes.cast(d, _u64p) for _, d in pairs])
    _check(lib, lib.gpu_ntt_download(n, dst, src, (ctypes.c_size_t * n)(*[x.size for x, _ in pairs])))


def decimal_leaves(leaves) -> bytes:
    """Device decimal digits: leaves = [(x, width)], x a numpy uint64 limb array
    (value below 10^width, width <= 4096).  Uploads, converts all leaves into
    one buffer (leaf i at the sum of the earlier widths), downloads it."""
    lib = load_library()
    n = len(leaves)
    xs = [np.ascontiguousarray(x, np.uint64) for x, _ in leaves]
    widths = [int(w) for _, w in leaves]
    offsets = list(np.cumsum([0] + widths[:-1])) if n else []
    total = sum(widths)
    words = (total + 7) // 8
    dev = []
    try:
        def alloc(k):
            p = _u64p()
            _check(lib, lib.gpu_ntt_dev_alloc(max(k, 1), ctypes.byref(p)))
            dev.append(p)
            return p
        dx = [alloc(x.size) for x in xs]
        dout = alloc(words)
        arr = lambda ps: (_u64p * len(ps))(*ps)
        sz = lambda v: (ctypes.c_size_t * len(v))(*[int(t) for t in v])
        _check(lib, lib.gpu_ntt_upload(n, arr(dx), arr([_ptr(x) for x in xs]), sz([x.size for x in xs])))
        _check(lib, lib.gpu_ntt_decimal_leaves_device(n, arr(dx), sz([x.size for x in xs]), sz(widths), sz(offsets),
                                                      ctypes.cast(dout, ctypes.c_void_p), total))
        host = np.zeros(max(words, 1), np.uint64)
        _check(lib, lib.gpu_ntt_download(1, arr([_ptr(host)]), arr([dout]), sz([words])))
        return host.tobytes()[:total]
    finally:
        for p in dev:
            lib.gpu_ntt_dev_free(p)


def add_batch(pairs, subtract, flags=False, enqueue=False):
    """Device add/subtract of magnitudes: [(x, y)] numpy uint64 limb arrays,
    subtract[i] -> x - y (needs x >= y).  Uploads, runs one batch, downloads.
    flags=True: x < y is allowed (the result wraps modulo 2^(64 w)) and the
    call returns (outputs, negative flags).  enqueue=True: the batch goes
    through the enqueue-only entry point and the call returns (outputs,
    whether any subtraction went negative)."""
    lib = load_library()
    n = len(pairs)
    xs = [np.ascontiguousarray(x, np.uint64) for x, _ in pairs]
    ys = [np.ascontiguousarray(y, np.uint64) for _, y in pairs]
    widths = [max(x.size, y.size) + (0 if s else 1) for (x, y), s in zip(zip(xs, ys), subtract)]
    dev = []
    try:
        def alloc(k):
            p = _u64p()
            _check(lib, lib.gpu_ntt_dev_alloc(k, ctypes.byref(p)))
            dev.append(p)
            return p
        dx = [alloc(x.size) for x in xs]
        dy = [alloc(y.size) for y in ys]
        do = [alloc(w) for w in widths]
        arr = lambda ps: (_u64p * len(ps))(*ps)
        sz = lambda v: (ctypes.c_size_t * len(v))(*v)
        _check(lib, lib.gpu_ntt_upload(n, arr(dx), arr([_ptr(x) for x in xs]), sz([x.size for x in xs])))
        _check(lib, lib.gpu_ntt_upload(n, arr(dy), arr([_ptr(y) for y in ys]), sz([y.size for y in ys])))
        subs = (ctypes.c_int * n)(*[1 if s else 0 for s in subtract])
        neg = (ctypes.c_int * n)()
        if enqueue:
            flag = alloc(1)
            zero = np.zeros(1, np.uint64)
            _check(lib, lib.gpu_ntt_upload(1, arr([flag]), arr([_ptr(zero)]), sz([1])))
            _check(lib, lib.gpu_ntt_add_batch_device_async(n, arr(dx), sz([x.size for x in xs]), arr(dy),
                                                           sz([y.size for y in ys]), subs, arr(do),
                                                           ctypes.cast(flag, ctypes.c_void_p)))
            got = np.zeros(1, np.uint64)
            _check(lib, lib.gpu_ntt_download(1, arr([_ptr(got)]), arr([flag]), sz([1])))
            outs = [np.zeros(w, np.uint64) for w in widths]
            _check(lib, lib.gpu_ntt_download(n, arr([_ptr(o) for o in outs]), arr(do), sz(widths)))
            return outs, bool(got[0])
        if flags:
            _check(lib, lib.gpu_ntt_add_batch_device_flags(n, arr(dx), sz([x.size for x in xs]), arr(dy),
                                                           sz([y.size for y in ys]), subs, arr(do), neg))
        else:
            _check(lib, lib.gpu_ntt_add_batch_device(n, arr(dx), sz([x.size for x in xs]), arr(dy),
                                                     sz([y.size for y in ys]), subs, arr(do)))
        outs = [np.zeros(w, np.uint64) for w in widths]
        _check(lib, lib.gpu_ntt_download(n, arr([_ptr(o) for o in outs]), arr(do), sz(widths)))
        return (outs, [bool(v) for v in neg]) if flags else outs
    finally:
        for p in dev:
            lib.gpu_ntt_dev_free(p)


# ---- dispatch cost model (include/pintt_mlp.h) ----

def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32)


def _mlp_check(lib, status: int) -> None:
    if status != PINTT_OK:
        raise PinttError(status, lib.gpu_mlp_last_error().decode())


def mlp_forward(params, outputs: int, x) -> np.ndarray:
    """Device inference: x[n][4] -> y[n][outputs] (float32)."""
    lib = load_library()
    w, x = _f32(params), _f32(x)
    y = np.zeros((x.shape[0], outputs), np.float32)
    _mlp_check(lib, lib.gpu_mlp_forward(w.ctypes.data_as(_f32p), outputs, x.ctypes.data_as(_f32p), x.shape[0],
                                        y.ctypes.data_as(_f32p)))
    return y


def mlp_gradient(params, outputs: int, x, path, target):
    """Device gradient and loss of the masked squared error."""
    lib = load_library()
    w, x, t = _f32(params), _f32(x), _f32(target)
    p = np.ascontiguousarray(path, np.int32)
    g = np.zeros(w.size, np.float32)
    loss = np.zeros(1, np.float32)
    _mlp_check(lib, lib.gpu_mlp_gradient(w.ctypes.data_as(_f32p), outputs, x.ctypes.data_as(_f32p),
                                         p.ctypes.data_as(_i32p), t.ctypes.data_as(_f32p), x.shape[0],
                                         g.ctypes.data_as(_f32p), loss.ctypes.data_as(_f32p)))
    return g, float(loss[0])


def mlp_train(params, outputs: int, x, path, target, steps: int, lr: float):
    """Full-batch Adam on the device; returns (params, loss history)."""
    lib = load_library()
    w = _f32(params).copy()
    x, t = _f32(x), _f32(target)
    p = np.ascontiguousarray(path, np.int32)
    hist = np.zeros(max(steps, 1), np.float32)
    _mlp_check(lib, lib.gpu_mlp_train(w.ctypes.data_as(_f32p), outputs, x.ctypes.data_as(_f32p),
                                      p.ctypes.data_as(_i32p), t.ctypes.data_as(_f32p), x.shape[0], steps, lr,
                                      hist.ctypes.data_as(_f32p)))
    return w, hist[:steps]
