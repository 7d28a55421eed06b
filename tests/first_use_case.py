"""First product of a fresh process on a caller stream (run as a script by
tests/test_gpu_parity.py::test_first_product_on_caller_stream).

The tile plans are loaded from the shipped plan file, so nothing tunes (tuning
waits on the library stream and would hide a missing order), and the very
first product of the process is a 2^23-point product enqueued on the caller's
own stream: its twiddle tables are built on the library stream at that moment,
and the product must wait for them.  The result is compared with GMP mpz_mul.
Prints FIRST_USE_OK and exits 0 when the product is exact.
"""
import ctypes
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import pi_compute_amd as pc  # noqa: E402
from pi_compute_amd import hiprt  # noqa: E402


def main() -> int:
    mode = sys.argv[1] if len(sys.argv) > 1 else "product"
    lib = pc.load_library()
    plans = ROOT / "pi_compute_amd" / "data" / "plans_gfx950.txt"
    if lib.gpu_ntt_load_plans(str(plans).encode()) != 0:
        print("load_plans failed:", lib.gpu_ntt_last_error().decode())
        return 1
    # an + bn just above 2^21 limbs: N = 2^23 (wide prime set), one chunk.
    a, b = pc.operands(23, 64 * ((1 << 21) - 4096), 64 * 8192)
    orc = ctypes.CDLL(str(ROOT / "oracle" / "liboracle.so"))
    u64p = ctypes.POINTER(ctypes.c_uint64)
    orc.orc_gmp_mul.argtypes = [u64p, ctypes.c_size_t, u64p, ctypes.c_size_t, u64p]
    want = np.zeros(a.size + b.size, np.uint64)
    orc.orc_gmp_mul(a.ctypes.data_as(u64p), a.size, b.ctypes.data_as(u64p), b.size, want.ctypes.data_as(u64p))
    da, db = hiprt.DeviceBuffer.from_array(a), hiprt.DeviceBuffer.from_array(b)
    dout = hiprt.DeviceBuffer(8 * want.size)
    stream = hiprt.Stream()
    print(f"first use ({mode}): {a.size}x{b.size} limbs, N=2^{pc.transform_log2(a.size, b.size)}", flush=True)
    if mode == "graph":
        graph = pc.BatchGraph([da.ptr], [a.size], [db.ptr], [b.size], [dout.ptr])
        graph.launch(stream.handle)
        stream.synchronize()
        graph.close()
    else:
        pc.mul_device(da.ptr, a.size, db.ptr, b.size, dout.ptr, stream.handle)
        stream.synchronize()
    got = dout.to_array()
    stream.close()
    for x in (da, db, dout):
        x.close()
    if not np.array_equal(got, want):
        bad = np.nonzero(got != want)[0]
        print(f"MISMATCH: {bad.size} limbs differ, first at {bad[0]}")
        return 1
    print("FIRST_USE_OK", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
