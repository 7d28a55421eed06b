"""gpu_ntt_mul_device_noalloc (residues, carry maps and twiddles in a static
device array; no allocation) against gpu_ntt_mul_device on the same device
operands, through the C ABI: bit-identical outputs on two calls in a row, the
second smaller (the scratch keeps the first call's contents), then a square, a
schoolbook-sized product, the size limit itself, and a product one limb past
it (refused, nothing launched).  Every output is also compared with mpz_mul."""
import ctypes

import numpy as np
import pytest

import pi_compute_amd as pc
from pi_compute_amd import hiprt

pytestmark = pytest.mark.gpu

MAX_LIMBS = 1 << 19  # PINTT_NOALLOC_MAX_LIMBS


def _operands(seed, an, bn, square):
    a, b = pc.operands(seed, 64 * an, 64 * bn)
    return (a, a) if square else (a, b)


def _upload(a, b):
    da = hiprt.DeviceBuffer.from_array(a)
    db = da if b is a else hiprt.DeviceBuffer.from_array(b)
    return da, db


def test_noalloc_matches_device_entry_point(engine, oracle):
    lib = pc.load_library()
    assert lib.gpu_ntt_transform_log2(MAX_LIMBS // 2, MAX_LIMBS // 2) == 20
    # (an, bn, square): the first call larger than the second, as the
    # scratch then still holds the first call's residues and maps.
    cases = [(300000, 120000, False), (5000, 3000, False), (200000, 200000, True), (300, 200, False),
             (MAX_LIMBS // 2 + 12345, MAX_LIMBS // 2 - 12345, False), (50000, 50000, True)]
    stream = hiprt.Stream()
    bufs = []
    try:
        work = []
        for i, (an, bn, sq) in enumerate(cases):
            a, b = _operands(4400 + i, an, bn, sq)
            da, db = _upload(a, b)
            d_na = hiprt.DeviceBuffer(8 * (a.size + b.size))
            d_ref = hiprt.DeviceBuffer(8 * (a.size + b.size))
            bufs += [da, d_na, d_ref] + ([] if db is da else [db])
            work.append((a, b, da, db, d_na, d_ref))
        # Two no-alloc calls in a row on one stream (the second smaller), then
        # the rest; the reference entry point on the same stream after them.
        for a, b, da, db, d_na, _ in work:
            print(f"noalloc {a.size}x{b.size}", flush=True)
            rc = lib.gpu_ntt_mul_device_noalloc(da.ptr, a.size, db.ptr, b.size, d_na.ptr, stream.handle)
            assert rc == 0, lib.gpu_ntt_last_error().decode()
        for a, b, da, db, _, d_ref in work:
            rc = lib.gpu_ntt_mul_device(da.ptr, a.size, db.ptr, b.size, d_ref.ptr, stream.handle)
            assert rc == 0, lib.gpu_ntt_last_error().decode()
        stream.synchronize()
        for a, b, _, _, d_na, d_ref in work:
            got, ref = d_na.to_array(), d_ref.to_array()
            assert np.array_equal(got, ref), f"noalloc differs from gpu_ntt_mul_device at {a.size}x{b.size}"
            assert np.array_equal(got, oracle.gmp_mul(a, b)), f"noalloc differs from mpz_mul at {a.size}x{b.size}"
    finally:
        stream.close()
        for x in bufs:
            x.close()


def test_noalloc_second_call_smaller_after_sync(engine, oracle):
    """The same two calls with the host waiting in between: the second call's
    output must not depend on what the first left in the scratch."""
    lib = pc.load_library()
    stream = hiprt.Stream()
    try:
        for seed, (an, bn) in ((4501, (400000, 100000)), (4502, (70000, 9000))):
            a, b = pc.operands(seed, 64 * an, 64 * bn)
            da, db = _upload(a, b)
            d_na = hiprt.DeviceBuffer(8 * (an + bn))
            d_ref = hiprt.DeviceBuffer(8 * (an + bn))
            try:
                assert lib.gpu_ntt_mul_device_noalloc(da.ptr, an, db.ptr, bn, d_na.ptr, stream.handle) == 0
                assert lib.gpu_ntt_mul_device(da.ptr, an, db.ptr, bn, d_ref.ptr, stream.handle) == 0
                stream.synchronize()
                got = d_na.to_array()
                assert np.array_equal(got, d_ref.to_array())
                assert np.array_equal(got, oracle.gmp_mul(a, b))
            finally:
                for x in (da, db, d_na, d_ref):
                    x.close()
    finally:
        stream.close()


def test_noalloc_rejects_products_past_its_limit(engine):
    lib = pc.load_library()
    x = hiprt.DeviceBuffer(8)
    try:
        # Sizes are checked before anything is read or launched.
        rc = lib.gpu_ntt_mul_device_noalloc(x.ptr, MAX_LIMBS // 2 + 1, x.ptr, MAX_LIMBS // 2, x.ptr, None)
        assert rc == 2
        assert "limit" in lib.gpu_ntt_last_error().decode()
        rc = lib.gpu_ntt_mul_device_noalloc(x.ptr, 0, x.ptr, 1, x.ptr, None)
        assert rc == 5
        rc = lib.gpu_ntt_mul_device_noalloc(ctypes.c_void_p(0), 1, x.ptr, 1, x.ptr, None)
        assert rc == 5
    finally:
        x.close()
