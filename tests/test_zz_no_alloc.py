"""The gate entry point that allocates nothing (tfhe_amd_gates_device_noalloc)
against tfhe_amd_gates_device, through the C-ABI: two calls in a row (the
second with a smaller batch, so the shared static workspace holds the first
call's contents), outputs bit-identical to the allocating entry point's and
decrypting to the plaintext gates; a count above the static workspace is
refused with an error and nothing launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind(T):
    L = T.lib()
    f = L.tfhe_amd_gates_device_noalloc
    f.restype = ctypes.c_int32
    f.argtypes = [T.CKP, ctypes.c_int32] + [ctypes.c_void_p] * 7
    return f


def _truth(op, a, b, c):
    res = [1 - (a & b), a & b, a | b, a ^ b, 1 - (a ^ b), 1 - (a | b), (1 - a) & b, a & (1 - b), (1 - a) | b,
           a | (1 - b), np.where(a == 1, b, c)]
    return np.choose(op, res)


def test_noalloc_matches_allocating_entry_point(gpu):
    T = gpu
    noalloc = _bind(T)
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=41)
    sk.upload()
    rng = np.random.default_rng(41)
    w = P.n + 1
    try:
        for count in (1500, 333):
            print(f"[test] noalloc batch of {count} gates", flush=True)
            bits = rng.integers(0, 2, size=3 * count).astype(np.int32)
            pool = np.zeros((4 * count, w), np.int32)
            pool[:3 * count] = sk.encrypt(bits)
            op = rng.integers(0, 11, count).astype(np.int32)
            ins = [np.arange(count, dtype=np.int32) + j * count for j in range(3)]
            outw = np.arange(count, dtype=np.int32) + 3 * count
            d_op = T.DeviceArray.from_numpy(op)
            d_in = [T.DeviceArray.from_numpy(x) for x in ins]
            d_out = T.DeviceArray.from_numpy(outw)
            d_a = T.DeviceArray.from_numpy(pool)
            d_b = T.DeviceArray.from_numpy(pool)
            T.gates_device(sk.cloud, d_op, d_in[0], d_in[1], d_in[2], d_out, d_a)
            rc = noalloc(sk.cloud, count, d_op.ptr, d_in[0].ptr, d_in[1].ptr, d_in[2].ptr, d_out.ptr, d_b.ptr, None)
            assert rc == 0, T.lib().tfhe_amd_last_error().decode()
            T.synchronize()
            ra, rb = d_a.numpy(), d_b.numpy()
            assert np.array_equal(ra, rb)
            exp = _truth(op, bits[:count], bits[count:2 * count], bits[2 * count:])
            assert np.array_equal(sk.decrypt(rb[3 * count:]), exp)
            for x in [d_op, d_out, d_a, d_b] + d_in:
                x.free()
        # above the static workspace: refused, with the limit in the message
        big = 8192
        d_op = T.DeviceArray(big)
        rc = noalloc(sk.cloud, big, d_op.ptr, d_op.ptr, d_op.ptr, d_op.ptr, d_op.ptr, d_op.ptr, None)
        assert rc != 0
        assert "static" in T.lib().tfhe_amd_last_error().decode()
        d_op.free()
    finally:
        sk.close()


def test_noalloc_needs_an_uploaded_key(gpu):
    T = gpu
    noalloc = _bind(T)
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=42)
    try:
        rc = noalloc(sk.cloud, 1, None, None, None, None, None, None, None)
        assert rc != 0
        assert "not on the device" in T.lib().tfhe_amd_last_error().decode()
    finally:
        sk.close()
