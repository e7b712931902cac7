"""Client-side device encryption / decryption (tfhe_amd_encrypt_device,
tfhe_amd_decrypt_device) and the error-rate run: 10^6 chained gates with
zero decryption errors at the default parameters."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def key(gpu):
    T = gpu
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=21)
    sk.upload()
    yield T, P, sk
    sk.close()


def rand_i32(rng, shape):
    return rng.integers(-2 ** 31, 2 ** 31, size=shape, dtype=np.int64).astype(np.int32)


def test_decrypt_device_bitexact(key):
    """Phases are b - <a, s> exactly (host phases as the check); bits = phase > 0."""
    T, P, sk = key
    rng = np.random.default_rng(1)
    cts = np.concatenate([rand_i32(rng, (3000, P.n + 1)), sk.encrypt(rng.integers(0, 2, 1000))])
    cts[0] = 0  # phase 0 decrypts to 0
    d = T.DeviceArray.from_numpy(cts)
    db, dp = T.DeviceArray(len(cts)), T.DeviceArray(len(cts))
    T.decrypt_device(sk, d, db, dp)
    T.synchronize()
    ph = dp.numpy()
    assert np.array_equal(ph, sk.phases(cts))
    assert np.array_equal(db.numpy(), (ph > 0).astype(np.int32))
    assert np.array_equal(db.numpy()[3000:], sk.decrypt(cts[3000:]))
    # oracle phase on a few rows
    keys = O.Keys(O.p128(), **sk.export())
    for g in (0, 1, 2999, 3999):
        assert ph[g] == O.lib().orc_lwe_phase(O.ptr(cts[g]), O.ptr(keys.lwe_key), P.n)
    for x in (d, db, dp):
        x.free()


def test_encrypt_device_statistics_and_reproducibility(key):
    T, P, sk = key
    rng = np.random.default_rng(2)
    count = 65536
    bits = rng.integers(0, 2, count).astype(np.int32)
    d_bits = T.DeviceArray.from_numpy(bits)
    d_cts = T.DeviceArray((count, P.n + 1))
    T.encrypt_device(sk, d_bits, d_cts, seed=0x1234_5678_9ABC, first_index=0)
    T.synchronize()
    cts = d_cts.numpy()
    # decrypts to the bits; noise e = phase - mu has stdev alpha = 2^-15, mean 0
    assert np.array_equal(sk.decrypt(cts), bits)
    mu = np.where(bits == 1, 1 << 29, -(1 << 29)).astype(np.int64)
    e = (sk.phases(cts).astype(np.int64) - mu) / 2.0 ** 32
    alpha = 2.0 ** -15
    assert abs(e.std() / alpha - 1) < 0.02
    assert abs(e.mean()) < 4 * alpha / np.sqrt(count)
    assert abs(np.mean(np.abs(e) < alpha) - 0.6827) < 0.01
    # mask: uniform 32-bit words (mean, and every bit balanced)
    a = cts[:, :P.n].view(np.uint32).astype(np.float64)
    assert abs(a.mean() / 2.0 ** 31 - 1) < 1e-3
    ubits = np.unpackbits(cts[:4096, :P.n].view(np.uint8)).reshape(-1, 8).mean(axis=0)
    assert np.all(np.abs(ubits - 0.5) < 0.01)
    # same (seed, index) -> same ciphertext; a shifted first_index reproduces the tail
    d_tail = T.DeviceArray((1000, P.n + 1))
    d_tb = T.DeviceArray.from_numpy(bits[5000:6000])
    T.encrypt_device(sk, d_tb, d_tail, seed=0x1234_5678_9ABC, first_index=5000)
    T.encrypt_device(sk, d_bits, d_cts, seed=0x1234_5678_9ABD, count=1000)
    T.synchronize()
    assert np.array_equal(d_tail.numpy(), cts[5000:6000])
    assert not np.array_equal(d_cts.numpy()[:1000, :P.n], cts[:1000, :P.n])
    for x in (d_bits, d_cts, d_tail, d_tb):
        x.free()


def test_encrypt_device_default_index_never_repeats(key):
    """Calls that leave the index to the library never reuse a (seed, index)
    pair under one key: two default calls with the same seed give different
    masks, and a default call after explicit indices starts past all of them
    (equal masks would reveal whether two plaintext bits differ)."""
    T, P, sk = key
    w = P.n + 1
    bits = np.ones(64, np.int32)
    d_bits = T.DeviceArray.from_numpy(bits)
    d1, d2, d3, d4 = (T.DeviceArray((64, w)) for _ in range(4))
    T.encrypt_device(sk, d_bits, d1, seed=777)
    T.encrypt_device(sk, d_bits, d2, seed=777)
    T.encrypt_device(sk, d_bits, d3, seed=777, first_index=10 ** 9)
    T.encrypt_device(sk, d_bits, d4, seed=777)
    T.synchronize()
    c1, c2, c3, c4 = (d.numpy() for d in (d1, d2, d3, d4))
    for x, y in ((c1, c2), (c1, c4), (c2, c4), (c3, c4)):
        assert not np.any(np.all(x[:, :P.n][:, None, :] == y[:, :P.n][None, :, :], axis=2))
    # the default call after the explicit one continued at 10^9 + 64: re-encrypting
    # there explicitly reproduces it
    d5 = T.DeviceArray((64, w))
    T.encrypt_device(sk, d_bits, d5, seed=777, first_index=10 ** 9 + 64)
    T.synchronize()
    assert np.array_equal(d5.numpy(), c4)
    for b in (d1, d2, d3, d4, d5):
        assert np.array_equal(sk.decrypt(b.numpy()), bits)
        b.free()
    d_bits.free()


def _truth_all(op, a, b, c):
    res = [1 - (a & b), a & b, a | b, a ^ b, 1 - (a ^ b), 1 - (a | b), (1 - a) & b, a & (1 - b), (1 - a) | b,
           a | (1 - b), np.where(a == 1, b, c)]
    return np.choose(op, res)


def test_million_gates_zero_decryption_errors(key):
    """123 generations of 8192 gates (1,007,616 gates, all 11 types): each
    generation reads random wires of the previous one, so every input after
    the first generation is a bootstrapped output. Every output is decrypted
    on the device and compared with the plaintext evaluation."""
    T, P, sk = key
    rng = np.random.default_rng(3)
    W, gens = 8192, 123
    w = P.n + 1
    d_pool = T.DeviceArray((2 * W, w))
    cur = rng.integers(0, 2, W).astype(np.int32)
    d_b = T.DeviceArray.from_numpy(cur)
    T.encrypt_device(sk, d_b, d_pool, seed=99, count=W)
    d_out_bits = T.DeviceArray(W)
    errors = total = 0
    for t in range(gens):
        src, dst = (t % 2) * W, ((t + 1) % 2) * W
        op = rng.integers(0, 11, W).astype(np.int32)
        ins = rng.integers(0, W, (3, W)).astype(np.int32)
        d_op = T.DeviceArray.from_numpy(op)
        d_in = [T.DeviceArray.from_numpy(x + src) for x in ins]
        d_o = T.DeviceArray.from_numpy(np.arange(W, dtype=np.int32) + dst)
        T.gates_device(sk.cloud, d_op, d_in[0], d_in[1], d_in[2], d_o, d_pool)
        T.decrypt_device(sk, d_pool, d_out_bits, count=W, offset_rows=dst)
        got = d_out_bits.numpy()
        exp = _truth_all(op, cur[ins[0]], cur[ins[1]], cur[ins[2]]).astype(np.int32)
        errors += int((got != exp).sum())
        total += W
        cur = exp
        for x in [d_op, d_o] + d_in:
            x.free()
    for x in (d_pool, d_b, d_out_bits):
        x.free()
    assert total >= 10 ** 6
    assert errors == 0, f"{errors} decryption errors in {total} gates"
