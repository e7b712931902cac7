"""Key generation on the device (tfhe_amd_keygen_device): reproducible from the
seed; secret keys binary and balanced; every keyswitch row decrypts to its
message h * t_i / base^(j+1) with noise of the stated standard deviation;
sampled bootstrapping-key rows decrypt exactly (oracle negacyclic product) to
mu * h_p on their component plus noise; and gates bootstrapped with the key
decrypt correctly and equal the oracle's on the same key material."""
import ctypes
import time

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dkey(gpu):
    T = gpu
    P = T.default_params(110)
    t0 = time.perf_counter()
    sk = T.SecretKey.generate_on_device(P, seed=77)
    dt = time.perf_counter() - t0
    yield T, P, sk, sk.export(), dt
    sk.close()


def _signed(u32):
    return (u32.astype(np.int64) + 2 ** 31) % 2 ** 32 - 2 ** 31


def test_reproducible_and_fast(dkey):
    T, P, sk, K, dt = dkey
    print(f"device keygen: {dt * 1e3:.1f} ms")
    sk2 = T.SecretKey.generate_on_device(P, seed=77)
    sk3 = T.SecretKey.generate_on_device(P, seed=78)
    try:
        K2, K3 = sk2.export(), sk3.export()
        for name in K:
            assert np.array_equal(K[name], K2[name]), name
        assert not np.array_equal(K["lwe_key"], K3["lwe_key"])
        assert not np.array_equal(K["bk"], K3["bk"])
    finally:
        sk2.close()
        sk3.close()
    assert dt < 5.0


def test_secret_keys_and_keyswitch_rows(dkey):
    T, P, sk, K, dt = dkey
    s, z = K["lwe_key"], K["tlwe_key"]
    assert set(np.unique(s)) == {0, 1} and set(np.unique(z)) == {0, 1}
    assert abs(s.mean() - 0.5) < 0.08 and abs(z.mean() - 0.5) < 0.05
    ksk = K["ksk"]  # [kN][t][base][n+1]
    assert not ksk[:, :, 0, :].any()  # h = 0 rows are zero
    rows = ksk[:, :, 1:, :].reshape(-1, P.n + 1).astype(np.int64)
    phase = (rows[:, P.n] - rows[:, :P.n] @ s.astype(np.int64)) % 2 ** 32
    i = np.repeat(np.arange(P.k * P.N), P.ks_t * 3)
    j = np.tile(np.repeat(np.arange(P.ks_t), 3), P.k * P.N)
    h = np.tile(np.arange(1, 4), P.k * P.N * P.ks_t)
    msg = (z[i].astype(np.int64) * h << (32 - (j + 1) * P.ks_basebit)) % 2 ** 32
    e = _signed((phase - msg) % 2 ** 32) / 2.0 ** 32
    alpha = 2.0 ** -15
    assert abs(e.std() / alpha - 1) < 0.02
    assert abs(e.mean()) < 4 * alpha / np.sqrt(len(e))


def test_sampled_bootstrapping_key_rows(dkey):
    T, P, sk, K, dt = dkey
    s, z = K["lwe_key"], K["tlwe_key"]
    bk = K["bk"]  # [n][kpl][k+1][N]
    rng = np.random.default_rng(1)
    errs = []
    for i in rng.choice(P.n, 6, replace=False):
        for r in range(P.kpl):
            a, b = bk[i, r, 0].copy(), bk[i, r, 1].copy()
            prod = np.zeros(P.N, np.int32)
            O.lib().orc_negacyclic_mul_add(O.ptr(a), O.ptr(z.astype(np.int32)), P.N, O.ptr(prod))
            phase = (b.astype(np.int64) - prod) % 2 ** 32
            blk, p = divmod(r, P.l)
            hp = 1 << (32 - (p + 1) * P.Bgbit)
            if blk == 1:  # message on b: phase = e + mu h_p at coefficient 0
                phase[0] = (phase[0] - int(s[i]) * hp) % 2 ** 32
            else:  # message on a_0: phase = e - mu h_p s(X)
                phase = (phase + int(s[i]) * hp * z.astype(np.int64)) % 2 ** 32
            errs.append(_signed(phase) / 2.0 ** 32)
    e = np.concatenate(errs)
    alpha = 2.0 ** -25
    assert np.abs(e).max() < 8 * alpha
    assert abs(e.std() / alpha - 1) < 0.05


def test_gates_with_device_key(dkey):
    T, P, sk, K, dt = dkey
    rng = np.random.default_rng(2)
    count = 1024
    ops = rng.integers(0, 11, size=count).astype(np.int32)
    bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    got = T.run_gates(sk.cloud, ops, a, b, c)
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, w) for o, x, y, w in zip(ops, *bits)])
    assert np.array_equal(sk.decrypt(got), exp)
    keys = O.Keys(O.p128(), **K)
    sel = np.array([0, 1, 2, int(np.argmax(ops == 10))])
    assert np.array_equal(got[sel], O.gates(keys, ops[sel], a[sel], b[sel], c[sel], threads=8))
