"""GPU parity for the N = 2048 parameter set (n = 750, l = 2, Bgbit = 10,
keyswitch t = 8 / basebit = 2): keyswitch and blind rotation bit-exact against
the oracle, gates bit-exact against the oracle on the same keys, and decrypted
bits equal to the plaintext gates over a large mixed batch."""
import ctypes

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pbig(gpu):
    T = gpu
    P = T.big_params()
    sk = T.SecretKey(P, seed=3)
    keys = O.Keys(O.pbig(), **sk.export())
    sk.upload()
    yield T, P, sk, keys
    sk.close()


def rand_i32(rng, shape):
    return rng.integers(-2 ** 31, 2 ** 31, size=shape, dtype=np.int64).astype(np.int32)


def test_pbig_keyswitch_bitexact(pbig):
    T, P, sk, keys = pbig
    rng = np.random.default_rng(41)
    ext = rand_i32(rng, (33, P.N + 1))
    got = T.keyswitch(sk.cloud, P, ext)
    for g in (0, 1, 17, 32):
        assert np.array_equal(got[g], O.keyswitch(keys, ext[g])), f"row {g}"


@pytest.mark.parametrize("row,nrows", [(0, 1), (749, 1), (300, 3)])
def test_pbig_blind_rotate_exact(pbig, row, nrows):
    T, P, sk, keys = pbig
    L = T.lib()
    rng = np.random.default_rng(row + 13 * nrows)
    acc0 = rand_i32(rng, (2, P.N))
    bara = rng.integers(1, 2 * P.N, size=nrows).astype(np.int32)
    if nrows > 1:
        bara[1] = P.N
    acc = L.new_TLweSample(P.ptr.contents.tgsw_params.contents.tlwe_params)
    base = sk.cloud.contents.bkFFT.contents.bkFFT
    row_ptr = ctypes.cast(ctypes.addressof(base.contents) + row * ctypes.sizeof(base._type_), type(base))
    try:
        for c in range(2):
            ctypes.memmove(acc.contents.a[c].coefsT, acc0[c].ctypes.data, 4 * P.N)
        L.tfhe_blindRotate_FFT(acc, row_ptr, bara.ctypes.data_as(T.i32p), nrows, P.ptr.contents.tgsw_params)
        got = np.zeros((2, P.N), np.int32)
        for c in range(2):
            ctypes.memmove(got[c].ctypes.data, acc.contents.a[c].coefsT, 4 * P.N)
    finally:
        L.delete_TLweSample(acc)
    ref = acc0.copy()
    OP = O.pbig()
    OP.n = nrows
    O.lib().orc_blind_rotate(ctypes.byref(OP), O.ptr(np.ascontiguousarray(keys.bk[row:row + nrows])), O.ptr(bara),
                             O.ptr(ref))
    assert np.array_equal(got, ref), f"max diff {np.abs(got.astype(np.int64) - ref).max()}"


def test_pbig_gates_bitexact_vs_oracle(pbig):
    T, P, sk, keys = pbig
    rng = np.random.default_rng(43)
    ops = np.arange(11, dtype=np.int32)
    bits = rng.integers(0, 2, size=(3, len(ops))).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    got = T.run_gates(sk.cloud, ops, a, b, c)
    ref = O.gates(keys, ops, a, b, c, threads=16)
    assert np.array_equal(got, ref)
    dec = sk.decrypt(got)
    for g, op in enumerate(ops):
        assert dec[g] == O.TRUTH[O.GATES[op]](bits[0][g], bits[1][g], bits[2][g])


def test_pbig_mixed_batch_truth(pbig):
    """1024 MUX + 1024 two-input gates in one batch decrypt correctly."""
    T, P, sk, keys = pbig
    rng = np.random.default_rng(44)
    ops = np.concatenate([np.full(1024, 10), rng.integers(0, 10, size=1024)]).astype(np.int32)
    rng.shuffle(ops)
    bits = rng.integers(0, 2, size=(3, len(ops))).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    got = T.run_gates(sk.cloud, ops, a, b, c)
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, z) for o, x, y, z in zip(ops, *bits)])
    assert np.array_equal(sk.decrypt(got), exp)


def test_pbig_pair_kernel_matches_one_wave_kernel(pbig):
    """The pair-transform kernel (four wavefronts per bootstrap, the default for
    N = 2048) and the one-wavefront kernel give identical ciphertexts, MUX
    included, for an odd batch."""
    T, P, sk, keys = pbig
    rng = np.random.default_rng(45)
    ops = rng.integers(0, 11, size=37).astype(np.int32)
    ops[:3] = 10
    bits = rng.integers(0, 2, size=(3, len(ops))).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    T.set_launch_config(1)
    try:
        one = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    pair = T.run_gates(sk.cloud, ops, a, b, c)
    assert np.array_equal(one, pair)
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, z) for o, x, y, z in zip(ops, *bits)])
    assert np.array_equal(sk.decrypt(pair), exp)


@pytest.mark.parametrize("count", [1, 37, 600])
def test_pbig_pair_kernel_forms_bitexact(pbig, count):
    """The pair kernel on the fused-butterfly transform (the default, mode 4),
    on the radix-8 stages (mode 8) and the two-wavefront kernel (mode 6) give
    identical ciphertexts, equal to the oracle's (first gates) and decrypting to
    the plaintext gates; 600 gates with MUX take more workgroups than the GPU
    holds at once."""
    T, P, sk, keys = pbig
    rng = np.random.default_rng(47 + count)
    ops = rng.integers(0, 11, size=count).astype(np.int32)
    ops[:min(3, count)] = 10
    bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    print(f"[test] N=2048 pair kernel forms, {count} gates", flush=True)
    got = {}
    for mode in (4, 8, 6):
        T.set_launch_config(mode)
        try:
            got[mode] = T.run_gates(sk.cloud, ops, a, b, c)
        finally:
            T.set_launch_config(0)
    assert np.array_equal(got[4], got[8])
    assert np.array_equal(got[4], got[6])
    k = min(count, 6)
    assert np.array_equal(got[4][:k], O.gates(keys, ops[:k], a[:k], b[:k], c[:k], threads=16))
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, z) for o, x, y, z in zip(ops, *bits)])
    assert np.array_equal(sk.decrypt(got[4]), exp)


@pytest.fixture(scope="module")
def n2048_l3(gpu):
    """N = 2048 with three gadget digits (n = 630, Bgbit = 7, P128's noise): the l = 3
    instantiation of the pair-transform kernel."""
    T = gpu
    P = T.Params(T.lib().new_gate_bootstrapping_parameters(630, 2048, 1, 3, 7, 8, 2, 2.0 ** -15, 2.0 ** -25,
                                                           0.012467))
    sk = T.SecretKey(P, seed=6)
    keys = O.Keys(O.OrcParams(630, 2048, 1, 3, 7, 8, 2, 2.0 ** -15, 2.0 ** -25), **sk.export())
    sk.upload()
    yield T, P, sk, keys
    sk.close()


def test_n2048_l3_pair_kernel_bitexact(n2048_l3):
    """Every gate type (and extra MUX, odd batch) through the l = 3 pair kernel
    (fused-butterfly and radix-8 forms) and the two-wavefront kernel: bit-exact
    against the oracle and the one-wavefront kernel, and decrypting to the
    plaintext gate."""
    T, P, sk, keys = n2048_l3
    rng = np.random.default_rng(46)
    ops = np.array(list(range(11)) + [10, 10], np.int32)
    bits = rng.integers(0, 2, size=(3, len(ops))).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    pair = T.run_gates(sk.cloud, ops, a, b, c)
    T.set_launch_config(1)
    try:
        one = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    T.set_launch_config(8)
    try:
        radix8 = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    T.set_launch_config(6)
    try:
        two_wave = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    assert np.array_equal(pair, one)
    assert np.array_equal(pair, radix8)
    assert np.array_equal(pair, two_wave)
    assert np.array_equal(pair, O.gates(keys, ops, a, b, c, threads=16))
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, z) for o, x, y, z in zip(ops, *bits)])
    assert np.array_equal(sk.decrypt(pair), exp)


def test_pbig_autotune_chooses_kernel_shape(pbig):
    """The autotuner on an N = 2048 key times the pair kernel (four wavefronts
    per bootstrap) and the two-wavefront kernel over its ladder of batch sizes
    and records the faster one per size; automatic mode then gives the same
    ciphertexts as either kernel forced, on both sides of the crossover and
    beyond the ladder."""
    T, P, sk, keys = pbig
    lat, pair_max, table = T.autotune(sk.cloud)
    assert lat == 0
    assert pair_max == 0 or pair_max in T.AUTOTUNE_SIZES
    assert set(table) == set(T.AUTOTUNE_MODES_N2048)
    assert all(v > 0 for row in table.values() for v in row.values())
    # the crossover is where the measured ladder says it is
    sizes = [s for s in T.AUTOTUNE_SIZES if s <= pair_max]
    assert all(table["pair_four_wave"][s] <= table["two_wave"][s] for s in sizes)
    # automatic mode: the faster kernel of the smallest ladder size that holds the batch,
    # the two-wavefront kernel beyond the ladder
    for size in T.AUTOTUNE_SIZES:
        want = 4 if table["pair_four_wave"][size] <= table["two_wave"][size] else 6
        assert T.kernel_for(sk.cloud, size) == want, size
    assert T.kernel_for(sk.cloud, max(T.AUTOTUNE_SIZES) + 1) == 6
    rng = np.random.default_rng(48)
    for count in (max(pair_max, 1), pair_max + 130, max(T.AUTOTUNE_SIZES) + 100):
        ops = rng.integers(0, 11, size=count).astype(np.int32)
        bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
        a, b, c = (sk.encrypt(x) for x in bits)
        auto = T.run_gates(sk.cloud, ops, a, b, c)
        for mode in (4, 6):
            T.set_launch_config(mode)
            try:
                assert np.array_equal(T.run_gates(sk.cloud, ops, a, b, c), auto), (count, mode)
            finally:
                T.set_launch_config(0)
        exp = np.array([O.TRUTH[O.GATES[o]](x, y, z) for o, x, y, z in zip(ops, *bits)])
        assert np.array_equal(sk.decrypt(auto), exp)
