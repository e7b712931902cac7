"""Level scheduler on the GPU: ripple-carry adders (test/test-addition-boot.cpp)
and a random layered circuit, with and without hipGraph replay, checked against
plaintext evaluation and, on a sample, bit-exactly against the CPU oracle."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def key(gpu):
    T = gpu
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=5)
    sk.upload()
    yield T, P, sk
    sk.close()


def _pool(T, sk, P, C, bits_in):
    pool = np.zeros((C.n_wires, P.n + 1), np.int32)
    pool[:len(bits_in)] = sk.encrypt(bits_in)
    return pool


@pytest.mark.parametrize("schedule", [0, 1])
def test_adders(key, schedule):
    T, P, sk = key
    from tfhe_amd.circuits import Netlist, ripple_carry_adders
    n_adders, bits = 24, 8
    C, sums, zero = ripple_carry_adders(n_adders, bits)
    rng = np.random.default_rng(schedule)
    x = rng.integers(0, 2 ** bits, size=n_adders)
    y = rng.integers(0, 2 ** bits, size=n_adders)
    bits_in = np.zeros(2 * n_adders * bits + 1, np.int32)
    for k in range(n_adders):
        for i in range(bits):
            bits_in[k * bits + i] = (x[k] >> i) & 1
            bits_in[n_adders * bits + k * bits + i] = (y[k] >> i) & 1
    pool = _pool(T, sk, P, C, bits_in)
    # carry-in: trivial encryption of 0 (bootsCONSTANT)
    pool[zero, :] = 0
    pool[zero, P.n] = -(1 << 29)
    nl = Netlist(sk.cloud, C, schedule)
    widths = nl.levels()
    assert widths.sum() == 5 * n_adders * bits
    assert len(widths) == 2 * bits + 1
    d = T.DeviceArray.from_numpy(pool)
    nl.run(d, use_graph=True)
    T.synchronize()
    out1 = d.numpy()
    # replay the captured graph on fresh inputs (outputs overwritten, inputs unchanged)
    d2 = T.DeviceArray.from_numpy(pool)
    nl.run(d2, use_graph=False)
    T.synchronize()
    out2 = d2.numpy()
    assert np.array_equal(out1, out2)
    d.free()
    d2.free()
    nl.close()
    dec = sk.decrypt(out1)
    for k in range(n_adders):
        got = sum(int(dec[sums[k, i]]) << i for i in range(bits + 1))
        assert got == x[k] + y[k], (k, x[k], y[k], got)
    # every wire equals the plaintext evaluation
    assert np.array_equal(dec[2 * n_adders * bits + 1:], C.evaluate(bits_in)[2 * n_adders * bits + 1:])


def test_graph_replay_repeats(key):
    """The same graph replayed several times on the same pool gives identical
    results each time (the replays interleaved with host work and syncs)."""
    T, P, sk = key
    from tfhe_amd.circuits import Netlist, ripple_carry_adders
    C, sums, zero = ripple_carry_adders(8, 4)
    rng = np.random.default_rng(9)
    bits_in = rng.integers(0, 2, size=2 * 8 * 4 + 1).astype(np.int32)
    bits_in[zero] = 0
    pool = _pool(T, sk, P, C, bits_in)
    pool[zero, :] = 0
    pool[zero, P.n] = -(1 << 29)
    nl = Netlist(sk.cloud, C, 0)
    d = T.DeviceArray.from_numpy(pool)
    outs = []
    for rep in range(3):
        nl.run(d, use_graph=True)
        T.synchronize()
        outs.append(d.numpy())
        _ = sk.decrypt(outs[-1][:4])  # host work between replays
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    d.free()
    nl.close()


def test_random_layered_circuit_vs_oracle(key):
    """Random layered circuit with every gate type (MUX included); all wires
    decrypt to the plaintext evaluation, and a few gates of the first level are
    bit-exact against the oracle."""
    T, P, sk = key
    from tfhe_amd.circuits import Circuit, Netlist
    rng = np.random.default_rng(13)
    n_in, width, depth = 40, 30, 6
    C = Circuit(n_in)
    layer = list(range(n_in))
    for _ in range(depth):
        new = []
        for _ in range(width):
            op = int(rng.integers(0, 11))
            a, b, c = (int(v) for v in rng.choice(layer, 3))
            new.append(C.gate(op, a, b, c if op == 10 else None))
        layer = new + list(rng.choice(layer, 10))
    bits_in = rng.integers(0, 2, size=n_in).astype(np.int32)
    pool = _pool(T, sk, P, C, bits_in)
    nl = Netlist(sk.cloud, C, 0)
    d = T.DeviceArray.from_numpy(pool)
    nl.run(d)
    T.synchronize()
    out = d.numpy()
    d.free()
    nl.close()
    assert np.array_equal(sk.decrypt(out), C.evaluate(bits_in))
    keys = O.Keys(O.p128(), **sk.export())
    op, in0, in1, in2, outw = C.arrays()
    sel = np.arange(6)
    ref = O.gates(keys, op[sel], pool[in0[sel]], pool[in1[sel]], pool[in2[sel]], threads=8)
    assert np.array_equal(out[outw[sel]], ref)


def test_netlist_rejects_bad_input(key):
    T, P, sk = key
    from tfhe_amd.circuits import Circuit, Netlist
    # a cycle: gate writes wire 2 and reads wire 2
    C2 = Circuit(2)
    C2.gate(0, 0, 1)
    C2.in1[0] = 2
    with pytest.raises(RuntimeError, match="cycle"):
        Netlist(sk.cloud, C2, 0)
    C3 = Circuit(2)
    C3.gate(0, 0, 7)
    with pytest.raises(RuntimeError, match="out of range"):
        Netlist(sk.cloud, C3, 0)
