"""The reference's integration circuits, restated on the GPU path gate by gate
through the reference gate API (one gate per call on host LweSample arrays, as
the reference's test programs call it):

  the 64-input NAND tree          test/test-gate-bootstrapping.cpp:45-93
  16-bit adder, MUX carries       test/test-addition-boot.cpp:30-69 (full_adder_MUX)
  16-bit adder, AND/XOR carries   test/test-addition-boot.cpp:72-97 (full_adder)
  16-bit comparison with MUX      test/test-addition-boot.cpp:100-117 (comparison_MUX)

Each result is checked the way the reference checks it (decrypted bits against
the plaintext circuit, test-addition-boot.cpp:175-240), and sampled gates are
compared bit-exactly with the oracle on the same inputs."""
import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def key(gpu):
    T = gpu
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=31)
    keys = O.Keys(O.p128(), **sk.export())
    sk.upload()
    yield T, P, sk, keys
    sk.close()


def _gate(T, L, sk, name, out, *ins):
    getattr(L, "boots" + name)(out, *ins, sk.cloud)


@pytest.mark.parametrize("trial", [0, 1])
def test_nand_tree(key, trial):
    """test-gate-bootstrapping.cpp:45-93: samples 64..127 are inputs, sample i =
    NAND(2i, 2i+1) for i = 63..1; every node is checked against its children."""
    T, P, sk, keys = key
    L = T.lib()
    rng = np.random.default_rng(100 + trial)
    nb = 64
    H = T.HostCiphertexts(P, 2 * nb)
    try:
        init = np.zeros((2 * nb, P.n + 1), np.int32)
        init[nb:] = sk.encrypt(rng.integers(0, 2, size=nb).astype(np.int32))
        init[0] = sk.encrypt(np.array([rng.integers(0, 2)], np.int32))[0]
        H.set(init)
        for i in range(nb - 1, 0, -1):
            _gate(T, L, sk, "NAND", H.item(i), H.item(2 * i), H.item(2 * i + 1))
        got = H.get()
    finally:
        H.close()
    bits = sk.decrypt(got)
    for i in range(nb - 1, 0, -1):
        assert bits[i] == 1 - (bits[2 * i] & bits[2 * i + 1]), i
    # the root and two inner nodes, bit-exact against the oracle on the same inputs
    sel = np.array([1, 5, 40])
    ref = O.gates(keys, np.zeros(len(sel), np.int32), got[2 * sel], got[2 * sel + 1], threads=8)
    assert np.array_equal(got[sel], ref)


def _bits_of(v, nb):
    return np.array([(v >> i) & 1 for i in range(nb)], np.int32)


class _Regs:
    """Named host ciphertexts: x[0..nb), y[0..nb), then scratch registers."""

    def __init__(self, T, P, sk, x, y, extra):
        self.nb = len(x)
        self.H = T.HostCiphertexts(P, 2 * self.nb + extra)
        init = np.zeros((2 * self.nb + extra, P.n + 1), np.int32)
        init[:self.nb] = sk.encrypt(x)
        init[self.nb:2 * self.nb] = sk.encrypt(y)
        self.H.set(init)
        self.sk = sk

    def x(self, i):
        return self.H.item(i)

    def y(self, i):
        return self.H.item(self.nb + i)

    def r(self, k):
        return self.H.item(2 * self.nb + k)

    def encrypt_into(self, k, bit):
        cur = self.H.get()
        cur[2 * self.nb + k] = self.sk.encrypt(np.array([bit], np.int32))[0]
        self.H.set(cur)

    def bits(self):
        return self.sk.decrypt(self.H.get())

    def close(self):
        self.H.close()


@pytest.mark.parametrize("variant", ["mux", "and_xor"])
def test_sixteen_bit_adder(key, variant):
    """full_adder_MUX / full_adder (test-addition-boot.cpp:30-97), 16 bits: sum
    bits and the carry out decrypt to x + y (test-addition-boot.cpp:175-215)."""
    T, P, sk, keys = key
    L = T.lib()
    nb = 16
    rng = np.random.default_rng(7 if variant == "mux" else 8)
    xv, yv = (int(v) for v in rng.integers(0, 1 << nb, size=2))
    R = _Regs(T, P, sk, _bits_of(xv, nb), _bits_of(yv, nb), extra=nb + 1 + 2 + 3)
    SUM, CARRY, TEMP = 0, nb + 1, nb + 3  # scratch register offsets
    try:
        R.encrypt_into(CARRY, 0)  # first carry 0
        for i in range(nb):
            _gate(T, L, sk, "XOR", R.r(TEMP), R.x(i), R.y(i))
            _gate(T, L, sk, "XOR", R.r(SUM + i), R.r(TEMP), R.r(CARRY))
            _gate(T, L, sk, "AND", R.r(TEMP + 1), R.x(i), R.y(i))
            if variant == "mux":
                L.bootsMUX(R.r(CARRY + 1), R.r(TEMP), R.r(CARRY), R.r(TEMP + 1), sk.cloud)
            else:
                _gate(T, L, sk, "AND", R.r(TEMP + 2), R.r(CARRY), R.r(TEMP))
                _gate(T, L, sk, "XOR", R.r(CARRY + 1), R.r(TEMP + 1), R.r(TEMP + 2))
            L.bootsCOPY(R.r(CARRY), R.r(CARRY + 1), sk.cloud)
        L.bootsCOPY(R.r(SUM + nb), R.r(CARRY), sk.cloud)
        bits = R.bits()
    finally:
        R.close()
    got = sum(int(b) << i for i, b in enumerate(bits[2 * nb + SUM:2 * nb + SUM + nb + 1]))
    assert got == xv + yv, (variant, xv, yv, got)


def test_sixteen_bit_comparison(key):
    """comparison_MUX (test-addition-boot.cpp:100-117): carry starts at 1 and
    becomes MUX(x_i XOR y_i, y_i, carry) from the low bit up; the result is
    [x <= y] as the reference's check computes it (:222-233). Bit-exact
    against the oracle for the last MUX."""
    T, P, sk, keys = key
    L = T.lib()
    nb = 16
    for trial, (xv, yv) in enumerate([(40000, 40000), (513, 40000), (65535, 2), (12345, 12346), (12346, 12345)]):
        R = _Regs(T, P, sk, _bits_of(xv, nb), _bits_of(yv, nb), extra=3)
        CARRY, TEMP = 0, 2
        try:
            R.encrypt_into(CARRY, 1)
            for i in range(nb):
                _gate(T, L, sk, "XOR", R.r(TEMP), R.x(i), R.y(i))
                if i == nb - 1:
                    before = R.H.get()
                L.bootsMUX(R.r(CARRY + 1), R.r(TEMP), R.y(i), R.r(CARRY), sk.cloud)
                L.bootsCOPY(R.r(CARRY), R.r(CARRY + 1), sk.cloud)
            after = R.H.get()
            bits = R.bits()
        finally:
            R.close()
        exp = 1
        for i in range(nb):
            xi, yi = (xv >> i) & 1, (yv >> i) & 1
            exp = yi if xi ^ yi else exp
        assert bits[2 * nb + CARRY] == exp == int(xv <= yv), (trial, xv, yv)
        if trial == 0:
            o = 2 * nb
            ref = O.gates(keys, np.array([T.MUX], np.int32), before[o + TEMP][None], before[nb + nb - 1][None],
                          before[o + CARRY][None], threads=8)
            assert np.array_equal(after[o + CARRY], ref[0])
