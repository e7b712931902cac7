"""Host-side levelisation of the level scheduler (no GPU)."""
import numpy as np
import pytest

from tfhe_amd.circuits import Circuit, levelize, ripple_carry_adders


def test_adder_levels_asap_and_alap():
    n_adders, bits = 512, 32
    C, sums, zero = ripple_carry_adders(n_adders, bits)
    n, lv = levelize(C, 0)
    assert n == 2 * bits + 1
    widths = np.bincount(lv, minlength=n + 1)[1:]
    # ASAP: every a^b and a&b first, then per bit (s, t3) and the carry
    assert widths[0] == 2 * n_adders * bits
    assert all(w == 2 * n_adders for w in widths[1::2]) and all(w == n_adders for w in widths[2::2])
    n2, lv2 = levelize(C, 1)
    w2 = np.bincount(lv2, minlength=n2 + 1)[1:]
    assert n2 == n and w2.sum() == 5 * n_adders * bits
    assert w2.max() <= 3 * n_adders  # t1, t2 of a bit move next to the carry that needs them


def test_levels_respect_dependencies():
    rng = np.random.default_rng(0)
    C = Circuit(10)
    wires = list(range(10))
    for _ in range(200):
        op = int(rng.integers(0, 11))
        a, b, c = (int(v) for v in rng.choice(wires, 3))
        wires.append(C.gate(op, a, b, c if op == 10 else None))
    n0 = levelize(C, 0)[0]
    for sched in (0, 1, 2, 64, 4096):
        n, lv = levelize(C, sched)
        assert n == n0
        prod = {w: g for g, w in enumerate(C.out)}
        for g in range(len(C.op)):
            ins = [C.in0[g], C.in1[g]] + ([C.in2[g]] if C.op[g] == 10 else [])
            for w in ins:
                if w in prod:
                    assert lv[prod[w]] < lv[g]
        assert lv.min() >= 1 and lv.max() == n


def _deadlines(C, depth):
    """As-late-as-possible level of every gate, gates nobody reads at the last level."""
    prod = {w: g for g, w in enumerate(C.out)}
    users = [[] for _ in C.op]
    for g in range(len(C.op)):
        ins = [C.in0[g], C.in1[g]] + ([C.in2[g]] if C.op[g] == 10 else [])
        for w in ins:
            if w in prod:
                users[prod[w]].append(g)
    dl = np.full(len(C.op), depth)
    for g in range(len(C.op) - 1, -1, -1):  # gates are appended after their inputs
        for u in users[g]:
            dl[g] = min(dl[g], dl[u] - 1)
    return dl


@pytest.mark.parametrize("base", [64, 1024, 2048])
def test_balanced_schedule_widths(base):
    """Balanced levels on the 512 adders and on a random circuit with MUX gates:
    the as-soon-as-possible depth, no gate after its deadline, and every level's
    bootstraps (MUX = 2) within the smallest base * 2^k that holds the gates whose
    deadline it is. On the adders with base 1024 (the SIMD count of a 256-CU GPU)
    the carry chains fill 64 levels of exactly 1024 bootstraps and the sum bits,
    which nothing reads, fill the last level."""
    rng = np.random.default_rng(base)
    R = Circuit(16)
    wires = list(range(16))
    for _ in range(3000):
        op = int(rng.integers(0, 11))
        a, b, c = (int(v) for v in rng.choice(wires[-200:], 3))
        wires.append(R.gate(op, a, b, c if op == 10 else None))
    A = ripple_carry_adders(512, 32)[0]
    for C in (A, R):
        depth = levelize(C, 0)[0]
        n, lv = levelize(C, base)
        assert n == depth
        dl = _deadlines(C, depth)
        assert (lv <= dl).all()
        wt = np.where(np.asarray(C.op) == 10, 2, 1)
        for level in range(1, n + 1):
            sel = lv == level
            forced = int(wt[sel & (dl == level)].sum())
            cap = base
            while cap < forced:
                cap *= 2
            assert wt[sel].sum() <= cap, (level, forced, wt[sel].sum(), cap)
    if base == 1024:
        w = np.bincount(levelize(A, base)[1])[1:]
        assert (w[:-1] == 1024).all() and w[-1] == 512 * 32


def test_levelize_errors():
    C = Circuit(2)
    C.gate(0, 0, 1)
    C.in1[0] = 2  # reads its own output
    with pytest.raises(RuntimeError, match="cycle"):
        levelize(C)
    C = Circuit(2)
    C.gate(0, 0, 5)
    with pytest.raises(RuntimeError, match="out of range"):
        levelize(C)
    C = Circuit(2)
    C.gate(0, 0, 1)
    C.gate(1, 0, 1)
    C.out[1] = C.out[0]
    with pytest.raises(RuntimeError, match="two gates"):
        levelize(C)
    C = Circuit(2)
    C.gate(0, 0, 1)
    C.op[0] = 17
    with pytest.raises(RuntimeError, match="bad gate code"):
        levelize(C)
    C = Circuit(2)
    C.gate(0, 0, 1)
    for bad in (-1, 3, 63):
        with pytest.raises(RuntimeError, match="unknown schedule"):
            levelize(C, bad)


def test_mux_without_third_input_is_rejected():
    """A MUX gate in a netlist whose in2 array is NULL is an error (not MUX(a, b, a))."""
    import ctypes
    from tfhe_amd import i32p, lib
    from tfhe_amd.circuits import _bind
    L = _bind()
    op = np.array([10], np.int32)
    in0, in1, out = (np.array([v], np.int32) for v in (0, 1, 3))
    p = lambda a: a.ctypes.data_as(i32p)
    lv = np.zeros(1, np.int32)
    n = L.tfhe_amd_netlist_levelize(4, 1, p(op), p(in0), p(in1), None, p(out), 0, p(lv))
    assert n < 0
    assert "MUX needs a third input" in lib().tfhe_amd_last_error().decode()
    # the same netlist with in2 given levelises
    in2 = np.array([2], np.int32)
    assert L.tfhe_amd_netlist_levelize(4, 1, p(op), p(in0), p(in1), p(in2), p(out), 0, p(lv)) == 1


class _FakeDevice:
    """Stands in for a DeviceArray in argument checks that fail before any device call."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.ptr = None


def test_python_wrappers_check_buffer_shapes():
    import tfhe_amd as T
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=3)
    try:
        with pytest.raises(ValueError, match="rows of n\\+1"):
            T.encrypt_device(sk, _FakeDevice((10,)), _FakeDevice((9, P.n + 1)), seed=1)
        with pytest.raises(ValueError, match="exceeds"):
            T.encrypt_device(sk, _FakeDevice((10,)), _FakeDevice((20, P.n + 1)), seed=1, count=11)
        with pytest.raises(ValueError, match="outside d_cts"):
            T.decrypt_device(sk, _FakeDevice((10, P.n + 1)), _FakeDevice((20,)), count=5, offset_rows=6)
        with pytest.raises(ValueError, match="d_bits holds fewer"):
            T.decrypt_device(sk, _FakeDevice((10, P.n + 1)), _FakeDevice((4,)))
    finally:
        sk.close()
