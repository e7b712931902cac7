"""Gate netlists for the level scheduler (tfhe_amd_netlist_*): circuit builders
and a thin wrapper. A netlist is a list of bootstrapped gates over wires; wire w
is slot w of a device ciphertext pool."""
import ctypes

import numpy as np

from . import CKP, DeviceArray, GATE_CODE, check, i32, i32p, lib, vp

XOR, AND, OR, MUX = GATE_CODE["XOR"], GATE_CODE["AND"], GATE_CODE["OR"], GATE_CODE["MUX"]

_SIGS = {
    "tfhe_amd_netlist_create": (vp, [CKP, i32, i32, i32p, i32p, i32p, i32p, i32p, i32]),
    "tfhe_amd_netlist_levels": (i32, [vp, i32p]),
    "tfhe_amd_netlist_levelize": (i32, [i32, i32, i32p, i32p, i32p, i32p, i32p, i32, i32p]),
    "tfhe_amd_netlist_run": (i32, [vp, vp, vp, i32]),
    "tfhe_amd_netlist_destroy": (None, [vp]),
}


def _bind():
    L = lib()
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


class Circuit:
    """Gates over wires; inputs are wires no gate writes."""

    def __init__(self, n_inputs):
        self.n_wires = n_inputs
        self.op, self.in0, self.in1, self.in2, self.out = [], [], [], [], []

    def gate(self, op, a, b, c=None):
        w = self.n_wires
        self.n_wires += 1
        self.op.append(op)
        self.in0.append(a)
        self.in1.append(b)
        self.in2.append(a if c is None else c)
        self.out.append(w)
        return w

    def arrays(self):
        return [np.ascontiguousarray(x, np.int32) for x in (self.op, self.in0, self.in1, self.in2, self.out)]

    def evaluate(self, inputs):
        """Plaintext evaluation (bits) of every wire."""
        v = np.zeros(self.n_wires, np.int64)
        v[:len(inputs)] = inputs
        order = np.argsort(self.out)
        truth = {0: lambda a, b, c: 1 - (a & b), 1: lambda a, b, c: a & b, 2: lambda a, b, c: a | b,
                 3: lambda a, b, c: a ^ b, 4: lambda a, b, c: 1 - (a ^ b), 5: lambda a, b, c: 1 - (a | b),
                 6: lambda a, b, c: (1 - a) & b, 7: lambda a, b, c: a & (1 - b), 8: lambda a, b, c: (1 - a) | b,
                 9: lambda a, b, c: a | (1 - b), 10: lambda a, b, c: b if a else c}
        for g in order:  # gates are appended in topological order, outputs increase
            v[self.out[g]] = truth[self.op[g]](v[self.in0[g]], v[self.in1[g]], v[self.in2[g]])
        return v


def ripple_carry_adders(n_adders, bits):
    """n_adders independent bits-wide adders (test/test-addition-boot.cpp:30-111):
    per bit t1 = a^b, s = t1^c, t2 = a&b, t3 = c&t1, c' = t2|t3 (5 bootstraps).
    Input wires: a[k][i] = k*bits + i, b[k][i] = n_adders*bits + k*bits + i, and one
    constant-0 carry-in wire. Returns the circuit and the sum / carry-out wires."""
    nb = n_adders * bits
    C = Circuit(2 * nb + 1)
    zero = 2 * nb
    sums = np.zeros((n_adders, bits + 1), np.int64)
    carry = [zero] * n_adders
    for i in range(bits):
        for k in range(n_adders):
            a, b = k * bits + i, nb + k * bits + i
            t1 = C.gate(XOR, a, b)
            s = C.gate(XOR, t1, carry[k])
            t2 = C.gate(AND, a, b)
            t3 = C.gate(AND, carry[k], t1)
            carry[k] = C.gate(OR, t2, t3)
            sums[k, i] = s
    for k in range(n_adders):
        sums[k, bits] = carry[k]
    return C, sums, zero


def levelize(circuit, schedule=0):
    """Host-only levelisation: (number of levels, level of every gate)."""
    L = _bind()
    op, in0, in1, in2, out = circuit.arrays()
    lv = np.zeros(len(op), np.int32)
    p = lambda a: a.ctypes.data_as(i32p)
    n = L.tfhe_amd_netlist_levelize(circuit.n_wires, len(op), p(op), p(in0), p(in1), p(in2), p(out), schedule,
                                    p(lv))
    if n < 0:
        raise RuntimeError(lib().tfhe_amd_last_error().decode())
    return n, lv


class Netlist:
    def __init__(self, cloud, circuit, schedule=0):
        L = _bind()
        op, in0, in1, in2, out = circuit.arrays()
        self.n_wires = circuit.n_wires
        self.n_gates = len(op)
        p = lambda a: a.ctypes.data_as(i32p)
        self.ptr = L.tfhe_amd_netlist_create(cloud, self.n_wires, self.n_gates, p(op), p(in0), p(in1), p(in2),
                                             p(out), schedule)
        if not self.ptr:
            raise RuntimeError(lib().tfhe_amd_last_error().decode())

    def levels(self):
        L = _bind()
        n = L.tfhe_amd_netlist_levels(self.ptr, None)
        w = np.zeros(n, np.int32)
        L.tfhe_amd_netlist_levels(self.ptr, w.ctypes.data_as(i32p))
        return w

    def run(self, d_pool, use_graph=True, stream=None):
        check(_bind().tfhe_amd_netlist_run(self.ptr, d_pool.ptr, stream, 1 if use_graph else 0), "netlist_run")

    def close(self):
        if self.ptr:
            _bind().tfhe_amd_netlist_destroy(self.ptr)
            self.ptr = None
