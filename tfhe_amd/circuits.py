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


# level schedules of tfhe_amd_netlist_create / tfhe_amd_netlist_levelize; a
# value of 64 or more is the balanced schedule with that base width
ASAP, ALAP, BALANCED = 0, 1, 2


def levelize(circuit, schedule=ASAP):
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
    def __init__(self, cloud, circuit, schedule=ASAP):
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


def random_layered_circuit(rng, depth, width, ops=tuple(range(11))):
    """Layered random circuit (the shape of test/test-long-run.cpp:61-185's
    random gate sequences, arranged in levels): `width` input wires, then
    `depth` layers of `width` gates whose inputs are drawn uniformly from the
    previous layer, gate types uniform over `ops` (MUX takes three inputs).
    Returns the circuit and the wires of the last layer."""
    C = Circuit(width)
    layer = np.arange(width)
    ops = np.asarray(ops, np.int64)
    for _ in range(depth):
        op = rng.choice(ops, size=width)
        ins = rng.choice(layer, size=(3, width))
        first = C.n_wires
        C.op.extend(op.tolist())
        C.in0.extend(ins[0].tolist())
        C.in1.extend(ins[1].tolist())
        C.in2.extend(np.where(op == MUX, ins[2], ins[0]).tolist())
        C.out.extend(range(first, first + width))
        C.n_wires += width
        layer = np.arange(first, first + width)
    return C, layer


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("gates", ctypes.c_int64), ("launches", ctypes.c_int64), ("max_launch_gates", ctypes.c_int64),
                ("queued", ctypes.c_int32), ("active", ctypes.c_int32), ("clients", ctypes.c_int32),
                ("n_streams", ctypes.c_int32), ("wire_bytes_budget", ctypes.c_size_t),
                ("wire_bytes_in_use", ctypes.c_size_t), ("wire_bytes_peak", ctypes.c_size_t),
                ("key_bytes", ctypes.c_size_t)]


_SERVER_SIGS = {
    "tfhe_amd_server_create": (vp, [i32, i32, i32, ctypes.c_size_t]),
    "tfhe_amd_server_add_client": (i32, [vp, CKP]),
    "tfhe_amd_server_submit": (ctypes.c_int64, [vp, i32, i32, i32, i32p, i32p, i32p, i32p, i32p, i32, i32p, i32p,
                                                i32, i32p, i32p]),
    "tfhe_amd_server_wait": (i32, [vp, ctypes.c_int64, i32, ctypes.POINTER(ctypes.c_double)]),
    "tfhe_amd_server_get_stats": (i32, [vp, ctypes.POINTER(ServerStats)]),
    "tfhe_amd_server_destroy": (None, [vp]),
    "tfhe_amd_server_set_packing": (i32, [vp, i32]),
    "tfhe_amd_server_packed_launches": (ctypes.c_int64, [vp]),
}


class Server:
    """Continuous-batching request loop (tfhe_amd_server_*). submit() returns a
    request id; wait() returns (output ciphertexts [n_out][n+1], latency ms)."""

    def __init__(self, n_streams=8, depth=2, max_batch=4096, wire_budget_bytes=1 << 30, pack_min_gates=0):
        L = lib()
        for name, (res, args) in _SERVER_SIGS.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.L = L
        self.ptr = L.tfhe_amd_server_create(n_streams, depth, max_batch, wire_budget_bytes)
        if not self.ptr:
            raise RuntimeError(L.tfhe_amd_last_error().decode())
        if pack_min_gates and L.tfhe_amd_server_set_packing(self.ptr, pack_min_gates) != 0:
            raise RuntimeError(L.tfhe_amd_last_error().decode())
        self._live = {}
        self.width = None

    def add_client(self, cloud, n):
        cid = self.L.tfhe_amd_server_add_client(self.ptr, cloud)
        if cid < 0:
            raise RuntimeError(self.L.tfhe_amd_last_error().decode())
        self.width = n + 1
        return cid

    def submit(self, client, circuit, input_wires, input_cts, output_wires):
        op, in0, in1, in2, out = circuit.arrays()
        iw = np.ascontiguousarray(input_wires, np.int32).reshape(-1)
        ic = np.ascontiguousarray(input_cts, np.int32)
        ow = np.ascontiguousarray(output_wires, np.int32).reshape(-1)
        if self.width is None:
            raise ValueError("submit: no client registered")
        if ic.shape != (len(iw), self.width):
            raise ValueError(f"submit: input_cts has shape {ic.shape}, expected ({len(iw)}, {self.width})")
        dst = np.zeros((len(ow), self.width), np.int32)
        p = lambda a: a.ctypes.data_as(i32p)
        rid = self.L.tfhe_amd_server_submit(self.ptr, client, circuit.n_wires, len(op), p(op), p(in0), p(in1),
                                            p(in2), p(out), len(iw), p(iw), p(ic), len(ow), p(ow), p(dst))
        if rid < 0:
            raise RuntimeError(self.L.tfhe_amd_last_error().decode())
        self._live[rid] = dst
        return rid

    def wait(self, rid, timeout_ms=-1):
        lat = ctypes.c_double()
        rc = self.L.tfhe_amd_server_wait(self.ptr, rid, timeout_ms, ctypes.byref(lat))
        if rc == 1:
            return None
        if rc != 0:
            self._live.pop(rid, None)
            raise RuntimeError(self.L.tfhe_amd_last_error().decode())
        return self._live.pop(rid), lat.value

    def stats(self):
        st = ServerStats()
        check(self.L.tfhe_amd_server_get_stats(self.ptr, ctypes.byref(st)), "server_get_stats")
        out = {name: getattr(st, name) for name, _ in ServerStats._fields_}
        out["packed_launches"] = int(self.L.tfhe_amd_server_packed_launches(self.ptr))
        return out

    def close(self):
        if self.ptr:
            self.L.tfhe_amd_server_destroy(self.ptr)
            self.ptr = None
