"""ctypes wrapper around oracle/liboracle.so (the CPU oracle, test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.environ.get("TFHE_AMD_ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")
REF_HARNESS = os.path.join(REF_DIR, "ref_harness")

GATES = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR", "ANDNY", "ANDYN", "ORNY", "ORYN", "MUX"]
GATE_CODE = {g: i for i, g in enumerate(GATES)}

# plaintext truth tables (libtfhe/boot-gates.cpp comments; test/boots_gates_test.cpp:167-217)
TRUTH = {
    "NAND": lambda a, b, c: 1 - (a & b),
    "AND": lambda a, b, c: a & b,
    "OR": lambda a, b, c: a | b,
    "XOR": lambda a, b, c: a ^ b,
    "XNOR": lambda a, b, c: 1 - (a ^ b),
    "NOR": lambda a, b, c: 1 - (a | b),
    "ANDNY": lambda a, b, c: (1 - a) & b,
    "ANDYN": lambda a, b, c: a & (1 - b),
    "ORNY": lambda a, b, c: (1 - a) | b,
    "ORYN": lambda a, b, c: a | (1 - b),
    "MUX": lambda a, b, c: b if a else c,
}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("N", ctypes.c_int32), ("k", ctypes.c_int32), ("l", ctypes.c_int32),
        ("Bgbit", ctypes.c_int32), ("ks_t", ctypes.c_int32), ("ks_basebit", ctypes.c_int32),
        ("lwe_stdev", ctypes.c_double), ("bk_stdev", ctypes.c_double),
    ]


def p128():
    """Default 128-bit set (libtfhe/tfhe_gate_bootstrapping.cpp:51-76)."""
    return OrcParams(630, 1024, 1, 3, 7, 8, 2, 2.0 ** -15, 2.0 ** -25)


def pbig():
    """N=2048 set: n=750, l=2, Bgbit=10, stdevs 2^-17 / 2^-44; keyswitch t=8, basebit=2."""
    return OrcParams(750, 2048, 1, 2, 10, 8, 2, 2.0 ** -17, 2.0 ** -44)


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        P = ctypes.POINTER
        i32p = P(ctypes.c_int32)
        pp = P(OrcParams)
        sig = {
            "orc_set_seed": (None, [P(ctypes.c_uint32), ctypes.c_int32]),
            "orc_keygen": (None, [pp, i32p, i32p, i32p, i32p]),
            "orc_lwe_encrypt": (None, [i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, i32p]),
            "orc_boots_encrypt": (None, [pp, i32p, ctypes.c_int32, i32p]),
            "orc_lwe_phase": (ctypes.c_int32, [i32p, i32p, ctypes.c_int32]),
            "orc_dtot32": (ctypes.c_int32, [ctypes.c_double]),
            "orc_modswitch_from_torus32": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
            "orc_modswitch_to_torus32": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
            "orc_gate_linear": (None, [ctypes.c_int32, ctypes.c_int32, i32p, i32p, i32p, ctypes.c_int32, i32p]),
            "orc_decompose": (None, [pp, i32p, i32p]),
            "orc_mul_xai": (None, [ctypes.c_int32, i32p, ctypes.c_int32, i32p]),
            "orc_mul_xai_minus_one": (None, [ctypes.c_int32, i32p, ctypes.c_int32, i32p]),
            "orc_negacyclic_mul_add": (None, [i32p, i32p, ctypes.c_int32, i32p]),
            "orc_extprod": (None, [pp, i32p, i32p]),
            "orc_blind_rotate": (None, [pp, i32p, i32p, i32p]),
            "orc_blind_rotate_and_extract": (None, [pp, i32p, i32p, ctypes.c_int32, i32p, i32p]),
            "orc_bootstrap_woks": (None, [pp, i32p, ctypes.c_int32, i32p, i32p]),
            "orc_keyswitch": (None, [pp, i32p, i32p, i32p]),
            "orc_gate": (None, [pp, i32p, i32p, ctypes.c_int32, i32p, i32p, i32p, i32p]),
            "orc_gates": (None, [pp, i32p, i32p, ctypes.c_int32, i32p, i32p, i32p, i32p, i32p, ctypes.c_int32]),
            "orc_fnv1a": (ctypes.c_uint64, [ctypes.c_void_p, ctypes.c_uint64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def fnv1a(a):
    a = np.ascontiguousarray(a)
    return int(lib().orc_fnv1a(a.ctypes.data_as(ctypes.c_void_p), a.nbytes))


def set_seed(seed):
    arr = (ctypes.c_uint32 * 1)(seed)
    lib().orc_set_seed(arr, 1)


class Keys:
    """Secret + cloud key material in flat int32 arrays (oracle layouts)."""

    def __init__(self, P, lwe_key, tlwe_key, ksk, bk):
        self.P, self.lwe_key, self.tlwe_key, self.ksk, self.bk = P, lwe_key, tlwe_key, ksk, bk


def key_shapes(P):
    kpl = (P.k + 1) * P.l
    return dict(
        lwe_key=(P.n,), tlwe_key=(P.k * P.N,),
        ksk=(P.k * P.N, P.ks_t, 1 << P.ks_basebit, P.n + 1),
        bk=(P.n, kpl, P.k + 1, P.N),
    )


def keygen(P, seed=None):
    if seed is not None:
        set_seed(seed)
    sh = key_shapes(P)
    arrs = {k: np.zeros(v, np.int32) for k, v in sh.items()}
    lib().orc_keygen(ctypes.byref(P), ptr(arrs["lwe_key"]), ptr(arrs["tlwe_key"]), ptr(arrs["ksk"]), ptr(arrs["bk"]))
    return Keys(P, **arrs)


def load_ref_keys(P, d):
    sh = key_shapes(P)
    arrs = {k: np.fromfile(os.path.join(d, k + ".bin"), np.int32).reshape(v) for k, v in sh.items()}
    return Keys(P, **arrs)


def boots_encrypt(keys, bit):
    out = np.zeros(keys.P.n + 1, np.int32)
    lib().orc_boots_encrypt(ctypes.byref(keys.P), ptr(keys.lwe_key), int(bit), ptr(out))
    return out


def phase(ct, key):
    ct = np.ascontiguousarray(ct, np.int32)
    return int(lib().orc_lwe_phase(ptr(ct), ptr(key), key.shape[0]))


def phases(cts, key):
    """Vectorised b - <a, s> mod 2^32 for a [count][n+1] array."""
    cts = np.asarray(cts).astype(np.int64)
    ax = (cts[:, :-1] * key.astype(np.int64)[None, :]).sum(axis=1)
    return ((cts[:, -1] - ax) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def gates(keys, ops, in0, in1, in2=None, threads=8):
    P = keys.P
    count = len(ops)
    ops = np.ascontiguousarray(ops, np.int32)
    out = np.zeros((count, P.n + 1), np.int32)
    i2 = ptr(np.ascontiguousarray(in2, np.int32)) if in2 is not None else None
    lib().orc_gates(ctypes.byref(P), ptr(keys.bk), ptr(keys.ksk), count, ptr(ops),
                    ptr(np.ascontiguousarray(in0, np.int32)), ptr(np.ascontiguousarray(in1, np.int32)), i2,
                    ptr(out), threads)
    return out


def keyswitch(keys, u):
    out = np.zeros(keys.P.n + 1, np.int32)
    lib().orc_keyswitch(ctypes.byref(keys.P), ptr(keys.ksk), ptr(np.ascontiguousarray(u, np.int32)), ptr(out))
    return out


def read_ref_gates(path):
    """Parse `ref_harness gates` output."""
    raw = np.fromfile(path, np.uint8)
    count, n = np.frombuffer(raw[:8], np.int32)
    rec = 4 * 4 + 4 * 4 * (n + 1) + 4 + 8
    body = raw[8:8 + count * rec].reshape(count, rec)
    hdr = body[:, :16].copy().view(np.int32)
    cts = body[:, 16:16 + 16 * (n + 1)].copy().view(np.int32).reshape(count, 4, n + 1)
    ph = body[:, 16 + 16 * (n + 1):20 + 16 * (n + 1)].copy().view(np.int32)[:, 0]
    ns = body[:, 20 + 16 * (n + 1):].copy().view(np.int64)[:, 0]
    return dict(ops=hdr[:, 0], bits=hdr[:, 1:4], cts=cts, phase=ph, ns=ns)
