"""Host-side checks of libtfhe_amd.so that need no GPU: every symbol the
public header declares is exported, key generation and encryption replay the
reference RNG bit-exactly (same as the oracle), the non-bootstrapped gates,
argument errors, and a C program built against include/tfhe_amd.h."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_py as O
import tfhe_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tfhe_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^TFHE_AMD_EXPORT[^;(]*?\b(\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) > 60
    L = T.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_keygen_matches_oracle_bit_exact():
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=0)
    K = sk.export()
    R = O.keygen(O.p128(), 0)
    for k in ("lwe_key", "tlwe_key", "ksk", "bk"):
        assert np.array_equal(K[k], getattr(R, k)), k
    # encryption continues the same generator
    bits = np.array([1, 0, 0, 1, 1], np.int32)
    cts = sk.encrypt(bits)
    for b, c in zip(bits, cts):
        assert np.array_equal(c, O.boots_encrypt(R, b))
    assert np.array_equal(sk.decrypt(cts), bits)
    sk.close()


def test_parameter_presets():
    P = T.default_params(110)
    g = P.ptr.contents.tgsw_params.contents
    assert (P.n, P.N, P.k, P.l, P.Bgbit, P.ks_t, P.ks_basebit) == (630, 1024, 1, 3, 7, 8, 2)
    assert [g.h[i] for i in range(3)] == [1 << 25, 1 << 18, 1 << 11]
    assert g.offset == (1 << 31) + (1 << 24) + (1 << 17)
    P80 = T.default_params(80)
    assert (P80.n, P80.l, P80.Bgbit) == (500, 2, 10)
    Pb = T.big_params()
    assert (Pb.n, Pb.N, Pb.l, Pb.Bgbit) == (750, 2048, 2, 10)


def test_lambda_above_128_aborts():
    code = ("import sys; sys.path.insert(0, %r); import tfhe_amd as T; "
            "T.lib().new_default_gate_bootstrapping_parameters(200)") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
    assert r.returncode != 0 and "only implemented for 80bit and 128bit" in r.stderr


def test_batch_rejects_bad_gate_code_without_touching_the_gpu():
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=1)
    H = T.HostCiphertexts(P, 2)
    ops = np.array([0, 42], np.int32)
    arr = (T.LSP * 2)(H.item(0), H.item(1))
    rc = T.lib().bootsGate_batch(2, ops.ctypes.data_as(T.i32p), arr, arr, arr, None, sk.cloud)
    assert rc == -1 and b"bad gate code" in T.lib().tfhe_amd_last_error()
    H.close()
    sk.close()


def test_server_calls_reject_a_missing_server():
    """The serving entry points check their handle before anything else: a null
    server is an error with a reason, not a device call."""
    L = T.lib()
    L.tfhe_amd_server_set_packing.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.tfhe_amd_server_set_packing.restype = ctypes.c_int32
    L.tfhe_amd_server_packed_launches.argtypes = [ctypes.c_void_p]
    L.tfhe_amd_server_packed_launches.restype = ctypes.c_int64
    assert L.tfhe_amd_server_set_packing(None, 1024) == -1
    assert b"set_packing" in L.tfhe_amd_last_error()
    assert L.tfhe_amd_server_packed_launches(None) == -1
    assert b"packed_launches" in L.tfhe_amd_last_error()


def test_modswitch_and_dtot32_match_oracle():
    L = T.lib()
    rng = np.random.default_rng(2)
    for v in rng.integers(-2 ** 31, 2 ** 31, size=300):
        for m in (8, 2048, 4096, 7):
            assert L.modSwitchFromTorus32(int(v), m) == O.lib().orc_modswitch_from_torus32(int(v), m)
    for d in (0.0, 0.5, -0.5, 0.25, -0.25, 0.123456, -3.75):
        assert L.dtot32(d) == O.lib().orc_dtot32(d)


def test_c_program_host_part(tmp_path):
    exe = str(tmp_path / "c_binding")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-o", exe, os.path.join(ROOT, "tests", "c", "c_binding.c"),
                    "-I", os.path.join(ROOT, "include"), "-L", os.path.join(ROOT, "tfhe_amd"), "-ltfhe_amd",
                    "-Wl,-rpath," + os.path.join(ROOT, "tfhe_amd")], check=True)
    r = subprocess.run([exe, "host"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_philox_known_answers(tmp_path):
    """The device encryption's generator reproduces Philox4x32-10's published
    known-answer vectors (compiled for the host from the same header)."""
    exe = str(tmp_path / "philox_kat")
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "philox_kat.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, src], check=True)
    assert subprocess.run([exe], capture_output=True, text=True, check=True).stdout.strip() == "OK"


def test_struct_layout_matches_reference_headers(tmp_path):
    """Every public struct has the size and field offsets the reference's own
    headers give it (include/*.h of tfhe: lweparams.h, tlwe.h, tgsw.h:78-96,
    lwebootstrappingkey.h:36-59, ...), so a program compiled against those
    headers finds every field where it expects it. The expected table was written
    from the reference headers (tests/golden/make_golden.py layout); where they
    are present it is rebuilt from them here too."""
    src = os.path.join(ROOT, "tests", "c", "struct_layout.c")
    exe = str(tmp_path / "struct_layout")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                   check=True)
    ours = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    golden = open(os.path.join(ROOT, "tests", "golden", "ref_struct_layout.txt")).read()
    assert ours == golden
    ref_inc = "/root/reference/src/include"
    if os.path.isdir(ref_inc):
        exr = str(tmp_path / "struct_layout_ref")
        subprocess.run(["g++", "-std=c++11", "-x", "c++", "-DREFERENCE_HEADERS", "-I", ref_inc, src, "-o", exr],
                       check=True)
        assert subprocess.run([exr], check=True, capture_output=True, text=True).stdout == golden
    sizes = dict(line.split(" size ") for line in golden.splitlines() if " size " in line)
    assert sizes["TGswSampleFFT"] == "24" and sizes["LweBootstrappingKeyFFT"] == "48"
