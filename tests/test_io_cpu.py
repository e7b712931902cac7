"""Key, parameter and ciphertext files in the reference's format
(libtfhe/tfhe_io.cpp). Pinned three ways:
  - files the reference wrote (tests/golden/ref_io_p128_seed11: parameters,
    LWE key, ciphertexts of recorded bits, a NAND of two of them) parse to the
    right values, decrypt to the recorded bits, and are rewritten byte for byte;
  - where the reference build is present (oracle/_ref), its full secret and
    cloud key files load into keys equal to the reference's own key dump and
    are rewritten byte for byte, and the reference's readers accept this
    library's files (decryptions, a bootstrapped NAND and key hashes agree);
  - this library's own files round-trip for both parameter sets.
No GPU is used."""
import ctypes
import filecmp
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import tfhe_amd as T

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_io_p128_seed11")
needs_ref = pytest.mark.skipif(not os.path.exists(O.REF_HARNESS), reason="reference build (oracle/_ref) not present")


class LweKeyS(ctypes.Structure):
    _fields_ = [("params", ctypes.POINTER(T.LweParams)), ("key", ctypes.POINTER(ctypes.c_int32))]


def _load_lwe_key(path):
    with T._CFile(path, "rb") as f:
        kp = T.lib().new_lweKey_fromFile(f)
    assert kp, T.lib().tfhe_amd_last_error().decode()
    k = ctypes.cast(kp, ctypes.POINTER(LweKeyS)).contents
    n = k.params.contents.n
    key = np.ctypeslib.as_array(k.key, shape=(n,)).copy()
    return kp, key


def _phases(cts, key):
    n = len(key)
    return (cts[:, n].astype(np.uint32) - (cts[:, :n].astype(np.uint32) * key.astype(np.uint32)).sum(
        axis=1, dtype=np.uint64).astype(np.uint32)).astype(np.int32)


def test_reference_written_files_parse_and_rewrite(tmp_path):
    P = T.load_params(os.path.join(GOLD, "params.gb"))
    assert (P.n, P.N, P.k, P.l, P.Bgbit, P.ks_t, P.ks_basebit) == (630, 1024, 1, 3, 7, 8, 2)
    pp = P.ptr.contents
    # the reference prints the noise parameters with 8 decimals (2^-15 -> 0.00003052)
    assert pp.in_out_params.contents.alpha_min == 3.052e-05
    assert pp.tgsw_params.contents.tlwe_params.contents.alpha_min == 3e-08
    assert pp.in_out_params.contents.alpha_max == 0.012467
    T.save_params(P, tmp_path / "params.gb")
    assert filecmp.cmp(os.path.join(GOLD, "params.gb"), tmp_path / "params.gb", shallow=False)

    kp, key = _load_lwe_key(os.path.join(GOLD, "lwekey.bin"))
    try:
        assert set(np.unique(key)) <= {0, 1}
        with T._CFile(tmp_path / "lwekey.bin", "wb") as f:
            T.lib().export_lweKey_toFile(f, kp)
        assert filecmp.cmp(os.path.join(GOLD, "lwekey.bin"), tmp_path / "lwekey.bin", shallow=False)
    finally:
        T.lib().delete_LweKey(kp)

    bits = np.array([int(c) for c in open(os.path.join(GOLD, "bits.txt")).read().strip()])
    cts, var = T.load_ciphertexts(P, os.path.join(GOLD, "cts.bin"), len(bits))
    assert np.array_equal((_phases(cts, key) > 0).astype(int), bits)
    assert np.allclose(var, 2.0 ** -30)  # alpha^2 of a fresh encryption
    nand, _ = T.load_ciphertexts(P, os.path.join(GOLD, "nand.bin"), 1)
    assert int(_phases(nand, key)[0] > 0) == 1 - (bits[0] & bits[1])
    T.save_ciphertexts(P, cts, tmp_path / "cts.bin", var)
    assert filecmp.cmp(os.path.join(GOLD, "cts.bin"), tmp_path / "cts.bin", shallow=False)


@pytest.mark.parametrize("which", ["p128", "pbig"])
def test_own_files_round_trip(tmp_path, which):
    P = T.default_params(110) if which == "p128" else T.big_params()
    sk = T.SecretKey(P, seed=4)
    try:
        sk.save(tmp_path / "secret.key")
        T.save_cloud_key(sk.cloud, tmp_path / "cloud.key")
        T.save_params(P, tmp_path / "params.gb")
        bits = np.arange(20) % 2
        cts = sk.encrypt(bits)
        T.save_ciphertexts(P, cts, tmp_path / "cts.bin")
        sk2 = T.SecretKey.load(tmp_path / "secret.key")
        a, b = sk.export(), sk2.export()
        for k in a:
            assert np.array_equal(a[k], b[k]), k
        P2 = T.load_params(tmp_path / "params.gb")
        assert (P2.n, P2.N, P2.k, P2.l, P2.Bgbit, P2.ks_t, P2.ks_basebit) == \
            (P.n, P.N, P.k, P.l, P.Bgbit, P.ks_t, P.ks_basebit)
        cts2, _ = T.load_ciphertexts(P2, tmp_path / "cts.bin", len(bits))
        assert np.array_equal(cts2, cts)
        assert np.array_equal(sk2.decrypt(cts2), bits)
        ck = T.load_cloud_key(tmp_path / "cloud.key")
        T.save_cloud_key(ck, tmp_path / "cloud2.key")
        assert filecmp.cmp(tmp_path / "cloud.key", tmp_path / "cloud2.key", shallow=False)
        T.lib().delete_gate_bootstrapping_cloud_keyset(ck)
        sk2.close()
    finally:
        sk.close()


def test_malformed_files_are_rejected(tmp_path):
    P = T.default_params(110)
    T.save_params(P, tmp_path / "params.gb")
    text = open(tmp_path / "params.gb").read()
    (tmp_path / "wrong.gb").write_text(text.replace("GATEBOOTSPARAMS", "LWEPARAMS", 2))
    with pytest.raises(T.TfheAmdError, match="expected section GATEBOOTSPARAMS"):
        T.load_params(tmp_path / "wrong.gb")
    (tmp_path / "short.gb").write_text(text[:len(text) // 2])
    with pytest.raises(T.TfheAmdError):
        T.load_params(tmp_path / "short.gb")
    # a cloud key file cut inside the keyswitch key
    sk = T.SecretKey(P, seed=5)
    try:
        T.save_cloud_key(sk.cloud, tmp_path / "cloud.key")
    finally:
        sk.close()
    data = open(tmp_path / "cloud.key", "rb").read()
    (tmp_path / "cut.key").write_bytes(data[:len(data) // 3])
    with pytest.raises(T.TfheAmdError, match="truncated"):
        T.load_cloud_key(tmp_path / "cut.key")


@needs_ref
def test_reference_key_files_load_and_rewrite_bytes(tmp_path):
    ref = tmp_path / "ref"
    dump = tmp_path / "dump"
    ref.mkdir()
    dump.mkdir()
    subprocess.run([O.REF_HARNESS, "files", "11", str(ref), "16"], check=True)
    subprocess.run([O.REF_HARNESS, "keys", "11", str(dump)], check=True)
    R = O.load_ref_keys(O.p128(), str(dump))
    sk = T.SecretKey.load(ref / "secret.key")
    try:
        K = sk.export()
        assert np.array_equal(K["lwe_key"], R.lwe_key)
        assert np.array_equal(K["tlwe_key"], R.tlwe_key)
        assert np.array_equal(K["bk"], R.bk)
        assert np.array_equal(K["ksk"], R.ksk)
        sk.save(tmp_path / "secret.key")
        assert filecmp.cmp(ref / "secret.key", tmp_path / "secret.key", shallow=False)
        T.save_cloud_key(sk.cloud, tmp_path / "cloud.key")
        assert filecmp.cmp(ref / "cloud.key", tmp_path / "cloud.key", shallow=False)
        bits = np.array([int(c) for c in open(ref / "bits.txt").read().strip()])
        cts, _ = T.load_ciphertexts(sk.params, ref / "cts.bin", 16)
        assert np.array_equal(sk.decrypt(cts), bits)
    finally:
        sk.close()


@needs_ref
def test_reference_reads_this_librarys_files(tmp_path):
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=12)
    try:
        T.save_params(P, tmp_path / "params.gb")
        sk.save(tmp_path / "secret.key")
        T.save_cloud_key(sk.cloud, tmp_path / "cloud.key")
        bits = np.array([1, 1, 0, 1, 0, 0, 1, 0, 1, 1, 1, 0, 0, 0, 1, 0])
        T.save_ciphertexts(P, sk.encrypt(bits), tmp_path / "cts.bin", np.full(16, 2.0 ** -30))
        K = sk.export()
    finally:
        sk.close()
    out = subprocess.run([O.REF_HARNESS, "readfiles", "0", str(tmp_path), "16"], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    vals = dict(line.split(" ", 1) for line in out if " " in line)
    assert vals["bits"] == "".join(map(str, bits))
    assert int(vals["nand"]) == 1 - (bits[0] & bits[1])
    assert int(vals["ks_hash"]) == O.fnv1a(K["ksk"])
    assert int(vals["bk_hash"]) == O.fnv1a(K["bk"])
    assert vals["params"].split()[:7] == ["630", "1024", "1", "3", "7", "8", "2"]


# ---------------------------------------- every object of the format (n = 16 key set)

GOLD16 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_io_n16_seed7")


@pytest.fixture(scope="module")
def io_components(tmp_path_factory):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path_factory.mktemp("ioc") / "io_components")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-o", exe,
                    os.path.join(root, "tests", "c", "io_components.cpp"), "-I", os.path.join(root, "include"),
                    "-L", os.path.join(root, "tfhe_amd"), "-ltfhe_amd",
                    "-Wl,-rpath," + os.path.join(root, "tfhe_amd")], check=True)
    return exe


def test_every_object_file_and_stream_form_matches_reference_bytes(io_components, tmp_path):
    """Files the reference wrote for one key set (parameters of each kind, LWE /
    TLWE / TGSW keys and samples, the secret key set): each is read through the
    FILE* reader and written through the std::ostream writer, and read through the
    std::istream reader and written through the FILE* writer; both copies equal
    the reference's bytes. The keyswitch key, bootstrapping key and cloud key
    set this library writes from the loaded secret key set have the SHA-256
    digests of the reference's own files."""
    import hashlib
    import json
    out = subprocess.run([io_components, GOLD16, str(tmp_path)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    assert out.stdout.strip().endswith("done")
    assert out.stdout.count("ok ") >= 40
    digests = json.load(open(os.path.join(GOLD16, "digests.json")))
    for name, d in digests.items():
        data = open(tmp_path / name, "rb").read()
        assert len(data) == d["bytes"] and hashlib.sha256(data).hexdigest() == d["sha256"], name


def test_n16_reference_gates_decrypt_under_the_loaded_key():
    """The reference's gate outputs under the n = 16 key set decrypt, with the
    secret key set read from its file, to the plaintext gates."""
    sk = T.SecretKey.load(os.path.join(GOLD16, "secret.key"))
    try:
        g = np.load(os.path.join(GOLD16, "gates.npz"))
        assert sk.params.n == 16
        bits = g["bits"]
        assert np.array_equal(sk.decrypt(g["inputs"].reshape(-1, 17)).reshape(-1, 3), bits)
        exp = np.array([O.TRUTH[O.GATES[o]](*b) for o, b in zip(g["ops"], bits)])
        assert np.array_equal(sk.decrypt(g["outputs"]), exp)
        assert np.array_equal(sk.phases(g["outputs"]), g["phase"])
    finally:
        sk.close()


@pytest.mark.parametrize("field,value,reason", [
    ("N", "1000", "power of two"),
    ("N", "1048576", "out of range"),
    ("k", "0", "out of range"),
    ("l", "9", "out of range"),
    ("n", "-3", "out of range"),
    ("n", "99999999999", "out of range"),
])
def test_malformed_headers_are_rejected_before_allocating(tmp_path, field, value, reason):
    """Parameter values outside the readers' limits (a non-power-of-two or huge
    N, k = 0, l Bgbit > 32, a negative or huge n) are rejected with a reason
    and NULL, before any key material is allocated."""
    import re
    text = open(os.path.join(GOLD16, "secret.key"), "rb").read()
    head_end = text.index(b"-----END TGSWPARAMS-----\n") + len(b"-----END TGSWPARAMS-----\n")
    head = text[:head_end].decode()
    section = {"N": "TLWEPARAMS", "k": "TLWEPARAMS", "l": "TGSWPARAMS", "n": "LWEPARAMS"}[field]
    a = head.index(f"-----BEGIN {section}-----")
    b = head.index(f"-----END {section}-----")
    body = re.sub(rf"^{field}: .*$", f"{field}: {value}", head[a:b], flags=re.M)
    bad = head[:a] + body + head[b:]
    (tmp_path / "bad.key").write_bytes(bad.encode() + text[head_end:])
    with T._CFile(tmp_path / "bad.key", "rb") as f:
        p = T.lib().new_tfheGateBootstrappingSecretKeySet_fromFile(f)
    assert not p
    assert reason in T.lib().tfhe_amd_last_error().decode()
