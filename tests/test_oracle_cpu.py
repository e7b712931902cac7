"""CPU checks of the oracle (oracle/tfhe_oracle.cpp) against the reference's
golden vectors (tests/golden/, written by tests/golden/make_golden.py from the
reference's own build) and against closed forms, plus a lane-by-lane CPU
emulation of the GPU transform. No GPU needed.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
L = O.lib()


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def i32(x):
    return np.asarray(x, dtype=np.int64).astype(np.uint32).view(np.int32)


# ---------------------------------------------------------------- closed forms


def test_dtot32_known_answers():
    # test/arithmetic_test.cpp:44-50
    assert L.orc_dtot32(0.0) == 0
    assert L.orc_dtot32(0.5) == -(1 << 31)
    assert L.orc_dtot32(-0.5) == -(1 << 31)
    assert L.orc_dtot32(0.25) == 1 << 30
    assert L.orc_dtot32(-0.25) == -(1 << 30)


def test_modswitch_is_round_to_nearest():
    # test/arithmetic_test.cpp:66-86, and the exact closed form round(v*M/2^32) mod M
    rng = np.random.default_rng(0)
    for msize in [2, 3, 7, 8, 199, 2048, 4096]:
        for v in rng.integers(-2 ** 31, 2 ** 31, size=200):
            w = L.orc_modswitch_from_torus32(int(v), msize)
            exact = (int(v) % 2 ** 32) * msize / 2 ** 32
            assert 0 <= w < msize
            d = abs(((exact - w + msize / 2) % msize) - msize / 2)
            assert d <= 0.5 + 1e-9
        for j in range(min(msize, 50)):
            t = L.orc_modswitch_to_torus32(j, msize)
            assert abs(((t % 2 ** 32) / 2 ** 32 - j / msize + 0.5) % 1 - 0.5) <= 1 / (2 * msize)
    assert L.orc_modswitch_to_torus32(1, 8) == 1 << 29
    assert L.orc_modswitch_to_torus32(-1, 8) == -(1 << 29)
    assert L.orc_modswitch_to_torus32(1, 4) == 1 << 30


@pytest.mark.parametrize("Pf", [O.p128, O.pbig])
def test_decomposition(Pf):
    # recomposition bound of test/tgsw_test.cpp:581-612, digit range, and the truncation direction
    P = Pf()
    rng = np.random.default_rng(1)
    poly = i32(rng.integers(-2 ** 31, 2 ** 31, size=P.N))
    poly[:4] = [0, -1, 2 ** 31 - 1, -2 ** 31]
    dig = np.zeros(P.l * P.N, np.int32)
    L.orc_decompose(ctypes.byref(P), O.ptr(poly), O.ptr(dig))
    dig = dig.reshape(P.l, P.N).astype(np.int64)
    half = 1 << (P.Bgbit - 1)
    assert dig.min() >= -half and dig.max() < half
    h = np.array([1 << (32 - (p + 1) * P.Bgbit) for p in range(P.l)], np.int64)
    rec = (dig * h[:, None]).sum(axis=0)
    err = (poly.astype(np.int64) - rec) % 2 ** 32
    assert err.max() < (1 << (32 - P.l * P.Bgbit))  # truncation: x - sum d h in [0, 2^(32 - l Bgbit))


def _negacyclic_bruteforce(a, b):
    N = len(a)
    out = [0] * N
    for i in range(N):
        for j in range(N):
            k = i + j
            if k < N:
                out[k] += int(a[i]) * int(b[j])
            else:
                out[k - N] -= int(a[i]) * int(b[j])
    return i32([v % 2 ** 32 for v in out])


@pytest.mark.parametrize("N", [8, 16, 64])
def test_negacyclic_product_bruteforce(N):
    rng = np.random.default_rng(N)
    a = rng.integers(-64, 64, size=N).astype(np.int32)
    b = i32(rng.integers(-2 ** 31, 2 ** 31, size=N))
    acc = np.zeros(N, np.int32)
    L.orc_negacyclic_mul_add(O.ptr(a), O.ptr(b), N, O.ptr(acc))
    assert np.array_equal(acc, _negacyclic_bruteforce(a, b))


def test_rotation_identities():
    # test/polynomial_test.cpp:304-390: X^a X^b = X^(a+b mod 2N), X^N = -1, (X^a - 1) p = X^a p - p
    N = 64
    rng = np.random.default_rng(3)
    p = i32(rng.integers(-2 ** 31, 2 ** 31, size=N))
    def xai(a, q):
        out = np.zeros(N, np.int32)
        L.orc_mul_xai(int(a), O.ptr(np.ascontiguousarray(q)), N, O.ptr(out))
        return out
    for a, b in [(0, 5), (3, 70), (63, 64), (100, 127), (64, 64)]:
        assert np.array_equal(xai(b, xai(a, p)), xai((a + b) % (2 * N), p))
    assert np.array_equal(xai(N, p), i32(-p.astype(np.int64)))
    mono = np.zeros(N, np.int32)
    mono[0] = 1
    assert np.array_equal(xai(5, p), _negacyclic_bruteforce(xai(5, mono), p))
    out = np.zeros(N, np.int32)
    L.orc_mul_xai_minus_one(37, O.ptr(p), N, O.ptr(out))
    assert np.array_equal(out, i32(xai(37, p).astype(np.int64) - p))


def _small_params(n):
    P = O.p128()
    P.n = n
    return P


def test_blind_rotation_noiseless_closed_form():
    # BK_i = s_i * H (no mask, no noise): ACC = X^{sum bara_i s_i} * ACC exactly when the
    # accumulator's coefficients are multiples of 2^(32 - l*Bgbit) (test/bootstrapping_test_fft.cpp:122-136)
    P = _small_params(6)
    N, l, kpl = P.N, P.l, 6
    rng = np.random.default_rng(4)
    s = rng.integers(0, 2, size=P.n).astype(np.int32)
    bk = np.zeros((P.n, kpl, 2, N), np.int32)
    for r in range(kpl):
        blk, p = divmod(r, l)
        bk[:, r, blk, 0] = s * (1 << (32 - (p + 1) * P.Bgbit))
    acc = i32(rng.integers(-2 ** 20, 2 ** 20, size=(2, N)) * 2 ** 11)
    bara = rng.integers(0, 2 * N, size=P.n).astype(np.int32)
    bara[0] = 0
    out = acc.copy()
    L.orc_blind_rotate(ctypes.byref(P), O.ptr(bk), O.ptr(bara), O.ptr(out))
    e = int((bara.astype(np.int64) * s).sum() % (2 * N))
    for c in range(2):
        exp = np.zeros(N, np.int32)
        L.orc_mul_xai(e, O.ptr(np.ascontiguousarray(acc[c])), N, O.ptr(exp))
        assert np.array_equal(out[c], exp)


def test_external_product_with_gadget_identity():
    # a noiseless TGSW(1) row (H itself) returns the gadget recomposition of the input
    P = O.p128()
    N, l = P.N, P.l
    row = np.zeros((6, 2, N), np.int32)
    for r in range(6):
        blk, p = divmod(r, l)
        row[r, blk, 0] = 1 << (32 - (p + 1) * P.Bgbit)
    rng = np.random.default_rng(5)
    acc = i32(rng.integers(-2 ** 31, 2 ** 31, size=(2, N)))
    out = acc.copy()
    L.orc_extprod(ctypes.byref(P), O.ptr(row), O.ptr(out))
    for c in range(2):
        dig = np.zeros(l * N, np.int32)
        L.orc_decompose(ctypes.byref(P), O.ptr(np.ascontiguousarray(acc[c])), O.ptr(dig))
        h = np.array([1 << (32 - (p + 1) * P.Bgbit) for p in range(l)], np.int64)
        rec = i32((dig.reshape(l, N).astype(np.int64) * h[:, None]).sum(axis=0))
        assert np.array_equal(out[c], rec)


def test_keyswitch_noiseless_closed_form():
    # test/lwekeyswitch_test.cpp:96-152: with noiseless rows (a = 0, b = h s_i / base^(j+1))
    # the result is b - sum_i s_i * round_t(a_i), round_t keeping the top basebit*t bits
    P = O.p128()
    n, nin, t = P.n, P.N, P.ks_t
    rng = np.random.default_rng(6)
    s = rng.integers(0, 2, size=nin)
    ksk = np.zeros((nin, t, 4, n + 1), np.int64)
    for j in range(t):
        for hh in range(4):
            ksk[:, j, hh, n] = (s * hh * (1 << (32 - 2 * (j + 1)))) % 2 ** 32
    keys = O.Keys(P, None, None, i32(ksk), None)
    u = i32(rng.integers(-2 ** 31, 2 ** 31, size=nin + 1))
    out = O.keyswitch(keys, u)
    assert not out[:n].any()
    rounded = ((u[:nin].astype(np.int64) % 2 ** 32 + (1 << 15)) >> 16 << 16) % 2 ** 32
    assert out[n] == i32(int(u[nin]) - int((rounded * s).sum()))


# ------------------------------------------------------- reference fixtures


def test_keygen_replays_reference_rng():
    g = gold("ref_p128_seed0_keys.npz")
    K = O.keygen(O.p128(), 0)
    assert O.fnv1a(K.lwe_key) == int(g["lwe_key_hash"])
    assert O.fnv1a(K.tlwe_key) == int(g["tlwe_key_hash"])
    assert O.fnv1a(K.ksk) == int(g["ksk_hash"])
    assert O.fnv1a(np.ascontiguousarray(K.bk[:, :, 0, :])) == int(g["bk_mask_hash"])
    # the reference builds b with its truncating FFT product: |difference| <= 1 unit
    d = K.bk[0, 0, 1, :].astype(np.int64) - g["bk_b_row0"]
    assert np.abs(d).max() <= 1
    assert int(g["bk_b_diff_max"]) <= 1


def test_keyswitch_matches_reference():
    g = gold("ref_p128_seed0_keyswitch.npz")
    K = O.keygen(O.p128(), 0)
    for x, y in zip(g["inputs"], g["outputs"]):
        assert np.array_equal(O.keyswitch(K, x), y)


@pytest.mark.parametrize("name,take", [("nand1000", 16), ("rand6_256", 16), ("mux64", 6)])
def test_gates_match_reference(name, take):
    """Encryptions replay bit-exactly; decrypted outputs equal the reference's and
    the truth table; output phases lie within 2^-5 of the reference's (their
    difference is noise: a one-unit rounding difference in the transform
    re-randomises the bootstrapping noise, whose standard deviation here is
    about 2^-8.5)."""
    g = gold(f"ref_p128_seed0_{name}.npz")
    P = O.p128()
    K = O.keygen(P, 0)
    ops, bits = g["ops"].astype(np.int32), g["bits"].astype(np.int32)
    nin = 3 if name.startswith("mux") else 2
    cts = np.stack([np.stack([O.boots_encrypt(K, bits[gi][j]) for j in range(nin)]) for gi in range(len(ops))])
    assert O.fnv1a(np.ascontiguousarray(cts)) == int(g["inputs_hash"])
    sel = np.arange(take)
    out = O.gates(K, ops[sel], cts[sel, 0], cts[sel, 1], cts[sel, 2] if nin == 3 else None, threads=8)
    ph = O.phases(out, K.lwe_key)
    ref = g["phase"][sel]
    for gi, o in zip(sel, ops[sel]):
        exp = O.TRUTH[O.GATES[o]](*bits[gi])
        assert int(ph[gi] > 0) == exp == int(ref[gi] > 0)
    diff = (ph.astype(np.int64) - ref.astype(np.int64) + 2 ** 31) % 2 ** 32 - 2 ** 31
    assert np.abs(diff).max() < 2 ** 27


def test_reference_phases_all_decrypt():
    # every reference output in the fixtures decrypts to its truth-table bit
    for name in ("nand1000", "rand6_256", "mux64"):
        g = gold(f"ref_p128_seed0_{name}.npz")
        for o, b, ph in zip(g["ops"], g["bits"], g["phase"]):
            assert int(ph > 0) == O.TRUTH[O.GATES[o]](*[int(x) for x in b])


# ------------------------------------------------------ transform emulation


@pytest.fixture(scope="module")
def fft_emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "fft_emu")
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_emu.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, src], check=True)
    return exe


@pytest.mark.parametrize("N,bits", [(1024, 7), (2048, 10), (-2048, 10), (-1024, 7), (1, 7), (1, 10), (2, 10)])
def test_wavefront_transform_emulation(fft_emu, N, bits):
    """The GPU transform's per-lane code, run lane by lane on the CPU, gives the
    exact negacyclic product after rounding (digit x torus polynomials). N = 1 is
    the N = 1024 transform with fused butterflies (FftCT), N = 2 the N = 2048
    pair-of-wavefronts transform with fused butterflies (FftCTPair)."""
    out = subprocess.run([fft_emu, str(N), "6", "7", str(bits)], capture_output=True, text=True, check=True).stdout
    f = out.split()
    assert f[0] == "maxerr" and float(f[1]) < 0.1
    assert int(f[3]) == 0
    assert float(f[5]) < 1e-4


@pytest.fixture(scope="module")
def brq_emu(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "brq_emu")
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "brq_emu.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, src], check=True)
    return exe


@pytest.mark.parametrize("seed,abar", [(1, 0), (2, 1), (3, 511), (4, 777), (5, 2047), (6, 2048), (7, 3000),
                                       (8, 4095)])
def test_two_wave_kernel_row_emulation(brq_emu, seed, abar):
    """One key row of the two-wavefront N = 2048 blind rotation, its per-lane
    rotation (one mask and a sign bit-field per value), decomposition, pair
    transform, multiply-accumulate, handover of the other component's half,
    inverse halves and accumulator update, emulated lane by lane on the CPU:
    equal to the exact external product mod 2^32 for rotations at and across
    the wrap points (0, N - 1, N, 2N - 1)."""
    out = subprocess.run([brq_emu, str(seed), str(abar)], capture_output=True, text=True).stdout.split()
    assert out[:4] == ["abar", str(abar), "mismatches", "0"], out


@pytest.mark.parametrize("N,kpl,bits,worst,exact", [
    (1, 6, 7, 0, True), (1, 6, 7, 1, True),          # P128, fused transform: typical and sign-aligned worst case
    (1024, 6, 7, 1, True), (-1024, 6, 7, 1, True),   # P128, radix-8 and half transforms, worst case
    (-2048, 4, 10, 0, True),                         # P_big, pair transform: typical
    (-2048, 4, 10, 1, False),                        # P_big worst case: |sum| = 2^53, past FP64 exactness
    (2, 4, 10, 0, True), (2, 4, 10, 1, False),       # P_big, pair transform with fused butterflies
])
def test_transform_exactness_bound(fft_emu, N, kpl, bits, worst, exact):
    """Where the FP64 product stops being exact. The kernels sum kpl digit x key
    products in the transform domain before one inverse; the emulation does the
    same lane by lane. With uniform digits and keys (worst = 0) outputs stay
    within 0.07 of an integer; with every digit and key value at its extreme and
    signs aligned on one output coefficient (worst = 1) P128 reaches 2^49.6 and
    is still exact (within 0.13), while P_big reaches 2^53 and is not. The key
    is uniformly random and secret, so that alignment has probability about
    2^-N."""
    out = subprocess.run([fft_emu, "mac", str(N), str(kpl), "2", "3", str(bits), str(worst)], capture_output=True,
                         text=True, check=True).stdout.split()
    err, miss, mag = float(out[1]), int(out[3]), float(out[5])
    if exact:
        assert miss == 0 and err < 0.2, out
    else:
        assert miss > 0 and mag > 52.9, out
