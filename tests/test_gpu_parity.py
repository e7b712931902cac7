"""GPU parity: the HIP path (through the C-ABI in libtfhe_amd.so) against the
CPU oracle (oracle/tfhe_oracle.cpp) and closed forms, on the same seeded keys.

Bars:
  keyswitch                 bit-exact (pure int32)
  one external product      bit-exact vs the exact integer negacyclic product
                            (FP64 transform with round-to-nearest; the
                            transform's rounding error is < 0.04 units, so the
                            rounded result equals the exact one)
  blind rotation / gates    bit-exact ciphertexts vs the oracle on the same keys
                            and inputs; decrypted bits equal the plaintext gate
  noiseless key             closed form X^{sum bara_i s_i - barb} * testvector
"""
import ctypes

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def p128(gpu):
    T = gpu
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=0)
    arrs = sk.export()
    keys = O.Keys(O.p128(), **arrs)
    sk.upload()
    yield T, P, sk, keys
    sk.close()


def rand_i32(rng, shape):
    return rng.integers(-2 ** 31, 2 ** 31, size=shape, dtype=np.int64).astype(np.int32)


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("count", [1, 3, 40, 600])
def test_keyswitch_bitexact(p128, count, path):
    """Both keyswitch kernels (int8 matrix cores; lanes over gates) against the oracle."""
    T, P, sk, keys = p128
    rng = np.random.default_rng(100 + count)
    ext = rand_i32(rng, (count, P.N + 1))
    ext[0, :] = 0  # all-zero digits row: the output must be (0, b) = 0
    T.lib().tfhe_amd_set_keyswitch_path(path)
    try:
        got = T.keyswitch(sk.cloud, P, ext)
    finally:
        T.lib().tfhe_amd_set_keyswitch_path(0)
    idx = list(range(count)) if count <= 40 else list(rng.choice(count, 40, replace=False))
    for g in idx:
        assert np.array_equal(got[g], O.keyswitch(keys, ext[g])), f"gate {g}"


@pytest.mark.parametrize("count", [1, 64, 65, 257, 1500])
def test_sliced_keyswitch_matches_integer_path(p128, count):
    """The matrix-core keyswitch slices its input coefficients over several
    workgroups (integer atomics into rows preset to (0, b)) while the grid is
    small, and runs one slice for large batches: every output row equals the
    integer lanes-over-gates kernel's, across the slicing thresholds."""
    T, P, sk, keys = p128
    rng = np.random.default_rng(300 + count)
    ext = rand_i32(rng, (count, P.N + 1))
    got = T.keyswitch(sk.cloud, P, ext)
    T.lib().tfhe_amd_set_keyswitch_path(1)
    try:
        ref = T.keyswitch(sk.cloud, P, ext)
    finally:
        T.lib().tfhe_amd_set_keyswitch_path(0)
    assert np.array_equal(got, ref)
    assert np.array_equal(got[0], O.keyswitch(keys, ext[0]))


def _row_ptr(cloud, i):
    base = cloud.contents.bkFFT.contents.bkFFT
    addr = ctypes.addressof(base.contents) + i * ctypes.sizeof(base._type_)
    return ctypes.cast(addr, type(base))


def _tlwe_params(P):
    return P.ptr.contents.tgsw_params.contents.tlwe_params


@pytest.mark.parametrize("row,nrows", [(0, 1), (17, 1), (629, 1), (100, 4)])
def test_blind_rotate_exact(p128, row, nrows):
    """tfhe_blindRotate_FFT on a random accumulator vs the exact product."""
    T, P, sk, keys = p128
    L = T.lib()
    rng = np.random.default_rng(row * 7 + nrows)
    acc0 = rand_i32(rng, (2, P.N))
    bara = rng.integers(1, 2 * P.N, size=nrows).astype(np.int32)
    if nrows > 1:
        bara[1] = P.N  # rotation by exactly N (pure negation)
    acc = L.new_TLweSample(_tlwe_params(P))
    try:
        for c in range(2):
            ctypes.memmove(acc.contents.a[c].coefsT, acc0[c].ctypes.data, 4 * P.N)
        L.tfhe_blindRotate_FFT(acc, _row_ptr(sk.cloud, row), bara.ctypes.data_as(T.i32p), nrows,
                               P.ptr.contents.tgsw_params)
        got = np.zeros((2, P.N), np.int32)
        for c in range(2):
            ctypes.memmove(got[c].ctypes.data, acc.contents.a[c].coefsT, 4 * P.N)
    finally:
        L.delete_TLweSample(acc)
    ref = acc0.copy()
    OP = O.p128()
    OP.n = nrows
    bk_rows = np.ascontiguousarray(keys.bk[row:row + nrows])
    O.lib().orc_blind_rotate(ctypes.byref(OP), O.ptr(bk_rows), O.ptr(bara), O.ptr(ref))
    assert np.array_equal(got, ref), f"max diff {np.abs(got.astype(np.int64) - ref).max()}"


def test_noiseless_key_closed_form(p128):
    """With BK_i = s_i * H (no mask, no noise) the rotation is exact:
    extracted sample = (0, +-MU) with sign from X^{sum bara_i s_i - barb}."""
    T, P, sk, keys = p128
    n, N, l = P.n, P.N, P.l
    bk = np.zeros((n, P.kpl, 2, N), np.int32)
    for r in range(P.kpl):
        blk, p = divmod(r, l)
        bk[:, r, blk, 0] = keys.lwe_key * (1 << (32 - (p + 1) * P.Bgbit))
    ck = T.cloud_from_arrays(P, bk, keys.ksk)
    try:
        rng = np.random.default_rng(5)
        count = 64
        a = rand_i32(rng, (count, n + 1))
        b = rand_i32(rng, (count, n + 1))
        ops = rng.integers(0, 10, size=count).astype(np.int32)
        ext = T.run_gates_woks(ck, P, ops, a, b)
        mu = 1 << 29
        for g in range(count):
            lin = np.zeros(n + 1, np.int32)
            O.lib().orc_gate_linear(int(ops[g]), 0, O.ptr(a[g]), O.ptr(b[g]), O.ptr(a[g]), n, O.ptr(lin))
            bara = [O.lib().orc_modswitch_from_torus32(int(v), 2 * N) for v in lin]
            e = (sum(int(x) * int(s) for x, s in zip(bara[:n], keys.lwe_key)) - bara[n]) % (2 * N)
            idx = (-e) % (2 * N)
            expect_b = mu if idx < N else -mu
            assert not ext[g, :N].any(), f"gate {g}: mask not zero"
            assert ext[g, N] == expect_b, f"gate {g}"
    finally:
        T.lib().delete_gate_bootstrapping_cloud_keyset(ck)


def _encrypt_inputs(sk, rng, count):
    bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
    return bits, sk.encrypt(bits[0]), sk.encrypt(bits[1]), sk.encrypt(bits[2])


def test_gates_bitexact_vs_oracle(p128):
    """Every gate type (MUX included): ciphertexts equal the exact oracle's."""
    T, P, sk, keys = p128
    rng = np.random.default_rng(7)
    ops = np.array(list(range(11)) + [10, 0, 3, 5], np.int32)
    bits, a, b, c = _encrypt_inputs(sk, rng, len(ops))
    got = T.run_gates(sk.cloud, ops, a, b, c)
    ref = O.gates(keys, ops, a, b, c, threads=16)
    exact = (got == ref).all(axis=1)
    dph = sk.phases(got).astype(np.int64) - sk.phases(ref).astype(np.int64)
    assert exact.all(), f"bit-exact {exact.mean():.3f}, max phase diff {np.abs(dph).max()} units"
    dec = sk.decrypt(got)
    for g, op in enumerate(ops):
        assert dec[g] == O.TRUTH[O.GATES[op]](bits[0][g], bits[1][g], bits[2][g])


def test_gates_truth_tables_4096(p128):
    T, P, sk, keys = p128
    rng = np.random.default_rng(11)
    count = 4096
    ops = rng.integers(0, 6, size=count).astype(np.int32)
    bits, a, b, c = _encrypt_inputs(sk, rng, count)
    got = T.run_gates(sk.cloud, ops, a, b)
    dec = sk.decrypt(got)
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, 0) for o, x, y in zip(ops, bits[0], bits[1])])
    assert (dec == exp).all(), f"{(dec != exp).sum()} decryption errors"
    # a sample of them bit-exact against the oracle
    sel = rng.choice(count, 8, replace=False)
    ref = O.gates(keys, ops[sel], a[sel], b[sel], threads=16)
    assert np.array_equal(got[sel], ref)


def test_c_abi_single_gates_match_batch(p128):
    """bootsNAND/.../bootsMUX (one gate per call, host LweSample) equal the
    batched device path on the same inputs, including result aliasing an input."""
    T, P, sk, keys = p128
    L = T.lib()
    rng = np.random.default_rng(3)
    ops = np.arange(11, dtype=np.int32)
    bits, a, b, c = _encrypt_inputs(sk, rng, len(ops))
    ref = T.run_gates(sk.cloud, ops, a, b, c)
    H = T.HostCiphertexts(P, 4)
    try:
        for g, op in enumerate(ops):
            H.set(np.stack([a[g], b[g], c[g], np.zeros_like(a[g])]))
            out, x, y, z = H.item(3), H.item(0), H.item(1), H.item(2)
            if op == T.MUX:
                L.bootsMUX(out, x, y, z, sk.cloud)
            else:
                getattr(L, "boots" + T.GATES[op])(out, x, y, sk.cloud)
            assert np.array_equal(H.get()[3], ref[g]), T.GATES[op]
        # aliasing: result is the first input
        H.set(np.stack([a[0], b[0], c[0], np.zeros_like(a[0])]))
        L.bootsNAND(H.item(0), H.item(0), H.item(1), sk.cloud)
        assert np.array_equal(H.get()[0], ref[0])
    finally:
        H.close()


def test_bootsgate_batch_dedup_and_alias(p128):
    """bootsGate_batch with shared inputs and outputs aliasing inputs."""
    T, P, sk, keys = p128
    L = T.lib()
    rng = np.random.default_rng(9)
    bits, a, b, c = _encrypt_inputs(sk, rng, 6)
    H = T.HostCiphertexts(P, 6)
    try:
        H.set(a)
        ops = np.array([0, 3, 10, 1, 2, 5], np.int32)
        # gate g reads items (g, (g+1)%6[, (g+2)%6]) and writes item g: every output aliases an input
        i0 = [H.item(g) for g in range(6)]
        i1 = [H.item((g + 1) % 6) for g in range(6)]
        i2 = [H.item((g + 2) % 6) for g in range(6)]
        arr = lambda xs: (T.LSP * 6)(*xs)
        before = H.get()
        rc = L.bootsGate_batch(6, ops.ctypes.data_as(T.i32p), arr(i0), arr(i0), arr(i1), arr(i2), sk.cloud)
        assert rc == 0, L.tfhe_amd_last_error()
        got = H.get()
        ref = T.run_gates(sk.cloud, ops, before, before[[1, 2, 3, 4, 5, 0]], before[[2, 3, 4, 5, 0, 1]])
        assert np.array_equal(got, ref)
    finally:
        H.close()


def _gold(name):
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name), allow_pickle=False)


def test_keyswitch_matches_reference_vectors(p128):
    """Reference lweKeySwitch outputs (seed-0 keys, which this library's key
    generation reproduces bit-exactly) on the GPU, through the device API and
    through lweKeySwitch on host samples, whose current_variance equals the
    reference's bit for bit (it depends on the input: one addition of the key
    rows' variance per nonzero digit)."""
    T, P, sk, keys = p128
    L = T.lib()
    g = _gold("ref_p128_seed0_keyswitch.npz")
    got = T.keyswitch(sk.cloud, P, g["inputs"])
    assert np.array_equal(got, g["outputs"])
    ks = sk.cloud.contents.bkFFT.contents.ks
    ext_params = ctypes.pointer(P.ptr.contents.tgsw_params.contents.tlwe_params.contents.extracted_lweparams)
    nx = P.k * P.N
    assert len(set(g["variance"].tolist())) > 1
    for r in range(len(g["inputs"])):
        src = L.new_LweSample(ext_params)
        dst = L.new_LweSample(P.ptr.contents.in_out_params)
        try:
            u = np.ascontiguousarray(g["inputs"][r])
            ctypes.memmove(src.contents.a, u.ctypes.data, 4 * nx)
            src.contents.b = int(u[nx])
            L.lweKeySwitch(dst, ks, src)
            out = np.zeros(P.n + 1, np.int32)
            ctypes.memmove(out.ctypes.data, dst.contents.a, 4 * P.n)
            out[P.n] = dst.contents.b
            assert np.array_equal(out, g["outputs"][r])
            assert dst.contents.current_variance == g["variance"][r], r
        finally:
            L.delete_LweSample(src)
            L.delete_LweSample(dst)


@pytest.mark.parametrize("name", ["nand1000", "rand6_256", "mux64"])
def test_gates_match_reference_outputs(name):
    """All reference gate outputs of the fixtures: same decrypted bits (and
    truth table) from a fresh seed-0 key set and the replayed encryptions;
    phases within 2^-5 of the reference's (the difference is bootstrapping
    noise, see test_oracle_cpu.test_gates_match_reference)."""
    import tfhe_amd as T
    g = _gold(f"ref_p128_seed0_{name}.npz")
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=0)
    ops, bits = g["ops"].astype(np.int32), g["bits"].astype(np.int32)
    nin = 3 if name.startswith("mux") else 2
    cts = np.stack([sk.encrypt(bits[gi][:nin]) for gi in range(len(ops))])
    assert O.fnv1a(np.ascontiguousarray(cts)) == int(g["inputs_hash"])
    out = T.run_gates(sk.cloud, ops, cts[:, 0], cts[:, 1], cts[:, 2] if nin == 3 else None)
    ph = sk.phases(out)
    exp = np.array([O.TRUTH[O.GATES[o]](*b) for o, b in zip(ops, bits)])
    assert np.array_equal((ph > 0).astype(int), exp)
    assert np.array_equal((g["phase"] > 0).astype(int), exp)
    diff = (ph.astype(np.int64) - g["phase"].astype(np.int64) + 2 ** 31) % 2 ** 32 - 2 ** 31
    assert np.abs(diff).max() < 2 ** 27
    # the noise distributions agree: output phase spread around +-1/8
    dev_gpu = np.abs(ph.astype(np.int64)) - 2 ** 29
    dev_ref = np.abs(g["phase"].astype(np.int64)) - 2 ** 29
    if len(ops) >= 256:
        ratio = dev_gpu.std() / dev_ref.std()
        assert 0.8 < ratio < 1.25, ratio
    sk.close()


def test_c_program_gpu(tmp_path):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "c_binding")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-o", exe, os.path.join(root, "tests", "c", "c_binding.c"),
                    "-I", os.path.join(root, "include"), "-L", os.path.join(root, "tfhe_amd"), "-ltfhe_amd",
                    "-Wl,-rpath," + os.path.join(root, "tfhe_amd")], check=True)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("mode", [1, 2, 3, 5, 7])
def test_all_blind_rotation_kernels_bitexact(p128, mode):
    """The blind-rotation kernels (1: one wavefront per bootstrap, 2: one
    bootstrap per one-wavefront workgroup on the fused-butterfly transform, 7: the
    same on the radix-8 transform, 3: two wavefronts per bootstrap, 5: 4l
    wavefronts per bootstrap on half transforms) give the oracle's ciphertexts
    bit-exactly, MUX included, for odd batch sizes."""
    T, P, sk, keys = p128
    L = T.lib()
    rng = np.random.default_rng(20 + mode)
    ops = np.array([10, 0, 3, 10, 5, 1, 2], np.int32)  # 7 gates, 9 bootstraps
    bits, a, b, c = _encrypt_inputs(sk, rng, len(ops))
    L.tfhe_amd_set_launch_config(mode, -1, -1)
    try:
        got = T.run_gates(sk.cloud, ops, a, b, c)
        big_ops = rng.integers(0, 11, size=301).astype(np.int32)
        bb, ba, bbb, bc = _encrypt_inputs(sk, rng, len(big_ops))
        big = T.run_gates(sk.cloud, big_ops, ba, bbb, bc)
    finally:
        L.tfhe_amd_set_launch_config(0, -1, -1)
    ref = O.gates(keys, ops, a, b, c, threads=16)
    assert np.array_equal(got, ref)
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, z) for o, x, y, z in zip(big_ops, bb[0], bb[1], bb[2])])
    assert np.array_equal(sk.decrypt(big), exp)
    sel = [0, 150, 300]
    assert np.array_equal(big[sel], O.gates(keys, big_ops[sel], ba[sel], bbb[sel], bc[sel], threads=8))


def test_autotune_crossovers_and_auto_mode(p128):
    """The autotuner returns ordered crossovers within its ladder, and the
    automatic kernel choice then gives the same ciphertexts as a forced kernel
    on both sides of each crossover."""
    T, P, sk, keys = p128
    lat, small, table = T.autotune(sk.cloud)
    try:
        assert 0 <= lat <= small <= max(T.AUTOTUNE_SIZES)
        assert all(v > 0 for t in table.values() for v in t.values())
        assert T.get_launch_config()[1:] == (lat, small)
        # automatic mode takes the fastest measured kernel of the smallest ladder size that
        # holds the batch (modes 5, 3, 2 in the table's order), the batched kernel beyond it
        names = list(T.AUTOTUNE_MODES)
        for k, size in enumerate(T.AUTOTUNE_SIZES):
            t5, t3, t2 = (table[nm][size] for nm in names)
            want = 5 if (t5 <= t3 and t5 <= t2) else (3 if t3 < t2 else 2)
            assert T.kernel_for(sk.cloud, size) == want, size
            if k:
                assert T.kernel_for(sk.cloud, T.AUTOTUNE_SIZES[k - 1] + 1) == want, size
        assert T.kernel_for(sk.cloud, max(T.AUTOTUNE_SIZES) + 1) == 2
        rng = np.random.default_rng(31)
        for count in sorted({max(1, lat), lat + 1, small, small + 1}):
            ops = rng.integers(0, 11, size=count).astype(np.int32)
            bits, a, b, c = _encrypt_inputs(sk, rng, count)
            auto = T.run_gates(sk.cloud, ops, a, b, c)
            T.set_launch_config(2)
            try:
                forced = T.run_gates(sk.cloud, ops, a, b, c)
            finally:
                T.set_launch_config(0)
            assert np.array_equal(auto, forced), count
        # explicit crossovers replace the per-size choice
        T.set_launch_config(0, 100, 900)
        assert [T.kernel_for(sk.cloud, c) for c in (1, 100, 101, 900, 901)] == [5, 5, 3, 3, 2]
    finally:
        T.set_launch_config(0, 256, 1024)


def _ks_variance_expected(ext, P, row_var):
    """current_variance the reference gives a keyswitched sample: one addition of
    the key rows' variance per nonzero digit of the input's mask
    (libtfhe/lwe-keyswitch-functions.cpp:114-129, lwe-functions.cpp:140-145),
    counted here from the keyswitch input rows ext [count][kN+1]."""
    t, bb = P.ptr.contents.ks_t, P.ptr.contents.ks_basebit
    prec = np.uint64(1 << (32 - (1 + bb * t)))
    ab = (ext[:, :-1].astype(np.int64).astype(np.uint64) + prec) & np.uint64(0xFFFFFFFF)
    nz = np.zeros(len(ext), np.int64)
    for j in range(t):
        nz += (((ab >> np.uint64(32 - (j + 1) * bb)) & np.uint64((1 << bb) - 1)) != 0).sum(axis=1)
    out = []
    for c in nz:
        v = 0.0
        for _ in range(int(c)):
            v += row_var
        out.append(v)
    return np.array(out), nz


def test_bootstrapped_current_variance_follows_keyswitch_digits(p128):
    """Outputs of bootsGate_batch, bootsNAND/bootsMUX, tfhe_bootstrap_FFT and
    lweKeySwitch carry the reference's current_variance: zero, plus the key
    rows' variance once per nonzero keyswitch digit, so it depends on the
    input; the expected counts come from the keyswitch inputs the GPU's
    bootstrap without keyswitch produces (bit-exact with the oracle)."""
    T, P, sk, keys = p128
    L = T.lib()
    rng = np.random.default_rng(31)
    ops = np.array(list(range(11)) + [10, 3], np.int32)
    bits, a, b, c = _encrypt_inputs(sk, rng, len(ops))
    rows = ctypes.cast(sk.cloud.contents.bkFFT.contents.ks.contents.ks0_raw, ctypes.POINTER(T.LweSample))
    row_var = rows[1].current_variance  # ks[0][0][1]; every row carries the same variance
    assert row_var > 0
    ext = T.run_gates_woks(sk.cloud, P, ops, a, b, c)
    exp, nz = _ks_variance_expected(ext, P, row_var)
    assert len(set(nz.tolist())) > 1  # the count depends on the input
    n = len(ops)
    H = T.HostCiphertexts(P, 4 * n)
    try:
        H.set(np.concatenate([a, b, c, np.zeros_like(a)]))
        arr = lambda xs: (T.LSP * n)(*xs)
        rc = L.bootsGate_batch(n, ops.ctypes.data_as(T.i32p), arr([H.item(3 * n + g) for g in range(n)]),
                               arr([H.item(g) for g in range(n)]), arr([H.item(n + g) for g in range(n)]),
                               arr([H.item(2 * n + g) for g in range(n)]), sk.cloud)
        assert rc == 0, L.tfhe_amd_last_error()
        got = np.array([H.item(3 * n + g).contents.current_variance for g in range(n)])
        assert np.array_equal(got, exp), (got, exp)
        # one gate per call
        L.bootsNAND(H.item(3 * n), H.item(0), H.item(n), sk.cloud)
        assert H.item(3 * n).contents.current_variance == exp[0]
        L.bootsMUX(H.item(3 * n + 10), H.item(10), H.item(n + 10), H.item(2 * n + 10), sk.cloud)
        assert H.item(3 * n + 10).contents.current_variance == exp[10]
    finally:
        H.close()
    # lweKeySwitch on random inputs of the extracted dimension
    ks = sk.cloud.contents.bkFFT.contents.ks
    ext_params = ctypes.pointer(P.ptr.contents.tgsw_params.contents.tlwe_params.contents.extracted_lweparams)
    nx = P.k * P.N
    u = rng.integers(-2 ** 31, 2 ** 31, size=(2, nx + 1), dtype=np.int64).astype(np.int32)
    u[1, :nx] = 0  # zero mask: every digit of the offset mask is zero, variance 0
    exp_u, _ = _ks_variance_expected(u, P, row_var)
    assert exp_u[1] == 0 and exp_u[0] > 0
    for r in range(2):
        src = L.new_LweSample(ext_params)
        dst = L.new_LweSample(P.ptr.contents.in_out_params)
        try:
            ctypes.memmove(src.contents.a, u[r].ctypes.data, 4 * nx)
            src.contents.b = int(u[r][nx])
            L.lweKeySwitch(dst, ks, src)
            assert dst.contents.current_variance == exp_u[r], r
        finally:
            L.delete_LweSample(src)
            L.delete_LweSample(dst)
    # tfhe_bootstrap_FFT: the keyswitch input is what tfhe_bootstrap_woKS_FFT returns
    mu = 1 << 29
    x = L.new_LweSample(P.ptr.contents.in_out_params)
    wo = L.new_LweSample(ext_params)
    res = L.new_LweSample(P.ptr.contents.in_out_params)
    try:
        ctypes.memmove(x.contents.a, a[0].ctypes.data, 4 * P.n)
        x.contents.b = int(a[0][P.n])
        L.tfhe_bootstrap_woKS_FFT(wo, sk.cloud.contents.bkFFT, mu, x)
        w = np.zeros((1, nx + 1), np.int32)
        ctypes.memmove(w.ctypes.data, wo.contents.a, 4 * nx)
        w[0, nx] = wo.contents.b
        assert wo.contents.current_variance == 0.0
        L.tfhe_bootstrap_FFT(res, sk.cloud.contents.bkFFT, mu, x)
        assert res.contents.current_variance == _ks_variance_expected(w, P, row_var)[0][0]
    finally:
        for smp in (x, wo, res):
            L.delete_LweSample(smp)
