"""GPU parity of the bootstrapping subsystem entry points that the reference's
gate functions are built from (include/tfhe.h:34-45, lwe-functions.h), called
through the C ABI on host samples: tfhe_bootstrap_woKS_FFT,
tfhe_bootstrap_FFT, tfhe_blindRotateAndExtract_FFT and lweKeySwitch, each
bit-exact against the oracle on the same keys; and the 80-bit preset
(n = 500, l = 2, Bgbit = 10), which runs the l = 2 instantiations of every
N = 1024 kernel, bit-exact against the oracle in every kernel mode."""
import ctypes

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


def rand_i32(rng, shape):
    return rng.integers(-2 ** 31, 2 ** 31, size=shape, dtype=np.int64).astype(np.int32)


@pytest.fixture(scope="module", params=["p128", "pbig"])
def keyset(gpu, request):
    T = gpu
    if request.param == "p128":
        P, OP, seed = T.default_params(110), O.p128(), 8
    else:
        P, OP, seed = T.big_params(), O.pbig(), 9
    sk = T.SecretKey(P, seed=seed)
    keys = O.Keys(OP, **sk.export())
    sk.upload()
    yield T, P, OP, sk, keys
    sk.close()


def _sample(T, params_ptr, values):
    s = T.lib().new_LweSample(params_ptr)
    n = len(values) - 1
    ctypes.memmove(s.contents.a, values.ctypes.data, 4 * n)
    s.contents.b = int(values[n])
    return s


def _read(s, n):
    out = np.zeros(n + 1, np.int32)
    ctypes.memmove(out.ctypes.data, s.contents.a, 4 * n)
    out[n] = s.contents.b
    return out


def test_bootstrap_without_and_with_keyswitch(keyset):
    T, P, OP, sk, keys = keyset
    L = T.lib()
    rng = np.random.default_rng(3)
    bkfft = sk.cloud.contents.bkFFT
    in_params = P.ptr.contents.in_out_params
    ext_params = ctypes.pointer(P.ptr.contents.tgsw_params.contents.tlwe_params.contents.extracted_lweparams)
    nx = P.k * P.N
    for trial in range(3):
        x = rand_i32(rng, P.n + 1)
        mu = int(rng.integers(-2 ** 31, 2 ** 31))
        xs = _sample(T, in_params, x)
        wo = L.new_LweSample(ext_params)
        full = L.new_LweSample(in_params)
        try:
            L.tfhe_bootstrap_woKS_FFT(wo, bkfft, mu, xs)
            L.tfhe_bootstrap_FFT(full, bkfft, mu, xs)
            got_wo, got_full = _read(wo, nx), _read(full, P.n)
        finally:
            for s in (xs, wo, full):
                L.delete_LweSample(s)
        ref_wo = np.zeros(nx + 1, np.int32)
        O.lib().orc_bootstrap_woks(ctypes.byref(OP), O.ptr(keys.bk), mu, O.ptr(x), O.ptr(ref_wo))
        assert np.array_equal(got_wo, ref_wo), trial
        assert np.array_equal(got_full, O.keyswitch(keys, ref_wo)), trial


def test_blind_rotate_and_extract(keyset):
    T, P, OP, sk, keys = keyset
    L = T.lib()
    rng = np.random.default_rng(4)
    nrows = 5
    rows = sk.cloud.contents.bkFFT.contents.bkFFT
    ext_params = ctypes.pointer(P.ptr.contents.tgsw_params.contents.tlwe_params.contents.extracted_lweparams)
    nx = P.k * P.N
    for barb in (0, 1, P.N, 2 * P.N - 3):
        v = rand_i32(rng, P.N)
        bara = rng.integers(0, 2 * P.N, size=nrows).astype(np.int32)
        bara[0] = 0  # a skipped row
        tp = L.new_TorusPolynomial(P.N)
        out = L.new_LweSample(ext_params)
        try:
            ctypes.memmove(tp.contents.coefsT, v.ctypes.data, 4 * P.N)
            L.tfhe_blindRotateAndExtract_FFT(out, tp, rows, barb, bara.ctypes.data_as(T.i32p), nrows,
                                             P.ptr.contents.tgsw_params)
            got = _read(out, nx)
        finally:
            L.delete_TorusPolynomial(tp)
            L.delete_LweSample(out)
        OPn = type(OP).from_buffer_copy(OP)
        OPn.n = nrows
        ref = np.zeros(nx + 1, np.int32)
        O.lib().orc_blind_rotate_and_extract(ctypes.byref(OPn), O.ptr(np.ascontiguousarray(keys.bk[:nrows])),
                                            O.ptr(v), barb, O.ptr(bara), O.ptr(ref))
        assert np.array_equal(got, ref), barb


def test_lwe_keyswitch(keyset):
    T, P, OP, sk, keys = keyset
    L = T.lib()
    rng = np.random.default_rng(5)
    ks = sk.cloud.contents.bkFFT.contents.ks
    ext_params = ctypes.pointer(P.ptr.contents.tgsw_params.contents.tlwe_params.contents.extracted_lweparams)
    nx = P.k * P.N
    for trial in range(3):
        u = rand_i32(rng, nx + 1)
        src = _sample(T, ext_params, u)
        dst = L.new_LweSample(P.ptr.contents.in_out_params)
        try:
            L.lweKeySwitch(dst, ks, src)
            got = _read(dst, P.n)
        finally:
            L.delete_LweSample(src)
            L.delete_LweSample(dst)
        assert np.array_equal(got, O.keyswitch(keys, u)), trial


@pytest.fixture(scope="module")
def p80(gpu):
    T = gpu
    P = T.default_params(80)
    sk = T.SecretKey(P, seed=10)
    OP = O.OrcParams(500, 1024, 1, 2, 10, 8, 2, 2.44e-5, 7.18e-9)
    keys = O.Keys(OP, **sk.export())
    sk.upload()
    yield T, P, sk, keys
    sk.close()


@pytest.mark.parametrize("mode", [2, 3, 5, 7])
def test_80bit_preset_every_kernel(p80, mode):
    T, P, sk, keys = p80
    assert (P.n, P.N, P.l, P.Bgbit) == (500, 1024, 2, 10)
    rng = np.random.default_rng(60 + mode)
    ops = np.array(list(range(11)) + [10, 3], np.int32)
    bits = rng.integers(0, 2, size=(3, len(ops))).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    T.set_launch_config(mode)
    try:
        got = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    assert np.array_equal(got, O.gates(keys, ops, a, b, c, threads=16))
    for g, op in enumerate(ops):
        assert sk.decrypt(got[g:g + 1])[0] == O.TRUTH[O.GATES[op]](bits[0][g], bits[1][g], bits[2][g])


@pytest.mark.parametrize("mode,count", [(2, 2048), (2, 1536), (3, 1024)])
def test_80bit_preset_whole_round_workgroups(p80, mode, count):
    """One whole device round of two-input gates under the 80-bit preset (l = 2): mode 2
    runs it on eight- or six-bootstrap workgroups, mode 3 on four-bootstrap ones, mode 9
    on one-bootstrap workgroups; bit-identical, decrypting correctly, samples equal the
    oracle."""
    T, P, sk, keys = p80
    rng = np.random.default_rng(90 + mode)
    ops = rng.integers(0, 10, size=count).astype(np.int32)
    bits = rng.integers(0, 2, size=(2, count)).astype(np.int32)
    a, b = (sk.encrypt(x) for x in bits)
    got = {}
    try:
        for m in (mode, 9):
            T.set_launch_config(m)
            got[m] = T.run_gates(sk.cloud, ops, a, b)
    finally:
        T.set_launch_config(0)
    assert np.array_equal(got[mode], got[9])
    exp = np.array([O.TRUTH[O.GATES[o]](x, y, 0) for o, x, y in zip(ops, bits[0], bits[1])])
    assert np.array_equal(sk.decrypt(got[mode]), exp)
    sel = [0, count // 2, count - 1]
    assert np.array_equal(got[mode][sel], O.gates(keys, ops[sel], a[sel], b[sel], threads=8))


@pytest.mark.parametrize("path", [0, 1])
def test_80bit_preset_keyswitch(p80, path):
    T, P, sk, keys = p80
    rng = np.random.default_rng(70 + path)
    ext = rand_i32(rng, (9, P.N + 1))
    T.lib().tfhe_amd_set_keyswitch_path(path)
    try:
        got = T.keyswitch(sk.cloud, P, ext)
    finally:
        T.lib().tfhe_amd_set_keyswitch_path(0)
    for g in range(len(ext)):
        assert np.array_equal(got[g], O.keyswitch(keys, ext[g])), g
