"""Edge cases of the batched gate path on the GPU: empty batches through every
entry point, a batch of MUX gates only (the job count at its maximum of two
bootstraps per gate) in every N = 1024 kernel, a ragged batch whose last
workgroup is partly idle, and one very large launch (131,072 gates reading a
shared pool of 4096 inputs) — each against the plaintext gates, and sampled
gates bit-exact against the oracle."""
import ctypes

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def key(gpu):
    T = gpu
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=21)
    keys = O.Keys(O.p128(), **sk.export())
    sk.upload()
    yield T, P, sk, keys
    sk.close()


def _truth(ops, bits):
    return np.array([O.TRUTH[O.GATES[o]](x, y, z) for o, x, y, z in zip(ops, *bits)])


def test_empty_batches(key):
    T, P, sk, keys = key
    L = T.lib()
    d_pool = T.DeviceArray((4, P.n + 1))
    d_idx = T.DeviceArray(4)
    try:
        assert L.tfhe_amd_gates_device(sk.cloud, 0, d_idx.ptr, d_idx.ptr, d_idx.ptr, d_idx.ptr, d_idx.ptr,
                                       d_pool.ptr, None) == 0
        assert L.tfhe_amd_keyswitch_device(sk.cloud, 0, d_pool.ptr, d_pool.ptr, None) == 0
        assert L.bootsGate_batch(0, None, None, None, None, None, sk.cloud) == 0
        T.synchronize()
    finally:
        d_pool.free()
        d_idx.free()
    # a negative count is an error, not a launch
    assert L.tfhe_amd_gates_device(sk.cloud, -1, None, None, None, None, None, None, None) != 0


@pytest.mark.parametrize("mode", [1, 2, 3, 5, 7])
def test_mux_only_batch_every_kernel(key, mode):
    T, P, sk, keys = key
    rng = np.random.default_rng(70 + mode)
    count = 515  # 1030 bootstraps: not a multiple of any kernel's bootstraps per workgroup
    ops = np.full(count, 10, np.int32)
    bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    T.set_launch_config(mode)
    try:
        got = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    assert np.array_equal(sk.decrypt(got), _truth(ops, bits))
    sel = [0, 257, 514]
    assert np.array_equal(got[sel], O.gates(keys, ops[sel], a[sel], b[sel], c[sel], threads=8))


def test_ragged_batch_in_the_batched_kernel(key):
    """4097 gates in the batched one-bootstrap-per-wavefront kernel: one gate past
    a full 4096-gate launch, with MUX gates appended as extra jobs and the idle
    tail of the grid leaving at once."""
    T, P, sk, keys = key
    rng = np.random.default_rng(80)
    count = 4097
    ops = rng.integers(0, 11, size=count).astype(np.int32)
    ops[-1] = 10
    bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    T.set_launch_config(2)
    try:
        got = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    assert np.array_equal(sk.decrypt(got), _truth(ops, bits))
    sel = [0, 2048, 4095, 4096]
    assert np.array_equal(got[sel], O.gates(keys, ops[sel], a[sel], b[sel], c[sel], threads=8))


@pytest.mark.parametrize("mode,count,kind", [(2, 8300, "mixed"), (2, 4133, "mixed"), (2, 2048, "mux"),
                                             (2, 1280, "two_input"), (2, 1536, "two_input"),
                                             (2, 1792, "two_input"), (2, 3072, "two_input"),
                                             (2, 1100, "two_input"), (2, 1901, "two_input"), (2, 1300, "mixed"),
                                             (3, 1024, "two_input"), (3, 512, "mux"), (3, 1100, "mixed"),
                                             (3, 1000, "two_input"), (3, 900, "mixed")])
def test_whole_round_workgroups_match_one_bootstrap_workgroups(key, mode, count, kind):
    """Mode 2 runs whole device rounds on workgroups of 5 to 8 bootstraps (the batch's
    bootstraps per CU when that is 5 to 8, else 8) with a barrier per key row, mode 3 (1024 jobs) on workgroups of four two-wavefront bootstraps, and the
    rest on one-bootstrap workgroups; mode 9 runs every job on the batched kernel's
    one-bootstrap workgroups. Mode 2 on 256 CUs: 8300 gates with MUX gates (about 9050
    jobs: three rounds of eight split off), 4133 (about 4500 jobs: too few whole rounds
    to split), 2048 MUX gates (4096 jobs, two whole rounds of eight, second bootstraps
    included), 1280 / 1536 / 1792 two-input gates (one round of five / six / seven) and
    3072 (one and a half rounds of eight: all on one-bootstrap workgroups), and partial
    single rounds that still give every CU w - 1 or more: 1100 two-input gates (five per
    workgroup), 1901 (eight, the last workgroup with five) and 1300 gates with MUX gates
    (six); mode 3: 1024 two-input gates and 512 MUX gates (one whole
    round each), 1100 gates with MUX gates (not whole rounds), 1000 two-input gates
    (one partial round of four per workgroup) and 900 gates with MUX gates (about 980
    jobs: the same). Bit-identical, and
    gates from every part equal the oracle."""
    T, P, sk, keys = key
    rng = np.random.default_rng(81 + count)
    if kind == "mux":
        ops = np.full(count, 10, np.int32)
    else:
        ops = rng.integers(0, 11 if kind == "mixed" else 10, size=count).astype(np.int32)
    sel = sorted({0, count // 2, count - 1} | ({2047, 6143, 6144} if count == 8300 else set()))
    if kind == "mixed":
        ops[sel] = 10
    bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    got = {}
    try:
        for m in (mode, 9):
            T.set_launch_config(m)
            assert T.kernel_for(sk.cloud, count) == m
            got[m] = T.run_gates(sk.cloud, ops, a, b, c)
    finally:
        T.set_launch_config(0)
    assert np.array_equal(got[mode], got[9])
    assert np.array_equal(sk.decrypt(got[mode]), _truth(ops, bits))
    assert np.array_equal(got[mode][sel], O.gates(keys, ops[sel], a[sel], b[sel], c[sel], threads=8))


def test_very_large_launch(key):
    """131,072 gates in one launch over a pool of 4096 inputs (so every index and
    workspace size is well past 2^16 gates)."""
    T, P, sk, keys = key
    rng = np.random.default_rng(90)
    n_in, count = 4096, 131072
    w = P.n + 1
    in_bits = rng.integers(0, 2, size=n_in).astype(np.int32)
    cts = sk.encrypt(in_bits)
    ops = rng.integers(0, 11, size=count).astype(np.int32)
    i0, i1, i2 = (rng.integers(0, n_in, size=count).astype(np.int32) for _ in range(3))
    out = (n_in + np.arange(count)).astype(np.int32)
    pool = np.zeros((n_in + count, w), np.int32)
    pool[:n_in] = cts
    arrs = [T.DeviceArray.from_numpy(x) for x in (pool, ops, i0, i1, i2, out)]
    d_pool, d_ops, d_i0, d_i1, d_i2, d_o = arrs
    try:
        T.gates_device(sk.cloud, d_ops, d_i0, d_i1, d_i2, d_o, d_pool)
        T.synchronize()
        res = d_pool.numpy()[n_in:]
    finally:
        for d in arrs:
            d.free()
    exp = _truth(ops, (in_bits[i0], in_bits[i1], in_bits[i2]))
    assert np.array_equal(sk.decrypt(res), exp)
    sel = np.array([0, 65535, 65536, 131071])
    assert np.array_equal(res[sel], O.gates(keys, ops[sel], cts[i0[sel]], cts[i1[sel]], cts[i2[sel]], threads=8))


def test_host_batch_large_and_repeated_outputs(key):
    """bootsGate_batch on 1100 host LweSample gates (the staging copies split across
    host threads) equals the device path; with one output named twice the later
    gate's result stays, as in a serial loop."""
    T, P, sk, keys = key
    L = T.lib()
    rng = np.random.default_rng(95)
    count = 1100
    ops = rng.integers(0, 11, size=count).astype(np.int32)
    bits = rng.integers(0, 2, size=(3, count)).astype(np.int32)
    a, b, c = (sk.encrypt(x) for x in bits)
    ref = T.run_gates(sk.cloud, ops, a, b, c)
    H = T.HostCiphertexts(P, 4 * count)
    try:
        H.set(np.concatenate([a, b, c, np.zeros_like(a)]))
        arr = lambda idx: (T.LSP * count)(*[H.item(i) for i in idx])
        i0, i1, i2 = arr(range(count)), arr(range(count, 2 * count)), arr(range(2 * count, 3 * count))
        outs = list(range(3 * count, 4 * count))
        rc = L.bootsGate_batch(count, ops.ctypes.data_as(T.i32p), arr(outs), i0, i1, i2, sk.cloud)
        assert rc == 0, L.tfhe_amd_last_error()
        assert np.array_equal(H.get()[3 * count:], ref)
        outs[1000] = outs[5]  # gates 5 and 1000 write the same sample
        H.set(np.concatenate([a, b, c, np.zeros_like(a)]))
        rc = L.bootsGate_batch(count, ops.ctypes.data_as(T.i32p), arr(outs), i0, i1, i2, sk.cloud)
        assert rc == 0, L.tfhe_amd_last_error()
        got = H.get()[3 * count:]
        assert np.array_equal(got[5], ref[1000])
        keep = np.ones(count, bool)
        keep[[5, 1000]] = False
        assert np.array_equal(got[keep], ref[keep])
    finally:
        H.close()


def test_device_api_on_two_streams(key):
    """Two tfhe_amd_gates_device calls queued back to back on two different HIP
    streams, with no host synchronisation between them, each in the stream's own
    workspace (so they overlap on the device): both results equal runs on the
    library's own stream."""
    T, P, sk, keys = key
    hip = ctypes.CDLL("libamdhip64.so")
    s1, s2 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(s1)) == 0 and hip.hipStreamCreate(ctypes.byref(s2)) == 0
    rng = np.random.default_rng(97)
    count, w = 3000, P.n + 1
    ops = [rng.integers(0, 11, size=count).astype(np.int32) for _ in range(2)]
    cts = [sk.encrypt(rng.integers(0, 2, size=3 * count).astype(np.int32)) for _ in range(2)]
    refs = [T.run_gates(sk.cloud, ops[k], cts[k][:count], cts[k][count:2 * count], cts[k][2 * count:])
            for k in range(2)]
    # pool: batch k's inputs at [4k count, 4k count + 3 count), outputs after them
    pool = np.zeros((8 * count, w), np.int32)
    for k in range(2):
        pool[4 * k * count:4 * k * count + 3 * count] = cts[k]
    base = [np.arange(count, dtype=np.int32) + 4 * k * count for k in range(2)]
    arrs = [T.DeviceArray.from_numpy(pool)]
    try:
        d_pool = arrs[0]
        calls = []
        for k in range(2):
            d = [T.DeviceArray.from_numpy(x) for x in (ops[k], base[k], base[k] + count, base[k] + 2 * count,
                                                       base[k] + 3 * count)]
            arrs.extend(d)
            calls.append(d)
        T.synchronize()
        for k, s in enumerate((s1, s2)):
            T.gates_device(sk.cloud, *calls[k], d_pool, stream=s.value)
        assert hip.hipStreamSynchronize(s1) == 0 and hip.hipStreamSynchronize(s2) == 0
        got = d_pool.numpy()
    finally:
        for d in arrs:
            d.free()
        hip.hipStreamDestroy(s1)
        hip.hipStreamDestroy(s2)
    for k in range(2):
        assert np.array_equal(got[4 * k * count + 3 * count:4 * (k + 1) * count], refs[k]), k


def test_device_api_on_ten_streams(key):
    """Ten streams, one small batch each, queued back to back: the first eight
    streams get workspaces of their own, the last two share the library stream's
    (each call waits on the device for the workspace's previous use). Every result
    equals the run on the library's own stream."""
    T, P, sk, keys = key
    hip = ctypes.CDLL("libamdhip64.so")
    ns, count, w = 10, 300, P.n + 1
    hs = [ctypes.c_void_p() for _ in range(ns)]
    for h in hs:
        assert hip.hipStreamCreate(ctypes.byref(h)) == 0
    rng = np.random.default_rng(98)
    ops = rng.integers(0, 11, size=(ns, count)).astype(np.int32)
    cts = sk.encrypt(rng.integers(0, 2, size=3 * count).astype(np.int32))
    refs = [T.run_gates(sk.cloud, ops[k], cts[:count], cts[count:2 * count], cts[2 * count:]) for k in range(ns)]
    pool = np.zeros((3 * count + ns * count, w), np.int32)
    pool[:3 * count] = cts
    idx = np.arange(count, dtype=np.int32)
    arrs = [T.DeviceArray.from_numpy(pool), T.DeviceArray.from_numpy(idx), T.DeviceArray.from_numpy(idx + count),
            T.DeviceArray.from_numpy(idx + 2 * count)]
    try:
        d_pool, d_i0, d_i1, d_i2 = arrs
        d_ops = [T.DeviceArray.from_numpy(ops[k]) for k in range(ns)]
        d_out = [T.DeviceArray.from_numpy(idx + (3 + k) * count) for k in range(ns)]
        arrs += d_ops + d_out
        T.synchronize()
        for k in range(ns):
            T.gates_device(sk.cloud, d_ops[k], d_i0, d_i1, d_i2, d_out[k], d_pool, stream=hs[k].value)
        for h in hs:
            assert hip.hipStreamSynchronize(h) == 0
        got = d_pool.numpy()
    finally:
        for d in arrs:
            d.free()
        for h in hs:
            hip.hipStreamDestroy(h)
    for k in range(ns):
        assert np.array_equal(got[(3 + k) * count:(4 + k) * count], refs[k]), k
