"""The netlist and serving workloads at their full sizes on the GPU.

- 512 independent 32-bit ripple-carry adders (81,920 bootstraps; the reference's
  adder circuit, test/test-addition-boot.cpp:30-111), levelled as soon as
  possible, as late as possible and balanced, run from a captured hipGraph: the level
  widths equal the ones derived from the circuit by hand, every sum equals the
  plaintext addition, and gates of the first and of the deepest level are
  bit-exact against the CPU oracle on the GPU's own input ciphertexts.
- The same netlist after its cloud key was released: the next run uploads the
  key again and re-captures the graph (the old graph's kernels point at freed
  memory); the outputs are bit-identical to the first run.
- Serving: 32 clients, each with its own key resident in HBM, Poisson arrivals of
  random layered circuits of depth U[1,64] x width U[1,512] (the shape of
  test/test-long-run.cpp:61-185's random circuits, in levels) for a few
  seconds; every output decrypts to the plaintext evaluation, requests re-run
  alone are bit-identical, and level-1 gates are bit-exact against the oracle.
"""
import threading
import time

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

N_ADDERS, BITS = 512, 32


def _expected_widths(schedule):
    """Level widths of N_ADDERS independent BITS-bit ripple-carry adders (per bit
    t1 = a^b, s = t1^c, t2 = a&b, t3 = c&t1, c' = t2|t3, carry-in an input).
    As soon as possible: t1, t2 of every bit at level 1; s_i, t3_i at 2i+2;
    c_{i+1} at 2i+3. As late as possible (circuit outputs s_i and c_BITS keep
    their earliest level): t1_i at 2i+1, s_i, t3_i and t2_i at 2i+2, c_{i+1} at
    2i+3."""
    depth = 2 * BITS + 1
    if schedule == 1024:
        # balanced with base 1024: the carry chains and the gates they need fill every
        # level but the last to 1024; the sum bits, which nothing reads, fill the last
        return np.array([1024] * (depth - 1) + [N_ADDERS * BITS], np.int32)
    w = np.zeros(depth + 1, np.int64)
    for i in range(BITS):
        if schedule == 0:
            w[1] += 2
        else:
            w[2 * i + 1] += 1
            w[2 * i + 2] += 1
        w[2 * i + 2] += 2
        w[2 * i + 3] += 1
    return (w[1:] * N_ADDERS).astype(np.int32)


@pytest.fixture(scope="module")
def adders(gpu):
    T = gpu
    from tfhe_amd.circuits import ripple_carry_adders
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=31)
    sk.upload()
    C, sums, zero = ripple_carry_adders(N_ADDERS, BITS)
    rng = np.random.default_rng(31)
    x = rng.integers(0, 2 ** BITS, size=N_ADDERS, dtype=np.int64)
    y = rng.integers(0, 2 ** BITS, size=N_ADDERS, dtype=np.int64)
    nb = N_ADDERS * BITS
    sh = np.arange(BITS)
    bits_in = np.zeros(2 * nb + 1, np.int32)
    bits_in[:nb] = ((x[:, None] >> sh) & 1).reshape(-1)
    bits_in[nb:2 * nb] = ((y[:, None] >> sh) & 1).reshape(-1)
    pool = np.zeros((C.n_wires, P.n + 1), np.int32)
    pool[:2 * nb] = sk.encrypt(bits_in[:2 * nb])
    pool[zero, P.n] = -(1 << 29)  # trivial encryption of 0 (bootsCONSTANT) as the carry-in
    yield T, P, sk, C, sums, x, y, pool
    sk.close()


def _check_sums(sk, out, sums, x, y):
    dec = sk.decrypt(out[sums.reshape(-1)]).reshape(sums.shape).astype(np.int64)
    got = (dec << np.arange(BITS + 1)).sum(axis=1)
    return int((got != x + y).sum())


@pytest.mark.parametrize("schedule", [0, 1, 1024], ids=["asap", "alap", "balanced"])
def test_512_adders_32bit_full_size(adders, schedule):
    T, P, sk, C, sums, x, y, pool = adders
    from tfhe_amd.circuits import Netlist
    print(f"[test] 512 x 32-bit adders, schedule {schedule}", flush=True)
    nl = Netlist(sk.cloud, C, schedule)
    widths = nl.levels()
    assert np.array_equal(widths, _expected_widths(schedule))
    assert widths.sum() == 5 * N_ADDERS * BITS == 81920
    d = T.DeviceArray.from_numpy(pool)
    try:
        nl.run(d, use_graph=True)  # capture + replay
        T.synchronize()
        out = d.numpy()
        nl.run(d, use_graph=True)  # second replay of the same graph: same outputs
        T.synchronize()
        assert np.array_equal(d.numpy(), out)
    finally:
        d.free()
        nl.close()
    assert _check_sums(sk, out, sums, x, y) == 0
    # bit-exact against the oracle, on the GPU's own input ciphertexts: level-1
    # gates (t1 and t2 of bit 0, fresh inputs) and gates of the deepest level
    # (c_32 = t2_31 | t3_31, inputs produced 64 levels deep)
    keys = O.Keys(O.p128(), **sk.export())
    op, in0, in1, in2, outw = C.arrays()
    g_first = np.array([0, 2, 5 * 7 + 0, 5 * 200 + 2])  # gate indices of bit 0: t1, t2 of adders 0, 7, 200
    g_last = (BITS - 1) * 5 * N_ADDERS + 5 * np.array([0, 99, 511]) + 4  # c_32 of adders 0, 99, 511
    sel = np.concatenate([g_first, g_last])
    ref = O.gates(keys, op[sel], out[in0[sel]], out[in1[sel]], out[in2[sel]], threads=8)
    assert np.array_equal(out[outw[sel]], ref)


def test_netlist_graph_after_key_release(adders):
    """Releasing the cloud key frees its device copy; the netlist's next graph run
    uploads the key again and re-captures (the first capture's kernels read the
    freed copy), with outputs bit-identical to a run before the release."""
    T, P, sk, C, sums, x, y, pool = adders
    from tfhe_amd.circuits import Netlist
    nl = Netlist(sk.cloud, C, 0)
    d = T.DeviceArray.from_numpy(pool)
    try:
        nl.run(d, use_graph=True)
        T.synchronize()
        before = d.numpy()
        T.lib().tfhe_amd_cloud_key_release(sk.cloud)
        assert T.lib().tfhe_amd_cloud_key_device_bytes(sk.cloud) == 0
        d.free()
        d = T.DeviceArray.from_numpy(pool)
        nl.run(d, use_graph=True)
        T.synchronize()
        after = d.numpy()
        assert T.lib().tfhe_amd_cloud_key_device_bytes(sk.cloud) > 0
    finally:
        d.free()
        nl.close()
    assert np.array_equal(before, after)
    assert _check_sums(sk, after, sums, x, y) == 0


def test_server_keeps_key_after_release_and_delete(gpu):
    """A server holds its clients' device keys: releasing a client's cloud key, or
    deleting the key set, while the server runs leaves its requests correct."""
    T = gpu
    from tfhe_amd.circuits import Server, random_layered_circuit
    P = T.default_params(110)
    sks = [T.SecretKey.generate_on_device(P, seed=700 + i) for i in range(2)]
    rng = np.random.default_rng(70)
    reqs = []
    for c in (0, 1, 0, 1):
        C, last = random_layered_circuit(rng, 3, 40)
        bits = rng.integers(0, 2, size=40).astype(np.int32)
        reqs.append((c, C, bits, sks[c].encrypt(bits), last))
    S = Server(n_streams=2, depth=2, max_batch=1024, wire_budget_bytes=64 << 20)
    try:
        cids = [S.add_client(sk.cloud, P.n) for sk in sks]
        outs = [S.wait(S.submit(cids[c], C, np.arange(40), ct, lw), 60000)[0] for c, C, b, ct, lw in reqs[:2]]
        T.lib().tfhe_amd_cloud_key_release(sks[0].cloud)
        outs.append(S.wait(S.submit(cids[0], reqs[2][1], np.arange(40), reqs[2][3], reqs[2][4]), 60000)[0])
        # the key set of client 1 is deleted; its lwe key is kept for the check
        lwe1 = sks[1].export()["lwe_key"]
        sks[1].close()
        outs.append(S.wait(S.submit(cids[1], reqs[3][1], np.arange(40), reqs[3][3], reqs[3][4]), 60000)[0])
    finally:
        S.close()
    for (c, C, bits, ct, lw), o in zip(reqs, outs):
        exp = C.evaluate(bits)[lw]
        if c == 0:
            assert np.array_equal(sks[0].decrypt(o), exp)
        else:
            ph = (o[:, P.n].astype(np.int64) - (o[:, :P.n].astype(np.int64) * lwe1).sum(axis=1)).astype(np.int32)
            assert np.array_equal((ph > 0).astype(np.int32), exp)
    sks[0].close()


def test_serving_32_clients_full_request_shape(gpu):
    """32 clients with their own resident keys; Poisson arrivals (4e4 gates/s
    offered, 3 s) of random layered circuits with depth U[1,64] and width
    U[1,512]. Each request also returns four of its level-1 wires for the oracle
    check."""
    T = gpu
    from tfhe_amd.circuits import Server, random_layered_circuit
    P = T.default_params(110)
    n_cl = 32
    print("[test] serving: generating 32 client keys on the device", flush=True)
    keys = [T.SecretKey.generate_on_device(P, seed=2000 + i) for i in range(n_cl)]
    rng = np.random.default_rng(2024)
    rate, window = 4e4, 3.0
    mean_gates = 32.5 * 256.5
    arr, t = [], 0.0
    while t < window:
        t += rng.exponential(mean_gates / rate)
        arr.append(t)
    reqs = []
    for a in arr:
        c = int(rng.integers(0, n_cl))
        depth, width = int(rng.integers(1, 65)), int(rng.integers(1, 513))
        C, last = random_layered_circuit(rng, depth, width)
        bits = rng.integers(0, 2, size=width).astype(np.int32)
        # outputs: the last layer, then up to four wires of the first layer
        first = np.arange(width, width + min(4, width))
        reqs.append((a, c, C, bits, keys[c].encrypt(bits), np.concatenate([last, first]), len(last)))
    assert len(reqs) >= 8
    S = Server(n_streams=8, depth=2, max_batch=4096, wire_budget_bytes=2 << 30, pack_min_gates=1024)
    try:
        cids = [S.add_client(sk.cloud, P.n) for sk in keys]
        ids = [None] * len(reqs)
        start = time.perf_counter()

        def feed():
            for j, (a, c, C, bits, cts, ow, _) in enumerate(reqs):
                dly = start + a - time.perf_counter()
                if dly > 0:
                    time.sleep(dly)
                ids[j] = S.submit(cids[c], C, np.arange(len(bits)), cts, ow)

        th = threading.Thread(target=feed)
        th.start()
        outs = []
        for j in range(len(reqs)):
            while ids[j] is None:
                time.sleep(0.001)
            outs.append(S.wait(ids[j], 120000)[0])
        th.join()
        st = S.stats()
        print(f"[test] serving: {len(reqs)} requests, {st['gates']} gates, {st['launches']} launches, "
              f"{st['packed_launches']} packed", flush=True)
        # three requests again, alone on the idle server: bit-identical
        for j in (0, len(reqs) // 2, len(reqs) - 1):
            a, c, C, bits, cts, ow, _ = reqs[j]
            alone, _ = S.wait(S.submit(cids[c], C, np.arange(len(bits)), cts, ow), 120000)
            assert np.array_equal(alone, outs[j]), j
    finally:
        S.close()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    assert st["clients"] == n_cl
    for (a, c, C, bits, cts, ow, nl), o in zip(reqs, outs):
        assert np.array_equal(keys[c].decrypt(o), C.evaluate(bits)[ow])
    # level-1 gates of two requests of different clients, bit-exact vs the oracle
    checked = 0
    for j in (0, len(reqs) - 1):
        a, c, C, bits, cts, ow, nl = reqs[j]
        width = len(bits)
        k = min(4, width)
        op, in0, in1, in2, outw = C.arrays()
        kk = O.Keys(O.p128(), **keys[c].export())
        ref = O.gates(kk, op[:k], cts[in0[:k]], cts[in1[:k]], cts[in2[:k]], threads=8)
        assert np.array_equal(outs[j][nl:nl + k], ref)
        checked += k
    assert checked >= 2
    for sk in keys:
        sk.close()
