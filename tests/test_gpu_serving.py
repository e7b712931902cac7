"""Serving loop on the GPU (tfhe_amd_server_*): continuous batching of circuit
requests from several clients with their own keys. A request's outputs are
bit-identical alone, in a mix of concurrent requests, through the netlist
scheduler and with several clients packed into one launch; every output
decrypts to the plaintext evaluation; admission keeps the wire arena within its
budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def keys(gpu):
    T = gpu
    P = T.default_params(110)
    sks = [T.SecretKey(P, seed=100 + i) for i in range(3)]
    yield T, P, sks
    for sk in sks:
        sk.close()


def _request(sk, rng, depth, width):
    from tfhe_amd.circuits import random_layered_circuit
    C, last = random_layered_circuit(rng, depth, width)
    bits = rng.integers(0, 2, size=width).astype(np.int32)
    return C, bits, sk.encrypt(bits), last


def _server(T, P, sks, **kw):
    from tfhe_amd.circuits import Server
    S = Server(**kw)
    cids = [S.add_client(sk.cloud, P.n) for sk in sks]
    return S, cids


def test_same_alone_and_in_the_mix(keys):
    T, P, sks = keys
    rng = np.random.default_rng(7)
    C, bits, cts, last = _request(sks[0], rng, 9, 37)
    others = []
    for i in range(24):
        c = i % 3
        others.append((c,) + _request(sks[c], rng, int(rng.integers(1, 12)), int(rng.integers(1, 200))))
    S, cids = _server(T, P, sks, n_streams=4, depth=2, max_batch=4096, wire_budget_bytes=256 << 20)
    try:
        rid = S.submit(cids[0], C, np.arange(len(bits)), cts, last)
        alone, _ = S.wait(rid, 60000)
        ids, tid = [], None
        for j, (c, Co, b, ct, lw) in enumerate(others):
            if j == 12:
                tid = S.submit(cids[0], C, np.arange(len(bits)), cts, last)
            ids.append((c, Co, b, lw, S.submit(cids[c], Co, np.arange(len(b)), ct, lw)))
        mixed, _ = S.wait(tid, 120000)
        results = [(c, Co, b, lw, S.wait(r, 120000)[0]) for c, Co, b, lw, r in ids]
        st = S.stats()
    finally:
        S.close()
    assert np.array_equal(alone, mixed)
    assert np.array_equal(sks[0].decrypt(alone), C.evaluate(bits)[last])
    for c, Co, b, lw, out in results:
        assert np.array_equal(sks[c].decrypt(out), Co.evaluate(b)[lw])
    assert st["completed"] == 26 and st["failed"] == 0
    # levels of concurrent requests of one client were merged into shared launches
    levels = 2 * 9 + sum(len(Co.op) // len(b) for _, Co, b, _lw, _o in results)
    assert st["launches"] < levels
    assert st["gates"] == 2 * len(C.op) + sum(len(Co.op) for _, Co, *_ in results)

    # the same request through the netlist scheduler (one level per launch)
    from tfhe_amd.circuits import Netlist
    pool = np.zeros((C.n_wires, P.n + 1), np.int32)
    pool[:len(bits)] = cts
    d = T.DeviceArray.from_numpy(pool)
    nl = Netlist(sks[0].cloud, C, 0)
    nl.run(d)
    T.synchronize()
    out = d.numpy()
    d.free()
    nl.close()
    assert np.array_equal(out[last], alone)


def test_packed_launches_bit_identical(keys):
    """With packing on, several clients' levels share launches (each client's
    gates with its own key as a segment of the batched blind-rotation kernel,
    each segment's keyswitch with its own key), and every output is bit-identical
    to the unpacked server's and decrypts to the plaintext evaluation."""
    T, P, sks = keys
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(18):
        c = i % 3
        reqs.append((c,) + _request(sks[c], rng, int(rng.integers(1, 10)), int(rng.integers(1, 300))))
    outs = {}
    for pack in (0, 1):
        S, cids = _server(T, P, sks, n_streams=4, depth=2, max_batch=4096, wire_budget_bytes=256 << 20,
                          pack_min_gates=pack)
        try:
            ids = [S.submit(cids[c], C, np.arange(len(b)), ct, lw) for c, C, b, ct, lw in reqs]
            outs[pack] = [S.wait(r, 120000)[0] for r in ids]
            st = S.stats()
        finally:
            S.close()
        assert st["failed"] == 0 and st["completed"] == len(reqs)
        assert (st["packed_launches"] > 0) if pack else (st["packed_launches"] == 0)
    for (c, C, b, ct, lw), o0, o1 in zip(reqs, outs[0], outs[1]):
        assert np.array_equal(o0, o1)
        assert np.array_equal(sks[c].decrypt(o1), C.evaluate(b)[lw])


def test_packing_more_clients_than_segments(keys):
    """34 clients with their own device-generated keys submit together: packed
    launches fill the 32-entry segment table and the rest of the ready clients
    launch on their own streams or in the next packed launch; every output is
    bit-identical to the unpacked server's and decrypts to the plaintext."""
    T, P, _ = keys
    sks = [T.SecretKey.generate_on_device(P, seed=500 + i) for i in range(34)]
    try:
        rng = np.random.default_rng(12)
        reqs = [(i,) + _request(sks[i], rng, int(rng.integers(2, 6)), int(rng.integers(20, 120))) for i in range(34)]
        outs, packed = {}, {}
        for pack in (0, 1):
            S, cids = _server(T, P, sks, n_streams=4, depth=2, max_batch=4096, wire_budget_bytes=256 << 20,
                              pack_min_gates=pack)
            try:
                ids = [S.submit(cids[c], C, np.arange(len(b)), ct, lw) for c, C, b, ct, lw in reqs]
                outs[pack] = [S.wait(r, 120000)[0] for r in ids]
                st = S.stats()
            finally:
                S.close()
            assert st["failed"] == 0 and st["completed"] == len(reqs)
            packed[pack] = st["packed_launches"]
        assert packed[0] == 0 and packed[1] > 0
        for (c, C, b, ct, lw), o0, o1 in zip(reqs, outs[0], outs[1]):
            assert np.array_equal(o0, o1)
            assert np.array_equal(sks[c].decrypt(o1), C.evaluate(b)[lw])
    finally:
        for sk in sks:
            sk.close()


def test_merging_and_overlap(keys):
    """Many small requests of one client submitted at once run in fewer launches
    than levels submitted, and the largest launch holds several requests."""
    T, P, sks = keys
    rng = np.random.default_rng(8)
    reqs = [_request(sks[1], rng, 4, 16) for _ in range(12)]
    S, cids = _server(T, P, sks[1:2], n_streams=2, depth=2, max_batch=4096, wire_budget_bytes=64 << 20)
    try:
        ids = [S.submit(cids[0], C, np.arange(len(b)), ct, lw) for C, b, ct, lw in reqs]
        outs = [S.wait(r, 60000)[0] for r in ids]
        st = S.stats()
    finally:
        S.close()
    for (C, b, ct, lw), out in zip(reqs, outs):
        assert np.array_equal(sks[1].decrypt(out), C.evaluate(b)[lw])
    assert st["launches"] < 12 * 4
    assert st["max_launch_gates"] > 16


def test_admission_by_wire_budget(keys):
    T, P, sks = keys
    rng = np.random.default_rng(9)
    w_bytes = 4 * (P.n + 1)
    # depth 4, width 64: 320 wires + 64 staged outputs + 1 row-list slot
    per_req = (5 * 64 + 64 + 1) * w_bytes
    budget = max(1 << 20, int(2.5 * per_req))
    reqs = [_request(sks[2], rng, 4, 64) for _ in range(8)]
    S, cids = _server(T, P, sks[2:3], n_streams=2, depth=2, max_batch=1024, wire_budget_bytes=budget)
    try:
        ids = [S.submit(cids[0], C, np.arange(len(b)), ct, lw) for C, b, ct, lw in reqs]
        outs = [S.wait(r, 60000)[0] for r in ids]
        st = S.stats()
        big = _request(sks[2], rng, 30, 300)[0]
        with pytest.raises(RuntimeError, match="wire budget"):
            S.submit(cids[0], big, np.arange(300), np.zeros((300, P.n + 1), np.int32), [big.n_wires - 1])
    finally:
        S.close()
    for (C, b, ct, lw), out in zip(reqs, outs):
        assert np.array_equal(sks[2].decrypt(out), C.evaluate(b)[lw])
    assert st["completed"] == 8
    assert per_req <= st["wire_bytes_peak"] <= st["wire_bytes_budget"] <= budget


def test_submit_rejects_bad_requests(keys):
    T, P, sks = keys
    from tfhe_amd.circuits import Circuit
    S, cids = _server(T, P, sks[:1], n_streams=1, depth=1, max_batch=64, wire_budget_bytes=8 << 20)
    try:
        C = Circuit(2)
        C.n_wires = 4  # wire 2 is neither an input nor driven
        C.op, C.in0, C.in1, C.in2, C.out = [0], [0], [2], [0], [3]
        with pytest.raises(RuntimeError, match="neither an input nor driven"):
            S.submit(cids[0], C, [0, 1], np.zeros((2, P.n + 1), np.int32), [3])
        C2 = Circuit(2)
        C2.gate(0, 0, 1)
        with pytest.raises(RuntimeError, match="driven by a gate"):
            S.submit(cids[0], C2, [0, 2], np.zeros((2, P.n + 1), np.int32), [2])
        with pytest.raises(RuntimeError, match="unknown client"):
            S.submit(7, C2, [0, 1], np.zeros((2, P.n + 1), np.int32), [2])
    finally:
        S.close()
