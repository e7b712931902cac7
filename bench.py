#!/usr/bin/env python3
"""Gate-bootstrapping throughput on one MI355X: batched vs one gate per launch.

Workload (default): one launch of 4096 independent gate bootstraps with the
default 128-bit parameters (n=630, N=1024, k=1, l=3, Bgbit=7, keyswitch t=8,
basebit=2), gate type uniform over NAND/AND/OR/XOR/XNOR/NOR, uniform input
bits, one cloud key from seed 0; keys and ciphertexts resident in HBM when the
timed region starts. A step is one such launch (blind rotation + sample
extraction + keyswitch for all 4096 gates).

Prints one JSON line (see the keys below). `value` is batched gates/s.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_BOOTSTRAP_P128 = 630 * 258048  # n*[(k+1)(l+1)(5 M log2 M + 6 M) + 8 (k+1) kpl M], M = 512
FP64_PEAK_TFLOPS = 78.6
# algorithmic bytes, P128: the transform-domain key (630 rows x 12 polynomials x 512 complex FP64),
# the keyswitch key's non-trivial rows (1024 x 8 x 3 rows of 631 int32), one gate's expected share
# of keyswitch rows (3/4 of the 8192 digits non-zero) and one gate's ciphertext traffic
BK_BYTES_P128 = 630 * 12 * 512 * 16
KSK_BYTES_P128 = 1024 * 8 * 3 * 631 * 4
KSK_ROWS_PER_GATE_BYTES = 6144 * 631 * 4
KS_IO_BYTES_PER_GATE = 4100 + 2524
GATE_IO_BYTES = 3 * 2524
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--per-gate", type=int, default=64, help="one-gate-per-launch samples")
    ap.add_argument("--cpu-gates", type=int, default=1500, help="reference CPU sample size, about 15 s of one core (0 = skip)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--adders", type=int, default=512, help="ripple-carry adders in the netlist section (0 = skip)")
    ap.add_argument("--adder-bits", type=int, default=32)
    ap.add_argument("--client-bits", type=int, default=1 << 20,
                    help="client-side section: bits encrypted and decrypted on the device (0 = skip)")
    ap.add_argument("--scaling", default="1,16,64,128,256,512,1024,2048,8192",
                    help="batch sizes for the per-kernel batch-size sweep ('' = skip)")
    ap.add_argument("--serve-clients", type=int, default=32, help="serving section: clients, one key each (0 = skip)")
    ap.add_argument("--serve-rates", default="1e3,1e4,1e5", help="serving section: offered gates/s")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="serving section: arrival window per rate")
    ap.add_argument("--serve-min-requests", type=int, default=4)
    ap.add_argument("--serve-streams", type=int, default=16, help="serving section: streams the clients share")
    ap.add_argument("--serve-depth", type=int, default=2, help="serving section: launches in flight per client")
    ap.add_argument("--serve-max-batch", type=int, default=4096, help="serving section: gates per launch at most")
    ap.add_argument("--serve-pack", type=int, default=1024,
                    help="serving section: pack several clients' levels into one launch from this many ready "
                         "gates on (0 = off)")
    ap.add_argument("--serve-max-requests", type=int, default=160)
    ap.add_argument("--pbig-launches", type=int, default=64,
                    help="N=2048 mixed-batch section: launches (0 = skip)")
    ap.add_argument("--host-api", type=int, default=4096,
                    help="host gate API section: gates per bootsGate_batch call on host LweSample arrays (0 = skip)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default small-batch crossover")
    ap.add_argument("--crossovers", default="",
                    help="LAT,SMALL: use the latency kernel up to LAT gates and the two-wavefront kernel up to "
                         "SMALL instead of autotuning")
    ap.add_argument("--no-compare-modes", action="store_true",
                    help="skip timing the batched kernel on the radix-8 transform beside the default one")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one plain copy kernel of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def profiled_traffic():
    """HBM bytes per batch launch of the blind-rotation kernel, from the
    calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE passes summarised in profiles/."""
    path = os.path.join(ROOT, "profiles", "br4w_kernel_p128_batch4096.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def bench_scaling(T, sk, P, sizes, seed):
    """gates/s of each blind-rotation kernel across batch sizes (input for the
    automatic kernel choice): 2 = one bootstrap per wavefront, 3 = two
    wavefronts per bootstrap, 0 = automatic."""
    rng = np.random.default_rng(seed + 3)
    maxb = max(sizes)
    ops = rng.integers(0, 6, size=maxb).astype(np.int32)
    bits = rng.integers(0, 2, size=(2, maxb)).astype(np.int32)
    w = P.n + 1
    pool = np.concatenate([sk.encrypt(bits[0]), sk.encrypt(bits[1]), np.zeros((maxb, w), np.int32)])
    d_pool = T.DeviceArray.from_numpy(pool)
    idx = np.arange(maxb, dtype=np.int32)
    d_ops, d_i0 = T.DeviceArray.from_numpy(ops), T.DeviceArray.from_numpy(idx)
    d_i1, d_o = T.DeviceArray.from_numpy(idx + maxb), T.DeviceArray.from_numpy(idx + 2 * maxb)
    out = {}
    for B in sizes:
        row = {}
        for mode in (5, 3, 2, 0):
            T.lib().tfhe_amd_set_launch_config(mode, -1, -1)
            T.check(T.lib().tfhe_amd_gates_device(sk.cloud, B, d_ops.ptr, d_i0.ptr, d_i1.ptr, d_i0.ptr, d_o.ptr,
                                                  d_pool.ptr, None), "gates")
            T.synchronize()
            reps = max(1, min(8, 8192 // B))
            t0 = time.perf_counter()
            for _ in range(reps):
                T.check(T.lib().tfhe_amd_gates_device(sk.cloud, B, d_ops.ptr, d_i0.ptr, d_i1.ptr, d_i0.ptr, d_o.ptr,
                                                      d_pool.ptr, None), "gates")
            T.synchronize()
            dt = (time.perf_counter() - t0) / reps
            row[{2: "one_per_wavefront_batched", 3: "two_wave_split", 5: "digit_parallel", 0: "auto"}[mode]] = {
                "gates_per_s": round(B / dt, 1), "ms": round(dt * 1e3, 3)}
        out[str(B)] = row
        print(f"[bench] batch {B}: {row}", file=sys.stderr, flush=True)
    T.lib().tfhe_amd_set_launch_config(0, -1, -1)
    dec = sk.decrypt(d_pool.numpy()[2 * maxb:2 * maxb + max(sizes)])
    truth = {0: lambda x, y: 1 - (x & y), 1: lambda x, y: x & y, 2: lambda x, y: x | y, 3: lambda x, y: x ^ y,
             4: lambda x, y: 1 - (x ^ y), 5: lambda x, y: 1 - (x | y)}
    exp = np.array([truth[o](x, y) for o, x, y in zip(ops, bits[0], bits[1])])
    out["decrypt_errors"] = int((dec != exp).sum())
    for d in (d_pool, d_ops, d_i0, d_i1, d_o):
        d.free()
    return out


def bench_serving(T, P, args):
    """Poisson stream of random layered circuits (depth U[1,64], width U[1,512],
    gate types uniform over all 11) from clients with their own keys resident in
    HBM, served by the continuous-batching request loop. Per offered rate:
    request latency p50/p99 (submit to outputs on the host) and gates/s."""
    import threading
    from tfhe_amd.circuits import Server, random_layered_circuit
    n_cl = args.serve_clients
    t0 = time.perf_counter()
    keys = []
    for i in range(n_cl):
        sk = T.SecretKey.generate_on_device(P, seed=1000 + i)
        sk.upload()
        keys.append(sk)
        if (i + 1) % 8 == 0:
            print(f"[bench] serving: {i + 1} client keys generated on the device "
                  f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    keygen_s = time.perf_counter() - t0
    S = Server(n_streams=args.serve_streams, depth=args.serve_depth, max_batch=args.serve_max_batch,
               wire_budget_bytes=4 << 30, pack_min_gates=args.serve_pack)
    cids = [S.add_client(sk.cloud, P.n) for sk in keys]
    mean_gates = 32.5 * 256.5
    out = {"workload": f"{n_cl} clients (own P128 key each, generated on the device, resident), Poisson arrivals, "
                       "random layered circuits "
                       "depth U[1,64] x width U[1,512], 11 gate types uniform", "clients": n_cl,
           "streams": args.serve_streams, "launches_in_flight_per_client": args.serve_depth,
           "pack_min_gates": args.serve_pack, "max_batch": args.serve_max_batch,
           "key_bytes_resident": S.stats()["key_bytes"], "keygen_and_upload_s": round(keygen_s, 2),
           "rates": {}}
    rng = np.random.default_rng(args.seed + 5)
    for rate in [float(x) for x in args.serve_rates.split(",") if x.strip()]:
        # arrival times (s) of a Poisson process of requests at rate / mean_gates
        arr, t = [], 0.0
        while len(arr) < args.serve_max_requests and (t < args.serve_seconds or len(arr) < args.serve_min_requests):
            t += rng.exponential(mean_gates / rate)
            arr.append(t)
        reqs = []
        for a in arr:
            c = int(rng.integers(0, n_cl))
            C, last = random_layered_circuit(rng, int(rng.integers(1, 65)), int(rng.integers(1, 513)))
            bits = rng.integers(0, 2, size=len(last)).astype(np.int32)
            reqs.append((a, c, C, bits, keys[c].encrypt(bits), last))
        ids = [None] * len(reqs)
        start = time.perf_counter()

        def feed():
            for j, (a, c, C, bits, cts, last) in enumerate(reqs):
                d = start + a - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[j] = S.submit(cids[c], C, np.arange(len(bits)), cts, last)

        th = threading.Thread(target=feed)
        th.start()
        lat, outs = [], []
        for j in range(len(reqs)):
            while ids[j] is None:
                time.sleep(0.0005)
            o, l = S.wait(ids[j], -1)
            outs.append(o)
            lat.append(l)
        end = time.perf_counter()
        th.join()
        gates = sum(len(r[2].op) for r in reqs)
        errors = 0
        for (a, c, C, bits, cts, last), o in zip(reqs, outs):
            errors += int((keys[c].decrypt(o) != C.evaluate(bits)[last]).sum())
        lat = np.array(lat)
        st = S.stats()
        row = {"requests": len(reqs), "gates": gates, "offered_gates_per_s": rate,
               "gates_per_s": round(gates / (end - start), 1), "window_s": round(end - start, 2),
               "p50_ms": round(float(np.percentile(lat, 50)), 2),
               # a 99th percentile needs at least 100 samples; below that only the maximum is reported
               "p99_ms": round(float(np.percentile(lat, 99)), 2) if len(lat) >= 100 else None,
               "max_ms": round(float(lat.max()), 2), "output_bit_errors": errors,
               "launches_total": st["launches"], "packed_launches_total": st["packed_launches"],
               "max_launch_gates": st["max_launch_gates"],
               "wire_bytes_peak": st["wire_bytes_peak"]}
        out["rates"][f"{rate:g}"] = row
        print(f"[bench] serving rate {rate:g}: {row}", file=sys.stderr, flush=True)
    S.close()
    for sk in keys:
        sk.close()
    return out


def bench_pbig(T, args):
    """N = 2048 parameter set (n=750, l=2, Bgbit=10, stdevs 2^-17 / 2^-44,
    keyswitch t=8 basebit=2): a mixed list of 16384 MUX + 16384 two-input gates
    (uniform over the 10 two-input types), shuffled, fresh encryptions resident
    in HBM, run as `launches` consecutive launches with sizes drawn uniformly
    from [1, 8192]. The drawn sizes sum to more than 32768, so two splits are
    reported: "drawn" takes the gates cyclically from the list (every launch has
    its drawn size), "scaled" scales the sizes to sum to exactly 32768. Plus
    one launch of 8192 gates from the same list."""
    P = T.big_params()
    sk = T.SecretKey(P, seed=args.seed)
    sk.upload()
    # kernel shapes per batch size for this key: the four-wavefront pair kernel and the
    # two-wavefront kernel timed on a ladder of sizes; the crossover goes to automatic mode
    tune = None
    if not args.no_autotune:
        _, pair_max, table = T.autotune(sk.cloud)
        tune = {"pair_kernel_up_to": pair_max, "two_wave_kernel_above": pair_max,
                "ms": table}
        print(f"[bench] N=2048 autotune: pair kernel up to {pair_max} gates, two-wavefront above", file=sys.stderr,
              flush=True)
    rng = np.random.default_rng(args.seed + 4)
    G = 32768
    ops = np.concatenate([np.full(G // 2, 10), rng.integers(0, 10, size=G // 2)]).astype(np.int32)
    rng.shuffle(ops)
    drawn = rng.integers(1, 8193, size=args.pbig_launches).astype(np.int64)
    scaled = np.maximum(1, np.floor(drawn * G / drawn.sum())).astype(np.int64)
    scaled[-1] += G - scaled.sum()
    while scaled[-1] < 1:  # keep every launch non-empty
        j = int(np.argmax(scaled[:-1]))
        scaled[j] -= 1
        scaled[-1] += 1
    bits = rng.integers(0, 2, size=(3, G)).astype(np.int32)
    w = P.n + 1
    pool = np.concatenate([sk.encrypt(bits[0]), sk.encrypt(bits[1]), sk.encrypt(bits[2]), np.zeros((G, w), np.int32)])
    d_pool = T.DeviceArray.from_numpy(pool)
    # index tables twice over so a launch can wrap around the end of the list
    idx = np.arange(2 * G, dtype=np.int32) % G
    ops2 = np.concatenate([ops, ops])
    d_ops = T.DeviceArray.from_numpy(ops2)
    d_idx = T.DeviceArray.from_numpy(np.concatenate([idx, idx + G, idx + 2 * G, idx + 3 * G]))
    i32 = 4
    truth = [lambda a, b, c: 1 - (a & b), lambda a, b, c: a & b, lambda a, b, c: a | b, lambda a, b, c: a ^ b,
             lambda a, b, c: 1 - (a ^ b), lambda a, b, c: 1 - (a | b), lambda a, b, c: (1 - a) & b,
             lambda a, b, c: a & (1 - b), lambda a, b, c: (1 - a) | b, lambda a, b, c: a | (1 - b),
             lambda a, b, c: b if a else c]
    exp = np.array([truth[o](x, y, z) for o, x, y, z in zip(ops, bits[0], bits[1], bits[2])])

    def run(sizes, streams=(None,)):
        off = 0
        for k, sz in enumerate(sizes):
            sz = int(sz)
            T.check(T.lib().tfhe_amd_gates_device(
                sk.cloud, sz, d_ops.ptr + i32 * off, d_idx.ptr + i32 * off, d_idx.ptr + i32 * (2 * G + off),
                d_idx.ptr + i32 * (4 * G + off), d_idx.ptr + i32 * (6 * G + off), d_pool.ptr,
                streams[k % len(streams)]), "gates")
            off = (off + sz) % G

    # the same launches spread round-robin over four HIP streams: every stream has a
    # workspace of its own, so consecutive launches overlap on the device (wall time only)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hs = [ctypes.c_void_p() for _ in range(4)]
    for h in hs:
        T.check(hip.hipStreamCreate(ctypes.byref(h)), "hipStreamCreate")
    streams = tuple(h.value for h in hs)

    def timed_streams(sizes):
        T.synchronize()
        t0 = time.perf_counter()
        run(sizes, streams)
        for h in hs:
            T.check(hip.hipStreamSynchronize(h), "hipStreamSynchronize")
        dt = time.perf_counter() - t0
        gates = int(np.sum(sizes))
        return {"gates": gates, "gates_per_s": round(gates / dt, 1), "ms_total": round(dt * 1e3, 2), "streams": 4}

    def timed(sizes):
        T.synchronize()
        T.profile_read()
        T.profile_enable(True)
        t0 = time.perf_counter()
        run(sizes)
        T.synchronize()
        dt = time.perf_counter() - t0
        br_ms, ks_ms, _ = T.profile_read()
        T.profile_enable(False)
        gates = int(np.sum(sizes))
        # bootstraps: MUX gates count twice
        mux = 0
        off = 0
        for sz in sizes:
            sel = (np.arange(off, off + int(sz)) % G)
            mux += int((ops[sel] == 10).sum())
            off = (off + int(sz)) % G
        boots = gates + mux
        return {"gates": gates, "gates_per_s": round(gates / dt, 1), "bootstraps_per_s": round(boots / dt, 1),
                "ms_total": round(dt * 1e3, 2), "blind_rotate_ms": round(br_ms, 2), "keyswitch_ms": round(ks_ms, 2),
                "fp64_tflops_blind_rotate": round(750 * 409600 * boots / (br_ms / 1e3) / 1e12, 3) if br_ms > 0 else None}

    run(scaled)  # warm-up, and every gate of the list written once
    T.synchronize()
    dec = sk.decrypt(d_pool.numpy()[3 * G:])
    errors = int((dec != exp).sum())
    out = {"workload": "N=2048 n=750 l=2 Bgbit=10 ks t=8 basebit=2: 16384 MUX + 16384 two-input gates, "
                       f"{len(drawn)} launches with sizes U[1,8192]",
           "drawn": dict(timed(drawn), launch_sizes={"min": int(drawn.min()), "median": int(np.median(drawn)),
                                                     "max": int(drawn.max())}),
           "scaled": dict(timed(scaled), launch_sizes={"min": int(scaled.min()), "median": int(np.median(scaled)),
                                                       "max": int(scaled.max())}),
           "batch_8192": timed([8192]),
           "decrypt_errors": errors,
           "autotune": tune}
    if not args.no_compare_modes:
        # the same 8192-gate launch on the pair kernel (four wavefronts per bootstrap) on the
        # radix-8 stages (mode 8) and on the fused butterflies (mode 4), and on the
        # two-wavefront kernel (mode 6, the automatic choice at this size), alternating
        cmp = {"pair_radix8_transform": [], "pair_fused_butterfly": [], "two_wave_fused_butterfly": []}
        for _ in range(2):
            for mode, name in ((8, "pair_radix8_transform"), (4, "pair_fused_butterfly"),
                               (6, "two_wave_fused_butterfly")):
                T.set_launch_config(mode)
                try:
                    cmp[name].append(timed([8192])["gates_per_s"])
                finally:
                    T.set_launch_config(0)
        out["batch_8192_kernel_variants_gates_per_s"] = cmp
    run(scaled, streams)  # warm-up of the streams' workspaces
    out["drawn_4_streams"] = timed_streams(drawn)
    out["scaled_4_streams"] = timed_streams(scaled)
    T.synchronize()
    for h in hs:
        hip.hipStreamSynchronize(h)
    out["decrypt_errors_4_streams"] = int((sk.decrypt(d_pool.numpy()[3 * G:]) != exp).sum())
    for h in hs:
        hip.hipStreamDestroy(h)
    print(f"[bench] N=2048 mixed: {out}", file=sys.stderr, flush=True)
    for d in (d_pool, d_ops, d_idx):
        d.free()
    sk.close()
    return out


def bench_adders(T, sk, P, n_adders, bits, reps, seed):
    """512 independent 32-bit ripple-carry adders as one netlist (5 bootstraps per
    bit), run level by level from a captured hipGraph; bit-exact check of every sum."""
    from tfhe_amd.circuits import Netlist, levelize, ripple_carry_adders
    C, sums, zero = ripple_carry_adders(n_adders, bits)
    rng = np.random.default_rng(seed + 7)
    x = rng.integers(0, 2 ** bits, size=n_adders, dtype=np.int64)
    y = rng.integers(0, 2 ** bits, size=n_adders, dtype=np.int64)
    nb = n_adders * bits
    ib = np.zeros(2 * nb + 1, np.int32)
    sh = np.arange(bits)
    ib[:nb] = ((x[:, None] >> sh) & 1).reshape(-1)
    ib[nb:2 * nb] = ((y[:, None] >> sh) & 1).reshape(-1)
    pool = np.zeros((C.n_wires, P.n + 1), np.int32)
    pool[:2 * nb] = sk.encrypt(ib[:2 * nb])
    pool[zero, P.n] = -(1 << 29)  # trivial encryption of 0 as carry-in
    out = {}
    d = T.DeviceArray.from_numpy(pool)
    # balanced: the as-soon-as-possible depth, levels filled to widths of the GPU's SIMD
    # count times a power of two
    for sched, name in ((0, "asap"), (1, "alap"), (2, "balanced")):
        n_levels, lv = levelize(C, sched)
        widths = np.bincount(lv, minlength=n_levels + 1)[1:]
        nl = Netlist(sk.cloud, C, sched)
        nl.run(d)  # capture + first replay
        T.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            nl.run(d)
        T.synchronize()
        dt = (time.perf_counter() - t0) / reps
        res = d.numpy()
        nl.close()
        dec = sk.decrypt(res[sums.reshape(-1)]).reshape(sums.shape).astype(np.int64)
        got = (dec << np.arange(bits + 1)).sum(axis=1)
        out[name] = {"gates_per_s": round(len(C.op) / dt, 1), "ms_per_run": round(dt * 1e3, 3),
                     "levels": int(n_levels), "width_min": int(widths.min()), "width_max": int(widths.max()),
                     "width_median": int(np.median(widths)), "wrong_sums": int((got != x + y).sum())}
        print(f"[bench] adders {name}: {out[name]}", file=sys.stderr, flush=True)
    d.free()
    return {"workload": f"{n_adders} independent {bits}-bit ripple-carry adders, {len(C.op)} bootstraps, "
                        "level by level from a captured hipGraph", **out}


def bench_client(T, sk, P, count, seed):
    """Client-side work on the device: batch encryption of `count` bits
    (Philox noise and masks) and batch decryption, both on HBM-resident data,
    and the generation of one secret key set on the device."""
    rng = np.random.default_rng(seed + 11)
    bits = rng.integers(0, 2, size=count).astype(np.int32)
    d_bits = T.DeviceArray.from_numpy(bits)
    d_cts = T.DeviceArray((count, P.n + 1))
    d_dec = T.DeviceArray(count)
    try:
        T.encrypt_device(sk, d_bits, d_cts, seed=seed + 12)  # warm-up
        T.decrypt_device(sk, d_cts, d_bits=d_dec)
        T.synchronize()
        t0 = time.perf_counter()
        T.encrypt_device(sk, d_bits, d_cts, seed=seed + 13)
        T.synchronize()
        t1 = time.perf_counter()
        T.decrypt_device(sk, d_cts, d_bits=d_dec)
        T.synchronize()
        t2 = time.perf_counter()
        errors = int((d_dec.numpy() != bits).sum())
    finally:
        for d in (d_bits, d_cts, d_dec):
            d.free()
    t3 = time.perf_counter()
    k = T.SecretKey.generate_on_device(P, seed=seed + 14)
    keygen_s = time.perf_counter() - t3
    k.close()
    out = {"bits": count,
           "encrypt_per_s": round(count / (t1 - t0), 1),
           "decrypt_per_s": round(count / (t2 - t1), 1),
           "encrypt_ms": round((t1 - t0) * 1e3, 3), "decrypt_ms": round((t2 - t1) * 1e3, 3),
           "decrypt_errors": errors,
           "keygen_device_s": round(keygen_s, 3)}
    print(f"[bench] client side: {out}", file=sys.stderr, flush=True)
    return out


def bench_host_api(T, sk, P, B, seed):
    """The reference-style host entry points, PCIe included: bootsGate_batch on host
    LweSample arrays (the library stages the inputs to HBM, runs the batch and copies
    the outputs back) and bootsNAND, one gate per call."""
    L = T.lib()
    rng = np.random.default_rng(seed + 7)
    ops = rng.integers(0, 6, size=B).astype(np.int32)
    bits = rng.integers(0, 2, size=(2, B)).astype(np.int32)
    H = T.HostCiphertexts(P, 3 * B)
    try:
        H.set(np.concatenate([sk.encrypt(bits[0]), sk.encrypt(bits[1]), np.zeros((B, P.n + 1), np.int32)]))
        arr = lambda idx: (T.LSP * B)(*[H.item(i) for i in idx])
        res, i0, i1 = arr(range(2 * B, 3 * B)), arr(range(B)), arr(range(B, 2 * B))
        run = lambda: L.bootsGate_batch(B, ops.ctypes.data_as(T.i32p), res, i0, i1, i0, sk.cloud)
        T.check(run(), "bootsGate_batch")  # warm-up
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            T.check(run(), "bootsGate_batch")
        dt = (time.perf_counter() - t0) / reps
        lat = []
        for g in range(16):
            t1 = time.perf_counter()
            L.bootsNAND(H.item(2 * B + g), H.item(g), H.item(B + g), sk.cloud)
            lat.append(time.perf_counter() - t1)
        out = H.get()[2 * B:]
    finally:
        H.close()
    truth = {0: lambda x, y: 1 - (x & y), 1: lambda x, y: x & y, 2: lambda x, y: x | y, 3: lambda x, y: x ^ y,
             4: lambda x, y: 1 - (x ^ y), 5: lambda x, y: 1 - (x | y)}
    exp = np.array([truth[o](x, y) for o, x, y in zip(ops, bits[0], bits[1])])
    exp[:16] = 1 - (bits[0][:16] & bits[1][:16])  # the first 16 outputs were rewritten by bootsNAND
    lat_ms = np.array(lat) * 1e3
    res = {"batch_gates": B, "batch_ms": round(dt * 1e3, 3), "batch_gates_per_s": round(B / dt, 1),
           "pcie_bytes_per_batch": B * 3 * (P.n + 1) * 4,
           "boots_nand_p50_ms": round(float(np.percentile(lat_ms, 50)), 4),
           "decrypt_errors": int((sk.decrypt(out) != exp).sum())}
    print(f"[bench] host API: {res}", file=sys.stderr, flush=True)
    return res


def usable_cpus():
    """Cores this process may run on: the affinity mask (what nproc prints), capped
    by a cgroup v2 CPU quota when one is set; plus the CPU model and flags."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    model, flags = "", ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and not model:
                model = line.split(":", 1)[1].strip()
            if line.startswith("flags") and not flags:
                flags = line.split(":", 1)[1]
    except OSError:
        pass
    return {"affinity_cpus": aff, "cgroup_quota_cpus": quota, "usable": min(aff, quota) if quota else aff,
            "model": model, "avx512": " avx512f " in f" {flags} ", "flags": flags}


def cpu_baseline(count, seed):
    """Reference spqlios-FMA CPU path (oracle/_ref, built from the reference's
    own sources) timed on this host; falls back to the oracle port."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if count <= 0:
        return None
    if os.path.exists(harness):
        out = subprocess.run([harness, "bench", str(seed), str(count), "1"], capture_output=True, text=True,
                             timeout=600, check=True).stdout.strip().splitlines()[-1]
        r = json.loads(out)
        res = {"value": round(r["gates_per_s"], 3), "unit": "gates/s", "cores": 1, "kind": "reference",
               "sample": f"{count} bootsNAND on fresh encryptions, P128 keys from seed {seed}, "
                         f"reference spqlios-FMA build, 1 thread",
               "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"]}
        # the same reference build on every core this process may use (the nproc of the
        # affinity mask, capped by a cgroup CPU quota if one is set), independent gates
        # per thread sharing one cloud key; a short sample of 100 gates per thread
        cpus = usable_cpus()
        threads = cpus["usable"]
        res["host"] = cpus
        if threads > 1:
            out = subprocess.run([harness, "bench", str(seed), str(100 * threads), str(threads)],
                                 capture_output=True, text=True, timeout=600,
                                 check=True).stdout.strip().splitlines()[-1]
            m = json.loads(out)
            res["all_threads"] = {"threads": threads, "gates_per_s": round(m["gates_per_s"], 3),
                                  "p50_ms": m["p50_ms"], "p99_ms": m["p99_ms"], "gates": 100 * threads}
        # the reference built for AVX-512 hosts (-march=x86-64-v4, closer to its own
        # -march=native build) beside the portable AVX2/FMA build, one thread, where
        # the host CPU has AVX-512
        h4 = os.path.join(ROOT, "oracle", "_ref", "v4", "ref_harness")
        if os.path.exists(h4) and cpus["avx512"]:
            out = subprocess.run([h4, "bench", str(seed), str(count), "1"], capture_output=True, text=True,
                                 timeout=600, check=True).stdout.strip().splitlines()[-1]
            r4 = json.loads(out)
            res["avx512_build"] = {"flags": "-O2 -march=x86-64-v4", "gates_per_s": round(r4["gates_per_s"], 3),
                                   "p50_ms": r4["p50_ms"], "threads": 1}
        res["host"] = {k: v for k, v in cpus.items() if k != "flags"}
        return res
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    P = O.p128()
    keys = O.keygen(P, seed)
    n = 4
    a = np.stack([O.boots_encrypt(keys, 1) for _ in range(n)])
    t0 = time.perf_counter()
    O.gates(keys, np.zeros(n, np.int32), a, a, threads=1)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 3), "unit": "gates/s", "cores": 1, "kind": "port",
            "sample": f"{n} NAND gates, exact-integer oracle port, 1 thread"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: independent gates shard across GPUs with a key replica each "
              "and no exchange, so no multi-GPU run is made here.", file=sys.stderr)
        sys.exit(2)
    import tfhe_amd as T

    if T.lib().tfhe_amd_device_count() < 1:
        raise SystemExit("no GPU visible to libtfhe_amd.so")
    if args.copy_calib_mb:
        nb = args.copy_calib_mb << 20
        src, dst = T.DeviceArray(nb // 4), T.DeviceArray(nb // 4)
        T.check(T.lib().tfhe_amd_copy_device(dst.ptr, src.ptr, nb, None), "copy")
        T.synchronize()
        print(f"[bench] copy kernel: {nb} bytes read, {nb} bytes written", file=sys.stderr, flush=True)
        src.free()
        dst.free()
    P = T.default_params(110)
    sk = T.SecretKey(P, seed=args.seed)
    sk.upload()
    B = args.batch
    rng = np.random.default_rng(args.seed + 1)
    ops = rng.integers(0, 6, size=B).astype(np.int32)
    bits = rng.integers(0, 2, size=(2, B)).astype(np.int32)
    a, b = sk.encrypt(bits[0]), sk.encrypt(bits[1])
    w = P.n + 1
    pool = np.concatenate([a, b, np.zeros((B, w), np.int32)])
    idx = np.arange(B, dtype=np.int32)
    d_pool = T.DeviceArray.from_numpy(pool)
    d_ops = T.DeviceArray.from_numpy(ops)
    d_i0, d_i1 = T.DeviceArray.from_numpy(idx), T.DeviceArray.from_numpy(idx + B)
    d_o = T.DeviceArray.from_numpy(idx + 2 * B)

    def step():
        T.gates_device(sk.cloud, d_ops, d_i0, d_i1, d_i0, d_o, d_pool)

    print(f"[bench] warmup {args.warmup} x {B} gates", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    T.synchronize()
    T.profile_read()
    T.profile_enable(True)
    T.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    T.synchronize()
    dt = time.perf_counter() - t0
    br_ms, ks_ms, launches = T.profile_read()
    T.profile_enable(False)
    out = d_pool.numpy()[2 * B:]
    dec = sk.decrypt(out)
    truth = {0: lambda x, y: 1 - (x & y), 1: lambda x, y: x & y, 2: lambda x, y: x | y, 3: lambda x, y: x ^ y,
             4: lambda x, y: 1 - (x ^ y), 5: lambda x, y: 1 - (x | y)}
    exp = np.array([truth[o](x, y) for o, x, y in zip(ops, bits[0], bits[1])])
    errors = int((dec != exp).sum())
    ms_per_step = 1e3 * dt / args.steps
    batched = B * args.steps / dt
    print(f"[bench] batched {batched:.1f} gates/s, {ms_per_step:.3f} ms/step, errors {errors}", file=sys.stderr,
          flush=True)

    mode_cmp = None
    if not args.no_compare_modes:
        # the batched kernel on both transforms, and on the fused-butterfly transform without
        # the row-synchronised eight-bootstrap workgroups, alternating, same batch (gates/s)
        mode_cmp = {"radix8_transform": [], "fused_butterfly_one_bootstrap_workgroups": [],
                    "fused_butterfly_transform": []}
        for _ in range(2):
            for mode, name in ((7, "radix8_transform"), (9, "fused_butterfly_one_bootstrap_workgroups"),
                               (2, "fused_butterfly_transform")):
                T.lib().tfhe_amd_set_launch_config(mode, -1, -1)
                step()
                T.synchronize()
                t1 = time.perf_counter()
                for _ in range(args.steps):
                    step()
                T.synchronize()
                mode_cmp[name].append(round(B * args.steps / (time.perf_counter() - t1), 1))
        T.lib().tfhe_amd_set_launch_config(0, -1, -1)
        print(f"[bench] batched kernel variants: {mode_cmp}", file=sys.stderr, flush=True)

    # one gate per launch (same kernels, count = 1, synchronised each time)
    lat = []
    one = T.DeviceArray.from_numpy(np.zeros(1, np.int32))
    for g in range(args.per_gate):
        d1_ops = T.DeviceArray.from_numpy(ops[g:g + 1])
        d1_i0 = T.DeviceArray.from_numpy(np.array([g], np.int32))
        d1_i1 = T.DeviceArray.from_numpy(np.array([g + B], np.int32))
        d1_o = T.DeviceArray.from_numpy(np.array([2 * B + g], np.int32))
        T.synchronize()
        t1 = time.perf_counter()
        T.gates_device(sk.cloud, d1_ops, d1_i0, d1_i1, d1_i0, d1_o, d_pool)
        T.synchronize()
        lat.append(time.perf_counter() - t1)
    one.free()
    lat_ms = np.array(lat) * 1e3
    per_gate = 1e3 / lat_ms.mean()
    print(f"[bench] per-gate {per_gate:.1f} gates/s, p50 {np.percentile(lat_ms, 50):.3f} ms", file=sys.stderr,
          flush=True)

    autotune = None
    if args.crossovers:
        lat, small = (int(x) for x in args.crossovers.split(","))
        T.set_launch_config(0, lat, small)
        autotune = {"latency_kernel_up_to": lat, "two_wave_kernel_up_to": small, "forced": True}
        print(f"[bench] crossovers forced: latency kernel up to {lat} gates, two-wavefront up to {small}",
              file=sys.stderr, flush=True)
    elif not args.no_autotune:
        lat, small, tune = T.autotune(sk.cloud)
        autotune = {"latency_kernel_up_to": lat, "two_wave_kernel_up_to": small,
                    "ms": {name: {str(k): round(v, 3) for k, v in t.items()} for name, t in tune.items()}}
        print(f"[bench] autotune: latency kernel (4l wavefronts) up to {lat} gates, two-wavefront up to {small}",
              file=sys.stderr, flush=True)
    sizes = [int(x) for x in args.scaling.split(",") if x.strip()]
    scaling = bench_scaling(T, sk, P, sizes, args.seed) if sizes else None
    client = bench_client(T, sk, P, args.client_bits, args.seed) if args.client_bits > 0 else None
    adders = bench_adders(T, sk, P, args.adders, args.adder_bits, 2, args.seed) if args.adders > 0 else None
    serving = bench_serving(T, P, args) if args.serve_clients > 0 else None
    pbig = bench_pbig(T, args) if args.pbig_launches > 0 else None
    host_api = bench_host_api(T, sk, P, args.host_api, args.seed) if args.host_api > 0 else None

    br_avg_s = br_ms / max(launches, 1) / 1e3
    achieved_tflops = FLOP_PER_BOOTSTRAP_P128 * B / br_avg_s / 1e12 if br_avg_s > 0 else None
    ks_avg_s = ks_ms / max(launches, 1) / 1e3
    p50_s = float(np.percentile(lat_ms, 50)) / 1e3

    def hbm_view(nbytes, secs):
        gbs = nbytes / secs / 1e9
        return {"algorithmic_bytes": nbytes, "ms": round(secs * 1e3, 4), "GB_s": round(gbs, 1),
                "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 5)}

    hbm = {
        "keyswitch_batched": hbm_view(KSK_BYTES_P128 + B * KS_IO_BYTES_PER_GATE, ks_avg_s) if ks_avg_s > 0 else None,
        "one_gate_path": hbm_view(BK_BYTES_P128 + KSK_ROWS_PER_GATE_BYTES + GATE_IO_BYTES, p50_s),
        "peak_GB_s": HBM_PEAK_GBS,
    }
    res = {
        "metric": "gates/s bootstrapped on 1 MI355X, batched vs per-gate; p50/p99 latency",
        "value": round(batched, 2),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{B} independent gate bootstraps per launch (NAND/AND/OR/XOR/XNOR/NOR uniform), "
                               f"P128 n=630 N=1024 k=1 l=3 Bgbit=7 ks t=8 basebit=2, one cloud key (seed {args.seed})",
                   "batch": B, "params": "P128"},
        "per_gate": {"gates_per_s": round(per_gate, 2), "samples": args.per_gate,
                     "p50_ms": round(float(np.percentile(lat_ms, 50)), 4),
                     "p99_ms": round(float(np.percentile(lat_ms, 99)), 4)},
        "batched_over_per_gate": round(batched / per_gate, 2),
        "batch_latency_ms": {"p50": round(ms_per_step, 4), "p99": round(ms_per_step, 4)},
        "decrypt_errors": errors,
        "kernel_variants_gates_per_s": mode_cmp,
        "autotune": autotune,
        "batch_size_sweep": scaling,
        "netlist_adders": adders,
        "client_side": client,
        "serving": serving,
        "pbig_mixed": pbig,
        "host_api": host_api,
        "kernel_ms": {"blind_rotate_avg": round(br_ms / max(launches, 1), 4),
                      "keyswitch_avg": round(ks_ms / max(launches, 1), 4)},
        "roofline": {"bound": "fp64-valu", "achieved": round(achieved_tflops, 3) if achieved_tflops else None,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / FP64_PEAK_TFLOPS, 4) if achieved_tflops else None,
                     "traffic": profiled_traffic() if B == 4096 else None,
                     "kernel": "br4w_kernel (blind rotation, FP64); traffic from profiles/br4w_kernel_p128_batch4096.md",
                     "note": "the blind rotation runs on the FP64 vector ALUs (no matrix-core instruction); "
                             "MI355X FP64 vector peak 78.6 TFLOP/s"},
        "hbm_views": hbm,
    }
    # everything this run put on the device is released before the line is printed;
    # the only child processes are the reference harness runs of cpu_baseline, each
    # awaited by subprocess.run
    for x in (d_pool, d_ops, d_i0, d_i1, d_o):
        x.free()
    sk.close()
    T.synchronize()
    try:
        res["cpu_baseline"] = cpu_baseline(args.cpu_gates, args.seed)
    except Exception as e:  # the GPU result stands on its own
        res["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
