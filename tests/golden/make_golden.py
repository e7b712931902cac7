#!/usr/bin/env python3
"""Writes the golden vectors in tests/golden/ from the reference build.

Needs oracle/_ref/ref_harness (make -C oracle ref, from the reference's own
sources); run in a container where those sources exist. The vectors are data
only: hashes of the reference's seed-0 keys, a few reference keyswitch
input/output pairs, the reference's output phases for gates on encryptions
that the oracle replays from the same seed, and the small files of the
reference's key-file format (parameters, LWE key, ciphertexts) for seed 11,
and every object of that format for a key set with n = 16 (seed 7) with the
reference's outputs of 22 gates under it.

    python tests/golden/make_golden.py
    python tests/golden/make_golden.py layout   # ref_struct_layout.txt only
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_py as O  # noqa: E402


def run(*args):
    subprocess.run([O.REF_HARNESS] + [str(a) for a in args], check=True)


def main():
    if not os.path.exists(O.REF_HARNESS):
        subprocess.run(["make", "-s", "-j8", "-C", O.ORACLE_DIR, "ref"], check=True)
    P = O.p128()
    with tempfile.TemporaryDirectory() as tmp:
        # keys (seed 0)
        run("keys", 0, tmp)
        R = O.load_ref_keys(P, tmp)
        K = O.keygen(P, 0)
        bdiff = K.bk[:, :, 1, :].astype(np.int64) - R.bk[:, :, 1, :]
        np.savez_compressed(
            os.path.join(HERE, "ref_p128_seed0_keys.npz"),
            lwe_key_hash=np.uint64(O.fnv1a(R.lwe_key)),
            tlwe_key_hash=np.uint64(O.fnv1a(R.tlwe_key)),
            ksk_hash=np.uint64(O.fnv1a(R.ksk)),
            bk_mask_hash=np.uint64(O.fnv1a(np.ascontiguousarray(R.bk[:, :, 0, :]))),
            bk_b_row0=np.ascontiguousarray(R.bk[0, 0, 1, :]),
            bk_b_diff_max=np.int64(np.abs(bdiff).max()),
            bk_b_diff_frac=np.float64((bdiff != 0).mean()),
        )
        # keyswitch pairs
        path = os.path.join(tmp, "ks.bin")
        run("ks", 0, 4, path)
        raw = np.fromfile(path, np.int32)
        cnt, nx, n = raw[:3]
        body = raw[3:].reshape(cnt, nx + 1 + n + 1 + 2)
        np.savez_compressed(os.path.join(HERE, "ref_p128_seed0_keyswitch.npz"),
                            inputs=body[:, :nx + 1].copy(), outputs=body[:, nx + 1:nx + 1 + n + 1].copy(),
                            variance=body[:, nx + 1 + n + 1:].copy().view(np.float64).reshape(cnt))
        # gates: 1000 NAND, 256 uniform over the six two-input gates, 64 MUX
        for name, count, ops in (("nand1000", 1000, "a"), ("rand6_256", 256, "rand6"), ("mux64", 64, "k")):
            path = os.path.join(tmp, name + ".bin")
            run("gates", 0, count, ops, path)
            G = O.read_ref_gates(path)
            nin = 3 if ops == "k" else 2
            inputs = np.ascontiguousarray(G["cts"][:, :nin, :])
            np.savez_compressed(os.path.join(HERE, f"ref_p128_seed0_{name}.npz"),
                                ops=G["ops"].astype(np.int8), bits=G["bits"].astype(np.int8),
                                phase=G["phase"], inputs_hash=np.uint64(O.fnv1a(inputs)),
                                ref_ns=G["ns"])
        # key-file format: the small files the reference writes for seed 11
        # (parameters, LWE key, 16 ciphertexts of recorded bits, NAND of the
        # first two); the key sets themselves are too large to keep
        io_dir = os.path.join(HERE, "ref_io_p128_seed11")
        os.makedirs(io_dir, exist_ok=True)
        run("files", 11, tmp, 16)
        for f in ("params.gb", "lwekey.bin", "cts.bin", "nand.bin", "bits.txt"):
            with open(os.path.join(tmp, f), "rb") as src, open(os.path.join(io_dir, f), "wb") as dst:
                dst.write(src.read())
        # every object of the reference's file functions, from a key set at the
        # 110-bit parameters with n = 16 (3 MB), and 22 gates (each type twice) with
        # their reference outputs; the three large derived files (keyswitch key,
        # bootstrapping key, cloud key, byte-identical sections of secret.key) are
        # kept as SHA-256 digests
        sk_dir = os.path.join(HERE, "ref_io_n16_seed7")
        os.makedirs(sk_dir, exist_ok=True)
        run("smallkeys", 7, tmp, 22)
        for f in ("secret.key", "params.gb", "lweparams.txt", "lwekey.bin", "lwesample.bin", "tlweparams.txt",
                  "tlwekey.bin", "tlwesample.bin", "tgswparams.txt", "tgswkey.bin", "tgswsample.bin"):
            with open(os.path.join(tmp, f), "rb") as src, open(os.path.join(sk_dir, f), "wb") as dst:
                dst.write(src.read())
        digests = {}
        for f in ("kskey.bin", "bk.bin", "cloud.key"):
            with open(os.path.join(tmp, f), "rb") as src:
                data = src.read()
            digests[f] = {"bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
        with open(os.path.join(sk_dir, "digests.json"), "w") as dst:
            json.dump(digests, dst, indent=1, sort_keys=True)
        raw = np.fromfile(os.path.join(tmp, "gates.bin"), np.int32)
        count, n = raw[:2]
        rec = raw[2:].reshape(count, 4 + 4 * (n + 1) + 1)
        np.savez_compressed(os.path.join(sk_dir, "gates.npz"), ops=rec[:, 0].copy(), bits=rec[:, 1:4].copy(),
                            inputs=rec[:, 4:4 + 3 * (n + 1)].reshape(count, 3, n + 1).copy(),
                            outputs=rec[:, 4 + 3 * (n + 1):4 + 4 * (n + 1)].copy(), phase=rec[:, -1].copy())
        # one-row blind rotations on the reference's own key rows (its transform)
        path = os.path.join(tmp, "extprod.bin")
        run("extprod", 7, path, 12)
        raw = np.fromfile(path, np.int32)
        count, k, N = raw[:3]
        rec = raw[3:].reshape(count, 2 + 2 * (k + 1) * N)
        np.savez_compressed(os.path.join(sk_dir, "blind_rotate_rows.npz"), row=rec[:, 0].copy(),
                            abar=rec[:, 1].copy(), acc_in=rec[:, 2:2 + (k + 1) * N].reshape(count, k + 1, N).copy(),
                            acc_out=rec[:, 2 + (k + 1) * N:].reshape(count, k + 1, N).copy())
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))


def struct_layout():
    """Sizes and field offsets of the gate API's structs as the reference's own
    headers lay them out (tests/c/struct_layout.c built as C++ against
    src/include), for the binary-layout test."""
    exe = os.path.join(tempfile.mkdtemp(), "struct_layout_ref")
    subprocess.run(["g++", "-std=c++11", "-x", "c++", "-DREFERENCE_HEADERS", "-I", "/root/reference/src/include",
                    os.path.join(os.path.dirname(HERE), "c", "struct_layout.c"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    open(os.path.join(HERE, "ref_struct_layout.txt"), "w").write(out)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "layout":
        struct_layout()
        sys.exit(0)
    main()
