"""Keys the reference wrote, loaded from its files into HBM and bootstrapped on
the GPU (tests/golden/ref_io_n16_seed7: a key set at the 110-bit parameters
with n = 16, written by the reference's export functions, libtfhe/tfhe_io.cpp).

- The cloud key file (the leading part of secret.key, exactly what
  export_tfheGateBootstrappingCloudKeySet_toFile writes) loads with
  new_tfheGateBootstrappingCloudKeySet_fromFile, uploads, and the GPU gates on
  the reference's own input ciphertexts equal the CPU oracle's on the loaded
  keys bit for bit and decrypt to the reference's recorded results; their
  phases lie within 2^27 units (2^-5 of the torus) of the reference's outputs.
  They are not bit-equal to them: the reference truncates its FP64 product
  (fft_processors/spqlios/fft_processor_spqlios.cpp:114) and this library
  rounds the exact one, so single units of the accumulator differ and, through
  the gadget digits, re-randomise the noise (measured spread on these 22
  gates: std 2^23.8 units).
- One-row blind rotations on the reference's own key rows: the GPU result is
  the exact product's and lies within 1 unit per coefficient of the
  reference's spqlios output (the reference's own bar,
  test/test-multiplication.cpp:67), with both -1 and +1 differences present.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_io_n16_seed7")
PHASE_BOUND = 1 << 27


@pytest.fixture(scope="module")
def ref_keys(gpu, tmp_path_factory):
    T = gpu
    sk = T.SecretKey.load(os.path.join(GOLD, "secret.key"))
    # the cloud key file: the part of secret.key the cloud key writer produces
    tmp = tmp_path_factory.mktemp("refcloud")
    data = open(os.path.join(GOLD, "secret.key"), "rb").read()
    T.save_cloud_key(sk.cloud, tmp / "mine.key")
    cloud_bytes = open(tmp / "mine.key", "rb").read()
    assert data.startswith(cloud_bytes)
    (tmp / "cloud.key").write_bytes(data[:len(cloud_bytes)])
    ck = T.load_cloud_key(tmp / "cloud.key")
    OP = O.OrcParams(16, 1024, 1, 3, 7, 8, 2, 2.0 ** -15, 2.0 ** -25)
    keys = O.Keys(OP, **sk.export())
    yield T, sk, ck, keys
    T.lib().delete_gate_bootstrapping_cloud_keyset(ck)
    sk.close()


def test_reference_cloud_key_file_bootstraps_on_gpu(ref_keys):
    T, sk, ck, keys = ref_keys
    L = T.lib()
    assert L.tfhe_amd_cloud_key_upload(ck) == 0, L.tfhe_amd_last_error().decode()
    assert L.tfhe_amd_cloud_key_device_bytes(ck) > 0
    g = np.load(os.path.join(GOLD, "gates.npz"))
    ops = g["ops"].astype(np.int32)
    inp = g["inputs"]
    a, b, c = (np.ascontiguousarray(inp[:, j]) for j in range(3))
    print(f"[test] {len(ops)} reference gates on the reference's n = 16 cloud key", flush=True)
    got = T.run_gates(ck, ops, a, b, c)
    ref = O.gates(keys, ops, a, b, c, threads=8)
    assert np.array_equal(got, ref)
    exp = np.array([O.TRUTH[O.GATES[o]](*bits) for o, bits in zip(ops, g["bits"])])
    assert np.array_equal(sk.decrypt(got), exp)
    assert np.array_equal(sk.decrypt(g["outputs"]), exp)
    d = (sk.phases(got).astype(np.int64) - g["phase"].astype(np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)
    assert np.abs(d).max() <= PHASE_BOUND, np.abs(d).max()


def test_blind_rotate_rows_within_one_unit_of_spqlios(ref_keys):
    T, sk, ck, keys = ref_keys
    L = T.lib()
    f = np.load(os.path.join(GOLD, "blind_rotate_rows.npz"))
    P = sk.params
    tp = P.ptr.contents.tgsw_params.contents.tlwe_params
    base = ck.contents.bkFFT.contents.bkFFT
    OP = O.OrcParams(1, 1024, 1, 3, 7, 8, 2, 2.0 ** -15, 2.0 ** -25)
    diffs = []
    for case in range(len(f["row"])):
        row, abar = int(f["row"][case]), np.array([f["abar"][case]], np.int32)
        acc0 = f["acc_in"][case]
        acc = L.new_TLweSample(tp)
        try:
            for comp in range(2):
                ctypes.memmove(acc.contents.a[comp].coefsT, np.ascontiguousarray(acc0[comp]).ctypes.data, 4 * 1024)
            rowp = ctypes.cast(ctypes.addressof(base.contents) + row * ctypes.sizeof(base._type_), type(base))
            L.tfhe_blindRotate_FFT(acc, rowp, abar.ctypes.data_as(T.i32p), 1, P.ptr.contents.tgsw_params)
            got = np.zeros((2, 1024), np.int32)
            for comp in range(2):
                ctypes.memmove(got[comp].ctypes.data, acc.contents.a[comp].coefsT, 4 * 1024)
        finally:
            L.delete_TLweSample(acc)
        exact = acc0.copy()
        O.lib().orc_blind_rotate(ctypes.byref(OP), O.ptr(np.ascontiguousarray(keys.bk[row:row + 1])), O.ptr(abar),
                                 O.ptr(exact))
        assert np.array_equal(got, exact), case
        d = (got.astype(np.int64) - f["acc_out"][case] + (1 << 31)) % (1 << 32) - (1 << 31)
        assert np.abs(d).max() <= 1, (case, np.abs(d).max())
        diffs.append(d)
    d = np.concatenate([x.ravel() for x in diffs])
    # the reference truncates toward zero, so both signs occur, and most coefficients agree
    assert (d == 1).any() and (d == -1).any() and (d == 0).mean() > 0.4
