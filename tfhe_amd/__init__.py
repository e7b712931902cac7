"""tfhe_amd: MI355X gate-bootstrapping engine behind the tfhe/tfhe C gate API.

This module is a thin ctypes mirror of the C-ABI in include/tfhe_amd.h
(libtfhe_amd.so). Every bootstrap and keyswitch runs in the library's HIP
kernels; there is no CPU fallback here: `lib()` raises if the shared library
was not built, and gate calls raise if the library reports an error.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
# TFHE_AMD_LIB selects another build of the same library (the sanitizer build of
# the host code, tfhe_amd/_build_san/libtfhe_amd.so, for the CPU tests)
LIB_PATH = os.environ.get("TFHE_AMD_LIB") or os.path.join(PKG, "libtfhe_amd.so")

GATES = ["NAND", "AND", "OR", "XOR", "XNOR", "NOR", "ANDNY", "ANDYN", "ORNY", "ORYN", "MUX"]
GATE_CODE = {g: i for i, g in enumerate(GATES)}
MUX = GATE_CODE["MUX"]

i32 = ctypes.c_int32
i32p = ctypes.POINTER(ctypes.c_int32)
vp = ctypes.c_void_p
f32p = ctypes.POINTER(ctypes.c_float)


class LweParams(ctypes.Structure):
    _fields_ = [("n", i32), ("alpha_min", ctypes.c_double), ("alpha_max", ctypes.c_double)]


class TLweParams(ctypes.Structure):
    _fields_ = [("N", i32), ("k", i32), ("alpha_min", ctypes.c_double), ("alpha_max", ctypes.c_double),
                ("extracted_lweparams", LweParams)]


class TGswParams(ctypes.Structure):
    _fields_ = [("l", i32), ("Bgbit", i32), ("Bg", i32), ("halfBg", i32), ("maskMod", ctypes.c_uint32),
                ("tlwe_params", ctypes.POINTER(TLweParams)), ("kpl", i32), ("h", i32p), ("offset", ctypes.c_uint32)]


class ParameterSet(ctypes.Structure):
    _fields_ = [("ks_t", i32), ("ks_basebit", i32), ("in_out_params", ctypes.POINTER(LweParams)),
                ("tgsw_params", ctypes.POINTER(TGswParams))]


class LweSample(ctypes.Structure):
    _fields_ = [("a", i32p), ("b", i32), ("current_variance", ctypes.c_double)]


class TorusPolynomial(ctypes.Structure):
    _fields_ = [("N", i32), ("coefsT", i32p)]


class TLweSample(ctypes.Structure):
    _fields_ = [("a", ctypes.POINTER(TorusPolynomial)), ("b", ctypes.POINTER(TorusPolynomial)),
                ("current_variance", ctypes.c_double), ("k", i32)]


class TGswSampleFFT(ctypes.Structure):
    _fields_ = [("all_samples", vp), ("sample", vp), ("k", i32), ("l", i32)]


class LweKeySwitchKey(ctypes.Structure):
    _fields_ = [("n", i32), ("t", i32), ("basebit", i32), ("base", i32), ("out_params", vp),
                ("ks0_raw", vp), ("ks1_raw", vp), ("ks", vp)]


class LweBootstrappingKeyFFT(ctypes.Structure):
    _fields_ = [("in_out_params", vp), ("bk_params", ctypes.POINTER(TGswParams)), ("accum_params", vp),
                ("extract_params", vp), ("bkFFT", ctypes.POINTER(TGswSampleFFT)),
                ("ks", ctypes.POINTER(LweKeySwitchKey))]


class CloudKeySet(ctypes.Structure):
    _fields_ = [("params", ctypes.POINTER(ParameterSet)), ("bk", vp),
                ("bkFFT", ctypes.POINTER(LweBootstrappingKeyFFT))]


class SecretKeySet(ctypes.Structure):
    _fields_ = [("params", ctypes.POINTER(ParameterSet)), ("lwe_key", vp), ("tgsw_key", vp), ("cloud", CloudKeySet)]


PP = ctypes.POINTER(ParameterSet)
SKP = ctypes.POINTER(SecretKeySet)
CKP = ctypes.POINTER(CloudKeySet)
LSP = ctypes.POINTER(LweSample)

# name: (restype, argtypes)
_SIGS = {
    "new_default_gate_bootstrapping_parameters": (PP, [i32]),
    "new_gate_bootstrapping_parameters": (PP, [i32, i32, i32, i32, i32, i32, i32, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double]),
    "delete_gate_bootstrapping_parameters": (None, [PP]),
    "new_random_gate_bootstrapping_secret_keyset": (SKP, [PP]),
    "delete_gate_bootstrapping_secret_keyset": (None, [SKP]),
    "delete_gate_bootstrapping_cloud_keyset": (None, [CKP]),
    "new_gate_bootstrapping_cloud_keyset_from_arrays": (CKP, [PP, i32p, i32p]),
    "new_gate_bootstrapping_ciphertext": (LSP, [PP]),
    "new_gate_bootstrapping_ciphertext_array": (LSP, [i32, PP]),
    "delete_gate_bootstrapping_ciphertext": (None, [LSP]),
    "delete_gate_bootstrapping_ciphertext_array": (None, [i32, LSP]),
    "bootsSymEncrypt": (None, [LSP, i32, SKP]),
    "bootsSymDecrypt": (i32, [LSP, SKP]),
    "bootsCONSTANT": (None, [LSP, i32, CKP]),
    "bootsNOT": (None, [LSP, LSP, CKP]),
    "bootsCOPY": (None, [LSP, LSP, CKP]),
    "bootsMUX": (None, [LSP, LSP, LSP, LSP, CKP]),
    "bootsGate_batch": (i32, [i32, i32p, ctypes.POINTER(LSP), ctypes.POINTER(LSP), ctypes.POINTER(LSP),
                              ctypes.POINTER(LSP), CKP]),
    "tfhe_random_generator_setSeed": (None, [ctypes.POINTER(ctypes.c_uint32), i32]),
    "tfhe_blindRotate_FFT": (None, [ctypes.POINTER(TLweSample), ctypes.POINTER(TGswSampleFFT), i32p, i32,
                                    ctypes.POINTER(TGswParams)]),
    "tfhe_blindRotateAndExtract_FFT": (None, [LSP, ctypes.POINTER(TorusPolynomial), ctypes.POINTER(TGswSampleFFT),
                                              i32, i32p, i32, ctypes.POINTER(TGswParams)]),
    "tfhe_bootstrap_woKS_FFT": (None, [LSP, ctypes.POINTER(LweBootstrappingKeyFFT), i32, LSP]),
    "tfhe_bootstrap_FFT": (None, [LSP, ctypes.POINTER(LweBootstrappingKeyFFT), i32, LSP]),
    "lweKeySwitch": (None, [LSP, ctypes.POINTER(LweKeySwitchKey), LSP]),
    "lwePhase": (i32, [LSP, vp]),
    "new_LweSample": (LSP, [ctypes.POINTER(LweParams)]),
    "delete_LweSample": (None, [LSP]),
    "new_TLweSample": (ctypes.POINTER(TLweSample), [ctypes.POINTER(TLweParams)]),
    "delete_TLweSample": (None, [ctypes.POINTER(TLweSample)]),
    "new_TorusPolynomial": (ctypes.POINTER(TorusPolynomial), [i32]),
    "delete_TorusPolynomial": (None, [ctypes.POINTER(TorusPolynomial)]),
    "modSwitchFromTorus32": (i32, [i32, i32]),
    "modSwitchToTorus32": (i32, [i32, i32]),
    "dtot32": (i32, [ctypes.c_double]),
    "tfhe_amd_encrypt_bits": (i32, [SKP, i32, i32p, i32p]),
    "tfhe_amd_phases": (i32, [SKP, i32, i32p, i32p]),
    "tfhe_amd_export_keys": (i32, [SKP, i32p, i32p, i32p, i32p]),
    "tfhe_amd_cloud_of": (CKP, [SKP]),
    "tfhe_amd_cloud_key_upload": (i32, [CKP]),
    "tfhe_amd_cloud_key_release": (None, [CKP]),
    "tfhe_amd_cloud_key_device_bytes": (ctypes.c_size_t, [CKP]),
    "tfhe_amd_gates_device": (i32, [CKP, i32, vp, vp, vp, vp, vp, vp, vp]),
    "tfhe_amd_gates_device_woks": (i32, [CKP, i32, vp, vp, vp, vp, vp, vp, vp]),
    "tfhe_amd_gates_device_noalloc": (i32, [CKP, i32, vp, vp, vp, vp, vp, vp, vp]),
    "tfhe_amd_keyswitch_device": (i32, [CKP, i32, vp, vp, vp]),
    "export_tfheGateBootstrappingParameterSet_toFile": (None, [vp, PP]),
    "new_tfheGateBootstrappingParameterSet_fromFile": (PP, [vp]),
    "export_tfheGateBootstrappingCloudKeySet_toFile": (None, [vp, CKP]),
    "new_tfheGateBootstrappingCloudKeySet_fromFile": (CKP, [vp]),
    "export_tfheGateBootstrappingSecretKeySet_toFile": (None, [vp, SKP]),
    "new_tfheGateBootstrappingSecretKeySet_fromFile": (SKP, [vp]),
    "export_gate_bootstrapping_ciphertext_toFile": (None, [vp, LSP, PP]),
    "import_gate_bootstrapping_ciphertext_fromFile": (None, [vp, LSP, PP]),
    "export_lweKey_toFile": (None, [vp, vp]),
    "new_lweKey_fromFile": (vp, [vp]),
    "delete_LweKey": (None, [vp]),
    "tfhe_amd_encrypt_device": (i32, [SKP, i32, vp, vp, ctypes.c_uint64, ctypes.c_uint64, vp]),
    "tfhe_amd_keygen_device": (SKP, [PP, ctypes.c_uint64]),
    "tfhe_amd_decrypt_device": (i32, [SKP, i32, vp, vp, vp, vp]),
    "tfhe_amd_set_launch_config": (None, [i32, i32, i32]),
    "tfhe_amd_set_keyswitch_path": (None, [i32]),
    "tfhe_amd_get_launch_config": (i32, [i32p, i32p, i32p]),
    "tfhe_amd_kernel_for": (i32, [CKP, i32]),
    "tfhe_amd_autotune": (i32, [CKP, i32p, i32p, f32p]),
    "tfhe_amd_profile_enable": (None, [i32]),
    "tfhe_amd_profile_read": (i32, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), i32p]),
    "tfhe_amd_copy_device": (i32, [vp, vp, ctypes.c_size_t, vp]),
    "tfhe_amd_malloc": (vp, [ctypes.c_size_t]),
    "tfhe_amd_free": (i32, [vp]),
    "tfhe_amd_memcpy_h2d": (i32, [vp, vp, ctypes.c_size_t]),
    "tfhe_amd_memcpy_d2h": (i32, [vp, vp, ctypes.c_size_t]),
    "tfhe_amd_synchronize": (i32, []),
    "tfhe_amd_stream": (vp, []),
    "tfhe_amd_device_count": (i32, []),
    "tfhe_amd_last_error": (ctypes.c_char_p, []),
}
for _g in GATES[:-1]:
    _SIGS["boots" + _g] = (None, [LSP, LSP, LSP, CKP])

_lib = None


class TfheAmdError(RuntimeError):
    pass


def lib():
    """Loads libtfhe_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TfheAmdError(f"{LIB_PATH} not built: run python -m tfhe_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what="tfhe_amd call"):
    if rc != 0:
        raise TfheAmdError(f"{what} failed: {lib().tfhe_amd_last_error().decode()}")


def _p(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(i32p)


def set_seed(seed):
    lib().tfhe_random_generator_setSeed((ctypes.c_uint32 * 1)(seed), 1)


class Params:
    def __init__(self, ptr):
        self.ptr = ptr
        p = ptr.contents
        g = p.tgsw_params.contents
        t = g.tlwe_params.contents
        self.n = p.in_out_params.contents.n
        self.N, self.k, self.l, self.Bgbit = t.N, t.k, g.l, g.Bgbit
        self.ks_t, self.ks_basebit = p.ks_t, p.ks_basebit
        self.kpl = g.kpl


def default_params(minimum_lambda=110):
    return Params(lib().new_default_gate_bootstrapping_parameters(minimum_lambda))


def big_params():
    """N = 2048 set: n = 750, l = 2, Bgbit = 10, stdevs 2^-17 / 2^-44, keyswitch t = 8, basebit = 2."""
    return Params(lib().new_gate_bootstrapping_parameters(750, 2048, 1, 2, 10, 8, 2, 2.0 ** -17, 2.0 ** -44,
                                                          0.012467))


class SecretKey:
    """Wraps a TFheGateBootstrappingSecretKeySet (host keys; device key uploaded on first use)."""

    def __init__(self, params, seed=None, ptr=None):
        if ptr is None:
            if seed is not None:
                set_seed(seed)
            ptr = lib().new_random_gate_bootstrapping_secret_keyset(params.ptr)
        self.params = params
        self.ptr = ptr
        self.cloud = lib().tfhe_amd_cloud_of(self.ptr)

    @classmethod
    def generate_on_device(cls, params, seed):
        """Key set generated on the device (Philox4x32-10 keyed by seed)."""
        ptr = lib().tfhe_amd_keygen_device(params.ptr, seed)
        if not ptr:
            raise TfheAmdError(lib().tfhe_amd_last_error().decode())
        return cls(params, ptr=ptr)

    @classmethod
    def load(cls, path):
        """Secret key set from a key file in the reference's format."""
        with _CFile(path, "rb") as f:
            ptr = lib().new_tfheGateBootstrappingSecretKeySet_fromFile(f)
        if not ptr:
            raise TfheAmdError(lib().tfhe_amd_last_error().decode())
        return cls(Params(ptr.contents.params), ptr=ptr)

    def save(self, path):
        with _CFile(path, "wb") as f:
            lib().export_tfheGateBootstrappingSecretKeySet_toFile(f, self.ptr)

    def encrypt(self, bits):
        bits = np.ascontiguousarray(bits, np.int32)
        out = np.zeros((len(bits), self.params.n + 1), np.int32)
        check(lib().tfhe_amd_encrypt_bits(self.ptr, len(bits), _p(bits), _p(out)), "encrypt")
        return out

    def phases(self, cts):
        cts = np.ascontiguousarray(cts, np.int32)
        ph = np.zeros(len(cts), np.int32)
        check(lib().tfhe_amd_phases(self.ptr, len(cts), _p(cts), _p(ph)), "phases")
        return ph

    def decrypt(self, cts):
        return (self.phases(cts) > 0).astype(np.int32)

    def export(self):
        P = self.params
        arrs = dict(lwe_key=np.zeros(P.n, np.int32), tlwe_key=np.zeros(P.k * P.N, np.int32),
                    bk=np.zeros((P.n, P.kpl, P.k + 1, P.N), np.int32),
                    ksk=np.zeros((P.k * P.N, P.ks_t, 1 << P.ks_basebit, P.n + 1), np.int32))
        check(lib().tfhe_amd_export_keys(self.ptr, _p(arrs["lwe_key"]), _p(arrs["tlwe_key"]), _p(arrs["bk"]),
                                         _p(arrs["ksk"])), "export_keys")
        return arrs

    def upload(self):
        check(lib().tfhe_amd_cloud_key_upload(self.cloud), "cloud key upload")

    def close(self):
        if self.ptr:
            lib().delete_gate_bootstrapping_secret_keyset(self.ptr)
            self.ptr = None


class _CFile:
    """A C FILE* for the key-file functions (opened with the C library's fopen)."""

    _libc = None

    def __init__(self, path, mode):
        if _CFile._libc is None:
            c = ctypes.CDLL(None)
            c.fopen.restype, c.fopen.argtypes = vp, [ctypes.c_char_p, ctypes.c_char_p]
            c.fclose.restype, c.fclose.argtypes = ctypes.c_int, [vp]
            _CFile._libc = c
        self.f = _CFile._libc.fopen(os.fsencode(path), mode.encode())
        if not self.f:
            raise OSError(f"cannot open {path}")

    def __enter__(self):
        return self.f

    def __exit__(self, *exc):
        _CFile._libc.fclose(self.f)


def save_params(params, path):
    with _CFile(path, "wb") as f:
        lib().export_tfheGateBootstrappingParameterSet_toFile(f, params.ptr)


def load_params(path):
    with _CFile(path, "rb") as f:
        ptr = lib().new_tfheGateBootstrappingParameterSet_fromFile(f)
    if not ptr:
        raise TfheAmdError(lib().tfhe_amd_last_error().decode())
    return Params(ptr)


def save_cloud_key(cloud, path):
    with _CFile(path, "wb") as f:
        lib().export_tfheGateBootstrappingCloudKeySet_toFile(f, cloud)


def load_cloud_key(path):
    with _CFile(path, "rb") as f:
        ck = lib().new_tfheGateBootstrappingCloudKeySet_fromFile(f)
    if not ck:
        raise TfheAmdError(lib().tfhe_amd_last_error().decode())
    return ck


def save_ciphertexts(params, cts, path, variances=None):
    """Ciphertext rows [m][n+1] as consecutive gate-bootstrapping ciphertext
    records (export_gate_bootstrapping_ciphertext_toFile)."""
    cts = np.ascontiguousarray(cts, np.int32)
    n = params.n
    s = lib().new_gate_bootstrapping_ciphertext(params.ptr)
    try:
        with _CFile(path, "wb") as f:
            for g in range(len(cts)):
                ctypes.memmove(s.contents.a, cts[g].ctypes.data, 4 * n)
                s.contents.b = int(cts[g, n])
                s.contents.current_variance = 0.0 if variances is None else float(variances[g])
                lib().export_gate_bootstrapping_ciphertext_toFile(f, s, params.ptr)
    finally:
        lib().delete_gate_bootstrapping_ciphertext(s)


def load_ciphertexts(params, path, count):
    """(rows [count][n+1], variances) read with import_gate_bootstrapping_ciphertext_fromFile."""
    n = params.n
    out = np.zeros((count, n + 1), np.int32)
    var = np.zeros(count)
    s = lib().new_gate_bootstrapping_ciphertext(params.ptr)
    try:
        with _CFile(path, "rb") as f:
            for g in range(count):
                lib().import_gate_bootstrapping_ciphertext_fromFile(f, s, params.ptr)
                ctypes.memmove(out[g].ctypes.data, s.contents.a, 4 * n)
                out[g, n] = s.contents.b
                var[g] = s.contents.current_variance
    finally:
        lib().delete_gate_bootstrapping_ciphertext(s)
    return out, var


def cloud_from_arrays(params, bk, ksk):
    bk = np.ascontiguousarray(bk, np.int32)
    ksk = np.ascontiguousarray(ksk, np.int32)
    ck = lib().new_gate_bootstrapping_cloud_keyset_from_arrays(params.ptr, _p(bk), _p(ksk))
    if not ck:
        raise TfheAmdError(lib().tfhe_amd_last_error().decode())
    return ck


class DeviceArray:
    """int32 device buffer allocated through the library."""

    def __init__(self, shape):
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (shape,)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = lib().tfhe_amd_malloc(max(self.nbytes, 4))
        if not self.ptr:
            raise TfheAmdError(lib().tfhe_amd_last_error().decode())

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, np.int32)
        d = cls(a.shape)
        check(lib().tfhe_amd_memcpy_h2d(d.ptr, a.ctypes.data_as(vp), d.nbytes), "h2d")
        return d

    def numpy(self):
        a = np.zeros(self.shape, np.int32)
        check(lib().tfhe_amd_memcpy_d2h(a.ctypes.data_as(vp), self.ptr, self.nbytes), "d2h")
        return a

    def free(self):
        if self.ptr:
            lib().tfhe_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gates_device(cloud, ops, in0, in1, in2, out, pool, stream=None):
    """Asynchronous device gates over a device pool (all DeviceArray)."""
    count = ops.shape[0]
    check(lib().tfhe_amd_gates_device(cloud, count, ops.ptr, in0.ptr, in1.ptr, in2.ptr, out.ptr, pool.ptr, stream),
          "gates_device")


NEXT_INDEX = (1 << 64) - 1  # TFHE_AMD_NEXT_INDEX: the key's next unused ciphertext indices


def encrypt_device(sk, d_bits, d_cts, seed, first_index=None, count=None, stream=None):
    """Device encryption of count bits (DeviceArray) into ciphertext rows d_cts.

    first_index=None takes the key's next unused ciphertext indices, so calls
    never repeat a (seed, index) pair; pass an explicit index only to reproduce
    earlier ciphertexts."""
    count = d_bits.shape[0] if count is None else count
    w = sk.params.n + 1
    if count < 0 or count > int(np.prod(d_bits.shape)):
        raise ValueError(f"encrypt_device: count {count} exceeds the {int(np.prod(d_bits.shape))} bits given")
    if count * w > int(np.prod(d_cts.shape)):
        raise ValueError(f"encrypt_device: d_cts {d_cts.shape} holds fewer than {count} rows of n+1 = {w} words")
    idx = NEXT_INDEX if first_index is None else int(first_index)
    check(lib().tfhe_amd_encrypt_device(sk.ptr, count, d_bits.ptr, d_cts.ptr, seed, idx, stream),
          "encrypt_device")


def decrypt_device(sk, d_cts, d_bits=None, d_phases=None, count=None, stream=None, offset_rows=0):
    """Device decryption of count ciphertext rows starting at row offset_rows of d_cts."""
    w = sk.params.n + 1
    rows = int(np.prod(d_cts.shape)) // w
    count = rows - offset_rows if count is None else count
    if count < 0 or offset_rows < 0 or offset_rows + count > rows:
        raise ValueError(f"decrypt_device: rows [{offset_rows}, {offset_rows + count}) outside d_cts ({rows} rows)")
    for d, name in ((d_bits, "d_bits"), (d_phases, "d_phases")):
        if d is not None and count > int(np.prod(d.shape)):
            raise ValueError(f"decrypt_device: {name} holds fewer than {count} entries")
    src = d_cts.ptr + 4 * offset_rows * w
    check(lib().tfhe_amd_decrypt_device(sk.ptr, count, src, d_bits.ptr if d_bits else None,
                                        d_phases.ptr if d_phases else None, stream), "decrypt_device")


AUTOTUNE_SIZES = (64, 128, 256, 384, 512, 640, 768, 896, 1024, 1280, 1536, 2048, 3072, 4096)
AUTOTUNE_MODES = ("digit_parallel", "two_wave_split", "one_per_wavefront_batched")


AUTOTUNE_MODES_N2048 = ("pair_four_wave", "two_wave")


def autotune(cloud):
    """Measures the blind-rotation kernels of this key's ring size over
    AUTOTUNE_SIZES and sets the automatic mode's crossovers. Returns
    (latency_max, small_batch_max, {mode name: {size: ms}}); for an N = 2048
    key small_batch_max is the pair kernel's crossover, latency_max 0, and the
    table has AUTOTUNE_MODES_N2048."""
    lat, small = ctypes.c_int32(), ctypes.c_int32()
    ms = (ctypes.c_float * (3 * len(AUTOTUNE_SIZES)))()
    check(lib().tfhe_amd_autotune(cloud, ctypes.byref(lat), ctypes.byref(small), ms), "autotune")
    big = cloud.contents.params.contents.tgsw_params.contents.tlwe_params.contents.N == 2048
    names = AUTOTUNE_MODES_N2048 if big else AUTOTUNE_MODES
    table = {name: {sz: ms[m * len(AUTOTUNE_SIZES) + i] for i, sz in enumerate(AUTOTUNE_SIZES)}
             for m, name in enumerate(names)}
    return lat.value, small.value, table


def kernel_for(cloud, count):
    """The blind-rotation kernel (mode) a batch of `count` gates runs with now."""
    m = lib().tfhe_amd_kernel_for(cloud, int(count))
    if m < 0:
        check(m, "kernel_for")
    return m


def set_launch_config(mode, latency_max=-1, small_batch_max=-1):
    lib().tfhe_amd_set_launch_config(int(mode), int(latency_max), int(small_batch_max))


def get_launch_config():
    m, lat, sb = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lib().tfhe_amd_get_launch_config(ctypes.byref(m), ctypes.byref(lat), ctypes.byref(sb))
    return m.value, lat.value, sb.value


def profile_enable(on=True):
    lib().tfhe_amd_profile_enable(1 if on else 0)


def profile_read():
    """(blind-rotation ms, keyswitch ms, batches) summed since the last read."""
    br, ks, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
    check(lib().tfhe_amd_profile_read(ctypes.byref(br), ctypes.byref(ks), ctypes.byref(n)), "profile_read")
    return br.value, ks.value, n.value


def synchronize():
    check(lib().tfhe_amd_synchronize(), "synchronize")


def run_gates(cloud, ops, a, b, c=None):
    """Evaluate gates on host ciphertext arrays [count][n+1] through the device
    API (one pool upload, one launch sequence, one download)."""
    ops = np.ascontiguousarray(ops, np.int32)
    count = len(ops)
    w = a.shape[1]
    c = a if c is None else c
    pool = np.concatenate([a, b, c, np.zeros((count, w), np.int32)]).astype(np.int32)
    idx = np.arange(count, dtype=np.int32)
    d_pool = DeviceArray.from_numpy(pool)
    d_ops = DeviceArray.from_numpy(ops)
    d_i0 = DeviceArray.from_numpy(idx)
    d_i1 = DeviceArray.from_numpy(idx + count)
    d_i2 = DeviceArray.from_numpy(idx + 2 * count)
    d_o = DeviceArray.from_numpy(idx + 3 * count)
    gates_device(cloud, d_ops, d_i0, d_i1, d_i2, d_o, d_pool)
    synchronize()
    out = d_pool.numpy()[3 * count:]
    for d in (d_pool, d_ops, d_i0, d_i1, d_i2, d_o):
        d.free()
    return out


def run_gates_woks(cloud, params, ops, a, b, c=None):
    """Bootstrap without keyswitch: extracted samples [count][k*N+1]."""
    ops = np.ascontiguousarray(ops, np.int32)
    count = len(ops)
    c = a if c is None else c
    pool = np.concatenate([a, b, c]).astype(np.int32)
    idx = np.arange(count, dtype=np.int32)
    d_pool = DeviceArray.from_numpy(pool)
    d_ops = DeviceArray.from_numpy(ops)
    d_i0 = DeviceArray.from_numpy(idx)
    d_i1 = DeviceArray.from_numpy(idx + count)
    d_i2 = DeviceArray.from_numpy(idx + 2 * count)
    d_ext = DeviceArray((count, params.k * params.N + 1))
    check(lib().tfhe_amd_gates_device_woks(cloud, count, d_ops.ptr, d_i0.ptr, d_i1.ptr, d_i2.ptr, d_pool.ptr,
                                           d_ext.ptr, None), "gates_device_woks")
    synchronize()
    return d_ext.numpy()


def keyswitch(cloud, params, ext):
    ext = np.ascontiguousarray(ext, np.int32)
    count = ext.shape[0]
    d_in = DeviceArray.from_numpy(ext)
    d_out = DeviceArray((count, params.n + 1))
    check(lib().tfhe_amd_keyswitch_device(cloud, count, d_in.ptr, d_out.ptr, None), "keyswitch_device")
    synchronize()
    return d_out.numpy()


class HostCiphertexts:
    """An LweSample array allocated by the library, with numpy views."""

    def __init__(self, params, count):
        self.params, self.count = params, count
        self.ptr = lib().new_gate_bootstrapping_ciphertext_array(count, params.ptr)

    def item(self, i):
        return ctypes.pointer(self.ptr[i])

    def set(self, arr):
        n = self.params.n
        for i in range(self.count):
            s = self.ptr[i]
            ctypes.memmove(s.a, arr[i].ctypes.data, 4 * n)
            s.b = int(arr[i][n])

    def get(self):
        n = self.params.n
        out = np.zeros((self.count, n + 1), np.int32)
        for i in range(self.count):
            s = self.ptr[i]
            ctypes.memmove(out[i].ctypes.data, s.a, 4 * n)
            out[i][n] = s.b
        return out

    def close(self):
        if self.ptr:
            lib().delete_gate_bootstrapping_ciphertext_array(self.count, self.ptr)
            self.ptr = None
