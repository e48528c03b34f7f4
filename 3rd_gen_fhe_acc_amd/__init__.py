"""MI355X gate-bootstrapping library: Python binding of libfheacc.so.

The compute path is the HIP library (lib/libfheacc.so, built by build_lib.py);
this module is a thin ctypes layer over its C ABI (include/fheacc.h).  There is
no CPU fallback: if the library is missing, loading fails with an error.
Device buffers are torch tensors (torch is used for device memory and streams
only).

Names follow the reference host API (FHE_ACC_runtime/include/{LWE,RLWE}.h):
gates OR, AND, NOR, NAND, XOR, XNOR (params.h:107); LWE ciphertexts are
[n+1] words (a then b); RLWE ciphertexts [2][N] (a then b, NTT domain).
"""
import ctypes as C
import os

import numpy as np

try:
    # torch bundles its own HIP runtime; loading it before libfheacc.so makes
    # both resolve to that one runtime in this process.
    import torch  # noqa: F401
except ImportError:  # the C ABI works without torch (host-pointer calls)
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
# FHEACC_LIBRARY=checked selects the build with device-side bounds checks
# (lib/libfheacc_checked.so); anything else, or unset, the default library
CHECKED = os.environ.get("FHEACC_LIBRARY", "") == "checked"
LIB_PATH = os.path.join(HERE, "lib", "libfheacc_checked.so" if CHECKED else "libfheacc.so")

GATES = {"OR": 0, "AND": 1, "NOR": 2, "NAND": 3, "XOR": 4, "XNOR": 5}
# blind-rotation orders (include/fheacc.h FHEACC_EVAL_*)
EVAL_SOFTWARE, EVAL_FPGA_X1, EVAL_FPGA_XN = 0, 1, 2


class FheAccError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [("n", C.c_int32), ("q", C.c_uint64), ("N", C.c_int32), ("Q", C.c_uint64),
                ("nbit", C.c_int32), ("BGbits", C.c_uint32), ("Br", C.c_uint64), ("Bks", C.c_uint64),
                ("Bks_R_bits", C.c_uint32), ("stddev", C.c_double), ("E_rlwe", C.c_uint64)]


class _Info(C.Structure):
    _fields_ = [("Q", C.c_uint64), ("iN", C.c_uint64), ("ROU", C.c_uint64), ("digitG", C.c_uint32),
                ("digitR", C.c_uint32), ("digitKS", C.c_uint32), ("digitKS_R", C.c_uint32),
                ("narrow", C.c_uint32), ("harvey", C.c_uint32), ("gate_constant", C.c_uint64 * 6),
                ("bsk_words", C.c_uint64),
                ("ksk_words", C.c_uint64), ("subs_words", C.c_uint64), ("bsk_device_bytes", C.c_uint64),
                ("ksk_device_bytes", C.c_uint64), ("subs_device_bytes", C.c_uint64),
                ("lazy_inverse_stages", C.c_uint32), ("reserved", C.c_uint32)]


class _Capture(C.Structure):
    _fields_ = [("acc", C.c_void_p), ("ks", C.c_void_p)]


class _ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_int32), ("max_inflight", C.c_int32),
                ("min_fill", C.c_int32), ("fair_share", C.c_int32)]


class _ServerStats(C.Structure):
    _fields_ = [("jobs_submitted", C.c_int64), ("jobs_done", C.c_int64), ("jobs_failed", C.c_int64),
                ("jobs_queued", C.c_int64), ("jobs_active", C.c_int64), ("bootstraps", C.c_int64),
                ("batches", C.c_int64), ("max_batch_seen", C.c_int64), ("wall_s", C.c_double),
                ("gpu_busy_s", C.c_double), ("gate_ms_p50", C.c_double), ("gate_ms_p99", C.c_double),
                ("job_ms_p50", C.c_double), ("job_ms_p99", C.c_double), ("pool_bytes", C.c_uint64),
                ("peak_bytes", C.c_uint64)]


EXPORTS = {
    "fheacc_ctx_create": [C.POINTER(_Params), C.c_int, C.POINTER(C.c_void_p)],
    "fheacc_ctx_destroy": [C.c_void_p],
    "fheacc_ctx_info": [C.c_void_p, C.POINTER(_Info)],
    "fheacc_ctx_set_stream": [C.c_void_p, C.c_void_p],
    "fheacc_sync": [C.c_void_p],
    "fheacc_last_error": [],
    "fheacc_device_checks_enabled": [],
    "fheacc_device_check_read": [C.c_int, C.POINTER(C.c_uint32)],
    "fheacc_ctx_profile": [C.c_void_p, C.c_int],
    "fheacc_ctx_kernel_ms": [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "fheacc_ctx_set_tuning": [C.c_void_p, C.c_int, C.c_int],
    "fheacc_ctx_get_tuning": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "fheacc_ctx_set_schedule": [C.c_void_p, C.c_int, C.c_int],
    "fheacc_ctx_get_schedule": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "fheacc_ctx_set_eval_mode": [C.c_void_p, C.c_int],
    "fheacc_ctx_get_eval_mode": [C.c_void_p, C.POINTER(C.c_int)],
    "fheacc_autotune": [C.c_void_p, C.c_void_p, C.c_int64, C.c_char_p, C.c_size_t],
    "fheacc_keys_create": [C.c_void_p, C.POINTER(C.c_void_p)],
    "fheacc_keys_destroy": [C.c_void_p],
    "fheacc_keys_upload_bsk": [C.c_void_p, C.c_void_p],
    "fheacc_keys_upload_ksk": [C.c_void_p, C.c_void_p],
    "fheacc_keys_upload_subs": [C.c_void_p, C.c_void_p],
    "fheacc_keys_download_bsk": [C.c_void_p, C.c_void_p],
    "fheacc_keys_download_ksk": [C.c_void_p, C.c_void_p],
    "fheacc_keys_download_subs": [C.c_void_p, C.c_void_p],
    "fheacc_keys_export_fpga_bsk": [C.c_void_p, C.c_void_p],
    "fheacc_keys_import_fpga_bsk": [C.c_void_p, C.c_void_p],
    "fheacc_keys_export_fpga_subs": [C.c_void_p, C.c_void_p],
    "fheacc_keys_import_fpga_subs": [C.c_void_p, C.c_void_p],
    "fheacc_keys_generate": [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p],
    "fheacc_keys_generate_subs": [C.c_void_p, C.c_uint64, C.c_void_p],
    "fheacc_keys_digest": [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "fheacc_keys_save": [C.c_void_p, C.c_char_p],
    "fheacc_keys_load": [C.c_void_p, C.c_char_p],
    "fheacc_lwe_secret": [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p],
    "fheacc_rlwe_secret": [C.c_void_p, C.c_uint64, C.c_void_p],
    "fheacc_lwe_encrypt": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64, C.c_uint64, C.c_void_p],
    "fheacc_lwe_decrypt": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
    "fheacc_bootstrap_batch": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_int64, C.POINTER(_Capture)],
    "fheacc_bootstrap_batch_dev": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int64, C.POINTER(_Capture)],
    "fheacc_bootstrap_batch_dev_noalloc": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_int64, C.POINTER(_Capture)],
    "fheacc_graph_capture_bootstrap": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.POINTER(C.c_void_p)],
    "fheacc_graph_capture_mixed": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)],
    "fheacc_graph_capture_interleaved": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)],
    "fheacc_graph_launch": [C.c_void_p],
    "fheacc_graph_destroy": [C.c_void_p],
    "fheacc_ntt_dev": [C.c_void_p, C.c_void_p, C.c_int64, C.c_int],
    "fheacc_external_product_dev": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64],
    "fheacc_subs_keyswitch_dev": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64],
    "fheacc_expand_rlwe_dev": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64],
    "fheacc_cmux_dev": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64],
    "fheacc_blind_rotate_rlwe_dev": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int64],
    "fheacc_blind_rotate_full_dev": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                     C.c_int64],
    "fheacc_expand_rlwe": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64],
    "fheacc_subs_keyswitch": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64],
    "fheacc_homo_expand_dev": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64],
    "fheacc_homo_expand": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64],
    "fheacc_cmux_tree_dev": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64],
    "fheacc_cmux_tree": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64],
    "fheacc_rlwe_encrypt": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64, C.c_uint64, C.c_void_p],
    "fheacc_rgsw_encrypt": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64, C.c_uint64, C.c_void_p],
    "fheacc_rlwe_decrypt": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64, C.c_void_p],
    "fheacc_external_product": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64],
    "fheacc_cmux": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64],
    "fheacc_blind_rotate_rlwe": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int64],
    "fheacc_blind_rotate_full": [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                 C.c_int64],
    "fheacc_server_create": [C.c_void_p, C.POINTER(_ServerConfig), C.POINTER(C.c_void_p)],
    "fheacc_server_destroy": [C.c_void_p],
    "fheacc_server_add_client": [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)],
    "fheacc_server_submit": [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                             C.c_int64, C.c_void_p, C.POINTER(C.c_int64)],
    "fheacc_server_wait": [C.c_void_p, C.c_int64, C.POINTER(C.c_double)],
    "fheacc_server_get_stats": [C.c_void_p, C.POINTER(_ServerStats)],
    "fheacc_server_reset_stats": [C.c_void_p],
}

_lib = None


def lib():
    """Load libfheacc.so; raise if it was not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FheAccError(os.path.basename(LIB_PATH) + " not built: run __graft_entry__.build() or "
                              "python 3rd_gen_fhe_acc_amd/build_lib.py")
        L = C.CDLL(LIB_PATH)
        for name, args in EXPORTS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "fheacc_last_error" else C.c_int
        _lib = L
    return _lib


# bits of device_check_read (include/fheacc.h FHEACC_DEVCHECK_*)
DEVCHECK_LDS, DEVCHECK_CIPHERTEXT = 1, 2


def device_checks_enabled():
    """True when the loaded library is the checked build."""
    return bool(lib().fheacc_device_checks_enabled())


def device_check_read(device=0):
    """Synchronize `device`; return and clear the device-check bits set since the last read."""
    bits = C.c_uint32()
    _check(lib().fheacc_device_check_read(device, C.byref(bits)))
    return bits.value


def _check(rc):
    if rc != 0:
        raise FheAccError("fheacc error %d: %s" % (rc, lib().fheacc_last_error().decode()))


def _np(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(C.c_void_p)


def _dptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """One parameter set on one GPU (the reference's global parameter block)."""

    def __init__(self, n, q, N, Q=0, nbit=27, BGbits=7, Br=32, Bks=32, Bks_R_bits=9, stddev=3.19,
                 E_rlwe=18, device=0, use_torch_stream=True):
        self._p = _Params(n, q, N, Q, nbit, BGbits, Br, Bks, Bks_R_bits, stddev, E_rlwe)
        h = C.c_void_p()
        _check(lib().fheacc_ctx_create(C.byref(self._p), device, C.byref(h)))
        self.h = h
        self.device = device
        info = _Info()
        _check(lib().fheacc_ctx_info(self.h, C.byref(info)))
        self.info = {f: getattr(info, f) for f, _ in _Info._fields_ if f not in ("gate_constant", "reserved")}
        self.info["gate_constant"] = list(info.gate_constant)
        self.n, self.q, self.N, self.Q = n, q, N, info.Q
        self.digitG, self.digitR, self.digitKS, self.digitKS_R = info.digitG, info.digitR, info.digitKS, info.digitKS_R
        self.narrow = bool(info.narrow)
        if use_torch_stream:
            import torch
            torch.cuda.set_device(device)
            self.set_stream(torch.cuda.current_stream().cuda_stream)

    def close(self):
        if getattr(self, "h", None):
            lib().fheacc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        _check(lib().fheacc_ctx_set_stream(self.h, C.c_void_p(stream) if stream else None))

    def sync(self):
        _check(lib().fheacc_sync(self.h))

    KERNEL_CLASSES = ("blind_rotate", "lwe_keyswitch", "external_product", "ntt")

    def profile(self, enable=True):
        """Start (and clear) or stop per-kernel HIP-event timing."""
        _check(lib().fheacc_ctx_profile(self.h, 1 if enable else 0))

    def kernel_ms(self):
        """{kernel class: (total ms, launches)} since profiling started."""
        ms = (C.c_double * 4)()
        n = (C.c_int64 * 4)()
        _check(lib().fheacc_ctx_kernel_ms(self.h, ms, n))
        return {k: (ms[i], n[i]) for i, k in enumerate(self.KERNEL_CLASSES)}

    def set_tuning(self, threads, polys_per_batch=0):
        _check(lib().fheacc_ctx_set_tuning(self.h, threads, polys_per_batch))

    def get_tuning(self):
        t, p = C.c_int(), C.c_int()
        _check(lib().fheacc_ctx_get_tuning(self.h, C.byref(t), C.byref(p)))
        return t.value, p.value

    def set_schedule(self, persist=0, sync_window=0):
        """Blind-rotation schedule: persist workgroups per CU (0: one per
        ciphertext), lock-step window in steps (0: free running)."""
        _check(lib().fheacc_ctx_set_schedule(self.h, int(persist), int(sync_window)))

    def get_schedule(self):
        p, w = C.c_int(), C.c_int()
        _check(lib().fheacc_ctx_get_schedule(self.h, C.byref(p), C.byref(w)))
        return p.value, w.value

    @property
    def eval_mode(self):
        """Blind-rotation order of the gate bootstrap: EVAL_SOFTWARE (residue-0
        steps skipped, RLWE.cpp:754-767), EVAL_FPGA_X1 (step (0,0) always
        multiplies, myfpga.cpp:1030-1071) or EVAL_FPGA_XN (every step
        multiplies, btkey[0] = RGSW(1) for residue 0, myfpga.cpp:1180-1225)."""
        m = C.c_int()
        _check(lib().fheacc_ctx_get_eval_mode(self.h, C.byref(m)))
        return m.value

    @eval_mode.setter
    def eval_mode(self, mode):
        _check(lib().fheacc_ctx_set_eval_mode(self.h, int(mode)))

    def autotune(self, keys, batch):
        """Time every blind-rotation launch shape on `batch` synthetic
        ciphertexts, keep the fastest; returns the tuner's report."""
        import json

        buf = C.create_string_buffer(16384)
        _check(lib().fheacc_autotune(self.h, keys.h, batch, buf, len(buf)))
        return json.loads(buf.value.decode())

    # host API -------------------------------------------------------------
    def lwe_secret(self, seed, binary=True):
        out = np.zeros(self.n, dtype=np.uint64)
        _check(lib().fheacc_lwe_secret(self.h, seed, 1 if binary else 0, _np(out)[1]))
        return out

    def rlwe_secret(self, seed):
        out = np.zeros(self.N, dtype=np.uint64)
        _check(lib().fheacc_rlwe_secret(self.h, seed, _np(out)[1]))
        return out

    def lwe_encrypt(self, sk, bits, seed, first_index=0):
        bits = np.ascontiguousarray(np.asarray(bits, dtype=np.uint8).reshape(-1))
        out = np.zeros((bits.size, self.n + 1), dtype=np.uint64)
        sk, psk = _np(sk)
        _check(lib().fheacc_lwe_encrypt(self.h, psk, bits.ctypes.data_as(C.c_void_p), bits.size, seed,
                                        first_index, _np(out)[1]))
        return out

    def lwe_decrypt(self, sk, c):
        c = np.ascontiguousarray(c, dtype=np.uint64).reshape(-1, self.n + 1)
        out = np.zeros(c.shape[0], dtype=np.uint8)
        sk, psk = _np(sk)
        _check(lib().fheacc_lwe_decrypt(self.h, psk, _np(c)[1], c.shape[0], out.ctypes.data_as(C.c_void_p)))
        return out

    def keys(self):
        return Keys(self)

    # gate path ------------------------------------------------------------
    def bootstrap(self, keys, gates, c1, c2, capture=False):
        """eval_bootstrap over host arrays; returns out (and acc, ks captures)."""
        c1 = np.ascontiguousarray(c1, dtype=np.uint64).reshape(-1, self.n + 1)
        c2 = np.ascontiguousarray(c2, dtype=np.uint64).reshape(-1, self.n + 1)
        B = c1.shape[0]
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gates, dtype=np.int32), (B,)))
        out = np.zeros((B, self.n + 1), dtype=np.uint64)
        cap = None
        acc = ks = None
        if capture:
            acc = np.zeros((B, 2, self.N), dtype=np.uint64)
            ks = np.zeros((B, self.n + 1), dtype=np.uint64)
            cap = _Capture(acc.ctypes.data, ks.ctypes.data)
        _check(lib().fheacc_bootstrap_batch(self.h, keys.h, g.ctypes.data_as(C.c_void_p), _np(c1)[1],
                                            _np(c2)[1], _np(out)[1], B, C.byref(cap) if cap else None))
        return (out, acc, ks) if capture else out

    def bootstrap_dev(self, keys, gates, c1, c2, out, cap_acc=None, cap_ks=None):
        """Asynchronous on the context stream; arguments are device tensors."""
        B = c1.shape[0]
        cap = _Capture(cap_acc.data_ptr() if cap_acc is not None else None,
                       cap_ks.data_ptr() if cap_ks is not None else None)
        _check(lib().fheacc_bootstrap_batch_dev(self.h, keys.h, _dptr(gates), _dptr(c1), _dptr(c2), _dptr(out),
                                                B, C.byref(cap)))

    def bootstrap_dev_noalloc(self, keys, gates, c1, c2, out, cap_acc=None, cap_ks=None):
        """bootstrap_dev with its scratch in the library's static device array:
        no device allocation (fheacc_bootstrap_batch_dev_noalloc)."""
        B = c1.shape[0]
        cap = _Capture(cap_acc.data_ptr() if cap_acc is not None else None,
                       cap_ks.data_ptr() if cap_ks is not None else None)
        _check(lib().fheacc_bootstrap_batch_dev_noalloc(self.h, keys.h, _dptr(gates), _dptr(c1), _dptr(c2),
                                                        _dptr(out), B, C.byref(cap)))

    def capture_bootstrap(self, keys, gates, c1, c2, out):
        """Capture the gate path over fixed device buffers into a HIP graph."""
        return BootstrapGraph(self, keys, gates, c1, c2, out)

    def ntt_dev(self, polys, inverse=False):
        _check(lib().fheacc_ntt_dev(self.h, _dptr(polys), polys.numel() // self.N, 1 if inverse else 0))

    def external_product_dev(self, acc, rgsw, out, per_ciphertext_key=False):
        count = acc.numel() // (2 * self.N)
        stride = self.digitG * 4 * self.N if per_ciphertext_key else 0
        _check(lib().fheacc_external_product_dev(self.h, _dptr(acc), _dptr(rgsw), stride, _dptr(out), count))

    def expand_rlwe_dev(self, keys, rlwe, out):
        """expand_RLWE (app.cpp:26-97): rlwe [B][2][N] -> out [B][N][2][N]."""
        B = rlwe.numel() // (2 * self.N)
        _check(lib().fheacc_expand_rlwe_dev(self.h, keys.h, _dptr(rlwe), _dptr(out), B))

    def cmux_dev(self, c0, c1, rgsw, out, per_ciphertext_key=False):
        count = c0.numel() // (2 * self.N)
        stride = self.digitG * 4 * self.N if per_ciphertext_key else 0
        _check(lib().fheacc_cmux_dev(self.h, _dptr(c0), _dptr(c1), _dptr(rgsw), stride, _dptr(out), count))

    def blind_rotate_rlwe_dev(self, rlwe, forward, rgsws, out):
        count = rlwe.numel() // (2 * self.N)
        steps = rgsws.numel() // (self.digitG * 4 * self.N)
        _check(lib().fheacc_blind_rotate_rlwe_dev(self.h, _dptr(rlwe), 1 if forward else 0, _dptr(rgsws), steps,
                                                  _dptr(out), count))

    def blind_rotate_full_dev(self, rlwe, b, a, rgsws, out):
        count = rlwe.numel() // (2 * self.N)
        a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
        _check(lib().fheacc_blind_rotate_full_dev(self.h, _dptr(rlwe), int(b), a.ctypes.data_as(C.c_void_p),
                                                  _dptr(rgsws) if a.size else None, a.size, _dptr(out), count))

    def expand_rlwe(self, keys, rlwe):
        """Host arrays: rlwe [B][2][N] -> [B][N][2][N] (fheacc_expand_rlwe)."""
        rlwe, p = _np(np.asarray(rlwe).reshape(-1, 2, self.N))
        out = np.zeros((rlwe.shape[0], self.N, 2, self.N), dtype=np.uint64)
        _check(lib().fheacc_expand_rlwe(self.h, keys.h, p, _np(out)[1], rlwe.shape[0]))
        return out

    def cmux(self, c0, c1, rgsw):
        c0, p0 = _np(np.asarray(c0).reshape(-1, 2, self.N))
        c1, p1 = _np(np.asarray(c1).reshape(-1, 2, self.N))
        rgsw, pg = _np(rgsw)
        out = np.zeros(c0.shape, dtype=np.uint64)
        stride = 0 if rgsw.size == self.digitG * 4 * self.N else self.digitG * 4 * self.N
        _check(lib().fheacc_cmux(self.h, p0, p1, pg, stride, _np(out)[1], c0.shape[0]))
        return out

    # RLWE / RGSW encryption and RLWE decryption (host arrays, device work)
    def rlwe_encrypt(self, sk, msgs, seed, first_index=0):
        """RLWE_encrypt (RLWE.cpp:324-344) of each row of msgs (time domain)."""
        msgs, pm = _np(np.asarray(msgs).reshape(-1, self.N))
        out = np.zeros((msgs.shape[0], 2, self.N), dtype=np.uint64)
        _check(lib().fheacc_rlwe_encrypt(self.h, _np(sk)[1], pm, msgs.shape[0], seed, first_index, _np(out)[1]))
        return out

    def rgsw_encrypt(self, sk, msgs, seed, first_index=0):
        """RGSW encryption (RLWE.cpp:561-577) of each row of msgs (time domain);
        reference-layout RGSWs [count][digitG * 4 * N]."""
        msgs, pm = _np(np.asarray(msgs).reshape(-1, self.N))
        out = np.zeros((msgs.shape[0], self.digitG * 4 * self.N), dtype=np.uint64)
        _check(lib().fheacc_rgsw_encrypt(self.h, _np(sk)[1], pm, msgs.shape[0], seed, first_index, _np(out)[1]))
        return out

    def rlwe_decrypt(self, sk, c, t):
        """RLWE_decrypt(s, c, t) (RLWE.cpp:362-376): [count][N] in [0, t]."""
        c, pc = _np(np.asarray(c).reshape(-1, 2, self.N))
        out = np.zeros((c.shape[0], self.N), dtype=np.uint64)
        _check(lib().fheacc_rlwe_decrypt(self.h, _np(sk)[1], pc, c.shape[0], t, _np(out)[1]))
        return out

    def subs_keyswitch(self, keys, rlwe, sub_factor):
        """Host arrays: RLWE_substitute + RLWE_keyswitch of [count][2][N]."""
        rlwe, pr = _np(np.asarray(rlwe).reshape(-1, 2, self.N))
        out = np.zeros(rlwe.shape, dtype=np.uint64)
        _check(lib().fheacc_subs_keyswitch(self.h, keys.h, pr, sub_factor, _np(out)[1], rlwe.shape[0]))
        return out

    def homo_expand_dev(self, keys, packed, enc_sec, out):
        """homo_expand (app.cpp:78-97): packed [count][digitG][2][N] ->
        out [count][N][digitG * 4 * N] (device tensors)."""
        count = packed.numel() // (self.digitG * 2 * self.N)
        _check(lib().fheacc_homo_expand_dev(self.h, keys.h, _dptr(packed), _dptr(enc_sec), _dptr(out), count))

    def homo_expand(self, keys, packed, enc_sec):
        packed, pp = _np(np.asarray(packed).reshape(-1, self.digitG, 2, self.N))
        enc_sec, pe = _np(enc_sec)
        out = np.zeros((packed.shape[0], self.N, self.digitG * 4 * self.N), dtype=np.uint64)
        _check(lib().fheacc_homo_expand(self.h, keys.h, pp, pe, _np(out)[1], packed.shape[0]))
        return out

    def cmux_tree_dev(self, leaves, rgsws, out, levels):
        """CMUX tree (test.cpp:1576-1582): leaves [count][2^levels][2][N],
        rgsws [levels] reference-layout RGSWs -> out [count][2][N]."""
        count = leaves.numel() // ((2 << levels) * self.N)
        _check(lib().fheacc_cmux_tree_dev(self.h, _dptr(leaves), levels, _dptr(rgsws) if levels else None, 0,
                                          _dptr(out), count))

    def cmux_tree(self, leaves, rgsws):
        rgsws, pg = _np(np.asarray(rgsws).reshape(-1, self.digitG * 4 * self.N))
        levels = rgsws.shape[0]
        leaves, pl = _np(np.asarray(leaves).reshape(-1, 1 << levels, 2, self.N))
        out = np.zeros((leaves.shape[0], 2, self.N), dtype=np.uint64)
        _check(lib().fheacc_cmux_tree(self.h, pl, levels, pg, 0, _np(out)[1], leaves.shape[0]))
        return out

    def subs_keyswitch_dev(self, keys, rlwe, sub_factor, out):
        _check(lib().fheacc_subs_keyswitch_dev(self.h, keys.h, _dptr(rlwe), sub_factor, _dptr(out),
                                               rlwe.numel() // (2 * self.N)))


class BootstrapGraph:
    """A captured gate-bootstrap batch; launch() replays it on the context stream."""

    def __init__(self, ctx, keys, gates, c1, c2, out):
        self.ctx, self.keys = ctx, keys
        self._bufs = (gates, c1, c2, out)  # the graph reads and writes these
        h = C.c_void_p()
        _check(lib().fheacc_graph_capture_bootstrap(ctx.h, keys.h, _dptr(gates), _dptr(c1), _dptr(c2),
                                                    _dptr(out), c1.shape[0], C.byref(h)))
        self.h = h

    def launch(self):
        _check(lib().fheacc_graph_launch(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().fheacc_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MixedGraph(BootstrapGraph):
    """Several parameter sets' batches captured into one HIP graph, set
    after set; launch() replays them as one unit.
    sets: [(ctx, keys, gates, c1, c2, out)] with device tensors."""

    def __init__(self, sets):
        self.ctx, self.keys = sets[0][0], [s[1] for s in sets]
        self._bufs = [s[2:] for s in sets]
        k = len(sets)
        arr = lambda vals: (C.c_void_p * k)(*vals)
        ctxs = arr([s[0].h.value for s in sets])
        keys = arr([s[1].h.value for s in sets])
        gates, c1, c2, out = (arr([s[i].data_ptr() for s in sets]) for i in (2, 3, 4, 5))
        B = (C.c_int64 * k)(*[s[3].shape[0] for s in sets])
        h = C.c_void_p()
        _check(lib().fheacc_graph_capture_mixed(k, ctxs, keys, gates, c1, c2, out, B, C.byref(h)))
        self.h = h


class InterleavedGraph(BootstrapGraph):
    """A mixed batch as it arrives: rows of several parameter sets interleaved
    in one buffer (set_of[r] names row r's set), each set's kernels gathering
    their rows on the device.  contexts/keys: one per set; gates [total],
    c1/c2/out [total][row_words] device tensors."""

    def __init__(self, contexts, keys, set_of, gates, c1, c2, out):
        self.ctx, self.keys = contexts[0], list(keys)
        set_of = np.ascontiguousarray(np.asarray(set_of, dtype=np.int32).reshape(-1))
        self._bufs = (gates, c1, c2, out, set_of)
        k = len(contexts)
        ctxs = (C.c_void_p * k)(*[c.h.value for c in contexts])
        ks = (C.c_void_p * k)(*[x.h.value for x in keys])
        h = C.c_void_p()
        _check(lib().fheacc_graph_capture_interleaved(k, ctxs, ks, set_of.ctypes.data_as(C.c_void_p), _dptr(gates),
                                                      _dptr(c1), _dptr(c2), _dptr(out), set_of.size, c1.shape[1],
                                                      C.byref(h)))
        self.h = h


class Keys:
    """A client's device-resident keys (bootstrap, LWE key switch, subs)."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().fheacc_keys_create(ctx.h, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fheacc_keys_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_bsk(self, bsk):
        bsk, p = _np(bsk)
        _check(lib().fheacc_keys_upload_bsk(self.h, p))

    def upload_ksk(self, ksk):
        ksk, p = _np(ksk)
        _check(lib().fheacc_keys_upload_ksk(self.h, p))

    def upload_subs(self, subs):
        subs, p = _np(subs)
        _check(lib().fheacc_keys_upload_subs(self.h, p))

    def download_bsk(self):
        out = np.zeros(self.ctx.info["bsk_words"], dtype=np.uint64)
        _check(lib().fheacc_keys_download_bsk(self.h, _np(out)[1]))
        return out

    def download_ksk(self):
        out = np.zeros(self.ctx.info["ksk_words"], dtype=np.uint64)
        _check(lib().fheacc_keys_download_ksk(self.h, _np(out)[1]))
        return out

    def download_subs(self):
        out = np.zeros(self.ctx.info["subs_words"], dtype=np.uint64)
        _check(lib().fheacc_keys_download_subs(self.h, _np(out)[1]))
        return out

    # the accelerator's DDR key images (FPGA::dma_write_bootstrap_key /
    # dma_write_subs_key, myfpga.cpp:596-712)
    def export_fpga_bsk(self):
        out = np.zeros(self.ctx.info["bsk_words"], dtype=np.uint64)
        _check(lib().fheacc_keys_export_fpga_bsk(self.h, _np(out)[1]))
        return out

    def import_fpga_bsk(self, image):
        image, p = _np(image)
        if image.size != self.ctx.info["bsk_words"]:
            raise ValueError("bootstrap key image must hold %d words" % self.ctx.info["bsk_words"])
        _check(lib().fheacc_keys_import_fpga_bsk(self.h, p))

    def export_fpga_subs(self):
        out = np.zeros(self.ctx.info["subs_words"], dtype=np.uint64)
        _check(lib().fheacc_keys_export_fpga_subs(self.h, _np(out)[1]))
        return out

    def import_fpga_subs(self, image):
        image, p = _np(image)
        if image.size != self.ctx.info["subs_words"]:
            raise ValueError("subs key image must hold %d words" % self.ctx.info["subs_words"])
        _check(lib().fheacc_keys_import_fpga_subs(self.h, p))

    def generate(self, seed, lwe_sk, rlwe_sk):
        lwe_sk, p1 = _np(lwe_sk)
        rlwe_sk, p2 = _np(rlwe_sk)
        _check(lib().fheacc_keys_generate(self.h, seed, p1, p2))

    def generate_subs(self, seed, rlwe_sk):
        rlwe_sk, p = _np(rlwe_sk)
        _check(lib().fheacc_keys_generate_subs(self.h, seed, p))

    def save(self, path):
        """Checkpoint the resident keys to a file (fheacc_keys_save)."""
        _check(lib().fheacc_keys_save(self.h, os.fsencode(path)))

    def load(self, path):
        """Restore keys written by save() for this parameter set (fheacc_keys_load)."""
        _check(lib().fheacc_keys_load(self.h, os.fsencode(path)))

    def digest(self):
        a, b = C.c_uint64(), C.c_uint64()
        _check(lib().fheacc_keys_digest(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value


class Server:
    """The serving loop (fheacc_server_*): gate-circuit jobs of many clients,
    admitted by HBM footprint and coalesced into shared gate-path launches."""

    def __init__(self, ctx, hbm_budget=1 << 30, max_batch=2048, max_inflight=2, min_fill=0, fair_share=True):
        self.ctx = ctx
        cfg = _ServerConfig(int(hbm_budget), int(max_batch), int(max_inflight), int(min_fill), int(bool(fair_share)))
        h = C.c_void_p()
        _check(lib().fheacc_server_create(ctx.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self._clients = []
        self._jobs = {}

    def add_client(self, keys):
        cid = C.c_int32()
        _check(lib().fheacc_server_add_client(self.h, keys.h, C.byref(cid)))
        self._clients.append(keys)
        return cid.value

    def submit(self, client, inputs, ops, outputs):
        """inputs: [n_in, n+1] words mod q; ops: [n_gates, 3] (gate, src1, src2);
        outputs: node indices.  Returns a job id for wait()."""
        inputs, pin = _np(np.asarray(inputs).reshape(-1, self.ctx.n + 1))
        ops = np.ascontiguousarray(np.asarray(ops, dtype=np.int32).reshape(-1, 3))
        outputs = np.ascontiguousarray(np.asarray(outputs, dtype=np.int64).reshape(-1))
        out = np.zeros((outputs.size, self.ctx.n + 1), dtype=np.uint64)
        jid = C.c_int64()
        _check(lib().fheacc_server_submit(self.h, client, pin, inputs.shape[0],
                                          ops.ctypes.data_as(C.c_void_p), ops.shape[0],
                                          outputs.ctypes.data_as(C.c_void_p), outputs.size,
                                          out.ctypes.data_as(C.c_void_p), C.byref(jid)))
        self._jobs[jid.value] = (inputs, ops, outputs, out)
        return jid.value

    def wait(self, job):
        """Returns (output words [n_out, n+1], latency in ms)."""
        lat = C.c_double()
        try:
            _check(lib().fheacc_server_wait(self.h, job, C.byref(lat)))
        finally:
            bufs = self._jobs.pop(job, None)
        return bufs[3], lat.value

    def stats(self):
        st = _ServerStats()
        _check(lib().fheacc_server_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}

    def reset_stats(self):
        _check(lib().fheacc_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().fheacc_server_destroy(self.h)
            self.h = None
            self._jobs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
