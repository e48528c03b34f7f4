"""ctypes binding of oracle/_ref/libfheref.so (the reference's own CPU path,
compiled from its sources in this container by `make -C oracle ref`).

TEST INFRASTRUCTURE ONLY.  Used to cross-check the oracle, and as the
"reference" kind of CPU baseline in bench.py when the library was built.
"""
import ctypes as C
import os

import numpy as np

from . import REF_LIB_PATH

_lib = None


def available():
    return os.path.exists(REF_LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(REF_LIB_PATH)
        U = C.POINTER(C.c_uint64)
        I32 = C.POINTER(C.c_int32)
        sig = {
            "ref_set_params": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.c_uint64, C.c_uint32,
                                         C.c_uint64, C.c_uint64, C.c_uint32]),
            "ref_get_params": (None, [U]),
            "ref_ntt": (None, [U, C.c_int64]),
            "ref_intt": (None, [U, C.c_int64]),
            "ref_mult_rgsw": (None, [U, U, U]),
            "ref_acc_init": (None, [C.c_uint64, C.c_uint64, C.c_uint64, U]),
            "ref_lwe_evaluate": (None, [U, U, C.c_int, U]),
            "ref_modswitch": (None, [U, U]),
            "ref_scale_round": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "ref_load_keys": (None, [U, U]),
            "ref_lwe_keyswitch": (None, [U, U]),
            "ref_eval_bootstrap": (C.c_double, [U, U, I32, C.c_int64, U]),
            "ref_eval_bootstrap_fpga": (C.c_int, [U, U, I32, C.c_int64, C.c_int, U, U]),
            "ref_substitute": (None, [U, C.c_int, U]),
            "ref_rlwe_keyswitch": (None, [U, U, U]),
            "ref_expand_rlwe": (None, [U, U, U]),
            "ref_rotate_freq": (None, [U, C.c_int, C.c_int, U]),
            "ref_cmux": (None, [U, U, U, U]),
            "ref_blind_rotate_rlwe": (C.c_int, [U, C.c_int, U, U]),
            "ref_blind_rotate_full": (None, [U, C.c_int, I32, C.c_int, U, U]),
            "ref_homo_expand": (None, [U, U, U, U]),
            "ref_cmux_tree": (None, [U, C.c_int, U, U]),
            "ref_rlwe_decrypt": (None, [U, U, C.c_uint64, U]),
            "ref_root_of_unity": (C.c_uint64, [C.c_int, C.c_uint64]),
            "ref_first_prime": (C.c_uint64, [C.c_int, C.c_int]),
            "ref_previous_prime": (C.c_uint64, [C.c_uint64, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


class Reference:
    """The reference's global parameter block set to one parameter set."""

    def __init__(self, n, q, N, Q, BGbits, Br, Bks, Bks_R_bits=9):
        lib().ref_set_params(n, q, N, Q, BGbits, Br, Bks, Bks_R_bits)
        self.n, self.q, self.N, self.Q = n, q, N, Q
        v = np.zeros(6, dtype=np.uint64)
        lib().ref_get_params(_p(v))
        self.ROU, self.iN = int(v[0]), int(v[1])
        self.digitG, self.digitR, self.digitKS, self.digitKS_R = (int(x) for x in v[2:])

    def ntt(self, a):
        a = np.array(a, dtype=np.uint64, copy=True)
        lib().ref_ntt(_p(a), a.size // self.N)
        return a

    def intt(self, a):
        a = np.array(a, dtype=np.uint64, copy=True)
        lib().ref_intt(_p(a), a.size // self.N)
        return a

    def mult_rgsw(self, acc, rgsw):
        acc = np.ascontiguousarray(acc, dtype=np.uint64)
        rgsw = np.ascontiguousarray(rgsw, dtype=np.uint64)
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_mult_rgsw(_p(acc), _p(rgsw), _p(out))
        return out

    def acc_init(self, b1, b2, value):
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_acc_init(b1, b2, value, _p(out))
        return out

    def evaluate(self, c1, c2, gate):
        c1 = np.ascontiguousarray(c1, dtype=np.uint64)
        c2 = np.ascontiguousarray(c2, dtype=np.uint64)
        out = np.zeros(self.n + 1, dtype=np.uint64)
        lib().ref_lwe_evaluate(_p(c1), _p(c2), gate, _p(out))
        return out

    def modswitch(self, c):
        c = np.ascontiguousarray(c, dtype=np.uint64)
        out = np.zeros(self.n + 1, dtype=np.uint64)
        lib().ref_modswitch(_p(c), _p(out))
        return out

    def load_keys(self, bsk, ksk):
        self._bsk = np.ascontiguousarray(bsk, dtype=np.uint64)
        self._ksk = np.ascontiguousarray(ksk, dtype=np.uint64)
        lib().ref_load_keys(_p(self._bsk), _p(self._ksk))

    def lwe_keyswitch(self, acc):
        acc = np.ascontiguousarray(acc, dtype=np.uint64)
        out = np.zeros(self.n + 1, dtype=np.uint64)
        lib().ref_lwe_keyswitch(_p(acc), _p(out))
        return out

    def bootstrap(self, c1, c2, gates):
        c1 = np.ascontiguousarray(c1, dtype=np.uint64).reshape(-1, self.n + 1)
        c2 = np.ascontiguousarray(c2, dtype=np.uint64).reshape(-1, self.n + 1)
        B = c1.shape[0]
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gates, dtype=np.int32), (B,)))
        out = np.zeros((B, self.n + 1), dtype=np.uint64)
        secs = lib().ref_eval_bootstrap(_p(c1), _p(c2), g.ctypes.data_as(C.POINTER(C.c_int32)), B, _p(out))
        return out, secs

    def bootstrap_fpga(self, c1, c2, gates, mode):
        """The accelerator's x1 (mode 1) or x4..x16 (mode 2) step order over the
        reference's own routines; returns (out, acc)."""
        c1 = np.ascontiguousarray(c1, dtype=np.uint64).reshape(-1, self.n + 1)
        c2 = np.ascontiguousarray(c2, dtype=np.uint64).reshape(-1, self.n + 1)
        B = c1.shape[0]
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gates, dtype=np.int32), (B,)))
        out = np.zeros((B, self.n + 1), dtype=np.uint64)
        acc = np.zeros((B, 2, self.N), dtype=np.uint64)
        rc = lib().ref_eval_bootstrap_fpga(_p(c1), _p(c2), g.ctypes.data_as(C.POINTER(C.c_int32)), B, mode, _p(out),
                                           _p(acc))
        if rc:
            raise ValueError("mode must be 1 or 2")
        return out, acc

    def substitute(self, rlwe, f):
        rlwe = np.ascontiguousarray(rlwe, dtype=np.uint64)
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_substitute(_p(rlwe), f, _p(out))
        return out

    def rlwe_keyswitch(self, rlwe, ks):
        rlwe = np.ascontiguousarray(rlwe, dtype=np.uint64)
        ks = np.ascontiguousarray(ks, dtype=np.uint64)
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_rlwe_keyswitch(_p(rlwe), _p(ks), _p(out))
        return out

    def expand_rlwe(self, rlwe, subs):
        rlwe = np.ascontiguousarray(rlwe, dtype=np.uint64)
        subs = np.ascontiguousarray(subs, dtype=np.uint64)
        out = np.zeros((self.N, 2, self.N), dtype=np.uint64)
        lib().ref_expand_rlwe(_p(rlwe), _p(subs), _p(out))
        return out

    def rotate_freq(self, rlwe, log_factor, forward):
        rlwe = np.ascontiguousarray(rlwe, dtype=np.uint64)
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_rotate_freq(_p(rlwe), log_factor, 1 if forward else 0, _p(out))
        return out

    def cmux(self, c0, c1, rgsw):
        c0, c1, rgsw = (np.ascontiguousarray(x, dtype=np.uint64) for x in (c0, c1, rgsw))
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_cmux(_p(c0), _p(c1), _p(rgsw), _p(out))
        return out

    def blind_rotate_full(self, rlwe, b, a, rgsws):
        rlwe = np.ascontiguousarray(rlwe, dtype=np.uint64)
        a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
        rgsws = np.ascontiguousarray(rgsws, dtype=np.uint64).reshape(-1)
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_blind_rotate_full(_p(rlwe), int(b), a.ctypes.data_as(C.POINTER(C.c_int32)), a.size, _p(rgsws),
                                    _p(out))
        return out

    def blind_rotate_rlwe(self, rlwe, forward, rgsws):
        rlwe, rgsws = (np.ascontiguousarray(x, dtype=np.uint64) for x in (rlwe, rgsws))
        out = np.zeros((2, self.N), dtype=np.uint64)
        if lib().ref_blind_rotate_rlwe(_p(rlwe), 1 if forward else 0, _p(rgsws), _p(out)) != 0:
            raise ValueError("the reference's blind_rotate needs N = 2048 (num_rotate_poly = 11)")
        return out

    def homo_expand(self, packed, enc_sec, subs):
        packed, enc_sec, subs = (np.ascontiguousarray(x, dtype=np.uint64) for x in (packed, enc_sec, subs))
        out = np.zeros((self.N, self.digitG * 4 * self.N), dtype=np.uint64)
        lib().ref_homo_expand(_p(packed), _p(enc_sec), _p(subs), _p(out))
        return out

    def cmux_tree(self, leaves, rgsws):
        leaves, rgsws = (np.ascontiguousarray(x, dtype=np.uint64) for x in (leaves, rgsws))
        levels = rgsws.size // (self.digitG * 4 * self.N)
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().ref_cmux_tree(_p(leaves), levels, _p(rgsws), _p(out))
        return out

    def rlwe_decrypt(self, sk_time, c, t):
        sk_time, c = (np.ascontiguousarray(x, dtype=np.uint64) for x in (sk_time, c))
        out = np.zeros(self.N, dtype=np.uint64)
        lib().ref_rlwe_decrypt(_p(sk_time), _p(c), t, _p(out))
        return out
