"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the parity reference for the HIP library.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The C sources (fhe_oracle.c) restate the reference's software path
(FHE_ACC_runtime/src/{nbt,params,LWE,RLWE,app}.cpp); each C function cites the
reference lines it follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libfheref.so")

GATES = {"OR": 0, "AND": 1, "NOR": 2, "NAND": 3, "XOR": 4, "XNOR": 5}


class Params(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("q", C.c_uint64), ("N", C.c_int32), ("Q", C.c_uint64),
        ("iN", C.c_uint64), ("ROU", C.c_uint64), ("BG", C.c_uint64),
        ("BGbits", C.c_uint32), ("digitG", C.c_uint32), ("Br", C.c_uint64),
        ("digitR", C.c_uint32), ("Bks", C.c_uint64), ("digitKS", C.c_uint32),
        ("Bks_R", C.c_uint64), ("Bks_R_bits", C.c_uint32), ("digitKS_R", C.c_uint32),
        ("E", C.c_uint64), ("E_rlwe", C.c_uint64), ("stddev", C.c_double),
        ("gate_constant", C.c_uint64 * 6), ("rou", C.POINTER(C.c_uint64)),
        ("irou", C.POINTER(C.c_uint64)), ("cdt", C.c_uint64 * 64), ("cdt_len", C.c_int32),
        ("eval_mode", C.c_int32),
    ]


# blind-rotation orders (fhe_oracle.h ORC_EVAL_*)
EVAL_SOFTWARE, EVAL_FPGA_X1, EVAL_FPGA_XN = 0, 1, 2


def build(force=False):
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "all"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(Params)
        U = C.POINTER(C.c_uint64)
        I32 = C.POINTER(C.c_int32)
        sig = {
            "orc_params_init": (C.c_int, [P, C.c_int, C.c_uint64, C.c_int, C.c_uint64, C.c_int,
                                          C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32,
                                          C.c_double, C.c_uint64]),
            "orc_params_free": (None, [P]),
            "orc_reverse_bit": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "orc_power_mod": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_inverse_mod": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "orc_is_prime": (C.c_int, [C.c_uint64]),
            "orc_first_prime": (C.c_uint64, [C.c_int, C.c_int]),
            "orc_previous_prime": (C.c_uint64, [C.c_uint64, C.c_int]),
            "orc_root_of_unity": (C.c_uint64, [C.c_int, C.c_uint64]),
            "orc_ntt_batch": (None, [P, U, C.c_int64]),
            "orc_intt_batch": (None, [P, U, C.c_int64]),
            "orc_poly_substitute": (None, [U, C.c_int, C.c_int, C.c_uint64, U]),
            "orc_digit_decompose": (None, [P, U, C.c_uint32, C.c_uint32, U]),
            "orc_rlwe_mult_rgsw": (None, [P, U, U, U]),
            "orc_rlwe_mult_rgsw_batch": (None, [P, U, U, C.c_int64, U, C.c_int64]),
            "orc_acc_init": (None, [P, C.c_uint64, C.c_uint64, C.c_uint64, U]),
            "orc_rlwe_substitute": (None, [P, U, C.c_int, U]),
            "orc_rlwe_keyswitch": (None, [P, U, U, U]),
            "orc_rotate_poly": (None, [P, C.c_int, C.c_int, U]),
            "orc_rlwe_rotate_freq": (None, [P, U, C.c_int, C.c_int, U]),
            "orc_expand_rlwe": (None, [P, U, U, U]),
            "orc_cmux": (None, [P, U, U, U, U]),
            "orc_blind_rotate_rlwe": (None, [P, U, C.c_int, U, C.c_int, U]),
            "orc_blind_rotate_full": (None, [P, U, C.c_int, I32, C.c_int, U, U]),
            "orc_lwe_evaluate": (C.c_int, [P, U, U, C.c_int, U]),
            "orc_lwe_eval_not": (None, [P, U, U]),
            "orc_lwe_keyswitch": (None, [P, U, U, U]),
            "orc_scale_round": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_lwe_modswitch": (None, [P, U, U]),
            "orc_lwe_decrypt": (C.c_uint64, [P, U, U]),
            "orc_blind_rotate": (C.c_int64, [P, U, U, U]),
            "orc_eval_bootstrap": (C.c_int, [P, U, U, C.c_int, U, U, U, U, U]),
            "orc_eval_bootstrap_batch": (C.c_int, [P, U, U, I32, C.c_int64, U, U, U, U, U, C.c_int]),
            "orc_rng_word": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_gauss": (C.c_int64, [P, C.c_uint64]),
            "orc_lwe_secret": (None, [P, C.c_uint64, C.c_int, U]),
            "orc_rlwe_secret": (None, [P, C.c_uint64, U]),
            "orc_rlwe_encrypt_zero": (None, [P, U, C.c_uint64, C.c_uint64, U]),
            "orc_rgsw_encrypt_monomial": (None, [P, U, C.c_uint64, C.c_uint64, C.c_uint64, U]),
            "orc_bsk_gen": (None, [P, U, U, C.c_uint64, U, C.c_int]),
            "orc_ksk_gen": (None, [P, U, U, C.c_uint64, U, C.c_int]),
            "orc_rlwe_ksk_gen": (None, [P, U, U, C.c_uint64, C.c_uint64, U]),
            "orc_subs_key_gen": (None, [P, U, C.c_uint64, U]),
            "orc_lwe_encrypt": (None, [P, U, C.c_uint64, C.c_uint64, C.c_uint64, U]),
            "orc_rlwe_encrypt": (None, [P, U, U, C.c_uint64, C.c_uint64, U]),
            "orc_rlwe_decrypt": (None, [P, U, U, C.c_uint64, U]),
            "orc_rgsw_encrypt": (None, [P, U, U, C.c_uint64, C.c_uint64, U]),
            "orc_homo_expand": (None, [P, U, U, U, U]),
            "orc_cmux_tree": (None, [P, U, C.c_int, U, C.c_size_t, U]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint64))


class Oracle:
    """One parameter set of the oracle (RLWE.cpp/LWE.cpp semantics)."""

    def __init__(self, n, q, N, Q=0, nbit=27, BGbits=7, Br=32, Bks=32, Bks_R_bits=9,
                 stddev=3.19, E_rlwe=18):
        self.p = Params()
        rc = lib().orc_params_init(C.byref(self.p), n, q, N, Q, nbit, BGbits, Br, Bks,
                                   Bks_R_bits, stddev, E_rlwe)
        if rc != 0:
            raise ValueError("invalid parameter set")
        p = self.p
        self.n, self.q, self.N, self.Q = p.n, p.q, p.N, p.Q
        self.digitG, self.digitR, self.digitKS, self.digitKS_R = p.digitG, p.digitR, p.digitKS, p.digitKS_R
        self.Br, self.Bks, self.BGbits, self.Bks_R_bits = p.Br, p.Bks, p.BGbits, p.Bks_R_bits
        self.iN, self.ROU = p.iN, p.ROU

    def __del__(self):
        try:
            lib().orc_params_free(C.byref(self.p))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self.p)

    @property
    def eval_mode(self):
        """Blind-rotation order used by blind_rotate / bootstrap (EVAL_*)."""
        return self.p.eval_mode

    @eval_mode.setter
    def eval_mode(self, mode):
        if mode not in (EVAL_SOFTWARE, EVAL_FPGA_X1, EVAL_FPGA_XN):
            raise ValueError("unknown evaluation mode %r" % (mode,))
        self.p.eval_mode = mode

    def rou_table(self):
        return np.ctypeslib.as_array(self.p.rou, shape=(self.N,)).copy()

    def irou_table(self):
        return np.ctypeslib.as_array(self.p.irou, shape=(self.N,)).copy()

    # sizes in u64 words
    @property
    def rgsw_words(self):
        return self.digitG * 4 * self.N

    @property
    def bsk_words(self):
        return self.Br * self.digitR * self.n * self.rgsw_words

    @property
    def ksk_words(self):
        return self.Bks * self.digitKS * self.N * (self.n + 1)

    # transforms
    def ntt(self, a):
        a, pa = _u64(np.array(a, dtype=np.uint64, copy=True))
        lib().orc_ntt_batch(self.ref, pa, a.size // self.N)
        return a

    def intt(self, a):
        a, pa = _u64(np.array(a, dtype=np.uint64, copy=True))
        lib().orc_intt_batch(self.ref, pa, a.size // self.N)
        return a

    def decompose(self, rlwe, d, bits):
        rlwe, pr = _u64(rlwe)
        out = np.zeros((d, 2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_digit_decompose(self.ref, pr, d, bits, po)
        return out

    def mult_rgsw(self, acc, rgsw):
        acc, pa = _u64(acc)
        rgsw, pr = _u64(rgsw)
        cnt = acc.size // (2 * self.N)
        stride = rgsw.size // cnt if rgsw.size != self.rgsw_words else 0
        out = np.zeros(acc.shape, dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_rlwe_mult_rgsw_batch(self.ref, pa, pr, stride, po, cnt)
        return out

    def acc_init(self, bound1, bound2, value):
        out = np.zeros((2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_acc_init(self.ref, bound1, bound2, value, po)
        return out

    def substitute(self, rlwe, f):
        rlwe, pr = _u64(rlwe)
        out = np.zeros((2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_rlwe_substitute(self.ref, pr, f, po)
        return out

    def rlwe_keyswitch(self, rlwe, ks):
        rlwe, pr = _u64(rlwe)
        ks, pk = _u64(ks)
        out = np.zeros((2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_rlwe_keyswitch(self.ref, pr, pk, po)
        return out

    def rotate_poly(self, log_factor, forward):
        out = np.zeros(self.N, dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_rotate_poly(self.ref, log_factor, 1 if forward else 0, po)
        return out

    def rotate_freq(self, rlwe, log_factor, forward):
        rlwe, pr = _u64(rlwe)
        out = np.zeros((2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_rlwe_rotate_freq(self.ref, pr, log_factor, 1 if forward else 0, po)
        return out

    def expand_rlwe(self, rlwe, subs):
        """expand_RLWE (app.cpp:26-97): [N][2][N] from one RLWE (NTT domain)."""
        rlwe, pr = _u64(rlwe)
        subs, pk = _u64(subs)
        out = np.zeros((self.N, 2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_expand_rlwe(self.ref, pr, pk, po)
        return out

    def cmux(self, c0, c1, rgsw):
        c0, p0 = _u64(c0)
        c1, p1 = _u64(c1)
        rgsw, pg = _u64(rgsw)
        out = np.zeros((2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_cmux(self.ref, p0, p1, pg, po)
        return out

    def blind_rotate_rlwe(self, rlwe, forward, rgsws):
        """blind_rotate (RLWE.cpp:822-838) over len(rgsws) RGSWs."""
        rlwe, pr = _u64(rlwe)
        rgsws = np.ascontiguousarray(rgsws, dtype=np.uint64).reshape(-1, self.rgsw_words)
        _, pg = _u64(rgsws)
        out = np.zeros((2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_blind_rotate_rlwe(self.ref, pr, 1 if forward else 0, pg, rgsws.shape[0], po)
        return out

    def blind_rotate_full(self, rlwe, b, a, rgsws):
        """blind_rotate_full (RLWE.cpp:795-820)."""
        rlwe, pr = _u64(rlwe)
        a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
        rgsws = np.ascontiguousarray(rgsws, dtype=np.uint64).reshape(-1)
        _, pg = _u64(rgsws)
        out = np.zeros((2, self.N), dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_blind_rotate_full(self.ref, pr, int(b), a.ctypes.data_as(C.POINTER(C.c_int32)), a.size, pg, po)
        return out

    def lwe_keyswitch(self, acc, ksk):
        acc, pa = _u64(acc)
        ksk, pk = _u64(ksk)
        out = np.zeros(self.n + 1, dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_lwe_keyswitch(self.ref, pa, pk, po)
        return out

    def modswitch(self, lwe):
        lwe, pl = _u64(lwe)
        out = np.zeros(self.n + 1, dtype=np.uint64)
        _, po = _u64(out)
        lib().orc_lwe_modswitch(self.ref, pl, po)
        return out

    def evaluate(self, c1, c2, gate):
        c1, p1 = _u64(c1)
        c2, p2 = _u64(c2)
        out = np.zeros(self.n + 1, dtype=np.uint64)
        _, po = _u64(out)
        rc = lib().orc_lwe_evaluate(self.ref, p1, p2, gate, po)
        if rc:
            raise ValueError("identical operands")
        return out

    def blind_rotate(self, lwe, bsk, acc):
        lwe, pl = _u64(lwe)
        bsk, pb = _u64(bsk)
        acc, pa = _u64(np.array(acc, dtype=np.uint64, copy=True))
        cnt = lib().orc_blind_rotate(self.ref, pl, pb, pa)
        return acc, cnt

    def bootstrap(self, c1, c2, gates, bsk, ksk, threads=0, capture=False):
        """Batched eval_bootstrap; returns out (and acc, ks captures)."""
        c1 = np.ascontiguousarray(c1, dtype=np.uint64).reshape(-1, self.n + 1)
        c2 = np.ascontiguousarray(c2, dtype=np.uint64).reshape(-1, self.n + 1)
        B = c1.shape[0]
        gates = np.ascontiguousarray(np.broadcast_to(np.asarray(gates, dtype=np.int32), (B,)))
        out = np.zeros((B, self.n + 1), dtype=np.uint64)
        cap_acc = np.zeros((B, 2, self.N), dtype=np.uint64) if capture else None
        cap_ks = np.zeros((B, self.n + 1), dtype=np.uint64) if capture else None
        U = C.POINTER(C.c_uint64)
        nul = C.cast(None, U)
        rc = lib().orc_eval_bootstrap_batch(
            self.ref, _u64(c1)[1], _u64(c2)[1], gates.ctypes.data_as(C.POINTER(C.c_int32)), B,
            _u64(bsk)[1], _u64(ksk)[1], _u64(out)[1],
            _u64(cap_acc)[1] if capture else nul, _u64(cap_ks)[1] if capture else nul, threads)
        if rc:
            raise ValueError("identical operands")
        return (out, cap_acc, cap_ks) if capture else out

    def decrypt(self, sk, c):
        sk, ps = _u64(sk)
        c = np.ascontiguousarray(c, dtype=np.uint64).reshape(-1, self.n + 1)
        return np.array([lib().orc_lwe_decrypt(self.ref, ps, _u64(row)[1]) for row in c])

    # keygen
    def lwe_secret(self, seed, binary=True):
        out = np.zeros(self.n, dtype=np.uint64)
        lib().orc_lwe_secret(self.ref, seed, 1 if binary else 0, _u64(out)[1])
        return out

    def rlwe_secret(self, seed):
        out = np.zeros(self.N, dtype=np.uint64)
        lib().orc_rlwe_secret(self.ref, seed, _u64(out)[1])
        return out

    def bsk_gen(self, lwe_sk, rlwe_sk, seed, threads=0):
        out = np.zeros(self.bsk_words, dtype=np.uint64)
        lib().orc_bsk_gen(self.ref, _u64(lwe_sk)[1], _u64(rlwe_sk)[1], seed, _u64(out)[1], threads)
        return out

    def ksk_gen(self, rlwe_sk, lwe_sk, seed, threads=0):
        out = np.zeros(self.ksk_words, dtype=np.uint64)
        lib().orc_ksk_gen(self.ref, _u64(rlwe_sk)[1], _u64(lwe_sk)[1], seed, _u64(out)[1], threads)
        return out

    def subs_key_gen(self, rlwe_sk, seed):
        logN = self.N.bit_length() - 1
        out = np.zeros(logN * self.digitKS_R * 2 * self.N, dtype=np.uint64)
        lib().orc_subs_key_gen(self.ref, _u64(rlwe_sk)[1], seed, _u64(out)[1])
        return out

    def rlwe_encrypt_zero(self, sk_ntt, seed, stream):
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().orc_rlwe_encrypt_zero(self.ref, _u64(sk_ntt)[1], seed, stream, _u64(out)[1])
        return out

    def lwe_encrypt(self, sk, bits, seed, first_index=0):
        bits = np.asarray(bits).reshape(-1)
        out = np.zeros((bits.size, self.n + 1), dtype=np.uint64)
        psk = _u64(sk)
        for i, m in enumerate(bits):
            lib().orc_lwe_encrypt(self.ref, psk[1], int(m), seed, first_index + i, _u64(out[i])[1])
        return out

    def gauss(self, word):
        return lib().orc_gauss(self.ref, word)

    # RLWE / RGSW encryption of message polynomials (time domain, mod Q)
    def rlwe_encrypt(self, sk_time, msgs, seed, first_index=0):
        """RLWE_encrypt (RLWE.cpp:324-344) of each row of msgs; ciphertext i
        draws from stream (TAG_RLWE_ENC, first_index + i)."""
        msgs = np.ascontiguousarray(msgs, dtype=np.uint64).reshape(-1, self.N)
        sk_ntt = self.ntt(np.asarray(sk_time, dtype=np.uint64))
        out = np.zeros((msgs.shape[0], 2, self.N), dtype=np.uint64)
        for i in range(msgs.shape[0]):
            lib().orc_rlwe_encrypt(self.ref, _u64(sk_ntt)[1], _u64(msgs[i])[1], seed,
                                   stream(TAG_RLWE_ENC, first_index + i), _u64(out[i])[1])
        return out

    def rlwe_decrypt(self, sk_time, c, t):
        """RLWE_decrypt(s, c, t) (RLWE.cpp:362-376): [count][N] time domain."""
        c = np.ascontiguousarray(c, dtype=np.uint64).reshape(-1, 2, self.N)
        sk_ntt = self.ntt(np.asarray(sk_time, dtype=np.uint64))
        out = np.zeros((c.shape[0], self.N), dtype=np.uint64)
        for i in range(c.shape[0]):
            lib().orc_rlwe_decrypt(self.ref, _u64(sk_ntt)[1], _u64(c[i])[1], t, _u64(out[i])[1])
        return out

    def rgsw_encrypt(self, sk_time, msgs, seed, first_index=0):
        """RGSW encryption (RLWE.cpp:561-577) of each row of msgs; RGSW i draws
        from streams (TAG_RGSW_ENC, (first_index + i) 2 digitG + 2h + g)."""
        msgs = np.ascontiguousarray(msgs, dtype=np.uint64).reshape(-1, self.N)
        sk_ntt = self.ntt(np.asarray(sk_time, dtype=np.uint64))
        out = np.zeros((msgs.shape[0], self.rgsw_words), dtype=np.uint64)
        for i in range(msgs.shape[0]):
            base = stream(TAG_RGSW_ENC, (first_index + i) * 2 * self.digitG)
            lib().orc_rgsw_encrypt(self.ref, _u64(sk_ntt)[1], _u64(msgs[i])[1], seed, base, _u64(out[i])[1])
        return out

    def homo_expand(self, packed, enc_sec, subs):
        """homo_expand (app.cpp:78-97): packed [digitG][2][N] -> [N] RGSWs."""
        packed, pp = _u64(np.asarray(packed).reshape(self.digitG, 2, self.N))
        enc_sec, pe = _u64(enc_sec)
        subs, pk = _u64(subs)
        out = np.zeros((self.N, self.rgsw_words), dtype=np.uint64)
        lib().orc_homo_expand(self.ref, pp, pe, pk, _u64(out)[1])
        return out

    def cmux_tree(self, leaves, rgsws):
        """The tree selection's CMUX tree (test.cpp:1576-1582): 2^L leaves,
        L RGSWs (level l selects with rgsws[l])."""
        rgsws = np.ascontiguousarray(rgsws, dtype=np.uint64).reshape(-1, self.rgsw_words)
        levels = rgsws.shape[0]
        leaves = np.ascontiguousarray(leaves, dtype=np.uint64).reshape(1 << levels, 2, self.N)
        out = np.zeros((2, self.N), dtype=np.uint64)
        lib().orc_cmux_tree(self.ref, _u64(leaves)[1], levels, _u64(rgsws)[1], self.rgsw_words, _u64(out)[1])
        return out


def rng_word(seed, stream, ctr):
    return lib().orc_rng_word(seed, stream, ctr)


def stream(tag, idx):
    return (tag << 56) | idx


# stream tags shared with the device sampler (fhe_oracle.h ORC_TAG_*)
TAG_RLWE_ENC, TAG_RGSW_ENC = 7, 8
