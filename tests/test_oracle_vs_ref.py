"""Cross-check the oracle against the reference's own CPU path compiled from
its sources in this container (oracle/_ref/libfheref.so, `make -C oracle ref`).

These pin the parts no reference-held vector covers: the full eval_bootstrap
word stream (blind rotation with the zero-digit skip, LWE_keyswitch with its
Q/8+1 offset and d=0 rows, LWE_modswitch), acc_initialization for every gate
and b, and external products at other (BG, d, Q, N).  Keys are explicit
inputs from the oracle's seeded generator (the reference's own keygen draws
from std::random_device, RNG.cpp:5-6).  Skipped when _ref was not built (where the reference sources are absent).
"""
import numpy as np
import pytest

from oracle import GATES, Oracle, refbind

pytestmark = [pytest.mark.ref, pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built")]

SETS = [
    # n, q, N, nbit/Q, BGbits, Br, Bks
    dict(n=16, q=512, N=256, Q=0, nbit=27, BGbits=9, Br=23, Bks=25),
    dict(n=12, q=1024, N=512, Q=0, nbit=27, BGbits=7, Br=32, Bks=32),
    dict(n=10, q=256, N=128, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=16, Bks=32),
    dict(n=8, q=1024, N=512, Q=0, nbit=27, BGbits=6, Br=32, Bks=32),
]


def pair(cfg):
    o = Oracle(n=cfg["n"], q=cfg["q"], N=cfg["N"], Q=cfg["Q"], nbit=cfg["nbit"],
               BGbits=cfg["BGbits"], Br=cfg["Br"], Bks=cfg["Bks"])
    r = refbind.Reference(cfg["n"], cfg["q"], cfg["N"], o.Q, cfg["BGbits"], cfg["Br"], cfg["Bks"])
    return o, r


@pytest.mark.parametrize("cfg", SETS)
def test_params_and_ntt(cfg):
    o, r = pair(cfg)
    assert (o.ROU, o.iN, o.digitG, o.digitR, o.digitKS, o.digitKS_R) == \
        (r.ROU, r.iN, r.digitG, r.digitR, r.digitKS, r.digitKS_R)
    rng = np.random.default_rng(cfg["N"])
    a = rng.integers(0, o.Q, (3, o.N), dtype=np.uint64)
    assert np.array_equal(o.ntt(a), r.ntt(a))
    assert np.array_equal(o.intt(a), r.intt(a))


@pytest.mark.parametrize("N,nbit,Qx", [(1024, 27, 0), (1024, 54, 0x3FFFFFFFFED001), (2048, 54, 0)])
def test_prime_search(N, nbit, Qx):
    from oracle import lib
    rl = refbind.lib()
    L = lib()
    p = L.orc_first_prime(nbit, N)
    assert p == rl.ref_first_prime(nbit, N)
    assert L.orc_previous_prime(p, N) == rl.ref_previous_prime(p, N)
    Q = Qx or L.orc_previous_prime(p, N)
    assert L.orc_root_of_unity(N, Q) == rl.ref_root_of_unity(N, Q)


@pytest.mark.parametrize("cfg", SETS)
def test_external_product(cfg):
    o, r = pair(cfg)
    rng = np.random.default_rng(7)
    for _ in range(2):
        acc = rng.integers(0, o.Q, (2, o.N), dtype=np.uint64)
        rgsw = rng.integers(0, o.Q, o.rgsw_words, dtype=np.uint64)
        assert np.array_equal(o.mult_rgsw(acc, rgsw), r.mult_rgsw(acc, rgsw))


@pytest.mark.parametrize("cfg", SETS[:2])
def test_acc_init_all_gates(cfg):
    o, r = pair(cfg)
    q = cfg["q"]
    for g in range(6):
        b1 = o.p.gate_constant[g] % q
        b2 = (b1 + q // 2) % q
        for value in (0, 1, q // 8, q // 2 + 3, q - 1):
            assert np.array_equal(o.acc_init(b1, b2, value), r.acc_init(b1, b2, value))


@pytest.mark.parametrize("cfg", SETS)
def test_bootstrap_words(cfg):
    o, r = pair(cfg)
    s = o.lwe_secret(5, binary=(cfg["N"] != 512))
    z = o.rlwe_secret(5)
    bsk = o.bsk_gen(s, z, 5)
    ksk = o.ksk_gen(z, s, 5)
    r.load_keys(bsk, ksk)
    m1 = np.array([0, 1, 1, 0, 1, 0])
    m2 = np.array([0, 0, 1, 1, 1, 1])
    c1 = o.lwe_encrypt(s, m1, seed=9)
    c2 = o.lwe_encrypt(s, m2, seed=9, first_index=50)
    gates = np.array([GATES[g] for g in ("NAND", "AND", "OR", "NOR", "XOR", "XNOR")], dtype=np.int32)
    out, acc, ks = o.bootstrap(c1, c2, gates, bsk, ksk, capture=True)
    rout, _ = r.bootstrap(c1, c2, gates)
    assert np.array_equal(out, rout)
    for i in range(len(gates)):
        assert np.array_equal(ks[i], r.lwe_keyswitch(acc[i]))
        assert np.array_equal(out[i], r.modswitch(ks[i]))
        ev = o.evaluate(c1[i], c2[i], int(gates[i]))
        assert np.array_equal(ev, r.evaluate(c1[i], c2[i], int(gates[i])))


@pytest.mark.parametrize("cfg", [SETS[0], SETS[2]])
def test_subs_and_rlwe_keyswitch(cfg):
    o, r = pair(cfg)
    rng = np.random.default_rng(3)
    x = rng.integers(0, o.Q, (2, o.N), dtype=np.uint64)
    logN = o.N.bit_length() - 1
    for f in range(logN):
        assert np.array_equal(o.substitute(x, f), r.substitute(x, f))
    ks = rng.integers(0, o.Q, o.digitKS_R * 2 * o.N, dtype=np.uint64)
    assert np.array_equal(o.rlwe_keyswitch(x, ks), r.rlwe_keyswitch(x, ks))


@pytest.mark.parametrize("N,Q,Bks_R_bits", [(64, 0x3FFFFFFFFED001, 9), (128, 0, 7)])
def test_expand_rlwe(N, Q, Bks_R_bits):
    # application::expand_RLWE (app.cpp:26-97) on random words and a random
    # substitution key: every word of all N outputs
    cfg = dict(n=4, q=2 * N, N=N, Q=Q, nbit=27, BGbits=9, Br=16, Bks=32)
    o = Oracle(**cfg, Bks_R_bits=Bks_R_bits)
    r = refbind.Reference(cfg["n"], cfg["q"], N, o.Q, cfg["BGbits"], cfg["Br"], cfg["Bks"], Bks_R_bits)
    rng = np.random.default_rng(N)
    x = rng.integers(0, o.Q, (2, N), dtype=np.uint64)
    logN = N.bit_length() - 1
    subs = rng.integers(0, o.Q, logN * o.digitKS_R * 2 * N, dtype=np.uint64)
    assert np.array_equal(o.expand_rlwe(x, subs), r.expand_rlwe(x, subs))
    for i in range(logN):
        for fwd in (True, False):
            assert np.array_equal(o.rotate_freq(x, i, fwd), r.rotate_freq(x, i, fwd))


def test_cmux_and_leveled_blind_rotate():
    # RLWE::CMUX and RLWE::blind_rotate (RLWE.cpp:788-838); the reference's
    # blind_rotate takes num_rotate_poly = 11 RGSWs, so N = 2048
    cfg = dict(n=4, q=1024, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=32, Bks=32)
    o, r = pair(cfg)
    rng = np.random.default_rng(11)
    c0, c1 = (rng.integers(0, o.Q, (2, o.N), dtype=np.uint64) for _ in range(2))
    rgsw = rng.integers(0, o.Q, (11, o.rgsw_words), dtype=np.uint64)
    assert np.array_equal(o.cmux(c0, c1, rgsw[0]), r.cmux(c0, c1, rgsw[0]))
    for fwd in (True, False):
        assert np.array_equal(o.blind_rotate_rlwe(c0, fwd, rgsw), r.blind_rotate_rlwe(c0, fwd, rgsw))


@pytest.mark.parametrize("N,Q,BGbits", [(64, 0x3FFFFFFFFED001, 18), (128, 0, 7)])
def test_blind_rotate_full(N, Q, BGbits):
    # RLWE::blind_rotate_full (RLWE.cpp:795-820): public and private rotation
    # factors of both signs and beyond N (poly_rotate_time reduces mod N)
    cfg = dict(n=4, q=2 * N, N=N, Q=Q, nbit=27, BGbits=BGbits, Br=16, Bks=32)
    o, r = pair(cfg)
    rng = np.random.default_rng(N + BGbits)
    x = rng.integers(0, o.Q, (2, N), dtype=np.uint64)
    a = np.array([3, -5, N + 7, -(N + 1), 0, 2 * N - 1], dtype=np.int32)
    rgsw = rng.integers(0, o.Q, (a.size, o.rgsw_words), dtype=np.uint64)
    for b in (0, 9, -4, N):
        assert np.array_equal(o.blind_rotate_full(x, b, a, rgsw), r.blind_rotate_full(x, b, a, rgsw)), b


@pytest.mark.parametrize("N,Q,BGbits,Bks_R_bits", [(64, 0x3FFFFFFFFED001, 18, 9), (128, 0, 9, 7)])
def test_homo_expand(N, Q, BGbits, Bks_R_bits):
    # application::homo_expand (app.cpp:78-97) on random words: every word of
    # all N output RGSWs (expansions, then products with RGSW(-s))
    cfg = dict(n=4, q=2 * N, N=N, Q=Q, nbit=27, BGbits=BGbits, Br=16, Bks=32)
    o = Oracle(**cfg, Bks_R_bits=Bks_R_bits)
    r = refbind.Reference(cfg["n"], cfg["q"], N, o.Q, BGbits, cfg["Br"], cfg["Bks"], Bks_R_bits)
    rng = np.random.default_rng(N + 1)
    packed = rng.integers(0, o.Q, (o.digitG, 2, N), dtype=np.uint64)
    enc_sec = rng.integers(0, o.Q, o.rgsw_words, dtype=np.uint64)
    logN = N.bit_length() - 1
    subs = rng.integers(0, o.Q, logN * o.digitKS_R * 2 * N, dtype=np.uint64)
    assert np.array_equal(o.homo_expand(packed, enc_sec, subs), r.homo_expand(packed, enc_sec, subs))


@pytest.mark.parametrize("levels", [1, 3])
def test_cmux_tree(levels):
    # the tree selection's CMUX loop (test.cpp:1576-1582) over the reference's
    # RLWE::CMUX, on random words
    cfg = dict(n=4, q=256, N=128, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=16, Bks=32)
    o, r = pair(cfg)
    rng = np.random.default_rng(40 + levels)
    leaves = rng.integers(0, o.Q, (1 << levels, 2, o.N), dtype=np.uint64)
    rgsw = rng.integers(0, o.Q, (levels, o.rgsw_words), dtype=np.uint64)
    assert np.array_equal(o.cmux_tree(leaves, rgsw), r.cmux_tree(leaves, rgsw))


@pytest.mark.parametrize("cfg", SETS[:3])
def test_rlwe_decrypt(cfg):
    # RLWE_decrypt(s, c, t) (RLWE.cpp:362-376) with its denoise(t) rounding,
    # including words that round across Q
    o, r = pair(cfg)
    rng = np.random.default_rng(cfg["N"] + 3)
    sk = rng.integers(0, o.Q, o.N, dtype=np.uint64)
    for t in (2, 4, 1024):
        c = rng.integers(0, o.Q, (2, o.N), dtype=np.uint64)
        assert np.array_equal(o.rlwe_decrypt(sk, c, t)[0], r.rlwe_decrypt(sk, c, t)), t
