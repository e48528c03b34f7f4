"""Pin the oracle with what the mathematics fixes (closed forms, invariants,
brute force on tiny inputs).  Each check is chosen so that a plausible mistake
(a dropped term, a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest

from oracle import GATES, Oracle, lib


def brute_min_root(N, Q):
    m = 2 * N
    best = None
    for x in range(2, Q):
        if pow(x, N, Q) == Q - 1:  # order exactly 2N (2N a power of two)
            best = x
            break
    return best


@pytest.mark.parametrize("N", [4, 8, 16, 32])
def test_min_root_brute_force(N):
    L = lib()
    # small NTT-friendly primes = 1 mod 2N
    primes = [p for p in range(2 * N + 1, 20000, 2 * N) if L.orc_is_prime(p)][:4]
    for Q in primes:
        assert L.orc_root_of_unity(N, Q) == brute_min_root(N, Q)


def test_is_prime_small():
    L = lib()
    sieve = np.ones(5000, dtype=bool)
    sieve[:2] = False
    for i in range(2, 71):
        if sieve[i]:
            sieve[i * i::i] = False
    for x in range(5000):
        assert bool(L.orc_is_prime(x)) == bool(sieve[x])


def schoolbook_negacyclic(a, b, Q):
    N = len(a)
    out = [0] * N
    for i in range(N):
        for j in range(N):
            k = i + j
            if k < N:
                out[k] = (out[k] + a[i] * b[j]) % Q
            else:
                out[k - N] = (out[k - N] - a[i] * b[j]) % Q
    return out


@pytest.mark.parametrize("N,nbit", [(16, 27), (64, 27), (32, 54)])
def test_ntt_convolution(N, nbit):
    o = Oracle(n=4, q=2 * N if 2 * N <= 1024 else 1024, N=N, nbit=nbit)
    rng = np.random.default_rng(N + nbit)
    a = rng.integers(0, o.Q, N, dtype=np.uint64)
    b = rng.integers(0, o.Q, N, dtype=np.uint64)
    A, Bv = o.ntt(a), o.ntt(b)
    prod = (A.astype(object) * Bv.astype(object)) % o.Q
    c = o.intt(prod.astype(np.uint64))
    assert [int(x) for x in c] == schoolbook_negacyclic([int(x) for x in a], [int(x) for x in b], o.Q)
    assert np.array_equal(o.intt(o.ntt(a)), a)


@pytest.mark.parametrize("N", [16, 1024])
def test_ntt_monomial_closed_form(N):
    # NTT(X^k)[i] = psi^((2 bitrev(i) + 1) k): the output is in bit-reversed order
    o = Oracle(n=4, q=min(2 * N, 1024), N=N, nbit=27)
    L = lib()
    logN = N.bit_length() - 1
    for k in (0, 1, 3, N - 1):
        x = np.zeros(N, dtype=np.uint64)
        x[k] = 1
        X = o.ntt(x)
        for i in (0, 1, 2, N // 2, N - 1):
            br = int(L.orc_reverse_bit(N, i))
            assert int(X[i]) == pow(o.ROU, (2 * br + 1) * k, o.Q)
    assert logN > 0


def test_decompose_recomposes():
    o = Oracle(n=4, q=128, N=64, nbit=27, BGbits=7)
    rng = np.random.default_rng(1)
    acc = o.ntt(rng.integers(0, o.Q, (2, 64), dtype=np.uint64))
    dec = o.decompose(acc, o.digitG, o.BGbits)
    time = o.intt(acc)
    digits = o.intt(dec.reshape(-1, 64)).reshape(dec.shape)
    assert digits.max() < (1 << o.BGbits)
    rec = sum(digits[j].astype(object) << (j * o.BGbits) for j in range(o.digitG))
    assert np.array_equal(np.array(rec, dtype=object).astype(np.uint64), time)


def test_transpose_invariant():
    # test.cpp:425-437: after the NTT-domain reversal, a'[0] = a[0] and
    # a'[N-i] = Q - a[i] in the time domain.
    o = Oracle(n=4, q=128, N=64, nbit=27)
    rng = np.random.default_rng(2)
    a = rng.integers(1, o.Q, 64, dtype=np.uint64)
    A = o.ntt(a)
    at = o.intt(A[::-1].copy())
    assert at[0] == a[0]
    for i in range(1, 64):
        assert int(at[64 - i]) == (o.Q - int(a[i])) % o.Q


def test_substitute_expansion_invariant():
    # test.cpp:1837-1914: repeated x + subs(x) halves the support; a constant
    # term survives scaled by N.
    N = 32
    o = Oracle(n=4, q=64, N=N, nbit=27)
    L = lib()
    import ctypes as C
    x = np.zeros(N, dtype=np.uint64)
    x[0] = 5
    x[3] = 7
    x[16] = 11
    cur = x.copy()
    logN = 5
    for f in range(logN):
        out = np.zeros(N, dtype=np.uint64)
        L.orc_poly_substitute(cur.ctypes.data_as(C.POINTER(C.c_uint64)), f, N, o.Q,
                              out.ctypes.data_as(C.POINTER(C.c_uint64)))
        cur = ((cur.astype(object) + out.astype(object)) % o.Q).astype(np.uint64)
    assert int(cur[0]) == 5 * N % o.Q
    assert not cur[1:].any()


def test_substitute_is_automorphism():
    # sigma(a*b) = sigma(a)*sigma(b) for X -> X^(N/2^f + 1)
    N = 16
    o = Oracle(n=4, q=32, N=N, nbit=27)
    rng = np.random.default_rng(3)
    a = rng.integers(0, o.Q, (2, N), dtype=np.uint64)
    A = o.ntt(a)
    for f in range(4):
        sa = o.substitute(A, f)
        prod = (A[0].astype(object) * A[1].astype(object)) % o.Q
        sp = o.substitute(np.stack([prod.astype(np.uint64), prod.astype(np.uint64)]), f)
        rhs = (sa[0].astype(object) * sa[1].astype(object)) % o.Q
        assert np.array_equal(sp[0], rhs.astype(np.uint64))


def test_scale_round_textbook():
    L = lib()
    Q, q = 0x7FFF801, 1024
    rng = np.random.default_rng(4)
    xs = list(rng.integers(0, Q, 2000)) + [0, Q - 1, Q // 2, Q - Q // (2 * q)]
    for x in xs:
        x = int(x)
        expect = (2 * x * q + Q) // (2 * Q)  # round-half-up of x*q/Q
        assert L.orc_scale_round(x, q, Q) == expect
    # not reduced mod q: the top of the range maps to q itself
    assert L.orc_scale_round(Q - 1, q, Q) == q


def test_lwe_evaluate_rules():
    o = Oracle(n=8, q=1024, N=512, nbit=27)
    rng = np.random.default_rng(5)
    c1 = rng.integers(0, 1024, 9, dtype=np.uint64)
    c2 = rng.integers(0, 1024, 9, dtype=np.uint64)
    for g, gi in GATES.items():
        out = o.evaluate(c1, c2, gi)
        if g in ("XOR", "XNOR"):
            exp = (2 * (c1.astype(np.int64) - c2.astype(np.int64))) % 1024
        else:
            exp = (c1 + c2) % 1024
        assert np.array_equal(out, exp.astype(np.uint64))
    with pytest.raises(ValueError):
        o.evaluate(c1, c1, GATES["NAND"])


def test_gauss_shape():
    o = Oracle(n=8, q=1024, N=512, nbit=27)
    from oracle import rng_word
    xs = np.array([o.gauss(rng_word(9, 1, i)) for i in range(20000)])
    assert abs(xs.mean()) < 0.1
    assert 2.72 < xs.std() < 2.90  # trunc(N(0, 3.19)) has std ~2.82
    assert np.count_nonzero(xs == 0) / xs.size == pytest.approx(0.24, abs=0.02)


def small_keys(o, seed=11):
    s = o.lwe_secret(seed, binary=True)
    z = o.rlwe_secret(seed)
    bsk = o.bsk_gen(s, z, seed)
    ksk = o.ksk_gen(z, s, seed)
    return s, z, bsk, ksk


@pytest.mark.parametrize("gate", list(GATES))
def test_bootstrap_truth_tables(gate):
    # test.cpp:496-625: decrypted gate outputs match the truth table
    o = Oracle(n=16, q=512, N=256, nbit=27, BGbits=9, Br=23, Bks=25)
    s, z, bsk, ksk = small_keys(o)
    m1 = np.array([0, 0, 1, 1])
    m2 = np.array([0, 1, 0, 1])
    c1 = o.lwe_encrypt(s, m1, seed=3, first_index=0)
    c2 = o.lwe_encrypt(s, m2, seed=3, first_index=100)
    out = o.bootstrap(c1, c2, GATES[gate], bsk, ksk)
    got = o.decrypt(s, out)
    table = {"OR": m1 | m2, "AND": m1 & m2, "NOR": 1 - (m1 | m2), "NAND": 1 - (m1 & m2),
             "XOR": m1 ^ m2, "XNOR": 1 - (m1 ^ m2)}[gate]
    assert list(got) == list(table)


def test_fresh_encryptions_decrypt():
    o = Oracle(n=64, q=1024, N=512, nbit=27)
    s = o.lwe_secret(4)
    bits = np.arange(200) % 2
    c = o.lwe_encrypt(s, bits, seed=8)
    assert list(o.decrypt(s, c)) == list(bits)
    # RLWE encryption of zero: b - a*s is the small error in the time domain
    z = o.rlwe_secret(4)
    zn = o.ntt(z)
    ct = o.rlwe_encrypt_zero(zn, 5, 77)
    e = (ct[1].astype(object) - ct[0].astype(object) * zn.astype(object)) % o.Q
    et = o.intt(e.astype(np.uint64)).astype(object)
    et = [x if x < o.Q // 2 else x - o.Q for x in et]
    assert max(abs(x) for x in et) < 18


def _rlwe_decrypt_phase(o, c, s_ntt):
    """Time-domain phase b - a s of an RLWE ciphertext (NTT domain in)."""
    Q = o.Q
    ph = (c[1].astype(object) - c[0].astype(object) * s_ntt.astype(object)) % Q
    return [int(v) for v in o.intt(np.array(ph, dtype=np.uint64))]


def test_rotate_freq_inverse_pair():
    # X^(2^i) X^(-2^i) = 1: forward then backward rotation is the identity
    o = Oracle(n=4, q=64, N=32, nbit=27)
    rng = np.random.default_rng(5)
    x = rng.integers(0, o.Q, (2, 32), dtype=np.uint64)
    for i in range(5):
        assert np.array_equal(o.rotate_freq(o.rotate_freq(x, i, True), i, False), x)
    # rotating by X^1 shifts the time-domain coefficients negacyclically
    t = o.intt(o.rotate_freq(x, 0, True)[0])
    a = o.intt(x[0])
    assert int(t[0]) == (o.Q - int(a[31])) % o.Q and np.array_equal(t[1:], a[:31])


@pytest.mark.parametrize("N", [32, 64])
def test_expand_rlwe_decrypts_to_scaled_coefficients(N):
    # an encryption of m(X) expands into N ciphertexts; slot k decrypts to
    # N m_k in its constant term (and ~0 elsewhere) -- the PSI query
    # expansion app.cpp:26-97 exists for
    o = Oracle(n=4, q=2 * N, N=N, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Bks_R_bits=9)
    z = o.rlwe_secret(77)
    s_ntt = o.ntt(z)
    subs = o.subs_key_gen(z, 78)
    rng = np.random.default_rng(N)
    m = rng.integers(0, 2, N)
    delta = o.Q // (4 * N)
    c = o.rlwe_encrypt_zero(s_ntt, 79, 1)
    mt = np.array([(int(v) * delta) % o.Q for v in m], dtype=np.uint64)
    c[1] = ((c[1].astype(object) + o.ntt(mt).astype(object)) % o.Q).astype(np.uint64)
    out = o.expand_rlwe(c, subs)
    scale = N * delta
    for k in range(N):
        ph = _rlwe_decrypt_phase(o, out[k], s_ntt)
        centered = [v if v < o.Q // 2 else v - o.Q for v in ph]
        assert round(centered[0] / scale) == m[k], k
        assert max(abs(v) for v in centered[1:]) < scale // 4, k


def test_cmux_selects():
    # CMUX(c0, c1, RGSW(b)) decrypts to c_b's message
    N = 64
    o = Oracle(n=4, q=128, N=N, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18)
    z = o.rlwe_secret(81)
    s_ntt = o.ntt(z)
    rng = np.random.default_rng(9)
    delta = o.Q // 8
    msgs = [rng.integers(0, 2, N) for _ in range(2)]
    cts = []
    for t, mm in enumerate(msgs):
        c = o.rlwe_encrypt_zero(s_ntt, 82, 10 + t)
        mt = np.array([(int(v) * delta) % o.Q for v in mm], dtype=np.uint64)
        c[1] = ((c[1].astype(object) + o.ntt(mt).astype(object)) % o.Q).astype(np.uint64)
        cts.append(c)
    import ctypes as C
    P = C.POINTER(C.c_uint64)
    # RGSW(1): the monomial X^0; RGSW(0): every row an encryption of zero
    one = np.zeros(o.rgsw_words, dtype=np.uint64)
    lib().orc_rgsw_encrypt_monomial(o.ref, s_ntt.ctypes.data_as(P), 0, 83, 20, one.ctypes.data_as(P))
    zero = np.concatenate([o.rlwe_encrypt_zero(s_ntt, 84, 200 + r).reshape(-1) for r in range(2 * o.digitG)])
    for bit, sel in ((0, zero), (1, one)):
        out = o.cmux(cts[0], cts[1], sel)
        ph = _rlwe_decrypt_phase(o, out, s_ntt)
        got = [round((v if v < o.Q // 2 else v - o.Q) / delta) % 2 for v in ph]
        assert got == list(msgs[bit]), bit


def _phase(o, c, s_ntt):
    return np.array(_rlwe_decrypt_phase(o, c, s_ntt), dtype=object)


def _noise(o, seed, stream_id):
    # the sampler's error term of an encryption of zero on stream_id
    from oracle import rng_word
    out = []
    for k in range(o.N):
        g = int(o.gauss(rng_word(seed, stream_id, o.N + k)))
        out.append(int(np.fmod(g, 18)))
    return np.array(out, dtype=object)


def test_rlwe_encrypt_identity():
    # RLWE_encrypt (RLWE.cpp:324-344): b - a s is exactly the sampled error
    # plus the message polynomial, coefficient by coefficient
    from oracle import TAG_RLWE_ENC, stream
    o = Oracle(n=4, q=64, N=32, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18)
    z = o.rlwe_secret(91)
    rng = np.random.default_rng(91)
    m = rng.integers(0, o.Q, (2, o.N), dtype=np.uint64)
    c = o.rlwe_encrypt(z, m, 92, first_index=5)
    s_ntt = o.ntt(z)
    for i in range(2):
        e = _noise(o, 92, stream(TAG_RLWE_ENC, 5 + i))
        assert list(_phase(o, c[i], s_ntt)) == [int(v) % o.Q for v in (e + m[i].astype(object))], i


def test_rgsw_encrypt_identity():
    # RGSW encryption (RLWE.cpp:561-577): row h column 0 carries m BG^h on a,
    # so its phase is e - m BG^h s; column 1 carries it on b, phase e + m BG^h
    from oracle import TAG_RGSW_ENC, stream
    o = Oracle(n=4, q=64, N=32, Q=0, nbit=27, BGbits=7)
    z = o.rlwe_secret(93)
    s_ntt = o.ntt(z)
    rng = np.random.default_rng(93)
    m = rng.integers(0, o.Q, o.N, dtype=np.uint64)
    rg = o.rgsw_encrypt(z, m, 94, first_index=2)[0].reshape(o.digitG, 2, 2, o.N)
    Q = o.Q
    ms = o.intt(((o.ntt(m).astype(object) * s_ntt.astype(object)) % Q).astype(np.uint64)).astype(object)
    for h in range(o.digitG):
        sc = pow(2, o.BGbits * h, Q)
        for g in range(2):
            e = _noise(o, 94, stream(TAG_RGSW_ENC, 2 * 2 * o.digitG) + 2 * h + g)
            want = e - sc * ms if g == 0 else e + sc * m.astype(object)
            assert list(_phase(o, rg[h, g], s_ntt)) == [int(v) % Q for v in want], (h, g)


def test_homo_expand_then_cmux_tree_selects():
    # homo_expand (app.cpp:78-97) turns packed bits into RGSW(b_i); a CMUX
    # tree over those RGSWs (test.cpp:1576-1582) selects leaf sum b_l 2^l
    N = 64
    o = Oracle(n=4, q=128, N=N, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Bks_R_bits=9)
    Q = o.Q
    z = o.rlwe_secret(95)
    subs = o.subs_key_gen(z, 96)
    neg_s = np.array([(Q - int(v)) % Q for v in z], dtype=np.uint64)
    enc_sec = o.rgsw_encrypt(z, neg_s, 97)[0]
    rng = np.random.default_rng(95)
    bits = rng.integers(0, 2, N)
    packed = np.zeros((o.digitG, N), dtype=np.uint64)
    for h in range(o.digitG):
        packed[h] = [(int(b) * pow(2, o.BGbits * h, Q) * o.iN) % Q for b in bits]
    rg = o.homo_expand(o.rlwe_encrypt(z, packed, 98), enc_sec, subs)
    levels = 3
    pats = rng.integers(0, 2, (1 << levels, N))
    leaves = o.rlwe_encrypt(z, np.array([[int(v) * (Q // 2) for v in p] for p in pats], dtype=np.uint64), 99)
    out = o.cmux_tree(leaves, rg[:levels])
    want = sum(int(bits[l]) << l for l in range(levels))
    assert list(o.rlwe_decrypt(z, out, 2)[0]) == list(pats[want])
