"""Cases for the accelerator's blind-rotation orders (myfpga.cpp:1003-1660),
shared by the CPU and GPU tests.

The closed form used as the pin: a noiseless RGSW of a monomial X^e (zero
masks, zero noise: row h is (BG^h X^e, 0) and (0, BG^h X^e) in the NTT
domain, RLWE.cpp:561-577 without the encryptions of zero) makes the external
product exact: the gadget digits recompose exactly (digitG * BGbits >= log2 Q,
RLWE.cpp:623-640), so acc (x) RGSW(X^e) = X^e * acc word for word.  A
bootstrap key made only of such RGSWs, each (r, j, i) with its own random
exponent, turns the whole blind rotation into one negacyclic rotation of the
test polynomial by the sum of the exponents of the steps actually taken.
Which steps those are is exactly what the evaluation order decides, and the
residue-0 exponents are what tell a wrong order or a wrong residue-0 key
apart."""
import numpy as np

EVAL_SOFTWARE, EVAL_FPGA_X1, EVAL_FPGA_XN = 0, 1, 2
MODES = (EVAL_SOFTWARE, EVAL_FPGA_X1, EVAL_FPGA_XN)


def monomial_ntt_rows(o, coef):
    """rows[e] = NTT(coef * X^e) for e in [0, 2N): X^e = -X^(e-N) for e >= N."""
    N, Q = o.N, o.Q
    t = np.zeros((2 * N, N), dtype=np.uint64)
    e = np.arange(2 * N)
    t[e, e % N] = np.where(e < N, coef % Q, (Q - coef % Q) % Q).astype(np.uint64)
    return o.ntt(t).reshape(2 * N, N)


def trivial_bsk(o, seed):
    """Reference-layout bootstrap key btkey[Br][digitR][n] of noiseless RGSWs
    of X^e[r, j, i], e uniform in [0, 2N).  Returns (bsk words, e)."""
    rng = np.random.default_rng(seed)
    N, d = o.N, o.digitG
    exps = rng.integers(0, 2 * N, (o.Br, o.digitR, o.n))
    bsk = np.zeros((o.Br, o.digitR, o.n, d, 2, 2, N), dtype=np.uint64)
    for h in range(d):
        rows = monomial_ntt_rows(o, pow(2, o.BGbits * h, o.Q))
        m = rows[exps]
        bsk[:, :, :, h, 0, 0, :] = m  # row (h, g=0): message on a
        bsk[:, :, :, h, 1, 1, :] = m  # row (h, g=1): message on b
    return bsk.reshape(-1), exps


def steps_taken(o, a_eval, mode):
    """(i, j, r) of every external product the order performs for the
    evaluated LWE mask a_eval (RLWE.cpp:754-767; myfpga.cpp:1030-1071 for x1,
    1180-1225 for the batched forms)."""
    out = []
    for i in range(o.n):
        nega = (o.q - int(a_eval[i])) % o.q
        for j in range(o.digitR):
            r = nega % o.Br
            if r or mode == EVAL_FPGA_XN or (mode == EVAL_FPGA_X1 and i == 0 and j == 0):
                out.append((i, j, r))
            nega //= o.Br
    return out


def negacyclic_rotate(t, E, N, Q):
    """X^E * t(X) mod (X^N + 1, Q) for E in [0, 2N)."""
    i = np.arange(N)
    idx = i + E
    neg = (idx // N) % 2 == 1
    out = np.zeros(N, dtype=np.uint64)
    vals = np.where(neg, (Q - t.astype(object)) % Q, t.astype(object)).astype(np.uint64)
    out[idx % N] = vals
    return out


def expected_acc(o, c_eval, gate, exps, mode):
    """Accumulator after the blind rotation under a trivial_bsk key: the test
    polynomial (RLWE.cpp:706-744) rotated by the applied exponents."""
    N, Q, q = o.N, o.Q, o.q
    E = sum(int(exps[r, j, i]) for i, j, r in steps_taken(o, c_eval, mode)) % (2 * N)
    b1 = int(o.p.gate_constant[gate]) % q
    b2 = (b1 + q // 2) % q
    init = o.acc_init(b1, b2, int(c_eval[o.n]))
    t = o.intt(init[1])
    acc = np.zeros((2, N), dtype=np.uint64)
    acc[1] = o.ntt(negacyclic_rotate(t, E, N, Q))
    return acc


def lwe_pairs(o, B, seed, zero_first=2):
    """B operand pairs with uniform words mod q.  The first zero_first pairs
    have a[0] = 0 in both operands, so the evaluated a[0] is 0 for every gate
    and step (0, 0) has residue 0 (the x1 order's fused INIT case)."""
    rng = np.random.default_rng(seed)
    c1 = rng.integers(0, o.q, (B, o.n + 1), dtype=np.uint64)
    c2 = rng.integers(0, o.q, (B, o.n + 1), dtype=np.uint64)
    c1[:zero_first, 0] = 0
    c2[:zero_first, 0] = 0
    return c1, c2
