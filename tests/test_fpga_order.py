"""The accelerator's blind-rotation orders in the oracle (fhe_oracle.c
orc_blind_rotate with eval_mode): fpga_eval_bootstrap_x1 multiplies at step
(0, 0) whatever the residue, the x4..x16 forms at every step with btkey[0] =
RGSW(1) for residue 0 (myfpga.cpp:1030-1071, 1180-1225).

Pins: the trivial-key closed form (fpga_order_cases.py), the same order
driven through the reference's own routines compiled from its sources
(oracle/_ref), and decryption of the gate outputs."""
import numpy as np
import pytest

from fpga_order_cases import MODES, EVAL_FPGA_X1, EVAL_FPGA_XN, EVAL_SOFTWARE, expected_acc, lwe_pairs, \
    steps_taken, trivial_bsk
from oracle import GATES, Oracle, refbind

SETS = {
    # Br = 4: digitR = 5, a quarter of all steps have residue 0
    "n16_N256_q27_br4": dict(n=16, q=512, N=256, nbit=27, BGbits=9, Br=4, Bks=25),
    # wide words, Br = 16
    "n10_N128_q54": dict(n=10, q=256, N=128, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=16, Bks=32),
}


def eval_rows(o, c1, c2, gates):
    return np.stack([o.evaluate(c1[b], c2[b], int(gates[b])) for b in range(len(gates))])


@pytest.mark.parametrize("name", list(SETS))
@pytest.mark.parametrize("mode", MODES)
def test_trivial_keys_closed_form(name, mode):
    o = Oracle(**SETS[name])
    o.eval_mode = mode
    bsk, exps = trivial_bsk(o, 3)
    ksk = np.zeros(o.ksk_words, dtype=np.uint64)
    B = 6
    c1, c2 = lwe_pairs(o, B, 4)
    gates = np.arange(B, dtype=np.int32) % 6
    _, acc, _ = o.bootstrap(c1, c2, gates, bsk, ksk, capture=True)
    ev = eval_rows(o, c1, c2, gates)
    for b in range(B):
        assert np.array_equal(acc[b], expected_acc(o, ev[b], int(gates[b]), exps, mode)), b


def test_orders_differ_on_these_inputs():
    # the closed-form test above can tell the three orders apart: the pairs
    # with a[0] = 0 take a residue-0 step at (0, 0), and every pair takes
    # residue-0 steps later on
    o = Oracle(**SETS["n16_N256_q27_br4"])
    c1, c2 = lwe_pairs(o, 6, 4)
    ev = eval_rows(o, c1, c2, np.arange(6) % 6)
    for b in range(2):
        s0, s1, s2 = (steps_taken(o, ev[b], m) for m in MODES)
        assert (0, 0, 0) in s1 and (0, 0, 0) not in s0
        assert len(s0) < len(s1) < len(s2) == o.n * o.digitR
    o.eval_mode = EVAL_SOFTWARE
    assert o.eval_mode == 0
    with pytest.raises(ValueError):
        o.eval_mode = 3


def small_keys(o, seed):
    s = o.lwe_secret(seed, binary=True)
    z = o.rlwe_secret(seed)
    return s, z, o.bsk_gen(s, z, seed), o.ksk_gen(z, s, seed)


@pytest.mark.parametrize("name", list(SETS))
@pytest.mark.parametrize("mode", [EVAL_FPGA_X1, EVAL_FPGA_XN])
def test_fpga_orders_decrypt(name, mode):
    # the extra products are by encryptions of 1: outputs still decrypt to
    # the gate's truth table (test.cpp:496-625)
    o = Oracle(**SETS[name])
    o.eval_mode = mode
    s, _, bsk, ksk = small_keys(o, 12)
    m1 = np.array([0, 0, 1, 1] * 3)
    m2 = np.array([0, 1, 0, 1] * 3)
    c1 = o.lwe_encrypt(s, m1, seed=5, first_index=0)
    c2 = o.lwe_encrypt(s, m2, seed=5, first_index=100)
    c1[:2, 0] = 0
    c2[:2, 0] = 0
    for gate in GATES:
        got = o.decrypt(s, o.bootstrap(c1, c2, GATES[gate], bsk, ksk))
        table = {"OR": m1 | m2, "AND": m1 & m2, "NOR": 1 - (m1 | m2), "NAND": 1 - (m1 & m2),
                 "XOR": m1 ^ m2, "XNOR": 1 - (m1 ^ m2)}[gate]
        assert list(got) == list(table), gate


@pytest.mark.ref
@pytest.mark.skipif(not refbind.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", list(SETS))
@pytest.mark.parametrize("mode", [EVAL_FPGA_X1, EVAL_FPGA_XN])
def test_fpga_orders_match_reference_routines(name, mode):
    kw = SETS[name]
    o = Oracle(**kw)
    o.eval_mode = mode
    r = refbind.Reference(kw["n"], kw["q"], kw["N"], o.Q, kw["BGbits"], kw["Br"], kw["Bks"])
    _, _, bsk, ksk = small_keys(o, 13)
    r.load_keys(bsk, ksk)
    B = 6
    c1, c2 = lwe_pairs(o, B, 14)
    gates = np.arange(B, dtype=np.int32) % 6
    out, acc, _ = o.bootstrap(c1, c2, gates, bsk, ksk, capture=True)
    rout, racc = r.bootstrap_fpga(c1, c2, gates, mode)
    assert np.array_equal(acc, racc)
    assert np.array_equal(out, rout)
    # and the order matters: the software order's words differ
    o.eval_mode = EVAL_SOFTWARE
    assert not np.array_equal(o.bootstrap(c1, c2, gates, bsk, ksk), out)
