"""Pin the oracle against the reference's own known-answer vectors.

Fixtures (tests/golden/, written by make_golden.py from the reference's RTL
verification dumps):
  rlwe_x_rgsw_n1024.npz  -- one external product, N=1024, Q=0x7FFF801, BG=2^9, d=3
  rlwesubs_n2048.npz     -- RLWE_substitute + RLWE_keyswitch, N=2048, 54-bit Q
  params.json            -- Q, N^-1, psi, Barrett m for both primes
"""
import json
import os

import numpy as np
import pytest

from oracle import Oracle, lib


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def test_params_1k_2k(golden_dir):
    with open(os.path.join(golden_dir, "params.json")) as f:
        pr = json.load(f)
    L = lib()
    # param_1k: N=1024, nbit=27; the reference derivation (fhe_wo_fpga.cpp:89)
    Q = L.orc_previous_prime(L.orc_first_prime(27, 1024), 1024)
    assert Q == pr["param_1k.txt"]["Q"] == 0x7FFF801
    assert L.orc_inverse_mod(1024, Q) == pr["param_1k.txt"]["iN"]
    assert L.orc_root_of_unity(1024, Q) == pr["param_1k.txt"]["ROU"]
    assert (1 << 54) // Q == pr["param_1k.txt"]["m"]
    # param_2k: N=2048, nbit=54
    Q2 = L.orc_previous_prime(L.orc_first_prime(54, 2048), 2048)
    assert Q2 == pr["param_2k.txt"]["Q"] == 0x3FFFFFFFFED001
    assert L.orc_inverse_mod(2048, Q2) == pr["param_2k.txt"]["iN"]
    assert L.orc_root_of_unity(2048, Q2) == pr["param_2k.txt"]["ROU"]
    assert (1 << 108) // Q2 == pr["param_2k.txt"]["m"]


def ext_oracle(g):
    # n, q, Br and Bks do not enter the external product
    return Oracle(n=8, q=512, N=int(g["N"]), Q=int(g["Q"]), BGbits=int(g["BGbits"]))


def test_rlwe_x_rgsw_transforms(golden_dir):
    g = load(golden_dir, "rlwe_x_rgsw_n1024.npz")
    o = ext_oracle(g)
    assert o.digitG == int(g["digitG"]) == 3
    assert np.array_equal(o.intt(g["acc"]), g["acc_time"])
    assert np.array_equal(o.ntt(g["acc_time"]), g["acc"])


def test_rlwe_x_rgsw_acc_init(golden_dir):
    # The dump's input accumulator is acc_initialization with b=0, q=512,
    # bound1 = 3q/8 (NAND), bound2 = bound1 + q/2 (mem_init_content README).
    g = load(golden_dir, "rlwe_x_rgsw_n1024.npz")
    o = Oracle(n=8, q=512, N=1024, Q=int(g["Q"]), BGbits=9)
    b1 = 3 * (512 >> 3)
    acc = o.acc_init(b1, (b1 + 256) % 512, 0)
    assert np.array_equal(acc, g["acc"])


def test_rlwe_x_rgsw_digits_and_mac(golden_dir):
    g = load(golden_dir, "rlwe_x_rgsw_n1024.npz")
    o = ext_oracle(g)
    dec = o.decompose(g["acc"], 3, 9)
    assert np.array_equal(dec, g["dec"])
    # running sums: loop over a-digits (column 0), then b-digits (column 1)
    Q = int(g["Q"])
    acc = np.zeros((2, 1024), dtype=object)
    for which in range(2):
        for h in range(3):
            d = dec[h, which].astype(object)
            for half in range(2):
                acc[half] = (acc[half] + g["rgsw"][h, which, half].astype(object) * d) % Q
            assert np.array_equal(acc.astype(np.uint64), g["running"][which, h])
    out = o.mult_rgsw(g["acc"], g["rgsw"])
    assert np.array_equal(out, g["out"])


def test_rlwesubs(golden_dir):
    g = load(golden_dir, "rlwesubs_n2048.npz")
    o = Oracle(n=8, q=512, N=2048, Q=int(g["Q"]), BGbits=9, Bks_R_bits=int(g["Bks_R_bits"]))
    assert o.ROU == int(g["ROU"])
    assert o.digitKS_R == int(g["digitKS_R"]) == 6
    assert np.array_equal(o.intt(g["inp"]), g["inp_time"])
    assert np.array_equal(o.ntt(g["inp_time"]), g["inp"])
    subs = o.substitute(g["inp"], int(g["subs_factor"]))
    assert np.array_equal(subs, g["subs"])
    dec = o.decompose(subs, 6, 9)
    assert np.array_equal(dec[:, 0], g["dec"])
    Q = int(g["Q"])
    run = np.zeros((2, 2048), dtype=object)
    for l in range(6):
        for half in range(2):
            m = (g["kskey"][l, half].astype(object) * dec[l, 0].astype(object)) % Q
            assert np.array_equal(m.astype(np.uint64), g["mult"][l, half])
            run[half] = (run[half] + m) % Q
        assert np.array_equal(run.astype(np.uint64), g["running"][l])
    out = o.rlwe_keyswitch(subs, g["kskey"])
    assert np.array_equal(out, g["out"])
