"""Extract the reference's known-answer dumps into compact fixtures.

Reads (in this container only) the two text dumps the reference keeps for its
RTL verification and the two parameter files beside them:

  cl_fhe_acc/verif/tests/mem_init_content/top_verify/bootstrap/bk/RLWExRGSW_output.txt
      one RLWE x RGSW external product, N=1024, Q=0x7FFF801, BG=2^9, d=3
      (written by test::RLWE_mult_RGSW, FHE_ACC_runtime/src/test.cpp:627-740,
      with the dump prints of RLWE.cpp:649-700 enabled)
  cl_fhe_acc/verif/tests/mem_init_content/top_verify/rlwesubs/bk/RLWESUBS_output.txt
      one RLWE_substitute(c0, 0) + RLWE_keyswitch, N=2048, Q=0x3FFFFFFFFED001,
      Bks_R=2^9, digitKS_R=6 (test.cpp:1288-1345 and RLWE.cpp:480-504 prints)
  cl_fhe_acc/verif/tests/mem_init_content/bk/param_{1k,2k}.txt

and writes tests/golden/{rlwe_x_rgsw_n1024.npz, rlwesubs_n2048.npz,
params.json}.  The fixtures are data only (inputs and expected outputs).
"""
import json
import os
import re
import sys

import numpy as np

REF = "/root/reference/cl_fhe_acc/verif/tests/mem_init_content"
HERE = os.path.dirname(os.path.abspath(__file__))
HEX = re.compile(r"^[0-9A-F]{14}$")


def blocks(path):
    """Split a dump into (label, [words]) blocks: every run of 14-hex-digit
    lines is attached to the nearest preceding text line."""
    out = []
    label, words = None, []
    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if HEX.match(line):
                words.append(int(line, 16))
            else:
                if words:
                    out.append((label, words))
                label, words = line, []
    if words:
        out.append((label, words))
    return out


def take(bl, label, n, start=0):
    for i in range(start, len(bl)):
        if bl[i][0] == label and len(bl[i][1]) == n:
            return i, np.array(bl[i][1], dtype=np.uint64)
    raise KeyError(label)


def rlwe_x_rgsw():
    path = os.path.join(REF, "top_verify/bootstrap/bk/RLWExRGSW_output.txt")
    bl = blocks(path)
    N = 1024
    i, a = take(bl, "RLWE poly a", N)
    i, b = take(bl, "RLWE poly b", N, i + 1)
    i, at = take(bl, "RLWE poly a time domain", N, i + 1)
    i, bt = take(bl, "RLWE poly b time domain", N, i + 1)
    # input RGSW: "RLWE h, g" headers precede the poly a/b blocks, g outer
    rgsw = np.zeros((3, 2, 2, N), dtype=np.uint64)
    pos = i + 1
    for g in range(2):
        for h in range(3):
            pos, pa = take(bl, "RLWE poly a", N, pos)
            pos, pb = take(bl, "RLWE poly b", N, pos + 1)
            rgsw[h, g, 0], rgsw[h, g, 1] = pa, pb
            pos += 1
    dec = np.zeros((3, 2, N), dtype=np.uint64)     # [h][a|b] NTT-domain digits
    run = np.zeros((2, 3, 2, N), dtype=np.uint64)  # [loop a|b][h][acc a|b]
    for which, tag in ((0, "a"), (1, "b")):
        for h in range(3):
            pos, d = take(bl, "decomposed %d RLWE poly %s" % (h, tag), N, pos)
            dec[h, which] = d
            pos, ka = take(bl, "RGSW %d, %d poly a" % (h, which), N, pos + 1)
            pos, kb = take(bl, "RGSW %d, %d poly b" % (h, which), N, pos + 1)
            assert np.array_equal(ka, rgsw[h, which, 0]) and np.array_equal(kb, rgsw[h, which, 1])
            pos, ra = take(bl, "tmp acc%d, 0 poly a" % h, N, pos + 1)
            pos, rb = take(bl, "tmp acc%d, 0 poly b" % h, N, pos + 1)
            run[which, h, 0], run[which, h, 1] = ra, rb
            pos += 1
    pos, oa = take(bl, "RLWE poly a", N, pos)
    pos, ob = take(bl, "RLWE poly b", N, pos + 1)
    np.savez_compressed(
        os.path.join(HERE, "rlwe_x_rgsw_n1024.npz"),
        Q=np.uint64(0x7FFF801), N=np.int64(N), BGbits=np.int64(9), digitG=np.int64(3),
        acc=np.stack([a, b]), acc_time=np.stack([at, bt]), rgsw=rgsw, dec=dec,
        running=run, out=np.stack([oa, ob]))


def rlwesubs():
    path = os.path.join(REF, "top_verify/rlwesubs/bk/RLWESUBS_output.txt")
    bl = blocks(path)
    N = 2048
    pos = 0
    ks = np.zeros((6, 2, N), dtype=np.uint64)
    for l in range(6):
        pos, ka = take(bl, "poly a", N, pos)
        pos, kb = take(bl, "poly b", N, pos + 1)
        ks[l, 0], ks[l, 1] = ka, kb
        pos += 1
    pos, ia = take(bl, "poly a", N, pos)
    pos, ib = take(bl, "poly b", N, pos + 1)
    pos, ta = take(bl, "poly a", N, pos + 1)
    pos, tb = take(bl, "poly b", N, pos + 1)
    pos, sa = take(bl, "poly a", N, pos + 1)
    pos, sb = take(bl, "poly b", N, pos + 1)
    dec = np.zeros((6, N), dtype=np.uint64)
    mult = np.zeros((6, 2, N), dtype=np.uint64)
    acc = np.zeros((6, 2, N), dtype=np.uint64)
    pos += 1
    for l in range(6):
        pos, d = take(bl, "decomposed poly a %d" % l, N, pos)
        pos, ma = take(bl, "multed RLWE poly a %d" % l, N, pos + 1)
        pos, mb = take(bl, "multed RLWE poly b %d" % l, N, pos + 1)
        pos, aa = take(bl, "acced RLWE poly a %d" % l, N, pos + 1)
        pos, ab = take(bl, "acced RLWE poly b %d" % l, N, pos + 1)
        dec[l], mult[l, 0], mult[l, 1], acc[l, 0], acc[l, 1] = d, ma, mb, aa, ab
        pos += 1
    pos, oa = take(bl, "poly a", N, pos)
    pos, ob = take(bl, "poly b", N, pos + 1)
    np.savez_compressed(
        os.path.join(HERE, "rlwesubs_n2048.npz"),
        Q=np.uint64(0x3FFFFFFFFED001), N=np.int64(N), ROU=np.uint64(2604308523238),
        Bks_R_bits=np.int64(9), digitKS_R=np.int64(6), subs_factor=np.int64(0),
        kskey=ks, inp=np.stack([ia, ib]), inp_time=np.stack([ta, tb]),
        subs=np.stack([sa, sb]), dec=dec, mult=mult, running=acc, out=np.stack([oa, ob]))


def params():
    res = {}
    for name in ("param_1k.txt", "param_2k.txt"):
        vals = {}
        with open(os.path.join(REF, "bk", name)) as f:
            for line in f:
                m = re.match(r"\s*(\w+)\s*=\s*([0-9A-Fa-f]+)", line)
                if m:
                    key, v = m.group(1), m.group(2)
                    vals[key] = int(v, 10) if key in ("k",) or (key == "m" and name == "param_2k.txt") else int(v, 16)
        res[name] = vals
    with open(os.path.join(HERE, "params.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference dumps not present; fixtures are already committed")
    rlwe_x_rgsw()
    rlwesubs()
    params()
    print("fixtures written to", HERE)
