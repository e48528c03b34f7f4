"""The checked library, lib/libfheacc_checked.so, over every product form of
the gate path and over the standalone NTT and external product.  That build
checks each LDS access of the transforms and external products against the
workgroup's allocation, and each gate workgroup's ciphertext against its
launch.  No check may fire, and every output word must equal the default
library's (the checks change no result; the default library's words are
pinned to the oracle by the parity tests).  The checked library runs in one
child process (FHEACC_LIBRARY=checked), so that its handles never mix with
the default library's in this process."""
import os
import subprocess
import sys

import numpy as np
import pytest

from fheacc_testlib import ROOT, SETS, fhe

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# parameter set -> launch shapes (threads, digit polys per LDS batch): pb = 2d
# is the in-place product, pb = d the split one, any other pb the separate
# digit buffer; (256, 8) at N = 1024 is the benchmark kernel (its forward
# transform's last stages in the multiply-accumulate)
CASES = [
    ("n12_N1024_q27", [(256, 8), (1024, 8), (256, 4), (256, 2)]),
    ("n8_N2048_q54", [(1024, 6), (512, 3), (256, 2)]),
    ("n8_N512_q27_bg6", [(256, 10)]),
    ("n12_N512_q27", [(512, 8), (128, 3)]),
    ("n8_N2048_q29", [(256, 6), (1024, 6)]),
    ("n16_N256_q27", [None]),
    ("n10_N128_q54", [None]),
]
B = 24


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).cuda()


def _host(t):
    return t.cpu().numpy().view(np.uint64)


def run_cases():
    """Every case above through the loaded library: name -> output words."""
    out = {}
    for name, shapes in CASES:
        ctx = fhe.Context(**SETS[name])
        s, z = ctx.lwe_secret(21), ctx.rlwe_secret(21)
        keys = ctx.keys()
        keys.generate(21, s, z)
        rng = np.random.default_rng(21)
        m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
        c1, c2 = ctx.lwe_encrypt(s, m1, 21, 0), ctx.lwe_encrypt(s, m2, 21, B)
        gates = rng.integers(0, 6, B).astype(np.int32)
        for shape in shapes:
            print("device checks:", name, shape, flush=True)
            if shape:
                ctx.set_tuning(*shape)
            res, acc, ks = ctx.bootstrap(keys, gates, c1, c2, capture=True)
            tag = "%s_%s" % (name, "x".join(map(str, shape)) if shape else "default")
            out[tag], out[tag + "_acc"], out[tag + "_ks"] = res, acc, ks
        keys.close()
        ctx.close()
    # the standalone transforms (an odd and an even count: the even one takes
    # the two-polys-per-workgroup kernel from N = 1024) and external product
    for N, nbit, Q, BGbits in ((1024, 54, 0x3FFFFFFFFED001, 9), (2048, 54, 0x3FFFFFFFFED001, 18), (1024, 27, 0, 7)):
        print("device checks: ntt / external product", N, nbit, flush=True)
        ctx = fhe.Context(n=4, q=1024, N=N, Q=Q, nbit=nbit, BGbits=BGbits)
        rng = np.random.default_rng(N + nbit)
        for cnt in (37, 36):
            x = rng.integers(0, ctx.Q, (cnt, N), dtype=np.uint64)
            t = _dev(x)
            ctx.ntt_dev(t)
            out["ntt_%d_%d_%d" % (N, nbit, cnt)] = _host(t)
            ctx.ntt_dev(t, inverse=True)
            out["intt_%d_%d_%d" % (N, nbit, cnt)] = _host(t)
        acc = rng.integers(0, ctx.Q, (64, 2, N), dtype=np.uint64)
        rg = rng.integers(0, ctx.Q, ctx.digitG * 4 * N, dtype=np.uint64)
        o = torch.zeros((64, 2, N), dtype=torch.int64, device="cuda")
        ctx.external_product_dev(_dev(acc), _dev(rg), o)
        ctx.sync()
        out["xprod_%d_%d" % (N, nbit)] = _host(o)
        ctx.close()
    # the leveled operations: RLWE expansion (the substitution permutes LDS
    # words by a caller-given factor), one substitution key switch, a CMUX
    for kw in (dict(n=4, q=512, N=256, Q=0, nbit=27, BGbits=7, Br=16, Bks=32, Bks_R_bits=7),
               dict(n=512, q=512, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=9, Br=23, Bks=25, Bks_R_bits=9)):
        N = kw["N"]
        print("device checks: expansion / subs / cmux", N, kw["nbit"], flush=True)
        ctx = fhe.Context(**kw)
        keys = ctx.keys()
        keys.generate_subs(33, ctx.rlwe_secret(32))
        rng = np.random.default_rng(N)
        x = rng.integers(0, ctx.Q, (1, 2, N), dtype=np.uint64)
        out["expand_%d" % N] = ctx.expand_rlwe(keys, x)
        out["subs_%d" % N] = ctx.subs_keyswitch(keys, x, 3)
        c0, c1 = (rng.integers(0, ctx.Q, (2, 2, N), dtype=np.uint64) for _ in range(2))
        rg = rng.integers(0, ctx.Q, ctx.digitG * 4 * N, dtype=np.uint64)
        out["cmux_%d" % N] = ctx.cmux(c0, c1, rg)
        keys.close()
        ctx.close()
    return out


def child_main(path):
    """Run in the child process, on the checked library: the words, then the check bits."""
    assert fhe.CHECKED and fhe.device_checks_enabled()
    fhe.device_check_read(0)  # start from a clear word
    res = run_cases()
    res["check_bits"] = np.array([fhe.device_check_read(0)], dtype=np.uint64)
    np.savez(path, **res)
    print("device checks: child done", flush=True)


def test_checked_library_fires_no_check_and_matches(tmp_path):
    if not os.path.exists(os.path.join(ROOT, "3rd_gen_fhe_acc_amd", "lib", "libfheacc_checked.so")):
        pytest.fail("libfheacc_checked.so not built: run __graft_entry__.build()")
    assert not fhe.device_checks_enabled()
    assert fhe.device_check_read(0) == 0  # the default library has no checks
    path = str(tmp_path / "checked.npz")
    env = dict(os.environ, FHEACC_LIBRARY="checked")
    code = ("import sys; sys.path.insert(0, %r); import test_gpu_device_checks as t; t.child_main(%r)"
            % (os.path.dirname(os.path.abspath(__file__)), path))
    print("device checks: starting the checked-library child", flush=True)
    r = subprocess.run([sys.executable, "-u", "-c", code], env=env, timeout=600)
    assert r.returncode == 0, "checked-library child exited %d" % r.returncode
    got = dict(np.load(path))
    assert int(got.pop("check_bits")[0]) == 0, "a device check fired"
    exp = run_cases()
    assert set(got) == set(exp)
    for k in sorted(exp):
        assert np.array_equal(got[k], exp[k]), k
