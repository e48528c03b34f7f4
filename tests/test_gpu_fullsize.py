"""GPU checks at the benchmark's full size (n=500, N=1024, 27-bit Q, Br=32,
batch 4096): every decrypted bit, and a sample of ciphertexts word for word
against the oracle (the oracle takes ~1 s per gate here, so it checks a
sample; the rest is pinned by the size-independent truth-table property).
Also runs the reference-shaped C++ host API (host/gate_demo.cpp)."""
import os
import subprocess

import numpy as np
import pytest

from fheacc_testlib import CFG0, CFG1, ROOT, fhe
from oracle import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def cfg0():
    ctx = fhe.Context(**CFG0)
    s, z = ctx.lwe_secret(123), ctx.rlwe_secret(123)
    keys = ctx.keys()
    keys.generate(123, s, z)
    return ctx, keys, s


def test_full_batch_truth_table_and_sample_parity(cfg0):
    ctx, keys, s = cfg0
    B = 4096
    rng = np.random.default_rng(7)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    c1 = ctx.lwe_encrypt(s, m1, 9, 0)
    c2 = ctx.lwe_encrypt(s, m2, 9, B)
    out, acc, ks = ctx.bootstrap(keys, gates, c1, c2, capture=True)
    truth = {0: m1 | m2, 1: m1 & m2, 2: 1 - (m1 | m2), 3: 1 - (m1 & m2), 4: m1 ^ m2, 5: 1 - (m1 ^ m2)}
    exp = np.array([truth[int(g)][i] for i, g in enumerate(gates)])
    assert np.array_equal(ctx.lwe_decrypt(s, out), exp)
    # outputs are LWE mod q words; the mod switch may return q itself (LWE.cpp:255-260)
    assert out.max() <= ctx.q
    o = Oracle(**CFG0)
    idx = np.array([0, 1, 2, 3, 1000, 2047, 4094, 4095])
    bsk, ksk = keys.download_bsk(), keys.download_ksk()
    eo, ea, ek = o.bootstrap(c1[idx], c2[idx], gates[idx], bsk, ksk, capture=True)
    assert np.array_equal(acc[idx], ea)
    assert np.array_equal(ks[idx], ek)
    assert np.array_equal(out[idx], eo)


def test_nand_1000_every_word(cfg0):
    # the benchmark parameters' parity case: NAND on 1000 seeded uniform bit
    # pairs, every accumulator word after the blind rotation, every word after
    # the key switch and after the mod switch, against the oracle (16 host
    # threads, the GPU host's CPU share), and every decrypted bit
    ctx, keys, s = cfg0
    B = 1000
    rng = np.random.default_rng(1000)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    c1 = ctx.lwe_encrypt(s, m1, 1001, 0)
    c2 = ctx.lwe_encrypt(s, m2, 1001, B)
    gates = np.full(B, 3, dtype=np.int32)
    out, acc, ks = ctx.bootstrap(keys, gates, c1, c2, capture=True)
    assert np.array_equal(ctx.lwe_decrypt(s, out), 1 - (m1 & m2))
    o = Oracle(**CFG0)
    eo, ea, ek = o.bootstrap(c1, c2, gates, keys.download_bsk(), keys.download_ksk(), threads=16, capture=True)
    assert np.array_equal(acc, ea)
    assert np.array_equal(ks, ek)
    assert np.array_equal(out, eo)


def test_microbench_shape_every_word():
    # the 54-bit microbench shape at its own size: 4096 RLWE at N = 1024,
    # BG = 2^9 d = 6, one shared RGSW; every word of the forward NTT of all
    # 8192 polys, the inverse round trip, and all 4096 external products
    # against the oracle
    kw = CFG1
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    rng = np.random.default_rng(4096)
    B, N = 4096, o.N
    x = rng.integers(0, o.Q, (B, 2, N), dtype=np.uint64)
    t = torch.from_numpy(x.view(np.int64).copy()).cuda()
    ctx.ntt_dev(t)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint64), o.ntt(x))
    ctx.ntt_dev(t, inverse=True)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint64), x)
    rgsw = rng.integers(0, o.Q, o.rgsw_words, dtype=np.uint64)
    out = torch.zeros_like(t)
    ctx.external_product_dev(t, torch.from_numpy(rgsw.view(np.int64).copy()).cuda(), out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint64), o.mult_rgsw(x, rgsw))


def test_batch_equals_single(cfg0):
    # the same gate gives the same words alone and inside a batch
    ctx, keys, s = cfg0
    B = 64
    rng = np.random.default_rng(8)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    c1, c2 = ctx.lwe_encrypt(s, m1, 10, 0), ctx.lwe_encrypt(s, m2, 10, B)
    gates = np.full(B, 3, dtype=np.int32)
    full = ctx.bootstrap(keys, gates, c1, c2)
    for i in (0, 17, 63):
        one = ctx.bootstrap(keys, gates[i:i + 1], c1[i:i + 1], c2[i:i + 1])
        assert np.array_equal(one[0], full[i])


def test_host_api_gate_demo():
    exe = os.path.join(ROOT, "3rd_gen_fhe_acc_amd", "lib", "gate_demo")
    r = subprocess.run([exe, "2"], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL GATES PASSED" in r.stdout


# the reference's own application parameter sets, full size
APP_SETS = {
    # apps/fhe_wo_fpga.cpp:35 and apps/bootstrap_fpga_128_ap.cpp:43 (binary LWE secret here)
    "app_128": dict(n=512, q=512, N=1024, Q=0, nbit=27, BGbits=9, Br=23, Bks=25),
    # apps/bootstrap_fpga_medium.cpp:43
    "app_medium": dict(n=256, q=512, N=1024, Q=0, nbit=27, BGbits=9, Br=23, Bks=25),
    # apps/bootstrap_fpga_256.cpp:43-135: 29-bit Q at N = 2048, 32-bit words with Harvey's butterflies
    "app_256": dict(n=1024, q=1024, N=2048, Q=0, nbit=29, BGbits=10, Br=32, Bks=25),
}


@pytest.mark.parametrize("name", list(APP_SETS))
def test_reference_app_sets(name):
    # every gate's truth table over a batch, and every word of a few
    # bootstraps against the oracle, at the apps' own sizes
    kw = APP_SETS[name]
    ctx = fhe.Context(**kw)
    s, z = ctx.lwe_secret(321), ctx.rlwe_secret(321)
    keys = ctx.keys()
    keys.generate(322, s, z)
    B = 1024
    rng = np.random.default_rng(323)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    c1, c2 = ctx.lwe_encrypt(s, m1, 324, 0), ctx.lwe_encrypt(s, m2, 324, B)
    out, acc, ks = ctx.bootstrap(keys, gates, c1, c2, capture=True)
    truth = {0: m1 | m2, 1: m1 & m2, 2: 1 - (m1 | m2), 3: 1 - (m1 & m2), 4: m1 ^ m2, 5: 1 - (m1 ^ m2)}
    exp = np.array([truth[int(g)][i] for i, g in enumerate(gates)])
    assert np.array_equal(ctx.lwe_decrypt(s, out), exp)
    o = Oracle(**kw)
    idx = np.array([0, 511, 1023])
    eo, ea, ek = o.bootstrap(c1[idx], c2[idx], gates[idx], keys.download_bsk(), keys.download_ksk(), capture=True)
    assert np.array_equal(acc[idx], ea) and np.array_equal(ks[idx], ek) and np.array_equal(out[idx], eo)


def test_empty_and_ragged_batches(cfg0):
    # B = 0 is a no-op; odd sizes around the key-switch staging group (256)
    # and the gather threshold (32) give the words of the full batch
    ctx, keys, s = cfg0
    assert ctx.bootstrap(keys, np.zeros(0, dtype=np.int32), np.zeros((0, ctx.n + 1), dtype=np.uint64),
                         np.zeros((0, ctx.n + 1), dtype=np.uint64)).shape == (0, ctx.n + 1)
    B = 515
    rng = np.random.default_rng(9)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    c1, c2 = ctx.lwe_encrypt(s, m1, 11, 0), ctx.lwe_encrypt(s, m2, 11, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    full = ctx.bootstrap(keys, gates, c1, c2)
    for lo, hi in ((0, 1), (1, 33), (33, 290), (290, 515)):
        assert np.array_equal(ctx.bootstrap(keys, gates[lo:hi], c1[lo:hi], c2[lo:hi]), full[lo:hi]), (lo, hi)
