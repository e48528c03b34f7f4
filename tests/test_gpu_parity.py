"""GPU parity: the HIP library through its C ABI against the oracle, word
for word, on seeded inputs at sizes the oracle finishes in seconds, plus the
reference's golden dumps.  Integer arithmetic: the bar is bit-exact."""
import os

import numpy as np
import pytest

from fheacc_testlib import SETS, fhe
from oracle import GATES, Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def pair(name):
    kw = SETS[name]
    return fhe.Context(**kw), Oracle(**kw)


@pytest.mark.parametrize("N,nbit,Q", [(16, 27, 0), (256, 27, 0), (512, 27, 0), (1024, 27, 0), (2048, 27, 0),
                                      (1024, 29, 0), (2048, 29, 0),  # 32-bit words, Harvey's butterflies
                                      (1024, 28, 0xBEB8801),  # lazy 3-stage and strict 4-stage inverse passes
                                      (128, 54, 0x3FFFFFFFFED001), (1024, 54, 0x3FFFFFFFFED001),
                                      (2048, 54, 0x3FFFFFFFFED001)])
def test_ntt_matches_oracle(N, nbit, Q):
    kw = dict(n=4, q=min(2 * N, 1024), N=N, Q=Q, nbit=nbit)
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    rng = np.random.default_rng(N + nbit)
    x = rng.integers(0, o.Q, (37, N), dtype=np.uint64)
    t = dev(x)
    ctx.ntt_dev(t)
    assert np.array_equal(host(t), o.ntt(x))
    ctx.ntt_dev(t, inverse=True)
    assert np.array_equal(host(t), x)
    t2 = dev(x)
    ctx.ntt_dev(t2, inverse=True)
    assert np.array_equal(host(t2), o.intt(x))
    # an even count: from N = 1024 every workgroup then takes two polys
    # (the compile-time-count kernel)
    x4 = x[:36]
    t4 = dev(x4)
    ctx.ntt_dev(t4)
    assert np.array_equal(host(t4), o.ntt(x4))
    ctx.ntt_dev(t4, inverse=True)
    assert np.array_equal(host(t4), x4)


@pytest.mark.parametrize("name,narrow,harvey,lazy_inv", [("n12_N1024_q27", 1, 0, 4), ("n8_N1024_q28", 1, 0, 3),
                                                        ("n8_N2048_q29", 1, 1, 2), ("n8_N1024_q30", 1, 1, 1),
                                                        ("n8_N2048_q54", 0, 0, 8)])
def test_word_width_selection(name, narrow, harvey, lazy_inv):
    # 27- and 28-bit Q: fully lazy 32-bit butterflies; 29/30-bit Q: 32-bit
    # words with Harvey's butterflies; 54-bit Q: 64-bit words.  Inverse NTT
    # passes of up to lazy_inv stages leave their sums unreduced (the
    # 1024-point inverse runs passes of 3, 3 and 4 stages)
    info = fhe.Context(**SETS[name]).info
    assert (info["narrow"], info["harvey"], info["lazy_inverse_stages"]) == (narrow, harvey, lazy_inv)


def test_external_product_golden(golden_dir):
    g = np.load(os.path.join(golden_dir, "rlwe_x_rgsw_n1024.npz"))
    ctx = fhe.Context(n=8, q=512, N=1024, Q=int(g["Q"]), BGbits=9)
    acc = dev(g["acc"])
    out = torch.zeros_like(acc)
    ctx.external_product_dev(acc, dev(g["rgsw"]), out)
    assert np.array_equal(host(out), g["out"])


@pytest.mark.parametrize("N,Q,BGbits", [(1024, 0x3FFFFFFFFED001, 9),   # microbench shape, d=6
                                        (1024, 0, 7),                   # benchmark shape, d=4
                                        (2048, 0x3FFFFFFFFED001, 18),   # set B, d=3
                                        (512, 0, 6),                    # set C, d=5
                                        (2048, 0x1FFF2001, 10),         # 29-bit, Harvey's butterflies
                                        (1024, 0x3FFF7801, 10),         # 30-bit, Harvey's butterflies
                                        (1024, 0xBEB8801, 7)])          # 28-bit, lazy and strict inverse passes
@pytest.mark.parametrize("per_ct", [False, True])
def test_external_product_matches_oracle(N, Q, BGbits, per_ct):
    kw = dict(n=4, q=1024, N=N, Q=Q, nbit=27, BGbits=BGbits)
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    rng = np.random.default_rng(N + BGbits + per_ct)
    B = 6
    acc = rng.integers(0, o.Q, (B, 2, N), dtype=np.uint64)
    rgsw = rng.integers(0, o.Q, ((B if per_ct else 1), o.rgsw_words), dtype=np.uint64)
    out = torch.zeros((B, 2, N), dtype=torch.int64, device="cuda")
    ctx.external_product_dev(dev(acc), dev(rgsw), out, per_ciphertext_key=per_ct)
    exp = o.mult_rgsw(acc, rgsw if per_ct else rgsw[0])
    assert np.array_equal(host(out), exp)


@pytest.mark.parametrize("name", ["n16_N256_q27", "n12_N512_q27", "n10_N128_q54"])
def test_device_keygen_matches_oracle(name):
    ctx, o = pair(name)
    s, z = ctx.lwe_secret(21), ctx.rlwe_secret(21)
    assert np.array_equal(s, o.lwe_secret(21)) and np.array_equal(z, o.rlwe_secret(21))
    keys = ctx.keys()
    keys.generate(77, s, z)
    # every residue, 0 (RGSW(1)) included, word for word
    assert np.array_equal(keys.download_bsk(), o.bsk_gen(s, z, 77))
    assert np.array_equal(keys.download_ksk(), o.ksk_gen(z, s, 77))


@pytest.mark.parametrize("name", ["n16_N256_q27", "n12_N512_q27", "n10_N128_q54"])
def test_key_upload_roundtrip(name):
    ctx, o = pair(name)
    rng = np.random.default_rng(5)
    bsk = rng.integers(0, o.Q, o.bsk_words, dtype=np.uint64)
    ksk = rng.integers(0, o.Q, o.ksk_words, dtype=np.uint64)
    keys = ctx.keys()
    keys.upload_bsk(bsk)
    keys.upload_ksk(ksk)
    assert np.array_equal(keys.download_bsk(), bsk)
    assert np.array_equal(keys.download_ksk(), ksk)
    d1 = keys.digest()
    keys.upload_bsk(bsk)
    assert keys.digest() == d1


def run_bootstrap_case(name, B, seed, upload=False):
    ctx, o = pair(name)
    s = ctx.lwe_secret(seed, binary=(name != "n16_N256_q27"))
    z = ctx.rlwe_secret(seed)
    keys = ctx.keys()
    if upload:
        keys.upload_bsk(o.bsk_gen(s, z, seed))
        keys.upload_ksk(o.ksk_gen(z, s, seed))
    else:
        keys.generate(seed, s, z)
    rng = np.random.default_rng(seed)
    m1 = rng.integers(0, 2, B)
    m2 = rng.integers(0, 2, B)
    c1 = ctx.lwe_encrypt(s, m1, seed, 0)
    c2 = ctx.lwe_encrypt(s, m2, seed, 10 ** 6)
    assert np.array_equal(c1, o.lwe_encrypt(s, m1, seed, 0))
    gates = rng.integers(0, 6, B).astype(np.int32)
    out, acc, ks = ctx.bootstrap(keys, gates, c1, c2, capture=True)
    if upload:
        bsk, ksk = o.bsk_gen(s, z, seed), o.ksk_gen(z, s, seed)
    else:
        bsk, ksk = keys.download_bsk(), keys.download_ksk()
    eo, ea, ek = o.bootstrap(c1, c2, gates, bsk, ksk, capture=True)
    assert np.array_equal(acc, ea)
    assert np.array_equal(ks, ek)
    assert np.array_equal(out, eo)
    truth = {0: m1 | m2, 1: m1 & m2, 2: 1 - (m1 | m2), 3: 1 - (m1 & m2), 4: m1 ^ m2, 5: 1 - (m1 ^ m2)}
    exp_bits = np.array([truth[int(g)][i] for i, g in enumerate(gates)])
    assert np.array_equal(ctx.lwe_decrypt(s, out), exp_bits)


@pytest.mark.parametrize("name", ["n16_N256_q27", "n12_N512_q27", "n10_N128_q54", "n8_N2048_q54",
                                  "n8_N512_q27_bg6", "n8_N2048_q29", "n8_N1024_q30", "n8_N1024_q28"])
def test_bootstrap_matches_oracle(name):
    run_bootstrap_case(name, B=24, seed=3)


@pytest.mark.parametrize("name,shapes", [("n12_N512_q27", (128, 256, 512)), ("n12_N1024_q27", (256, 1024)),
                                         ("n8_N2048_q54", (256, 512, 1024))])
def test_every_launch_shape_bit_exact(name, shapes):
    # the autotuner may pick any shape: all must give the same words
    ctx, o = pair(name)
    s, z = ctx.lwe_secret(12), ctx.rlwe_secret(12)
    keys = ctx.keys()
    keys.generate(12, s, z)
    rng = np.random.default_rng(12)
    m1, m2 = rng.integers(0, 2, 40), rng.integers(0, 2, 40)
    c1, c2 = ctx.lwe_encrypt(s, m1, 12, 0), ctx.lwe_encrypt(s, m2, 12, 40)
    gates = rng.integers(0, 6, 40).astype(np.int32)
    exp = o.bootstrap(c1, c2, gates, keys.download_bsk(), keys.download_ksk())
    # the CMUX chains share the launch shape
    x0, x1 = (rng.integers(0, o.Q, (3, 2, o.N), dtype=np.uint64) for _ in range(2))
    rg = rng.integers(0, o.Q, o.rgsw_words, dtype=np.uint64)
    cm = np.stack([o.cmux(x0[i], x1[i], rg) for i in range(3)])
    for threads in shapes:
        # pb = 2d: the in-place product; pb = d: in place in two batches (the
        # digits of a, then of b); any other pb: a separate digit buffer
        for pb in sorted({1, 2, 3, 4, ctx.digitG, 2 * ctx.digitG}):
            ctx.set_tuning(threads, pb)
            assert np.array_equal(ctx.bootstrap(keys, gates, c1, c2), exp), (threads, pb)
            out = torch.zeros((3, 2, o.N), dtype=torch.int64, device="cuda")
            ctx.cmux_dev(dev(x0), dev(x1), dev(rg), out)
            assert np.array_equal(host(out), cm), ("cmux", threads, pb)
    rep = ctx.autotune(keys, 256)
    assert rep["chosen"]["threads"] in (128, 256, 512, 1024)
    assert np.array_equal(ctx.bootstrap(keys, gates, c1, c2), exp)


def test_graph_replays_match_eager():
    # replay a captured batch several times, with new inputs written into the
    # captured buffers and eager work in between; every replay must equal the
    # eager result for the same inputs
    ctx, o = pair("n12_N512_q27")
    s, z = ctx.lwe_secret(13), ctx.rlwe_secret(13)
    keys = ctx.keys()
    keys.generate(13, s, z)
    B = 64
    rng = np.random.default_rng(13)
    g = torch.zeros(B, dtype=torch.int32, device="cuda")
    d1 = torch.zeros((B, ctx.n + 1), dtype=torch.int64, device="cuda")
    d2 = torch.zeros_like(d1)
    out = torch.zeros_like(d1)
    graph = ctx.capture_bootstrap(keys, g, d1, d2, out)
    for rep in range(3):
        m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
        c1, c2 = ctx.lwe_encrypt(s, m1, 100 + rep, 0), ctx.lwe_encrypt(s, m2, 100 + rep, B)
        gates = rng.integers(0, 6, B).astype(np.int32)
        g.copy_(torch.from_numpy(gates))
        d1.copy_(dev(c1))
        d2.copy_(dev(c2))
        eager = ctx.bootstrap(keys, gates, c1, c2)
        graph.launch()
        ctx.sync()
        assert np.array_equal(host(out), eager), rep
    graph.close()


def test_bootstrap_uploaded_keys():
    run_bootstrap_case("n12_N512_q27", B=9, seed=4, upload=True)


def test_bootstrap_staged_keyswitch_ragged():
    # above the gather threshold and not a multiple of the key-switch tile
    run_bootstrap_case("n12_N512_q27", B=613, seed=6)


def test_bootstrap_single():
    run_bootstrap_case("n16_N256_q27", B=1, seed=8)


def test_identical_operands_rejected():
    ctx, o = pair("n16_N256_q27")
    s, z = ctx.lwe_secret(1), ctx.rlwe_secret(1)
    keys = ctx.keys()
    keys.generate(1, s, z)
    c = ctx.lwe_encrypt(s, [1, 0], 1)
    with pytest.raises(fhe.FheAccError):
        ctx.bootstrap(keys, [GATES["NAND"]] * 2, c, c)
    # the device path flags it at the next sync
    g = torch.tensor([3, 3], dtype=torch.int32, device="cuda")
    dc = dev(c)
    out = torch.zeros_like(dc)
    ctx.bootstrap_dev(keys, g, dc, dc, out)
    with pytest.raises(fhe.FheAccError):
        ctx.sync()
    ctx.sync()  # flag cleared


def test_subs_keyswitch_golden(golden_dir):
    g = np.load(os.path.join(golden_dir, "rlwesubs_n2048.npz"))
    ctx = fhe.Context(n=8, q=1024, N=2048, Q=int(g["Q"]), BGbits=18, Bks_R_bits=int(g["Bks_R_bits"]))
    assert ctx.digitKS_R == 6
    subs = np.random.default_rng(1).integers(0, ctx.Q, (11, 6, 2, 2048), dtype=np.uint64)
    subs[0] = g["kskey"]
    keys = ctx.keys()
    keys.upload_subs(subs)
    assert np.array_equal(keys.download_subs(), subs.reshape(-1))
    x = dev(g["inp"])
    out = torch.zeros_like(x)
    ctx.subs_keyswitch_dev(keys, x, 0, out)
    assert np.array_equal(host(out), g["out"])


@pytest.mark.parametrize("name", ["n16_N256_q27", "n10_N128_q54", "n8_N2048_q29"])
def test_subs_keyswitch_matches_oracle(name):
    ctx, o = pair(name)
    z = ctx.rlwe_secret(9)
    keys = ctx.keys()
    keys.generate_subs(9, z)
    subs = keys.download_subs()
    assert np.array_equal(subs, o.subs_key_gen(z, 9))
    rng = np.random.default_rng(2)
    x = rng.integers(0, o.Q, (3, 2, o.N), dtype=np.uint64)
    per = o.digitKS_R * 2 * o.N
    for f in range(o.N.bit_length() - 1):
        out = torch.zeros((3, 2, o.N), dtype=torch.int64, device="cuda")
        ctx.subs_keyswitch_dev(keys, dev(x), f, out)
        got = host(out)
        for i in range(3):
            exp = o.rlwe_keyswitch(o.substitute(x[i], f), subs[f * per:(f + 1) * per])
            assert np.array_equal(got[i], exp)


def test_key_checkpoint_roundtrip(tmp_path):
    # fheacc_keys_save / fheacc_keys_load: the restored key set is word for
    # word the saved one and gives the same gate outputs; a corrupted or
    # truncated file, or one of another parameter set, is refused
    ctx, o = pair("n12_N512_q27")
    s, z = ctx.lwe_secret(51), ctx.rlwe_secret(51)
    keys = ctx.keys()
    keys.generate(51, s, z)
    keys.generate_subs(52, z)
    path = tmp_path / "keys.bin"
    keys.save(path)
    k2 = ctx.keys()
    k2.load(path)
    assert k2.digest() == keys.digest()
    assert np.array_equal(k2.download_bsk(), keys.download_bsk())
    assert np.array_equal(k2.download_ksk(), keys.download_ksk())
    assert np.array_equal(k2.download_subs(), keys.download_subs())
    rng = np.random.default_rng(53)
    m1, m2 = rng.integers(0, 2, 16), rng.integers(0, 2, 16)
    c1, c2 = ctx.lwe_encrypt(s, m1, 54, 0), ctx.lwe_encrypt(s, m2, 54, 16)
    gates = rng.integers(0, 6, 16).astype(np.int32)
    assert np.array_equal(ctx.bootstrap(k2, gates, c1, c2), ctx.bootstrap(keys, gates, c1, c2))
    data = bytearray(path.read_bytes())
    data[len(data) // 2] ^= 1
    bad = tmp_path / "bad.bin"
    bad.write_bytes(bytes(data))
    k3 = ctx.keys()
    with pytest.raises(fhe.FheAccError, match="digest"):
        k3.load(bad)
    (tmp_path / "short.bin").write_bytes(bytes(data[: len(data) // 3]))
    with pytest.raises(fhe.FheAccError, match="truncated"):
        k3.load(tmp_path / "short.bin")
    other = fhe.Context(**SETS["n16_N256_q27"])
    with pytest.raises(fhe.FheAccError, match="parameter set"):
        other.keys().load(path)
