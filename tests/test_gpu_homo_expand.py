"""GPU parity for homomorphic expansion (application::homo_expand,
app.cpp:78-97; the accelerator's fpga_homo_expand, myfpga.cpp:2278-2452), the
tree selection's CMUX tree (test.cpp:1510-1600, 3984-4575), and the RLWE /
RGSW encryption and RLWE decryption that feed them (RLWE.cpp:324-376,
561-577): the HIP library through its C ABI against the oracle, word for
word.  The oracle is pinned to the reference's own homo_expand, CMUX and
RLWE_decrypt compiled from its sources (tests/test_oracle_vs_ref.py) and to
exact encryption identities (tests/test_oracle_math.py)."""
import numpy as np
import pytest

from fheacc_testlib import fhe
from oracle import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SETS = {
    "N256_q27": dict(n=4, q=512, N=256, Q=0, nbit=27, BGbits=7, Br=16, Bks=32, Bks_R_bits=7),
    "N128_q54": dict(n=4, q=256, N=128, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=16, Bks=32, Bks_R_bits=9),
    "N256_q29": dict(n=4, q=512, N=256, Q=0, nbit=29, BGbits=10, Br=16, Bks=32, Bks_R_bits=8),
}
# the reference's expansion app (apps/rlwe_expand_fpga.cpp:43-144)
EXPAND_APP = dict(n=512, q=512, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=9, Br=23, Bks=25, Bks_R_bits=9)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("name", list(SETS))
def test_encrypt_decrypt_match_oracle(name):
    kw = SETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    z = ctx.rlwe_secret(51)
    rng = np.random.default_rng(51)
    m = rng.integers(0, o.Q, (3, o.N), dtype=np.uint64)
    c = ctx.rlwe_encrypt(z, m, 52, first_index=7)
    assert np.array_equal(c, o.rlwe_encrypt(z, m, 52, first_index=7))
    assert np.array_equal(ctx.rgsw_encrypt(z, m, 53, first_index=3), o.rgsw_encrypt(z, m, 53, first_index=3))
    for t in (2, 4, 1 << 10):
        assert np.array_equal(ctx.rlwe_decrypt(z, c, t), o.rlwe_decrypt(z, c, t)), t
    # a message on the t grid decrypts back to itself
    msg = rng.integers(0, 4, (2, o.N), dtype=np.uint64)
    enc = ctx.rlwe_encrypt(z, (msg * (o.Q // 4)) % o.Q, 54)
    assert np.array_equal(ctx.rlwe_decrypt(z, enc, 4) % 4, msg)


@pytest.mark.parametrize("name", list(SETS))
def test_homo_expand_matches_oracle(name):
    kw = SETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    z = ctx.rlwe_secret(55)
    keys = ctx.keys()
    keys.generate_subs(56, z)
    subs = keys.download_subs()
    rng = np.random.default_rng(55)
    count, N = 2, o.N
    packed = rng.integers(0, o.Q, (count, o.digitG, 2, N), dtype=np.uint64)
    enc_sec = rng.integers(0, o.Q, o.rgsw_words, dtype=np.uint64)
    out = torch.zeros((count, N, o.rgsw_words), dtype=torch.int64, device="cuda")
    ctx.homo_expand_dev(keys, dev(packed), dev(enc_sec), out)
    got = host(out)
    for q in range(count):
        assert np.array_equal(got[q], o.homo_expand(packed[q], enc_sec, subs)), q
    # the host-buffer form gives the same words
    assert np.array_equal(ctx.homo_expand(keys, packed[1:], enc_sec)[0], got[1])


@pytest.mark.parametrize("name", ["N256_q27", "N128_q54"])
@pytest.mark.parametrize("levels", [1, 2, 4])
def test_cmux_tree_matches_oracle(name, levels):
    kw = SETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    rng = np.random.default_rng(57 + levels)
    count, N = 3, o.N
    leaves = rng.integers(0, o.Q, (count, 1 << levels, 2, N), dtype=np.uint64)
    rgsw = rng.integers(0, o.Q, (levels, o.rgsw_words), dtype=np.uint64)
    out = torch.zeros((count, 2, N), dtype=torch.int64, device="cuda")
    ctx.cmux_tree_dev(dev(leaves), dev(rgsw), out, levels)
    got = host(out)
    for t in range(count):
        assert np.array_equal(got[t], o.cmux_tree(leaves[t], rgsw)), t
    assert np.array_equal(ctx.cmux_tree(leaves, rgsw), got)


def test_homo_expansion_and_tree_selection_reference_app():
    # test.cpp:1510-1600 at the reference app's size (N = 2048, 54-bit Q,
    # BG = 2^9, d = 6): packed bits b_i = i & 1, homo_expand with RGSW(-s),
    # a CMUX tree over 2^(18 - 11) leaves with RGSWs 11..17 (selects leaf 85),
    # then blind_rotate backward with RGSWs 0..10 (by 682)
    kw = EXPAND_APP
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    N, Q, d = o.N, o.Q, o.digitG
    z = ctx.rlwe_secret(71)
    keys = ctx.keys()
    keys.generate_subs(72, z)
    subs = keys.download_subs()
    neg_s = np.array([(Q - int(v)) % Q for v in z], dtype=np.uint64)
    enc_sec = ctx.rgsw_encrypt(z, neg_s, 73)[0]
    assert np.array_equal(enc_sec, o.rgsw_encrypt(z, neg_s, 73)[0])
    msgs = np.zeros((d, N), dtype=np.uint64)
    for h in range(d):
        msgs[h, 1::2] = (pow(2, o.BGbits * h, Q) * o.iN) % Q
    packed = ctx.rlwe_encrypt(z, msgs, 74)
    assert np.array_equal(packed, o.rlwe_encrypt(z, msgs, 74))
    print("homo_expand at N = 2048, d = 6", flush=True)
    rg = ctx.homo_expand(keys, packed, enc_sec)[0]
    rows = rg.reshape(N, d, 2, 2, N)
    for j in (0, d - 1):  # column 1 of rows j is expansion j, every slot
        assert np.array_equal(rows[:, j, 1], o.expand_rlwe(packed[j], subs)), j
    rng = np.random.default_rng(71)
    for i, j in zip(rng.integers(0, N, 16), rng.integers(0, d, 16)):  # column 0: the product with RGSW(-s)
        assert np.array_equal(rows[i, j, 0], o.mult_rgsw(rows[i, j, 1][None], enc_sec)[0]), (i, j)
    total, digitN = 18, 11
    leaf_msgs = np.zeros((1 << (total - digitN), N), dtype=np.uint64)
    for i in range(leaf_msgs.shape[0]):
        leaf_msgs[i, :i + 1] = Q // 2
    leaves = ctx.rlwe_encrypt(z, leaf_msgs, 75)
    print("CMUX tree of 128 leaves", flush=True)
    sel = ctx.cmux_tree(leaves, rg[digitN:total])[0]
    assert np.array_equal(sel, o.cmux_tree(leaves, rg[digitN:total]))
    out = torch.zeros((1, 2, N), dtype=torch.int64, device="cuda")
    ctx.blind_rotate_rlwe_dev(dev(sel[None]), False, dev(rg[:digitN]), out)
    rot = host(out)[0]
    assert np.array_equal(rot, o.blind_rotate_rlwe(sel, False, rg[:digitN]))
    plain = ctx.rlwe_decrypt(z, rot, 2)[0]
    want = np.zeros(N, dtype=np.uint64)
    want[N - 682:N - 682 + 86] = 1  # leaf 85 rotated by X^-682, negacyclic
    assert np.array_equal(plain, want)


@pytest.mark.parametrize("N", [256, 2048])
def test_cpp_tree_selection_demo(N):
    # test.cpp:1510-1600 through the reference-shaped C++ layer
    # (application::homo_expand, FPGA::fpga_homo_expand, RLWE::CMUX,
    # blind_rotate, RLWE_encrypt / RGSW_encrypt / RLWE_decrypt): lib/tree_demo
    # checks the decrypted selection itself
    import os
    import subprocess
    from fheacc_testlib import ROOT
    exe = os.path.join(ROOT, "3rd_gen_fhe_acc_amd", "lib", "tree_demo")
    r = subprocess.run([exe, str(N)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "TREE DEMO OK" in r.stdout
