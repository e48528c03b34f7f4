"""GPU parity for RLWE expansion (expand_RLWE, app.cpp:26-97) and the leveled
operations CMUX, blind_rotate and blind_rotate_full (RLWE.cpp:788-838):
the HIP library through its C ABI against the oracle, word for word. The
oracle itself is pinned to the reference's own code compiled from its
sources (tests/test_oracle_vs_ref.py)."""
import numpy as np
import pytest

from fheacc_testlib import fhe
from oracle import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SETS = {
    # 27-bit Q, narrow device words, Bks_R = 2^7
    "N256_q27": dict(n=4, q=512, N=256, Q=0, nbit=27, BGbits=7, Br=16, Bks=32, Bks_R_bits=7),
    # 54-bit Q, wide device words, Bks_R = 2^9
    "N128_q54": dict(n=4, q=256, N=128, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=16, Bks=32, Bks_R_bits=9),
    # 29-bit Q, 32-bit words with Harvey's butterflies, Bks_R = 2^8
    "N256_q29": dict(n=4, q=512, N=256, Q=0, nbit=29, BGbits=10, Br=16, Bks=32, Bks_R_bits=8),
}
# the reference's expansion app (apps/rlwe_expand_fpga.cpp:43-144)
EXPAND_APP = dict(n=512, q=512, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=9, Br=23, Bks=25, Bks_R_bits=9)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def subs_pair(kw, seed):
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    z = ctx.rlwe_secret(seed)
    keys = ctx.keys()
    keys.generate_subs(seed + 1, z)
    return ctx, o, keys


@pytest.mark.parametrize("name", list(SETS))
def test_expand_rlwe_matches_oracle(name):
    ctx, o, keys = subs_pair(SETS[name], 31)
    subs = keys.download_subs()
    rng = np.random.default_rng(31)
    B, N = 3, o.N
    x = rng.integers(0, o.Q, (B, 2, N), dtype=np.uint64)
    out = torch.zeros((B, N, 2, N), dtype=torch.int64, device="cuda")
    ctx.expand_rlwe_dev(keys, dev(x), out)
    got = host(out)
    for b in range(B):
        assert np.array_equal(got[b], o.expand_rlwe(x[b], subs)), b


def test_expand_rlwe_reference_app_size():
    # N = 2048, 54-bit Q, Bks_R = 2^9 (digitKS_R = 6): one full expansion
    ctx, o, keys = subs_pair(EXPAND_APP, 32)
    rng = np.random.default_rng(32)
    x = rng.integers(0, o.Q, (1, 2, 2048), dtype=np.uint64)
    out = torch.zeros((1, 2048, 2, 2048), dtype=torch.int64, device="cuda")
    ctx.expand_rlwe_dev(keys, dev(x), out)
    assert np.array_equal(host(out)[0], o.expand_rlwe(x[0], keys.download_subs()))


@pytest.mark.parametrize("name", list(SETS))
@pytest.mark.parametrize("per_ct", [False, True])
def test_cmux_matches_oracle(name, per_ct):
    kw = SETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    rng = np.random.default_rng(33 + per_ct)
    B, N = 5, o.N
    c0, c1 = (rng.integers(0, o.Q, (B, 2, N), dtype=np.uint64) for _ in range(2))
    rgsw = rng.integers(0, o.Q, ((B if per_ct else 1), o.rgsw_words), dtype=np.uint64)
    out = torch.zeros((B, 2, N), dtype=torch.int64, device="cuda")
    ctx.cmux_dev(dev(c0), dev(c1), dev(rgsw), out, per_ciphertext_key=per_ct)
    got = host(out)
    for b in range(B):
        assert np.array_equal(got[b], o.cmux(c0[b], c1[b], rgsw[b if per_ct else 0])), b


@pytest.mark.parametrize("name", list(SETS))
@pytest.mark.parametrize("forward", [True, False])
def test_blind_rotate_rlwe_matches_oracle(name, forward):
    kw = SETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    rng = np.random.default_rng(35 + forward)
    B, N = 3, o.N
    steps = N.bit_length() - 1
    x = rng.integers(0, o.Q, (B, 2, N), dtype=np.uint64)
    rgsw = rng.integers(0, o.Q, (steps, o.rgsw_words), dtype=np.uint64)
    out = torch.zeros((B, 2, N), dtype=torch.int64, device="cuda")
    ctx.blind_rotate_rlwe_dev(dev(x), forward, dev(rgsw), out)
    got = host(out)
    for b in range(B):
        assert np.array_equal(got[b], o.blind_rotate_rlwe(x[b], forward, rgsw)), b


@pytest.mark.parametrize("name", list(SETS))
def test_blind_rotate_full_matches_oracle(name):
    kw = SETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    rng = np.random.default_rng(37)
    B, N = 2, o.N
    x = rng.integers(0, o.Q, (B, 2, N), dtype=np.uint64)
    a = np.array([1, -3, N + 5, -(N + 2), 0, 2 * N - 7, 17], dtype=np.int32)
    rgsw = rng.integers(0, o.Q, (a.size, o.rgsw_words), dtype=np.uint64)
    for b in (0, 11, -6):
        out = torch.zeros((B, 2, N), dtype=torch.int64, device="cuda")
        ctx.blind_rotate_full_dev(dev(x), b, a, dev(rgsw), out)
        got = host(out)
        for i in range(B):
            assert np.array_equal(got[i], o.blind_rotate_full(x[i], b, a, rgsw)), (b, i)
    # no private steps: the public rotation alone
    out = torch.zeros((B, 2, N), dtype=torch.int64, device="cuda")
    ctx.blind_rotate_full_dev(dev(x), 5, [], None, out)
    assert np.array_equal(host(out)[0], o.blind_rotate_full(x[0], 5, [], np.zeros(0, dtype=np.uint64)))


def test_leveled_argument_errors():
    ctx = fhe.Context(**SETS["N256_q27"])
    x = torch.zeros((1, 2, 256), dtype=torch.int64, device="cuda")
    rgsw = torch.zeros(ctx.digitG * 4 * 256, dtype=torch.int64, device="cuda")
    with pytest.raises(fhe.FheAccError, match="steps"):
        fhe._check(fhe.lib().fheacc_blind_rotate_rlwe_dev(ctx.h, x.data_ptr(), 1, rgsw.data_ptr(), 9,
                                                          x.data_ptr(), 1))
    keys = ctx.keys()
    with pytest.raises(fhe.FheAccError, match="subs key"):
        ctx.expand_rlwe_dev(keys, x, torch.zeros((1, 256, 2, 256), dtype=torch.int64, device="cuda"))


@pytest.mark.parametrize("N,Q,BGbits,BksR", [(256, 0, 7, 7), (128, 0x3FFFFFFFFED001, 18, 9)])
def test_cpp_rlwe_api_demo(tmp_path, N, Q, BGbits, BksR):
    # the reference-shaped C++ layer (host/fhe_acc.hpp: application::
    # expand_RLWE, RLWE::RLWE_mult_RGSW, CMUX, blind_rotate, blind_rotate_full)
    # through lib/rlwe_demo on words from files, against the oracle
    import os
    import subprocess
    from fheacc_testlib import ROOT
    kw = dict(n=8, q=min(2 * N, 1024), N=N, Q=Q, nbit=54 if Q else 27, BGbits=BGbits, Br=32, Bks=32,
              Bks_R_bits=BksR)
    o = Oracle(**kw)
    ctx = fhe.Context(**kw)
    z = ctx.rlwe_secret(61)
    keys = ctx.keys()
    keys.generate_subs(62, z)
    subs = keys.download_subs()
    rng = np.random.default_rng(N + BGbits)
    logN = N.bit_length() - 1
    x = rng.integers(0, o.Q, (2, 2, N), dtype=np.uint64)
    rg = rng.integers(0, o.Q, (logN, o.rgsw_words), dtype=np.uint64)
    a = np.array([(-1) ** i * (3 * i + 1) for i in range(logN)], dtype=np.int32)
    b = -7
    x.tofile(tmp_path / "in.bin")
    rg.tofile(tmp_path / "rgsw.bin")
    a.tofile(tmp_path / "a.bin")
    exe = os.path.join(ROOT, "3rd_gen_fhe_acc_amd", "lib", "rlwe_demo")
    r = subprocess.run([exe, str(tmp_path), str(N), hex(o.Q), str(BGbits), str(BksR), "61", "62", str(b)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "RLWE DEMO DONE" in r.stdout
    rd = lambda name, cnt: np.fromfile(tmp_path / name, dtype=np.uint64).reshape(cnt, 2, N)
    assert np.array_equal(rd("expand.bin", N), o.expand_rlwe(x[0], subs))
    assert np.array_equal(rd("mult.bin", 1)[0], o.mult_rgsw(x[0][None], rg[0])[0])
    assert np.array_equal(rd("cmux.bin", 1)[0], o.cmux(x[0], x[1], rg[0]))
    assert np.array_equal(rd("br_fwd.bin", 1)[0], o.blind_rotate_rlwe(x[0], True, rg))
    assert np.array_equal(rd("br_bwd.bin", 1)[0], o.blind_rotate_rlwe(x[0], False, rg))
    assert np.array_equal(rd("brf.bin", 1)[0], o.blind_rotate_full(x[0], b, a, rg))


def test_host_buffer_forms():
    # fheacc_expand_rlwe / fheacc_cmux (host buffers) equal the device forms
    ctx, o, keys = subs_pair(SETS["N128_q54"], 63)
    rng = np.random.default_rng(63)
    x = rng.integers(0, o.Q, (2, 2, o.N), dtype=np.uint64)
    out = torch.zeros((2, o.N, 2, o.N), dtype=torch.int64, device="cuda")
    ctx.expand_rlwe_dev(keys, dev(x), out)
    assert np.array_equal(ctx.expand_rlwe(keys, x), host(out))
    rg = rng.integers(0, o.Q, (2, o.rgsw_words), dtype=np.uint64)
    got = ctx.cmux(x[0:2], x[1::-1], rg)
    for i in range(2):
        assert np.array_equal(got[i], o.cmux(x[i], x[1 - i], rg[i]))
