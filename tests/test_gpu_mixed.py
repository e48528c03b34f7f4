"""Mixed parameter sets in one launch (fheacc_graph_capture_mixed, and
fheacc_graph_capture_interleaved for rows of several sets interleaved in one
buffer): one HIP graph holding every set's kernels, replayed as one unit.
Each set's words must be bit-exact:
- at reduced sizes against the oracle, and against the same set run on its
  own;
- at full size (60 % benchmark set, 30 % N=2048 54-bit d=3 n=630,
  10 % N=512 27-bit d=5; 8192 ciphertexts, uniformly interleaved) against
  each set's own dense launch, which the parity tests pin to the oracle at
  the same kernel shapes; plus every decrypted bit."""
import numpy as np
import pytest

from fheacc_testlib import MIXED, MIXED_SHARE, SETS, fhe
from oracle import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TRUTH = {0: lambda a, b: a | b, 1: lambda a, b: a & b, 2: lambda a, b: 1 - (a | b), 3: lambda a, b: 1 - (a & b),
         4: lambda a, b: a ^ b, 5: lambda a, b: 1 - (a ^ b)}


def dev(a):
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(np.int64) if a.dtype == np.uint64 else a).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def interleave(total, shares, seed):
    """Uniform random interleaving: position -> set name."""
    names = list(shares)
    rng = np.random.default_rng(seed)
    counts = [int(round(total * shares[k])) for k in names]
    counts[0] += total - sum(counts)
    lab = np.concatenate([np.full(c, i) for i, c in enumerate(counts)])
    rng.shuffle(lab)
    return [names[i] for i in lab]


def build(kw, B, seed, full_keys=True):
    ctx = fhe.Context(**kw)
    s, z = ctx.lwe_secret(seed), ctx.rlwe_secret(seed)
    keys = ctx.keys()
    keys.generate(seed, s, z)
    rng = np.random.default_rng(seed)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    c1, c2 = ctx.lwe_encrypt(s, m1, seed, 0), ctx.lwe_encrypt(s, m2, seed, B)
    exp = np.array([TRUTH[int(g)](a, b) for g, a, b in zip(gates, m1, m2)])
    return dict(ctx=ctx, keys=keys, s=s, gates=gates, c1=c1, c2=c2, exp=exp)


def fresh_inputs(st, seed):
    """New operands and gates for a set built by build(), same keys."""
    B = st["gates"].size
    rng = np.random.default_rng(seed)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    c1, c2 = st["ctx"].lwe_encrypt(st["s"], m1, seed, 0), st["ctx"].lwe_encrypt(st["s"], m2, seed, B)
    exp = np.array([TRUTH[int(g)](a, b) for g, a, b in zip(gates, m1, m2)])
    return dict(gates=gates, c1=c1, c2=c2, exp=exp)


def test_mixed_graph_matches_oracle_small():
    names = ["n16_N256_q27", "n10_N128_q54", "n8_N512_q27_bg6"]
    lab = interleave(60, dict(zip(names, (0.6, 0.3, 0.1))), 5)
    sets = {}
    for i, nm in enumerate(names):
        B = lab.count(nm)
        st = build(SETS[nm], B, 40 + i)
        st["dg"], st["d1"], st["d2"] = dev(st["gates"]), dev(st["c1"]), dev(st["c2"])
        st["out"] = torch.zeros_like(st["d1"])
        sets[nm] = st
    g = fhe.MixedGraph([(st["ctx"], st["keys"], st["dg"], st["d1"], st["d2"], st["out"]) for st in sets.values()])
    for rep in range(2):
        for st in sets.values():
            st["out"].zero_()
        g.launch()
        for nm, st in sets.items():
            st["ctx"].sync()
            got = host(st["out"])
            o = Oracle(**SETS[nm])
            exp = o.bootstrap(st["c1"], st["c2"], st["gates"], st["keys"].download_bsk(), st["keys"].download_ksk())
            assert np.array_equal(got, exp), (nm, rep)
            assert np.array_equal(got, st["ctx"].bootstrap(st["keys"], st["gates"], st["c1"], st["c2"])), nm
            assert np.array_equal(st["ctx"].lwe_decrypt(st["s"], got), st["exp"]), nm
    g.close()


def test_mixed_full_size():
    lab = interleave(8192, MIXED_SHARE, 6)
    sets = {}
    for i, nm in enumerate(MIXED):
        B = lab.count(nm)
        st = build(MIXED[nm], B, 50 + i)
        st["ctx"].autotune(st["keys"], min(B, 1024))
        st["dg"], st["d1"], st["d2"] = dev(st["gates"]), dev(st["c1"]), dev(st["c2"])
        st["out"] = torch.zeros_like(st["d1"])
        st["dense"] = torch.zeros_like(st["d1"])
        sets[nm] = st
    assert sum(st["gates"].size for st in sets.values()) == 8192
    g = fhe.MixedGraph([(st["ctx"], st["keys"], st["dg"], st["d1"], st["d2"], st["out"]) for st in sets.values()])
    g.launch()
    for nm, st in sets.items():
        st["ctx"].sync()
        st["ctx"].bootstrap_dev(st["keys"], st["dg"], st["d1"], st["d2"], st["dense"])
        st["ctx"].sync()
        got = host(st["out"])
        assert np.array_equal(got, host(st["dense"])), nm
        assert np.array_equal(st["ctx"].lwe_decrypt(st["s"], got), st["exp"]), nm
    g.close()
    # set B at full size (n=630, N=2048, 54-bit Q, BG=2^18 d=3), three gates
    # against the oracle: every accumulator word after the blind rotation,
    # every key-switched word and every output word
    st = sets["B"]
    idx = np.array([0, st["gates"].size // 2, st["gates"].size - 1])
    out, acc, ks = st["ctx"].bootstrap(st["keys"], st["gates"][idx], st["c1"][idx], st["c2"][idx], capture=True)
    assert np.array_equal(out, host(st["out"])[idx])
    o = Oracle(**MIXED["B"])
    eo, ea, ek = o.bootstrap(st["c1"][idx], st["c2"][idx], st["gates"][idx], st["keys"].download_bsk(),
                             st["keys"].download_ksk(), threads=3, capture=True)
    assert np.array_equal(acc, ea)
    assert np.array_equal(ks, ek)
    assert np.array_equal(out, eo)


def test_interleaved_graph_matches_oracle():
    # the mixed batch consumed as it arrives: three parameter sets' rows
    # interleaved in one buffer (row stride = the largest n + 1), each set's
    # kernels gathering their rows on the device; every row equals the oracle
    # on its set, over two replays with new inputs in between
    names = ["n12_N512_q27", "n8_N2048_q54", "n8_N512_q27_bg6"]
    lab = np.array(interleave(90, dict(zip(names, (0.5, 0.3, 0.2))), 7))
    sets = [build(SETS[nm], int(np.count_nonzero(lab == nm)), 60 + i) for i, nm in enumerate(names)]
    rw = max(st["ctx"].n for st in sets) + 1
    total = lab.size
    set_of = np.array([names.index(x) for x in lab], dtype=np.int32)
    g = torch.zeros(total, dtype=torch.int32, device="cuda")
    d1 = torch.zeros((total, rw), dtype=torch.int64, device="cuda")
    d2, out = torch.zeros_like(d1), torch.zeros_like(d1)
    graph = fhe.InterleavedGraph([st["ctx"] for st in sets], [st["keys"] for st in sets], set_of, g, d1, d2, out)
    for rep in range(2):
        c1 = np.zeros((total, rw), dtype=np.uint64)
        c2, gates = np.zeros_like(c1), np.zeros(total, dtype=np.int32)
        for i, st in enumerate(sets):
            if rep:
                st.update(fresh_inputs(st, 70 + i))
            rows = np.flatnonzero(set_of == i)
            n1 = st["ctx"].n + 1
            c1[rows, :n1], c2[rows, :n1], gates[rows] = st["c1"], st["c2"], st["gates"]
        d1.copy_(torch.from_numpy(c1.view(np.int64)))
        d2.copy_(torch.from_numpy(c2.view(np.int64)))
        g.copy_(torch.from_numpy(gates))
        torch.cuda.synchronize()
        graph.launch()
        got = host(out)
        for i, st in enumerate(sets):
            st["ctx"].sync()
            rows = np.flatnonzero(set_of == i)
            n1 = st["ctx"].n + 1
            o = Oracle(**SETS[names[i]])
            exp = o.bootstrap(st["c1"], st["c2"], st["gates"], st["keys"].download_bsk(), st["keys"].download_ksk())
            assert np.array_equal(got[rows, :n1], exp), (names[i], rep)
            assert np.array_equal(st["ctx"].lwe_decrypt(st["s"], got[rows, :n1]), st["exp"]), (names[i], rep)
    graph.close()
