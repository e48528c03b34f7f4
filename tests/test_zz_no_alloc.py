"""The allocation-free gate path (fheacc_bootstrap_batch_dev_noalloc) against
the allocating one (fheacc_bootstrap_batch_dev): the same words on two calls
in a row, the second smaller than the first, with the intermediate captures;
a batch larger than the static scratch is refused before any launch."""
import numpy as np
import pytest

from fheacc_testlib import SETS, fhe

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to(dev)


def _run(ctx, keys, gates, c1, c2, noalloc):
    dev = "cuda:%d" % ctx.device
    d1, d2, dg = _dev(c1, dev), _dev(c2, dev), torch.from_numpy(gates).to(dev)
    out = torch.zeros_like(d1)
    acc = torch.zeros((c1.shape[0], 2, ctx.N), dtype=torch.int64, device=dev)
    ks = torch.zeros_like(d1)
    f = ctx.bootstrap_dev_noalloc if noalloc else ctx.bootstrap_dev
    f(keys, dg, d1, d2, out, cap_acc=acc, cap_ks=ks)
    ctx.sync()
    return [t.cpu().numpy().view(np.uint64) for t in (out, acc, ks)]


def test_noalloc_equals_allocating_path_on_two_calls():
    kw = SETS["n12_N1024_q27"]
    ctx = fhe.Context(**kw)
    s, z = ctx.lwe_secret(11), ctx.rlwe_secret(11)
    keys = ctx.keys()
    keys.generate(11, s, z)
    rng = np.random.default_rng(11)
    for B in (300, 37):  # the second call is smaller and reuses the array
        print("noalloc vs allocating, B =", B, flush=True)
        m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
        c1 = ctx.lwe_encrypt(s, m1, 11, 0)
        c2 = ctx.lwe_encrypt(s, m2, 11, B)
        gates = rng.integers(0, 6, B).astype(np.int32)
        ref = _run(ctx, keys, gates, c1, c2, noalloc=False)
        got = _run(ctx, keys, gates, c1, c2, noalloc=True)
        for r, g in zip(ref, got):
            assert np.array_equal(r, g)
        truth = {0: m1 | m2, 1: m1 & m2, 2: 1 - (m1 | m2), 3: 1 - (m1 & m2), 4: m1 ^ m2, 5: 1 - (m1 ^ m2)}
        exp = np.array([truth[g][i] for i, g in enumerate(gates)])
        assert np.array_equal(ctx.lwe_decrypt(s, got[0]), exp)


def test_noalloc_refuses_a_batch_beyond_its_scratch():
    kw = SETS["n12_N1024_q27"]
    ctx = fhe.Context(**kw)
    s, z = ctx.lwe_secret(12), ctx.rlwe_secret(12)
    keys = ctx.keys()
    keys.generate(12, s, z)
    B = 6000  # 6000 * N * digitKS = 36.9 MB of key-switch digits > 32 MiB
    dev = "cuda:%d" % ctx.device
    c = torch.zeros((B, ctx.n + 1), dtype=torch.int64, device=dev)
    g = torch.zeros(B, dtype=torch.int32, device=dev)
    with pytest.raises(fhe.FheAccError, match="scratch"):
        ctx.bootstrap_dev_noalloc(keys, g, c, c, torch.zeros_like(c))
