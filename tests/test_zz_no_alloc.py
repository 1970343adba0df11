"""cg_dev_lookup_noalloc: the lookup whose directory and build buffers live
in a fixed device array instead of hipMalloc'd memory, called through the C
ABI and checked bit for bit against cg_dev_lookup (and against a numpy
search) on calls in a row that reuse the array: a layer large enough for the
coarser fixed-size directory, then a smaller one, then an empty one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DPD_PREFIX = 0xFFF0FFF000000000


@pytest.fixture(scope="module")
def cg():
    import cablegen_amd
    cablegen_amd.lib()
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    return cablegen_amd


def layer(rng, n):
    """sorted unique keys with the DPDGap prefix and 40 varying bits"""
    k = rng.integers(0, 1 << 40, size=n, dtype=np.uint64) | np.uint64(DPD_PREFIX)
    return np.unique(k)


def run_both(cg, keys, vals, q, canon):
    L = cg.lib()
    dk = torch.from_numpy(keys.view(np.int64)).cuda() if len(keys) else \
        torch.empty(1, dtype=torch.int64, device="cuda")
    dv = torch.from_numpy(vals).cuda() if len(keys) else torch.empty(1, dtype=torch.float64, device="cuda")
    dq = torch.from_numpy(q.view(np.int64)).cuda()
    a = torch.full((len(q),), -1.0, dtype=torch.float64, device="cuda")
    b = torch.full((len(q),), -2.0, dtype=torch.float64, device="cuda")
    print(f"no-alloc lookup: n={len(keys)} nq={len(q)} canonicalize={canon}", flush=True)
    assert L.cg_dev_lookup(dk.data_ptr(), dv.data_ptr(), len(keys), dq.data_ptr(), len(q),
                           a.data_ptr(), canon) == 0, cg.last_error()
    assert L.cg_dev_lookup_noalloc(dk.data_ptr(), dv.data_ptr(), len(keys), dq.data_ptr(), len(q),
                                   b.data_ptr(), canon) == 0, cg.last_error()
    return a.cpu().numpy(), b.cpu().numpy()


def test_no_alloc_lookup_matches_lookup(cg):
    rng = np.random.default_rng(11)
    for n, canon in ((3_000_000, 0), (100_000, 1), (0, 0)):
        keys = layer(rng, n) if n else np.zeros(0, np.uint64)
        vals = rng.random(len(keys)) + 0.5
        hits = keys[rng.integers(0, len(keys), size=200_000)] if len(keys) else np.zeros(0, np.uint64)
        q = np.concatenate([hits, layer(rng, 50_000)])
        rng.shuffle(q)
        ref, got = run_both(cg, keys, vals, q, canon)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), n
        if not canon:   # and the plain search
            exp = np.zeros(len(q))
            if len(keys):
                pos = np.clip(np.searchsorted(keys, q), 0, len(keys) - 1)
                exp = np.where(keys[pos] == q, vals[pos], 0.0)
            assert np.array_equal(got, exp)
            assert (got > 0).sum() >= len(hits)


def test_no_alloc_lookup_rejects_2_32(cg):
    """n >= 2^32 fails with a message before any launch"""
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    assert cg.lib().cg_dev_lookup_noalloc(0, 0, 1 << 32, 0, 0, out.data_ptr(), 0) == -1
    assert "2^32" in cg.last_error()
