"""The allocation-free crawler block body (ilc_crawl_block_static) through the
C ABI: its output equals ilc_crawl_block's bit for bit, on two calls in a row,
the second with a smaller block (the static scratch keeps the first call's
contents), then a larger one again; blocks above its maximum are refused
before any device work."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from illcrawl_amd import backend
    return backend


L = np.array([10000.0, 10000.0, 12000.0])
VOL = ((4000.0, 6000.0, 5000.0), (3000.0, 2500.0, 4000.0), 250.0)


def block(n, ncols, seed):
    rng = np.random.default_rng(seed)
    xyz = (rng.uniform(-0.3, 1.3, (n, 3)) * L).astype(np.float32)
    cols = [rng.random(n).astype(np.float32) for _ in range(ncols)]
    return xyz, cols


def check_same(B, xyz, cols):
    c, e, bd = VOL
    p4, outs, idx = B.crawl_block(c, e, L, bd, xyz, cols, with_index=True)
    s4, souts, sidx = B.crawl_block(c, e, L, bd, xyz, cols, with_index=True,
                                    static_scratch=True)
    torch.cuda.synchronize()
    say(f"  kept {p4.shape[0]} of {xyz.shape[0]}")
    assert s4.shape[0] == p4.shape[0] > 0
    assert torch.equal(sidx, idx)
    assert np.array_equal(s4.cpu().numpy().view(np.uint32), p4.cpu().numpy().view(np.uint32))
    for a, b in zip(souts, outs):
        assert torch.equal(a, b)


def test_static_block_equals_allocating_block(B):
    # two calls in a row, the second smaller, then larger again
    for n, ncols, seed in ((1_500_000, 3, 11), (70_001, 2, 12), (2_000_000, 1, 13)):
        say(f"case: crawl_block_static n={n} ncols={ncols} seed={seed}")
        check_same(B, *block(n, ncols, seed))


def test_static_block_refuses_oversized_blocks(B):
    import ctypes
    from illcrawl_amd import _native as N
    say("case: crawl_block_static above its maximum")
    x = torch.zeros((16, 3), dtype=torch.float32, device="cuda")
    out = torch.zeros((16, 4), dtype=torch.float32, device="cuda")
    vol = N.VolumeC((ctypes.c_double * 3)(0, 0, 0), (ctypes.c_double * 3)(1, 1, 1),
                    (ctypes.c_double * 3)(10, 10, 10), 0.0)
    m = ctypes.c_uint64(7)
    # the host refuses the size before it reads a pointer or launches a kernel
    st = N.lib().ilc_crawl_block_static(ctypes.byref(vol), ctypes.c_void_p(x.data_ptr()),
                                        (1 << 29) + 1, None, 0, ctypes.c_void_p(out.data_ptr()),
                                        None, None, ctypes.byref(m), None)
    assert st == N.ILC_ERANGE
    assert b"allocation-free maximum" in N.lib().ilc_last_error()
