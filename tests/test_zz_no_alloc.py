"""siSearchNoAlloc (the search whose plan lives in the library's static device array)
against the allocating entry points siSearchOffsets + siSearchFill, through the C-ABI:
two calls in a row, the second with a smaller batch (the array keeps the first call's
contents), both orders, both coordinate widths, the capacity cut and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import superintervals_amd as sia
    import superintervals_amd._native as N
    L = N.lib
    L.siSearchNoAlloc.restype = C.c_int
    L.siSearchNoAlloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    return sia, N, torch


def _inputs(dtype, n, nq, seed):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, 2_000_000, n)
    e = s + rng.geometric(1 / 400, n) - 1
    qs = rng.integers(0, 2_000_000, nq)
    qe = qs + rng.integers(0, 3000, nq)
    base = (np.int64(7) << 36) if dtype == np.int64 else 0
    return tuple((x + base).astype(dtype) for x in (s, e, qs, qe))


def _allocating(ix, N, torch, tqs, tqe, order):
    nq = tqs.numel()
    off = torch.empty(nq + 1, dtype=torch.int64, device="cuda")
    cnt = torch.empty(nq, dtype=torch.int32, device="cuda")
    N.check(N.lib.siSearchOffsets(ix.handle, tqs.data_ptr(), tqe.data_ptr(), nq, cnt.data_ptr(), off.data_ptr(),
                                  None, 0, None), "siSearchOffsets")
    torch.cuda.synchronize()
    tot = int(off[-1].item())
    idx = torch.empty(max(tot, 1), dtype=torch.int32, device="cuda")
    N.check(N.lib.siSearchFill(ix.handle, tqs.data_ptr(), tqe.data_ptr(), nq, order, off.data_ptr(), idx.data_ptr(),
                               None), "siSearchFill")
    torch.cuda.synchronize()
    return cnt, off, idx[:tot]


def _no_alloc(N, torch, ix, tqs, tqe, order, capacity):
    nq = tqs.numel()
    off = torch.full((nq + 1,), -1, dtype=torch.int64, device="cuda")
    cnt = torch.full((nq,), -1, dtype=torch.int32, device="cuda")
    idx = torch.full((max(capacity, 1) + 64,), -7, dtype=torch.int32, device="cuda")
    rc = N.lib.siSearchNoAlloc(ix.handle, tqs.data_ptr(), tqe.data_ptr(), nq, order, cnt.data_ptr(), off.data_ptr(),
                               idx.data_ptr(), capacity, None)
    torch.cuda.synchronize()
    return rc, cnt, off, idx


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_no_alloc_matches_allocating_search(env, dtype):
    sia, N, torch = env
    print(f"[gpu] siSearchNoAlloc dtype={np.dtype(dtype).name}", flush=True)
    s, e, qs, qe = _inputs(dtype, 400_000, 300_000, 11)
    ix = sia.DeviceIndex(dtype).build(s, e)
    for nq in (300_000, 40_000):            # a second call with a smaller batch
        tqs = torch.from_numpy(qs[:nq]).cuda()
        tqe = torch.from_numpy(qe[:nq]).cuda()
        for order in (N.SI_ORDER_DESC, N.SI_ORDER_VECTOR):
            cnt0, off0, idx0 = _allocating(ix, N, torch, tqs, tqe, order)
            tot = int(off0[-1].item())
            rc, cnt, off, idx = _no_alloc(N, torch, ix, tqs, tqe, order, tot)
            assert rc == N.SI_OK, N.lib.siGetLastError()
            assert torch.equal(cnt, cnt0) and torch.equal(off, off0)
            assert torch.equal(idx[:tot], idx0)
            assert bool((idx[tot:] == -7).all())          # nothing written past the output
    # capacity cut: only slots below the capacity are written
    nq = 40_000
    tqs = torch.from_numpy(qs[:nq]).cuda()
    tqe = torch.from_numpy(qe[:nq]).cuda()
    cnt0, off0, idx0 = _allocating(ix, N, torch, tqs, tqe, N.SI_ORDER_DESC)
    cap = int(off0[-1].item()) // 3
    rc, cnt, off, idx = _no_alloc(N, torch, ix, tqs, tqe, N.SI_ORDER_DESC, cap)
    assert rc == N.SI_OK and torch.equal(off, off0)
    assert torch.equal(idx[:cap], idx0[:cap]) and bool((idx[cap:] == -7).all())
    # counts and offsets only (capacity 0), an empty batch, and the size limit
    rc, cnt, off, idx = _no_alloc(N, torch, ix, tqs, tqe, N.SI_ORDER_DESC, 0)
    assert rc == N.SI_OK and torch.equal(off, off0) and bool((idx == -7).all())
    rc, cnt, off, idx = _no_alloc(N, torch, ix, tqs[:0], tqe[:0], N.SI_ORDER_DESC, 0)
    assert rc == N.SI_OK and int(off[0].item()) == 0
    big = 1 << 20
    rc = N.lib.siSearchNoAlloc(ix.handle, tqs.data_ptr(), tqe.data_ptr(), big + 1, N.SI_ORDER_DESC, None,
                               off.data_ptr(), None, 0, None)
    assert rc == N.SI_ERR_ARG and b"at most" in N.lib.siGetLastError()
    ix.close()
