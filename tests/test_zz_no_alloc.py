"""cvk_infections_noalloc (no allocation; per-block counts in the library's fixed device array)
gives exactly what cvk_infections gives, on consecutive calls of different sizes through the C ABI.
Runs last: it is the one entry point whose scratch persists between calls."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def call(entry, beta, p1, p2, lb, rt, rs, uniforms=None, legacy=0, t=0, v=0, l=0, capacity=None):
    import torch
    from covasim_amd import _lib
    E = p1.numel()
    cap = (E if legacy else 2 * E) if capacity is None else capacity
    src = torch.full((max(cap, 1),), -7, dtype=torch.int32, device='cuda')
    tgt = torch.full((max(cap, 1),), -7, dtype=torch.int32, device='cuda')
    prob = torch.full((max(cap, 1),), -1.0, dtype=torch.float32, device='cuda')
    count = C.c_int64(-1)
    rc = getattr(_lib.lib(), entry)(float(np.float32(beta)), p1.data_ptr(), p2.data_ptr(), lb.data_ptr(), E,
                                    rt.data_ptr(), rs.data_ptr(), rt.numel(),
                                    uniforms.data_ptr() if uniforms is not None else None, 77, t, v, l, legacy,
                                    src.data_ptr(), tgt.data_ptr(), prob.data_ptr(), cap, C.byref(count),
                                    _lib.stream_ptr())
    torch.cuda.synchronize()
    k = max(count.value, 0)
    return rc, count.value, src[:k].cpu().numpy(), tgt[:k].cpu().numpy(), prob[:k].cpu().numpy().view(np.uint32)


def inputs(seed, n, E):
    import torch
    rng = np.random.default_rng(seed)
    p1 = rng.integers(0, n, E).astype(np.int32)
    p2 = rng.integers(0, n, E).astype(np.int32)
    p1[:20] = p2[:20]
    lb = rng.choice([1.0, 0.0, 2.5], E, p=[0.9, 0.05, 0.05]).astype(np.float32)
    rt = np.where(rng.random(n) < 0.2, rng.random(n) * 3, 0).astype(np.float32)
    rs = np.where(rng.random(n) < 0.8, rng.random(n), 0).astype(np.float32)
    U = rng.random((2, E))
    dev = lambda a: torch.from_numpy(a).cuda()
    return dev(p1), dev(p2), dev(lb), dev(rt), dev(rs), dev(U)


def test_noalloc_equals_allocating_entry_point(gpu):
    from covasim_amd import _lib
    # the second call is smaller: block counts left in the scratch by the first must not leak into it
    for (seed, n, E) in [(1, 200_000, 3_000_000), (2, 5_000, 70_001)]:
        p1, p2, lb, rt, rs, U = inputs(seed, n, E)
        for legacy in (0, 1):
            for uni in (None, U):
                for beta in (0.3, 0.02):
                    a = call('cvk_infections', beta, p1, p2, lb, rt, rs, uni, legacy, t=3, v=1, l=2)
                    b = call('cvk_infections_noalloc', beta, p1, p2, lb, rt, rs, uni, legacy, t=3, v=1, l=2)
                    assert a[0] == b[0] == _lib.OK
                    assert a[1] == b[1] > 0, (seed, legacy, beta)
                    for x, y in zip(a[2:], b[2:]):
                        assert np.array_equal(x, y), (seed, legacy, beta)


def test_noalloc_errors(gpu):
    import torch
    from covasim_amd import _lib
    p1, p2, lb, rt, rs, U = inputs(3, 1000, 10_000)
    # output too small: the count is reported and CVK_E_CAPACITY returned, as by cvk_infections
    a = call('cvk_infections', 0.9, p1, p2, lb, rt, rs, capacity=1)
    b = call('cvk_infections_noalloc', 0.9, p1, p2, lb, rt, rs, capacity=1)
    assert a[0] == b[0] == _lib.E_CAPACITY and a[1] == b[1] > 1
    # beyond the fixed scratch: rejected before any device work, with a message naming the limit
    L = _lib.lib()
    big = _lib.INFECTIONS_NOALLOC_MAX_SLOTS // 2 + 1
    cnt = C.c_int64(-1)
    dummy = torch.zeros(4, dtype=torch.int32, device='cuda')
    rc = L.cvk_infections_noalloc(0.1, dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), big, dummy.data_ptr(),
                                  dummy.data_ptr(), 4, None, 1, 0, 0, 0, 0, dummy.data_ptr(), dummy.data_ptr(), None,
                                  0, C.byref(cnt), _lib.stream_ptr())
    assert rc == _lib.E_CAPACITY and cnt.value == 0
    assert b'limit' in L.cvk_last_error()
    # and it still works afterwards
    c = call('cvk_infections_noalloc', 0.9, p1, p2, lb, rt, rs)
    d = call('cvk_infections', 0.9, p1, p2, lb, rt, rs)
    assert c[1] == d[1] and all(np.array_equal(x, y) for x, y in zip(c[2:], d[2:]))
