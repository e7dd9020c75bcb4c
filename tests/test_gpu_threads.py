"""Two host threads, each on its own stream, factoring and solving at once.

The panel launches of an LU above order 512 and the single-column substitutions need all their workgroups resident
together (they exchange pivots and rows through one XCD's L2 with bounded waits). The library holds a process-wide
lock across those calls, so two callers on two streams take turns instead of splitting the XCD between them and
running out their waits (status 100000). Each thread's factors, pivots and solution must be bit for bit what the
same call gives alone.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_two_threads_factor_and_solve_order_9000_at_once():
    import torch
    torch.cuda.set_device(0)
    from rustymath_amd import device as dv
    n = 9000
    print(f'two threads, getrf + getrs, n={n} float64', flush=True)
    g = torch.Generator(device='cuda'); g.manual_seed(9)
    mats = [torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g) for _ in range(2)]
    rhs = [torch.randn(n, 1, dtype=torch.float64, device='cuda', generator=g) for _ in range(2)]

    def run(i):
        f, perm = dv.getrf(mats[i].clone())
        x = rhs[i].clone()
        dv.getrs(f, perm, x)
        return f, perm, x

    alone = [run(i) for i in range(2)]
    torch.cuda.synchronize()
    out, errors = [None, None], []

    def worker(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(2):
                    out[i] = run(i)
                s.synchronize()
        except Exception as e:  # reported after the join
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in threads: t.start()
    for t in threads: t.join()
    assert not errors, errors
    for i in range(2):
        for want, got in zip(alone[i], out[i]):
            assert torch.equal(want.view(torch.int64) if want.is_floating_point() else want,
                               got.view(torch.int64) if got.is_floating_point() else got)
        x = out[i][2].reshape(-1)
        r = rhs[i].reshape(-1) - mats[i] @ x
        assert float(r.norm() / (mats[i].norm() * x.norm())) <= 10 * n * np.finfo(np.float64).eps
