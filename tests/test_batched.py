"""GPU parity of the batched uneven solves against the oracle, through the C ABI (hs_dev_*gesv_batched_*).

Each cell of a batch must come out as the oracle's `solve` of that cell (rustymath's `LuWork::solve`,
core/src/linalg/factor.rs:207-226), bit for bit: the `regular` decision, the pivot order, the factors and X. The cases
cover every order where the reference's leaf and panel structure changes (16-column leaves, 32-row backward leaves),
empty systems and systems without right-hand sides, zero, subnormal, NaN and tied pivots, both dtypes, every tuning
choice, graph replays, and the full 10,000-system workload through size-independent checks (backward error of every
system, a sample against the oracle).
"""
import numpy as np
import pytest

import oracle
from test_oracle import bits_equal

pytestmark = pytest.mark.gpu

EDGE_NS = [0, 1, 2, 3, 7, 8, 9, 15, 16, 17, 23, 24, 31, 32, 33, 47, 48, 63, 64, 65, 96, 100, 127, 128, 129, 200, 255,
           256, 257, 300, 383, 384, 400, 500, 511, 512]


@pytest.fixture(scope='module')
def dv():
    import torch
    torch.cuda.set_device(0)
    from rustymath_amd import device
    return device


def solve_batch(dv, mats, rhs, dtype=np.float64, **knobs):
    "Runs one batch through BatchedSolve; returns (factors, perms, regular flags, solutions) on the host."
    import torch
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    a, b, ns, ks = dv.pack(mats, rhs, dtype=tdt)
    plan = dv.BatchedSolve(a, b, ns, ks, **knobs)
    plan.run()
    torch.cuda.synchronize()
    out = []
    reg = plan.regular.cpu().numpy()
    for i in range(len(ns)):
        f, perm = plan.factors(i)
        out.append((f.cpu().numpy(), perm.cpu().numpy(), bool(reg[i]), plan.x(i).cpu().numpy()))
    plan.close()
    return out


def check_cell(a, b, got, what):
    "One cell against the oracle: regular flag, pivots, factors (rows moved to their positions) and X, bit for bit."
    f, perm, regular, x = got
    n = a.shape[0]
    want = oracle.solve(a, b)
    assert regular == (want is not None), f'{what}: regular {regular}'
    if n == 0: return
    wf, wperm, _ = oracle.lu(a)
    assert (perm == wperm).all(), f'{what}: pivots'
    assert bits_equal(f[perm], wf.T), f'{what}: factors'
    if regular: assert bits_equal(x, want), f'{what}: solution'
    else: assert bits_equal(x, b), f'{what}: b must stay as it was'


# Each (panel width, workgroups per CU) form: 512 threads with one panel row each, or 256 with up to two.
FORMS = [(32, 1), (16, 1), (32, 2), (16, 2)]


@pytest.mark.parametrize('form', FORMS)
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_every_order_matches_the_oracle(dv, dtype, form):
    "Orders across the leaf (16) and backward-leaf (32) boundaries up to 512, 0 to 8 right-hand sides, N(0, 1) entries."
    rng = np.random.default_rng(11)
    mats = [rng.standard_normal((n, n)).astype(dtype) for n in EDGE_NS]
    rhs = [rng.standard_normal((n, i % 9)).astype(dtype) for i, n in enumerate(EDGE_NS)]
    got = solve_batch(dv, mats, rhs, dtype, panel_width=form[0], wgs_per_cu=form[1])
    for i, (a, b) in enumerate(zip(mats, rhs)):
        check_cell(a, b, got[i], f'n={a.shape[0]} k={b.shape[1]}')


def special_cells(dtype):
    "Matrices that exercise the reference's pivot rules: zero pivots, ties, subnormal pivots, NaN and infinity."
    rng = np.random.default_rng(5)
    cells = []
    z = np.zeros((20, 20), dtype); cells.append(('zero matrix', z))
    a = rng.standard_normal((40, 40)).astype(dtype); a[:, 5] = 0; cells.append(('zero column in a leaf', a))
    a = rng.standard_normal((70, 70)).astype(dtype); a[:, 33] = 0; a[:, 34] = 0; cells.append(('zero columns at a panel edge', a))
    a = np.sign(rng.standard_normal((64, 64))).astype(dtype); cells.append(('random signs (ties)', a))
    a = (rng.integers(-2, 3, size=(90, 90))).astype(dtype); cells.append(('small integers (ties, zero pivots)', a))
    tiny = np.ldexp(1.0, -1060 if dtype == np.float64 else -140)
    a = (rng.standard_normal((48, 48)) * tiny).astype(dtype); cells.append(('subnormal pivots', a))
    a = rng.standard_normal((36, 36)).astype(dtype); a[7, 3] = np.nan; cells.append(('a NaN', a))
    a = rng.standard_normal((36, 36)).astype(dtype); a[:, 2] = np.nan; cells.append(('a NaN column', a))
    a = rng.standard_normal((36, 36)).astype(dtype); a[10, 10] = np.inf; cells.append(('an infinity', a))
    a = np.eye(33, dtype=dtype)[::-1].copy(); cells.append(('reversed identity', a))
    a = np.triu(rng.standard_normal((50, 50))).astype(dtype); cells.append(('upper triangular', a))
    a = (rng.standard_normal((60, 5)) @ rng.standard_normal((5, 60))).astype(dtype); cells.append(('rank 5', a))
    # Above order 256, where the 256-thread form holds two rows of a panel per thread.
    a = (rng.integers(-2, 3, size=(300, 300))).astype(dtype); cells.append(('small integers, order 300', a))
    a = rng.standard_normal((400, 400)).astype(dtype); a[:, 100] = 0; a[:, 260] = 0; a[:, 261] = 0
    cells.append(('zero columns, order 400', a))
    return cells


@pytest.mark.parametrize('form', FORMS)
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_special_matrices_match_the_oracle(dv, dtype, form):
    rng = np.random.default_rng(6)
    cells = special_cells(dtype)
    mats = [a for _, a in cells]
    rhs = []
    for i, a in enumerate(mats):
        b = rng.standard_normal((a.shape[0], 1 + i % 8))
        if 'subnormal' in cells[i][0]: b = b * np.ldexp(1.0, -1060 if dtype == np.float64 else -140)
        rhs.append(b.astype(dtype))
    got = solve_batch(dv, mats, rhs, dtype, panel_width=form[0], wgs_per_cu=form[1])
    for i, (name, a) in enumerate(cells):
        check_cell(a, rhs[i], got[i], name)


def test_tuning_choices_and_graph_replays_give_the_same_bits(dv):
    """Panel width, workgroups per CU, a direct launch and repeated graph replays (with other GPU work between them)
    change nothing in the results."""
    import torch
    rng = np.random.default_rng(9)
    ns = rng.integers(1, 513, size=60)
    mats = [rng.standard_normal((n, n)) for n in ns]
    rhs = [rng.standard_normal((n, 1 + i % 8)) for i, n in enumerate(ns)]
    a0, b0, ns_, ks_ = dv.pack(mats, rhs)
    ref = None
    for pw in (16, 32):
        for wg in (1, 2):
            a, b = a0.clone(), b0.clone()
            plan = dv.BatchedSolve(a, b, ns_, ks_, panel_width=pw, wgs_per_cu=wg)
            for rep, graph in enumerate((True, True, False, True)):
                a.copy_(a0); b.copy_(b0)
                plan.run(graph=graph)
                noise = torch.randn(1000, 1000, device='cuda', dtype=torch.float64)
                (noise @ noise).sum().item()      # unrelated eager work and a synchronization between replays
                torch.cuda.synchronize()
                got = (a.cpu().numpy(), b.cpu().numpy(), plan.perm.cpu().numpy(), plan.regular.cpu().numpy())
                if ref is None:
                    ref = got
                    for i in range(0, len(ns), 7):
                        check_cell(mats[i], rhs[i], (plan.factors(i)[0].cpu().numpy(), plan.factors(i)[1].cpu().numpy(),
                                                      bool(got[3][i]), plan.x(i).cpu().numpy()), f'cell {i}')
                else:
                    assert all(bits_equal(x, y) for x, y in zip(got, ref)), (pw, wg, rep, graph)
            plan.close()


def test_full_workload_backward_errors_and_every_cell(dv):
    """The 10,000-system workload (n uniform in [8, 512], k in [1, 8], A = G + sqrt(n) I, seed 3): every system is
    regular and its normwise backward error ||B - A X|| / (||A|| ||X||) is at most 10 n eps64; and every cell's
    pivots, factors and solution equal the oracle's bit for bit."""
    import torch
    from bench import batched_workload
    a0, b0, ns, ks = batched_workload()
    a, b = a0.clone(), b0.clone()
    plan = dv.BatchedSolve(a, b, ns, ks)
    plan.run()
    torch.cuda.synchronize()
    assert bool((plan.regular[:len(ns)] == 1).all())
    eps = np.finfo(np.float64).eps
    worst = 0.0
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        for k in np.unique(ks[idx]):
            sel = idx[ks[idx] == k]
            A = torch.stack([a0[plan.a_off[i]:plan.a_off[i] + n * n].view(n, n) for i in sel])
            B = torch.stack([b0[plan.b_off[i]:plan.b_off[i] + n * k].view(n, k) for i in sel])
            X = torch.stack([b[plan.b_off[i]:plan.b_off[i] + n * k].view(n, k) for i in sel])
            r = (B - torch.bmm(A, X.contiguous())).flatten(1).norm(dim=1)
            e = r / (A.flatten(1).norm(dim=1) * X.flatten(1).norm(dim=1))
            worst = max(worst, float((e / (10 * n * eps)).max()))
    assert worst <= 1.0, worst
    a_h, b_h, f_h, x_h = a0.cpu().numpy(), b0.cpu().numpy(), a.cpu().numpy(), b.cpu().numpy()
    p_h = plan.perm.cpu().numpy()
    for i in range(len(ns)):
        n, k = int(ns[i]), int(ks[i])
        ao, bo, po = int(plan.a_off[i]), int(plan.b_off[i]), int(plan.p_off[i])
        check_cell(a_h[ao:ao + n * n].reshape(n, n), b_h[bo:bo + n * k].reshape(n, k),
                   (f_h[ao:ao + n * n].reshape(n, n), p_h[po:po + n], True, x_h[bo:bo + n * k].reshape(n, k)),
                   f'system {i} n={n} k={k}')
    plan.close()


def test_autotuner_picks_a_candidate_and_remembers_it(dv):
    import torch
    rng = np.random.default_rng(2)
    ns = rng.integers(8, 129, size=40)
    mats = [rng.standard_normal((n, n)) + np.sqrt(n) * np.eye(n) for n in ns]
    rhs = [rng.standard_normal((n, 2)) for n in ns]
    a, b, ns_, ks_ = dv.pack(mats, rhs)
    choice, times = dv.tune(a, b, ns_, ks_, reps=1)
    assert choice in times and len(times) == 4 and all(t > 0 for t in times.values())
    assert dv.tune(a, b, ns_, ks_)[0] == choice
    torch.cuda.synchronize()


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_factors_are_reused_for_solves_and_transposed_solves(dv, dtype):
    """A factor-and-solve run leaves factors and perm that the 'lu_solve' mode reuses for new right-hand sides (the
    oracle's lu_solve on the same factors, bit for bit) and the 'solve_t' mode for Aᵀ Y = W (the large-matrix
    transposed solve's bits, and a small residual)."""
    import torch
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    rng = np.random.default_rng(4)
    ns = [1, 5, 16, 33, 64, 100, 257, 512]
    mats = [rng.standard_normal((n, n)).astype(dtype) for n in ns]
    rhs = [rng.standard_normal((n, 1 + i % 8)).astype(dtype) for i, n in enumerate(ns)]
    a, b, ns_, ks_ = dv.pack(mats, rhs, dtype=tdt)
    plan = dv.BatchedSolve(a, b, ns_, ks_)
    plan.run()
    rhs2 = [rng.standard_normal((n, 8 - i % 8)).astype(dtype) for i, n in enumerate(ns)]
    _, b2, _, ks2 = dv.pack(mats, rhs2, dtype=tdt)
    again = dv.BatchedSolve(a, b2, ns_, ks2, mode='lu_solve', perm=plan.perm)
    again.run()
    _, b3, _, ks3 = dv.pack(mats, rhs2, dtype=tdt)
    trans = dv.BatchedSolve(a, b3, ns_, ks3, mode='solve_t', perm=plan.perm)
    trans.run()
    torch.cuda.synchronize()
    for i, n in enumerate(ns):
        wf, wperm, _ = oracle.lu(mats[i])
        assert bits_equal(again.x(i).cpu().numpy(), oracle.lu_solve(wf, wperm, rhs2[i])), f'lu_solve n={n}'
        f = torch.from_numpy(mats[i]).cuda()
        f, perm = dv.getrf(f)
        w = torch.from_numpy(rhs2[i]).cuda()
        dv.getrs(f, perm, w, trans=True)
        y = trans.x(i)
        assert bits_equal(y.cpu().numpy(), w.cpu().numpy()), f'solve_t n={n}'
        r = np.linalg.norm(mats[i].T.astype(np.float64) @ y.cpu().numpy() - rhs2[i])
        scale = np.linalg.norm(mats[i].astype(np.float64)) * np.linalg.norm(y.cpu().numpy())
        assert r <= 10 * n * np.finfo(dtype).eps * scale
    for p in (plan, again, trans): p.close()


def test_fp32_lu_with_fp64_refinement_on_the_full_workload(dv):
    """The second form of the batched workload (10,000 float64 systems, n uniform in [8, 512], k in [1, 8],
    A = G + sqrt(n) I, seed 3): FP32 LU on MFMA and FP64 refinement on the GPU. Every system is regular and ends with
    a normwise backward error ||B - A X||_F / (||A||_F ||X||_F) at most 10 n eps64, recomputed here in FP64 from the
    original float64 A and B; a graph replay gives the same bits; and the FP32 factors of a sample are the oracle's
    FP32 `lu` of float32(A_i), bit for bit (rustymath's own f32 factorization)."""
    import torch
    from bench import batched_workload
    a0, b0, ns, ks = batched_workload()
    plan = dv.BatchedMixedSolve(a0, b0, ns, ks)
    plan.run(graph=False)
    torch.cuda.synchronize()
    assert bool((plan.regular[:len(ns)] == 1).all())
    berr = plan.berr[:len(ns)].cpu().numpy()
    iters = plan.iters[:len(ns)].cpu().numpy()
    eps = np.finfo(np.float64).eps
    assert (berr <= 10 * ns * eps).all(), float((berr / (10 * ns * eps)).max())
    assert 1 <= iters.max() <= 10
    worst = 0.0
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        for k in np.unique(ks[idx]):
            sel = idx[ks[idx] == k]
            A = torch.stack([a0[plan.a_off[i]:plan.a_off[i] + n * n].view(n, n) for i in sel])
            B = torch.stack([b0[plan.b_off[i]:plan.b_off[i] + n * k].view(n, k) for i in sel])
            X = torch.stack([plan.x(i) for i in sel])
            r = (B - torch.bmm(A, X.contiguous())).flatten(1).norm(dim=1)
            e = (r / (A.flatten(1).norm(dim=1) * X.flatten(1).norm(dim=1))).cpu().numpy()
            assert (e <= 10 * n * eps).all(), (n, k, float(e.max()))
            worst = max(worst, float((e / (10 * n * eps)).max()))
    x1 = plan.xbuf.clone()
    plan.run(graph=True); plan.run(graph=True)
    torch.cuda.synchronize()
    assert bits_equal(plan.xbuf.cpu().numpy(), x1.cpu().numpy())
    for i in range(0, len(ns), 997):
        n = int(ns[i])
        f32, perm = plan.factors32(i)
        a32 = a0[plan.a_off[i]:plan.a_off[i] + n * n].view(n, n).float().cpu().numpy()
        wf, wperm, _ = oracle.lu(a32)
        assert (perm.cpu().numpy() == wperm).all(), f'cell {i}: pivots'
        assert bits_equal(f32.cpu().numpy()[wperm], wf.T), f'cell {i}: factors'
    plan.close()
    print(f'worst backward error / (10 n eps) = {worst:.3f}, corrections {np.bincount(iters)}')


def test_refinement_stops_at_its_cap_and_skips_singular_systems(dv):
    """max_iter = 0 applies no correction (the FP32 solve alone, its backward error reported); a singular system is
    marked not regular, with a NaN error and X left at zero, while the others in the batch converge."""
    import torch
    rng = np.random.default_rng(9)
    mats = [rng.standard_normal((n, n)) + np.sqrt(n) * np.eye(n) for n in (5, 100, 300)]
    mats.append(np.ones((50, 50)))
    rhs = [rng.standard_normal((m.shape[0], 3)) for m in mats]
    a, b, ns, ks = dv.pack(mats, rhs)
    p0 = dv.BatchedMixedSolve(a, b, ns, ks, max_iter=0)
    p0.run(); torch.cuda.synchronize()
    assert (p0.iters[:4].cpu().numpy() == 0).all()
    assert (p0.berr[:3].cpu().numpy() > 10 * np.array(ns[:3]) * np.finfo(np.float64).eps).any()
    p0.close()
    p = dv.BatchedMixedSolve(a, b, ns, ks)
    p.run(); torch.cuda.synchronize()
    reg = p.regular[:4].cpu().numpy()
    assert list(reg) == [1, 1, 1, 0]
    berr = p.berr[:4].cpu().numpy()
    assert (berr[:3] <= 10 * np.array(ns[:3]) * np.finfo(np.float64).eps).all() and np.isnan(berr[3])
    assert not p.x(3).abs().sum().item()
    p.close()


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_host_pointer_uneven_batch_matches_each_cell(dtype):
    """hs_*gesv_batched through the C ABI: systems of different orders from host pointers, each X_i the oracle's
    `solve` of that cell bit for bit (b's copy where it is not regular): orders past 512 and more than 8 right-hand
    sides run one by one, an empty system and a singular one count as not regular."""
    import ctypes
    import torch
    torch.cuda.set_device(0)
    from rustymath_amd import _lib
    from support import random
    lib = _lib.load()
    ns = [0, 1, 17, 64, 129, 300, 512, 530, 5, 200]
    ks = [1, 3, 0, 8, 2, 5, 1, 2, 9, 4]
    mats = [random(n, n, 100 + 7 * i, dtype) for i, n in enumerate(ns)]
    mats[5][40] = 0.0  # a zero row: singular, left to least squares
    rhs = [random(n, k, 300 + 11 * i, dtype) for i, (n, k) in enumerate(zip(ns, ks))]
    xs = [np.full((n, k), np.nan, dtype) for n, k in zip(ns, ks)]
    cnt = len(ns)
    n_arr, k_arr = np.array(ns, np.int64), np.array(ks, np.int64)
    ptrs = lambda arrs: (ctypes.c_void_p * cnt)(*[x.ctypes.data for x in arrs])
    regular = (ctypes.c_int * cnt)()
    f = getattr(lib, 'hs_dgesv_batched' if dtype == np.float64 else 'hs_sgesv_batched')
    print(f'host uneven batch {np.dtype(dtype).name}: {cnt} systems', flush=True)
    rc = f(cnt, n_arr.ctypes.data, k_arr.ctypes.data, ptrs(mats), ptrs(rhs), ptrs(xs), regular)
    assert rc == 0, _lib.load().hs_last_error()
    for i in range(cnt):
        want = oracle.solve(mats[i], rhs[i]) if ns[i] else None
        assert bool(regular[i]) == (want is not None), f'system {i}: regular {regular[i]}'
        if want is not None:
            assert bits_equal(xs[i], want), f'system {i}: solution'
        elif ns[i]:
            assert bits_equal(xs[i], rhs[i]), f'system {i}: b expected'
    assert not regular[0] and not regular[5]
    assert f(-1, None, None, None, None, None, None) == -1
    bad = np.array([3], np.int64)
    assert f(1, bad.ctypes.data, np.array([-1], np.int64).ctypes.data, None, None, None, None) == -3
