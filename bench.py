"""Benchmark: FP64-accurate solve GFLOP/s on one MI355X, low-precision LU + FP64 iterative refinement against the plain
FP64 LU, on the mixed-precision workload (n = 16384, A = U diag(s) Vᵀ with U, V products of 8 Householder reflectors
and s log-spaced from 1 to 1e-6, condition number 1e6; b iid N(0, 1)).

One step = one full solve from a resident A and b: cast A to FP32, LU on MFMA, solve, FP64 residuals and corrections
until the backward error ||b - A x|| / (||A|| ||x||) is at most 10·n·eps64. `value` is (2/3 n³ + 2 n²) / step time
(the HPL-MxP convention: the same FP64 flop count for every path, so speedups are time ratios). The plain FP64 path
(copy A, LU, solve) is timed on the same matrix for the comparison. Prints one JSON line.
"""
import argparse, json, math, os, sys, time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'FP64-accurate solve GFLOP/s: low-prec LU+IR vs plain FP64 LU; batched solves/s'
PEAK = {'fp64_mfma': 78.6e12, 'fp32_mfma': 157.3e12, 'hbm': 8.0e12}
# Measured per launch in one mixed solve of order n (profiles/mixed_solve_n16384_counters.md): HBM bytes from two
# rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) over `bench.py --steps 1 --warmup 0 --only-mixed --no-cpu`, each
# scaled by what the same counter reads on a 1 GiB device copy (`bench.py --copy-probe 1024`: FETCH_SIZE counts half the
# bytes read, WRITE_SIZE all the bytes written); and the MFMA units' busy fraction (SQ_VALU_MFMA_BUSY_CYCLES over the
# dispatch's GRBM_GUI_ACTIVE cycles times the 1024 SIMDs, rocprofv3's MfmaUtil) from a third pass.
TRAFFIC = {('panel_f32', 16384): 136345027, ('trail_f32', 16384): 489906952}
MFMA_BUSY = {('panel_f32', 16384): 0.2441, ('trail_f32', 16384): 0.6246}


def host_threads():
    """Host threads for the CPU baseline: the cores this process may run on, capped by OMP_NUM_THREADS where the host
    sets it (a shared GPU box grants each GPU a share of its cores); with the CPU model name, for the report."""
    visible = len(os.sched_getaffinity(0))
    cap = os.environ.get('OMP_NUM_THREADS')
    threads = min(visible, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else visible
    model = 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            model = next((l.split(':', 1)[1].strip() for l in f if l.startswith('model name')), model)
    except OSError:
        pass
    return threads, visible, model


def gaussian_system(n, seed):
    "The N(0, 1) variant of the mixed-precision workload: A (n×n) then b (n) iid N(0, 1) from torch's CUDA generator."
    import torch
    g = torch.Generator(device='cuda'); g.manual_seed(seed)
    a = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    return a, b


def householder_matrix(n, seed, cond=1e6, reflectors=8):
    """A = U diag(s) Vᵀ with U and V each a product of `reflectors` Householder reflectors H = I - 2 v vᵀ / vᵀv (v iid
    N(0, 1)) and s log-spaced from 1 to 1/cond, built in FP64 on the GPU by rank-1 updates, never forming U or V."""
    import torch
    g = torch.Generator(device='cuda'); g.manual_seed(seed)
    s = torch.logspace(0, -math.log10(cond), n, dtype=torch.float64, device='cuda')
    a = torch.diag(s)
    for _ in range(reflectors):       # A <- A H (columns): A - 2 (A v) vᵀ / vᵀv
        v = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
        a.addr_(a @ v, v, alpha=-2.0 / float(v @ v))
    for _ in range(reflectors):       # A <- H A (rows): A - 2 v (vᵀ A) / vᵀv
        v = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
        a.addr_(v, v @ a, alpha=-2.0 / float(v @ v))
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    return a.contiguous(), b


def batched_workload(count=10000, seed=3, nmin=8, nmax=512, kmax=8, dtype=None):
    """The batched uneven-solve workload: `count` float64 systems, n uniform in [nmin, nmax], k uniform in [1, kmax],
    A = G + sqrt(n) I with G iid N(0, 1), B iid N(0, 1). Orders and column counts come from numpy's generator (seed),
    the entries from torch's CUDA generator (seed), straight into the flat device buffers of BatchedSolve."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count).astype(np.int64)
    ks = rng.integers(1, kmax + 1, size=count).astype(np.int64)
    g = torch.Generator(device='cuda'); g.manual_seed(seed)
    dtype = dtype or torch.float64
    a = torch.randn(int((ns * ns).sum()), dtype=torch.float64, device='cuda', generator=g)
    b = torch.randn(int((ns * ks).sum()), dtype=torch.float64, device='cuda', generator=g)
    # The diagonals: offset of system i plus j (n_i + 1), for j < n_i.
    offs = np.concatenate([[0], np.cumsum(ns * ns)[:-1]])
    diag = np.concatenate([o + np.arange(n) * (n + 1) for o, n in zip(offs, ns)])
    shift = np.repeat(np.sqrt(ns.astype(np.float64)), ns)
    a[torch.from_numpy(diag).cuda()] += torch.from_numpy(shift).cuda()
    return a.to(dtype), b.to(dtype), ns, ks


def batched_bench(reps=3):
    """The batched uneven-solve workload on the GPU: the autotuned plan's graph replayed `reps` times on fresh copies
    of the systems, timed with HIP events; solves/s, GFLOP/s (Σ 2/3 n³ + 2 n² k) and the FP64 MFMA fraction."""
    import numpy as np
    import torch
    from rustymath_amd import device as dv
    a0, b0, ns, ks = batched_workload()
    flops = float((2 / 3 * ns.astype(np.float64) ** 3 + 2.0 * ns * ns * ks).sum())
    (pw, wg), tuned = dv.tune(a0, b0, ns, ks)
    a, b = a0.clone(), b0.clone()
    plan = dv.BatchedSolve(a, b, ns, ks, panel_width=pw, wgs_per_cu=wg)
    times = []
    for r in range(reps + 1):
        a.copy_(a0); b.copy_(b0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(); plan.run(); e1.record(); e1.synchronize()
        if r: times.append(e0.elapsed_time(e1) / 1e3)
    ok = bool((plan.regular[:len(ns)] == 1).all())
    plan.close()
    t = min(times)
    big = ns >= 256
    # The n >= 256 systems on their own, where the FP64 MFMA target applies: gathered from the same buffers.
    a_off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]]); b_off = np.concatenate([[0], np.cumsum(ns * ks)[:-1]])
    idx = np.nonzero(big)[0]
    sa0 = torch.cat([a0[int(a_off[i]):int(a_off[i] + ns[i] * ns[i])] for i in idx])
    sb0 = torch.cat([b0[int(b_off[i]):int(b_off[i] + ns[i] * ks[i])] for i in idx])
    sa, sb = sa0.clone(), sb0.clone()
    splan = dv.BatchedSolve(sa, sb, ns[idx], ks[idx], panel_width=pw, wgs_per_cu=wg)
    stimes = []
    for r in range(reps + 1):
        sa.copy_(sa0); sb.copy_(sb0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(); splan.run(); e1.record(); e1.synchronize()
        if r: stimes.append(e0.elapsed_time(e1) / 1e3)
    splan.close()
    st = min(stimes)
    sflops = float((2 / 3 * ns[idx].astype(np.float64) ** 3 + 2.0 * ns[idx] * ns[idx] * ks[idx]).sum())
    # The FP32-LU + FP64-refinement form on the same systems (inputs read, never written): one graph replay per run.
    mplan = dv.BatchedMixedSolve(a0, b0, ns, ks, panel_width=pw, wgs_per_cu=wg)
    mtimes = []
    for r in range(reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(); mplan.run(); e1.record(); e1.synchronize()
        if r: mtimes.append(e0.elapsed_time(e1) / 1e3)
    mt = min(mtimes)
    berr = mplan.berr[:len(ns)].cpu().numpy(); iters = mplan.iters[:len(ns)].cpu().numpy()
    mixed = {'ms': round(mt * 1e3, 3), 'solves_per_s': round(len(ns) / mt, 1), 'gflops': round(flops / mt / 1e9, 2),
             'frac_fp64_peak': round(flops / mt / PEAK['fp64_mfma'], 4),
             'all_regular': bool((mplan.regular[:len(ns)] == 1).all()),
             'worst_backward_error_over_10n_eps': round(float((berr / (10 * ns * 2.0 ** -52)).max()), 4),
             'corrections_histogram': [int(c) for c in np.bincount(iters)],
             'form': 'FP32 LU on MFMA (float32(A_i)), FP64 residuals and corrections on the GPU until the normwise '
                     'backward error <= 10 n_i eps64; same GFLOP count as the FP64 line'}
    mplan.close()
    return {'workload': f'{len(ns)} float64 systems, n uniform in [8, 512], k uniform in [1, 8], A = G + sqrt(n) I, '
                        'seed 3; one persistent launch replayed from a hipGraph', 'fp32_lu_fp64_refinement': mixed,
            'solves_per_s': round(len(ns) / t, 1), 'ms': round(t * 1e3, 3), 'gflops': round(flops / t / 1e9, 2),
            'frac_fp64_peak': round(flops / t / PEAK['fp64_mfma'], 4), 'all_regular': ok,
            'tuned': {'panel_width': pw, 'wgs_per_cu': wg},
            'tuning_ms': {f'{k[0]}x{k[1]}': round(v, 3) for k, v in tuned.items()},
            'systems_n_ge_256': int(big.sum()),
            'n_ge_256_alone': {'ms': round(st * 1e3, 3), 'solves_per_s': round(len(idx) / st, 1),
                               'gflops': round(sflops / st / 1e9, 2), 'frac_fp64_peak': round(sflops / st / PEAK['fp64_mfma'], 4)}}


def batched_cpu_baseline(sample=800, threads=None):
    """The oracle's per-cell solve (rustymath's LuWork::solve restated) over a sample of the batched workload's
    distribution, one system per host thread as rustymath's stacks run, in solves/s."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    threads, visible, model = host_threads() if threads is None else (threads, threads, 'unknown')
    oracle.set_threads(1)
    rng = np.random.default_rng(3)
    ns = rng.integers(8, 513, size=sample); ks = rng.integers(1, 9, size=sample)
    mats = [rng.standard_normal((n, n)) + np.sqrt(n) * np.eye(n) for n in ns]
    rhs = [rng.standard_normal((n, k)) for n, k in zip(ns, ks)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda i: oracle.solve(mats[i], rhs[i]), range(sample)))
    t = time.perf_counter() - t0
    oracle.set_threads(threads)
    flops = float((2 / 3 * ns.astype(np.float64) ** 3 + 2.0 * ns * ns * ks).sum())
    return {'value': round(sample / t, 1), 'unit': 'solves/s', 'gflops': round(flops / t / 1e9, 2), 'cores': threads,
            'host_cores_visible': visible, 'cpu_model': model, 'kind': 'port',
            'sample': f'{sample} systems of the same distribution (numpy seed 3), oracle solve, one system per thread '
                      f'on {threads} threads'}


def config1_bench(reps=5):
    """The reference's CPU-runnable case (its first configuration): rm.matmul of two 1024×1024 float64 N(0, 1) matrices
    (numpy seed 0, a then b) and rm.linalg.solve of a 1024×1024 N(0, 1) system with b 1024×1 (seed 1, A then b), through
    this library's Python surface (host arrays in and out, the PCIe copies included) and device-resident (inputs already
    in HBM), each the best of `reps` runs after one untimed; with the solve's backward error. Returns the timings and
    the results, which the CPU leg compares with the oracle."""
    import numpy as np
    import torch
    from rustymath_amd import linalg as la, device as dv
    n = 1024
    rng = np.random.default_rng(0); a = rng.standard_normal((n, n)); b = rng.standard_normal((n, n))
    rng = np.random.default_rng(1); A = rng.standard_normal((n, n)); v = rng.standard_normal((n, 1))

    def host_best(fn):
        ts = []
        for r in range(reps + 1):
            t0 = time.perf_counter(); fn(); t1 = time.perf_counter()
            if r: ts.append(t1 - t0)
        return min(ts)

    def dev_best(fn):
        ts = []
        for r in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(); fn(); e1.record(); e1.synchronize()
            if r: ts.append(e0.elapsed_time(e1) / 1e3)
        return min(ts)

    c = la.matmul(a, b); x = la.solve(A, v)
    t_mm_host = host_best(lambda: la.matmul(a, b)); t_solve_host = host_best(lambda: la.solve(A, v))
    ad, bd = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(); cd = torch.empty_like(ad)
    Ad, vd = torch.from_numpy(A).cuda(), torch.from_numpy(v).cuda()
    work, xd = torch.empty_like(Ad), torch.empty_like(vd)

    def dev_solve():
        work.copy_(Ad); xd.copy_(vd); dv.gesv(work, xd)
    t_mm_dev = dev_best(lambda: dv.gemm(ad, bd, cd)); t_solve_dev = dev_best(dev_solve)
    mm_flops, solve_flops = 2.0 * n ** 3, 2 / 3 * n ** 3 + 2.0 * n * n
    berr = float(np.linalg.norm(v - A @ x) / (np.linalg.norm(A) * np.linalg.norm(x)))
    line = {'workload': 'matmul 1024x1024x1024 float64 N(0,1) (numpy seed 0); solve 1024x1024 float64 N(0,1) with b '
                        '1024x1 (numpy seed 1, A then b); through linalg.matmul / linalg.solve (host arrays) and '
                        'device-resident',
            'matmul_host_ms': round(t_mm_host * 1e3, 3), 'matmul_host_gflops': round(mm_flops / t_mm_host / 1e9, 1),
            'matmul_device_ms': round(t_mm_dev * 1e3, 3), 'matmul_device_gflops': round(mm_flops / t_mm_dev / 1e9, 1),
            'solve_host_ms': round(t_solve_host * 1e3, 3), 'solve_host_gflops': round(solve_flops / t_solve_host / 1e9, 1),
            'solve_device_ms': round(t_solve_dev * 1e3, 3),
            'solve_device_gflops': round(solve_flops / t_solve_dev / 1e9, 1),
            'backward_error': berr, 'backward_error_bound': 10 * n * float(np.finfo(np.float64).eps)}
    return line, (a, b, c, A, v, x)


def config1_cpu_baseline(line, data, reps=5):
    """The CPU leg of the first configuration: the oracle (the C restatement of rustymath's CPU kernels) on the same
    1024 inputs, on one host thread and on every host thread, best of `reps`; and whether the GPU's product and solve
    equal the oracle's bit for bit."""
    import oracle
    a, b, c, A, v, x = data
    threads, visible, model = host_threads()
    out = {}
    for t in (1, threads):
        oracle.set_threads(t)
        tm = min(_timed(lambda: oracle.matmul(a, b)) for _ in range(reps))
        ts = min(_timed(lambda: oracle.solve(A, v)) for _ in range(reps))
        out[f'{t}_thread' + ('s' if t > 1 else '')] = {'matmul_gflops': round(2.0 * 1024 ** 3 / tm / 1e9, 2),
                                                       'solve_gflops': round((2 / 3 * 1024 ** 3 + 2 * 1024 ** 2) / ts / 1e9, 2)}
    same = lambda p, q: p.shape == q.shape and p.tobytes() == q.tobytes()
    line['bits_equal_oracle'] = {'matmul': same(c, oracle.matmul(a, b)), 'solve': same(x, oracle.solve(A, v))}
    line['cpu_baseline'] = dict(out, kind='port', cores=threads, host_cores_visible=visible, cpu_model=model,
                                sample='the same inputs; oracle matmul and solve, best of 5')
    return line


def _timed(fn):
    t0 = time.perf_counter(); fn(); return time.perf_counter() - t0


def adjoint_bench(reps=3):
    """loss = sum(w ⊙ solve(A, C D)) and its gradients w.r.t. A, C, D through the hand-written adjoint: 256 systems
    of order 256 and one of order 4096, C n×n, D and w n×16, A = G + sqrt(n) I; milliseconds per loss + gradient."""
    import torch
    from rustymath_amd import adjoint
    out = {}
    for n, batch in ((256, 256), (4096, 0)):
        g = torch.Generator(device='cuda'); g.manual_seed(3)
        shape = (batch,) if batch else ()
        a = torch.randn(*shape, n, n, dtype=torch.float64, device='cuda', generator=g)
        a += n ** 0.5 * torch.eye(n, dtype=torch.float64, device='cuda')
        c = torch.randn(*shape, n, n, dtype=torch.float64, device='cuda', generator=g)
        d = torch.randn(*shape, n, 16, dtype=torch.float64, device='cuda', generator=g)
        w = torch.randn(*shape, n, 16, dtype=torch.float64, device='cuda', generator=g)
        best = math.inf
        for r in range(reps + 1):
            xs = [x.clone().requires_grad_(True) for x in (a, c, d)]
            torch.cuda.synchronize(); t0 = time.perf_counter()
            (w * adjoint.solve(xs[0], adjoint.matmul(xs[1], xs[2]))).sum().backward()
            torch.cuda.synchronize()
            if r: best = min(best, time.perf_counter() - t0)
        flops = max(batch, 1) * (2 / 3 * n ** 3 + 12.0 * n * n * 16)
        out[f'{max(batch, 1)}x{n}_ms'] = round(best * 1e3, 3)
        out[f'{max(batch, 1)}x{n}_gflops'] = round(flops / best / 1e9, 1)
        out[f'{max(batch, 1)}x{n}_frac_fp64_peak'] = round(flops / best / PEAK['fp64_mfma'], 4)
    out['workload'] = 'loss = sum(w * solve(A, C D)) + gradients in A, C, D; C n x n, D and w n x 16; float64'
    out['flops'] = ('per system 2/3 n^3 (LU) + 12 n^2 k: the products C D, -Bbar X^T, Bbar D^T, C^T Bbar and the solve '
                    'and transposed solve, 2 n^2 k each (k = 16)')
    return out


def precision_variants(a, b, x64, n, reps=2):
    """The mixed-precision workload's variants: FP32 factors with classical refinement and BF16 trailing products with
    GMRES-based refinement, on the cond-1e6 matrix (a, b, its FP64 LU solution x64) and on an iid N(0, 1) matrix of the
    same order (seed 4); each the best of `reps` solves, with refinement and GMRES step counts, the backward error and
    the forward error against the plain FP64 LU solution."""
    import torch
    from rustymath_amd import device as dv
    tol = 10 * n * torch.finfo(torch.float64).eps
    flops = 2 / 3 * n ** 3 + 2 * n * n
    ag, bg = gaussian_system(n, 4)
    fg, pg = dv.getrf(ag.clone())
    xg = bg.reshape(n, 1).clone(); dv.getrs(fg, pg, xg); xg = xg.reshape(n)
    del fg
    out = {}
    for name, (A, B, X) in (('cond1e6', (a, b, x64)), ('gaussian', (ag, bg, xg))):
        for prec in ('fp32', 'bf16x3', 'bf16'):
            best, res = math.inf, None
            for _ in range(reps):
                torch.cuda.synchronize(); t0 = time.perf_counter()
                if prec == 'fp32': res = dv.dsgesv(A, B, tol=tol)
                else: res = dv.dbgesv(A, B, tol=tol, max_iter=10, terms=3 if prec == 'bf16x3' else 1)
                torch.cuda.synchronize(); best = min(best, time.perf_counter() - t0)
            x = res[0]
            line = {'ms': round(best * 1e3, 3), 'gflops': round(flops / best / 1e9, 1), 'refinement_steps': res[1],
                    'backward_error': res[-1], 'forward_error_vs_fp64_lu': float((x - X).norm() / X.norm())}
            if prec != 'fp32': line['gmres_steps'] = res[2]
            out[f'{name}_{prec}'] = line
    out['workload'] = (f'n={n}: the cond-1e6 matrix of the headline and an iid N(0,1) matrix (seed 4); fp32 = FP32 LU + '
                       'classical FP64 refinement; bf16x3 = FP32 panels with trailing products on split-BF16 operands '
                       '(hi + lo pairs, three BF16 MFMA products, FP32 accumulation) + FP64 GMRES-based refinement; bf16 = '
                       'the same with operands rounded once to BF16; restart 40, at most 10 refinement steps; tolerance '
                       '10 n eps64')
    return out


def plain_fp64_bench(n=8192, k=16, reps=3):
    """The plain FP64 workload on the GPU: matmul of two n×n float64 N(0, 1) matrices (2 n³ flops), and the LU solve
    of an n×n float64 N(0, 1) system with k right-hand sides (2/3 n³ + 2 n² k flops: copy A, factor, solve), each the
    best of `reps` runs timed with HIP events; with the product's error against the k·eps·|A||B| bound on a sample of
    rows and the solve's normwise backward error ||B - A X||_F / (||A||_F ||X||_F)."""
    import torch
    from rustymath_amd import device as dv
    g = torch.Generator(device='cuda'); g.manual_seed(0)
    a = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    b = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    c = torch.empty(n, n, dtype=torch.float64, device='cuda')

    def timed(fn):
        best = math.inf
        for r in range(reps + 1):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(); fn(); e1.record(); e1.synchronize()
            if r: best = min(best, e0.elapsed_time(e1) / 1e3)
        return best

    t_mm = timed(lambda: dv.gemm(a, b, c))
    rows = torch.arange(0, n, n // 8, device='cuda')
    cols = torch.arange(0, n, n // 256, device='cuda')
    ar, bc = a[rows].cpu().numpy(), b[:, cols].cpu().numpy()
    exact = ar.astype('longdouble') @ bc.astype('longdouble')
    bound = n * float(torch.finfo(torch.float64).eps) * (abs(ar) @ abs(bc))
    mm_ratio = float((abs(c[rows][:, cols].cpu().numpy() - exact) / bound).max())
    del b, c
    g.manual_seed(1)
    A = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    B = torch.randn(n, k, dtype=torch.float64, device='cuda', generator=g)
    work, X = torch.empty_like(A), torch.empty_like(B)

    def solve():
        work.copy_(A)
        f, perm = dv.getrf(work)
        X.copy_(B)
        dv.getrs(f, perm, X)
    t_lu = timed(solve)
    berr = float((B - A @ X).norm() / (A.norm() * X.norm()))
    lu_flops = 2 / 3 * n ** 3 + 2 * n * n * k
    return {'workload': f'matmul {n}x{n}x{n} float64 N(0,1) seed 0; LU solve {n}x{n} float64 N(0,1) with {k} right-hand '
                        'sides, seed 1 (copy, factor, solve)',
            'matmul_ms': round(t_mm * 1e3, 3), 'matmul_gflops': round(2 * n ** 3 / t_mm / 1e9, 1),
            'matmul_frac_fp64_peak': round(2 * n ** 3 / t_mm / PEAK['fp64_mfma'], 4),
            'matmul_max_error_over_bound': mm_ratio,
            'lu_solve_ms': round(t_lu * 1e3, 3), 'lu_solve_gflops': round(lu_flops / t_lu / 1e9, 1),
            'lu_solve_frac_fp64_peak': round(lu_flops / t_lu / PEAK['fp64_mfma'], 4),
            'backward_error': berr, 'backward_error_bound': 10 * n * float(torch.finfo(torch.float64).eps)}


def cpu_baseline(n=6144, runs=3, threads=None):
    """The oracle's LU solve (the C restatement of rustymath's blocked CPU kernels) on a bounded sample: an FP64 N(0, 1)
    system of order n on the host's threads (host_threads), best of `runs`, in GFLOP/s (2/3 n³ + 2 n²)."""
    import numpy as np
    import oracle
    threads, visible, model = host_threads() if threads is None else (threads, threads, 'unknown')
    oracle.set_threads(threads)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((n, n)); b = rng.standard_normal(n)
    best = math.inf
    for _ in range(runs):
        t0 = time.perf_counter(); oracle.solve(a, b); best = min(best, time.perf_counter() - t0)
    # Fidelity point beside the reference's published single-thread figure (lu(256).solve at 24.5 GFLOP/s on a
    # Ryzen 9 7950X3D, tests/speed.txt:149-150): the same restatement at n = 256 on one thread, best of 5.
    oracle.set_threads(1)
    a1 = rng.standard_normal((256, 256)); b1 = rng.standard_normal(256)
    one = math.inf
    for _ in range(5):
        t0 = time.perf_counter(); oracle.solve(a1, b1); one = min(one, time.perf_counter() - t0)
    oracle.set_threads(threads)
    return {'value': round((2 / 3 * n ** 3 + 2 * n * n) / best / 1e9, 2), 'unit': 'GFLOP/s', 'cores': threads,
            'host_cores_visible': visible, 'cpu_model': model, 'kind': 'port',
            'single_thread_n256_gflops': round((2 / 3 * 256 ** 3 + 2 * 256 * 256) / one / 1e9, 2),
            'sample': f'oracle (C restatement of rustymath lu_factor/lu_solve, AVX2 6x8 tiles) FP64 solve n={n}, 1 RHS, '
                      f'best of {runs}, {threads} threads'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=16384)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--only-mixed', action='store_true', help='the mixed-precision solve only (for profiling runs)')
    ap.add_argument('--part', choices=['plain', 'batched', 'adjoint', 'config1'], default=None,
                    help='one secondary line only, without the headline (for profiling runs)')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only copy MIB MiB between two device buffers 5 times: the known-size kernel that scales the '
                         'HBM byte counters of a profiling run')
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(5): dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_probe_bytes_per_launch': {'read': args.copy_probe << 20, 'written': args.copy_probe << 20}}))
        return
    if args.gpus != 1:
        print(f'bench.py runs on one GPU: this build measures the single-GPU path only (asked for --gpus {args.gpus})',
              file=sys.stderr)
        sys.exit(2)
    import torch
    from rustymath_amd import device as dv, _lib
    n = args.n
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    if args.part:
        parts = {'plain': plain_fp64_bench, 'batched': batched_bench, 'adjoint': adjoint_bench,
                 'config1': lambda: config1_bench()[0]}
        print(json.dumps({args.part: parts[args.part]()}))
        return
    a, b = householder_matrix(n, seed=3)
    torch.cuda.synchronize()
    tol = 10 * n * torch.finfo(torch.float64).eps
    flops = 2 / 3 * n ** 3 + 2 * n * n

    # Plain FP64: copy A (the LU works in place), factor, solve.
    work = torch.empty_like(a)
    def fp64_step():
        work.copy_(a)
        f, perm = dv.getrf(work)
        x = b.reshape(n, 1).clone()
        dv.getrs(f, perm, x)
        return x.reshape(n)
    # The same warm-up and steps as the mixed solve: mean and best.
    for _ in range(max(args.warmup, 1)):
        x64 = fp64_step()
    t64s = []
    for _ in range(args.steps):
        torch.cuda.synchronize(); t0 = time.perf_counter()
        x64 = fp64_step()
        torch.cuda.synchronize(); t64s.append(time.perf_counter() - t0)
    t64, t64_best = sum(t64s) / len(t64s), min(t64s)
    r64 = ((b - a @ x64).norm() / (a.norm() * x64.norm())).item()

    # Low precision + refinement: the timed steps. The LU's live profile (dispatch-bound events on the panel and
    # trail_rest launches) runs during the warm-up too, so that the timed steps reuse its events rather than create them.
    lib = _lib.load()
    import ctypes
    lib.hs_profile_lu(1, None)
    for _ in range(max(args.warmup, 1)):
        dv.dsgesv(a, b, tol=tol)
    torch.cuda.synchronize()
    def prof(fn):
        v = [ctypes.c_double(0), ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_double(0)]
        fn(0, *[ctypes.byref(x) for x in v])
        return v[0].value, v[1].value, v[2].value, v[3].value
    times, iters, berr, x = [], 0, 0.0, None
    # The events cost about 2 % of a step (a timestamp before and after each of 256 panel launches), so the last timed
    # step alone carries them: the launch durations are that step's, the step times every step's.
    lu_prof = (ctypes.c_double * 8)()
    lib.hs_profile_lu(0, None)  # clears the warm-up's records, keeping their events for reuse
    for step in range(args.steps):
        if step == args.steps - 1: lib.hs_profile_lu(1, None)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, iters, berr = dv.dsgesv(a, b, tol=tol)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    lib.hs_profile_lu(0, lu_prof)
    tms, tl, tf, tb, pms, pl, pfl, wfl = list(lu_prof)
    tl, pl = int(tl), int(pl)
    t = sum(times) / len(times)
    fwd = ((x - x64).norm() / x64.norm()).item()

    # The rooflines, from live timings of the timed steps: the library brackets every panel and trail_rest launch with
    # HIP events recorded at the dispatch's start and end on its stream (hs_profile_lu). The dominant kernel by time is
    # the panel kernel: its launches factor the 64-column panels on one XCD and, on the other XCDs, compute trailing tiles
    # of the previous block's product (look-ahead workers), so its flops are the panels' own and the workers' tiles
    # (2 rows cols k each), against the FP32 MFMA peak. trail_rest takes the tiles the workers left; the product
    # launches (gemm_mfma, the next block's columns) are timed over one more solve (hs_profile_gemm), as are the panel's
    # per-phase counters (workgroup 0; they slow it). `traffic` is each kernel's measured HBM bytes per launch (TRAFFIC).
    ticks = (ctypes.c_int64 * 8)()
    lib.hs_profile_gemm(1, None, None, None, None)
    lib.hs_profile_panel(1, None)
    dv.dsgesv(a, b, tol=tol); torch.cuda.synchronize()
    gms, gl, gf, gb = prof(lib.hs_profile_gemm)
    lib.hs_profile_panel(0, ticks)
    factor_flops = 2 / 3 * n ** 3
    panel_achieved = (pfl + wfl) / (pms / 1e3) / 1e12
    trail_achieved = tf / (tms / 1e3) / 1e12
    roof = {'bound': 'mfma', 'achieved': round(panel_achieved, 2), 'peak': PEAK['fp32_mfma'] / 1e12, 'unit': 'TFLOP/s',
            'frac': round(panel_achieved * 1e12 / PEAK['fp32_mfma'], 4), 'traffic': TRAFFIC.get(('panel_f32', n)),
            'kernel': 'panel_kernel<float> (the panels on one XCD, trailing tiles of the previous block on the other seven)',
            'launches': pl, 'flop_per_launch': round((pfl + wfl) / pl), 'avg_launch_us': round(pms / pl * 1e3, 2),
            'mfma_busy_rocprof': MFMA_BUSY.get(('panel_f32', n)),
            'share_of_step_time': round(pms / (t * 1e3), 4),
            'worker_flop_share': round(wfl / (pfl + wfl), 4),
            'share_of_factorization_flops': round((pfl + wfl) / factor_flops, 4),
            'other_kernels': [
                {'kernel': 'trail_rest<float> (the trailing tiles the panel launches left)', 'bound': 'mfma',
                 'achieved': round(trail_achieved, 2), 'frac': round(trail_achieved * 1e12 / PEAK['fp32_mfma'], 4),
                 'launches': tl, 'flop_per_launch': round(tf / tl), 'algorithmic_bytes_per_launch': round(tb / tl),
                 'avg_launch_us': round(tms / tl * 1e3, 2), 'traffic': TRAFFIC.get(('trail_f32', n)),
                 'mfma_busy_rocprof': MFMA_BUSY.get(('trail_f32', n)),
                 'share_of_step_time': round(tms / (t * 1e3), 4),
                 'share_of_factorization_flops': round(tf / factor_flops, 4)},
                {'kernel': 'gemm_mfma<float> (the next block\'s columns)', 'launches': gl, 'ms': round(gms, 3),
                 'achieved': round(gf / (gms / 1e3) / 1e12, 2), 'frac': round(gf / (gms / 1e3) / PEAK['fp32_mfma'], 4),
                 'share_of_factorization_flops': round(gf / factor_flops, 4)}]}
    panel_s = (sum(ticks[:5]) + ticks[7]) / 1e8
    panel = {'kernel': 'panel_kernel<float>', 'bound': 'latency (one pivot exchange per column)',
             'column_loop_ms': round(panel_s * 1e3, 3), 'us_per_column': round(panel_s / n * 1e6, 3),
             'phases_ms': {k: round(ticks[q] / 1e5, 3) for k, q in (('reduce', 0), ('publish', 7), ('update', 1),
                                                                     ('exchange', 2), ('barrier', 3), ('pivot_update', 4),
                                                                     ('row_loads', 5), ('row_stores', 6))}}
    batched = adj = plain = variants = cfg1 = None
    if not args.only_mixed:
        variants = precision_variants(a, b, x64, n)
        batched = batched_bench()
        adj = adjoint_bench()
        plain = plain_fp64_bench()
        cfg1, cfg1_data = config1_bench()

    out = {
        'metric': METRIC, 'value': round(flops / t / 1e9, 2), 'unit': 'GFLOP/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(t * 1e3, 3), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32 LU on MFMA + f64 refinement', 'data': 'synthetic',
        'config': {'workload': f'mixed precision solve, n={n}, A = U diag(s) V^T (8+8 Householder reflectors, '
                               f's = logspace(0, -6)), cond 1e6, b N(0,1), seed 3', 'n': n, 'rhs': 1},
        'fp64_lu_solve_gflops': round(flops / t64 / 1e9, 2), 'fp64_lu_solve_ms': round(t64 * 1e3, 3),
        'speedup_vs_fp64_lu': round(t64 / t, 3), 'best_ms_per_step': round(min(times) * 1e3, 3),
        'fp64_lu_solve_best_ms': round(t64_best * 1e3, 3), 'speedup_vs_fp64_lu_best': round(t64_best / min(times), 3),
        'refinement_solves': iters, 'backward_error': berr, 'backward_error_fp64_lu': r64, 'tolerance': tol,
        'forward_error_vs_fp64_lu': fwd, 'roofline': roof, 'panel': panel, 'precision_variants': variants,
        'batched': batched, 'adjoint': adj, 'plain_fp64': plain, 'config1': cfg1,
    }
    if not args.no_cpu and not args.only_mixed:
        out['cpu_baseline'] = cpu_baseline()
        out['batched']['cpu_baseline'] = batched_cpu_baseline()
        out['config1'] = config1_cpu_baseline(cfg1, cfg1_data)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
