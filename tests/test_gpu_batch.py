"""Batched CG engine: every system's solution matches the FP64 oracle CG, mixed systems
reach the true 1e-10 residual, and a system's result is bitwise the same alone, in a mix
and when admitted mid-flight (continuous batching)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import generators as gen
from tensors_amd.batch import BatchCG

pytestmark = pytest.mark.gpu


def systems(count, seed=4004):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        n_target = int(np.exp(rng.uniform(np.log(300), np.log(6000))))
        dims = 2 if rng.random() < 0.5 else 3
        side = max(2, int(round(n_target ** (1.0 / dims))))
        sides = (side,) * dims
        rp, ci, val = gen.weighted_grid(dims, sides, 1e-3, seed=seed + s)
        n = len(rp) - 1
        k = int(rng.integers(1, 9))
        B = gen.uniform(seed * 10 + s, n * k).reshape(n, k)
        d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
        out.append((rp, ci, val, d, B, n, k))
    return out


def to_dev(sysd, dev):
    rp, ci, val, d, B, n, k = sysd
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dt)
    return (T(rp, torch.int32), T(ci, torch.int32), T(val, torch.float64), T(d, torch.float64),
            T(B, torch.float64), torch.zeros(n, k, dtype=torch.float64, device=dev))


def run_batch(sysl, dev, mixed, tol, order=None, staged=False):
    bt = BatchCG()
    devs = [to_dev(s, dev) for s in sysl]
    ids = []
    idx = order if order is not None else range(len(sysl))
    half = len(sysl) // 2
    for j, i in enumerate(idx):
        rp, ci, val, d, B, X = devs[i]
        mode = mixed[i] if isinstance(mixed, (list, tuple)) else int(mixed)
        ids.append((i, bt.add(rp, ci, val, B, X, tol, 20000, diag=d, mixed=mode == 1, true_residual=mode == 2)))
        if staged and j == half:
            bt.step(3)  # admit the rest while the first half is running
    bt.run(8)
    res = {i: bt.result(sid) for i, sid in ids}
    X = {i: devs[i][5].cpu().numpy() for i, _ in ids}
    bt.close()
    return X, res


@pytest.mark.parametrize("mixed", [0, 1, 2], ids=["fp64", "mixed", "fp64_true_residual"])
def test_batch_matches_oracle(cuda, mixed):
    sysl = systems(10)
    X, res = run_batch(sysl, cuda, mixed, 1e-10)
    for i, (rp, ci, val, d, B, n, k) in enumerate(sysl):
        assert res[i]["succeeded"], (i, res[i])
        A = sp.csr_matrix((val, ci, rp), shape=(n, n))
        true = np.linalg.norm(B - A @ X[i], axis=0) / np.linalg.norm(B, axis=0)
        if mixed:
            # mixed systems and FP64 systems with the true-residual stop stop on the
            # FP64 true residual
            assert np.all(res[i]["rel_res"] <= 1e-10), (i, res[i])
            assert np.all(true <= 1.5e-10), (i, true)
            Xs = np.linalg.solve(A.toarray(), B) if n <= 3000 else None
            if Xs is not None:
                assert rel_err(X[i], Xs) < 1e-6
        else:
            # plain FP64 systems run the reference recurrences: same iteration count as the
            # oracle CG at the same tolerance, and the same iterate up to rounding
            Xo = np.zeros((n, k))
            ok, it, _ = oracle.cg(n, rp, ci, val, B, Xo, k, 20000, 1e-10, prec_diag=d, threads=4)
            assert ok
            assert abs(res[i]["iterations"] - it) <= 1, (i, res[i]["iterations"], it)
            assert rel_err(X[i], Xo) < 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_alone_equals_mix_and_midflight(cuda, mixed):
    sysl = systems(8, seed=777)
    Xmix, _ = run_batch(sysl, cuda, mixed, 1e-10)
    Xstaged, _ = run_batch(sysl, cuda, mixed, 1e-10, order=list(reversed(range(8))), staged=True)
    for j in (0, 3, 7):
        Xalone, _ = run_batch([sysl[j]], cuda, mixed, 1e-10)
        assert np.array_equal(Xalone[0], Xmix[j])
        assert np.array_equal(Xstaged[j], Xmix[j])


def test_batch_spmm_kernels_agree(cuda):
    """The streaming and the row-per-lane SpMM kernels of the engine run the same
    recurrences: same iteration counts (within one) and the same solutions to rounding."""
    from tensors_amd import _capi as capi

    sysl = systems(6, seed=99)
    before = capi.lib().tg_tune_get(b"batch_stream")
    try:
        capi.call("tg_tune_set", b"batch_stream", 1)
        Xs, rs = run_batch(sysl, cuda, False, 1e-10)
        capi.call("tg_tune_set", b"batch_stream", 0)
        Xr, rr = run_batch(sysl, cuda, False, 1e-10)
    finally:
        capi.call("tg_tune_set", b"batch_stream", before)
    for i in range(len(sysl)):
        assert abs(rs[i]["iterations"] - rr[i]["iterations"]) <= 1
        assert rel_err(Xs[i], Xr[i]) < 1e-8


@pytest.mark.parametrize("mixed", [False, True])
def test_split_and_fused_direction_forms_are_bitwise_equal(cuda, mixed):
    """The split iteration (direction pass, SpMM of p, update) and the fused one (direction
    rebuilt inside the SpMM) compute p, u and every partial with the same operations in the
    same order: identical iterates, iteration counts and replacements."""
    from tensors_amd import _capi as capi

    sysl = systems(6, seed=123)
    before = (capi.lib().tg_tune_get(b"batch_split"), capi.lib().tg_tune_get(b"batch_stream"))
    try:
        capi.call("tg_tune_set", b"batch_stream", 0)
        capi.call("tg_tune_set", b"batch_split", 1)
        Xs, rs = run_batch(sysl, cuda, mixed, 1e-10)
        capi.call("tg_tune_set", b"batch_split", 0)
        Xf, rf = run_batch(sysl, cuda, mixed, 1e-10)
    finally:
        capi.call("tg_tune_set", b"batch_split", before[0])
        capi.call("tg_tune_set", b"batch_stream", before[1])
    for i in range(len(sysl)):
        assert rs[i]["iterations"] == rf[i]["iterations"]
        assert np.array_equal(Xs[i], Xf[i])


def test_batch_modes_in_one_mix_equal_alone(cuda):
    """FP64 (reference stop), mixed and FP64-true-residual systems in one batch: each one's
    result is bitwise what it gives alone in its own mode."""
    sysl = systems(6, seed=31)
    modes = [0, 1, 2, 2, 1, 0]
    Xmix, rmix = run_batch(sysl, cuda, modes, 1e-10)
    for j in range(len(sysl)):
        Xa, ra = run_batch([sysl[j]], cuda, modes[j], 1e-10)
        assert np.array_equal(Xa[0], Xmix[j]), (j, modes[j])
        assert ra[0]["iterations"] == rmix[j]["iterations"]


def wide_system(n=70000, k=2, seed=9):
    """A tridiagonal SPD system plus a symmetric coupling of the first and the last row:
    a column offset of n - 1 > 32767, so the engine keeps its int32 column indices."""
    main = np.full(n, 2.5)
    off = np.full(n - 1, -1.0)
    A = sp.diags([off, main, off], [-1, 0, 1], format="lil")
    A[0, n - 1] = -0.25
    A[n - 1, 0] = -0.25
    A = A.tocsr()
    A.sort_indices()
    B = gen.uniform(seed, n * k).reshape(n, k)
    return (A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, 1.0 / A.diagonal(), B, n, k)


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_idx16_offsets_are_bitwise_neutral(cuda, mixed):
    """16-bit column offsets only change how the SpMM reads the pattern: results are bitwise
    those with int32 indices, and a system whose offsets do not fit falls back to int32."""
    from tensors_amd import _capi as capi

    sysl = systems(4, seed=55) + [wide_system()]
    before = capi.lib().tg_tune_get(b"batch_idx16")
    try:
        capi.call("tg_tune_set", b"batch_idx16", 1)
        X16, r16 = run_batch(sysl, cuda, mixed, 1e-10)
        capi.call("tg_tune_set", b"batch_idx16", 0)
        X32, r32 = run_batch(sysl, cuda, mixed, 1e-10)
    finally:
        capi.call("tg_tune_set", b"batch_idx16", before)
    for i in range(len(sysl)):
        assert r16[i]["succeeded"] and r16[i]["iterations"] == r32[i]["iterations"]
        assert np.array_equal(X16[i], X32[i])
    rp, ci, val, d, B, n, k = sysl[-1]
    A = sp.csr_matrix((val, ci, rp), shape=(n, n))
    Xw = X16[len(sysl) - 1]
    assert np.max(np.linalg.norm(B - A @ Xw, axis=0) / np.linalg.norm(B, axis=0)) < 1e-8


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_graph_replayed_steps_equal_eager_steps(cuda, mixed):
    """Steps run as captured hipGraphs (replayed every step, with the host's polls, parks and
    residual replacements between the replays, and captured again whenever the tile list
    changes) give bitwise the results of eager launches."""
    from tensors_amd import _capi as capi

    sysl = systems(6, seed=202)
    before = capi.lib().tg_tune_get(b"batch_graph")
    try:
        capi.call("tg_tune_set", b"batch_graph", 1)
        Xg, rg = run_batch(sysl, cuda, mixed, 1e-10, staged=True)
        capi.call("tg_tune_set", b"batch_graph", 0)
        Xe, re_ = run_batch(sysl, cuda, mixed, 1e-10, staged=True)
    finally:
        capi.call("tg_tune_set", b"batch_graph", before)
    for i in range(len(sysl)):
        assert rg[i]["iterations"] == re_[i]["iterations"] and rg[i]["outer"] == re_[i]["outer"]
        assert np.array_equal(Xg[i], Xe[i])


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_xcd_tile_order_is_bitwise_neutral(cuda, mixed):
    """The spread tile order (knob batch_xcd_order = 1: each XCD's tiles drawn from the
    whole list instead of a contiguous run) changes only which XCD runs a tile: every
    reduction adds the tile partials by tile index, so the results are bitwise those of the
    default order."""
    from tensors_amd import _capi as capi

    sysl = systems(9, seed=303)
    before = capi.lib().tg_tune_get(b"batch_xcd_order")
    try:
        capi.call("tg_tune_set", b"batch_xcd_order", 1)
        Xa, ra = run_batch(sysl, cuda, mixed, 1e-10, staged=True)
        capi.call("tg_tune_set", b"batch_xcd_order", 0)
        Xb, rb = run_batch(sysl, cuda, mixed, 1e-10, staged=True)
    finally:
        capi.call("tg_tune_set", b"batch_xcd_order", before)
    for i in range(len(sysl)):
        assert ra[i]["iterations"] == rb[i]["iterations"] and ra[i]["outer"] == rb[i]["outer"]
        assert np.array_equal(Xa[i], Xb[i])


def nine_point_system(side=60, k=2, seed=12):
    """An SPD 9-point grid operator (3 x 3 neighbourhood, -1 off the diagonal, 8 + 1e-2 on
    it): 9 nonzeros per interior row, more than the LDS-staged SpMM holds for a 256-row
    segment, so those segments take its global-load path."""
    n = side * side
    A = sp.lil_matrix((n, n))
    for a in range(side):
        for b in range(side):
            i = a * side + b
            A[i, i] = 8.01
            for da in (-1, 0, 1):
                for db in (-1, 0, 1):
                    if (da or db) and 0 <= a + da < side and 0 <= b + db < side:
                        A[i, (a + da) * side + b + db] = -1.0
    A = A.tocsr()
    A.sort_indices()
    B = gen.uniform(seed, n * k).reshape(n, k)
    return (A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, 1.0 / A.diagonal(), B, n, k)


@pytest.mark.parametrize("mixed", [0, 1, 2])
@pytest.mark.parametrize("tile_rows", [1024, 4096])
def test_spmm_kernel_forms_are_bitwise_equal(cuda, mixed, tile_rows):
    """The SpMM kernels of the split iteration -- the row loop, the software-pipelined row
    loop (row pointers through LDS in segments of 2048 rows, the next row's indices loaded
    during this row's gathers), the LDS-staged one (a 256-row segment's offsets and values
    copied with coalesced loads; segments with more nonzeros than it holds, and systems with
    int32 indices, on global loads) -- give bitwise the same results, with tiles of one
    pipelined segment and of two."""
    from tensors_amd import _capi as capi

    sysl = systems(5, seed=808) + [wide_system(n=9000, k=3), nine_point_system(), wide_system(n=40000, k=1)]
    before = (capi.lib().tg_tune_get(b"batch_spmm"), capi.lib().tg_tune_get(b"batch_tile_rows"))
    out = {}
    try:
        capi.call("tg_tune_set", b"batch_tile_rows", tile_rows)
        for form in (0, 1, 2):
            capi.call("tg_tune_set", b"batch_spmm", form)
            out[form] = run_batch(sysl, cuda, mixed, 1e-10)
    finally:
        capi.call("tg_tune_set", b"batch_spmm", before[0])
        capi.call("tg_tune_set", b"batch_tile_rows", before[1])
    X0, r0 = out[0]
    for form in (1, 2):
        X, r = out[form]
        for i in range(len(sysl)):
            assert r[i]["succeeded"] and r[i]["iterations"] == r0[i]["iterations"], (form, i)
            assert np.array_equal(X[i], X0[i]), (form, i)


def tiny_system(n, k, seed):
    """A small SPD tridiagonal system (n = 1 is the 1 x 1 matrix [2])."""
    A = sp.diags([np.full(n - 1, -1.0), np.full(n, 2.0 + 0.1 * seed), np.full(n - 1, -1.0)], [-1, 0, 1],
                 format="csr") if n > 1 else sp.csr_matrix(np.array([[2.0]]))
    A.sort_indices()
    B = gen.uniform(seed, n * k).reshape(n, k)
    return (A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, 1.0 / A.diagonal(), B, n, k)


@pytest.mark.parametrize("mixed", [0, 1, 2])
def test_batch_edge_cases(cuda, mixed):
    """Tiny systems (n = 1, 2, 3; 1 and 8 right-hand sides), a zero right-hand side (done at
    setup, X = 0), a right-hand side with a NaN (reported as failed, nothing iterated) and
    max_iter = 0, all in one batch, next to an ordinary system."""
    sysl = [tiny_system(1, 1, 1), tiny_system(2, 8, 2), tiny_system(3, 3, 3)] + systems(1, seed=5)
    zero = list(tiny_system(40, 2, 4))
    zero[4] = np.zeros_like(zero[4])
    nanb = list(tiny_system(40, 2, 5))
    nanb[4] = nanb[4].copy()
    nanb[4][7, 1] = np.nan
    capped = tiny_system(40, 2, 6)
    bt = BatchCG()
    devs = [to_dev(s, cuda) for s in sysl + [tuple(zero), tuple(nanb), capped]]
    ids = []
    for j, (rp, ci, val, d, B, X) in enumerate(devs):
        ids.append(bt.add(rp, ci, val, B, X, 1e-10, 0 if j == len(devs) - 1 else 20000, diag=d, mixed=mixed == 1,
                          true_residual=mixed == 2))
    bt.run(8)
    res = [bt.result(sid) for sid in ids]
    Xs = [dv[5].cpu().numpy() for dv in devs]
    bt.close()
    for j, (rp, ci, val, d, B, n, k) in enumerate(sysl):
        assert res[j]["succeeded"], (j, res[j])
        A = sp.csr_matrix((val, ci, rp), shape=(n, n))
        assert np.max(np.linalg.norm(B - A @ Xs[j], axis=0) / np.linalg.norm(B, axis=0)) < 1.5e-10
    z, nb, cap = len(sysl), len(sysl) + 1, len(sysl) + 2
    assert res[z]["succeeded"] and res[z]["iterations"] == 0 and not np.any(Xs[z])
    assert not res[nb]["succeeded"] and res[nb]["iterations"] == 0
    assert not res[cap]["succeeded"] and res[cap]["iterations"] == 0
