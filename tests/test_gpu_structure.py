"""GPU structure work against the oracle: triples -> CSR (FromTriples + SortInner +
Compress, symmetrize), GridLaplacian assembly, transpose and sparse x sparse.  Row
pointers and column indices must match bit for bit; values exactly where the sums are
exact (small integers), else within 1e-14 relative."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import rel_err
from tensors_amd import structure as st

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rand_triples(count, m, n, seed, integer=True):
    rng = np.random.default_rng(seed)
    I = rng.integers(0, m, count).astype(np.int32)
    J = rng.integers(0, n, count).astype(np.int32)
    V = rng.integers(-5, 6, count).astype(np.float64) if integer else rng.normal(size=count)
    return I, J, V


@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("symmetrize", [False, True])
def test_from_triples_matches_oracle(cuda, compress, symmetrize):
    m = n = 300
    I, J, V = rand_triples(5000, m, n, seed=1)
    grp, gci, gval = st.from_triples(I, J, V, m, n, compress, symmetrize)
    orp, oci, oval = oracle.from_triples(I, J, V, m, n, lists=3, compress=compress, symmetrize=symmetrize)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci)
    if compress:
        assert np.array_equal(gval, oval)  # integer sums are exact in any order
    else:  # duplicates kept: compare the multiset of values per (row, col)
        key = lambda rp, ci, v: sorted(zip(np.repeat(np.arange(m), np.diff(rp)), ci, v))
        assert key(grp, gci, gval) == key(orp, oci, oval)
    if symmetrize:  # one coo of all triples: sparse "+" would drop sums that cancel to 0
        off = I != J
        I, J, V = np.concatenate([I, J[off]]), np.concatenate([J, I[off]]), np.concatenate([V, V[off]])
    A = sp.coo_matrix((V, (I, J)), shape=(m, n)).tocsr()
    A.sum_duplicates()
    if compress:
        assert np.array_equal(grp, A.indptr) and np.array_equal(gci, A.indices)


def test_from_triples_float_values_and_empty(cuda):
    m, n = 500, 400
    I, J, V = rand_triples(20000, m, n, seed=2, integer=False)
    grp, gci, gval = st.from_triples(I, J, V, m, n)
    A = sp.coo_matrix((V, (I, J)), shape=(m, n)).tocsr()
    A.sum_duplicates()
    assert np.array_equal(grp, A.indptr) and np.array_equal(gci, A.indices)
    assert np.max(np.abs(gval - A.data)) <= 1e-14 * np.max(np.abs(A.data))
    rp0, ci0, v0 = st.from_triples(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), 7, 7)
    assert np.all(rp0 == 0) and len(ci0) == 0


def test_from_triples_rejects_out_of_range(cuda):
    from tensors_amd import TensorsError
    with pytest.raises(TensorsError):
        st.from_triples(np.array([0, 5], np.int32), np.array([0, 1], np.int32), np.ones(2), 5, 5)


def grid_triples(g, mass):
    """The push order of GridLaplacian.hpp:17-70."""
    I, J, V = [], [], []
    def push(i, j, v):
        I.append(i); J.append(j); V.append(v)
    for i in range(g - 1):
        for j in range(g - 1):
            v = g * i + j; l = v + g; r = v + 1
            for a, b, c in ((v, v, 1), (v, r, -1), (r, v, -1), (r, r, 1), (v, v, 1), (v, l, -1), (l, v, -1), (l, l, 1)):
                push(a, b, c)
    for i in range(g - 1):
        v = g * i + g - 1; l = v + g
        for a, b, c in ((v, v, 1), (v, l, -1), (l, v, -1), (l, l, 1)):
            push(a, b, c)
    for j in range(g - 1):
        v = g * (g - 1) + j; r = v + 1
        for a, b, c in ((v, v, 1), (v, r, -1), (r, v, -1), (r, r, 1)):
            push(a, b, c)
    for v in range(g * g):
        push(v, v, mass)
    return np.array(I, np.int32), np.array(J, np.int32), np.array(V, np.float64)


@pytest.mark.parametrize("g", [4, 16, 256])
def test_grid_laplacian_assembly_bit_exact(cuda, g):
    I, J, V = grid_triples(g, 1.0)
    grp, gci, gval = st.from_triples(I, J, V, g * g, g * g)
    orp, oci, oval = oracle.grid_laplacian(g, 1.0)
    assert np.array_equal(grp, orp) and np.array_equal(gci, oci) and np.array_equal(gval, oval)
    assert len(gci) == 5 * g * g - 4 * g


def test_transpose_matches_oracle(cuda):
    m, n = 230, 170
    I, J, V = rand_triples(3000, m, n, seed=3, integer=False)
    rp, ci, val = st.from_triples(I, J, V, m, n)
    trp, tci, tval = st.transpose(m, n, rp, ci, val)
    orp, oci, oval = oracle.transpose(m, n, rp, ci, val)
    assert np.array_equal(trp, orp) and np.array_equal(tci, oci) and np.array_equal(tval, oval)
    brp, bci, bval = st.transpose(n, m, trp, tci, tval)  # (A^T)^T = A
    assert np.array_equal(brp, rp) and np.array_equal(bci, ci) and np.array_equal(bval, val)


@pytest.mark.parametrize("values", [True, False])
def test_spgemm_matches_oracle(cuda, values):
    m, k, n = 120, 90, 110
    A = st.from_triples(*rand_triples(900, m, k, seed=4, integer=False), m, k)
    B = st.from_triples(*rand_triples(800, k, n, seed=5, integer=False), k, n)
    crp, cci, cval = st.spgemm(m, k, n, A, B, values=values)
    orp, oci, oval = oracle.spgemm(m, k, n, A, B, values=values)
    assert np.array_equal(crp, orp) and np.array_equal(cci, oci)
    Ad = sp.csr_matrix((A[2] if values else np.ones(len(A[1])), A[1], A[0]), shape=(m, k))
    Bd = sp.csr_matrix((B[2] if values else np.ones(len(B[1])), B[1], B[0]), shape=(k, n))
    Cd = (Ad @ Bd).tocsr()
    Cd.sort_indices()
    if values:
        assert np.max(np.abs(cval - oval)) <= 1e-14 * np.max(np.abs(oval))
        dense = Cd.toarray()
        got = sp.csr_matrix((cval, cci, crp), shape=(m, n)).toarray()
        assert np.max(np.abs(got - dense)) <= 1e-13 * np.max(np.abs(dense))
    else:  # pattern product == boolean product
        pattern = (Ad.astype(bool) @ Bd.astype(bool)).tocsr()
        pattern.sort_indices()
        assert np.array_equal(crp, pattern.indptr) and np.array_equal(cci, pattern.indices)


def test_cpp_dropin_program(cuda, tmp_path):
    """The header-only drop-in (include/Tensors) used as reference code would use it."""
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_test")
    assert os.path.exists(exe), "build tests/cpp/dropin_test first (make -C tests/cpp)"
    dump = tmp_path / "grid16.bin"
    sdir = tmp_path / "solvers"
    sdir.mkdir()
    r = subprocess.run([exe, str(dump), str(sdir)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr[-2000:])
    assert r.returncode == 0 and "ALL OK" in r.stdout
    raw = dump.read_bytes()
    m, nnz = np.frombuffer(raw[:16], np.int64)
    off = 16
    rp = np.frombuffer(raw[off:off + 8 * (m + 1)], np.int64); off += 8 * (m + 1)
    ci = np.frombuffer(raw[off:off + 4 * nnz], np.int32); off += 4 * nnz
    val = np.frombuffer(raw[off:off + 8 * nnz], np.float64)
    orp, oci, oval = oracle.grid_laplacian(16, 1.0)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(val, oval)
    # the drop-in's solver accessors after fixed-iteration runs (CG 20 steps; GMRES(10),
    # two cycles) against the oracle's values on the same right-hand sides, at 1e-12
    n, k, m = 256, 2, 10
    ld = lambda name, *shape: np.fromfile(sdir / name, np.float64).reshape(shape)
    B = ld("B.bin", n, k)
    Xo = np.zeros((n, k))
    _, _, rr_o, ar_o = oracle.cg_ex(n, orp, oci, oval, B, Xo, k, 20, 0.0)
    assert rel_err(ld("cg_X.bin", n, k), Xo) <= 1e-12
    # residuals in units of the initial norm (sqrt(b.b) for CG, ||b|| for GMRES)
    assert np.max(np.abs(ld("cg_res.bin", k) - ar_o) / (ar_o / rr_o)) <= 1e-12
    assert np.max(np.abs(ld("cg_rel.bin", k) - rr_o)) <= 1e-12
    Xg = np.zeros((n, k))
    _, it_o, _, grr_o, Q_o, H_o, beta_o, sol_o = oracle.gmres_ex(n, orp, oci, oval, B, Xg, k, m, 0.0, 1)
    assert rel_err(ld("gm_X.bin", n, k), Xg) <= 1e-12
    assert np.max(np.abs(ld("gm_res.bin", k) - np.abs(beta_o[it_o])) / (np.abs(beta_o[it_o]) / grr_o)) <= 1e-12
    assert np.max(np.abs(ld("gm_rel.bin", k) - grr_o)) <= 1e-12
    assert rel_err(ld("gm_sol.bin", n, k), sol_o) <= 1e-12
    # the second cycle's basis and Hessenberg matrix are built from the restart residual,
    # which has fallen to rr1 of ||b||: rounding differences of the first cycle's x
    # (~1e-15 ||b||) reach them divided by rr1, so their bar is 1e-12 / rr1
    _, _, _, rr1 = oracle.gmres(n, orp, oci, oval, B, np.zeros((n, k)), k, m, 0.0, 0)
    bar = 1e-12 / min(1.0, float(rr1.min()))
    assert rel_err(ld("gm_Q.bin", m + 1, n, k), Q_o) <= bar
    assert rel_err(ld("gm_H.bin", m + 1, m, k), H_o) <= bar


@pytest.mark.parametrize("symmetric", [False, True])
def test_matrix_market_load_matches_oracle(cuda, tmp_path, symmetric):
    """MatrixCSR.from_matrix_market (host parse, GPU assembly) against the oracle's
    FromTriples on the same entries read independently by numpy: the reference's
    LoadFromMatrixMarket (MatrixMarket.hpp:190-196) builds the CSR from the file's triples,
    summing duplicates and, for "symmetric" files, mirroring the off-diagonal entries.
    Then write -> load reproduces the matrix bit for bit and Dot works on it."""
    import torch

    from tensors_amd.sparse import MatrixCSR

    m = n = 2000
    rng = np.random.default_rng(12)
    k = 30000
    I = rng.integers(1, m + 1, k)
    J = rng.integers(1, n + 1, k)
    if symmetric:  # lower triangle and diagonal, as symmetric files store them
        I, J = np.maximum(I, J), np.minimum(I, J)
    V = rng.normal(size=k)
    f = tmp_path / "a.mtx"
    with open(f, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real " + ("symmetric" if symmetric else "general") + "\n")
        fh.write("% generated\n" + f"{m} {n} {k}\n")
        np.savetxt(fh, np.column_stack([I, J, V]), fmt=["%d", "%d", "%.17g"])
    A = MatrixCSR.from_matrix_market(f)
    T = np.loadtxt(f, comments="%", skiprows=3)  # banner, comment, size line
    orp, oci, oval = oracle.from_triples(T[:, 0].astype(np.int32) - 1, T[:, 1].astype(np.int32) - 1, T[:, 2], m, n,
                                         compress=True, symmetrize=symmetric)
    assert np.array_equal(A.Outer(), orp) and np.array_equal(A.Inner(), oci)
    assert np.array_equal(A.Values(), oval)  # duplicates summed in file order, as the oracle
    g = tmp_path / "b.mtx"
    A.write_matrix_market(g)
    B = MatrixCSR.from_matrix_market(g)
    assert np.array_equal(B.Outer(), orp) and np.array_equal(B.Inner(), oci) and np.array_equal(B.Values(), oval)
    X = torch.rand(n, 3, dtype=torch.float64, device=cuda)
    Y = torch.zeros(m, 3, dtype=torch.float64, device=cuda)
    assert B.Dot(1.0, X, 0.0, Y) == 0
    Yo = sp.csr_matrix((oval, oci, orp), shape=(m, n)) @ X.cpu().numpy()
    assert np.abs(Y.cpu().numpy() - Yo).max() <= 1e-12 * np.abs(Yo).max()
