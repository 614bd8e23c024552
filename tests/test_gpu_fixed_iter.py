"""GPU Krylov solvers against the oracle at a fixed iteration count, at 1e-12.

With tol = 0 every solver runs exactly the requested steps (no stop test can fire), so the
GPU iterate and the oracle iterate after k steps are the same function of the inputs and
differ only by rounding.  The oracle's iterates are pinned to the Krylov minimisers that
these steps compute (tests/test_oracle_krylov.py), and two oracle runs with different
reduction splits agree to ~1e-14, so 1e-12 is a rounding-level bar: a different beta
formula, a Givens sign slip or a dropped Gram-Schmidt term fails it by orders of
magnitude.  Compared: the iterate X (max |X - X_o| / max |X_o|), the relative residuals and
Residuals() (sqrt|rho| for CG, |beta_iter| for GMRES) in units of the initial
(preconditioned) norm, and, for GMRES, the basis, the rotated Hessenberg matrix and the
solution of the last cycle (GetBasis / GetHessenbergMatrix / Solution, GMRES.hpp:820-833).
References: ConjugateGradient.hpp:250-339, 447-456; GMRES.hpp:377-414, 526-658, 779-833.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from helpers import rel_err
from tensors_amd import _capi as capi
from tensors_amd import generators as gen
from tensors_amd.batch import BatchCG
from tensors_amd.solvers import GMRES, ConjugateGradient
from tensors_amd.sparse import MatrixCSR

pytestmark = pytest.mark.gpu

TOL = 1e-12


def dev(a, d, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(d).to(dtype)


def res_err(a, b, scale=1.0):
    """Difference of residual estimates in units of the initial (preconditioned) norm:
    relative residuals are already in those units; absolute ones are divided by `scale`
    (sqrt(b.Pb) for CG, ||P b|| for GMRES).  A residual that has fallen to 1e-8 of the
    initial one is known only to rounding in those units, not to 1e-12 of itself."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / scale))


@pytest.fixture
def fused_knob():
    before = int(capi.lib().tg_tune_get(b"cg_fused"))
    yield lambda v: capi.call("tg_tune_set", b"cg_fused", int(v))
    capi.call("tg_tune_set", b"cg_fused", before)


@pytest.mark.parametrize("form", [3, 0, 2], ids=["default", "four_kernels", "fused_direction"])
@pytest.mark.parametrize("steps", [1, 5, 20])
@pytest.mark.parametrize("prec", [False, True], ids=["identity", "jacobi"])
@pytest.mark.parametrize("k", [1, 8])
def test_cg_fixed_iterations_match_oracle(cuda, fused_knob, form, steps, prec, k):
    fused_knob(form)
    rp, ci, val = oracle.grid_laplacian(24, 0.1)
    n = 576
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    B = gen.uniform(4100 + steps + k, n * k).reshape(n, k)
    Xo = np.zeros((n, k))
    ok_o, it_o, rr_o, ar_o = oracle.cg_ex(n, rp, ci, val, B, Xo, k, steps, 0.0, prec_diag=d)
    cg = ConjugateGradient(n, steps, k)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok = cg(MatrixCSR(rp, ci, val, n, n), None if d is None else dev(d, cuda), 1.0, dev(B, cuda), 0.0, Xd, 0.0)
    print(f"CG form {form} steps {steps} prec {prec} k {k}")
    assert not ok and ok_o == 0 and cg.IterationCount() == it_o == steps
    assert rel_err(Xd.cpu().numpy(), Xo) <= TOL
    assert res_err(cg.RelativeResiduals(), rr_o) <= TOL
    assert res_err(cg.Residuals(), ar_o, ar_o / rr_o) <= TOL


@pytest.mark.parametrize("steps", [3, 30])
def test_cg_fixed_iterations_3d_wide(cuda, steps):
    """32 right-hand sides on the 3D 7-point Laplacian (16-lane groups, 2 columns per lane)."""
    g, k = 16, 32
    rp, ci, val = gen.laplacian_3d(g, 1e-2)
    n = g ** 3
    B = gen.uniform(3003, n * k).reshape(n, k)
    Xo = np.zeros((n, k))
    _, it_o, rr_o, ar_o = oracle.cg_ex(n, rp, ci, val, B, Xo, k, steps, 0.0, threads=8)
    cg = ConjugateGradient(n, steps, k)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    cg(MatrixCSR(rp.astype(np.int32), ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, Xd, 0.0)
    assert cg.IterationCount() == it_o == steps
    assert rel_err(Xd.cpu().numpy(), Xo) <= TOL
    assert res_err(cg.RelativeResiduals(), rr_o) <= TOL
    assert res_err(cg.Residuals(), ar_o, ar_o / rr_o) <= TOL


@pytest.fixture
def panel_knob():
    before = int(capi.lib().tg_tune_get(b"cg_panel"))
    yield lambda v: capi.call("tg_tune_set", b"cg_panel", int(v))
    capi.call("tg_tune_set", b"cg_panel", before)


@pytest.mark.parametrize("panel", [1, 0], ids=["panel_order", "contiguous"])
@pytest.mark.parametrize("prec", [False, True], ids=["identity", "jacobi"])
def test_cg_fixed_iterations_panel_order(cuda, panel_knob, panel, prec):
    """The SpMM of the default iteration with its rows in panel order: the 7-point
    Laplacian (+ 0.1) of a 128 x 64 x 8 grid, x fastest, has band 8192, so the rows are
    cut into 8 panels of one plane and every workgroup takes 4 offsets of each (knob
    cg_panel = 1, the default); the contiguous chunks (cg_panel = 0) for comparison.  Both
    at 1e-12 from the oracle after 20 steps (a well-conditioned operator, so that 20 steps
    do not amplify rounding past that bar)."""
    panel_knob(panel)

    def path(m):
        return sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])

    nx, ny, nz = 128, 64, 8
    A = (sp.kron(sp.identity(nz * ny), path(nx)) + sp.kron(sp.kron(sp.identity(nz), path(ny)), sp.identity(nx))
         + sp.kron(path(nz), sp.identity(ny * nx)) + 0.1 * sp.identity(nx * ny * nz)).tocsr()
    A.sort_indices()
    rp, ci, val = A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.astype(np.float64)
    assert int(np.max(np.abs(ci - np.repeat(np.arange(A.shape[0]), np.diff(rp))))) == nx * ny
    n, k, steps = A.shape[0], 8, 20
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    B = gen.uniform(809, n * k).reshape(n, k)
    Xo = np.zeros((n, k))
    _, it_o, rr_o, ar_o = oracle.cg_ex(n, rp, ci, val, B, Xo, k, steps, 0.0, prec_diag=d, threads=8)
    cg = ConjugateGradient(n, steps, k)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    cg(MatrixCSR(rp.astype(np.int32), ci, val, n, n), None if d is None else dev(d, cuda), 1.0, dev(B, cuda), 0.0,
       Xd, 0.0)
    print(f"CG panel {panel} prec {prec}")
    assert cg.IterationCount() == it_o == steps
    assert rel_err(Xd.cpu().numpy(), Xo) <= TOL
    assert res_err(cg.RelativeResiduals(), rr_o) <= TOL
    assert res_err(cg.Residuals(), ar_o, ar_o / rr_o) <= TOL


@pytest.fixture(params=[1, 0], ids=["delayed", "cgs2"])
def gmres_delayed(request):
    """Both Gram-Schmidt forms of the GMRES step (knob gmres_delayed)."""
    from tensors_amd import _capi as capi

    before = int(capi.lib().tg_tune_get(b"gmres_delayed"))
    capi.call("tg_tune_set", b"gmres_delayed", request.param)
    yield request.param
    capi.call("tg_tune_set", b"gmres_delayed", before)


@pytest.mark.parametrize("m", [5, 12])
@pytest.mark.parametrize("cycles", [1, 2])
@pytest.mark.parametrize("side", [0, 1], ids=["left", "right"])
@pytest.mark.parametrize("prec", [False, True], ids=["identity", "jacobi"])
@pytest.mark.parametrize("k", [1, 3])
def test_gmres_fixed_cycles_match_oracle(cuda, gmres_delayed, m, cycles, side, prec, k):
    rp, ci, val = gen.convection_diffusion_2d(24, 10.0)
    n = len(rp) - 1
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal() if prec else None
    B = gen.uniform(5100 + m + k, n * k).reshape(n, k)
    Xo = np.zeros((n, k))
    ok_o, it_o, rs_o, rr_o, Q_o, H_o, beta_o, sol_o = oracle.gmres_ex(
        n, rp, ci, val, B, Xo, k, m, 0.0, cycles - 1, prec_diag=d, side=side)
    gm = GMRES(n, m, k, side=side)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    ok = gm(MatrixCSR(rp.astype(np.int32), ci, val, n, n), None if d is None else dev(d, cuda), 1.0,
            dev(B, cuda), 0.0, Xd, 0.0, cycles - 1)
    print(f"GMRES delayed {gmres_delayed} m {m} cycles {cycles} side {side} prec {prec} k {k}")
    assert not ok and ok_o == 0
    assert gm.IterationCount() == it_o == m and gm.RestartCount() == rs_o == cycles - 1
    assert rel_err(Xd.cpu().numpy(), Xo) <= TOL
    assert res_err(gm.RelativeResiduals(), rr_o) <= TOL
    assert res_err(gm.Residuals(), np.abs(beta_o[it_o]), np.abs(beta_o[it_o]) / rr_o) <= TOL
    assert rel_err(gm.Solution().cpu().numpy(), sol_o) <= TOL
    assert rel_err(gm.GetBasis().cpu().numpy(), Q_o) <= TOL
    assert rel_err(gm.GetHessenbergMatrix().cpu().numpy(), H_o) <= TOL


@pytest.mark.parametrize("k", [2, 8, 16])
def test_gmres_fixed_cycles_wide_blocks(cuda, gmres_delayed, k):
    """Blocks of 2, 8 and 16 right-hand sides (two, eight and sixteen column lanes in the
    reductions over a wave) against the oracle after two cycles of 12 steps, left Jacobi."""
    rp, ci, val = gen.convection_diffusion_2d(24, 10.0)
    n, m, cycles = len(rp) - 1, 12, 2
    d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
    B = gen.uniform(5400 + k, n * k).reshape(n, k)
    Xo = np.zeros((n, k))
    ok_o, it_o, rs_o, rr_o, Q_o, H_o, beta_o, sol_o = oracle.gmres_ex(
        n, rp, ci, val, B, Xo, k, m, 0.0, cycles - 1, prec_diag=d, side=0)
    gm = GMRES(n, m, k, side=0)
    Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
    gm(MatrixCSR(rp.astype(np.int32), ci, val, n, n), dev(d, cuda), 1.0, dev(B, cuda), 0.0, Xd, 0.0, cycles - 1)
    print(f"GMRES delayed {gmres_delayed} k {k}")
    assert gm.IterationCount() == it_o == m and gm.RestartCount() == rs_o == cycles - 1
    assert rel_err(Xd.cpu().numpy(), Xo) <= TOL
    assert res_err(gm.RelativeResiduals(), rr_o) <= TOL
    assert rel_err(gm.GetBasis().cpu().numpy(), Q_o) <= TOL
    assert rel_err(gm.GetHessenbergMatrix().cpu().numpy(), H_o) <= TOL


@pytest.mark.parametrize("side", [0, 1], ids=["left", "right"])
def test_gmres_long_restart_falls_back_to_cgs2(cuda, side):
    """Restart lengths above 32 take the CGS2 step whatever knob gmres_delayed says: the
    two settings give bitwise the same solution, basis and Hessenberg matrix, and the
    solution after two cycles of 40 steps matches the oracle at 1e-12."""
    rp, ci, val = gen.convection_diffusion_2d(24, 10.0)
    n, m, k = len(rp) - 1, 40, 2
    B = gen.uniform(5300 + side, n * k).reshape(n, k)
    Xo = np.zeros((n, k))
    oracle.gmres_ex(n, rp, ci, val, B, Xo, k, m, 0.0, 1, side=side)
    before = int(capi.lib().tg_tune_get(b"gmres_delayed"))
    out = {}
    try:
        for mode in (1, 0):
            capi.call("tg_tune_set", b"gmres_delayed", mode)
            gm = GMRES(n, m, k, side=side)
            Xd = torch.zeros(n, k, dtype=torch.float64, device=cuda)
            gm(MatrixCSR(rp.astype(np.int32), ci, val, n, n), None, 1.0, dev(B, cuda), 0.0, Xd, 0.0, 1)
            out[mode] = (Xd.cpu(), gm.GetBasis().cpu(), gm.GetHessenbergMatrix().cpu(), gm.IterationCount())
    finally:
        capi.call("tg_tune_set", b"gmres_delayed", before)
    assert out[1][3] == out[0][3] == m
    for a, b in zip(out[1][:3], out[0][:3]):
        assert torch.equal(a, b)
    assert rel_err(out[1][0].numpy(), Xo) <= TOL


@pytest.mark.parametrize("steps", [1, 7, 25])
def test_batch_cg_fixed_iterations_match_oracle(cuda, steps):
    """The batched engine's FP64 mode (the reference recurrences, Jacobi) after a fixed
    number of iterations: every system of an uneven batch at 1e-12 from the oracle."""
    rng = np.random.default_rng(77)
    bt = BatchCG()
    cases = []
    for s in range(6):
        dims = 2 + (s % 2)
        side = int(rng.integers(6, 40 if dims == 2 else 12))
        rp, ci, val = gen.weighted_grid(dims, (side,) * dims, 1e-3, seed=900 + s)
        n = len(rp) - 1
        k = int(rng.integers(1, 9))
        B = gen.uniform(901 + s, n * k).reshape(n, k)
        d = 1.0 / sp.csr_matrix((val, ci, rp), shape=(n, n)).diagonal()
        T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(cuda).to(dt)
        devs = (T(rp, torch.int32), T(ci, torch.int32), T(val, torch.float64), T(d, torch.float64), T(B, torch.float64),
                torch.zeros(n, k, dtype=torch.float64, device=cuda))
        sid = bt.add(devs[0], devs[1], devs[2], devs[4], devs[5], 0.0, steps, diag=devs[3])
        cases.append((sid, rp, ci, val, d, B, n, k, devs))
    bt.run(4)
    for sid, rp, ci, val, d, B, n, k, devs in cases:
        r = bt.result(sid)
        Xo = np.zeros((n, k))
        _, it_o, rr_o = oracle.cg(n, rp, ci, val, B, Xo, k, steps, 0.0, prec_diag=d)
        assert r["iterations"] == it_o == steps and not r["succeeded"]
        assert rel_err(devs[5].cpu().numpy(), Xo) <= TOL
        assert res_err(r["rel_res"], rr_o) <= TOL
    bt.close()
