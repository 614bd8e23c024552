"""Shared test helpers: small matrices, the reference's error metric and bounds."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def rel_err(Y, Yref):
    """max|Y - Y_ref| / max|Y_ref|: the reference's metric (Unit_Test_SpMM/main.cpp:116-132)."""
    Y = np.asarray(Y, np.float64)
    Yref = np.asarray(Yref, np.float64)
    mx = np.max(np.abs(Yref)) if Yref.size else 0.0
    err = np.max(np.abs(Y - Yref)) if Y.size else 0.0
    return err / mx if mx > 0 else err


def random_csr_np(m, n, density, seed, empty_rows=0.0, dtype=np.float64, sorted_cols=True):
    """Small random CSR with optional empty rows, via scipy (test input only)."""
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, format="csr", random_state=rng, dtype=np.float64)
    if empty_rows > 0:
        mask = rng.random(m) < empty_rows
        A = sp.csr_matrix(sp.diags((~mask).astype(np.float64)) @ A)
        A.eliminate_zeros()
    A.sort_indices()
    rp = A.indptr.astype(np.int64)
    ci = A.indices.astype(np.int32)
    val = A.data.astype(dtype)
    return rp, ci, val


def dense_of(m, n, rp, ci, val):
    D = np.zeros((m, n))
    for i in range(m):
        for l in range(rp[i], rp[i + 1]):
            D[i, ci[l]] += 1.0 if val is None else val[l]
    return D


def gamma(k, u):
    return k * u / (1 - k * u)


def spmm_bound(m, n, rp, ci, val, alpha, X, beta, Y0, u):
    """Per-entry rounding bound for any summation order: gamma_{L+2} (|alpha| |A| |X| + |beta| |Y0|)."""
    A = sp.csr_matrix((np.abs(val) if val is not None else np.ones(len(ci)), ci, rp), shape=(m, n))
    lens = np.diff(rp)
    g = np.array([gamma(L + 3, u) for L in lens])[:, None]
    return g * (abs(alpha) * (A @ np.abs(X.reshape(n, -1))) + abs(beta) * np.abs(Y0.reshape(m, -1))) + 1e-300
