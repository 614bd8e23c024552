"""CPU checks of the C ABI: the library loads, exports every function the headers in
include/ declare, and the host-side generators are deterministic and well-formed.
No GPU compute is called here."""
import glob
import os
import re

import numpy as np
import scipy.sparse as sp

from tensors_amd import _capi as capi
from tensors_amd import generators as gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(tg_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.SIGNATURES) <= names


def test_version_and_error_slot():
    assert b"gfx950" in capi.lib().tg_version()
    assert isinstance(capi.last_error(), str)


def test_uniform_is_deterministic_and_offset_consistent():
    a = gen.uniform(5, 1000)
    b = gen.uniform(5, 1000)
    c = gen.uniform(5, 500, offset=500)
    assert np.array_equal(a, b) and np.array_equal(a[500:], c)
    assert a.min() >= -1 and a.max() <= 1 and abs(a.mean()) < 0.1


def test_random_csr_properties():
    m = n = 20000
    rp, ci, v64, v32 = gen.random_csr(m, n, 32.0, 1, 512, seed=2002, want32=True)
    lens = np.diff(rp)
    assert lens.min() >= 1 and lens.max() <= 512
    assert abs(lens.mean() - 32.0) < 0.3 and abs(lens.var() - 32.0) < 2.0  # Poisson(32)
    assert ci.min() >= 0 and ci.max() < n
    for i in range(0, m, 997):  # strictly increasing columns per row
        row = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)
    assert np.array_equal(v32, v64.astype(np.float32))
    assert abs(v64.mean()) < 0.01 and abs(v64.std() - 1.0) < 0.01
    rp2, ci2, v2, _ = gen.random_csr(m, n, 32.0, 1, 512, seed=2002)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(v64, v2)


def test_laplacian_3d_matches_kronecker_sum():
    g, shift = 5, 1e-2
    rp, ci, val = gen.laplacian_3d(g, shift)
    A = sp.csr_matrix((val, ci, rp), shape=(g ** 3, g ** 3)).toarray()
    T = sp.diags([-1, 0, -1], [-1, 0, 1], shape=(g, g)).toarray()  # path-graph adjacency (negated)
    I = np.eye(g)
    adj = -(np.kron(np.kron(T, I), I) + np.kron(np.kron(I, T), I) + np.kron(np.kron(I, I), T))
    L = np.diag(adj.sum(1)) - adj + shift * np.eye(g ** 3)
    assert np.array_equal(A, L)
    assert len(ci) == 7 * g ** 3 - 6 * g * g


def test_weighted_grid_is_spd_and_symmetric():
    for dims, sides in ((2, (7, 9)), (3, (4, 5, 3))):
        rp, ci, val = gen.weighted_grid(dims, sides, 1e-3, seed=4004)
        N = len(rp) - 1
        A = sp.csr_matrix((val, ci, rp), shape=(N, N)).toarray()
        assert np.allclose(A, A.T, rtol=0, atol=0)
        assert np.allclose(A.sum(1), 1e-3)  # Laplacian + shift
        assert np.linalg.eigvalsh(A).min() > 0
        for i in range(N):
            assert np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0)


def test_convection_diffusion_is_m_matrix():
    rp, ci, val = gen.convection_diffusion_2d(6, 10.0)
    N = 36
    A = sp.csr_matrix((val, ci, rp), shape=(N, N)).toarray()
    off = A - np.diag(np.diag(A))
    assert np.all(off <= 0) and np.all(A.sum(1) >= -1e-12)
    assert not np.allclose(A, A.T)
    assert np.all(np.linalg.inv(A) >= -1e-12)  # inverse-positive


def tuning_keys():
    """The knob keys the header documents for tg_tune_set (lines ` *   "key":` or
    ` *   "key1", "key2":` before the tg_tune_set declaration)."""
    text = open(os.path.join(ROOT, "include", "tensors_amd.h")).read()
    block = text[:text.index("int tg_tune_set(")]
    keys = []
    for m in re.finditer(r'^ \*   ((?:"[a-z0-9_]+"(?:, )?)+):', block, flags=re.M):
        keys += re.findall(r'"([a-z0-9_]+)"', m.group(1))
    return keys


def test_every_documented_knob_reads_back():
    """tg_tune_get knows every key tg_tune_set documents (a key missing from its dispatch
    reads -1, and a caller that saves and restores a knob then fails), and setting a knob
    to the value it reads is accepted.  Host globals only: no GPU call."""
    keys = tuning_keys()
    assert len(keys) >= 20 and "cg_xdefer" in keys and "gmres_delayed" in keys
    L = capi.lib()
    for key in keys:
        v = int(L.tg_tune_get(key.encode()))
        assert v >= 0, key
        assert L.tg_tune_set(key.encode(), v) == 0, (key, capi.last_error())
        assert int(L.tg_tune_get(key.encode())) == v, key
    assert int(L.tg_tune_get(b"no_such_knob")) == -1
