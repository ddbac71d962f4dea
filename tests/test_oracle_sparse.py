"""Pins the sparse oracle (oracle/sparse_oracle.c): against the reference's own
csr_matrix.cpp compiled as-is (oracle/_ref), scipy, dense linear algebra and
analytic ray lengths.  CPU only."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import sparse as osp

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libref.so"


def _ref():
    if not REF.exists():
        pytest.skip("reference build (oracle/_ref) not present")
    lib = C.CDLL(str(REF))
    lib.ref_csr_build.restype = C.c_void_p
    lib.ref_csr_nnz.restype = C.c_int64
    lib.ref_csr_frobenius.restype = C.c_double
    for f in ("ref_csr_nnz", "ref_csr_arrays", "ref_csr_spmv", "ref_csr_add_diagonal", "ref_csr_scale",
              "ref_csr_frobenius", "ref_csr_free"):
        getattr(lib, f).argtypes = None
    lib.ref_csr_nnz.argtypes = [C.c_void_p]
    lib.ref_csr_frobenius.argtypes = [C.c_void_p]
    lib.ref_csr_free.argtypes = [C.c_void_p]
    lib.ref_csr_add_diagonal.argtypes = [C.c_void_p, C.c_double]
    lib.ref_csr_scale.argtypes = [C.c_void_p, C.c_double]
    lib.ref_csr_arrays.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.ref_csr_spmv.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
    lib.ref_csr_build.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _d(a):
    return a.ctypes.data_as(C.c_void_p)


def ref_build(lib, rows, cols, r, c, v):
    h = lib.ref_csr_build(rows, cols, r.size, _d(r), _d(c), _d(v))
    assert h, "reference refuses inputs with rows of < 8 entries (its padding pass overflows)"
    nnz = lib.ref_csr_nnz(h)
    rp = np.zeros(rows + 1, np.int64)
    ci = np.zeros(nnz, np.int64)
    val = np.zeros(nnz)
    lib.ref_csr_arrays(h, _d(rp), _d(ci), _d(val))
    return h, rp, ci, val


def random_triplets(rng, rows, cols, per_row=12):
    r, c, v = [], [], []
    for i in range(rows):
        cs = rng.choice(cols, per_row, replace=False)
        r += [i] * per_row
        c += list(cs)
        v += list(rng.standard_normal(per_row))
        # pairwise duplicates (two terms: order-independent sum)
        for j in cs[:3]:
            r.append(i); c.append(j); v.append(rng.standard_normal())
    # out-of-range entries are dropped
    r += [-1, rows, 0, 1]
    c += [0, 0, -3, cols + 2]
    v += [9.0, 9.0, 9.0, 9.0]
    p = rng.permutation(len(r))
    return np.asarray(r, np.int64)[p], np.asarray(c, np.int64)[p], np.asarray(v)[p]


def test_assembly_matches_reference_bit_exact():
    lib = _ref()
    rng = np.random.default_rng(0)
    rows, cols = 60, 90
    r, c, v = random_triplets(rng, rows, cols)
    h, rp, ci, val = ref_build(lib, rows, cols, r, c, v)
    try:
        orp, oci, oval = osp.csr_from_triplets(rows, cols, r, c, v)
        np.testing.assert_array_equal(orp, rp)
        np.testing.assert_array_equal(oci, ci)
        np.testing.assert_array_equal(oval, val)
        x = rng.standard_normal(cols)
        y = np.zeros(rows)
        lib.ref_csr_spmv(h, _d(x), cols, _d(y))
        np.testing.assert_array_equal(osp.spmv(orp, oci, oval, x), y)      # same even/odd accumulation order
        assert lib.ref_csr_frobenius(h) == pytest.approx(np.sqrt((oval ** 2).sum()), rel=1e-15)
        lib.ref_csr_add_diagonal(h, 2.5)
        lib.ref_csr_arrays(h, _d(rp), _d(ci), _d(val))
        want = oval.copy()
        for i in range(rows):
            for k in range(orp[i], orp[i + 1]):
                if oci[k] == i:
                    want[k] += 2.5
                    break
        np.testing.assert_array_equal(val, want)
    finally:
        lib.ref_csr_free(h)


def test_assembly_matches_scipy_sum_duplicates():
    rng = np.random.default_rng(1)
    rows, cols = 40, 30
    r = rng.integers(0, rows, 500)
    c = rng.integers(0, cols, 500)
    v = rng.standard_normal(500)
    rp, ci, val = osp.csr_from_triplets(rows, cols, r, c, v)
    m = sp.coo_matrix((v, (r, c)), shape=(rows, cols)).tocsr()
    m.sum_duplicates()
    m.sort_indices()
    np.testing.assert_array_equal(rp, m.indptr)
    np.testing.assert_array_equal(ci, m.indices)
    np.testing.assert_allclose(val, m.data, rtol=1e-14, atol=1e-14)


def test_spmv_transpose_matches_dense():
    rng = np.random.default_rng(2)
    a = sp.random(30, 45, density=0.2, random_state=3, format="csr")
    x = rng.standard_normal(30)
    y = osp.spmv_t(a.indptr.astype(np.int64), a.indices.astype(np.int64), a.data, x, 45)
    np.testing.assert_allclose(y, a.toarray().T @ x, rtol=1e-13, atol=1e-13)


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    m = sp.random(n, n, density=0.05, random_state=seed)
    a = (m @ m.T + sp.identity(n) * 2.0).tocsr()
    a.sort_indices()
    return a, rng.standard_normal(n)


@pytest.mark.parametrize("solver", ["cg", "jacobi_pcg"])
def test_cg_solves_spd_system(solver):
    a, b = _spd(80, 4)
    x, it, rr = getattr(osp, solver)(a.indptr.astype(np.int64), a.indices.astype(np.int64), a.data, b, 1000, 1e-12)
    np.testing.assert_allclose(x, np.linalg.solve(a.toarray(), b), rtol=1e-9, atol=1e-10)
    assert rr < 1e-12 and it <= 80 + 5


def test_cg_exact_in_n_steps_on_tiny_system():
    a = sp.csr_matrix(np.array([[4.0, 1.0, 0.0], [1.0, 3.0, 1.0], [0.0, 1.0, 2.0]]))
    b = np.array([1.0, 2.0, 3.0])
    x, it, rr = osp.cg(a.indptr.astype(np.int64), a.indices.astype(np.int64), a.data, b, 3, 0.0)
    assert it == 3
    np.testing.assert_allclose(x, np.linalg.solve(a.toarray(), b), rtol=1e-12)


def test_cg_zero_rhs_returns_zero():
    a, _ = _spd(10, 5)
    x, it, rr = osp.cg(a.indptr.astype(np.int64), a.indices.astype(np.int64), a.data, np.zeros(10))
    assert it == 0 and not x.any()


@pytest.mark.parametrize("precond", [0, 1])
def test_cgls_matches_scipy_lsqr(precond):
    rng = np.random.default_rng(6)
    a = sp.random(120, 70, density=0.1, random_state=7, format="csr")
    a.sort_indices()
    b = rng.standard_normal(120)
    lam = 1e-2
    x, it, rr, hist = osp.cgls(a.indptr.astype(np.int64), a.indices.astype(np.int64), a.data, b, 70, lam, 1e-12,
                               2000, precond)
    ref = spla.lsqr(a, b, damp=np.sqrt(lam), atol=1e-15, btol=1e-15, iter_lim=10000)[0]
    np.testing.assert_allclose(x, ref, rtol=1e-8, atol=1e-9)
    dense = a.toarray()
    np.testing.assert_allclose((dense.T @ dense + lam * np.eye(70)) @ x, dense.T @ b, atol=1e-9)
    assert np.all(np.diff(hist) < 1e3)   # history is finite


def test_siddon_axis_aligned_and_diagonal():
    n = (5, 4, 3)
    cells, lens = osp.siddon(n, 1.0, [0.0, 1.5, 2.5], [5.0, 1.5, 2.5])
    np.testing.assert_array_equal(cells, [(ix * 4 + 1) * 3 + 2 for ix in range(5)])
    np.testing.assert_allclose(lens, 1.0, rtol=0, atol=1e-15)
    # main diagonal of a cube: sqrt(3) per cell
    cells, lens = osp.siddon((4, 4, 4), 2.0, [0.0, 0.0, 0.0], [8.0, 8.0, 8.0])
    np.testing.assert_array_equal(cells, [(i * 4 + i) * 4 + i for i in range(4)])
    np.testing.assert_allclose(lens, 2.0 * np.sqrt(3.0), rtol=1e-14)


def test_siddon_matches_fine_sampling():
    rng = np.random.default_rng(8)
    n = (6, 5, 7)
    for _ in range(20):
        p0 = rng.random(3) * n
        p1 = rng.random(3) * n
        cells, lens = osp.siddon(n, 1.0, p0, p1)
        L = np.linalg.norm(p1 - p0)
        assert lens.sum() == pytest.approx(L, rel=1e-12)
        m = 200000
        t = (np.arange(m) + 0.5) / m
        pts = p0 + t[:, None] * (p1 - p0)
        idx = np.floor(pts).astype(int)
        flat = (idx[:, 0] * n[1] + idx[:, 1]) * n[2] + idx[:, 2]
        est = np.bincount(flat, minlength=np.prod(n)) * (L / m)
        got = np.zeros(np.prod(n))
        np.add.at(got, cells, lens)
        np.testing.assert_allclose(got, est, atol=3 * L / m + 1e-12)


def test_smooth_gradient_closed_forms():
    """Constant fields are fixed points (renormalised weights); an interior
    delta returns the normalised truncated Gaussian, whose centre value is
    1 / (sum_{|d| <= r} exp(-d^2 / (2 sigma^2)))^3 with r = int(3 sigma);
    the filter is linear and symmetric (self-adjoint away from the edges)."""
    n = (13, 14, 13)
    V = int(np.prod(n))
    for sigma in (0.4, 1.0, 1.2):
        assert np.allclose(osp.smooth_gradient(np.full(V, 2.5), *n, sigma), 2.5, rtol=1e-14)
        r = int(3 * sigma)
        z1 = sum(np.exp(-d * d / (2 * sigma ** 2)) for d in range(-r, r + 1))
        d = np.zeros(n)
        d[6, 7, 6] = 1.0    # every cell within r of it has its whole cube inside the grid
        out = osp.smooth_gradient(d.ravel(), *n, sigma).reshape(n)
        assert out[6, 7, 6] == pytest.approx(1.0 / z1 ** 3, rel=1e-12)
        assert out[6, 7, 7] == pytest.approx(np.exp(-1 / (2 * sigma ** 2)) / z1 ** 3, rel=1e-12)
        assert out[7, 8, 7] == pytest.approx(np.exp(-3 / (2 * sigma ** 2)) / z1 ** 3, rel=1e-12)
        assert out.sum() == pytest.approx(1.0, rel=1e-12)
        assert np.count_nonzero(out) == (2 * r + 1) ** 3
    # sigma < 1/3: radius 0, identity
    x = np.random.default_rng(0).standard_normal(V)
    assert np.array_equal(osp.smooth_gradient(x, *n, 0.3), x)
    # short input -> zeros (sparse_ops.cpp:122)
    assert not osp.smooth_gradient(np.ones(V - 1), *n, 1.0).any()
