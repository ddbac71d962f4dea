"""The Python names the reference's own modules import from seismic_engine
(gradient.py:5, model.py:6, inversion.py:12-15), called the way those modules
call them, against the oracle."""
import numpy as np
import pytest
import torch

from oracle import fd as ofd
from oracle import sparse as osp

pytestmark = pytest.mark.gpu


def test_health_functions_follow_inversion_call_pattern(cuda):
    from seismic_engine import check_field_health, clamp_field_safe, repair_field_nan
    rng = np.random.default_rng(1)
    vp = rng.uniform(1500, 3000, (6, 5, 4))
    vp[2, 2, 2] = np.nan
    vp[0, 1, 1] = np.inf
    vp[3, 3, 1] = -np.inf
    # inversion.py:295-301
    assert int(check_field_health(np.ascontiguousarray(vp, dtype=np.float64))) == 3
    # inversion.py:303-317: repair a copy in place, with-halo extents = model extents, halo 1
    repaired = np.ascontiguousarray(vp.copy(), dtype=np.float64)
    out = repair_field_nan(repaired, 6, 5, 4, halo_width=1, replace_value=1500.0)
    ref = ofd.repair_field_nan(vp.ravel().copy(), 6, 5, 4, 1, 1500.0)
    assert np.array_equal(repaired.ravel(), ref) and np.array_equal(np.asarray(out).ravel(), ref)
    flat = vp.ravel().copy()
    c = clamp_field_safe(flat, lo=1600.0, hi=2900.0)     # in place, like the forcecast binding
    want = np.where(np.isfinite(vp.ravel()), np.clip(vp.ravel(), 1600.0, 2900.0), 0.0)
    assert c is flat and np.array_equal(flat, want)
    # default limits +-1e10 and torch tensors in place
    t = torch.tensor([1e11, -1e12, float("nan"), 3.0], dtype=torch.float64, device=cuda)
    clamp_field_safe(t)
    assert t.cpu().tolist() == [1e10, -1e10, 0.0, 3.0]


# radius int(3 sigma): <= 8 takes the windowed/LDS passes, 10 the general one;
# the larger boxes cross the 64-cell walk segments and the 256-cell z tiles
@pytest.mark.parametrize("sigma,n", [(0.5, (12, 10, 9)), (1.0, (12, 10, 9)), (2.0, (12, 10, 9)),
                                     (2.7, (12, 10, 9)), (3.5, (12, 10, 9)), (1.0, (70, 40, 300)),
                                     (2.7, (66, 8, 258))])
def test_smooth_gradient_matches_oracle(sigma, n, cuda):
    from seismic_engine import SparseOps
    g = np.random.default_rng(2).standard_normal(int(np.prod(n)))
    got = SparseOps.smooth_gradient(g, *n, sigma)
    ref = osp.smooth_gradient(g, *n, sigma)
    assert isinstance(got, np.ndarray) and got.dtype == np.float64
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14)
    assert not SparseOps.smooth_gradient(g[:-1], *n, sigma).any()


def test_fd_matrix_call_pattern_of_model_py(cuda):
    from seismic_engine import CSRMatrix, SparseOps, compute_fd_stencil
    n = (5, 4, 6)
    vp = np.random.default_rng(3).uniform(1500, 2500, n)
    A = compute_fd_stencil(*n, vp.flatten(), 10.0, 10.0, 10.0, 1e-3)     # model.py:96-105
    assert isinstance(A, CSRMatrix) and A.rows() == A.cols() == 120
    rp, ci, val = (np.asarray(a) for a in (A.row_ptr(), A.col_ind(), A.values()))
    V = 120
    dense = np.zeros((V, V))
    for i in range(V):
        dense[i, ci[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    c = vp.ravel() ** 2 * 1e-6 / 100.0
    assert np.allclose(np.diag(dense), 2 - 6 * c, rtol=1e-14)
    assert dense[0, 1] == pytest.approx(c[0]) and dense[1, 0] == pytest.approx(c[1])
    # dot_product with two arrays; a CSRMatrix argument is a type error (gradient.py:47-49 misuses it)
    x = np.arange(V, dtype=float)
    assert SparseOps.dot_product(x, x) == pytest.approx(float(x @ x), rel=1e-15)
    with pytest.raises(TypeError):
        SparseOps.dot_product(A, x)


def test_rtm_engine_and_adjoint_gradient(cuda):
    from seismic_engine import RTMEngine, compute_adjoint_gradient
    n, hw = (6, 5, 4), 2
    nh = tuple(k + 2 * hw for k in n)
    rng = np.random.default_rng(4)
    f = rng.standard_normal(nh)
    a = rng.standard_normal(nh)
    f[3, 3, 3] = np.nan
    eng = RTMEngine(*n, halo_width=hw)
    eng.apply_image_condition(f, a, nh[1], nh[2], 0)
    eng.finalize_image()
    r = eng.result()
    hs = tuple(slice(hw, hw + k) for k in n)
    fi = np.where(np.isfinite(f), f, 0)[hs].ravel()
    ai = a[hs].ravel()
    assert np.array_equal(r.image, fi * ai) and np.array_equal(r.illumination, fi * fi)
    assert np.allclose(r.reflectivity, np.where(fi * fi > 1e-10, ai / np.where(fi == 0, 1, fi), 0))
    # gradient.py:24-40, same operation order
    fw, ad, v = (rng.standard_normal(60) for _ in range(3))
    g = compute_adjoint_gradient(fw, ad, v, 1e-3)
    assert np.array_equal(g, -fw * ad * (2.0 * v * 1e-3 * 1e-3))
