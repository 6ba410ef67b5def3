"""The learned-start network's CPU oracle (oracle/mpnn.py) pinned by finite
differences: its analytic gradient of the subspace loss matches central
differences of the loss itself, so a dropped term, a transposed operand or a
wrong sign in the backward pass fails here."""
import importlib.util

import numpy as np
import pytest

from conftest import ROOT, load_pkg

pkg = load_pkg()
gen = pkg.gen
spec = importlib.util.spec_from_file_location("oracle_mpnn", ROOT / "oracle" / "mpnn.py")
om = importlib.util.module_from_spec(spec)
spec.loader.exec_module(om)


def _case():
    A = gen.grf_lap3d(5, seed=2, corr=2.0)
    D = gen.csc_to_dense(*A, symmetrize_upper=True)
    w, V = np.linalg.eigh(D)
    U = V[:, :4]
    H0 = om.features(*A, seed=9, gen=gen)
    rng = np.random.default_rng(3)
    params = (rng.standard_normal(3 * om.LAYER_P) * 0.2)
    return A, U, H0, params


def _loss(params, A, U, H0, ridge):
    Wb = om.unpack(params, 3)
    Hs, _ = om.forward(Wb, A, H0)
    return om.loss_and_dX(Hs[-1], U, ridge)[0]


def test_oracle_gradient_matches_finite_differences():
    A, U, H0, params = _case()
    ridge = om.ridge_of(om.forward(om.unpack(params, 3), A, H0)[0][-1])
    L, grads = om.loss_grad(om.unpack(params, 3), A, H0, U, ridge)
    g = om.flat(grads)
    assert 0.0 < L < 1.0
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(om.LAYER_P, 12, replace=False) + l * om.LAYER_P for l in range(3)])
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (_loss(p1, A, U, H0, ridge) - _loss(p2, A, U, H0, ridge)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.abs(g).max()), (i, fd, g[i])


def test_oracle_loss_closed_forms():
    """X spanning U gives loss 0 (up to the 1e-6 ridge); X orthogonal to U gives 1."""
    A, U, H0, params = _case()
    R = np.random.default_rng(1).standard_normal((U.shape[0], om.WIDTH - U.shape[1]))
    X = np.concatenate([U, R / np.linalg.norm(R, axis=0)], axis=1)
    assert abs(om.loss_and_dX(X, U)[0]) <= 1e-5
    P = np.eye(U.shape[0]) - U @ U.T
    Xo = P @ np.random.default_rng(2).standard_normal((U.shape[0], om.WIDTH))
    assert abs(om.loss_and_dX(Xo, U)[0] - 1.0) <= 1e-8
