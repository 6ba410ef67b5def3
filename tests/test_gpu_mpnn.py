"""The learned start on the GPU: the network's device forward pass, loss and
gradient against the CPU oracle (oracle/mpnn.py, itself pinned by finite
differences in test_mpnn_oracle.py), training progress, and the solver
contract -- a learned start changes the iteration count, not the answer."""
import importlib.util

import numpy as np
import pytest

from conftest import ROOT, load_pkg

pkg = load_pkg()
gen = pkg.gen
pytestmark = pytest.mark.gpu
spec = importlib.util.spec_from_file_location("oracle_mpnn", ROOT / "oracle" / "mpnn.py")
om = importlib.util.module_from_spec(spec)
spec.loader.exec_module(om)


def _oracle_weights(p):
    Wb = []
    for W, b in om.unpack(p.astype(np.float64), 3):
        Wb.append((om.bf16_round(W.astype(np.float32)).reshape(W.shape), b))
    return Wb


def _eigvecs(A, k):
    D = gen.csc_to_dense(*A, symmetrize_upper=True)
    w, V = np.linalg.eigh(D)
    return V[:, :k]


def test_forward_matches_oracle(hiplib):
    A = gen.grf_lap3d(9, seed=4, corr=3.0)
    m = pkg.StartModel(3, seed=7)
    X = m.infer(A, seed=11).astype(np.float64)
    H0 = om.features(*A, seed=11, gen=gen)
    Hs, _ = om.forward(_oracle_weights(m.get_params()), A, H0)
    err = np.abs(X - Hs[-1]).max() / np.abs(Hs[-1]).max()
    print("forward max rel err", err)
    assert err <= 1e-4
    m.close()


def test_gradient_matches_oracle(hiplib):
    A = gen.grf_lap3d(9, seed=5, corr=3.0)
    U = _eigvecs(A, 6)
    m = pkg.StartModel(3, seed=3)
    loss, g = m.grad(A, U, seed=21)
    H0 = om.features(*A, seed=21, gen=gen)
    Lo, grads = om.loss_grad(_oracle_weights(m.get_params()), A, H0, U)
    go = om.flat(grads)
    print("loss gpu", loss, "oracle", Lo)
    assert abs(loss - Lo) <= 1e-5
    for l in range(3):
        sl = slice(l * om.LAYER_P, (l + 1) * om.LAYER_P)
        rel = np.linalg.norm(g[sl] - go[sl]) / np.linalg.norm(go[sl])
        print("layer", l, "grad rel err", rel)
        assert rel <= 2e-3
    m.close()


def test_training_reduces_loss(hiplib):
    insts = [gen.grf_lap3d(10, seed=100 + i, corr=4.0) for i in range(4)]
    Us = [_eigvecs(A, 8) for A in insts]
    m = pkg.StartModel(3, seed=1)
    first = np.mean([m.grad(A, U, seed=i)[0] for i, (A, U) in enumerate(zip(insts, Us))])
    for it in range(60):
        i = it % 4
        m.train_step(insts[i], Us[i], seed=1000 + it, lr=3e-3)
    last = np.mean([m.grad(A, U, seed=i)[0] for i, (A, U) in enumerate(zip(insts, Us))])
    print("mean loss before", first, "after", last)
    assert last < first - 0.05
    m.close()


@pytest.mark.parametrize("precision", [0, 2])
def test_learned_start_same_answer(hiplib, precision):
    """any network (here an untrained one) only moves the start: same eigenvalues
    (1e-10 relative) as the uniform start, residuals at the tolerance."""
    A = gen.grf_lap3d(16, seed=9)
    m = pkg.StartModel(3, seed=5)
    r0 = pkg.speigs(A, 10, "sm", opts={"tol": 1e-11}, hip_opts={"precision": precision})
    r1 = pkg.speigs(A, 10, "sm", opts={"tol": 1e-11},
                    hip_opts={"precision": precision, "start": 2, "model": m.handle.value})
    print("uniform", r0.info["nops"], "learned", r1.info["nops"])
    assert r0.rc == 0 and r1.rc == 0 and r1.info["flag"] == 0
    assert np.abs(r1.lam - r0.lam).max() <= 1e-10 * np.abs(r0.lam).max()
    assert r1.resid.max() <= 1e-11 * r1.info["anorm"]
    m.close()


def test_trained_start_gain(hiplib, oracle):
    """The committed network (models/start_grf32.bin, trained on the GPU on
    plain-solver eigenvectors of 32^3 GRF Laplacians, seeds 10,000..11,999) on
    three held-out instances of its family (seeds 950,000..950,002), k = 16,
    tol 1e-10, on the chosen path: the eigenspace part its block captures is
    far above an untrained network's, the block eigensolver needs fewer
    iterations in total than with the uniform start, and the answer is the
    uniform start's and (first instance) the oracle's to 1e-10 relative."""
    m = pkg.StartModel(3, seed=1)
    m.load(ROOT / "sparse-eigs-mex_amd" / "models" / "start_grf32.bin")
    u = pkg.StartModel(3, seed=1)                      # untrained, same architecture
    it_u = it_l = 0
    for j in range(3):
        A = gen.grf_lap3d(32, seed=950_000 + j)
        print(f"[case] grf_lap3d(32, seed={950_000 + j}) k=16", flush=True)
        r0 = pkg.speigs(A, 16, "sm", opts={"tol": 1e-10}, hip_opts={"precision": 2})
        r1 = pkg.speigs(A, 16, "sm", opts={"tol": 1e-10},
                        hip_opts={"precision": 2, "start": 2, "model": m.handle.value})
        assert r0.rc == 0 and r1.rc == 0 and r0.info["flag"] == 0 and r1.info["flag"] == 0
        assert np.abs(r1.lam - r0.lam).max() <= 1e-10 * np.abs(r0.lam).max()
        assert r1.resid.max() <= 1e-10 * r1.info["anorm"]
        lt, _ = m.grad(A, r0.V, seed=88172645463325252)
        lu, _ = u.grad(A, r0.V, seed=88172645463325252)
        print(f"instance {j}: iterations uniform {r0.info['restarts']} learned {r1.info['restarts']}, "
              f"loss trained {lt:.3f} untrained {lu:.3f}")
        assert lt <= 0.5 and lu >= 0.7 and lt <= lu - 0.25
        it_u += r0.info["restarts"]
        it_l += r1.info["restarts"]
        if j == 0:
            o = pkg.speigs(A, 16, "sm", opts={"tol": 1e-10}, lib=oracle, prefix="oracle_")
            assert o.rc == 0
            assert np.abs(r1.lam - o.lam).max() <= 1e-10 * np.abs(o.lam).max()
    print("total iterations: uniform", it_u, "learned", it_l)
    assert it_l < it_u
    m.close()
    u.close()
