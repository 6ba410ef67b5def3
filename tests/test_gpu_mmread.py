"""Matrix Market files through the product library on the GPU: the fixtures
under tests/golden/ solved by speigs / speigs_z give their closed-form
spectra and the oracle's values, and a weighted Laplacian written as a
SuiteSparse-style lower-triangle file and read back solves to the same
eigenvalues as the generator's arrays (the reader is lossless)."""
import numpy as np
import pytest

from conftest import ROOT, load_pkg
from test_mmread import closed_form

pkg = load_pkg()
pytestmark = pytest.mark.gpu
GOLD = ROOT / "tests" / "golden"


def _case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("name,k,which", [
    ("lap1d_30_lower", 5, "sm"),
    ("lap2d_8_general", 3, "sm"),
    ("upper_strip_6", 2, "sm"),
    ("ring_12_pattern", 2, "lm"),
])
def test_fixture_speigs(hiplib, oracle, name, k, which):
    _case(f"{name} k={k} {which}")
    m = pkg.mmread(GOLD / f"{name}.mtx")
    r = pkg.speigs(m.csc, k, which)
    o = pkg.speigs(m.csc, k, which, lib=oracle, prefix="oracle_")
    assert r.rc == 0 and o.rc == 0 and r.info["flag"] == 0
    ex = closed_form(name)
    want = ex[np.argsort(np.abs(ex), kind="stable")]
    want = want[:k] if which == "sm" else want[::-1][:k]
    assert np.abs(np.sort(r.lam) - np.sort(want)).max() <= 1e-12
    assert np.abs(np.sort(r.lam) - np.sort(o.lam)).max() <= 1e-12


def test_fixture_speigs_z(hiplib, oracle):
    _case("flux_ring_16_herm k=4 sm")
    h = pkg.mmread(GOLD / "flux_ring_16_herm.mtx")
    r = pkg.speigs_z(h.Ap, h.Ai, h.Ax, h.Axi, 4, "sm")
    o = pkg.speigs_z(h.Ap, h.Ai, h.Ax, h.Axi, 4, "sm", lib=oracle, prefix="oracle_")
    assert r.rc == 0 and o.rc == 0
    assert np.abs(r.lam - closed_form("flux_ring_16_herm")[:4]).max() <= 1e-12
    assert np.abs(r.lam - o.lam).max() <= 1e-12
    H = np.zeros((h.n, h.n), dtype=complex)
    col = np.repeat(np.arange(h.n), np.diff(h.Ap))
    H[h.Ai, col] = h.Ax + 1j * h.Axi
    assert np.linalg.norm(H @ r.V - r.V * r.lam, axis=0).max() <= 1e-10


def test_weighted_lap_file_round_trip(hiplib, tmp_path):
    """30^3 weighted Laplacian + 1e-3 I (k = 12, tol 1e-10): the file read in
    full storage solves to the generator's eigenvalues within 1e-12 relative,
    lambda_0 = 1e-3."""
    _case("weighted_lap3d(30, seed=5) via .mtx, k=12")
    A = pkg.gen.weighted_lap3d(30, seed=5)
    f = tmp_path / "w30.mtx"
    pkg.mmwrite(f, *A, comment="weighted graph Laplacian 30^3 + 1e-3 I, seed 5")
    m = pkg.mmread(f, full=True)
    r0 = pkg.speigs(A, 12, "sm", opts={"tol": 1e-10})
    r1 = pkg.speigs(m.csc, 12, "sm", opts={"tol": 1e-10})
    print("lam diff", np.abs(r0.lam - r1.lam).max())
    assert r0.rc == 0 and r1.rc == 0
    assert np.abs(r0.lam - r1.lam).max() <= 1e-12 * np.abs(r0.lam).max()
    assert abs(r1.lam[0] - 1e-3) <= 1e-13
