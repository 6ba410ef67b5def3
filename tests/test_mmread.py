"""Matrix Market reader (sparse-eigs-mex_amd/mmread.py): the fixtures under
tests/golden/ (written by tests/golden/make_mtx.py) have closed-form spectra,
so each file is checked against mathematics (dense eigenvalues of what was
read), against scipy.io.mmread (an independent reader), and through the CPU
oracle's speigs / speigs_z, the reference's entry points restated
(src/speigs_core.c:783-1024, :660-768).  The GPU solve of the same files is
in tests/test_gpu_mmread.py."""
import gzip
import shutil

import numpy as np
import pytest
import scipy.io as sio

from conftest import ROOT, load_pkg

pkg = load_pkg()
GOLD = ROOT / "tests" / "golden"


def closed_form(name):
    if name == "lap1d_30_lower":
        return np.sort(2 - 2 * np.cos(np.arange(1, 31) * np.pi / 31))
    if name == "lap2d_8_general":
        t = 2 - 2 * np.cos(np.arange(1, 9) * np.pi / 9)
        return np.sort((t[:, None] + t[None, :]).ravel())
    if name == "ring_12_pattern":
        return np.sort(2 * np.cos(2 * np.pi * np.arange(12) / 12))
    if name == "flux_ring_16_herm":
        return np.sort(2 - 2 * np.cos(2 * np.pi * np.arange(16) / 16 + 0.3))
    if name == "upper_strip_6":
        return np.sort(2 - 2 * np.cos(np.arange(1, 7) * np.pi / 7))
    raise KeyError(name)


NAMES = ["lap1d_30_lower", "lap2d_8_general", "ring_12_pattern", "flux_ring_16_herm", "upper_strip_6"]


def dense_hermitian(m):
    """The full matrix a CSC read stands for (its upper triangle mirrored
    when only that is stored)."""
    z = m.Ax if m.Axi is None else m.Ax + 1j * m.Axi
    U = np.zeros((m.n, m.n), dtype=complex if m.Axi is not None else float)
    col = np.repeat(np.arange(m.n), np.diff(m.Ap))
    U[m.Ai, col] = z
    if m.full:
        return U
    D = np.diag(np.diag(U))
    return U + U.conj().T - D


@pytest.mark.parametrize("name", NAMES)
def test_fixture_structure_and_spectrum(name):
    m = pkg.mmread(GOLD / f"{name}.mtx", check=True)
    assert m.Ap.dtype == np.int64 and m.Ai.dtype == np.int64 and m.Ax.dtype == np.float64
    assert m.Ap[0] == 0 and m.Ap[-1] == len(m.Ai) == len(m.Ax) and np.all(np.diff(m.Ap) >= 0)
    col = np.repeat(np.arange(m.n), np.diff(m.Ap))
    assert m.full == (m.Axi is not None), "complex files default to full storage"
    if not m.full:
        assert np.all(m.Ai <= col), "upper triangle only"
    for c in range(m.n):
        rows = m.Ai[m.Ap[c]:m.Ap[c + 1]]
        assert np.all(np.diff(rows) > 0), "rows sorted and unique within a column"
    A = dense_hermitian(m)
    assert np.abs(np.linalg.eigvalsh(A) - closed_form(name)).max() <= 1e-13


@pytest.mark.parametrize("name", NAMES)
def test_equals_scipy(name):
    m = pkg.mmread(GOLD / f"{name}.mtx")
    S = sio.mmread(GOLD / f"{name}.mtx").toarray()
    if name == "lap2d_8_general":
        assert np.array_equal(np.triu(S), np.triu(dense_hermitian(m)))
    else:
        assert np.array_equal(S, dense_hermitian(m))


def test_duplicates_summed_and_mirrored_entries():
    m = pkg.mmread(GOLD / "lap2d_8_general.mtx")
    assert m.Ax[m.Ap[0]] == 4.0                       # (1,1) stored as 1 + 3
    u = pkg.mmread(GOLD / "upper_strip_6.mtx")          # entries in both triangles
    assert np.array_equal(u.Ap, [0, 1, 3, 5, 7, 9, 11])
    assert np.array_equal(u.Ai, [0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5])


def test_write_read_round_trip(tmp_path):
    """mmwrite stores the lower triangle (the SuiteSparse layout); reading it
    back gives the same CSC arrays bit for bit, real and Hermitian."""
    A = pkg.gen.weighted_lap3d(6, seed=3)
    f = tmp_path / "w.mtx"
    pkg.mmwrite(f, *A, comment="weighted Laplacian 6^3\nseed 3")
    m = pkg.mmread(f, check=True)
    assert m.symmetry == "symmetric"
    assert np.array_equal(m.Ap, A[0]) and np.array_equal(m.Ai, A[1]) and np.array_equal(m.Ax, A[2])
    S = sio.mmread(f).toarray()
    assert np.array_equal(S, dense_hermitian(m))
    h = pkg.mmread(GOLD / "flux_ring_16_herm.mtx")
    g = tmp_path / "h.mtx"
    pkg.mmwrite(g, h.Ap, h.Ai, h.Ax, h.Axi)
    h2 = pkg.mmread(g)
    assert h2.symmetry == "hermitian"
    for a, b in ((h.Ap, h2.Ap), (h.Ai, h2.Ai), (h.Ax, h2.Ax), (h.Axi, h2.Axi)):
        assert np.array_equal(a, b)


def test_full_storage(tmp_path):
    """full=True gives both triangles, as MATLAB stores a sparse symmetric
    matrix: the generator's full-storage 3-D Laplacian comes back bit for bit
    from a lower-triangle file; a Hermitian file read upper-only and then
    mirrored by hand equals its full read."""
    A = pkg.gen.lap3d(5, True)
    f = tmp_path / "l.mtx"
    pkg.mmwrite(f, *pkg.gen.upper_only(*A))
    m = pkg.mmread(f, full=True)
    assert m.full and np.array_equal(m.Ap, A[0]) and np.array_equal(m.Ai, A[1]) and np.array_equal(m.Ax, A[2])
    up = pkg.mmread(GOLD / "flux_ring_16_herm.mtx", full=False)
    fu = pkg.mmread(GOLD / "flux_ring_16_herm.mtx")
    assert not up.full and fu.full
    assert np.array_equal(dense_hermitian(up), dense_hermitian(fu))
    assert np.abs(dense_hermitian(fu) - dense_hermitian(fu).conj().T).max() == 0.0


def test_gzip(tmp_path):
    g = tmp_path / "lap1d.mtx.gz"
    with open(GOLD / "lap1d_30_lower.mtx", "rb") as src, gzip.open(g, "wb") as dst:
        shutil.copyfileobj(src, dst)
    a, b = pkg.mmread(g), pkg.mmread(GOLD / "lap1d_30_lower.mtx")
    assert np.array_equal(a.Ap, b.Ap) and np.array_equal(a.Ai, b.Ai) and np.array_equal(a.Ax, b.Ax)


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n0\n0\n1\n", "coordinate"),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1.0\n", "skew"),
    ("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1.0\n", "square"),
    ("%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n3 1 1.0\n", "range"),
    ("%%MatrixMarket matrix coordinate real symmetric\n2 2 2\n1 1 1.0\n", "entries"),
    ("%%MatrixMarket matrix coordinate complex symmetric\n2 2 1\n2 1 1.0 0.5\n", "Hermitian"),
    ("%%MatrixMarket matrix coordinate complex hermitian\n2 2 1\n1 1 1.0 0.5\n", "real diagonal"),
    ("%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 1 1.0\n", "complex"),
    ("%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1.0\n", "not a Matrix Market"),
])
def test_refused(tmp_path, text, msg):
    f = tmp_path / "bad.mtx"
    f.write_text(text)
    with pytest.raises(ValueError, match=msg):
        pkg.mmread(f)


def test_general_not_symmetric_checked(tmp_path):
    f = tmp_path / "ns.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 2.0\n1 2 1.0\n2 1 1.5\n")
    pkg.mmread(f)                                     # upper triangle read, as speigs reads it
    with pytest.raises(ValueError, match="not symmetric"):
        pkg.mmread(f, check=True)


def test_fixtures_through_oracle(oracle):
    """The files feed the reference's own entry points (restated in the
    oracle): closed-form values from speigs (SM, both storages) and from
    speigs_z on the Hermitian ring."""
    m = pkg.mmread(GOLD / "lap1d_30_lower.mtx")
    r = pkg.speigs(m.csc, 5, "sm", lib=oracle, prefix="oracle_")
    assert r.rc == 0 and np.abs(r.lam - closed_form("lap1d_30_lower")[:5]).max() <= 1e-12
    m = pkg.mmread(GOLD / "lap2d_8_general.mtx")
    r = pkg.speigs(m.csc, 3, "sm", lib=oracle, prefix="oracle_")
    assert r.rc == 0 and np.abs(r.lam - closed_form("lap2d_8_general")[:3]).max() <= 1e-12
    h = pkg.mmread(GOLD / "flux_ring_16_herm.mtx")
    r = pkg.speigs_z(h.Ap, h.Ai, h.Ax, h.Axi, 4, "sm", lib=oracle, prefix="oracle_")
    assert r.rc == 0 and np.abs(r.lam - closed_form("flux_ring_16_herm")[:4]).max() <= 1e-12
    # speigs_z's embedding needs both triangles (src/speigs_core.c:679-692):
    # the upper triangle alone makes [[R, -S], [S, R]] unsymmetric
    h = pkg.mmread(GOLD / "flux_ring_16_herm.mtx", full=False)
    r = pkg.speigs_z(h.Ap, h.Ai, h.Ax, h.Axi, 4, "sm", lib=oracle, prefix="oracle_")
    assert r.rc == -6
