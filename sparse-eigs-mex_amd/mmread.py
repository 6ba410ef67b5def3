"""Matrix Market reader for speigs' input format.

The reference's real-matrix protocol (docs/BENCHMARK-PROTOCOL.md:22-36) runs
on SuiteSparse Matrix Collection files (bench/fetch_matrices.sh:10-19), which
are Matrix Market coordinate files.  This reads one into what speigs() /
speigs_z() take: CSC arrays (int64 column pointers and row indices, rows
sorted within each column, float64 values; split real and imaginary values
for a Hermitian matrix), of the upper triangle or of both triangles.

speigs() reads the upper triangle of either storage (include/speigs.h:72-74),
so a real matrix is returned upper-only by default.  speigs_z() needs both
triangles: it embeds H = R + iS as [[R, -S], [S, R]] column by column from the
stored entries (src/speigs_core.c:679-692), and that embedding is symmetric
only when every stored (i, j) has its (j, i) -- an upper-only Hermitian
input makes the embedding fail the symmetry test -- so a complex matrix is
returned in full storage by default (``full=`` overrides either default).

Supported headers: ``%%MatrixMarket matrix coordinate {real|integer|pattern|complex}
{general|symmetric|hermitian}``.  Symmetric and Hermitian files
store one triangle (either one: an entry (i, j) stands for (j, i) too, the
Hermitian one conjugated); a general file must be symmetric (Hermitian) to be
solved, and its upper triangle is what speigs reads, as the reference reads
the upper triangle (include/speigs.h:72-74) -- use ``check=True`` to have the
mirror entries compared.  Duplicate entries are summed.  Pattern files get
value 1.  Skew-symmetric files (A^T = -A) are refused: speigs solves
symmetric and Hermitian problems only.  Files must already be local.
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class MMMatrix:
    n: int
    Ap: np.ndarray          # int64, n + 1
    Ai: np.ndarray          # int64 rows, sorted per column
    Ax: np.ndarray          # float64 (real part for a complex matrix)
    Axi: np.ndarray | None  # float64 imaginary parts, or None for a real matrix
    field: str
    symmetry: str
    full: bool              # both triangles stored (else the upper one only)

    @property
    def csc(self):
        return self.Ap, self.Ai, self.Ax


def _open(path: Path):
    if path.suffix == ".gz":
        return gzip.open(path, "rt")
    return open(path, "rt")


def mmread(path, check: bool = False, full: bool | None = None) -> MMMatrix:
    """Read a Matrix Market coordinate file into speigs' CSC input; ``full``
    None stores a complex matrix's both triangles, a real one's upper one."""
    path = Path(path)
    with _open(path) as f:
        head = f.readline().split()
        if len(head) != 5 or head[0].lower() != "%%matrixmarket" or head[1].lower() != "matrix":
            raise ValueError(f"{path}: not a Matrix Market matrix file")
        fmt, field, sym = (h.lower() for h in head[2:])
        if fmt != "coordinate":
            raise ValueError(f"{path}: only the coordinate format is read (got {fmt})")
        if field not in ("real", "integer", "pattern", "complex"):
            raise ValueError(f"{path}: unknown field {field}")
        if sym == "skew-symmetric":
            raise ValueError(f"{path}: a skew-symmetric matrix is neither symmetric nor Hermitian")
        if sym not in ("general", "symmetric", "hermitian"):
            raise ValueError(f"{path}: unknown symmetry {sym}")
        if sym == "hermitian" and field != "complex":
            raise ValueError(f"{path}: a Hermitian file has complex entries")
        line = f.readline()
        while line and (line.startswith("%") or not line.strip()):
            line = f.readline()
        m, n, nnz = (int(t) for t in line.split())
        if m != n:
            raise ValueError(f"{path}: speigs needs a square matrix ({m} x {n})")
        ncol = {"pattern": 2, "complex": 4}.get(field, 3)
        data = np.loadtxt(f, dtype=np.float64, ndmin=2, max_rows=nnz) if nnz else np.zeros((0, ncol))
    if data.shape != (nnz, ncol):
        raise ValueError(f"{path}: expected {nnz} entries of {ncol} fields, read {data.shape}")
    i = data[:, 0].astype(np.int64) - 1
    j = data[:, 1].astype(np.int64) - 1
    if nnz and (i.min() < 0 or j.min() < 0 or i.max() >= n or j.max() >= n):
        raise ValueError(f"{path}: index out of range")
    if field == "pattern":
        re, im = np.ones(nnz), None
    elif field == "complex":
        re, im = data[:, 2].copy(), data[:, 3].copy()
        if sym == "symmetric":
            # complex symmetric (A^T = A) is Hermitian only when it is real
            if np.any(im != 0.0):
                raise ValueError(f"{path}: a complex symmetric matrix with imaginary parts is not Hermitian")
            im = None
    else:
        re, im = data[:, 2].copy(), None
    if sym == "general":
        if check:
            _check_mirror(n, i, j, re, im, path)
        keep = i <= j
        i, j, re = i[keep], j[keep], re[keep]
        im = None if im is None else im[keep]
    else:
        # one triangle stored: move every entry to the upper one (i <= j);
        # a moved Hermitian entry is conjugated
        low = i > j
        if im is not None:
            if sym == "hermitian":
                im = np.where(low, -im, im)
            if sym == "hermitian" and np.any((i == j) & (im != 0.0)):
                raise ValueError(f"{path}: a Hermitian matrix has a real diagonal")
        i, j = np.where(low, j, i), np.where(low, i, j)
    # CSC by column, rows sorted, duplicates summed
    order = np.lexsort((i, j))
    i, j, re = i[order], j[order], re[order]
    im = None if im is None else im[order]
    if len(i):
        first = np.ones(len(i), dtype=bool)
        first[1:] = (i[1:] != i[:-1]) | (j[1:] != j[:-1])
        grp = np.cumsum(first) - 1
        re = np.bincount(grp, weights=re)
        if im is not None:
            im = np.bincount(grp, weights=im)
        i, j = i[first], j[first]
    if full is None:
        full = im is not None
    if full:
        # mirror the strictly upper entries, conjugated, and sort again
        off = i < j
        i, j = np.concatenate([i, j[off]]), np.concatenate([j, i[off]])
        re = np.concatenate([re, re[off]])
        im = None if im is None else np.concatenate([im, -im[off]])
        order = np.lexsort((i, j))
        i, j, re = i[order], j[order], re[order]
        im = None if im is None else im[order]
    Ap = np.zeros(n + 1, dtype=np.int64)
    np.add.at(Ap, j + 1, 1)
    Ap = np.cumsum(Ap)
    return MMMatrix(n, Ap, i.astype(np.int64), re.astype(np.float64),
                    None if im is None else im.astype(np.float64), field, sym, bool(full))


def _check_mirror(n, i, j, re, im, path):
    z = re if im is None else re + 1j * im
    up = {(a, b): v for a, b, v in zip(i.tolist(), j.tolist(), z.tolist()) if a < b}
    lo = {(b, a): v for a, b, v in zip(i.tolist(), j.tolist(), z.tolist()) if a > b}
    for key in set(up) | set(lo):
        u, w = up.get(key, 0.0), lo.get(key, 0.0)
        w = np.conj(w) if im is not None else w
        if abs(u - w) > 1e-12 * max(abs(u), abs(w), 1e-300):
            raise ValueError(f"{path}: not symmetric at {key}")


def mmwrite(path, Ap, Ai, Ax, Axi=None, comment: str = "") -> None:
    """Write speigs' CSC input (upper-only or full storage; the upper
    triangle is what is read) as a symmetric (Hermitian, given Axi) Matrix
    Market coordinate file, stored as the lower triangle the way the
    SuiteSparse Matrix Collection stores its symmetric matrices: the stored
    entry (j, i) of the upper entry (i, j) is conjugated."""
    Ap = np.asarray(Ap, dtype=np.int64)
    Ai = np.asarray(Ai, dtype=np.int64)
    Ax = np.asarray(Ax, dtype=np.float64)
    n = len(Ap) - 1
    col = np.repeat(np.arange(n, dtype=np.int64), np.diff(Ap))
    keep = Ai <= col                                   # full storage: its upper half
    r, c, x = Ai[keep], col[keep], Ax[keep]
    # lower-triangle entry (c, r), ordered by its column r then row c
    order = np.lexsort((c, r))
    r, c, x = r[order], c[order], x[order]
    with open(path, "wt") as f:
        if Axi is None:
            f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        else:
            f.write("%%MatrixMarket matrix coordinate complex hermitian\n")
        for line in comment.splitlines():
            f.write(f"% {line}\n")
        f.write(f"{n} {n} {len(r)}\n")
        if Axi is None:
            for a, b, v in zip((c + 1).tolist(), (r + 1).tolist(), x.tolist()):
                f.write(f"{a} {b} {v!r}\n")
        else:
            y = -np.asarray(Axi, dtype=np.float64)[keep][order]
            for a, b, v, w in zip((c + 1).tolist(), (r + 1).tolist(), x.tolist(), y.tolist()):
                f.write(f"{a} {b} {v!r} {w!r}\n")
