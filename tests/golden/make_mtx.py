"""Writes the Matrix Market fixtures under tests/golden/ (run from anywhere:
python tests/golden/make_mtx.py).  Every fixture has a closed-form spectrum,
so the reader is checked end to end against mathematics, not against itself:

  lap1d_30_lower.mtx     real symmetric, lower triangle stored (the SuiteSparse
                         convention): 1-D Dirichlet Laplacian, n = 30,
                         lambda_j = 2 - 2 cos(j pi / 31), j = 1..30
  lap2d_8_general.mtx    integer general (both triangles, entries shuffled,
                         one diagonal entry split into two duplicates that
                         sum): 2-D Dirichlet Laplacian on 8 x 8,
                         lambda = (2 - 2cos(p pi/9)) + (2 - 2cos(q pi/9))
  ring_12_pattern.mtx    pattern symmetric: cycle graph C_12 adjacency,
                         lambda = 2 cos(2 pi j / 12)
  flux_ring_16_herm.mtx  complex Hermitian, lower triangle: ring of 16 sites
                         with hopping -exp(i theta) and diagonal 2,
                         lambda_m = 2 - 2 cos(2 pi m / 16 + theta), theta = 0.3
  upper_strip_6.mtx      real symmetric with entries in BOTH triangles of a
                         symmetric file (allowed: each stands for its mirror),
                         tridiagonal n = 6 with diagonal 2 and off-diagonal -1
"""
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _write(name, head, n, rows, comment):
    with open(HERE / name, "wt") as f:
        f.write(f"%%MatrixMarket matrix coordinate {head}\n")
        f.write(f"% {comment}\n")
        f.write(f"{n} {n} {len(rows)}\n")
        for r in rows:
            f.write(" ".join(str(t) for t in r) + "\n")


def main():
    n = 30
    rows = []
    for j in range(1, n + 1):
        rows.append((j, j, "2.0"))
        if j < n:
            rows.append((j + 1, j, "-1.0"))
    _write("lap1d_30_lower.mtx", "real symmetric", n, rows, "1-D Dirichlet Laplacian, n = 30")

    m = 8
    rows = []
    for x in range(m):
        for y in range(m):
            a = x * m + y + 1
            if a == 1:
                rows += [(a, a, 1), (a, a, 3)]
            else:
                rows.append((a, a, 4))
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                if 0 <= x + dx < m and 0 <= y + dy < m:
                    rows.append((a, (x + dx) * m + y + dy + 1, -1))
    perm = np.random.default_rng(5).permutation(len(rows))
    _write("lap2d_8_general.mtx", "integer general", m * m, [rows[p] for p in perm],
           "2-D Dirichlet Laplacian on 8 x 8, both triangles, shuffled; entry (1,1) split as 1 + 3")

    n = 12
    rows = [(j + 2, j + 1) for j in range(n - 1)] + [(n, 1)]
    _write("ring_12_pattern.mtx", "pattern symmetric", n, rows, "cycle graph C_12 adjacency")

    n, th = 16, 0.3
    rows = [(j, j, "2.0", "0.0") for j in range(1, n + 1)]
    # upper entry H[j, j+1] = -exp(i th); stored lower (j+1, j) = conj = -exp(-i th)
    # the wrap entry H[n-1, 0] = -exp(i th) (site n-1 hops to site 0) is lower already
    for j in range(1, n):
        rows.append((j + 1, j, repr(float(-np.cos(th))), repr(float(np.sin(th)))))
    rows.append((n, 1, repr(float(-np.cos(th))), repr(float(-np.sin(th)))))
    _write("flux_ring_16_herm.mtx", "complex hermitian", n, rows,
           "ring of 16 sites, hopping -exp(i 0.3), diagonal 2")

    n = 6
    rows = [(j, j, "2") for j in range(1, n + 1)]
    rows += [(1, 2, "-1"), (3, 2, "-1"), (3, 4, "-1"), (5, 4, "-1"), (5, 6, "-1")]
    _write("upper_strip_6.mtx", "real symmetric", n, rows, "tridiagonal 6, entries in both triangles")


if __name__ == "__main__":
    main()
