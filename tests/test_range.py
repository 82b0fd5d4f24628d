"""range (GHOST fftp-3/fftp.f90:17-48), the block distribution that GHOST's
callers size their slabs with, as the Fortran drop-in computes it.  Integer
work, so the check is exact over a sweep of (n1, n2, nprocs, irank): the ranks'
ranges tile n1..n2 in rank order, the first mod(n2-n1+1, nprocs) ranks hold one
point more than the others, and ranks beyond the point count get empty ranges
(iend = ista - 1).  The 512-point slab counts are written out by hand."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ghost_amd", "fortran", "range_table")


@pytest.fixture(scope="module")
def table():
    if not os.path.exists(EXE):
        pytest.skip("build the Fortran host side first (__graft_entry__.build)")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=60, check=True).stdout
    return np.array([[int(x) for x in line.split()] for line in out.splitlines()], dtype=np.int64)


def test_range_tiles_the_interval(table):
    groups = {}
    for n1, n2, p, r, ista, iend in table:
        groups.setdefault((n1, n2, p), {})[r] = (ista, iend)
    assert len(groups) == 3 * 42 * 12 + 4 * 64
    for (n1, n2, p), ranks in groups.items():
        assert sorted(ranks) == list(range(p))
        npts = n2 - n1 + 1
        nxt = n1
        for r in range(p):
            ista, iend = ranks[r]
            size = iend - ista + 1
            assert ista == nxt, (n1, n2, p, r)
            assert size == npts // p + (1 if r < npts % p else 0), (n1, n2, p, r)
            nxt = iend + 1
        assert nxt == n2 + 1


def test_range_known_slabs(table):
    rows = {(int(n2), int(p), int(r)): (int(a), int(b)) for n1, n2, p, r, a, b in table if n1 == 1}
    # 257 kx columns of a 512^3 grid over 8 ranks: 33 on rank 0, 32 on the rest
    assert rows[(257, 8, 0)] == (1, 33)
    assert rows[(257, 8, 1)] == (34, 65)
    assert rows[(257, 8, 7)] == (226, 257)
    # 512 z planes over 8 ranks: 64 each
    assert [rows[(512, 8, r)] for r in range(8)] == [(1 + 64 * r, 64 * (r + 1)) for r in range(8)]
    # 256 z planes over 3 ranks: 86, 85, 85
    assert [rows[(256, 3, r)] for r in range(3)] == [(1, 86), (87, 171), (172, 256)]
    # one rank: the whole range, as GHOST's callers use it on one GPU
    assert rows[(129, 1, 0)] == (1, 129)
