"""Launch shapes of the FFT passes, checked on the host for every line length
with a compiled schedule, every tile width the GHX_LZ/GHX_LY knobs or the
ensemble autotuner can request, and both LDS limits the library may meet (64
KiB and gfx950's 160 KiB per block).  The launchers compute the same shapes
and refuse a launch that fails the same check, so no setting can launch more
lanes than a kernel's __launch_bounds__, a partial line, a tile wider than
the 32-bit line masks, or more LDS than the device allows."""
import ctypes as C

import pytest

import ghost_amd

SIZES = [2, 3, 4, 5, 6, 8, 10, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 1024]
X2 = [24, 32, 48, 64, 96, 128, 192, 256, 384, 512]
WANTS = [-7, 0, 1, 2, 3, 4, 5, 8, 9, 16, 31, 32, 33, 64, 512, 1 << 20]
LIMITS = [65536, 163840]
TILE_BOUND, ROW_BOUND = 512, 256


def shape(kind, n, nlines, want, limit):
    lib = ghost_amd.load()
    lines, block, lanes = C.c_int(), C.c_int(), C.c_int()
    lds = C.c_longlong()
    rc = lib.ghx_pass_shape(kind, n, nlines, want, C.c_longlong(limit), C.byref(lines), C.byref(block),
                            C.byref(lds), C.byref(lanes))
    return rc, lines.value, block.value, lds.value, lanes.value


@pytest.mark.parametrize("kind", [0, 1])
def test_tile_passes(kind):
    for n in SIZES:
        for nlines in (1, 2, 3, 5, 8, 17, 33, 96, 257, 513, 1024):
            for want in WANTS:
                for limit in LIMITS:
                    rc, L, block, lds, T = shape(kind, n, nlines, want, limit)
                    assert rc == 0, (kind, n, nlines, want, limit)
                    assert L >= 1 and L & (L - 1) == 0 and L <= 32
                    assert L <= max(1, nlines) and (want < 2 or L <= want)
                    assert block == L * T <= TILE_BOUND
                    rows = n + n // 8 + 1 + (n if kind == 1 else 0)
                    assert lds == L * rows * 16 <= limit


def test_wide_y_inverse():
    for n in SIZES:
        rc, L, block, lds, T = shape(2, n, 257, 8, 163840)
        if n == 512:
            assert rc == 0 and L == 8 and T == 32 and block == 256 and lds == 8 * 577 * 16
            # 73.9 KB: refused under a 64 KiB limit, where the launcher takes the tile pass
            assert shape(2, n, 257, 8, 65536)[0] == -2
        else:
            assert rc == -1  # no wide schedule: the launcher takes the tile pass


@pytest.mark.parametrize("kind", [3, 4])
def test_row_passes(kind):
    for n in SIZES:
        for limit in LIMITS:
            rc, R, block, lds, T = shape(kind, n, 1, 0, limit)
            if kind == 3 and n not in X2:
                assert rc == -1
                continue
            assert rc == 0, (kind, n, limit)
            assert R >= 1 and block == R * T <= ROW_BOUND
            assert lds == R * (n + n // 8 + 1) * 16 <= limit
            if (ROW_BOUND // T) * (n + n // 8 + 1) * 16 <= limit:
                assert block == ROW_BOUND or T > ROW_BOUND  # full blocks when LDS allows


def test_unknown_sizes_and_kinds():
    assert shape(0, 7, 8, 8, 65536)[0] == -1
    assert shape(9, 64, 8, 8, 65536)[0] == -1
    assert shape(0, 64, 0, 8, 65536)[0] == -1


def test_row_block_shrinks_under_lds_limit():
    """A full 256-lane x block of 192-point rows (6 points per lane, 32 lanes,
    8 rows) asks 8 padded rows of LDS, 27.8 KB; under a 16 KiB limit the
    block holds half the rows."""
    rc, R, block, lds, T = shape(4, 192, 1, 0, 163840)
    assert rc == 0 and T == 32 and R == 8 and lds == 8 * 217 * 16
    rc, R, block, lds, T = shape(4, 192, 1, 0, 16384)
    assert rc == 0 and T == 32 and R == 4 and block == 128 and lds == 4 * 217 * 16 <= 16384
