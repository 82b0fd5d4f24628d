"""randu, the random-phase forcing and the k^4 exp initial condition of the
oracle, pinned by the textbook Park-Miller generator (computed here with
64-bit products, not Schrage's factorisation) and by a hand-derived
known-answer test of the draw order of examples/initialfv.f90_random."""
import numpy as np
import pytest

from oracle import oracle as O

MASK = 123459876
M = 2147483647


def minstd(t):
    return (16807 * t) % M


def test_randu_is_masked_park_miller():
    # Park & Miller (1988): from 1, the 10000th MINSTD value is 1043618065
    s = 1 ^ MASK
    for _ in range(10000):
        _, s = O.randu(s)
    assert s ^ MASK == 1043618065
    # each value is (t * 2^-31 - 0.5) * 2 with t the next MINSTD state
    s, t = 42 ^ MASK, 42
    for _ in range(200):
        v, s = O.randu(s)
        t = minstd(t)
        assert s ^ MASK == t
        assert v == (t * 2.0 ** -31 - 0.5) * 2  # am = 1./im in single precision is exactly 2^-31


def draws(seed, n):
    out = []
    for _ in range(n):
        v, seed = O.randu(seed)
        out.append(v)
    return out, seed


def test_forcing_draw_order_known_answer():
    """8^3 grid, shell kdn = kup = 1.  By the loops of initialfv.f90_random the
    draws go, three per mode (C1, C2, C3): (kx,ky,kz) = (0,1,0) [j loop, mirror
    (0,-1,0) conjugated], (0,0,1) [k loop, mirror (0,0,-1)], then (1,0,0) in the
    kx>0 planes with a factor 2.  Amplitude 1/|k| = 1.  Then f = curl C,
    normalised to <f^2> = f0^2."""
    n, f0, seed = 8, 0.7, 4242
    g = O.Grid(n)
    f, s_end = g.initialfv_random(1.0, 1.0, f0, seed)
    r, s_chk = draws(seed, 9)
    assert s_end == s_chk  # exactly nine draws
    C = [g.czeros() for _ in range(3)]
    ph = [np.exp(1j * 2 * np.pi * x) for x in r]
    for q in range(3):
        C[q][0, 1, 0] = ph[q]
        C[q][0, n - 1, 0] = np.conj(ph[q])
        C[q][0, 0, 1] = ph[3 + q]
        C[q][0, 0, n - 1] = np.conj(ph[3 + q])
        C[q][1, 0, 0] = 2 * ph[6 + q]
    want = g.normalize(*g.curl(*C), f0)
    for k in range(3):
        assert np.max(np.abs(f[k] - want[k])) < 1e-13 * np.max(np.abs(want[k]))


def test_forcing_structure():
    g = O.Grid(32)
    f, seed = g.initialfv_random(2.0, 3.0, 0.5, 77)
    KX, KY, KZ = np.meshgrid(g.kx[: g.nxh], g.ky, g.kz, indexing="ij")
    nz = (np.abs(f[0]) + np.abs(f[1]) + np.abs(f[2])) > 0
    assert np.all((g.kk2[nz] >= 4.0 - 1e-12) & (g.kk2[nz] <= 9.0 + 1e-12))
    div = KX * f[0] + KY * f[1] + KZ * f[2]
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(f[0]))
    assert abs(g.energy(*f, 1) - 0.25) < 1e-14
    # number of draws: three per in-shell mode of the fill (kx=0 half plane incl. its
    # mirror lines drawn once, every kx>0 mode)
    shell = (g.kk2 >= 4.0) & (g.kk2 <= 9.0)
    m0 = shell[0]
    ny = nz_ = 32
    count0 = sum(m0[j, 0] for j in range(1, ny // 2 + 1)) + sum(m0[0, k] for k in range(1, nz_ // 2 + 1)) \
        + sum(m0[j, k] for j in range(1, ny) for k in range(1, nz_ // 2 + 1))
    ndraw = 3 * (int(count0) + int(shell[1:].sum()))
    _, s_chk = draws(77, ndraw)
    assert seed == s_chk


def test_k4exp_spectrum_shape_and_normalisation():
    g = O.Grid(64)
    k0 = 3.0
    v, _ = g.initialv_k4exp(k0, 1.0, 2024)
    assert abs(g.energy(*v, 1) - 1.0) < 1e-14  # <u^2> = u0^2, E = 0.5
    assert g.divergence(*v) < 1e-24
    # all modes outside the dealias set are zero
    kept = (g.kn2 <= g.kmax) & (g.kn2 >= g.tiny)
    assert all(np.all(x[~kept] == 0) for x in v)
    # shell spectrum follows k^4 exp(-2 (k/k0)^2) up to sampling noise
    Ek, _ = g.spectrumc(*v, kin=1, hel=0)
    k = np.arange(1, len(Ek) + 1, dtype=float)
    model = k ** 4 * np.exp(-2 * (k / k0) ** 2)
    sel = slice(1, 9)
    ratio = Ek[sel] / model[sel]
    assert np.std(ratio) / np.mean(ratio) < 0.35
    assert np.argmax(Ek) + 1 in (3, 4)


# Draw tables derived by hand from the four loops of
# examples/initialfv.f90_random:16-111 (GHOST's Fortran indices j = ky + 1 for
# ky >= 0, j = ny + ky + 1 for ky < 0, likewise k for kz):
#   loop 1: kx = 0, kz = 0, j = 2..ny/2+1         mirror (0, -ky, 0) conjugated
#   loop 2: kx = 0, ky = 0, k = 2..nz/2+1         mirror (0, 0, -kz) conjugated
#   loop 3: kx = 0, j = 2..ny (outer), k = 2..nz/2+1 (inner), mirror (0, -ky, -kz)
#   loop 4: i = 2..nx/2+1, j = 1..ny, k = 1..nz (k innermost), factor 2, no mirror
# Each row is one mode (loop, kx, ky, kz) in draw order; it takes three draws,
# for C1, C2 and C3.  Mirrors are written right after their mode, so a later
# mode whose mirror lands on an earlier entry overwrites it.
TABLE_8 = [  # 8^3, kdn = 1, kup = 2.25: kk2 in {1, 2, 4, 5}
    (1, 0, 1, 0), (1, 0, 2, 0),
    (2, 0, 0, 1), (2, 0, 0, 2),
    (3, 0, 1, 1), (3, 0, 1, 2), (3, 0, 2, 1), (3, 0, -2, 1), (3, 0, -1, 1), (3, 0, -1, 2),
    (4, 1, 0, 0), (4, 1, 0, 1), (4, 1, 0, 2), (4, 1, 0, -2), (4, 1, 0, -1),
    (4, 1, 1, 0), (4, 1, 1, 1), (4, 1, 1, -1),
    (4, 1, 2, 0),
    (4, 1, -2, 0),
    (4, 1, -1, 0), (4, 1, -1, 1), (4, 1, -1, -1),
    (4, 2, 0, 0), (4, 2, 0, 1), (4, 2, 0, -1), (4, 2, 1, 0), (4, 2, -1, 0),
]
TABLE_4 = [  # 4^3, kdn = 2, kup = 2.2361: kk2 in {4, 5}; ky = 2 and kz = 2 are Nyquist
    (1, 0, 2, 0),            # j = 3 is its own mirror: the entry ends as the conjugate
    (2, 0, 0, 2),            # likewise k = 3
    (3, 0, 1, 2),            # j = 2, k = 3: mirror (0, -1, 2) on the same Nyquist plane
    (3, 0, 2, 1),            # j = 3, k = 2: mirror (0, 2, -1)
    (3, 0, -1, 2),           # j = 4, k = 3: overwrites (0, -1, 2) and its mirror (0, 1, 2)
    (4, 1, 0, 2), (4, 1, 2, 0),   # j = 1 (ky = 0) comes before j = 3 (ky = 2)
    (4, 2, 0, 0), (4, 2, 0, 1), (4, 2, 0, -1), (4, 2, 1, 0), (4, 2, -1, 0),
]


def fill_from_table(g, table, seed, amp):
    """C1, C2, C3 from the draw table: GHOST's write sequence, entry by entry."""
    ny, nz = g.ny, g.nz
    C = [g.czeros() for _ in range(3)]
    r, seed = draws(seed, 3 * len(table))
    for m, (loop, kx, ky, kz) in enumerate(table):
        dump = amp(kx * kx + ky * ky + kz * kz)
        for q in range(3):
            val = np.exp(1j * 2 * np.pi * r[3 * m + q]) * dump
            if loop == 4:
                C[q][kx, ky % ny, kz % nz] = 2 * val
            else:
                C[q][0, ky % ny, kz % nz] = val
                C[q][0, -ky % ny, -kz % nz] = np.conj(val)
    return C, seed


def test_forcing_draw_table_8():
    """initialfv_random on 8^3 with a shell that puts 2, 2, 6 and 18 modes in
    the four loops: the oracle's field must be the curl of the table's fill,
    normalised, and it must take exactly 3 x 28 draws.  Swapping the nesting of
    loop 3 or loop 4, or starting a loop one row late, reorders the draws."""
    g = O.Grid(8)
    seed, f0 = 8642, 0.9
    f, s_end = g.initialfv_random(1.0, 2.25, f0, seed)
    C, s_chk = fill_from_table(g, TABLE_8, seed, lambda kk: 1.0 / np.sqrt(kk))
    assert s_end == s_chk
    want = g.normalize(*g.curl(*C), f0)
    for k in range(3):
        assert np.max(np.abs(f[k] - want[k])) < 1e-13 * np.max(np.abs(want[k]))


def test_forcing_draw_table_4_nyquist_overwrites():
    """4^3 with kk2 in {4, 5}: every kx = 0 mode sits on a Nyquist line or plane,
    where GHOST's mirror writes land on the loop's own entries."""
    g = O.Grid(4)
    seed, f0 = 1357, 0.4
    f, s_end = g.initialfv_random(2.0, 2.2361, f0, seed)
    C, s_chk = fill_from_table(g, TABLE_4, seed, lambda kk: 1.0 / np.sqrt(kk))
    assert s_end == s_chk
    want = g.normalize(*g.curl(*C), f0)
    for k in range(3):
        assert np.max(np.abs(f[k] - want[k])) < 1e-13 * np.max(np.abs(want[k]))


def test_k4exp_draw_table():
    """initialv_k4exp fills every dealiased mode in the same order with modulus
    exp(-k^2/k0^2).  On 4^3 the dealias set kn2 <= 16/9 is the six unit modes:
    one in loop 1, one in loop 2, none in loop 3 and (1, 0, 0) in loop 4."""
    g = O.Grid(4)
    k0, u0, seed = 1.7, 1.0, 2468
    v, s_end = g.initialv_k4exp(k0, u0, seed)
    table = [(1, 0, 1, 0), (2, 0, 0, 1), (4, 1, 0, 0)]
    C, s_chk = fill_from_table(g, table, seed, lambda kk: np.exp(-kk / k0 ** 2))
    assert s_end == s_chk
    want = g.normalize(*g.curl(*C), u0)
    for k in range(3):
        assert np.max(np.abs(v[k] - want[k])) < 1e-13 * max(np.max(np.abs(want[k])), 1e-300)
    # on 8^3 the fill covers the dealias set kn2 <= 64/9 in the same four loops
    g8 = O.Grid(8)
    table8 = []
    for j in range(2, 6):
        ky = j - 1
        if 1 <= ky * ky <= 7:
            table8.append((1, 0, ky, 0))
    for k in range(2, 6):
        kz = k - 1
        if 1 <= kz * kz <= 7:
            table8.append((2, 0, 0, kz))
    for j in range(2, 9):
        for k in range(2, 6):
            ky = j - 1 if j - 1 <= 4 else j - 9
            kz = k - 1
            if ky * ky + kz * kz <= 7:
                table8.append((3, 0, ky, kz))
    for i in range(2, 6):
        for j in range(1, 9):
            for k in range(1, 9):
                kx, ky, kz = i - 1, (j - 1 if j - 1 <= 4 else j - 9), (k - 1 if k - 1 <= 4 else k - 9)
                if 1 <= kx * kx + ky * ky + kz * kz <= 7:
                    table8.append((4, kx, ky, kz))
    v8, s8 = g8.initialv_k4exp(k0, u0, seed)
    C8, s8_chk = fill_from_table(g8, table8, seed, lambda kk: np.exp(-kk / k0 ** 2))
    assert s8 == s8_chk
    want8 = g8.normalize(*g8.curl(*C8), u0)
    for k in range(3):
        assert np.max(np.abs(v8[k] - want8[k])) < 1e-13 * np.max(np.abs(want8[k]))


# phaseshift (pseudospec3D_hd.f90:2084-2150), pinned by real fields.  A real
# Fourier mode cos(k.x + theta) whose wavevector lies in the half space GHOST
# multiplies by cdump (kx > 0; kx = 0 and kz > 0; kx = kz = 0 and ky > 0) must
# come back as cos(k.x + theta + phi).  On the kx = 0 plane GHOST's loops visit
# the Nyquist entries (ky = ny/2 on kz = 0, kz = nz/2 for every ky) once with
# cdump and once with its conjugate, so those modes come back unchanged.

def real_mode(g, k, theta):
    nz, ny, nx = g.rshape
    z, y, x = np.meshgrid(*(2 * np.pi * np.arange(n) / n for n in (nz, ny, nx)), indexing="ij")
    return np.cos(k[0] * x + k[1] * y + k[2] * z + theta)


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (8, 6, 12)])
def test_phaseshift_of_real_modes(shape):
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz)
    phi = 0.8123
    cd = np.cos(phi) + 1j * np.sin(phi)
    shifted = [(2, -3, 1), (1, 0, 0), (3, 2, -2),         # kx > 0 planes (below kx = nx/2)
               (0, -2, 3), (0, 1, 1), (0, 2, -nz // 2 + 1),  # kx = 0, kz of either sign
               (0, 2, 0), (0, 1, 0),                     # kz = 0 line, ky > 0
               (0, 0, 2), (0, 0, 1),                     # ky = kz = 0 line, kz > 0
               (0, ny // 2, 1), (0, ny // 2, -2)]        # the ky Nyquist line off the kz planes
    nyquist = [(0, 2, nz // 2), (0, -1, nz // 2), (0, ny // 2, 0), (0, 0, nz // 2), (0, ny // 2, nz // 2)]
    rng = np.random.default_rng(11)
    for k in shifted + nyquist:
        theta = rng.uniform(0, 2 * np.pi)
        u = real_mode(g, k, theta)
        c = g.phaseshift(g.r2c(u), cd)
        # the kx = 0 plane stays Hermitian: c(0, -ky, -kz) = conj c(0, ky, kz)
        p = c[0]
        mirror = np.conj(p[(-np.arange(ny)) % ny][:, (-np.arange(nz)) % nz])
        assert np.max(np.abs(p - mirror)) < 1e-12 * nx * ny * nz, k
        back = g.c2r(c) / (nx * ny * nz)
        # the representative with the sign GHOST shifts by +phi
        kx, ky, kz = k
        kzr = kz if 2 * abs(kz) != nz else 0
        kyr = ky if 2 * abs(ky) != ny else 0
        pos = kx > 0 or (kx == 0 and (kzr > 0 or (kzr == 0 and kyr > 0)))
        if k in nyquist:
            want = u
        elif pos:
            want = real_mode(g, k, theta + phi)
        else:
            want = real_mode(g, k, theta - phi)
        assert np.max(np.abs(back - want)) < 1e-13, k
