"""randu, the random-phase forcing and the k^4 exp initial condition of the
oracle, pinned by the textbook Park-Miller generator (computed here with
64-bit products, not Schrage's factorisation) and by a hand-derived
known-answer test of the draw order of examples/initialfv.f90_random."""
import numpy as np

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
