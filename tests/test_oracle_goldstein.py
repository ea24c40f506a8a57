"""Pins the oracle's Goldstein branch-cut unwrapping (oracle/goldstein.py).

* Closed forms: a wrapped vortex pair s * arg((p - z1) / (p - z2)) (p = d1 + i d2, vortex
  centres on loop centres) is unwrapped to that same function, whose discontinuity is the
  segment between the vortices, up to a constant, for pairs apart along d1 or d2 and both
  winding signs: that fixes the signs of residues, jumps and steps and where the branch
  (search + raster) is drawn.  A phase jumping by 3.5 rad across one edge makes a dipole
  in the two loops sharing it; its unwrap gives the phase back (a step when the edge is
  on a row path, none otherwise).  A diagonal pair differs
  from the closed form only between the segment and the rastered staircase.  A single vortex near a
  border is cut to the nearer border: along -d1 it is a step in the first column (step2),
  along -d2 a rastered column of step1; the closed form is arg of the rotated offset.
* Nearest lists: every entry is a residue of the same z at its true (truncated
  Euclidean) distance, ascending, ending in border entries (-1, border distance); where the
  box search can reach every ring, the list is the brute-force 8 nearest.  The reference's
  ring loop stops below max(sq - c, c) rings, one short of the far side for the last
  case; a 2 x 2 case grid shows the miss.
* The raster update x - step * 2 pi rounds once (a fused multiply-add), against exact
  rational arithmetic.
* Charge balance of the cut network on random maps: for every residue left by the dipole
  removal, sum of jumps of branches ending on it (as p1) minus those starting from it (as
  p2) equals its charge, and branches join residues of one z (or the border).
"""
import numpy as np
import pytest

from oracle import goldstein as G
from oracle import rytov as R


def _grid(n1, n2):
    x, y = np.meshgrid(np.arange(n1), np.arange(n2), indexing='ij')
    return x + 1j * y


def _wrap(phi):
    return np.angle(np.exp(1j * phi)).astype(np.float32)[:, :, None]


def _const_offset(u, expect):
    d = u[:, :, 0].astype(np.float64) - expect
    return np.max(np.abs(d - d[0, 0]))


@pytest.mark.parametrize('sign', [1, -1])
@pytest.mark.parametrize('z1,z2', [((19.5, 16.5), (19.5, 22.5)),     # far apart along d2
                                   ((19.5, 16.5), (27.5, 16.5)),     # far apart along d1
                                   ((12.5, 20.5), (20.5, 20.5))])
def test_vortex_pair_unwraps_to_cut_between(sign, z1, z2):
    p = _grid(40, 40)
    a, b = complex(*z1), complex(*z2)
    phi = sign * np.angle((p - a) / (p - b))
    w = _wrap(phi)
    u = G.unwrap(w)
    assert int(np.abs(R.residues(w)).sum()) == 2
    assert _const_offset(u, phi) < 1e-4


@pytest.mark.parametrize('sign', [1, -1])
@pytest.mark.parametrize('along', [1, 2])
def test_dipole_cut_crosses_the_jump_edge(sign, along):
    # a phase f that jumps by 3.5 rad across one edge (wrapped to 3.5 - 2 pi) has residues
    # of opposite sign in the two loops sharing that edge: a dipole.  Its cut crosses that
    # edge, and unwrapping around it gives f back: with a step in row 17 when the edge lies
    # on a row path (dipole along d2), with none when it does not (dipole along d1)
    f = np.zeros((40, 40))
    if along == 2:
        f[19, 17], f[20, 17] = -1.75 * sign, 1.75 * sign          # edge (19,17)-(20,17)
    else:
        f[20, 17], f[20, 18] = 1.75 * sign, -1.75 * sign          # edge (20,17)-(20,18)
    w = _wrap(f)
    res = R.residues(w)
    assert int(np.abs(res).sum()) == 2
    s1, s2, info = G.step_maps(w)
    assert not info['dipole_free'].any() and info['branches'].shape[1] == 0
    assert np.count_nonzero(s1) == (1 if along == 2 else 0) and not s2.any()
    assert _const_offset(R.raster_unwrap(w, s1, s2), f) < 1e-5


@pytest.mark.parametrize('sign', [1, -1])
def test_diagonal_branch_is_a_staircase_along_the_segment(sign):
    # the raster crosses each row path at the start of its step along d1, so the unwrapped
    # phase differs from the exact-segment closed form by 2 pi only on pixels next to the
    # segment, between it and the staircase
    p = _grid(40, 40)
    a, b = complex(15.5, 18.5), complex(23.5, 21.5)
    phi = sign * np.angle((p - a) / (p - b))
    u = G.unwrap(_wrap(phi))[:, :, 0].astype(np.float64)
    k = (u - phi - (u[0, 0] - phi[0, 0])) / (2 * np.pi)
    assert np.max(np.abs(k - np.round(k))) < 1e-4
    off = np.argwhere(np.round(k) != 0)
    assert 0 < len(off) <= 8
    t = np.clip(((off[:, 0] - a.real) * (b - a).real + (off[:, 1] - a.imag) * (b - a).imag) / abs(b - a) ** 2, 0, 1)
    dist = np.abs(off[:, 0] + 1j * off[:, 1] - (a + t * (b - a)))
    assert np.max(dist) < 1.0


@pytest.mark.parametrize('sign', [1, -1])
def test_vortex_near_border_cuts_to_border(sign):
    p = _grid(40, 40)
    # closer to the d1 = 0 border: cut along -d1, a step of the first column
    a = complex(3.5, 20.5)
    phi = sign * np.angle(p - a)                      # discontinuous on the -d1 ray
    s1, s2, info = G.step_maps(_wrap(phi))
    assert not s1.any() and np.count_nonzero(s2) == 1 and s2[21, 0] != 0
    assert _const_offset(R.raster_unwrap(_wrap(phi), s1, s2), phi) < 1e-4
    # closer to the d2 = 0 border: cut along -d2, rastered into step1 at d1 = 21, d2 = 0..3
    a = complex(20.5, 3.5)
    phi = sign * np.angle(-1j * (p - a))              # discontinuous on the -d2 ray
    s1, s2, info = G.step_maps(_wrap(phi))
    assert not s2.any()
    assert sorted(zip(*np.nonzero(s1[:, :, 0]))) == [(21, j) for j in range(4)]
    assert _const_offset(R.raster_unwrap(_wrap(phi), s1, s2), phi) < 1e-4


def _random_maps():
    rng = np.random.default_rng(1)
    maps = [rng.uniform(-np.pi, np.pi, s).astype(np.float32) for s in ((40, 40, 2), (33, 29, 1), (20, 50, 3), (9, 9, 2))]
    p = _grid(100, 100)
    c = rng.uniform(5, 95, (30, 2))
    phi = 0.005 * p.real * p.imag + sum(np.angle(p - complex(*cc)) * s for cc, s in zip(c, rng.choice([-1, 1], 30)))
    maps.append(_wrap(phi + 0.6 * rng.standard_normal(phi.shape)))
    return maps


def test_nearest_lists_against_brute_force():
    exhaustive = 0
    for w in _random_maps():
        _, _, info = G.step_maps(w)
        o, sh, sq_z = info['ordered'], info['shortest'], info['square']
        m1, m2, _ = info['dipole_free'].shape
        for i in range(o.shape[1]):
            z, x0, y0 = o[5, i], int(o[3, i]), int(o[4, i])
            others = np.flatnonzero((o[5] == z) & (np.arange(o.shape[1]) != i))
            d = np.sqrt(((o[3, others] - x0) ** 2 + (o[4, others] - y0) ** 2).astype(np.float32)).astype(int)
            border = min(x0, y0, m1 - x0, m2 - y0)
            got_d = [int(v) for v in sh[1, :, i]]
            assert got_d == sorted(got_d)
            for k in range(8):
                if sh[0, k, i] >= 0:
                    j = int(sh[0, k, i])
                    assert o[5, j] == z and j != i and d[list(others).index(j)] == got_d[k]
                else:
                    assert got_d[k] == border
            sq = int(sq_z[z])
            c1, c2 = int(o[2, i]) % sq, int(o[2, i]) // sq
            if max(sq - c1, sq - c2, c1, c2) >= sq:            # every ring reachable
                exhaustive += 1
                want = sorted(int(v) for v in d if v < border)[:8]
                assert got_d == want + [border] * (8 - len(want))
    assert exhaustive > 400


def test_nearest_ring_bound_quirk():
    # 2 x 2 cases (sq = ceil(sqrt((N + 1) / 16)) = 2 for N = 17..48): a residue in case
    # (1, 1) searches only its own case, so a residue 2 away in case (0, 1) is missed
    res = np.zeros((60, 60, 1), np.float32)
    res[31, 31, 0] = 1                       # case (1, 1) (case side 30)
    res[29, 31, 0] = -1                      # case (0, 1), distance 2
    k = 0
    for i in range(2, 58, 6):                # fill to N = 20, far away, alternating charges
        for j in (2, 57):
            if k < 18:
                res[i, j, 0] = 1 if k % 2 else -1
                k += 1
    o, sq = G.order_residues(res)
    assert list(sq) == [2]
    table, zs = G.lookup(o, sq)
    i = int(np.flatnonzero((o[3] == 31) & (o[4] == 31))[0])
    cur = G.nearest(o, table, zs, sq, res.shape, i)
    ids = [cur[2 * k] for k in range(8)]
    j = int(np.flatnonzero((o[3] == 29) & (o[4] == 31))[0])
    assert j not in ids and cur[1] > 2


def test_branch_network_balances_every_charge():
    for w in _random_maps():
        _, _, info = G.step_maps(w)
        res, br = info['dipole_free'], info['branches']
        net = {}
        for b in range(br.shape[1]):
            p1x, p1y, p2x, p2y, z, jump = (int(v) for v in br[:, b])
            assert res[p2x, p2y, z] != 0
            if p1x >= 0:
                assert res[p1x, p1y, z] != 0
                net[(p1x, p1y, z)] = net.get((p1x, p1y, z), 0) + jump
            net[(p2x, p2y, z)] = net.get((p2x, p2y, z), 0) - jump
        for pos in zip(*np.nonzero(res)):
            assert net.get(tuple(int(v) for v in pos), 0) == res[pos]


def test_dipole_removal_hand_cases():
    res = np.zeros((8, 9, 2), np.float32)
    res[3, 4, 0], res[3, 3, 0] = 1, -1       # + with - at d2 - 1: direction 1, step +1
    res[5, 2, 1], res[5, 3, 1] = 1, -1       # + with - at d2 + 1: direction 3, step -1
    res[1, 6, 1], res[0, 6, 1] = 1, -1       # along d1: no step
    out, s1, s2 = G.remove_dipoles(res)
    assert not out.any() and not s2.any()
    assert s1[4, 4, 0] == 1 and s1[6, 3, 1] == -1 and np.count_nonzero(s1) == 2
    # a positive between two negatives pairs with the d1 - 1 one (direction 0 runs first)
    res = np.zeros((8, 9, 1), np.float32)
    res[3, 4, 0], res[2, 4, 0], res[3, 5, 0] = 1, -1, -1
    out, s1, _ = G.remove_dipoles(res)
    assert out[2, 4, 0] == 0 and out[3, 5, 0] == -1 and not s1.any()


def test_residue_free_map_needs_no_cuts():
    p = _grid(20, 17)
    phi = 0.9 * p.real + 0.6 * p.imag + 0.02 * p.real * p.imag
    s1, s2, info = G.step_maps(_wrap(phi))
    assert not s1.any() and not s2.any() and info['branches'].shape[1] == 0
    np.testing.assert_array_equal(G.unwrap(_wrap(phi)), R.unwrap_residue_free(_wrap(phi)))


def test_integration_step_rounds_once():
    # the raster update x - step * 2 pi is a fused multiply-add: compare with the exact
    # rational value rounded to the nearest float32 (ties to even)
    from fractions import Fraction
    rng = np.random.default_rng(3)
    tp = np.float32(2.0) * R.PI32
    for _ in range(2000):
        x = np.float32(rng.uniform(-400, 400))
        s = int(rng.integers(-9, 10))
        exact = Fraction(float(x)) - s * Fraction(float(tp))
        f = np.float32(float(exact))
        cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
        errs = [abs(Fraction(float(c)) - exact) for c in cands]
        best = min(errs)
        want = [c for c, e in zip(cands, errs) if e == best]
        want = want[0] if len(want) == 1 else [c for c in want if (c.view(np.int32) & 1) == 0][0]
        assert R._fma_step(x, s, tp) == want
