"""Pins the oracle's Rytov reconstruction and phase unwrapping (oracle/rytov.py).

* Residues: a single phase vortex gives exactly one +1 residue, in the loop that holds
  it (its sign flips with the winding); a smooth map gives none; residues of a map are
  invariant to adding 2 pi multiples pixel by pixel.
* Raster unwrap: a wrapped smooth map (a plane plus a quadratic term) unwraps to the
  true phase up to the value at pixel (0, 0), the reference's starting pixel; a step map
  of +1 at one pixel lowers everything the row integration reaches after it by 2 pi.
* vector2scalarfield: a pure x-polarised field passes through; equal phase rotation of
  both polarisations factors out.
* Rytov: for a weak bead, measured fields from the CBS oracle reconstruct the bead's
  refractive index contrast to a few percent (this fixes the 4 pi, kz/i and pixel-size
  factors of :88 and :105-107); no scatterer gives n_m everywhere.
"""
import numpy as np
import pytest

from helpers import params, sphere_ri, rel
from oracle import rytov as R


def _vortex(n1=12, n2=10, c=(5.3, 4.6), sign=1):
    x, y = np.meshgrid(np.arange(n1), np.arange(n2), indexing='ij')
    return (sign * np.arctan2(y - c[1], x - c[0])).astype(np.float32)


def test_residues_of_vortex_and_smooth_map():
    signs = []
    for s in (1, -1):
        r = R.residues(_vortex(sign=s)[:, :, None])[:, :, 0]
        nz = np.argwhere(r != 0)
        assert len(nz) == 1
        assert tuple(nz[0]) == (5, 4)                  # the loop whose corners surround (5.3, 4.6)
        assert abs(r[5, 4]) == 1
        signs.append(r[5, 4])
    assert signs[0] == -signs[1]                       # opposite winding, opposite residue
    x, y = np.meshgrid(np.arange(16), np.arange(14), indexing='ij')
    smooth = np.angle(np.exp(1j * (0.9 * x + 0.7 * y + 0.01 * x * y))).astype(np.float32)
    assert not np.any(R.residues(smooth[:, :, None]))
    k = np.random.default_rng(0).integers(-3, 4, smooth.shape)
    np.testing.assert_array_equal(R.residues((smooth + 2 * np.pi * k).astype(np.float32)[:, :, None]),
                                  R.residues(smooth[:, :, None]))


def test_vortex_sign_convention_is_winding():
    # counter-clockwise increasing phase in the (dim 1, dim 2) plane gives the same sign for every vortex
    r1 = R.residues(_vortex(c=(3.5, 2.5))[:, :, None]).sum()
    r2 = R.residues(_vortex(c=(7.5, 6.5))[:, :, None]).sum()
    r3 = R.residues(_vortex(c=(7.5, 6.5), sign=-1)[:, :, None]).sum()
    assert r1 == r2 == -r3 and abs(r1) == 1


def test_raster_unwrap_recovers_smooth_phase():
    x, y = np.meshgrid(np.arange(20), np.arange(17), indexing='ij')
    true = (0.9 * x + 0.6 * y + 0.02 * x * y - 1.0)[:, :, None]
    wrapped = np.angle(np.exp(1j * true)).astype(np.float32)
    u = R.unwrap_residue_free(wrapped)
    expect = true - true[0, 0] + wrapped[0, 0]
    assert np.max(np.abs(u - expect)) < 2e-4
    # a +1 step at row pixel (4, 3) subtracts 2 pi from that pixel and the rest of its row
    s1 = np.zeros(wrapped.shape, np.int32)
    s1[4, 3, 0] = 1
    u2 = R.raster_unwrap(wrapped, s1, np.zeros((17, 1), np.int32))
    d = (u - u2)[:, :, 0] / (2 * np.pi)
    np.testing.assert_allclose(d[4:, 3], 1, atol=1e-5)
    np.testing.assert_allclose(np.delete(d, 3, axis=1), 0, atol=1e-5)
    np.testing.assert_allclose(d[:4, 3], 0, atol=1e-5)


def test_unwrap_refuses_residues():
    with pytest.raises(ValueError):
        R.unwrap_residue_free(_vortex()[:, :, None])


def test_vector2scalarfield_closed_forms():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((8, 8, 1, 3)) + 1j * rng.standard_normal((8, 8, 1, 3))
    b = rng.standard_normal((8, 8, 1, 3)) + 1j * rng.standard_normal((8, 8, 1, 3))
    s, t = R.vector2scalarfield(a, b)
    np.testing.assert_allclose(s, a[:, :, 0], atol=1e-15)
    a2 = np.concatenate([a, 0 * a], axis=2)
    b2 = np.concatenate([b, 0 * b], axis=2)
    s, t = R.vector2scalarfield(a2, b2)
    np.testing.assert_allclose(s, a[:, :, 0], atol=1e-12)
    np.testing.assert_allclose(t, b[:, :, 0], atol=1e-12)


def _measured(ri, n_s, radius, A, iters=30):
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    from inverse_solver_amd.phantoms import circle_illumination
    p = params(ri, iterations=iters, return_reflection=False, return_3D=False)
    o = Oracle(p)
    RI = sphere_ri(ri, radius, n_s=n_s)
    o.set_RI(RI)
    # illumination inside NA 0.9 so every angle's plane wave passes the NA 1.2 pupil
    u = circle_illumination(ri[:2], A, 0.9, 0.532, [0.1] * 3).astype(np.complex128)
    t, _, _ = o.solve(u)
    return p, RI, u, t


def test_rytov_reconstructs_weak_bead():
    ri = (24, 24, 25)
    p, RI, u, t = _measured(ri, 1.345, 0.6, 9)
    rp = dict(p, use_non_negativity=False)
    RI_rec, OR = R.rytov_solve(rp, u, t)
    d_true = RI - 1.336
    d_rec = np.real(RI_rec) - 1.336
    # the reconstruction is band limited (NA 1.2 and the missing cone): compare the bead's
    # integrated contrast and its correlation with the phantom
    ratio = d_rec.sum() / d_true.sum()
    corr = np.sum(d_rec * d_true) / np.sqrt(np.sum(d_rec ** 2) * np.sum(d_true ** 2))
    print('Rytov integrated contrast ratio %.3f  correlation %.3f' % (ratio, corr))
    assert 0.85 < ratio < 1.15
    assert corr > 0.8


def test_rytov_without_scatterer_is_background():
    ri = (16, 16, 17)
    p = params(ri, iterations=4)
    from inverse_solver_amd.phantoms import circle_illumination
    u = circle_illumination(ri[:2], 4, 0.9, 0.532, [0.1] * 3).astype(np.complex128)
    RI_rec, OR = R.rytov_solve(p, u, u.copy())
    # u / u is 1 up to rounding, and the single-precision unwrap adds up to an ulp of pi per
    # step (the reference's arithmetic): far below a 1 % amplitude change
    assert np.max(np.abs(RI_rec - 1.336)) < 1e-7
    assert np.max(np.abs(OR)) < 1e-4 * np.max(np.abs(R.rytov_solve(p, u, 1.01 * u)[1]))
