"""Pins the vector (dyadic Green's function, pole_num = 3) branch of the oracle.

* Dense brute force: on a tiny grid the dyadic operators G_s (I - c r r^T) are
  built from their closed form -- a direct sum over the explicitly shifted
  frequency sets, r = kappa / k0 per frequency, c = k^2 / (k^2 + i eps), no FFT
  and no phase ramp -- and the two-chain series on 3-component fields is compared
  with born_loop() at every iteration count.
* Normal incidence on a z-stratified potential with periodic xy: every field
  stays at k_xy = 0, where the transverse x component does not couple to the
  others, so the vector solve's x polarisation equals the scalar solve and the
  y polarisation stays zero (through the full 2D-in / 2D-out path).
* Polarisation transforms: the Ewald tangent against its closed form
  (K3 cos phi, K3 sin phi, -rho), transversality of the 3D field, energy
  preservation, and the 2D -> 3D -> 2D round trip (= NA filter).
"""
import numpy as np
import pytest

from helpers import params, random_field, rel
from oracle.cbs import ForwardSolverConvergentBorn as Oracle
from oracle.optics import polarisation_utility


def _tiny_vec(ri=(4, 3, 5), iters=6):
    p = params(ri, bt_px_target=1, iterations=iters, vector_simulation=True)
    rng = np.random.default_rng(0)
    RI = 1.336 + 0.08 * rng.random(ri)
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    return o, RI


def _dense_dyadic(o, sign):
    nx, ny, nz = o.V.shape
    res = o.utility_border['fourier_res']
    k0 = o.utility_border['k0_nm']
    km2 = (2 * np.pi * k0) ** 2
    eps = o.eps_imag
    c = km2 / (km2 + 1j * eps)
    sets = [sign * ((np.arange(n) - n // 2) + 0.25) for n in (nx, ny, nz)]
    KX, KY, KZ = np.meshgrid(*sets, indexing='ij')
    fx, fy, fz = KX * res[0], KY * res[1], KZ * res[2]                 # shifted frequencies (cycles/um)
    g = (1.0 / (4 * np.pi ** 2 * (fx ** 2 + fy ** 2 + fz ** 2) - km2 - 1j * eps)).ravel()
    r = np.stack([fx.ravel(), fy.ravel(), fz.ravel()], -1) / k0        # (N, 3)
    d = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij'), -1).reshape(-1, 3)
    kap = np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], -1)
    nn = np.array([nx, ny, nz], float)
    N = nx * ny * nz
    ph = np.exp(2j * np.pi * (d / nn) @ kap.T)                          # (N, N)
    G = np.zeros((3 * N, 3 * N), complex)
    for a in range(3):
        for b in range(3):
            w = g * ((1.0 if a == b else 0.0) - c * r[:, a] * r[:, b])
            G[a * N:(a + 1) * N, b * N:(b + 1) * N] = (ph * w[None, :]) @ ph.conj().T / N
    return G


@pytest.mark.parametrize('ri', [(4, 3, 5), (4, 4, 4)])
def test_vector_born_loop_equals_dense_dyadic_series(ri):
    iters = 6
    o, RI0 = _tiny_vec(ri, iters)
    shp = o.V.shape
    N = int(np.prod(shp))
    Gp = _dense_dyadic(o, +1)
    Gm = _dense_dyadic(o, -1)
    V3 = np.tile(o.V.ravel(order='C'), 3)
    b = o.boundary_thickness_pixel
    A = np.ones(shp)
    for j, n in enumerate(shp):
        dist = np.minimum(np.arange(n), n - 1 - np.arange(n)).astype(float)
        fac = np.minimum(dist / (b[j] - 0.5), 1.0) if b[j] > 0 else np.ones(n)
        s = [1, 1, 1]
        s[j] = -1
        A = A * fac.reshape(s)
    A3 = np.tile(A.ravel(order='C'), 3)
    eps = o.eps_imag
    gam = 1j * V3 / eps
    R = o.ROI
    roi = (slice(R[0] - 1, R[1]), slice(R[2] - 1, R[3]), slice(R[4] - 1, R[5]))
    u_in = random_field(tuple(o.expected_RI_size) + (3,), 5)
    S = np.zeros(shp + (3,), complex)
    S[roi] = (o.V[roi] + 1j * eps)[..., None] * u_in
    p0 = np.concatenate([((1j / eps) * S[..., j] / 2).ravel(order='C') for j in range(3)])
    F1 = A3 * V3 * (Gm @ p0)
    F2 = A3 * V3 * (Gp @ p0)
    M = {+1: lambda x: A3 * (x - gam * x + V3 * (Gp @ (gam * x))),
         -1: lambda x: A3 * (x - gam * x + V3 * (Gm @ (gam * x)))}
    for it in range(1, iters + 1):
        Field = F1.copy()
        if it >= 2:
            Field = Field + F2
            f = F2.copy()
        for jj in range(3, it + 1):
            sgn = -1 if jj % 2 else +1
            f = M[sgn](f)
            Field = Field + f
            if jj == 3:
                f = f + F1
        expect = np.stack([Field[j * N:(j + 1) * N].reshape(shp, order='C')[roi] for j in range(3)], -1) + u_in
        o.set_RI(RI0)
        o.Bornmax = it
        got = o.born_loop(u_in)
        assert rel(got, expect) < 1e-12, (ri, it)


def test_vector_equals_scalar_for_normal_incidence_on_layers():
    ri = (8, 8, 13)
    z = np.arange(ri[2])
    layer = 1.336 + 0.05 * np.exp(-((z - 6) / 2.5) ** 2)
    RI = np.broadcast_to(layer, ri).copy()
    kw = dict(bt_px_target=3, iterations=10)
    pv = params(ri, vector_simulation=True, **kw)
    ps = params(ri, vector_simulation=False, **kw)
    for p in (pv, ps):
        bt = p['boundary_thickness']
        p['boundary_thickness'] = [0, 0, bt[2]]          # periodic in x and y
    ov, os_ = Oracle(pv), Oracle(ps)
    ov.set_RI(RI)
    os_.set_RI(RI)
    assert ov.cyclic_boundary_xy
    amp = 0.8 - 0.3j
    uv = np.zeros((8, 8, 2, 1), complex)
    uv[:, :, 0, 0] = amp                                  # x-polarised normal-incidence plane wave
    us = np.full((8, 8, 1, 1), amp, complex)
    tv, rv, fv = ov.solve(uv)
    ts, rs, fs = os_.solve(us)
    assert rel(tv[:, :, 0, 0], ts[:, :, 0, 0]) < 1e-12
    assert rel(rv[:, :, 0, 0], rs[:, :, 0, 0]) < 1e-12
    assert rel(fv[:, :, :, 0, 0], fs[:, :, :, 0, 0]) < 1e-12
    assert np.abs(tv[:, :, 1, 0]).max() < 1e-12 * np.abs(tv).max()
    assert np.abs(fv[:, :, :, 1:, 0]).max() < 1e-12 * np.abs(fv).max()
    # and the layer does scatter (the check is not vacuous)
    assert rel(ts[:, :, 0, 0], np.full((8, 8), amp)) > 1e-3


def test_ewald_tangent_closed_form_and_transversality():
    n, lam, nm, dx = 16, 0.532, 1.336, [0.1, 0.1]
    R2, P2, T3, Km = polarisation_utility((n, n), nm, lam, dx)
    k = 2 * np.pi * nm / lam
    K1 = (2 * np.pi / (dx[0] * n) / k) * (np.arange(n) - n // 2)
    rng = np.random.default_rng(1)
    for _ in range(20):
        i, j = rng.integers(0, n, 2)
        rho = np.hypot(K1[i], K1[j])
        if rho >= 1 or rho == 0:
            continue
        k3 = np.sqrt(1 - rho ** 2)
        phi = np.arctan2(K1[j], K1[i])
        np.testing.assert_allclose(T3[i, j], [k3 * np.cos(phi), k3 * np.sin(phi), -rho], atol=1e-12)
        np.testing.assert_allclose(R2[i, j], [np.cos(phi), np.sin(phi)], atol=1e-12)
        np.testing.assert_allclose(P2[i, j], [np.sin(phi), -np.cos(phi)], atol=1e-12)
        assert abs(np.dot(T3[i, j], [K1[i], K1[j], k3])) < 1e-12
    assert Km[n // 2, n // 2] and not Km[0, 0]
    np.testing.assert_allclose(T3[n // 2, n // 2], [1, 0, 0], atol=1e-15)


def test_polarisation_round_trip_and_energy():
    ri = (16, 16, 5)
    o = Oracle(params(ri, iterations=2, vector_simulation=True), io_kernels=False)
    F = random_field((16, 16, 2, 3), 7)
    F3 = o.transform_field_3D(F)
    na = o.utility['NA_circle'][:, :, 0]
    cos_t = o.utility['cos_theta'][:, :, 0]
    # 3D field is transverse to the propagation direction at every NA point
    nm, lam = 1.336, 0.532
    k = 2 * np.pi * nm / lam
    K1 = (2 * np.pi / (0.1 * 16) / k) * (np.arange(16) - 8)
    KX, KY = np.meshgrid(K1, K1, indexing='ij')
    K3 = np.sqrt(np.maximum(1 - KX ** 2 - KY ** 2, 0))
    dot = F3[:, :, 0] * KX[..., None] + F3[:, :, 1] * KY[..., None] + F3[:, :, 2] * K3[..., None]
    assert np.abs(dot).max() < 1e-12 * np.abs(F3).max()
    # the basis change preserves |E|^2 (after the Abbe factor)
    w = np.where(na, 1 / np.sqrt(np.where(na, cos_t, 1)), 0)
    e_in = np.sum(np.abs(F) ** 2, axis=2) * (w ** 2)[..., None]
    np.testing.assert_allclose(np.sum(np.abs(F3) ** 2, axis=2), e_in, rtol=1e-12, atol=1e-14)
    # 2D -> 3D -> 2D returns the NA-filtered input (the Abbe factors cancel)
    back = o.transform_field_2D(F3)
    np.testing.assert_allclose(back, F * na[:, :, None, None], atol=1e-12)
    # reflection: mirror image in z of the transmission basis
    backr = o.transform_field_2D_reflection(F3 * np.array([1, 1, -1])[None, None, :, None])
    np.testing.assert_allclose(backr, F * na[:, :, None, None], atol=1e-12)


def test_vector_input_validation():
    o = Oracle(params((4, 4, 4), bt_px_target=1, iterations=2, vector_simulation=True), io_kernels=False)
    o.set_RI(np.full((4, 4, 4), 1.336) + 0.01 * random_field((4, 4, 4), 2).real)
    with pytest.raises(ValueError):
        o.solve(np.ones((4, 4, 1, 1)))
    with pytest.raises(ValueError):
        polarisation_utility((8, 6), 1.336, 0.532, [0.1, 0.1])
