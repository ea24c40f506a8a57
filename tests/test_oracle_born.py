"""Pins the oracle's Born loop (oracle/cbs.py) without reusing its formulas.

* Dense brute force: on a tiny non-cubic grid the shifted Green's operators are
  built from their closed form -- a direct sum over the explicitly shifted
  frequency set S+ = {m + 1/4} (S- = -S+), no FFT and no phase ramp -- and the
  two-chain series (F1 = A V G-(p0), F2 = A V G+(p0), f <- M_s f with f += F1
  after the third step) is evaluated with matrices and compared with
  born_loop() at every iteration count.
* G- = G+^T: follows from S- = -S+ and the evenness of the Green's function;
  it fixes how the flipped Greenp pairs with the conjugated ramp.
* The attenuation mask, the ramp, eps and Bornmax against their closed forms.
* Linearity in the incident field.
"""
import numpy as np
import pytest

from helpers import params, sphere_ri, random_field, rel
from oracle.cbs import ForwardSolverConvergentBorn as Oracle


def _tiny(ri=(4, 3, 5), iters=6, **kw):
    p = params(ri, bt_px_target=1, iterations=iters, **kw)
    rng = np.random.default_rng(0)
    RI = 1.336 + 0.08 * rng.random(ri)
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    return o, RI


def _freq_set(n, sign):
    """Shifted frequency indices: S+ = {m + 1/4 : m = -floor(n/2) .. n - floor(n/2) - 1}, S- = -S+."""
    m = np.arange(n) - n // 2
    return sign * (m + 0.25)


def _dense_green(o, sign):
    nx, ny, nz = o.V.shape
    res = o.utility_border['fourier_res']
    km2 = (2 * np.pi * o.utility_border['k0_nm']) ** 2
    eps = o.eps_imag
    kx, ky, kz = (_freq_set(n, sign) for n in (nx, ny, nz))
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing='ij')
    g = 1.0 / (4 * np.pi ** 2 * ((KX * res[0]) ** 2 + (KY * res[1]) ** 2 + (KZ * res[2]) ** 2) - km2 - 1j * eps)
    d = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij'), -1).reshape(-1, 3)
    N = nx * ny * nz
    kap = np.stack([KX.ravel(), KY.ravel(), KZ.ravel()], -1)            # (N, 3)
    nn = np.array([nx, ny, nz], float)
    # K(d, d') = (1/N) sum_kappa g(kappa) exp(2 pi i kappa . (d - d') / n)
    ph_d = np.exp(2j * np.pi * (d / nn) @ kap.T)                        # (N, N) over (d, kappa)
    return (ph_d * g.ravel()[None, :]) @ ph_d.conj().T / N


def test_green_minus_is_transpose_of_green_plus():
    o, _ = _tiny()
    Gp = _dense_green(o, +1)
    Gm = _dense_green(o, -1)
    np.testing.assert_allclose(Gm, Gp.T, rtol=0, atol=1e-12 * np.abs(Gp).max())


@pytest.mark.parametrize('ri', [(4, 3, 5), (4, 4, 4)])
def test_born_loop_equals_dense_series(ri):
    iters = 7
    o, RI0 = _tiny(ri, iters)
    shp = o.V.shape
    Gp = _dense_green(o, +1)
    Gm = _dense_green(o, -1)
    V = o.V.ravel(order='C')
    # closed-form mask: min(dist / (b - 0.5), 1) per axis, dist to the nearest edge
    b = o.boundary_thickness_pixel
    A = np.ones(shp)
    for j, n in enumerate(shp):
        dist = np.minimum(np.arange(n), n - 1 - np.arange(n)).astype(float)
        fac = np.minimum(dist / (b[j] - 0.5), 1.0) if b[j] > 0 else np.ones(n)
        s = [1, 1, 1]
        s[j] = -1
        A = A * fac.reshape(s)
    np.testing.assert_allclose(o.attenuation_mask * np.ones(shp), A, atol=1e-15)
    A = A.ravel(order='C')
    eps = o.eps_imag
    gam = 1j * V / eps
    R = o.ROI
    roi = (slice(R[0] - 1, R[1]), slice(R[2] - 1, R[3]), slice(R[4] - 1, R[5]))
    u_in = random_field(tuple(o.expected_RI_size), 3)
    S = np.zeros(shp, complex)
    S[roi] = (o.V[roi] + 1j * eps) * u_in
    p0 = ((1j / eps) * S / 2).ravel(order='C')
    F1 = A * V * (Gm @ p0)
    F2 = A * V * (Gp @ p0)
    M = {+1: lambda x: A * (x - gam * x + V * (Gp @ (gam * x))),
         -1: lambda x: A * (x - gam * x + V * (Gm @ (gam * x)))}
    for it in range(1, iters + 1):
        Field = F1.copy()
        if it >= 2:
            Field = Field + F2
            f = F2.copy()
        for jj in range(3, it + 1):
            s = -1 if jj % 2 else +1
            f = M[s](f)
            Field = Field + f
            if jj == 3:
                f = f + F1
        expect = Field.reshape(shp, order='C')[roi] + u_in
        o.set_RI(RI0)          # resets the persistent flip state
        o.Bornmax = it
        got = o.born_loop(u_in)
        assert rel(got, expect) < 1e-12, (ri, it)


def test_linear_in_incident_field():
    o, _ = _tiny((4, 4, 4), 5)
    R = o.ROI
    roi = (slice(R[0] - 1, R[1]), slice(R[2] - 1, R[3]), slice(R[4] - 1, R[5]))
    u1 = random_field((4, 4, 4), 1)
    u2 = random_field((4, 4, 4), 2)
    a, c = 0.7 - 0.2j, -1.3 + 0.5j
    RI = o.RI[roi]
    outs = []
    for u in (u1, u2, a * u1 + c * u2):
        o.set_RI(RI)            # resets the persistent flip state between calls
        outs.append(o.born_loop(u) - u)
    assert rel(outs[2], a * outs[0] + c * outs[1]) < 1e-12


def test_homogeneous_ri_is_rejected():
    p = params((4, 4, 4), bt_px_target=1, iterations=2)
    o = Oracle(p, io_kernels=False)
    with pytest.raises(ValueError):
        o.set_RI(np.full((4, 4, 4), 1.336))


def test_eps_ramp_and_bornmax_closed_forms():
    ri = (6, 5, 4)
    p = params(ri, bt_px_target=2, iterations=-2)
    RI = sphere_ri(ri, 0.2, n_s=1.40)
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    k = 2 * np.pi * 1.336 / 0.532
    vt_max = k ** 2 * (1.40 ** 2 / 1.336 ** 2 - 1)
    assert abs(o.eps_imag - 1.01 * vt_max) < 1e-9 * vt_max
    n = np.array(o.V.shape)
    step = 2 * k / o.eps_imag
    bm = int(np.ceil(np.linalg.norm(n * 0.1) / step / 2 + 1) * 2) * 2
    assert o.Bornmax == bm
    # ramp: exp(-i pi (d - floor(n/2)) / (2 n)) per axis, unit at the centre
    for j in range(3):
        d = np.arange(n[j])
        expect = np.exp(-1j * np.pi * (d - n[j] // 2) / (2 * n[j]))
        idx = [n[0] // 2, n[1] // 2, n[2] // 2]
        line = []
        for dd in d:
            ii = list(idx)
            ii[j] = dd
            line.append(o.phase_ramp[tuple(min(ii[q], o.phase_ramp.shape[q] - 1) for q in range(3))])
        np.testing.assert_allclose(np.array(line), expect, atol=1e-14)


def test_convergence_stop_truncates_the_series():
    """The oracle's FP64 convergence test (an option the reference does not have):
    with tol, the loop stops after the first iteration n >= 3 whose term norm is at
    most tol times the first-order field's norm, and returns exactly the fixed-length
    series of n iterations; the flip state then stands where a full Bornmax run leaves
    it (the next angle starts on the same shift as without the test).  No boundary
    (ROI = the whole grid), so the first-order field's norm is pinned by the returned
    2-iteration series."""
    ri = (8, 7, 9)
    M = 15
    p = params(ri, bt_px_target=0, iterations=M)
    RI = sphere_ri(ri, 0.25, n_s=1.40)
    u = random_field(ri, 3)
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    assert list(o.V.shape) == list(ri)
    inc, st = [], {}
    o.born_loop(u, increments=inc, tol=0.0, stats=st)
    assert st['iterations'] == M and len(inc) == M
    o2 = Oracle(dict(p, iterations_number=2), io_kernels=False)
    o2.set_RI(RI)
    F12 = o2.born_loop(u) - u
    np.testing.assert_allclose(st['fref'], np.sum(np.abs(F12) ** 2), rtol=1e-12)
    ratio = np.sqrt(np.array(inc) / st['fref'])
    for tol_n in (6, 9):
        tol = float(np.sqrt(ratio[tol_n - 1] * ratio[tol_n - 2]))
        want = next(n for n in range(3, M + 1) if ratio[n - 1] <= tol)
        o = Oracle(p, io_kernels=False)
        o.set_RI(RI)
        st2 = {}
        got = o.born_loop(u, tol=tol, stats=st2)
        assert st2['iterations'] == want < M
        on = Oracle(dict(p, iterations_number=want), io_kernels=False)
        on.set_RI(RI)
        np.testing.assert_array_equal(got, on.born_loop(u))
        # the next angle: the same as after a full-length first angle
        ofull = Oracle(p, io_kernels=False)
        ofull.set_RI(RI)
        ofull.born_loop(u)
        np.testing.assert_array_equal(o.born_loop(u), ofull.born_loop(u))


def test_snapshots_equal_shorter_series():
    """born_loop(snapshots={n: ...}) records after iteration n exactly what a series of
    Bornmax = n returns (used to state the error at several depths from one run)."""
    ri = (8, 7, 9)
    p = params(ri, bt_px_target=2, iterations=12)
    RI = sphere_ri(ri, 0.25, n_s=1.40)
    u = random_field(ri, 4)
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    snaps = {4: None, 7: None, 12: None}
    full = o.born_loop(u, snapshots=snaps)
    np.testing.assert_array_equal(snaps[12], full)
    for n in (4, 7):
        on = Oracle(dict(p, iterations_number=n), io_kernels=False)
        on.set_RI(RI)
        np.testing.assert_array_equal(snaps[n], on.born_loop(u))


def test_reciprocity_dot_product():
    """Adjoint (dot-product) pin of the series behind the MULTI gradient
    (Codes/BACKWARD_SOLVER_MULTI.m:30-54): the solution operator (I - G V)^-1 G of the
    Helmholtz equation is symmetric because G is, so the scattered field S x =
    Field[ROI] of an incident x obeys y^T Vt (S x) = (S y)^T Vt x (bilinear, no
    conjugate).  The converged series keeps it to its truncation and boundary
    error; the sesquilinear pairing (S y)^H Vt x, which an adjoint built with a
    conjugate too many would satisfy, does not hold."""
    ri = (10, 9, 8)
    p = params(ri, bt_px_target=0, iterations=300)
    RI = sphere_ri(ri, 0.3, n_s=1.40)
    x, y = random_field(ri, 1), random_field(ri, 2)
    outs = []
    for f in (x, y):
        o = Oracle(p, io_kernels=False)
        o.set_RI(RI)
        outs.append(o.born_loop(f) - f)
    R = o.ROI
    roi = tuple(slice(R[2 * j] - 1, R[2 * j + 1]) for j in range(3))
    Vt = o.V[roi] + 1j * o.eps_imag
    Sx, Sy = outs
    a = np.sum(y * Vt * Sx)
    b = np.sum(Sy * Vt * x)
    c = np.sum(np.conj(Sy) * Vt * x)
    assert abs(a - b) < 2e-3 * abs(a)
    assert abs(a - c) > 0.3 * abs(a)
