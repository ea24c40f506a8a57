"""TV proximal step and the MULTI gradient.

CPU: the product TV prox equals the oracle restatement of TV_FISTA_inner_v2;
the oracle prox is pinned by its special cases (lambda -> 0 returns the
projected input, a constant is a fixed point, the prox objective decreases) and
the oracle gradient by the sign/consistency of finite-difference directional
derivatives of the oracle's own misfit.
GPU: gradient and two FISTA outer iterations against the oracle."""
import numpy as np
import pytest

from helpers import params, rel, sphere_ri


def test_tv_product_matches_oracle():
    from inverse_solver_amd.tv import tv_fista_inner
    from oracle.backward import tv_fista_inner_v2
    rng = np.random.default_rng(0)
    x = rng.standard_normal((9, 8, 7)) * 0.1
    for args in [(0.05, -np.inf, np.inf, 0, False), (0.02, -0.02, 0.05, 0, True)]:
        a = tv_fista_inner(x, *args, False, 30)
        b = tv_fista_inner_v2(x, *args, False, 30)
        assert rel(a, b) < 1e-12
    xc = x + 1j * rng.standard_normal(x.shape) * 0.01
    a = tv_fista_inner(xc, 0.03, -np.inf, np.inf, 0.004, False, False, 20)
    b = tv_fista_inner_v2(xc, 0.03, -np.inf, np.inf, 0.004, False, False, 20)
    assert np.iscomplexobj(a) and rel(a, b) < 1e-12
    assert np.all(a.imag >= 0) and np.all(a.imag <= 0.004 + 1e-15)


def test_tv_prox_special_cases():
    from oracle.backward import tv_fista_inner_v2, tv_value
    rng = np.random.default_rng(1)
    x = rng.standard_normal((8, 8, 8)) * 0.1
    # lambda -> 0: the prox returns the (projected) input
    y = tv_fista_inner_v2(x, 1e-9, -np.inf, np.inf, 0, False, False, 20)
    assert rel(y, x) < 1e-6
    y = tv_fista_inner_v2(x, 1e-9, -0.05, 0.05, 0, True, False, 20)
    assert rel(y, np.clip(x, -0.05, 0.05)) < 1e-6
    # a constant is a fixed point
    c = np.full((6, 5, 7), 0.3)
    np.testing.assert_allclose(tv_fista_inner_v2(c, 0.1, -np.inf, np.inf, 0, False, False, 50), c, atol=1e-12)
    # the prox objective 1/2 |y - x|^2 + lam TV(y) is below its value at y = x
    lam = 0.05
    y = tv_fista_inner_v2(x, lam, -np.inf, np.inf, 0, False, False, 200)
    obj = lambda v: 0.5 * np.sum((v - x) ** 2) + lam * tv_value(v)
    assert obj(y) < 0.5 * obj(x)
    assert tv_value(y) < tv_value(x)


def test_oracle_gradient_matches_finite_differences_in_direction():
    """The MULTI gradient (-i mean(conj(u) w), w the conj-flipped adjoint solve)
    is the reciprocity gradient of the transmitted-field misfit, preconditioned
    by a positive, frequency-dependent weight: the adjoint injects the residual
    through the refocusing kernel while the forward projection uses the Green's
    kernel (extra 1/(4 pi i k3)), and the Abbe-sine factors are not mirrored.
    So only the sign and the rough size of directional derivatives can be
    pinned: along smooth real perturbations (kept off the potential maximum so
    eps = 1.01 max|Vt| is unchanged) finite differences of the oracle's own
    misfit and <Re grad, d> must agree in sign, within a bounded ratio spread."""
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    from oracle.backward import get_gradiant_static
    from inverse_solver_amd.phantoms import circle_illumination
    ri = (10, 10, 9)
    p = params(ri, iterations=10)
    o = Oracle(p)
    RI_true = sphere_ri(ri, 0.3, n_s=1.37)
    u = circle_illumination(ri[:2], 3, 1.2, 0.532, [0.1] * 3).astype(np.complex128)
    o.set_RI(RI_true)
    meas, _, _ = o.solve(u)
    RI0 = sphere_ri(ri, 0.25, n_s=1.355)
    n_m = 1.336
    pot0 = RI0 ** 2 / n_m ** 2 - 1
    g, _ = get_gradiant_static(RI0, u, meas, o)

    def misfit(pot):
        o.set_RI(n_m * np.sqrt(pot + 1))
        t, _, _ = o.solve(u)
        return float(np.sum(np.abs(t - meas) ** 2))
    rng = np.random.default_rng(3)
    ratios = []
    for _ in range(4):
        d = rng.standard_normal(ri)
        d = np.real(np.fft.ifftn(np.fft.fftn(d) * np.exp(-0.5 * sum(
            (np.fft.fftfreq(n)[tuple(slice(None) if q == j else None for q in range(3))] * 6) ** 2
            for j, n in enumerate(ri)))))
        d[pot0 > 0.5 * pot0.max()] = 0
        d *= 1e-4 / np.abs(d).max()
        fd = (misfit(pot0 + d) - misfit(pot0 - d)) / 2
        an = float(np.sum(np.real(g) * d))
        ratios.append(fd / an)
    ratios = np.array(ratios)
    assert np.all(ratios > 0), ratios
    assert ratios.max() / ratios.min() < 4, ratios


@pytest.mark.gpu
@pytest.mark.parametrize('nmax', [np.inf, 1.6])
def test_gpu_gradient_and_fista_match_oracle(nmax):
    # nmax = Inf (the default): Vimag_max = NaN, the complex TV prox; a finite nmax with
    # kappamax = 0: Vimag_max = 0, the real one.  The FISTA iterates, the gradient and the
    # RI stay in device memory between the two outer iterations.
    from inverse_solver_amd.backward_solver import BACKWARD_SOLVER_MULTI, get_default_parameters
    from inverse_solver_amd.phantoms import circle_illumination
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    from oracle.backward import get_gradiant_static, backward_multi_solve
    ri = (16, 16, 17)
    p = params(ri, iterations=12)
    RI_true = sphere_ri(ri, 0.5, n_s=1.39)
    u = circle_illumination(ri[:2], 4, 1.2, 0.532, [0.1] * 3)
    o = Oracle(p)
    o.set_RI(RI_true)
    meas, _, _ = o.solve(u)
    rng = np.random.default_rng(5)
    meas = meas + 0.1 * np.sqrt(np.mean(np.abs(meas) ** 2)) * (rng.standard_normal(meas.shape)
                                                               + 1j * rng.standard_normal(meas.shape)) / np.sqrt(2)
    RI0 = np.full(ri, 1.336) + 0.002 * (sphere_ri(ri, 0.4, n_s=1.337) - 1.336) / 0.001
    bp = get_default_parameters(dict(forward_solver_parameters=p, step=0.01, tv_param=0.001, inner_itt=20,
                                     itter_max=2, verbose=False, nmax=nmax))
    b = BACKWARD_SOLVER_MULTI(bp)
    g, err = b.get_gradiant_static(RI0, u, meas)
    g_ref, err_ref = get_gradiant_static(RI0, u.astype(np.complex128), meas, o)
    print('gradient rel err %.2e  misfit rel err %.2e' % (rel(g, g_ref), abs(err - err_ref) / err_ref))
    assert rel(g, g_ref) < 1e-4
    assert abs(err - err_ref) < 1e-4 * err_ref
    RI_gpu = b.solve(u, meas, init_RI=RI0)
    RI_ref, errs = backward_multi_solve(RI0, u.astype(np.complex128), meas, o, dict(
        RI_bg=1.336, step=0.01, tv_param=0.001, use_non_negativity=False, nmin=-np.inf, nmax=nmax, kappamax=0,
        inner_itt=20, itter_max=2))
    d = rel(RI_gpu - 1.336, RI_ref - 1.336)
    print('FISTA 2 outer iterations: RI contrast rel err %.2e' % d)
    assert d < 1e-4
    np.testing.assert_allclose(b.err_list, errs, rtol=1e-4)
    b.close()


@pytest.mark.gpu
def test_gpu_tv_prox_matches_oracle():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.tv import tv_fista_inner_device
    from oracle.backward import tv_fista_inner_v2
    ri = (16, 16, 17)
    s = FORWARD_SOLVER_CONVERGENT_BORN(params(ri, iterations=4))
    rng = np.random.default_rng(4)
    x = rng.standard_normal((20, 18, 17)) * 0.05
    xc = x + 1j * rng.standard_normal(x.shape) * 0.02
    nan = 2 * 0 * np.inf          # Vimag_max of the defaults nmax = Inf, kappamax = 0
    cases = [(x, 0.02, -np.inf, np.inf, 0, False), (x, 0.01, -0.03, 0.04, 0, True),
             (xc, 0.02, -np.inf, np.inf, nan, False), (xc, 0.01, -0.03, 0.04, 0.01, True),
             (xc, 0.02, -np.inf, np.inf, 0, False)]
    for (inp, lam, lo, hi, vim, nn) in cases:
        got = tv_fista_inner_device(s, inp, lam, lo, hi, vim, nn, False, 60)
        ref = tv_fista_inner_v2(inp, lam, lo, hi, vim, nn, False, 60)
        print('TV device vs oracle (complex=%s, Vimag_max=%s): %.2e' % (np.iscomplexobj(ref), vim, rel(got, ref)))
        assert np.iscomplexobj(got) == np.iscomplexobj(ref)
        assert rel(got, ref) < 2e-5
        if np.iscomplexobj(got):
            assert np.all(got.imag >= 0)
            if vim > 0:
                assert np.all(got.imag <= vim * (1 + 1e-6))
    s.close()


@pytest.mark.gpu
def test_gpu_field_gradient_kernel():
    """cbs_field_gradient against -i mean(conj(u) flip_z(conj(w))) (BACKWARD_SOLVER_MULTI.m:50-53),
    host and device memory, odd sizes."""
    import ctypes as C
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN, device as dv
    from inverse_solver_amd._lib import check, lib, CBS_MEM_HOST, CBS_MEM_DEVICE
    s = FORWARD_SOLVER_CONVERGENT_BORN(params((16, 16, 17), iterations=4))
    rng = np.random.default_rng(9)
    for dims, A in [((7, 5, 3), 1), ((20, 18, 17), 5)]:
        u = (rng.standard_normal(dims + (A,)) + 1j * rng.standard_normal(dims + (A,))).astype(np.complex64)
        w = (rng.standard_normal(dims + (A,)) + 1j * rng.standard_normal(dims + (A,))).astype(np.complex64)
        ref = -(1j * np.mean(np.conj(u.astype(np.complex128)) * np.flip(np.conj(w.astype(np.complex128)), axis=2),
                             axis=3))
        uf, wf = np.asfortranarray(u), np.asfortranarray(w)
        cd = (C.c_uint32 * 3)(*dims)
        g = np.empty(dims, np.complex128, order='F')
        check(lib.cbs_field_gradient(s._h, uf.ctypes.data_as(C.c_void_p), wf.ctypes.data_as(C.c_void_p), A, cd,
                                     g.ctypes.data_as(C.c_void_p), CBS_MEM_HOST))
        assert rel(g, ref) < 1e-12
        du = dv.DeviceArray.from_host(np.ascontiguousarray(uf.T))
        dw = dv.DeviceArray.from_host(np.ascontiguousarray(wf.T))
        dg = dv.DeviceArray(dims[::-1], np.complex128)
        check(lib.cbs_field_gradient(s._h, C.c_void_p(du.ptr), C.c_void_p(dw.ptr), A, cd, C.c_void_p(dg.ptr),
                                     CBS_MEM_DEVICE))
        assert rel(dg.to_host().T, ref) < 1e-12
    s.close()
