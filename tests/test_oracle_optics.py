"""Pins oracle/optics.py and oracle/hankel.py against closed forms, the
reference's own Bessel-zero table (Codes/sub_code/hankel_transform/dht.mat)
and the Hankel-transform known answers test_hankel.m uses."""
import os

import numpy as np
import pytest
import scipy.special as sp

from oracle.optics import derive_optical_tool, ri2potential, potential2ri, fft_flip, cfft2, cifft2
from oracle.hankel import jn_roots, DiscreteHankelCropedFourier, bin_search_nearest
from oracle.cbs import matlab_round

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'dht_bessel_zeros.npz')


def test_jn_roots_match_reference_table():
    # The reference table's J2 row holds a NaN at its first positive root (the
    # reference's own root finder does not resolve it from its starting guess),
    # so the rows it holds completely are compared: J0 (the order the solver
    # uses), J1, J3 and J4.
    c = np.load(GOLD)['c']
    assert np.isnan(c[2]).sum() == 1
    for n in (0, 1, 3, 4):
        r = jn_roots(n, c.shape[1])
        np.testing.assert_allclose(r, c[n], rtol=0, atol=1e-11)


def test_jn_roots_match_scipy_and_are_zeros():
    r = jn_roots(0, 40001)
    np.testing.assert_allclose(r[:2000], sp.jn_zeros(0, 2000), rtol=1e-13)
    assert np.max(np.abs(sp.j0(r))) < 1e-12
    assert np.all(np.diff(r) > 3.0)


def test_hankel_gaussian_pair_and_disk():
    # 2D Fourier transform of a radial function with angular frequency k:
    # exp(-r^2/s^2) -> pi s^2 exp(-k^2 s^2 / 4); disk radius a -> 2 pi a J1(k a)/k
    h = DiscreteHankelCropedFourier(1.0, 3000, 0, 0.25)
    s = 0.1
    F = h.forward(np.exp(-h.r ** 2 / s ** 2)[:, None])[:, 0]
    ana = np.pi * s * s * np.exp(-h.k ** 2 * s * s / 4)
    assert np.linalg.norm(F - ana) / np.linalg.norm(ana) < 1e-12
    a = 0.3
    F = h.forward((h.r < a).astype(float)[:, None])[:, 0]
    ana = 2 * np.pi * a * sp.j1(h.k * a) / h.k
    assert np.linalg.norm(F - ana) / np.linalg.norm(ana) < 1e-3


def test_hankel_roundtrip_uncropped():
    h = DiscreteHankelCropedFourier(1.0, 1500, 0, 1.0)
    f = np.exp(-h.r ** 2 / 0.05 ** 2)[:, None]
    assert h.pos == 1500
    assert np.linalg.norm(h.backward(h.forward(f)) - f) / np.linalg.norm(f) < 1e-12


def test_bin_search_nearest_bracket():
    rng = np.random.default_rng(1)
    pos = np.sort(rng.random(1000)) * 10
    val = rng.random((30, 40)) * 10
    n1, n2, w = bin_search_nearest(val, pos)
    # n1 = last index with pos < value (0 when none), n2 = n1 + 1 when pos[n1] < value
    expect = np.clip(np.searchsorted(pos, val, side='left') - 1, 0, None)
    np.testing.assert_array_equal(n1, expect)
    inner = (n2 != n1)
    assert np.all(pos[n1][inner] < val[inner]) and np.all(pos[n2][inner] >= val[inner])
    np.testing.assert_allclose(w[inner], ((val - pos[n1]) / (pos[n2] - pos[n1]))[inner])


def test_optical_tool_closed_forms():
    u = derive_optical_tool(dict(size=[8, 7, 6], resolution=[0.1, 0.2, 0.3], wavelength=0.5, NA=1.1, RI_bg=1.33))
    assert u['fourier_res'] == pytest.approx([1 / 0.8, 1 / 1.4, 1 / 1.8])
    np.testing.assert_allclose(u['fourier_coor'][0].ravel(), (np.arange(8) - 4) / 0.8)
    np.testing.assert_allclose(u['image_coor'][1].ravel(), (np.arange(7) - 3) * 0.2)
    assert u['k0_nm'] == pytest.approx(1.33 / 0.5)
    assert u['dV'] == pytest.approx(0.006)
    k3 = np.sqrt(np.maximum(u['k0_nm'] ** 2 - u['coorxy'] ** 2, 0))
    np.testing.assert_allclose(u['k3'], k3)
    assert np.array_equal(u['NA_circle'], u['coorxy'] < 1.1 / 0.5)


def test_potential_roundtrip_and_zero():
    RI = np.array([1.336, 1.40, 1.5 + 0.01j])
    V = ri2potential(RI, 0.532, 1.336)
    assert abs(V[0]) < 1e-12
    np.testing.assert_allclose(potential2ri(V, 0.532, 1.336), RI, rtol=1e-14)
    k = 2 * np.pi * 1.336 / 0.532
    assert V[1] == pytest.approx(k * k * (1.40 ** 2 / 1.336 ** 2 - 1))


@pytest.mark.parametrize('n', [5, 6])
def test_fft_flip_is_index_negation(n):
    a = np.random.default_rng(2).random((n, n + 1, n + 2))
    f = fft_flip(a, [1, 1, 1], False)
    for idx in [(0, 0, 0), (1, 2, 3), (n - 1, n, n + 1), (n // 2, 1, 2)]:
        neg = tuple((-i) % s for i, s in zip(idx, a.shape))
        assert f[idx] == a[neg]
    np.testing.assert_array_equal(fft_flip(f, [1, 1, 1], False), a)


def test_centered_fft_of_centered_delta_is_constant():
    x = np.zeros((6, 5), complex)
    x[3, 2] = 1
    np.testing.assert_allclose(cfft2(x), np.ones((6, 5)), atol=1e-15)
    y = np.random.default_rng(3).random((6, 5)) + 0j
    np.testing.assert_allclose(cifft2(cfft2(y)), y, atol=1e-14)


def test_matlab_round_half_away():
    np.testing.assert_array_equal(matlab_round([0.5, 1.5, 2.5, -0.5, 15.93, 3.49]), [1, 2, 3, -1, 16, 3])


@pytest.mark.parametrize('shape', [(8, 8), (9, 7), (16, 11), (13, 20), (1, 5)])
def test_centred_pupil_filter_equals_uncentred_with_shifted_pupil(shape):
    """The device pupil filter (cbs_field_pupil, cbs_solve_field) computes
    ifft2(ifftshift(P) * fft2(x)) without the centring shifts of the reference's
    fftshift(ifft2(ifftshift(P * fftshift(fft2(ifftshift(x)))))): a shift by n // 2 before
    the forward transform multiplies the spectrum by exp(2 pi i k (n // 2) / n) and the one
    after the inverse by its conjugate, for even and odd n alike."""
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape + (3,)) + 1j * rng.standard_normal(shape + (3,))
    P = rng.random(shape)
    ax = (0, 1)
    ref = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(
        P[:, :, None] * np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(x, axes=ax), axes=ax), axes=ax), axes=ax),
        axes=ax), axes=ax)
    got = np.fft.ifft2(np.fft.ifftshift(P)[:, :, None] * np.fft.fft2(x, axes=ax), axes=ax)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.max(np.abs(ref)))
    # a wrong shift (fftshift of P instead of ifftshift) differs for odd lengths
    if shape[0] % 2 or shape[1] % 2:
        bad = np.fft.ifft2(np.fft.fftshift(P)[:, :, None] * np.fft.fft2(x, axes=ax), axes=ax)
        assert np.max(np.abs(bad - ref)) > 1e-6 * np.max(np.abs(ref))
