"""Input generators against closed forms (PHANTOM 'bead', FIELD_GENERATOR 'circle',
seeded spheres, Gaussian random field)."""
import numpy as np

from inverse_solver_amd.phantoms import (bead, spheres, circle_illumination, circle_delta_positions,
                                         gaussian_random_field, circle_directions)


def test_bead_volume_and_normalisation():
    b = bead([40, 40, 40], [20, 20, 20])
    assert b.min() == 0 and b.max() == 1
    vol = b.sum()
    assert abs(vol - 4 / 3 * np.pi * 10 ** 3) / (4 / 3 * np.pi * 10 ** 3) < 0.02
    # symmetric about the centre voxel floor(n/2)
    np.testing.assert_allclose(b[20 + 3, 20, 20], b[20 - 3, 20, 20])


def test_spheres_seeded_non_overlapping_inside():
    RI, placed = spheres((64, 64, 64), 5, (0.4, 1.2), (1.36, 1.48), seed=7)
    RI2, placed2 = spheres((64, 64, 64), 5, (0.4, 1.2), (1.36, 1.48), seed=7)
    np.testing.assert_array_equal(RI, RI2)
    box = 6.4
    for i, (c, r, n) in enumerate(placed):
        assert 0.4 <= r <= 1.2 and 1.36 <= n <= 1.48
        assert np.all(c - r >= 0) and np.all(c + r <= box)
        for c2, r2, _ in placed[i + 1:]:
            assert np.linalg.norm(c - c2) > r + r2
    assert RI.min() >= 1.336 - 1e-12 and RI.max() <= 1.48 + 1e-12


def test_circle_illumination_plane_waves():
    N, A = 32, 9
    u = circle_illumination((N, N), A, 1.2, 0.532, [0.1] * 3)
    assert u.shape == (N, N, 1, A)
    np.testing.assert_allclose(np.abs(u), 1.0, atol=1e-6)          # unit plane waves
    np.testing.assert_allclose(u[..., 0], u[0, 0, 0, 0], atol=1e-6)   # normal incidence first
    d = circle_directions(A)
    assert d[0] == (0.0, 0.0)
    for (d1, d2) in d[1:]:
        assert abs(np.hypot(d1, d2) - 1) < 1e-12
    pos = circle_delta_positions((N, N), A, 1.2, 0.532, [0.1] * 3)
    kmax_px = 0.95 * 1.2 / 0.532 / (1 / (0.1 * N))
    for (p1, p2), (d1, d2) in zip(pos, d):
        assert p1 == N // 2 + int(np.sign(d1 * kmax_px) * np.floor(abs(d1 * kmax_px) + 0.5))
        assert p2 == N // 2 + int(np.sign(d2 * kmax_px) * np.floor(abs(d2 * kmax_px) + 0.5))
        spec = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(u[:, :, 0, pos.index((p1, p2))])))
        assert np.unravel_index(np.argmax(np.abs(spec)), spec.shape) == (p1, p2)


def test_gaussian_random_field_bounds():
    g = gaussian_random_field((32, 32, 16), 0.5, (1.33, 1.40), seed=3)
    inside = g != 1.336
    assert inside.mean() > 0.2
    assert g[inside].min() >= 1.33 - 1e-12 and g[inside].max() <= 1.40 + 1e-12
