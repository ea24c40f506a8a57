"""Seeded synthetic inputs: phantoms and illumination fields.

* bead(): PHANTOM.get with name='bead' (Codes/PHANTOM.m:10-54,94-98): 8 sub-samples
  at +-1/4 voxel, normalised to [0, 1].
* spheres(): several non-overlapping spheres with seeded radii and indices,
  anti-aliased with the same 8-sub-sample rule (the multi-sphere workloads).
* gaussian_random_field(): smooth RI texture clipped inside an ellipsoid.
* circle_illumination(): FIELD_GENERATOR 'circle' (Codes/FIELD_GENERATOR.m:25-72):
  normal incidence first, then A-1 directions on a circle at 0.95 NA, each a
  unit k-space delta turned into a plane wave by a centred fft2.
* incident_plane_waves(): the 3D plane waves those illuminations refocus to,
  evaluated analytically on the RI grid (input of the Born loop alone).

All seeds go through numpy.random.Generator(PCG64(seed)).
"""
from __future__ import annotations

import numpy as np

from .params import round_half_away, centered_fft2

_SUB = np.array([[1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1],
                 [1, 1, -1], [-1, 1, -1], [1, -1, -1], [-1, -1, -1]], float) / 4.0


def _centered(n):
    return np.arange(n, dtype=np.float64) - (n // 2)


def bead(outer_size, inner_size):
    """Codes/PHANTOM.m:10-54 ('bead', antialiasing on), normalised as :94-98."""
    d = [_centered(n) for n in outer_size]
    ph = np.zeros(outer_size)
    for s in _SUB:
        r2 = sum(((2 * (d[j] + s[j]) / inner_size[j]).reshape([-1 if q == j else 1 for q in range(3)])) ** 2
                 for j in range(3))
        ph += (np.sqrt(r2) < 1)
    ph -= ph.min()
    if ph.max() != 0:
        ph /= ph.max()
    return ph


def bead_ri(outer_size, diameter_px, n_s, n_m):
    return n_m + bead(outer_size, [diameter_px] * 3) * (n_s - n_m)


def spheres(shape, n_spheres, radius_um, n_range, seed, dx=0.1, n_m=1.336, max_tries=100000):
    """Non-overlapping spheres fully inside the box; radii ~ U(radius_um), RI ~ U(n_range).
    Returns (RI, list of (center_um, radius_um, n))."""
    rng = np.random.default_rng(seed)
    box = np.array(shape, float) * dx
    placed = []
    tries = 0
    while len(placed) < n_spheres:
        tries += 1
        if tries > max_tries:
            raise RuntimeError('could not place non-overlapping spheres')
        r = rng.uniform(*radius_um)
        if np.any(2 * r >= box):
            continue
        c = np.array([rng.uniform(r, b - r) for b in box])
        if all(np.linalg.norm(c - c2) > r + r2 for c2, r2, _ in placed):
            placed.append((c, r, rng.uniform(*n_range)))
    RI = np.full(shape, n_m, dtype=np.float64)
    coords = [(np.arange(n) + 0.5) * dx for n in shape]
    for c, r, n in placed:
        lo = [max(0, int((c[j] - r) / dx) - 1) for j in range(3)]
        hi = [min(shape[j], int((c[j] + r) / dx) + 2) for j in range(3)]
        sub = [coords[j][lo[j]:hi[j]] for j in range(3)]
        frac = np.zeros([hi[j] - lo[j] for j in range(3)])
        for s in _SUB:
            g = [(sub[j] + s[j] * dx - c[j]).reshape([-1 if q == j else 1 for q in range(3)]) for j in range(3)]
            frac += (g[0] ** 2 + g[1] ** 2 + g[2] ** 2) < r * r
        frac /= len(_SUB)
        blk = RI[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
        RI[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = blk + frac * (n - n_m)
    return RI, placed


def gaussian_random_field(shape, corr_um, n_range, seed, dx=0.1, n_m=1.336, semi_axes_frac=0.45):
    """White noise convolved with a Gaussian (correlation length corr_um), normalised,
    mapped to n_range and clipped to it inside an ellipsoid; n_m outside."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    k = np.meshgrid(*[np.fft.fftfreq(n, d=dx) for n in shape], indexing='ij', sparse=True)
    k2 = sum(kk ** 2 for kk in k)
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-2 * (np.pi * corr_um) ** 2 * k2)))
    f = (f - f.mean()) / (f.std() + 1e-30)
    lo, hi = n_range
    n = np.clip((lo + hi) / 2 + f * (hi - lo) / 4, lo, hi)
    g = np.meshgrid(*[(_centered(s)) / (semi_axes_frac * s) for s in shape], indexing='ij', sparse=True)
    inside = sum(gg ** 2 for gg in g) <= 1
    return np.where(inside, n, n_m)


def circle_directions(A, start_with_normal=True):
    """Direction pairs (d1, d2) of FIELD_GENERATOR 'circle' (:44-60)."""
    out = []
    for m in range(1, A + 1):
        div = 1.0 / (A - 1) if (start_with_normal and A > 1) else 1.0 / A
        d1 = np.sin((m - 1) * 2 * np.pi * div)
        d2 = np.cos((m - 1) * 2 * np.pi * div)
        if m == 1 and start_with_normal:
            d1 = d2 = 0.0
        out.append((d1, d2))
    return out


def circle_delta_positions(size_xy, A, NA, wavelength, resolution, pct=0.95):
    """0-based k-space delta positions of FIELD_GENERATOR (:62-65)."""
    kmax = NA / wavelength
    res = [1.0 / (resolution[j] * size_xy[j]) for j in range(2)]
    pos = []
    for d1, d2 in circle_directions(A):
        p1 = size_xy[0] // 2 + int(round_half_away(d1 * pct * kmax / res[0]))
        p2 = size_xy[1] // 2 + int(round_half_away(d2 * pct * kmax / res[1]))
        pos.append((p1, p2))
    return pos


def circle_illumination(size_xy, A, NA, wavelength, resolution, pct=0.95):
    """Scalar input fields (x, y, 1, A), complex64 (Codes/FIELD_GENERATOR.m:25-72)."""
    out = np.zeros((size_xy[0], size_xy[1], 1, A), dtype=np.complex128)
    for m, (p1, p2) in enumerate(circle_delta_positions(size_xy, A, NA, wavelength, resolution, pct)):
        out[p1, p2, 0, m] = 1
    return centered_fft2(out).astype(np.complex64)


def incident_plane_waves(ri_size, A, NA, wavelength, n_m, resolution, pct=0.95, field_xy=None):
    """3D plane waves (x, y, z, A) for the circle directions on the RI grid.

    The transverse wave vector is the delta position of circle_illumination on
    the field grid (sign as its centred fft2 produces); k_z follows from the
    medium wavenumber.  Complex64, unit amplitude."""
    fxy = field_xy or ri_size[:2]
    res = [1.0 / (resolution[j] * fxy[j]) for j in range(2)]
    k0 = n_m / wavelength
    x = [(_centered(n) * resolution[j]) for j, n in enumerate(ri_size)]
    out = np.empty(tuple(ri_size) + (A,), dtype=np.complex64)
    for m, (p1, p2) in enumerate(circle_delta_positions(fxy, A, NA, wavelength, resolution, pct)):
        kx = -(p1 - fxy[0] // 2) * res[0]
        ky = -(p2 - fxy[1] // 2) * res[1]
        kz = np.sqrt(max(k0 * k0 - kx * kx - ky * ky, 0.0))
        ph = (kx * x[0][:, None, None] + ky * x[1][None, :, None] + kz * x[2][None, None, :])
        out[..., m] = np.exp(2j * np.pi * ph)
    return out
