"""Shared builders for the parity tests (inputs only; the checker is the oracle)."""
import numpy as np


def params(ri_size, bt_px_target=4, iterations=20, NA=1.2, field_size=None, **kw):
    """Solver parameters for an RI grid; boundary_thickness chosen so that the
    boundary is bt_px_target pixels at dx = 0.1 um (round(bt*0.532/1.336/0.2))."""
    bt = bt_px_target * 0.2 * 1.336 / 0.532
    fs = field_size or ri_size[:2]
    p = dict(size=[fs[0], fs[1], ri_size[2]], wavelength=0.532, NA=NA, RI_bg=1.336, resolution=[0.1, 0.1, 0.1],
             vector_simulation=False, use_abbe_sine=True, return_transmission=True, return_reflection=True,
             return_3D=True, use_GPU=True, iterations_number=iterations, boundary_thickness=[bt, bt, bt],
             boundary_sharpness=1, verbose=False, acyclic=True, RI_xy_size=[ri_size[0], ri_size[1]],
             RI_center=[0, 0])
    p.update(kw)
    return p


def sphere_ri(shape, radius_um, n_s=1.46, n_m=1.336, center=None, dx=0.1):
    c = [(n // 2) if center is None else center[j] for j, n in enumerate(shape)]
    g = np.meshgrid(*[(np.arange(n) - c[j]) * dx for j, n in enumerate(shape)], indexing='ij')
    r = np.sqrt(sum(x ** 2 for x in g))
    return np.where(r < radius_um, n_s, n_m).astype(np.float64)


def random_field(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)


def plane_wave(shape, kx_frac=0.0, ky_frac=0.0, n_m=1.336, wl=0.532, dx=0.1):
    k = 2 * np.pi * n_m / wl
    kx, ky = k * kx_frac, k * ky_frac
    kz = np.sqrt(k * k - kx * kx - ky * ky)
    x, y, z = np.meshgrid(*[np.arange(n) * dx for n in shape], indexing='ij')
    return np.exp(1j * (kx * x + ky * y + kz * z))


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))
