"""The product's host construction of the refocusing / transmission kernels
(inverse_solver_amd/io_kernels.py) equals the oracle's line-by-line restatement
of truncated_green_plus and the constructor (oracle/hankel.py, oracle/cbs.py)."""
import numpy as np
import pytest

from helpers import params, rel


@pytest.mark.parametrize('cyclic,rc', [(False, [0, 0]), (True, [0, 0]), (False, [2, -1]), (False, [-3, 1])])
def test_io_kernels_match_oracle(cyclic, rc):
    from inverse_solver_amd.io_kernels import build_io_kernels
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri = (12, 10, 8)
    p = params(ri, bt_px_target=2, iterations=4, RI_center=rc)
    if cyclic:
        p['boundary_thickness'] = [0, 0, 1.0]
    o = Oracle(p)
    ref, kt, kr = build_io_kernels(o.parameters, o.expected_RI_size, o.cyclic_boundary_xy)
    assert rel(ref, o.refocusing_util) < 1e-10
    assert rel(kt, o.kernel_trans) < 1e-10
    assert rel(kr, o.kernel_ref) < 1e-10


def test_bessel_zeros_product_vs_reference_table():
    import os
    from inverse_solver_amd.io_kernels import bessel_j0_zeros
    c = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'dht_bessel_zeros.npz'))['c']
    np.testing.assert_allclose(bessel_j0_zeros(c.shape[1]), c[0], rtol=0, atol=1e-11)


def test_oracle_refuses_ri_center_in_solve():
    """crop_field2RI (:173-179) crops parameters.size from an array of that size, so a
    nonzero RI_center indexes outside it and the reference's 2D solve stops (:554)."""
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    from helpers import sphere_ri
    from inverse_solver_amd.phantoms import circle_illumination
    ri = (12, 12, 8)
    for rc in ([1, 0], [0, 2], [0, -2]):
        o = Oracle(params(ri, bt_px_target=2, iterations=2, RI_center=rc))
        o.set_RI(sphere_ri(ri, 0.3, n_s=1.4))
        # a positive offset stops in the crop (:554), a negative one already in the
        # refocusing product (:550), whose cropped kernel is smaller than the grid
        with pytest.raises((IndexError, ValueError)):
            o.solve(circle_illumination(ri[:2], 1, 1.2, 0.532, [0.1] * 3))


def octant_expand(oct_r, oct_t, cx, cy, rz):
    """The reference's reading of first-octant kernels (execution2.cu:195-275): centred
    index c -> octant index conv/2 - |c - conv/2| per axis, z -> RI_z/2 - |z - RI_z/2|;
    from the middle z plane on, the refocusing value is conjugated and the transmission
    value negated and conjugated; reflection = conj(transmission)."""
    g1 = np.array([cx // 2 - abs(c - cx // 2) for c in range(cx)])
    g2 = np.array([cy // 2 - abs(c - cy // 2) for c in range(cy)])
    g3 = np.array([rz // 2 - abs(z - rz // 2) for z in range(rz)])
    up = (np.arange(rz) >= rz // 2)[None, None, :]
    R = oct_r[g1][:, g2][:, :, g3]
    T = oct_t[g1][:, g2][:, :, g3]
    R = np.where(up, np.conj(R), R)
    T = np.where(up, -np.conj(T), T)
    return R, T, np.conj(T)


@pytest.mark.parametrize('ri,bt', [((16, 16, 17), 4), ((16, 16, 16), 4), ((12, 14, 9), 3), ((12, 12, 9), 0)])
def test_octant_symmetry_of_the_kernels(ri, bt):
    """The first octants the reference's CUDA subclass uploads
    (FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m:28-38) determine the whole refocusing,
    transmission and reflection kernels of the MATLAB class (:65-115) through that
    symmetry, to FP64 rounding (so cbs_set_io_kernels_octant, which expands them so,
    serves FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m unchanged)."""
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    o = Oracle(params(ri, bt_px_target=bt, iterations=4))
    R, T, K = o.refocusing_util, o.kernel_trans, o.kernel_ref
    cx, cy, rz = R.shape
    gx, gy, gz = cx // 2 + 1, cy // 2 + 1, rz // 2 + 1
    Re, Te, Ke = octant_expand(R[:gx, :gy, :gz], T[:gx, :gy, :gz], cx, cy, rz)
    for got, want in ((Re, R), (Te, T), (Ke, K)):
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
