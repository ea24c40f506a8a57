"""The product's host construction of the refocusing / transmission kernels
(inverse_solver_amd/io_kernels.py) equals the oracle's line-by-line restatement
of truncated_green_plus and the constructor (oracle/hankel.py, oracle/cbs.py)."""
import numpy as np
import pytest

from helpers import params, rel


@pytest.mark.parametrize('cyclic', [False, True])
def test_io_kernels_match_oracle(cyclic):
    from inverse_solver_amd.io_kernels import build_io_kernels
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri = (12, 10, 8)
    p = params(ri, bt_px_target=2, iterations=4)
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
