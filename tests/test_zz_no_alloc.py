"""GPU: cbs_solve_3d_noalloc, the solve that allocates no device memory, against
cbs_solve_3d (same kernels, same plan) through the C ABI.

Two calls in a row on the same handle, the second with fewer right-hand sides, each
bit for bit equal to the allocating entry point's result; the workspace shared by two
handles on one GPU; norms and the convergence test with their accumulators in the
workspace; a batch over the workspace's size refused with CBS_ERR_MEMORY."""
import ctypes as C

import numpy as np
import pytest

from helpers import params, sphere_ri, random_field

pytestmark = pytest.mark.gpu


def _abi_solve(s, fields, noalloc):
    """Runs one solve of fields (x, y, z, A) on device buffers; returns (x, y, z, A)."""
    from inverse_solver_amd import device as dv
    from inverse_solver_amd._lib import lib, check, CBS_MEM_DEVICE
    u = np.ascontiguousarray(np.transpose(fields, (3, 2, 1, 0))).astype(np.complex64)   # [rhs][z][y][x]
    d_in = dv.DeviceArray.from_host(u)
    d_out = dv.DeviceArray(u.shape, np.complex64)
    A = u.shape[0]
    if noalloc:
        check(lib.cbs_solve_3d_noalloc(s._h, C.c_void_p(d_in.ptr), A, C.c_void_p(d_out.ptr)))
    else:
        check(lib.cbs_solve_3d(s._h, C.c_void_p(d_in.ptr), A, C.c_void_p(d_out.ptr), CBS_MEM_DEVICE))
    out = np.transpose(d_out.to_host(), (3, 2, 1, 0))
    d_in.free()
    d_out.free()
    return out


def _pair(p, RI, opts=None):
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    a = FORWARD_SOLVER_CONVERGENT_BORN(p, options=opts)
    a.set_RI(RI)
    b = FORWARD_SOLVER_CONVERGENT_BORN(p, options=opts)
    b.set_RI(RI)
    return a, b


@pytest.mark.parametrize('form', [0, 1, 2])
def test_noalloc_equals_allocating_solve_twice(form):
    ri = (16, 16, 16)
    p = params(ri, iterations=14)
    RI = sphere_ri(ri, 0.5, n_s=1.42)
    fields = np.stack([random_field(ri, 40 + a) for a in range(3)], axis=-1)
    ref_h, na_h = _pair(p, RI, dict(form=form))
    print('form', form, ref_h.plan(), flush=True)
    for A in (3, 1):                       # the second call with a smaller input
        ref = _abi_solve(ref_h, fields[..., :A], False)
        got = _abi_solve(na_h, fields[..., :A], True)
        np.testing.assert_array_equal(got, ref)
    # the allocating entry point of the same handle still works after the workspace call
    np.testing.assert_array_equal(_abi_solve(na_h, fields[..., :2], False), _abi_solve(ref_h, fields[..., :2], False))
    ref_h.close()
    na_h.close()


def test_noalloc_shared_workspace_norms_and_convergence():
    """Two handles (different grids) take turns on the one workspace; the norm and
    convergence accumulators there start from zero on every call."""
    ri1, ri2 = (16, 16, 16), (16, 12, 10)
    p1 = params(ri1, iterations=15)
    p2 = params(ri2, iterations=12)
    RI1 = sphere_ri(ri1, 0.5, n_s=1.40)
    RI2 = sphere_ri(ri2, 0.4, n_s=1.44)
    f1 = np.stack([random_field(ri1, 50 + a) for a in range(2)], axis=-1)
    f2 = np.stack([random_field(ri2, 60 + a) for a in range(2)], axis=-1)
    r1, n1 = _pair(p1, RI1, dict(track_norms=1))
    r2, n2 = _pair(p2, RI2)
    for h in (r2, n2):
        h.set_convergence(0.26)         # stops these series at iteration 9 of 12
    for call in range(2):
        np.testing.assert_array_equal(_abi_solve(n1, f1, True), _abi_solve(r1, f1, False))
        np.testing.assert_allclose(n1.norms(2), r1.norms(2), rtol=1e-12)
        np.testing.assert_array_equal(_abi_solve(n2, f2[..., :2 - call], True), _abi_solve(r2, f2[..., :2 - call], False))
        assert list(n2.iterations(2 - call)) == list(r2.iterations(2 - call))
        assert min(r2.iterations(2 - call)) < 12
    for h in (r1, n1, r2, n2):
        h.close()


def test_noalloc_refuses_batches_over_the_workspace():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd import device as dv
    from inverse_solver_amd._lib import lib, CBS_ERR_MEMORY
    ri = (64, 64, 64)                                  # 96^3 simulation grid: 21.2 MB of state per rhs
    p = params(ri, iterations=4, boundary_thickness=[8, 8, 8])
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)
    s.set_RI(sphere_ri(ri, 1.0))
    assert s.simulation_size == [96, 96, 96]
    u = np.zeros((4, 64, 64, 64), np.complex64)
    d = dv.DeviceArray.from_host(u)
    st = lib.cbs_solve_3d_noalloc(s._h, C.c_void_p(d.ptr), 4, C.c_void_p(d.ptr))
    assert st == CBS_ERR_MEMORY
    assert 'no-allocation workspace' in lib.cbs_last_error().decode()
    # two rhs fit; the result equals the allocating solve
    f = np.stack([random_field(ri, 90 + a) for a in range(2)], axis=-1)
    ref = _abi_solve(s, f, False)
    np.testing.assert_array_equal(_abi_solve(s, f, True), ref)
    d.free()
    s.close()
