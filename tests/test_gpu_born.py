"""GPU parity of the Born loop (cbs_solve_3d) against the FP64 oracle.

Tolerance: relative L2 error <= 1e-5 (complex64 GPU vs complex128 oracle), the
plain-form bar of the solver's accuracy target.
"""
import numpy as np
import pytest

from helpers import params, sphere_ri, random_field, plane_wave, rel
from oracle.cbs import ForwardSolverConvergentBorn as Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _gpu():
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    return FORWARD_SOLVER_CONVERGENT_BORN


def _oracle_loop(p, RI, fields):
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    return np.stack([o.born_loop(f) for f in fields], axis=-1), o


CASES = [
    # (ri_size, iterations, options)
    ((16, 16, 16), 20, dict(form=0)),
    ((16, 16, 16), 20, dict(form=1)),
    ((16, 16, 16), 20, dict(form=0, split_a=6)),
    ((16, 12, 10), 12, dict(form=0)),
    ((16, 12, 10), 12, dict(form=1)),
    ((16, 16, 16), 20, dict(form=3)),            # per-axis form, every pass in its own launch
    ((16, 12, 10), 12, dict(form=3)),
    ((17, 17, 17), 10, dict(form=0)),
    ((17, 17, 19), 10, dict(form=0)),
    ((16, 16, 16), 20, dict(form=2)),            # four-pass form (row blocks, y and z passes)
    ((16, 16, 16), 9, dict(form=2, threads_real=256)),
]


@pytest.mark.parametrize('ri_size,iters,opts', CASES)
def test_born_loop_matches_oracle(ri_size, iters, opts):
    p = params(ri_size, iterations=iters)
    RI = sphere_ri(ri_size, 0.5)
    u = plane_wave(ri_size, 0.2, -0.1)
    ref, o = _oracle_loop(p, RI, [u])
    s = _gpu()(p, options=opts)
    s.set_RI(RI)
    assert s.simulation_size == list(o.V.shape)
    got = s.solve_incident(u)
    err = rel(got, ref[..., 0])
    print(ri_size, opts, s.plan(), 'rel err', err)
    assert err < TOL
    s.close()


@pytest.mark.parametrize('opts', [dict(form=0, split_a=24, tile_t=8), dict(form=0, split_a=24, tile_t=16),
                                  dict(form=0, split_a=32, tile_t=8), dict(form=0, split_a=48, tile_t=8)])
def test_small_grid_plans_match_oracle(opts):
    # 64^3 RI + 16-px boundary = 96^3 simulation grid (the single-angle latency case):
    # the small-tile plans that give both passes more blocks than CUs
    ri = (64, 64, 64)
    p = params(ri, iterations=8, boundary_thickness=[8, 8, 8])
    RI = sphere_ri(ri, 1.5, n_s=1.40)
    u = plane_wave(ri, 0.15, 0.05)
    ref, o = _oracle_loop(p, RI, [u])
    s = _gpu()(p, options=opts)
    s.set_RI(RI)
    assert s.simulation_size == [96, 96, 96]
    got = s.solve_incident(u)
    plan = s.plan()
    err = rel(got, ref[..., 0])
    print(opts, plan, 'rel err', err, flush=True)
    assert plan['a'] == opts['split_a'] and plan['tile_t'] == opts['tile_t']
    assert err < TOL
    s.close()


def test_block_orders_give_identical_results():
    # every block order (rhs, row or storage-order numbering of each pass, cached or
    # nontemporal k-space loads) computes the same values
    ri = (16, 16, 16)
    p = params(ri, iterations=10)
    RI = sphere_ri(ri, 0.6, n_s=1.40)
    fields = np.stack([random_field(ri, 90 + j) for j in range(8)], -1)
    outs = []
    for order in range(16):
        s = _gpu()(p, options=dict(form=0, block_order=order))
        s.set_RI(RI)
        outs.append(s.solve_incident(fields))
        assert s.plan()['block_order'] == order
        s.close()
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.parametrize('form', [0, 1, 2, 3])
def test_batch_equals_individual_and_oracle(form):
    ri = (16, 16, 16)
    p = params(ri, iterations=16)
    RI = sphere_ri(ri, 0.6, n_s=1.40)
    fields = [random_field(ri, s) for s in range(3)]
    ref, _ = _oracle_loop(p, RI, fields)
    s = _gpu()(p, options=dict(form=form))
    s.set_RI(RI)
    batch = s.solve_incident(np.stack(fields, -1))
    for j in range(3):
        s.set_RI(RI)
        one = s.solve_incident(fields[j])
        assert rel(batch[..., j], ref[..., j]) < TOL
        np.testing.assert_allclose(one, batch[..., j], rtol=0, atol=1e-6 * np.abs(one).max())
    s.close()


def test_autotuned_plan_and_batch_width_match_oracle():
    """cbs_autotune times every fused kernel pair, the block orders of the fastest and
    then the batch width; whatever it keeps, the solve of the tuned handle equals the
    oracle (odd Bornmax, so a split batch must also carry the flip parity across its
    launches), and the chosen width is the full call or one of the narrower trials."""
    ri = (16, 16, 16)
    n = 5
    p = params(ri, iterations=11)
    RI = sphere_ri(ri, 0.6, n_s=1.40)
    fields = [random_field(ri, 20 + k) for k in range(n)]
    ref, _ = _oracle_loop(p, RI, fields)
    s = _gpu()(p)
    s.set_RI(RI)
    ms = s.autotune(n, 4)
    plan = s.plan()
    print('autotuned', plan, '%.3f ms' % ms, flush=True)
    assert ms > 0 and plan['form'] == 'fused'
    assert plan['max_batch'] in (0, 1, 2, 3, 4)
    s.set_RI(RI)                      # the flip state as at the oracle's start
    got = s.solve_incident(np.stack(fields, -1))
    for j in range(n):
        assert rel(got[..., j], ref[..., j]) < TOL, j
    s.close()


@pytest.mark.parametrize('form', [0, 1, 2, 3])
def test_odd_bornmax_alternation_persists_across_rhs(form):
    """With odd Bornmax the reference leaves Greenp flipped after each angle
    (FORWARD_SOLVER_CONVERGENT_BORN.m:579-585), so consecutive angles start on
    opposite shifts.  The oracle keeps that state; the GPU must match it."""
    ri = (16, 16, 16)
    p = params(ri, iterations=5)
    RI = sphere_ri(ri, 0.6)
    fields = [random_field(ri, 10 + s) for s in range(3)]
    ref, _ = _oracle_loop(p, RI, fields)
    s = _gpu()(p, options=dict(form=form))
    s.set_RI(RI)
    got = s.solve_incident(np.stack(fields, -1))
    for j in range(3):
        assert rel(got[..., j], ref[..., j]) < TOL
    s.close()


@pytest.mark.parametrize('form', [0, 1, 2, 3])
@pytest.mark.parametrize('iters', [1, 2, 3, 4, 5, 6, 7])
def test_short_series(iters, form):
    """Series lengths around the special first iterations and the paired Field
    accumulation (from jj = 4, an even jj below the last defers its increment to the
    odd jj after it; a series ending on an even jj adds its own)."""
    ri = (16, 16, 16)
    p = params(ri, iterations=iters)
    RI = sphere_ri(ri, 0.5)
    u = plane_wave(ri)
    ref, _ = _oracle_loop(p, RI, [u])
    s = _gpu()(p, options=dict(form=form))
    s.set_RI(RI)
    assert rel(s.solve_incident(u), ref[..., 0]) < TOL
    s.close()


def test_graph_replay_equals_eager():
    """use_graph captures the iteration loop once and replays it; every replay
    (with eager work and synchronisation in between) must equal the eager solve."""
    ri = (16, 16, 16)
    p = params(ri, iterations=14)
    RI = sphere_ri(ri, 0.5)
    fields = np.stack([random_field(ri, 20 + s) for s in range(2)], -1)
    s = _gpu()(p, options=dict(use_graph=0))
    s.set_RI(RI)
    eager = s.solve_incident(fields)
    g = _gpu()(p, options=dict(use_graph=1))
    g.set_RI(RI)
    for rep in range(4):
        got = g.solve_incident(fields)
        np.testing.assert_array_equal(got, eager)
        s.solve_incident(fields[..., :1])       # eager work on another handle in between
    s.close()
    g.close()


@pytest.mark.parametrize('form,P', [(0, 1), (1, 1), (2, 1), (0, 3)])
def test_increment_norms_match_oracle(form, P):
    """track_norms: the FP64 |f_n|^2 of every series term over the simulation grid
    (summed over the components of a vector run) equals the oracle's, and tracking
    them leaves the fields unchanged (the steady-state kernels accumulate them as
    the general ones do)."""
    ri = (16, 16, 16)
    M = 11
    p = params(ri, iterations=M, vector_simulation=(P == 3))
    RI = sphere_ri(ri, 0.5, n_s=1.42)
    if P == 1:
        fields = [random_field(ri, 70 + a) for a in range(2)]
    else:
        fields = [np.stack([random_field(ri, 80 + 3 * a + j) for j in range(3)], axis=-1) for a in range(2)]
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    inc_ref = []
    for f in fields:
        inc = []
        o.born_loop(f, increments=inc)
        inc_ref.append(inc)
    inc_ref = np.array(inc_ref)
    s = _gpu()(p, options=dict(form=form))
    s.set_RI(RI)
    plain = s.solve_incident(np.stack(fields, axis=-1))
    s.close()
    s = _gpu()(p, options=dict(form=form, track_norms=1))
    s.set_RI(RI)
    got = s.solve_incident(np.stack(fields, axis=-1))
    nrm = s.norms(2)
    print(form, P, s.plan()['form'], 'norm rel err', np.max(np.abs(nrm - inc_ref) / inc_ref), flush=True)
    assert nrm.shape == inc_ref.shape == (2, M)
    np.testing.assert_allclose(nrm, inc_ref, rtol=1e-5)
    assert rel(got, plain) < 1e-6
    s.close()


def test_host_buffer_pipeline_equals_device_path():
    """Host buffers with more rhs than one chunk take the overlapped path (chunks of
    8 here: 8, 8, 4, two staging buffers each way); it must give the device-resident
    batch's result bit for bit, keep the per-rhs flip parity of an odd series across
    chunks and across two calls, and fill the norms of every rhs (to rounding: they are
    accumulated by atomics)."""
    from inverse_solver_amd import device as dv
    ri = (16, 16, 16)
    M, A = 7, 20
    p = params(ri, iterations=M)
    RI = sphere_ri(ri, 0.5, n_s=1.42)
    fields = np.stack([random_field(ri, 100 + a) for a in range(A)], axis=-1).astype(np.complex64)
    host = _gpu()(p, options=dict(track_norms=1))
    host.set_RI(RI)
    dev = _gpu()(p, options=dict(track_norms=1))
    dev.set_RI(RI)
    u_abi = np.ascontiguousarray(np.transpose(fields, (3, 2, 1, 0)))       # [rhs][z][y][x]
    d_in = dv.DeviceArray.from_host(u_abi)
    d_out = dv.DeviceArray(u_abi.shape, np.complex64)
    for call in range(2):
        got = host.solve_incident(fields)
        dev.solve_incident_device(d_in.ptr, A, d_out.ptr)
        ref = np.transpose(d_out.to_host(), (3, 2, 1, 0))
        np.testing.assert_array_equal(got, ref)
        # the FP64 norms are summed with one atomic add per wave, in no fixed order
        np.testing.assert_allclose(host.norms(A), dev.norms(A), rtol=1e-12)
        assert np.all(host.norms(A) > 0)
    # two calls of 20 odd-length series leave the flip state where it started, so a third
    # call equals a fresh oracle run over the 20 angles in order (the oracle's flip state
    # also runs on from angle to angle)
    o = Oracle(p, io_kernels=False)
    o.set_RI(RI)
    refs = [o.born_loop(fields[..., a]) for a in range(A)]
    got3 = host.solve_incident(fields)
    for a in (0, 9, 19):
        assert rel(got3[..., a], refs[a]) < TOL, a
    host.close()
    dev.close()


@pytest.mark.parametrize('form', [0, 1, 2, 3])
def test_repeat_set_ri_on_one_handle(form):
    """A new potential on the same geometry keeps the handle's tables (they depend on the
    plan and the geometry scalars only); solves after each set_RI equal fresh handles', and
    a changed boundary sharpness rebuilds them."""
    ri = (16, 16, 16)
    p = params(ri, iterations=9)
    RIs = [sphere_ri(ri, 0.5, n_s=1.40), sphere_ri(ri, 0.35, n_s=1.46, center=(6, 9, 8))]
    u = plane_wave(ri, 0.1, 0.2)
    s = _gpu()(p, options=dict(form=form))
    for k in range(3):
        RI = RIs[k % 2]
        if k == 2:
            s.parameters['boundary_sharpness'] = 0.5
        s.set_RI(RI)
        got = s.solve_incident(u)
        pf = dict(p, boundary_sharpness=0.5) if k == 2 else p
        f = _gpu()(pf, options=dict(form=form))
        f.set_RI(RI)
        np.testing.assert_array_equal(got, f.solve_incident(u))
        f.close()
    s.close()
