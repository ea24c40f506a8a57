#!/usr/bin/env python3
"""Benchmark of the batched-angle convergent Born series forward solve on one MI355X.

Workload (one "step" = one complete solve):
  RI grid 128^3 at 0.1 um plus a 16-pixel absorbing boundary per side -> 160^3
  simulation grid; 8 seeded non-overlapping spheres (radii U[0.5, 2] um,
  n U[1.36, 1.48]) in n_m = 1.336 at lambda = 0.532 um; 49 plane-wave
  illuminations (normal incidence + 48 on a circle at 0.95 * NA 1.2) solved as
  right-hand sides sharing the potential; 200 Born iterations; complex64.
  The incident 3D fields are resident in HBM when the timed region starts.

Metric: Gvoxel-iter/s = 160^3 * 200 * 49 / seconds per solve (higher is better),
with p50/p99 solve latency, the relative L2 error against the FP64 oracle, the
HBM roofline of the dominant kernel (pass_real) and the complex64 CPU port of
the reference loop timed on this host's cores on a bounded sample.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# roofline.traffic is not counted in this run: it is the algorithmic bytes scaled by the
# ratio that the calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE runs of this same command gave
TRAFFIC_SOURCE = ('profiles/pass_real_traffic_160cubed_49rhs.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE '
                  'runs of bench.py on this workload and plan, calibrated on a device copy of known size')
METRIC = 'Gvoxel-iter/s at ≤1e-5 rel err vs FP64 oracle; p50/p99 solve latency'


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def workload_params(ri, iters, A_fields):
    from inverse_solver_amd import get_default_parameters
    return get_default_parameters(dict(
        size=[ri[0], ri[1], ri[2]], wavelength=0.532, NA=1.2, RI_bg=1.336, resolution=[0.1, 0.1, 0.1],
        vector_simulation=False, use_abbe_sine=True, iterations_number=iters,
        boundary_thickness=[8, 8, 8], boundary_sharpness=1, acyclic=True,
        return_transmission=True, return_reflection=False, return_3D=True))


def pass_real_bytes(nvox, A, M, v_bytes=8):
    """Algorithmic HBM bytes of every pass_real launch of one solve (complex64 = 8 B).
    v_bytes: bytes of V per voxel, 4 when the fused form reads a real RI's potential as its
    real part (the imaginary part is -eps everywhere).

    Per launch and voxel: the spectrum is read when the launch completes an
    iteration and written when it starts the next one; f is read by the full
    updates (jj >= 3) and as the source term p0 (jj = 0, 1) and written when the
    next p is gamma*f; Field is read and written by the accumulating updates,
    which from jj = 4 on come in pairs (an even jj below M leaves Field alone and
    the odd jj after it adds both increments); V (shared by the A right-hand
    sides) is read once per voxel."""
    def kind(jj):       # 0 none, 1 set, 2 add, 3 full3, 4 full, 5 defer, 6 pair
        if jj <= 3:
            return jj
        if jj % 2 == 0 and jj < M:
            return 5
        return 6 if (jj % 2 == 1 and jj >= 5) else 4

    def one(jj):
        upd = kind(jj)
        fwd = 0 if jj == M else (1 if jj <= 1 else 2)
        rd = (upd != 0) + (upd >= 3 or fwd == 1) + (upd in (2, 3, 4, 6))
        wr = (upd not in (0, 5)) + (fwd == 2) + (fwd != 0)
        return 8 * A * (rd + wr) + v_bytes
    return nvox * np.array([one(jj) for jj in range(0, M + 1)], dtype=np.float64)


def v_bytes_of(plan, RI):
    """Bytes of V per voxel that the real-space pass reads: 4 in the fused form for a real
    RI (its potential's real part as floats; the imaginary part is -eps everywhere), else 8."""
    return 4 if plan['form'] == 'fused' and np.isrealobj(RI) else 8


def measured_traffic(sim, A, plan, alg_per_launch, v_bytes=8):
    """HBM bytes per pass_real launch from the calibrated FETCH_SIZE/WRITE_SIZE runs
    recorded in profiles/ for this exact workload and plan (None otherwise)."""
    fn = os.path.join(ROOT, 'profiles', 'pass_real_traffic_160cubed_49rhs.json')
    if not os.path.exists(fn) or list(sim) != [160, 160, 160] or A != 49 or plan.get('a') != 40:
        return None
    rec = json.load(open(fn))
    if plan.get('block_order', 0) & 4 != rec.get('block_order', 0) & 4 or v_bytes != rec.get('v_bytes', 8):
        return None     # the block numbering and the form of V change how often V is fetched
    return round(alg_per_launch * rec['measured_over_algorithmic'])


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _line(metric, value, unit, workload, ms, args, extra, hib=True, dtype='complex64'):
    out = {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
           'ms_per_step': round(ms, 3), 'higher_is_better': hib, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': dtype, 'data': 'synthetic', 'config': {'workload': workload}}
    out.update(extra)
    return out


def workload_c1(args):
    """64^3 RI + 16-px boundary (96^3 sim), centred SiO2 bead (n 1.46, 30-voxel
    diameter), one normal-incidence plane wave, 100 iterations: the complete 2D-in /
    2D-out solve on the GPU against the FP64 oracle and the complex64 CPU port."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.phantoms import bead_ri, circle_illumination
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri = [64, 64, 64]
    p = workload_params(ri, 100, 1)
    p.update(return_reflection=True)
    RI = bead_ri(ri, 30, 1.46, 1.336)
    u = circle_illumination(ri[:2], 1, 1.2, 0.532, [0.1] * 3)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p)        # production settings: the iteration loop replays a hipGraph
    s.set_RI(RI)
    if not args.no_tune:
        s.autotune(1, 6)                         # one angle: the plan that fills the device best
    s.solve(u)                                   # kernels, plans, warm-up
    times = []
    for _ in range(max(args.steps, 3)):
        t0 = time.perf_counter()
        t, r, f = s.solve(u)
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    nvox = int(np.prod(s.simulation_size))
    cores = os.cpu_count() or 1
    t0 = time.perf_counter()
    o = Oracle(p, workers=cores)
    o.set_RI(RI)
    tr, rr, fr = o.solve(u)
    t_or = time.perf_counter() - t0
    oc = Oracle(p, dtype=np.complex64, workers=cores)
    oc.set_RI(RI)
    oc.solve(u)
    oc.set_RI(RI)
    t0 = time.perf_counter()
    oc.solve(u)
    t_cpu = time.perf_counter() - t0
    val = nvox * 100 / (ms / 1e3) / 1e9
    return _line('Gvoxel-iter/s at <=1e-5 rel err vs FP64 oracle (full 2D solve, single angle)', round(val, 3),
                 'Gvoxel-iter/s', 'C1: 64^3 RI + 16-px boundary = 96^3 sim, bead n=1.46 d=30 px, 1 angle, '
                 '100 iterations, full 2D in/out solve', ms, args,
                 {'latency_ms': {'p50': ms, 'p99': float(np.percentile(times, 99))},
                  'accuracy': {'rel_l2_err_3d': _rel(f, fr), 'rel_l2_err_trans': _rel(t, tr),
                               'rel_l2_err_ref': _rel(r, rr), 'tolerance': 1e-5},
                  'plan': s.plan(),
                  'cpu_baseline': {'value': round(nvox * 100 / t_cpu / 1e9, 5), 'unit': 'Gvoxel-iter/s',
                                   'cores': cores, 'kind': 'port',
                                   'sample': 'the whole C1 solve (complex64 NumPy port of solve/solve_raw), '
                                             '%.1f s; FP64 oracle %.1f s' % (t_cpu, t_or)}})


def workload_inverse(args):
    """C3: RI 96x96x97 (+16/16/19 px -> 128x128x135 sim), 5 seeded spheres,
    24 circle angles at NA 1.2, measured fields from the GPU forward solve plus
    10 % complex Gaussian noise; BACKWARD_SOLVER_MULTI from the Rytov reconstruction, with
    the device TV prox."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd.backward_solver import BACKWARD_SOLVER_MULTI, get_default_parameters
    from inverse_solver_amd.phantoms import spheres, circle_illumination
    ri = [96, 96, 97]
    bt_xy = 8.0
    bt_z = 19 / (0.532 / 1.336 / 0.2)            # 19 px at 0.1 um
    fp = workload_params(ri, -1, 24)
    fp.update(boundary_thickness=[bt_xy, bt_xy, bt_z], return_reflection=True, return_3D=True)
    RI_true, _ = spheres(tuple(ri), 5, (0.5, 2.0), (1.36, 1.48), seed=11)
    u = circle_illumination(ri[:2], 24, 1.2, 0.532, [0.1] * 3)
    f = FORWARD_SOLVER_CONVERGENT_BORN(fp)
    f.set_RI(RI_true)
    meas, _, _ = f.solve(u)
    sim = f.simulation_size
    f.close()
    rng = np.random.default_rng(12)
    sig = 0.1 * np.sqrt(np.mean(np.abs(meas) ** 2))
    meas = meas + sig * (rng.standard_normal(meas.shape) + 1j * rng.standard_normal(meas.shape)) / np.sqrt(2)
    bp = get_default_parameters(dict(forward_solver_parameters=fp, step=0.01, tv_param=0.001,
                                     inner_itt=args.inner, itter_max=args.outer, verbose=False))
    b = BACKWARD_SOLVER_MULTI(bp)
    b._share_io()          # the 2D projection kernels are built once, as the reference's constructor does
    t0 = time.perf_counter()
    # Initial guess: the Rytov reconstruction from the same fields, as BACKWARD_SOLVER_MULTI
    # starts (:94); --init lowpass uses the phantom low-passed at 70 % contrast instead.
    t_r = time.perf_counter()
    if args.init == 'rytov':
        from inverse_solver_amd.rytov import BACKWARD_SOLVER_RYTOV, get_default_parameters as rytov_defaults
        rp = rytov_defaults(dict(size=list(ri), wavelength=0.532, NA=1.2, RI_bg=1.336, resolution=[0.1] * 3))
        init = np.real(BACKWARD_SOLVER_RYTOV(rp).solve(u, meas)[0])
    else:
        from scipy.ndimage import gaussian_filter
        init = 1.336 + 0.7 * gaussian_filter(np.real(RI_true) - 1.336, 2.0)
    rytov_s = time.perf_counter() - t_r
    RI = b.solve(u, meas, init_RI=init)
    wall = time.perf_counter() - t0
    steady = b.timings[1:] if len(b.timings) > 1 else b.timings    # the first iteration allocates buffers
    g_t = float(np.mean([x[0] for x in steady]))
    tv_t = float(np.mean([x[1] for x in steady]))
    nvox = int(np.prod(sim))
    M = b.forward.Bornmax
    err = float(np.sum(np.abs(RI - RI_true) ** 2) / np.sum(np.abs(1.336 - RI_true) ** 2))
    b.close()
    return _line('seconds per gradient (forward + adjoint CBS over 24 angles)', round(g_t, 4), 's',
                 'C3: 96x96x97 RI -> 128x128x135 sim, 24 angles, %d iterations per solve, %d outer '
                 'iterations, TV inner %d' % (M, args.outer, args.inner), g_t * 1e3, args,
                 {'gradient_gvoxel_iter_per_s': round(2 * 24 * nvox * M / g_t / 1e9, 3),
                  'tv_prox_seconds': round(tv_t, 3), 'wall_seconds': round(wall, 2),
                  'misfit_first_last': [b.err_list[0], b.err_list[-1]],
                  'ri_rel_error_vs_phantom': err,
                  'initial_guess': args.init, 'initial_guess_seconds': round(rytov_s, 2),
                  'ri_rel_error_of_initial_guess': float(np.sum(np.abs(init - RI_true) ** 2)
                                                         / np.sum(np.abs(1.336 - RI_true) ** 2))},
                 hib=False)


def workload_serving(args):
    """C4: Poisson arrivals of forward/adjoint requests; grids RI {64^3, 96^3, 128^3,
    160^2x96, 256^2x128} + 16 px; angles U{1..61}; 1-20 spheres, contrast <= 0.15;
    admission by HBM footprint."""
    from inverse_solver_amd.serving import Scheduler, Request, SolverCache, estimate_footprint, summarize
    from inverse_solver_amd.phantoms import spheres, circle_illumination
    from inverse_solver_amd import device as dv
    rng = np.random.default_rng(21)
    grids = [(64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 96), (256, 256, 128)]
    reqs = []
    t = 0.0
    for i in range(args.requests):
        t += rng.exponential(1.0 / args.rate)
        ri = grids[rng.integers(len(grids))]
        A = int(rng.integers(1, 62))
        ns = int(rng.integers(1, 21))
        RI, _ = spheres(ri, ns, (0.3, 1.2), (1.336, 1.336 + 0.15), seed=1000 + i)
        p = workload_params(list(ri), -1, A)
        p.update(return_reflection=False, return_3D=False)
        sim = [n + 32 for n in ri]
        reqs.append(Request(arrival=t, rid=i, kind='adjoint' if rng.random() < 0.5 else 'forward', params=p,
                            RI=RI, fields=circle_illumination(ri[:2], A, 1.2, 0.532, [0.1] * 3),
                            footprint=estimate_footprint(sim, ri, [2 * ri[0], 2 * ri[1]], A)))
    free, total = dv.mem_info()
    cache = SolverCache(options=dict(max_batch=16))
    seen = {}
    for r in reqs:
        seen.setdefault(SolverCache.key(r.params), r.params)
    cache.prewarm(list(seen.values()), log=log)
    log('serving %d requests' % len(reqs))
    sched = Scheduler(cache, hbm_budget=int(free * 0.8), slots=4)
    done = sched.run(reqs, realtime=True)
    cache.close()

    def vi(r):
        from inverse_solver_amd.params import round_half_away
        sim = [n + 32 for n in r.RI.shape]
        k = 2 * np.pi * 1.336 / 0.532
        vt = np.max(np.abs(k * k * (r.RI ** 2 / 1.336 ** 2 - 1)))
        step = 2 * k / (1.01 * vt)
        M = np.ceil(np.linalg.norm(np.array(sim) * 0.1) / step / 2 + 1) * 2
        return float(np.prod(sim)) * M * r.fields.shape[3]
    summ = summarize(done, vi)
    lat = summ.get('latency_ms', {})
    return _line('serving throughput (Gvoxel-iter/s) with p50/p99 request latency', round(
        summ.get('gvoxel_iter_per_s', 0.0), 3), 'Gvoxel-iter/s',
        'C4: %d Poisson requests at %.1f/s, mixed forward/adjoint, 5 grid sizes, 1-61 angles' % (
            args.requests, args.rate), lat.get('p50', float('nan')), args,
        {'latency_ms': lat, 'queue_ms': summ.get('queue_ms'), 'completed': summ.get('completed'),
         'failed': summ.get('failed'), 'span_s': summ.get('span_s'), 'hbm_budget_gb': round(free * 0.8 / 2 ** 30, 1)})


def workload_large(args):
    """C5: RI 512x512x256 Gaussian random field in an ellipsoid (n in [1.33, 1.40]),
    32-px boundary -> 576x576x320 sim, 30 angles, 300 iterations (Born loop from
    the analytic incident plane waves)."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd import device as dv
    from inverse_solver_amd.phantoms import gaussian_random_field, incident_plane_waves
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri = [512, 512, 256]
    A, M = 30, args.iters if args.iters != 200 else 300
    RI = gaussian_random_field(tuple(ri), 0.5, (1.33, 1.40), seed=31)
    p = workload_params(ri, M, A)
    p.update(boundary_thickness=[16, 16, 16])
    s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(time_kernels=1))
    s.set_RI(RI)
    u = incident_plane_waves(ri, A, 1.2, 0.532, 1.336, [0.1] * 3)
    u_host = np.ascontiguousarray(np.transpose(u, (3, 2, 1, 0)))
    del u
    d_in = dv.DeviceArray.from_host(u_host)
    d_out = dv.DeviceArray(u_host.shape, np.complex64)
    t0 = time.perf_counter()
    s.solve_incident_device(d_in.ptr, A, d_out.ptr)
    dv.synchronize()
    sec = time.perf_counter() - t0
    nvox = int(np.prod(s.simulation_size))
    val = nvox * M * A / sec / 1e9
    kt = s.kernel_times()
    kinds = {}
    for k, (ms_, n_) in kt.items():
        if n_:
            kinds[k] = {'avg_launch_ms': round(ms_ / n_, 4), 'launches': int(n_)}
    # per-launch algorithmic bytes of the four-pass form: real-space pass 48 B/voxel/rhs (+V),
    # each y transform 16 B, the k-space pass 16 B
    if 'pass_real' in kinds:
        kinds['pass_real']['GB_s'] = round(float(pass_real_bytes(nvox, A, M).sum()) / kt['pass_real'][0] / 1e6, 1)
    if 'pass_kspace' in kinds:
        kinds['pass_kspace']['GB_s'] = round(16.0 * A * nvox * kt['pass_kspace'][1] / kt['pass_kspace'][0] / 1e6, 1)
    # the per-axis plain form on the same grid, 1 angle x 10 iterations, for comparison
    pp = workload_params(ri, 10, 1)
    pp.update(boundary_thickness=[16, 16, 16])
    spn = FORWARD_SOLVER_CONVERGENT_BORN(pp, options=dict(form=3, use_graph=0))
    spn.set_RI(RI)
    spn.solve_incident_device(d_in.ptr, 1, d_out.ptr)
    dv.synchronize()
    t1 = time.perf_counter()
    spn.solve_incident_device(d_in.ptr, 1, d_out.ptr)
    dv.synchronize()
    plain_val = nvox * 10 / (time.perf_counter() - t1) / 1e9
    spn.close()
    # parity sample: 1 angle x 6 iterations against the FP64 oracle
    pc = workload_params(ri, 6, 1)
    pc.update(boundary_thickness=[16, 16, 16])
    sg = FORWARD_SOLVER_CONVERGENT_BORN(pc)
    sg.set_RI(RI)
    u0 = u_host[0].transpose(2, 1, 0)
    got = sg.solve_incident(u0)
    sg.close()
    o = Oracle(pc, io_kernels=False, workers=os.cpu_count())
    o.set_RI(RI)
    err = _rel(got, o.born_loop(u0))
    return _line('Gvoxel-iter/s (large volume)', round(val, 3), 'Gvoxel-iter/s',
                 'C5: 512x512x256 RI + 32-px boundary = 576x576x320 sim, 30 angles, %d iterations' % M,
                 sec * 1e3, args, {'plan': s.plan(), 'kernels': kinds,
                                   'plain_form_single_angle': {'value': round(plain_val, 3), 'unit': 'Gvoxel-iter/s'},
                                   'accuracy': {'rel_l2_err_vs_fp64_oracle': err, 'angles': 1, 'iterations': 6}})


def hbm_ceiling(stream, log, gib=2, reps=10):
    """Practical HBM ceilings on this device, in the same run: the nontemporal 16-byte
    streaming probe (inverse_solver_amd/csrc/stream_probe.hip, libcbs_probe.so) over
    arrays of `gib` GiB -- a copy (1 read + 1 write stream) and a 3-read/2-write mix --
    best of three runs of `reps` launches, HIP events on the solver's stream; and the
    runtime's own hipMemcpy device-to-device (host timer around the synchronous copy).
    Rates count bytes read + written."""
    import ctypes as C
    from inverse_solver_amd import device as dv
    lib = C.CDLL(os.path.join(ROOT, 'inverse_solver_amd', 'libcbs_probe.so'))
    f = lib.cbs_probe_stream
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p,
                  C.POINTER(C.c_float)]
    nbytes = gib << 30
    arrs = [dv.DeviceArray((nbytes // 8,), np.complex64) for _ in range(3)]
    for d in arrs:
        d.upload(np.zeros(nbytes // 8, np.complex64))
    res = {}
    for kind, name, streams in ((0, 'stream_copy', 2), (1, 'stream_3r2w', 5)):
        best = 1e30
        for _ in range(3):
            ms = C.c_float(0.0)
            rc = f(kind, arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, nbytes, reps, C.c_void_p(stream), C.byref(ms))
            if rc != 0:
                raise RuntimeError('cbs_probe_stream(%d) failed: %d' % (kind, rc))
            best = min(best, ms.value)
        gbs = streams * nbytes / (best / 1e3) / 1e9
        res[name] = {'GB_s': round(gbs, 1), 'frac_of_peak': round(gbs / HBM_PEAK_GBS, 4), 'ms': round(best, 4)}
    arrs[1].copy_from(arrs[0])
    dv.synchronize()
    best = 1e30
    for _ in range(5):
        t0 = time.perf_counter()
        arrs[1].copy_from(arrs[0])
        dv.synchronize()
        best = min(best, time.perf_counter() - t0)
    gbs = 2.0 * nbytes / best / 1e9
    res['hipmemcpy_d2d'] = {'GB_s': round(gbs, 1), 'frac_of_peak': round(gbs / HBM_PEAK_GBS, 4)}
    for d in arrs:
        d.free()
    res['what'] = ('%d GiB arrays; stream_*: nontemporal 16-byte loads/stores, 256-thread blocks, best of 3 x %d '
                   'launches by HIP events; hipmemcpy_d2d: best of 5 synchronous copies' % (gib, reps))
    log('HBM ceilings: stream copy %.0f, stream 3R2W %.0f, hipMemcpy %.0f GB/s'
        % (res['stream_copy']['GB_s'], res['stream_3r2w']['GB_s'], gbs))
    return res


def pattern_ceilings(stream, log, dims, reps=10, v_float=False):
    """The fused passes' own access patterns with no arithmetic (cbs_probe_pattern in
    inverse_solver_amd/csrc/stream_probe.hip, libcbs_probe.so), on arrays of the solve's
    size: the k-space pass's in-place 128-byte row segments, and the real-space pass's tile
    streams of an even launch (spec, V, f) and an odd one (with Field), blocks in storage
    order.  Best of 3 x `reps` launches by HIP events on `stream`; rates count the
    algorithmic bytes of the pass (V once per launch).  dims = [nx, ny, nz, a, b, nrhs];
    v_float: V read as one float per voxel, as the pass reads a real RI's potential."""
    import ctypes as C
    from inverse_solver_amd import device as dv
    lib = C.CDLL(os.path.join(ROOT, 'inverse_solver_amd', 'libcbs_probe.so'))
    f = lib.cbs_probe_pattern
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_void_p,
                  C.POINTER(C.c_float)]
    nx, ny, nz, a, b, nrhs = [int(v) for v in dims]
    nvox = nx * ny * nz
    arrs = [dv.DeviceArray((nrhs * nvox,), np.complex64) for _ in range(3)] + [dv.DeviceArray((nvox,), np.complex64)]
    for d in arrs:
        dv._chk(dv.hip.hipMemset(C.c_void_p(d.ptr), 0, C.c_size_t(d.nbytes)), 'hipMemset')
    cd = (C.c_int * 6)(nx, ny, nz, a, b, nrhs)
    res = {}
    vb = 4.0 if v_float else 8.0
    ke, ko = (3, 4) if v_float else (1, 2)
    for kind, name, nbytes in ((0, 'kspace', 16.0 * nrhs * nvox), (ke, 'real_even', (32.0 * nrhs + vb) * nvox),
                               (ko, 'real_odd', (48.0 * nrhs + vb) * nvox)):
        best = 1e30
        for _ in range(3):
            ms = C.c_float(0.0)
            rc = f(kind, arrs[0].ptr, arrs[3].ptr, arrs[1].ptr, arrs[2].ptr, cd, reps, C.c_void_p(stream), C.byref(ms))
            if rc != 0:
                raise RuntimeError('cbs_probe_pattern(%d) failed: %d' % (kind, rc))
            best = min(best, ms.value)
        res[name] = {'ms': round(best, 4), 'GB_s': round(nbytes / (best / 1e3) / 1e9, 1)}
    for d in arrs:
        d.free()
    res['what'] = ('the passes\' access patterns without their arithmetic on %dx%dx%d x %d rhs (a %d, b %d): '
                   'k-space 128-byte in-place row segments, cached loads and nontemporal stores; real-space '
                   'storage-order tiles, nontemporal streams and cached V (%d bytes a voxel)'
                   % (nx, ny, nz, nrhs, a, b, int(vb)))
    log('pattern ceilings: k-space %.3f ms (%.0f GB/s), real-space even %.3f ms (%.0f GB/s), odd %.3f ms (%.0f GB/s)'
        % (res['kspace']['ms'], res['kspace']['GB_s'], res['real_even']['ms'], res['real_even']['GB_s'],
           res['real_odd']['ms'], res['real_odd']['GB_s']))
    return res


def workload_demo(args):
    """The reference's own demo geometry (SIMULATION_BEAD.m:11-19, 48): a 161x161x61 RI
    grid at 0.532/4/1.2 um holding a 5-um bead (n 1.4609 in 1.336).  With the default
    boundary_thickness 6 it is a 183x183x83 simulation grid (183 = 3 x 61 and the prime
    83: Bluestein transforms on every axis); the backward solver's [2 2 4] gives
    169x169x75 (169 = 13^2 on run-time Stockham stages, 75 compiled).  12 circle angles,
    20 iterations, the batched per-axis form, inputs and outputs in HBM.  Beside it: the
    reference's GPU loop restated op for op on rocFFT (one angle), and the FP64 oracle
    at 6 iterations on one angle."""
    import torch
    torch.zeros(1, device='cuda:0')                 # torch's HIP context first (reference-form loop)
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN, get_default_parameters
    from inverse_solver_amd import device as dv
    from inverse_solver_amd.phantoms import bead_ri, incident_plane_waves
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri, res, A, M = [161, 161, 61], 0.532 / 4 / 1.2, 12, 20
    RI = bead_ri(ri, 5.0 / res, 1.4609, 1.336)

    def prm(bt, iters):
        return get_default_parameters(dict(
            size=ri, wavelength=0.532, NA=1.2, RI_bg=1.336, resolution=[res] * 3, vector_simulation=False,
            use_abbe_sine=True, iterations_number=iters, boundary_thickness=bt, boundary_sharpness=1,
            acyclic=True, return_transmission=True, return_reflection=False, return_3D=True))
    u = incident_plane_waves(ri, A, 1.2, 0.532, 1.336, [res] * 3)
    u_host = np.ascontiguousarray(np.transpose(u, (3, 2, 1, 0))).astype(np.complex64)
    d_in = dv.DeviceArray.from_host(u_host)
    d_out = dv.DeviceArray(u_host.shape, np.complex64)
    out, rates = {}, {}
    for name, bt in (('forward', [6, 6, 6]), ('backward', [2, 2, 4])):
        s = FORWARD_SOLVER_CONVERGENT_BORN(prm(bt, M))
        s.set_RI(RI)
        s.solve_incident_device(d_in.ptr, A, d_out.ptr)
        dv.synchronize()
        times = []
        for _ in range(max(args.steps, 3)):
            t0 = time.perf_counter()
            s.solve_incident_device(d_in.ptr, A, d_out.ptr)
            dv.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        ms = float(np.median(times))
        sim = list(s.simulation_size)
        rates[name] = {'sim_grid': sim, 'value': round(float(np.prod(sim)) * M * A / (ms / 1e3) / 1e9, 3),
                       'unit': 'Gvoxel-iter/s', 'ms_per_solve': round(ms, 3), 'plan': s.plan()['form']}
        if name == 'forward':
            fwd_ms, fwd_sim = ms, sim
            ref_ms, _ = reference_form_gpu(s, u[..., 0], M, lambda *a: print(*a, file=sys.stderr, flush=True))
            out['reference_form_single_angle'] = {
                'value': round(float(np.prod(sim)) * M / (ref_ms / 1e3) / 1e9, 3), 'unit': 'Gvoxel-iter/s',
                'ms_per_angle': round(ref_ms, 3)}
        s.close()
    # accuracy: one angle, 6 iterations, against the FP64 oracle
    p6 = prm([6, 6, 6], 6)
    s = FORWARD_SOLVER_CONVERGENT_BORN(p6)
    s.set_RI(RI)
    got = s.solve_incident(u[..., :1])[..., 0]
    s.close()
    o = Oracle(p6, io_kernels=False, workers=os.cpu_count() or 1)
    o.set_RI(RI)
    out['accuracy'] = {'rel_l2_err_vs_fp64_oracle': _rel(got, o.born_loop(u[..., 0])), 'angles': 1, 'iterations': 6,
                       'grid': fwd_sim}
    d_in.free()
    d_out.free()
    out['geometries'] = rates
    return _line('Gvoxel-iter/s (reference demo geometry, per-axis form)', rates['forward']['value'], 'Gvoxel-iter/s',
                 'SIMULATION_BEAD: 161x161x61 RI + default boundary = 183x183x83 sim, 5-um bead, %d angles, '
                 '%d iterations' % (A, M), fwd_ms, args, out)


def workload_ceiling(args):
    """The practical HBM ceilings alone (hbm_ceiling on the null stream): what the
    passes' achieved GB/s are compared against."""
    lg = lambda *a: print(*a, file=sys.stderr, flush=True)
    c = hbm_ceiling(None, lg)
    pc = pattern_ceilings(None, lg, [160, 160, 160, 40, 4, 49], v_float=True)
    return _line('streaming HBM ceiling (3 read + 2 write streams)', c['stream_3r2w']['GB_s'], 'GB/s',
                 'nontemporal 16-byte streaming probe over 2 GiB arrays', c['stream_3r2w']['ms'], args,
                 {'hbm_ceiling': c, 'pattern_ceiling': pc}, dtype='f32')


def reference_form_gpu(s, u0, M, log_fn):
    """The reference's own GPU loop, op for op: FORWARD_SOLVER_CONVERGENT_BORN.solve_raw
    (:577-647) as MATLAB runs it on gpuArrays for one angle -- one device kernel per
    array expression, 3D FFTs from the vendor library (here torch.fft on rocFFT),
    fft_flip of Greenp and conj of the ramp every iteration, complex64.  A
    measurement baseline only (the product never runs it).  Tables follow
    :251-271 (mask), :347-351,384 (Greenp), :355-377 (ramp)."""
    import torch
    dev = torch.device('cuda:0')
    sim = list(s.simulation_size)
    b = [int(x) for x in s.boundary_thickness_pixel]
    p = s.parameters
    eps = float(s.eps_imag)
    # arrays held as (z, y, x) so x is the fastest axis, as in MATLAB's column-major order;
    # V formed on the host as set_RI does in MATLAB (:205-215, :238-241, RI2potential.m)
    from inverse_solver_amd.params import ri_to_potential
    Vt = ri_to_potential(np.asarray(s.RI).astype(np.complex128), p['wavelength'], p['RI_bg'])
    Vh = (np.pad(Vt, [(b[0],) * 2, (b[1],) * 2, (b[2],) * 2]) - 1j * eps).astype(np.complex64)
    V = torch.from_numpy(np.ascontiguousarray(Vh.T)).to(dev)
    def axis_shape(j):
        shp = [1, 1, 1]
        shp[2 - j] = -1
        return shp
    mask = torch.ones((1, 1, 1), dtype=torch.float32, device=dev)
    ramp = torch.ones((1, 1, 1), dtype=torch.complex64, device=dev)
    g = None
    fres = [1.0 / (p['resolution'][j] * sim[j]) for j in range(3)]
    k0 = p['RI_bg'] / p['wavelength']
    ks = []
    for j in range(3):
        n = sim[j]
        x = np.abs(np.arange(1, n + 1) - (np.floor(n / 2 + 1)) + 0.5) - 0.5
        x = np.roll(x, -(n // 2))
        if b[j] > 0:
            v0 = np.minimum(np.abs(x / (b[j] - 0.5)), 1.0)
            v0 = np.minimum(1 - v0, p['boundary_sharpness']) / p['boundary_sharpness']
        else:
            v0 = np.zeros(n)
        mask = mask * torch.from_numpy((1 - v0).astype(np.float32)).to(dev).reshape(axis_shape(j))
        r = np.exp(-1j * np.pi * np.arange(n) / n / 2)
        r = r / r[n // 2]
        ramp = ramp * torch.from_numpy(r.astype(np.complex64)).to(dev).reshape(axis_shape(j))
        ks.append(((np.arange(n) - n // 2) * fres[j] + fres[j] / 4).reshape(axis_shape(j)))
    g = 1.0 / (4 * np.pi ** 2 * (ks[0] ** 2 + ks[1] ** 2 + ks[2] ** 2) - (2 * np.pi * k0) ** 2 - 1j * eps)
    Greenp = torch.from_numpy(np.fft.ifftshift(g).astype(np.complex64)).to(dev)
    R = [int(v) for v in s.ROI]
    roi = (slice(R[4] - 1, R[5]), slice(R[2] - 1, R[3]), slice(R[0] - 1, R[1]))
    inc = torch.from_numpy(np.ascontiguousarray(np.asarray(u0).T).astype(np.complex64)).to(dev)
    dims = (0, 1, 2)

    def run():
        nonlocal Greenp, ramp
        psi = torch.zeros_like(V)
        Field = torch.zeros_like(V)
        Field_n = torch.zeros_like(V)
        source = (V[roi] + 1j * eps) * inc
        temp = None
        for jj in range(1, M + 1):
            Greenp = torch.roll(torch.flip(Greenp, dims), (1, 1, 1), dims)     # fft_flip(Greenp,[1 1 1],false)
            ramp = torch.conj(ramp)
            psi.zero_()
            if jj in (1, 2):
                Field_n.zero_()
                psi[roi] = (1j / eps) * source / 2
            else:
                psi = (1j / eps) * Field_n * V
            PSI = Greenp * torch.fft.fftn(psi * ramp, dim=dims)
            PSI = torch.fft.ifftn(PSI, dim=dims) * torch.conj(ramp)
            if jj not in (1, 2):
                Field_n = Field_n - psi
            Field_n = Field_n + V * PSI
            Field_n = Field_n * mask
            if jj == 2:
                temp = Field.clone()
            Field = Field + Field_n
            if jj == 3:
                Field_n = Field_n + temp
        return Field[roi] + inc

    G0, r0 = Greenp.clone(), ramp.clone()
    run()                                   # warm-up: FFT plans
    torch.cuda.synchronize()
    Greenp, ramp = G0, r0
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    out = run()
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    log_fn('reference-form GPU loop (torch/rocFFT, 1 angle): %.1f ms' % ms)
    return ms, out.cpu().numpy()


def workload_retrieval(args):
    """Experimental input pipeline (FIELD_EXPERIMENTAL_RETRIEVAL.get_fields) on synthetic
    off-axis holograms: 49 background + 49 sample images of 512^2 pixels (a phase bump of
    7 rad behind tilted plane waves, reference carrier at a quarter of the sampling rate),
    sideband filtering on the host, unwrapping on the GPU (cbs_phase_unwrap)."""
    from inverse_solver_amd.field_retrieval import FIELD_EXPERIMENTAL_RETRIEVAL, get_default_parameters
    n, A = 512, 49
    i = np.arange(n)
    x = i - n // 2
    phi = 7.0 * np.exp(-(x[:, None] ** 2 + (x[None, :] + 5) ** 2) / (2 * 44.0 ** 2))
    plane = lambda k: np.exp(2j * np.pi * (k[0] * i[:, None] + k[1] * i[None, :]) / n)   # noqa: E731
    R = plane((n // 4, n // 4))
    rng = np.random.default_rng(4)
    tilts = [(0, 0)] + [tuple(rng.integers(-20, 21, 2)) for _ in range(A - 1)]
    bg = np.stack([100 * np.abs(R + 0.5 * plane(t)) ** 2 for t in tilts], -1)
    sp = np.stack([100 * np.abs(R + 0.5 * plane(t) * np.exp(1j * phi)) ** 2 for t in tilts], -1)
    p = get_default_parameters(dict(size=[n, n, 1], resolution=[0.1] * 3, resolution_image=[0.1, 0.1]))
    FIELD_EXPERIMENTAL_RETRIEVAL(dict(p)).get_fields(bg[:, :, :2], sp[:, :, :2])      # warm-up
    times = []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        inp, out, _, k0s, _, _ = FIELD_EXPERIMENTAL_RETRIEVAL(dict(p)).get_fields(bg, sp)
        times.append(time.perf_counter() - t0)
    sec = float(np.median(times))
    ratio = out[:, :, 0, :] / inp[:, :, 0, :]
    err = _rel(ratio, np.broadcast_to(np.exp(1j * phi)[2:-2, 2:-2, None], ratio.shape))
    return _line('holograms per second (off-axis field retrieval)', round(A / sec, 2), 'holograms/s',
                 'FIELD_EXPERIMENTAL_RETRIEVAL.get_fields: %d background + %d sample holograms of %d^2 px' % (A, A, n),
                 sec * 1e3, args, {'accuracy': {'rel_l2_err_ratio_vs_object': err}}, dtype='complex128')


def workload_vector(args):
    """Vector (dyadic Green's function) Born loop: 128^3 RI + 16 px = 160^3 sim, 8 spheres
    (seed 2), p-polarised plane waves for 16 circle angles, 200 iterations; inputs
    resident in HBM; the same unit as the scalar line (voxel-iterations, each voxel
    carrying the three field components)."""
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd import device as dv
    from inverse_solver_amd.phantoms import spheres, incident_plane_waves, circle_delta_positions
    from oracle.cbs import ForwardSolverConvergentBorn as Oracle
    ri = list(args.ri)
    A = 16 if args.angles == 49 else args.angles
    M = args.iters
    RI, _ = spheres(tuple(ri), 8, (0.5, 2.0), (1.36, 1.48), seed=args.seed)
    u = incident_plane_waves(ri, A, 1.2, 0.532, 1.336, [0.1, 0.1, 0.1])
    # p polarisation: unit vector in the plane of incidence, perpendicular to k
    k0 = 1.336 / 0.532
    pol = np.zeros((3, A))
    for m, (p1, p2) in enumerate(circle_delta_positions(ri[:2], A, 1.2, 0.532, [0.1] * 3, 0.95)):
        kx = -(p1 - ri[0] // 2) / (0.1 * ri[0])
        ky = -(p2 - ri[1] // 2) / (0.1 * ri[1])
        kz = np.sqrt(max(k0 * k0 - kx * kx - ky * ky, 0.0))
        rho = np.hypot(kx, ky)
        c, sn = (kx / rho, ky / rho) if rho > 0 else (1.0, 0.0)
        pol[:, m] = [kz / k0 * c, kz / k0 * sn, -rho / k0]
    u3 = (u[:, :, :, None, :] * pol[None, None, None, :, :]).astype(np.complex64)   # (x, y, z, 3, A)
    del u
    p = workload_params(ri, M, A)
    p['vector_simulation'] = True
    s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(time_kernels=1))
    s.set_RI(RI)
    if not args.no_tune:
        s.autotune(A, 6)
    nvox = int(np.prod(s.simulation_size))
    u_host = np.ascontiguousarray(np.transpose(u3, (4, 3, 2, 1, 0)))
    d_in = dv.DeviceArray.from_host(u_host)
    d_out = dv.DeviceArray(u_host.shape, np.complex64)
    stream = s.stream()
    for _ in range(args.warmup):
        s.solve_incident_device(d_in.ptr, A, d_out.ptr)
    dv.synchronize()
    e0, e1 = dv.Event(), dv.Event()
    step_ms, pk = [], [0.0, 0.0, 0.0, 0.0]
    for k in range(args.steps):
        e0.record(stream)
        s.solve_incident_device(d_in.ptr, A, d_out.ptr)
        e1.record(stream)
        e1.synchronize()
        step_ms.append(e0.elapsed_ms(e1))
        kt = s.kernel_times()
        pk[0] += kt['pass_real'][0]; pk[1] += kt['pass_real'][1]
        pk[2] += kt['pass_kspace'][0]; pk[3] += kt['pass_kspace'][1]
        log('vector step', k + 1, '%.1f ms' % step_ms[-1])
    ms = float(np.mean(step_ms))
    pr_ms, pk_ms = pk[0] / max(pk[1], 1), pk[2] / max(pk[3], 1)
    vb = v_bytes_of(s.plan(), RI)
    bytes_pr = float(pass_real_bytes(nvox, 3 * A, M, vb).sum()) / (pk[1] / args.steps)   # 3A component slots share V
    bytes_pk = 16.0 * 3 * A * nvox
    out = d_out.to_host()
    # parity sample: angle 0, 8 iterations, against the FP64 oracle
    pc = workload_params(ri, 8, 1)
    pc['vector_simulation'] = True
    sg = FORWARD_SOLVER_CONVERGENT_BORN(pc)
    sg.set_RI(RI)
    u0 = u_host[0].transpose(3, 2, 1, 0)
    got = sg.solve_incident(u0)
    sg.close()
    o = Oracle(pc, io_kernels=False, workers=os.cpu_count())
    o.set_RI(RI)
    err = _rel(got, o.born_loop(u0))
    plan = s.plan()
    s.close()
    return _line('Gvoxel-iter/s (vector CBS, 3 field components per voxel)', round(nvox * M * A / (ms / 1e3) / 1e9, 3),
                 'Gvoxel-iter/s', 'vector Born loop: %dx%dx%d RI + 16-px boundary = %s sim, %d p-polarised plane '
                 'waves, %d iterations' % (ri[0], ri[1], ri[2], s.simulation_size, A, M), ms, args,
                 {'plan': plan, 'finite': bool(np.isfinite(out).all()),
                  'kernels': {'pass_real': {'avg_launch_ms': round(pr_ms, 4),
                                            'GB_s': round(bytes_pr / (pr_ms / 1e3) / 1e9, 1)},
                              'pass_kspace_vector': {'avg_launch_ms': round(pk_ms, 4),
                                                     'GB_s': round(bytes_pk / (pk_ms / 1e3) / 1e9, 1)}},
                  'accuracy': {'rel_l2_err_vs_fp64_oracle': err, 'angles': 1, 'iterations': 8}})


def main():
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--ri', type=int, nargs=3, default=[128, 128, 128])
    ap.add_argument('--angles', type=int, default=49)
    ap.add_argument('--iters', type=int, default=200)
    ap.add_argument('--seed', type=int, default=2)
    ap.add_argument('--check-depths', type=int, nargs='*', default=[24, 100],
                    help='series depths besides the full one at which the FP64-oracle check is stated '
                         '(1 angle, same grid and plan)')
    ap.add_argument('--cpu-iters', type=int, default=250,
                    help='iterations of the complex64 CPU-port sample (1 angle, same grid)')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline and the oracle check')
    ap.add_argument('--no-plain', action='store_true', help='skip the plain single-angle comparison')
    ap.add_argument('--no-refform', action='store_true',
                    help='skip the reference-form GPU loop (torch/rocFFT restatement of solve_raw)')
    ap.add_argument('--no-tune', action='store_true', help='use the planner default instead of the autotuner')
    ap.add_argument('--no-whole', action='store_true',
                    help='skip the whole 2D-in/2D-out solve (cbs_solve_2d) timing')
    ap.add_argument('--whole-steps', type=int, default=3, help='timed whole 2D-in/2D-out solves')
    ap.add_argument('--copy-probe-mb', type=int, default=0,
                    help='run one device-to-device copy of this many MiB first (calibrates HBM byte counters)')
    ap.add_argument('--workload', default='batched',
                    choices=['batched', 'c1', 'inverse', 'serving', 'large', 'vector', 'retrieval', 'ceiling', 'demo'],
                    help='batched: the headline 49-angle solve (default); the others are the secondary '
                         'configurations, each printing its own JSON line')
    ap.add_argument('--requests', type=int, default=200, help='serving: number of requests')
    ap.add_argument('--rate', type=float, default=4.0, help='serving: Poisson arrival rate (requests/s)')
    ap.add_argument('--outer', type=int, default=20, help='inverse: outer iterations')
    ap.add_argument('--inner', type=int, default=100, help='inverse: TV inner iterations')
    ap.add_argument('--init', default='rytov', choices=['rytov', 'lowpass'],
                    help='inverse: initial guess (Rytov reconstruction, or the low-passed phantom)')
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({'error': 'bench.py runs on one GPU; --gpus %d is not supported (the angle-parallel '
                                   'sharding across GPUs is designed but not built)' % args.gpus}))
        sys.exit(2)

    if args.workload != 'batched':
        out = {'c1': workload_c1, 'inverse': workload_inverse, 'serving': workload_serving,
               'large': workload_large, 'vector': workload_vector,
               'retrieval': workload_retrieval, 'ceiling': workload_ceiling,
               'demo': workload_demo}[args.workload](args)
        print(json.dumps(out), flush=True)
        return

    if not args.no_refform:
        # the reference-form baseline runs on torch; its HIP context is created before the
        # solver library's (in the other order torch finds no device on these machines)
        import torch
        torch.zeros(1, device='cuda:0')
    from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
    from inverse_solver_amd import device as dv
    from inverse_solver_amd.phantoms import spheres, incident_plane_waves

    ri = list(args.ri)
    A, M = args.angles, args.iters
    log('building phantom', ri, 'angles', A, 'iterations', M)
    RI, _ = spheres(tuple(ri), 8, (0.5, 2.0), (1.36, 1.48), seed=args.seed)
    u = incident_plane_waves(ri, A, 1.2, 0.532, 1.336, [0.1, 0.1, 0.1])
    p = workload_params(ri, M, A)
    # complex64 bulk with FP64 residuals: the per-iteration |f_n|^2 of every rhs is
    # accumulated in FP64 by the timed kernels; HIP events around every pass
    s = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(time_kernels=1, track_norms=1))
    s.set_RI(RI)
    sim = s.simulation_size
    nvox = int(np.prod(sim))
    if not args.no_tune:
        t0 = time.perf_counter()
        s.autotune(A, 6)
        log('autotune %.1f s' % (time.perf_counter() - t0))
    plan = s.plan()
    log('sim grid', sim, 'plan', plan, 'device', dv.device_name(0))

    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 8
        ca, cb = dv.DeviceArray((n,), np.complex64), dv.DeviceArray((n,), np.complex64)
        cb.copy_from(ca)
        dv.synchronize()
        log('copy probe: %d MiB device-to-device' % args.copy_probe_mb)
        ca.free()
        cb.free()

    # inputs resident in HBM: per-rhs x-fastest blocks
    u_host = np.ascontiguousarray(np.transpose(u, (3, 2, 1, 0)))
    d_in = dv.DeviceArray.from_host(u_host)
    d_out = dv.DeviceArray(u_host.shape, np.complex64)
    stream = s.stream()

    for w in range(args.warmup):
        s.solve_incident_device(d_in.ptr, A, d_out.ptr)
        log('warmup', w + 1, 'done')
    dv.synchronize()

    step_ms = []
    pr_ms = pr_n = pk_ms = pk_n = 0.0
    e0, e1 = dv.Event(), dv.Event()
    for k in range(args.steps):
        e0.record(stream)
        s.solve_incident_device(d_in.ptr, A, d_out.ptr)
        e1.record(stream)
        e1.synchronize()
        step_ms.append(e0.elapsed_ms(e1))
        kt = s.kernel_times()
        pr_ms += kt['pass_real'][0]
        pr_n += kt['pass_real'][1]
        pk_ms += kt['pass_kspace'][0]
        pk_n += kt['pass_kspace'][1]
        log('step', k + 1, '%.1f ms' % step_ms[-1])
    dv.synchronize()
    ms = float(np.mean(step_ms))
    vox_iters = float(nvox) * M * A
    value = vox_iters / (ms / 1e3) / 1e9
    p50, p99 = float(np.percentile(step_ms, 50)), float(np.percentile(step_ms, 99))

    # roofline of the dominant kernel
    vb = v_bytes_of(plan, RI)
    bytes_pr = float(pass_real_bytes(nvox, A, M, vb).sum())   # per solve
    pr_avg_ms = pr_ms / max(pr_n, 1)
    pr_launches = pr_n / args.steps
    achieved_pr = bytes_pr / pr_launches / (pr_avg_ms / 1e3) / 1e9
    bytes_pk_launch = 16.0 * A * nvox
    pk_avg_ms = pk_ms / max(pk_n, 1)
    achieved_pk = bytes_pk_launch / (pk_avg_ms / 1e3) / 1e9
    log('pass_real avg %.3f ms  %.0f GB/s; pass_kspace avg %.3f ms  %.0f GB/s'
        % (pr_avg_ms, achieved_pr, pk_avg_ms, achieved_pk))

    result_gpu = d_out.to_host()
    nrm = s.norms(A)           # FP64 |f_n|^2 per rhs and iteration, accumulated in the timed solves
    out = {
        'metric': METRIC, 'value': round(value, 3), 'unit': 'Gvoxel-iter/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(ms, 3), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'complex64', 'data': 'synthetic',
        'config': {'workload': 'batched-angle CBS forward solve: %dx%dx%d RI + %d-px boundary = %dx%dx%d sim, '
                               '%d plane-wave RHS sharing V, %d iterations, 8 seeded spheres'
                               % (ri[0], ri[1], ri[2], (sim[0] - ri[0]) // 2, sim[0], sim[1], sim[2], A, M),
                   'sim_grid': sim, 'ri_grid': ri, 'n_rhs': A, 'iterations': M, 'seed': args.seed,
                   'plan': plan, 'parallelism': 'single GPU, all angles in each launch'},
        'latency_ms': {'p50': round(p50, 3), 'p99': round(p99, 3)},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved_pr, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved_pr / HBM_PEAK_GBS, 4), 'traffic': measured_traffic(sim, A, plan, bytes_pr / pr_launches, vb),
                     'traffic_source': TRAFFIC_SOURCE,
                     'kernel': 'pass_real', 'avg_launch_ms': round(pr_avg_ms, 4),
                     'algorithmic_bytes_per_launch': round(bytes_pr / pr_launches),
                     'pass_kspace': {'achieved': round(achieved_pk, 1), 'avg_launch_ms': round(pk_avg_ms, 4),
                                     'algorithmic_bytes_per_launch': bytes_pk_launch}},
        'fp64_residuals': {'tracked_in_timed_region': True,
                           'last_term_over_first_max': float(np.sqrt(np.max(nrm[:, -1] / nrm[:, 0])))},
    }

    if not args.no_cpu:
        # accuracy of the timed result itself (rhs 0 of the last timed solve: the autotuned
        # plan, all M iterations) and of the same handle at shorter depths, against one
        # FP64 oracle run that records the series at each depth
        from oracle.cbs import ForwardSolverConvergentBorn as Oracle
        cores = os.cpu_count() or 1
        u0 = u[..., 0]
        depths = sorted(set([d for d in args.check_depths if 0 < d < M] + [M]))
        snaps = {d: None for d in depths}
        t0 = time.perf_counter()
        o = Oracle(workload_params(ri, M, 1), io_kernels=False, workers=cores)
        o.set_RI(RI)
        o.born_loop(u0, snapshots=snaps)
        t_or = time.perf_counter() - t0
        errs = {M: _rel(result_gpu[0].T, snaps[M])}
        for d in depths[:-1]:
            s.parameters['iterations_number'] = d
            s.set_RI(RI)                    # same handle: the autotuned plan stays
            errs[d] = _rel(s.solve_incident(u0), snaps[d])
        s.parameters['iterations_number'] = M
        s.set_RI(RI)
        err = errs[M]
        out['accuracy'] = {'rel_l2_err_vs_fp64_oracle': err, 'angles': 1, 'iterations': M,
                           'by_iterations': {str(d): errs[d] for d in depths},
                           'what': 'rhs 0 of the timed solve (autotuned plan, all %d iterations) and the same '
                                   'handle at the shorter depths, against the FP64 oracle' % M,
                           'grid': sim, 'tolerance': 1e-5, 'pass': all(e <= 1e-5 for e in errs.values()),
                           'oracle_seconds': round(t_or, 2)}
        log('accuracy rel err %s (oracle %.1f s)' % ({d: '%.2e' % e for d, e in errs.items()}, t_or))

    if not args.no_whole:
        # the whole 2D-in/2D-out solve on device buffers (cbs_solve_2d: refocus of the 2D
        # input into the 3D incident fields, the Born loop, the projection onto the 2D
        # transmitted fields; 3D fields kept), production settings (graph replay)
        import ctypes as C
        from inverse_solver_amd._lib import lib, check, CBS_MEM_DEVICE
        from inverse_solver_amd.phantoms import circle_illumination
        t0 = time.perf_counter()
        s.set_options(time_kernels=0)
        _, src = s.source_2d(circle_illumination(ri[:2], A, 1.2, 0.532, [0.1] * 3))
        log('I/O kernels and 2D sources %.1f s' % (time.perf_counter() - t0))
        d_src = dv.DeviceArray.from_host(np.ascontiguousarray(src.T))
        d_t = dv.DeviceArray(d_src.shape, np.complex64)
        whole = lambda: check(lib.cbs_solve_2d(s._h, C.c_void_p(d_src.ptr), A, C.c_void_p(d_t.ptr), None,
                                               C.c_void_p(d_out.ptr), CBS_MEM_DEVICE))
        whole()
        wms = []
        for k in range(args.whole_steps):
            e0.record(stream)
            whole()
            e1.record(stream)
            e1.synchronize()
            wms.append(e0.elapsed_ms(e1))
        wm = float(np.mean(wms))
        out['whole_solve_2d'] = {
            'value': round(vox_iters / (wm / 1e3) / 1e9, 3), 'unit': 'Gvoxel-iter/s', 'ms_per_step': round(wm, 3),
            'steps': args.whole_steps,
            'what': 'cbs_solve_2d with 2D inputs and outputs in HBM: refocus (A6), the Born loop, the transmitted '
                    '2D fields (A7) and the 3D fields, %d angles x %d iterations' % (A, M)}
        log('whole 2D solve %.1f ms -> %.2f Gvoxel-iter/s' % (wm, out['whole_solve_2d']['value']))
        d_src.free()
        d_t.free()
    # the ceilings on the solver's stream, before the handle (and its stream) is closed
    out['hbm_ceiling'] = hbm_ceiling(stream, log)
    c = out['hbm_ceiling']['stream_3r2w']['GB_s']
    out['roofline']['frac_of_stream_ceiling'] = round(achieved_pr / c, 4)
    out['roofline']['pass_kspace']['frac_of_stream_ceiling'] = round(achieved_pk / c, 4)
    if plan['form'] == 'fused':
        # each pass against its own access pattern without the arithmetic (pass_real
        # alternates even and odd launches)
        pc = pattern_ceilings(stream, log, list(sim) + [plan['a'], plan['b'], A], v_float=vb == 4)
        out['pattern_ceiling'] = pc
        pat_pr = 0.5 * (pc['real_even']['ms'] + pc['real_odd']['ms'])
        out['roofline']['frac_of_pattern_ceiling'] = round(pat_pr / pr_avg_ms, 4)
        out['roofline']['pass_kspace']['frac_of_pattern_ceiling'] = round(pc['kspace']['ms'] / pk_avg_ms, 4)
    s.close()

    if not args.no_plain:
        # plain single-angle form (unfused per-axis kernels, batch 1) on the same inputs
        sp_ = FORWARD_SOLVER_CONVERGENT_BORN(p, options=dict(form=3, use_graph=0))
        sp_.set_RI(RI)
        d1 = dv.DeviceArray.from_host(u_host[:1])
        d1o = dv.DeviceArray(u_host[:1].shape, np.complex64)
        sp_.solve_incident_device(d1.ptr, 1, d1o.ptr)
        dv.synchronize()
        st = sp_.stream()
        e0.record(st)
        sp_.solve_incident_device(d1.ptr, 1, d1o.ptr)
        e1.record(st)
        e1.synchronize()
        pms = e0.elapsed_ms(e1)
        plain_val = float(nvox) * M / (pms / 1e3) / 1e9
        out['plain_single_angle'] = {'value': round(plain_val, 3), 'unit': 'Gvoxel-iter/s',
                                     'ms_per_angle': round(pms, 3),
                                     'speedup_of_batched_fused': round(value / plain_val, 2)}
        diff = np.linalg.norm(d1o.to_host() - result_gpu[:1]) / np.linalg.norm(result_gpu[:1])
        out['plain_single_angle']['rel_diff_vs_batched'] = float(diff)
        sp_.close()
        log('plain single angle %.1f ms -> %.2f Gvoxel-iter/s' % (pms, plain_val))

    if not args.no_refform:
        # the reference's MATLAB gpuArray loop restated op for op on the GPU, one angle
        sr = FORWARD_SOLVER_CONVERGENT_BORN(p)
        sr.set_RI(RI)
        rms, rout = reference_form_gpu(sr, u[..., 0], M, log)
        sr.close()
        rval = float(nvox) * M / (rms / 1e3) / 1e9
        rdiff = float(np.linalg.norm(rout - result_gpu[0]) / np.linalg.norm(result_gpu[0]))
        out['reference_form_single_angle'] = {
            'value': round(rval, 3), 'unit': 'Gvoxel-iter/s', 'ms_per_angle': round(rms, 3),
            'what': 'solve_raw :577-647 op for op on the GPU (one kernel per MATLAB array expression, '
                    'torch.fft/rocFFT 3D FFTs, fft_flip + conj every iteration), complex64, one angle',
            'speedup_of_batched_fused': round(value / rval, 2), 'rel_diff_vs_batched': rdiff}

    if not args.no_cpu:
        # CPU baseline: complex64 op-for-op port of the reference loop on this host
        pcpu = workload_params(ri, args.cpu_iters, 1)
        ow = Oracle(workload_params(ri, 3, 1), dtype=np.complex64, io_kernels=False, workers=cores)
        ow.set_RI(RI)
        ow.born_loop(u0)          # warm-up (FFT plans, page-in)
        oc = Oracle(pcpu, dtype=np.complex64, io_kernels=False, workers=cores)
        oc.set_RI(RI)
        t0 = time.perf_counter()
        oc.born_loop(u0)
        t_cpu = time.perf_counter() - t0
        cpu_val = float(nvox) * args.cpu_iters / t_cpu / 1e9
        out['cpu_baseline'] = {'value': round(cpu_val, 5), 'unit': 'Gvoxel-iter/s', 'cores': cores, 'kind': 'port',
                               'sample': '1 angle x %d iterations on the %dx%dx%d grid (complex64 NumPy/SciPy port '
                                         'of solve_raw, scipy.fft workers=%d), %.1f s'
                                         % (args.cpu_iters, sim[0], sim[1], sim[2], cores, t_cpu)}
        log('cpu baseline %.4f Gvoxel-iter/s on %d cores' % (cpu_val, cores))

    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
