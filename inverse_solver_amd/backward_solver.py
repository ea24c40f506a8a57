"""BACKWARD_SOLVER_MULTI, MI355X backend: gradient of the transmitted-field
misfit through a forward and an adjoint CBS solve, and the TV-regularised FISTA
reconstruction loop (Codes/BACKWARD_SOLVER_MULTI.m).

The forward potential V and the adjoint potential V' (from conj(flip_z(RI)),
:48) live on the device in two solver handles that share one set of 2D
projection kernels.  A gradient uploads the RI grid once; both potentials are
formed from it on the device (the adjoint one by reading it conjugated and
reversed along z), which replaces the two host-side set_RI calls of the reference.
"""
from __future__ import annotations

import time

import numpy as np

from .forward_solver import FORWARD_SOLVER_CONVERGENT_BORN, get_default_parameters as fwd_defaults
from .params import update_struct
from .tv import tv_fista_inner


def get_default_parameters(init=None):
    """Codes/BACKWARD_SOLVER_MULTI.m:12-29 (on top of BACKWARD_SOLVER's optical defaults)."""
    p = fwd_defaults()
    p.update(forward_solver_parameters=fwd_defaults(dict(return_transmission=True, return_reflection=True,
                                                         return_3D=True, vector_simulation=False)),
             step=0.01, tv_param=0.001, use_non_negativity=False, nmin=-np.inf, nmax=np.inf, kappamax=0,
             inner_itt=100, itter_max=100, num_scan_per_iteration=0, verbose=True, tv_on_device=True)
    return update_struct(p, init) if init is not None else p


class BACKWARD_SOLVER_MULTI:
    get_default_parameters = staticmethod(get_default_parameters)

    def __init__(self, params, device=0, options=None, autotune=None):
        """autotune: measure the kernel plans once, at the first gradient, on the
        forward handle and give the adjoint handle the same plan (default: when no
        options force a plan)."""
        self.parameters = dict(params)
        self._autotune = (not options) if autotune is None else bool(autotune)
        self._tuned = False
        fp = dict(self.parameters['forward_solver_parameters'])
        if not (fp.get('return_transmission') and fp.get('return_reflection') and fp.get('return_3D')):
            raise ValueError('need to set all return parameter to true in the forward solver')   # :57-66
        fp['return_reflection'] = False          # the gradient reads only the transmitted and 3D fields
        self.forward = FORWARD_SOLVER_CONVERGENT_BORN(fp, device=device, options=options)
        self.adjoint = FORWARD_SOLVER_CONVERGENT_BORN(fp, device=device, options=options)
        self.RI_inter = None
        self.err_list = []

    def close(self):
        self._bufs = None
        self._nbuf = None
        self.forward.close()
        self.adjoint.close()

    def _share_io(self):
        self.forward._ensure_io()
        if not self.adjoint._io_ready:
            self.adjoint._io_arrays = self.forward._io_arrays
            from ._lib import check, lib, CBS_MEM_HOST
            from .forward_solver import _fptr
            check(lib.cbs_set_io_kernels(self.adjoint._h, *[_fptr(a) for a in self.adjoint._io_arrays],
                                         CBS_MEM_HOST))
            self.adjoint._io_ready = True

    def _fields(self, A):
        """Device buffers for the forward and adjoint 3D fields of A angles (all field
        components; kept between calls)."""
        from . import device as dv
        n = int(np.prod(self.forward.expected_RI_size)) * self.forward.pole_num * A
        if getattr(self, '_nbuf', None) != A:
            self._bufs = None
            self._bufs = (dv.DeviceArray((n,)), dv.DeviceArray((n,)),
                          dv.DeviceArray((int(np.prod(self.forward.expected_RI_size)),), np.complex128))
            self._nbuf = A
        return self._bufs

    def _ri_device(self, RI):
        """The RI grid as FP64 on the device (x fastest; complex only when it has an
        imaginary part), in a buffer kept between gradients."""
        from . import device as dv
        RI = np.asarray(RI)
        cplx = bool(np.iscomplexobj(RI) and np.any(np.imag(RI) != 0))
        arr = (np.asfortranarray(RI, dtype=np.complex128) if cplx
               else np.asfortranarray(np.real(RI), dtype=np.float64))
        buf = getattr(self, '_ri_buf', None)
        if buf is None or buf.nbytes != arr.nbytes:
            self._ri_buf = None
            self._ri_buf = dv.DeviceArray((arr.size,), arr.dtype)
        self._ri_buf.upload(arr)
        return self._ri_buf, cplx

    def _tune(self, A):
        from ._lib import CbsError
        self._tuned = True
        try:
            self.forward.autotune(A, 6)
        except CbsError:
            return                                   # no fused plan on this grid: keep the planner's choice
        pl = self.forward.plan()
        self.adjoint.set_options(form=0, split_a=pl['a'], tile_t=pl['tile_t'], threads_real=pl['threads_real'],
                                 threads_kspace=pl['threads_kspace'], block_order=pl['block_order'])

    def get_gradiant_static(self, RI, input_field, output_field):
        """Codes/BACKWARD_SOLVER_MULTI.m:30-54: forward solve, residual, adjoint solve on
        conj(flip_z(RI)) with conj(residual), gradient -i mean(conj(u) w).  Both 3D fields
        stay in device memory; cbs_field_gradient forms the gradient in FP64 there."""
        RI = np.asarray(RI)
        if list(RI.shape[:3]) != self.forward.expected_RI_size:
            raise ValueError('The refractiv index does not have the expected size : %d %d %d'
                             % tuple(self.forward.expected_RI_size))
        ri_dev, cplx = self._ri_device(RI.reshape(self.forward.expected_RI_size))
        dg, err = self._gradient_device(ri_dev, cplx, input_field, output_field)
        rx, ry, rz = self.forward.expected_RI_size
        grad = dg.to_host().reshape(rz, ry, rx).transpose(2, 1, 0)
        return grad, err

    def _gradient_device(self, ri_dev, cplx, input_field, output_field):
        """The gradient of get_gradiant_static for an RI grid already in device memory (FP64,
        x fastest; complex when cplx): returns the device buffer holding the complex128
        gradient (x fastest) and the misfit."""
        import ctypes as C
        from ._lib import check, lib, CBS_MEM_DEVICE
        self._share_io()
        inp = self.forward._inputs(input_field)
        A = inp.shape[3]
        du, dw, dg = self._fields(A)
        vmax = self.forward.set_RI_device(ri_dev.ptr, cplx)                      # :32
        if self._autotune and not self._tuned:
            self._tune(A)
        trans, _ = self.forward.solve_keep_3d(inp, du)                            # :34
        diff = trans.astype(np.complex128) - np.asarray(output_field)            # :36
        err = float(np.sum(np.abs(diff) ** 2))                                    # :37
        self.adjoint.set_RI_device(ri_dev.ptr, cplx, flip_conj=True, vmax=vmax)  # :48 conj(flip_z(RI))
        self.adjoint.solve_keep_3d(np.conj(diff), dw, return_2d=False)            # :49
        dims = (C.c_uint32 * 3)(*self.forward.expected_RI_size)
        # mean over polarisation components and angles (:53 averages dims 4 and 5)
        check(lib.cbs_field_gradient(self.adjoint._h, C.c_void_p(du.ptr), C.c_void_p(dw.ptr),
                                     A * self.forward.pole_num, dims, C.c_void_p(dg.ptr), CBS_MEM_DEVICE))   # :50-53
        return dg, err

    def _scan(self, A, rng=None):
        """Angles of one gradient (:80-91): all, or num_scan_per_iteration random ones."""
        n = self.parameters['num_scan_per_iteration']
        if n == 0:
            return np.arange(A)
        rng = rng or np.random.default_rng()
        return np.unique(np.concatenate([[0], rng.permutation(A)[:n]]))[:n]

    def get_gradiant(self, RI, input_field, output_field, rng=None):
        """:80-91 (random angle subsets when num_scan_per_iteration > 0)."""
        scan = self._scan(np.asarray(input_field).shape[3], rng)
        return self.get_gradiant_static(RI, np.asarray(input_field)[:, :, :, scan],
                                        np.asarray(output_field)[:, :, :, scan])

    def solve(self, input_field, output_field, init_RI=None, init_solver=None):
        """:92-196.  The reference starts from init_solver.solve(input, output), a Rytov
        reconstruction (inverse_solver_amd.rytov.BACKWARD_SOLVER_RYTOV); init_RI skips it."""
        p = self.parameters
        if init_RI is None:
            if init_solver is None:
                raise ValueError('give init_RI or an init_solver with .solve(input_field, output_field)')
            init_RI = init_solver.solve(input_field, output_field)
            if isinstance(init_RI, tuple):          # BACKWARD_SOLVER_RYTOV returns (RI, ORytov)
                init_RI = init_RI[0]
        # the iterates are kept x-fastest (Fortran order), the layout the device takes,
        # so no gradient re-lays the RI grid out before its upload
        RI = np.asfortranarray(np.asarray(init_RI).astype(np.complex128))
        n_m = p['RI_bg']
        alpha = 1.0 / p['step']
        t_np = 1.0
        u_n = RI ** 2 / n_m ** 2 - 1
        x_n = u_n.copy(order='F')
        if RI.shape[2] % 2 == 0:
            raise ValueError('need to be odd size')                                # :117-119
        Vmin = (p['nmin'] ** 2 - p['kappamax'] ** 2) / n_m ** 2 - 1
        Vmax = p['nmax'] ** 2 / n_m ** 2 - 1
        Vimag_max = 2 * p['kappamax'] * p['nmax'] / n_m ** 2
        if p.get('tv_on_device', True):
            return self._solve_device(RI, u_n, x_n, input_field, output_field, alpha, Vmin, Vmax, Vimag_max)
        self.err_list = []
        self.timings = []
        for ii in range(int(p['itter_max'])):
            t0 = time.perf_counter()
            grad, err = self.get_gradiant(RI, input_field, output_field)
            t1 = time.perf_counter()
            self.err_list.append(err)
            t_n = t_np
            s_n = tv_fista_inner(u_n - (1 / alpha) * grad, p['tv_param'] / alpha, Vmin, Vmax, Vimag_max,
                                 p['use_non_negativity'], False, p['inner_itt'])
            t_np = (1 + np.sqrt(1 + 4 * t_n ** 2)) / 2
            u_n = s_n + (t_n - 1) / t_np * (s_n - x_n)
            x_n = s_n
            RI = n_m * np.sqrt(u_n.astype(np.complex128) + 1)
            self.RI_inter = RI
            self.timings.append((t1 - t0, time.perf_counter() - t1))
            if p.get('verbose'):
                print('itter : %d  err %.6e  gradient %.2f s  TV %.2f s' % (ii + 1, err, t1 - t0,
                                                                             time.perf_counter() - t1))
        return RI

    def _solve_device(self, RI, u_n, x_n, input_field, output_field, alpha, Vmin, Vmax, Vimag_max):
        """The loop of solve (:120-161) with the iterates, the gradient and the RI in device
        memory: the proximal input (cbs_fista_prox_input), the TV prox of TV_FISTA_inner_v2
        (cbs_tv_prox, or cbs_tv_prox_complex when Vimag_max != 0) and the momentum and RI
        update (cbs_fista_momentum), in the host loop's order and precision; only the
        misfit's 2D fields cross to the host, and the RI once at the end."""
        import ctypes as C
        from . import device as dv
        from ._lib import check, lib, CBS_MEM_DEVICE
        p = self.parameters
        shape = tuple(RI.shape)
        n = int(np.prod(shape))
        cplx_tv = 0 if Vimag_max == 0 else 1
        tv_dt = np.complex64 if cplx_tv else np.float32
        bufs = [dv.DeviceArray.from_host(np.ravel(a, order='F').astype(np.complex128)) for a in (RI, u_n, x_n)]
        d_ri, d_u, d_x = bufs
        d_in, d_out = dv.DeviceArray((n,), tv_dt), dv.DeviceArray((n,), tv_dt)
        dims = (C.c_uint32 * 3)(*shape)
        h = self.forward._h
        lam = p['tv_param'] / alpha
        non_neg = int(bool(p['use_non_negativity']))
        inp_all, out_all = np.asarray(input_field), np.asarray(output_field)
        t_np = 1.0
        self.err_list = []
        self.timings = []
        try:
            for ii in range(int(p['itter_max'])):
                t0 = time.perf_counter()
                scan = self._scan(inp_all.shape[3])
                dg, err = self._gradient_device(d_ri, True, inp_all[:, :, :, scan], out_all[:, :, :, scan])
                t1 = time.perf_counter()
                self.err_list.append(err)
                t_n = t_np
                check(lib.cbs_fista_prox_input(h, C.c_void_p(d_u.ptr), C.c_void_p(dg.ptr), 1.0 / alpha, n, cplx_tv,
                                               C.c_void_p(d_in.ptr)))
                if cplx_tv:
                    check(lib.cbs_tv_prox_complex(h, C.c_void_p(d_in.ptr), C.c_void_p(d_out.ptr), dims, float(lam),
                                                  float(Vmin), float(Vmax), float(Vimag_max), non_neg, 0,
                                                  int(p['inner_itt']), CBS_MEM_DEVICE))
                else:
                    check(lib.cbs_tv_prox(h, C.c_void_p(d_in.ptr), C.c_void_p(d_out.ptr), dims, float(lam),
                                          float(Vmin), float(Vmax), non_neg, 0, int(p['inner_itt']), CBS_MEM_DEVICE))
                t_np = (1 + np.sqrt(1 + 4 * t_n ** 2)) / 2
                check(lib.cbs_fista_momentum(h, C.c_void_p(d_out.ptr), cplx_tv, C.c_void_p(d_x.ptr),
                                             C.c_void_p(d_u.ptr), C.c_void_p(d_ri.ptr), n, (t_n - 1) / t_np,
                                             float(p['RI_bg'])))
                self.timings.append((t1 - t0, time.perf_counter() - t1))
                if p.get('verbose'):
                    print('itter : %d  err %.6e  gradient %.2f s  TV %.2f s' % (ii + 1, err, t1 - t0,
                                                                                 time.perf_counter() - t1))
            RI = d_ri.to_host().reshape(shape, order='F')
        finally:
            for b in bufs + [d_in, d_out]:
                b.free()
        self.RI_inter = RI
        return RI
