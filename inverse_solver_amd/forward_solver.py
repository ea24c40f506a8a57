"""FORWARD_SOLVER_CONVERGENT_BORN, MI355X backend.

Python mirror of the reference's MATLAB class API
(Codes/FORWARD_SOLVER_CONVERGENT_BORN.m, Codes/FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m):
the same parameters (RI, wavelength, resolution, boundary_thickness,
iterations_number, ...), set_RI(RI) and solve(input_field) returning the
transmitted, reflected and 3D fields.  The host side here only prepares the
potential and the geometry in FP64, as the MATLAB class does before calling its
MEX; every solve runs in libcbs_hip.so on the GPU.

Arrays use MATLAB index order: RI is (x, y, z); 2D fields (x, y, P, A) and 3D
fields (x, y, z, P, A), with P = 1 for scalar runs and, for vector runs
(vector_simulation = true, the dyadic Green's function), P = 2 for the 2D input and
output fields and P = 3 (x, y, z components) for the 3D fields.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .params import (basic_optical_parameter, update_struct, round_half_away, pupil,
                     centered_fft2, centered_ifft2, polarisation_basis)
from .io_kernels import build_io_kernels


def get_default_parameters(init=None):
    """FORWARD_SOLVER.get_default_parameters (Codes/FORWARD_SOLVER.m:9-21) followed by
    FORWARD_SOLVER_CONVERGENT_BORN.get_default_parameters (:32-45)."""
    p = basic_optical_parameter()
    p.update(return_transmission=True, return_reflection=False, return_3D=False, use_GPU=True)
    p.update(iterations_number=-1, boundary_thickness=[6, 6, 6], boundary_sharpness=1, verbose=False,
             acyclic=True, RI_xy_size=[0, 0], RI_center=[0, 0])
    return update_struct(p, init) if init is not None else p


def _fptr(a):
    return a.ctypes.data_as(C.c_void_p)


class FORWARD_SOLVER_CONVERGENT_BORN:
    """CBS forward solver (scalar or vector) on one MI355X.

    ``options`` maps to cbs_options (form, split_a, tile_t, max_batch,
    track_norms, use_graph, time_kernels, threads_real, threads_kspace, block_order).
    """

    get_default_parameters = staticmethod(get_default_parameters)

    def __init__(self, params, device=0, options=None):
        p = dict(params)
        self.pole_num = 3 if p.get('vector_simulation', False) else 1        # :294-297
        rxy = list(p.get('RI_xy_size', [0, 0]))
        rxy[0] = rxy[0] or p['size'][0]                                     # :54-59
        rxy[1] = rxy[1] or p['size'][1]
        p['RI_xy_size'] = rxy
        self.parameters = p
        self.expected_RI_size = [int(rxy[0]), int(rxy[1]), int(p['size'][2])]   # :60
        bt = list(p['boundary_thickness'])
        if len(bt) != 3:                                                    # :126-132
            raise ValueError('Boundary in only one direction is not precise; enter "boundary_thickness" as an '
                             'array of size 3 use size 0 for ciclic boundary')
        self.cyclic_boundary_xy = (bt[0] == 0 and bt[1] == 0 and self.expected_RI_size[0] == p['size'][0]
                                   and self.expected_RI_size[1] == p['size'][1])     # :63
        self.boundary_thickness_pixel = round_half_away(
            np.array(bt, float) * p['wavelength'] / p['RI_bg'] / (np.array(p['resolution'], float) * 2)
        ).astype(int)                                                       # :129
        self.simulation_size = [self.expected_RI_size[j] + 2 * int(self.boundary_thickness_pixel[j])
                                for j in range(3)]
        if self.cyclic_boundary_xy:
            conv = [int(p['size'][0]), int(p['size'][1])]
        else:
            conv = [int(p['size'][0]) + self.expected_RI_size[0], int(p['size'][1]) + self.expected_RI_size[1]]
        self.conv_size = conv
        geo = _lib.Geometry()
        for j in range(3):
            geo.simulation_size[j] = self.simulation_size[j]
            geo.RI_size[j] = self.expected_RI_size[j]
        geo.field_size[0], geo.field_size[1], geo.field_size[2] = int(p['size'][0]), int(p['size'][1]), 1
        geo.conv_size[0], geo.conv_size[1], geo.conv_size[2] = conv[0], conv[1], 1
        geo.device = int(device)
        rc = list(p.get('RI_center', [0, 0]))
        geo.RI_center[0], geo.RI_center[1] = int(rc[0]), int(rc[1])
        geo.pole_num = self.pole_num
        h = C.c_void_p()
        check(lib.cbs_create(C.byref(geo), C.byref(h)))
        self._h = h
        self.device = int(device)
        self._io_ready = False
        if options:
            self.set_options(**options)

    # ---------------------------------------------------------------- lifecycle
    def close(self):
        if getattr(self, '_h', None):
            check(lib.cbs_destroy(self._h))
            self._h = None

    delete = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, **kw):
        o = _lib.Options()
        check(lib.cbs_get_options(self._h, C.byref(o)))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise KeyError(f'unknown option {k}')
            setattr(o, k, int(v))
        check(lib.cbs_set_options(self._h, C.byref(o)))

    def autotune(self, n_rhs, iterations=6):
        """Measure every fused kernel plan on the current potential and keep the fastest, then
        the batch width (rhs per launch) for n_rhs right-hand sides; returns ms per solve."""
        ms = C.c_double()
        check(lib.cbs_autotune(self._h, int(n_rhs), int(iterations), C.byref(ms)))
        return ms.value

    def plan(self):
        f, a, b, t = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        check(lib.cbs_get_plan(self._h, C.byref(f), C.byref(a), C.byref(b), C.byref(t)))
        o = _lib.Options()
        check(lib.cbs_get_options(self._h, C.byref(o)))
        form = {_lib.CBS_FORM_PLAIN: 'plain', _lib.CBS_FORM_ROWS: 'rows', _lib.CBS_FORM_AXES: 'axes'}.get(f.value,
                                                                                                  'fused')
        return {'form': form, 'a': a.value, 'b': b.value,
                'tile_t': t.value, 'threads_real': o.threads_real, 'threads_kspace': o.threads_kspace,
                'block_order': o.block_order, 'max_batch': o.max_batch}

    # ---------------------------------------------------------------- potential (:180-396)
    def set_RI(self, RI, flip_conj=False):
        """Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:180-198 (scalar RI).  The RI grid goes to
        the device once; RI2potential, eps = 1.01 max|Vt| and the padded V = Vt - i eps
        are formed there (cbs_potential_max, cbs_set_potential_ri).  flip_conj: use
        conj(RI) reversed along z (the adjoint potential of BACKWARD_SOLVER_MULTI.m:48)."""
        RI = np.asarray(RI)
        if list(RI.shape[:3]) != self.expected_RI_size or (RI.ndim > 3 and int(np.prod(RI.shape[3:])) != 1):
            raise ValueError('The refractiv index does not have the expected size : %d %d %d'
                             % tuple(self.expected_RI_size))
        RI = RI.reshape(self.expected_RI_size)
        cplx = bool(np.iscomplexobj(RI) and np.any(np.imag(RI) != 0))
        arr = (np.asfortranarray(RI, dtype=np.complex128) if cplx
               else np.asfortranarray(np.real(RI), dtype=np.float64))
        self.set_RI_device(arr.ctypes.data, cplx, flip_conj, mem=_lib.CBS_MEM_HOST)
        self.RI = RI

    def set_RI_device(self, ri_ptr, ri_complex, flip_conj=False, vmax=None, mem=None):
        """set_RI from an FP64 RI grid (x fastest; real, or interleaved complex when
        ri_complex) at ri_ptr, device memory unless mem says otherwise.  vmax: a max|Vt|
        already measured on the same grid (conj and flip do not change it).  Returns it."""
        p = self.parameters
        mem = _lib.CBS_MEM_DEVICE if mem is None else mem
        ptr = C.c_void_p(int(ri_ptr))
        if vmax is None:
            out = C.c_double()
            check(lib.cbs_potential_max(self._h, ptr, int(bool(ri_complex)), float(p['wavelength']),
                                        float(p['RI_bg']), mem, C.byref(out)))
            vmax = out.value
        if not np.isfinite(vmax):
            raise ValueError('the refractive index holds non-finite values')
        prm = self._potential_params(float(vmax * 1.01))                   # :205
        check(lib.cbs_set_potential_ri(self._h, C.byref(prm), ptr, int(bool(ri_complex)), int(bool(flip_conj)),
                                       float(p['wavelength']), float(p['RI_bg']), mem))
        return vmax

    def _potential_params(self, eps):
        """The scalars of set_RI / init for eps_imag = eps (:199-222, :280-396); sets
        eps_imag, ROI and Bornmax on the object."""
        p = self.parameters
        k2 = (2 * np.pi * p['RI_bg'] / p['wavelength']) ** 2
        if not eps > 1e-12 * k2:
            raise ValueError('homogeneous RI: eps_imag = 0 and the Born series is undefined '
                             '(the reference would divide by zero)')
        b = self.boundary_thickness_pixel
        rs = self.expected_RI_size
        self.eps_imag = eps
        self.ROI = [b[0] + 1, b[0] + rs[0], b[1] + 1, b[1] + rs[1], b[2] + 1, b[2] + rs[2]]
        k0_nm = p['RI_bg'] / p['wavelength']
        step = abs(2 * (2 * np.pi * k0_nm) / eps)                            # :326
        res = np.array(p['resolution'], float)
        bornmax_opt = np.ceil(np.linalg.norm(np.array(self.simulation_size) * res) / step / 2 + 1) * 2   # :327
        it = p['iterations_number']
        self.Bornmax = 0 if it == 0 else (int(bornmax_opt * abs(it)) if it < 0 else int(it))  # :330-336
        acyclic = bool(p['acyclic'])
        shift_axes = (0b100 if self.cyclic_boundary_xy else 0b111) if acyclic else 0        # :300-315
        ramp_axes = (0b100 if (b[0] == 0 and b[1] == 0) else 0b111) if acyclic else 0       # :355-377
        prm = _lib.PotentialParams()
        prm.k0_nm = k0_nm
        prm.eps_imag = eps
        for j in range(3):
            prm.fourier_res[j] = 1.0 / (res[j] * self.simulation_size[j])
        for j in range(6):
            prm.ROI[j] = int(self.ROI[j])
        prm.Bornmax = self.Bornmax
        prm.dV = float(np.prod(res))
        prm.boundary_sharpness = float(p['boundary_sharpness'])
        prm.acyclic = int(acyclic)
        prm.shift_axes = shift_axes
        prm.ramp_axes = ramp_axes
        return prm

    # ---------------------------------------------------------------- 2D in / 2D out (:397-520)
    def _ensure_io(self):
        if self._io_ready:
            return
        ref, kt, kr = build_io_kernels(self.parameters, self.expected_RI_size, self.cyclic_boundary_xy)
        want = (self.conv_size[0], self.conv_size[1], self.expected_RI_size[2])
        if ref.shape != want:
            # a negative RI_center shrinks the cropped kernels (:82-85), and the refocusing
            # product of :550 then stops on mismatched dimensions
            raise ValueError('refocusing kernel is %s but the convolution grid is %s (RI_center %s)'
                             % (ref.shape, want, list(self.parameters.get('RI_center', [0, 0]))))
        self._io_arrays = [np.asfortranarray(a.astype(np.complex64)) for a in (ref, kt, kr)]
        check(lib.cbs_set_io_kernels(self._h, *[_fptr(a) for a in self._io_arrays], _lib.CBS_MEM_HOST))
        self._io_ready = True

    def _pupil_factors(self):
        """NA circle and the Abbe-sine factors of the current parameters, computed once per
        parameter set (the per-solve filters below reuse them)."""
        p = self.parameters
        key = (tuple(int(v) for v in p['size'][:2]), tuple(float(v) for v in p['resolution'][:2]),
               float(p['wavelength']), float(p['NA']), float(p['RI_bg']))
        cache = getattr(self, '_pupil_cache', None)
        if cache is None or cache[0] != key:
            na, cos_t, _ = pupil(p['size'][:2], p['resolution'], p['wavelength'], p['NA'], p['RI_bg'])
            inv = np.where(na, 1.0 / np.sqrt(np.where(na, cos_t, 1.0)), 0.0)
            fwd = np.sqrt(cos_t) * na
            cache = (key, na, inv, fwd)
            self._pupil_cache = cache
        return cache[1:]

    def _pupil_filter(self, F, inverse):
        p = self.parameters
        na, inv, fwd = self._pupil_factors()
        shp = na.shape + (1,) * (F.ndim - 2)
        F = F * na.reshape(shp)
        if p['use_abbe_sine']:
            F = F * (inv if inverse else fwd).reshape(shp)
        return F

    def transform_field_3D(self, F):
        """FORWARD_SOLVER.m:38-68: NA circle, 1/sqrt(cos theta) (Abbe sine) and, for two
        polarisations (x, y, 2, ...), the change to the 3D field (x, y, 3, ...) through
        the radial / perpendicular basis and the Ewald tangent."""
        F = self._pupil_filter(F, True)
        if F.shape[2] == 1:
            return F
        if F.shape[2] != 2:
            raise ValueError('Far field has two polarisation')
        p = self.parameters
        rad, perp, tan, inside = polarisation_basis(F.shape[:2], p['resolution'], p['wavelength'], p['RI_bg'])
        ex = (slice(None), slice(None), None) + (None,) * (F.ndim - 3)
        F = F * inside[ex]
        a_rad = F[:, :, 0] * rad[:, :, 0][ex[:2] + ex[3:]] + F[:, :, 1] * rad[:, :, 1][ex[:2] + ex[3:]]
        a_perp = F[:, :, 0] * perp[:, :, 0][ex[:2] + ex[3:]] + F[:, :, 1] * perp[:, :, 1][ex[:2] + ex[3:]]
        out = np.empty(F.shape[:2] + (3,) + F.shape[3:], dtype=np.result_type(F, np.complex64))
        for j in range(3):
            out[:, :, j] = a_rad * tan[:, :, j][ex[:2] + ex[3:]]
        for j in range(2):
            out[:, :, j] += a_perp * perp[:, :, j][ex[:2] + ex[3:]]
        return out

    def _to_2d(self, F, reflection):
        F = self._pupil_filter(F, False)
        if F.shape[2] == 1:
            return F
        if F.shape[2] != 3:
            raise ValueError('Near field has three polarisation')
        p = self.parameters
        rad, perp, tan, inside = polarisation_basis(F.shape[:2], p['resolution'], p['wavelength'], p['RI_bg'])
        if reflection:
            tan = tan * np.array([1.0, 1.0, -1.0])                         # :114 (k3 reversed)
        ex = (slice(None), slice(None)) + (None,) * (F.ndim - 3)
        F = F * inside[ex][:, :, None]
        a_tan = sum(F[:, :, j] * tan[:, :, j][ex] for j in range(3))
        a_perp = sum(F[:, :, j] * perp[:, :, j][ex] for j in range(2))
        out = np.empty(F.shape[:2] + (2,) + F.shape[3:], dtype=np.result_type(F, np.complex64))
        for j in range(2):
            out[:, :, j] = a_tan * rad[:, :, j][ex] + a_perp * perp[:, :, j][ex]
        return out

    def transform_field_2D(self, F):
        """FORWARD_SOLVER.m:69-97 (3 -> 2 polarisations for vector fields)."""
        return self._to_2d(F, False)

    def transform_field_2D_reflection(self, F):
        """FORWARD_SOLVER.m:98-129: as transform_field_2D with the tangent's z part
        negated for the reversed propagation direction."""
        return self._to_2d(F, True)

    def solve(self, input_field):
        """Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:397-520 / _CUDA.m:69-148.

        input_field: (x, y, P, A) on the field grid (P = 1 scalar, 2 vector; (x, y, A) and
        (x, y) are accepted for scalar runs, (x, y, 2) for one vector angle).
        Returns (fields_trans (x, y, P, A) or None, fields_ref (x, y, P, A) or None,
        fields_3D (x, y, z, 1 or 3, A) or None) as complex64, following return_* flags."""
        trans, ref, f3 = self._solve(input_field, None)
        return trans, ref, f3

    def solve_keep_3d(self, input_field, field_3d, return_2d=True):
        """solve() with the 3D fields left in device memory: field_3d is a
        device.DeviceArray of n_rhs * components * RI-grid complex64 (x fastest, then
        y, z, component, angle), e.g. for cbs_field_gradient.  Returns (fields_trans,
        fields_ref), both None when return_2d is False (only the 3D fields are formed)."""
        trans, ref, _ = self._solve(input_field, field_3d, return_2d)
        return trans, ref

    def _inputs(self, input_field):
        inp = np.asarray(input_field)
        vec = self.pole_num == 3
        if inp.ndim == 2:
            inp = inp[:, :, None, None]
        elif inp.ndim == 3:
            inp = inp[:, :, :, None] if vec else inp[:, :, None, :]
        if inp.shape[2] > 1 and not vec:                                    # :403-407
            raise ValueError('the source is 2D but parameter indicate a vectorial simulation')
        if inp.shape[2] == 1 and vec:
            raise ValueError('the source is 3D but parameter indicate a non-vectorial simulation')
        if inp.shape[2] > 2:                                                # :411-413
            raise ValueError('Input field must be either a scalar or a 2D vector')
        return inp

    def source_2d(self, input_field):
        """The host half of solve before the library call (:415-417, _CUDA.m:122): the
        input spectrum U after transform_field_3D, and the filtered real-space source
        (x, y, P, A) complex64 that cbs_solve_2d takes.  Also uploads the I/O kernels."""
        p = self.parameters
        inp = self._inputs(input_field)
        if list(inp.shape[:2]) != [int(p['size'][0]), int(p['size'][1])]:
            raise ValueError('input field must have the field size %s' % (p['size'][:2],))
        if not hasattr(self, 'Bornmax'):
            raise RuntimeError('set_RI must be called before solve')
        self._ensure_io()
        U = self.transform_field_3D(centered_fft2(inp.astype(np.complex128)))          # :415-417
        src = np.asfortranarray(centered_ifft2(U).astype(np.complex64))                # _CUDA.m:122
        return U, src

    def _pupils32(self):
        """The input and output pupils of the scalar filters as float32 (x fastest, centred
        frequency order): NA circle, times 1/sqrt(cos theta) and sqrt(cos theta) under
        use_abbe_sine (FORWARD_SOLVER.m:45-50, :76-81)."""
        na, inv, fwd = self._pupil_factors()
        abbe = bool(self.parameters['use_abbe_sine'])
        cache = getattr(self, '_pupil32', None)
        if cache is None or cache[0] is not na or cache[1] != abbe:
            if abbe:
                pin, pout = na * inv, na * fwd
            else:
                pin = pout = na
            cache = (na, abbe, np.asfortranarray(pin, np.float32), np.asfortranarray(pout, np.float32))
            self._pupil32 = cache
        return cache[2], cache[3]

    def _solve_scalar(self, input_field, field_3d_dev, return_2d):
        """Scalar solve with the pupil filters and centred transforms on the device
        (cbs_solve_field): the input field goes in, the output fields come out."""
        p = self.parameters
        inp = self._inputs(input_field)
        fx, fy = int(p['size'][0]), int(p['size'][1])
        if list(inp.shape[:2]) != [fx, fy]:
            raise ValueError('input field must have the field size %s' % (p['size'][:2],))
        if not hasattr(self, 'Bornmax'):
            raise RuntimeError('set_RI must be called before solve')
        self._ensure_io()
        A = inp.shape[3]
        rx, ry, rz = self.expected_RI_size
        want_t = bool(p['return_transmission']) and return_2d
        want_r = bool(p['return_reflection']) and return_2d
        want_3 = bool(p['return_3D']) or field_3d_dev is not None
        pin, pout = self._pupils32()
        nul = C.c_void_p(None)
        fin = np.asfortranarray(inp, dtype=np.complex64)
        if field_3d_dev is None:
            t2 = np.empty((fx, fy, 1, A), np.complex64, order='F') if want_t else None
            r2 = np.empty((fx, fy, 1, A), np.complex64, order='F') if want_r else None
            f3 = np.empty((rx, ry, rz, 1, A), np.complex64, order='F') if want_3 else None
            check(lib.cbs_solve_field(self._h, _fptr(fin), A, _fptr(pin), _fptr(pout), _fptr(t2) if want_t else nul,
                                      _fptr(r2) if want_r else nul, _fptr(f3) if want_3 else nul, _lib.CBS_MEM_HOST))
            return t2, r2, f3
        from . import device as dv
        if field_3d_dev.nbytes != 8 * rx * ry * rz * A or field_3d_dev.dtype != np.complex64:
            raise ValueError('field_3d must hold %d x %s complex64 values' % (A, [rx, ry, rz]))
        d_in = dv.DeviceArray.from_host(fin.T)
        d_t = dv.DeviceArray((A, 1, fy, fx)) if want_t else None
        d_r = dv.DeviceArray((A, 1, fy, fx)) if want_r else None
        check(lib.cbs_solve_field(self._h, C.c_void_p(d_in.ptr), A, _fptr(pin), _fptr(pout),
                                  C.c_void_p(d_t.ptr) if want_t else nul, C.c_void_p(d_r.ptr) if want_r else nul,
                                  C.c_void_p(field_3d_dev.ptr), _lib.CBS_MEM_DEVICE))
        return (d_t.to_host().T if want_t else None), (d_r.to_host().T if want_r else None), None

    def _solve(self, input_field, field_3d_dev, return_2d=True):
        if self.pole_num == 1:
            return self._solve_scalar(input_field, field_3d_dev, return_2d)
        p = self.parameters
        U, src = self.source_2d(input_field)
        A = src.shape[3]
        P = self.pole_num
        fx, fy = int(p['size'][0]), int(p['size'][1])
        rx, ry, rz = self.expected_RI_size
        want_t = bool(p['return_transmission']) and return_2d
        want_r = bool(p['return_reflection']) and return_2d
        want_3 = bool(p['return_3D']) or field_3d_dev is not None
        nul = C.c_void_p(None)
        if field_3d_dev is None:
            t2 = np.empty((fx, fy, P, A), np.complex64, order='F') if want_t else None
            r2 = np.empty((fx, fy, P, A), np.complex64, order='F') if want_r else None
            f3 = np.empty((rx, ry, rz, P, A), np.complex64, order='F') if want_3 else None
            check(lib.cbs_solve_2d(self._h, _fptr(src), A, _fptr(t2) if want_t else nul,
                                   _fptr(r2) if want_r else nul, _fptr(f3) if want_3 else nul, _lib.CBS_MEM_HOST))
        else:
            from . import device as dv
            if field_3d_dev.nbytes != 8 * rx * ry * rz * P * A or field_3d_dev.dtype != np.complex64:
                raise ValueError('field_3d must hold %d x %d x %s complex64 values' % (A, P, [rx, ry, rz]))
            f3 = None
            d_src = dv.DeviceArray.from_host(np.ascontiguousarray(src.T))
            d_t = dv.DeviceArray((A, P, fy, fx)) if want_t else None
            d_r = dv.DeviceArray((A, P, fy, fx)) if want_r else None
            check(lib.cbs_solve_2d(self._h, C.c_void_p(d_src.ptr), A, C.c_void_p(d_t.ptr) if want_t else nul,
                                   C.c_void_p(d_r.ptr) if want_r else nul, C.c_void_p(field_3d_dev.ptr),
                                   _lib.CBS_MEM_DEVICE))
            t2 = d_t.to_host().T if want_t else None
            r2 = d_r.to_host().T if want_r else None
        trans = ref = None
        if want_t:                                                                      # :481-492
            T = centered_fft2(t2.astype(np.complex128)) + U
            trans = centered_ifft2(self.transform_field_2D(T)).astype(np.complex64)
        if want_r:                                                                      # :505-508
            R = centered_fft2(r2.astype(np.complex128))
            ref = centered_ifft2(self.transform_field_2D_reflection(R)).astype(np.complex64)
        return trans, ref, f3

    # ---------------------------------------------------------------- Born loop from a 3D incident field
    def solve_incident(self, u_in):
        """Born loop of :564-647 for incident fields on the RI grid: (x, y, z[, A]) for
        scalar runs, (x, y, z, 3[, A]) for vector runs; returns Field[ROI] + u_in with
        the same shape (complex64)."""
        u = np.asarray(u_in)
        P = self.pole_num
        squeeze = u.ndim == (3 if P == 1 else 4)
        if squeeze:
            u = u[..., None]
        if P == 1:
            u = u[:, :, :, None, :]
        if list(u.shape[:3]) != self.expected_RI_size or u.shape[3] != P:
            raise ValueError('incident field must have the RI grid size %s%s' % (
                self.expected_RI_size, '' if P == 1 else ' x 3 components'))
        A = u.shape[4]
        uf = np.asfortranarray(u, dtype=np.complex64)      # no copy when already complex64, x fastest
        out = np.empty_like(uf, order='F')
        check(lib.cbs_solve_3d(self._h, _fptr(uf), A, _fptr(out), _lib.CBS_MEM_HOST))
        if P == 1:
            out = out[:, :, :, 0, :]
        return out[..., 0] if squeeze else out

    def solve_incident_device(self, u_in_ptr, n_rhs, u_out_ptr):
        """Same as solve_incident on device-resident complex64 buffers (pointers)."""
        check(lib.cbs_solve_3d(self._h, C.c_void_p(int(u_in_ptr)), int(n_rhs), C.c_void_p(int(u_out_ptr)),
                               _lib.CBS_MEM_DEVICE))

    def norms(self, n_rhs):
        out = np.zeros(n_rhs * self.Bornmax, dtype=np.float64)
        check(lib.cbs_get_norms(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), n_rhs, self.Bornmax))
        return out.reshape(n_rhs, self.Bornmax)

    def set_convergence(self, tol):
        """FP64 convergence test (0 = off, the reference's fixed Bornmax): each rhs stops
        after the first iteration n >= 3 with |f_n| <= tol |F_1 + F_2| (cbs_set_convergence)."""
        check(lib.cbs_set_convergence(self._h, float(tol)))

    def iterations(self, n_rhs):
        """Iterations each rhs of the last solve ran."""
        out = np.zeros(n_rhs, dtype=np.uint32)
        check(lib.cbs_get_iterations(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32)), n_rhs))
        return out

    def kernel_times(self):
        ms = (C.c_double * 3)()
        cnt = (C.c_uint64 * 3)()
        check(lib.cbs_get_kernel_times(self._h, ms, cnt, 3))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(['pass_real', 'pass_kspace', 'other'])}

    def stream(self):
        s = C.c_void_p()
        check(lib.cbs_get_stream(self._h, C.byref(s)))
        return s.value

    def footprint(self, n_rhs):
        b = C.c_uint64()
        check(lib.cbs_footprint(self._h, int(n_rhs), C.byref(b)))
        return b.value
