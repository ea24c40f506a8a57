"""Scalar convergent Born series forward solver (FP64 restatement).

Test infrastructure only (see oracle/__init__.py).

Follows Codes/FORWARD_SOLVER_CONVERGENT_BORN.m and Codes/FORWARD_SOLVER.m line by
line for the scalar (vector_simulation=false) and the vector (dyadic Green's
function, pole_num = 3, scalar RI) branches.  ``dtype`` selects the
complex type of the solve: complex128 is the oracle; complex64 is the op-for-op
CPU port of the reference loop that ``bench.py`` times as the CPU baseline.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft as sfft

from .optics import (basic_optical_parameter, update_struct, derive_optical_tool, ri2potential,
                     potential2ri, fft_flip, cfft2, cifft2, cifftn3, polarisation_utility)
from .hankel import truncated_green_plus


def matlab_round(x):
    """MATLAB round(): half away from zero."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def forward_solver_default_parameters(init=None):
    """Codes/FORWARD_SOLVER.m:9-21 + Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:32-45."""
    p = basic_optical_parameter()
    p.update(return_transmission=True, return_reflection=False, return_3D=False, use_GPU=True)
    p.update(iterations_number=-1, boundary_thickness=[6, 6, 6], boundary_sharpness=1,
             verbose=False, acyclic=True, RI_xy_size=[0, 0], RI_center=[0, 0])
    if init is not None:
        p = update_struct(p, init)
    return p


class ForwardSolverConvergentBorn:
    """FORWARD_SOLVER_CONVERGENT_BORN restated in NumPy.

    Arrays use MATLAB index order: RI is (x, y, z); input fields (x, y, P, A) with
    P = 1 (scalar) or 2 (vector); 3D fields (x, y, z, 1 or 3, A).
    """

    def __init__(self, params, dtype=np.complex128, workers=None, io_kernels=True):
        self.parameters = dict(params)
        p = self.parameters
        self.pole_num = 3 if p.get('vector_simulation', False) else 1          # :294-297
        self.dtype = np.dtype(dtype)
        self.rdtype = np.float64 if self.dtype == np.complex128 else np.float32
        self.workers = workers if workers is not None else os.cpu_count()
        self.utility = derive_optical_tool(p)                                   # FORWARD_SOLVER.m:26
        rxy = list(p['RI_xy_size'])
        if rxy[0] == 0:                                                         # :54-59
            rxy[0] = p['size'][0]
        if rxy[1] == 0:
            rxy[1] = p['size'][1]
        p['RI_xy_size'] = rxy
        self.expected_RI_size = [int(rxy[0]), int(rxy[1]), int(p['size'][2])]   # :60
        bt = list(p['boundary_thickness'])
        self.cyclic_boundary_xy = (bt[0] == 0 and bt[1] == 0 and
                                   self.expected_RI_size[0] == p['size'][0] and
                                   self.expected_RI_size[1] == p['size'][1])   # :63
        u = self.utility
        if not io_kernels:
            # Born-loop-only use (solve_raw from a given incident field): skip the
            # I/O kernel construction of :62-118, which the loop does not read.
            pass
        elif self.cyclic_boundary_xy:                                           # :65-72
            ref = np.exp(u['refocusing_kernel'] * u['image_coor'][2])
            ref = ref * u['NA_circle']
            self.refocusing_util = ref
            fsg = ref / (1j * 4 * np.pi)
            fsg = fsg * u['NA_circle'] / (u['k3'] + (~u['NA_circle']))
            fsg = fsg / (u['res'][0] * u['res'][1])
            fsg = cifft2(fsg)
        else:                                                                   # :73-110
            ptg = dict(p)
            ptg['size'] = [p['size'][0] + self.expected_RI_size[0] + p['RI_center'][0],
                           p['size'][1] + self.expected_RI_size[1] + p['RI_center'][1],
                           p['size'][2]]
            rc = p['RI_center']
            ref = truncated_green_plus(ptg, True)                               # :80
            ref = self._crop_center_shift(ref, rc)                              # :82-85
            ref = cifftn3(ref)                                                  # :86
            ref = cfft2(ref)                                                    # :87
            self.refocusing_util = ref * (u['res'][0] * u['res'][1])           # :92
            fsg = truncated_green_plus(ptg)                                     # :97
            fsg = self._crop_center_shift(fsg, rc)                              # :103-106
            fsg = cifftn3(fsg)                                                  # :107
        if io_kernels:
            self.free_space_green = fsg
            self.kernel_trans = cfft2(np.conj(fsg))                             # :114
            self.kernel_ref = cfft2(fsg)                                        # :115
        if len(bt) != 3:                                                        # :126-132
            raise ValueError('boundary_thickness must have 3 entries')
        self.boundary_thickness_pixel = matlab_round(
            np.array(bt, dtype=np.float64) * p['wavelength'] / p['RI_bg'] /
            (np.array(p['resolution'], dtype=np.float64) * 2)).astype(int)     # :129
        self.flip_count = 0

    @staticmethod
    def _crop_center_shift(a, rc):
        """:82-85 crop by RI_center then circshift by -RI_center (xy)."""
        s0 = -min(0, rc[0]); e0 = a.shape[0] - max(0, rc[0])
        s1 = -min(0, rc[1]); e1 = a.shape[1] - max(0, rc[1])
        a = a[s0:e0, s1:e1]
        return np.roll(a, (-rc[0], -rc[1]), axis=(0, 1))

    # ---- padding / cropping between field, RI and convolution grids (:134-179)
    def _size_conv(self):
        p = self.parameters
        return [p['size'][0] + self.expected_RI_size[0], p['size'][1] + self.expected_RI_size[1]]

    @staticmethod
    def _pad_xy(m, conv, add_start):
        pad = [(int(add_start[0]), int(conv[0] - m.shape[0] - add_start[0])),
               (int(add_start[1]), int(conv[1] - m.shape[1] - add_start[1]))]
        pad += [(0, 0)] * (m.ndim - 2)
        return np.pad(m, pad)

    def padd_RI2conv(self, m):
        """:134-147."""
        if self.cyclic_boundary_xy:
            return m
        conv = self._size_conv()
        rc = self.parameters['RI_center']
        st = [conv[j] // 2 - m.shape[j] // 2 + rc[j] for j in range(2)]
        return self._pad_xy(m, conv, st)

    def padd_field2conv(self, m):
        """:148-160."""
        if self.cyclic_boundary_xy:
            return m
        conv = self._size_conv()
        st = [conv[j] // 2 - m.shape[j] // 2 for j in range(2)]
        return self._pad_xy(m, conv, st)

    def crop_conv2field(self, m):
        """:161-168."""
        if self.cyclic_boundary_xy:
            return m
        sz = self.parameters['size']
        st = [m.shape[j] // 2 - sz[j] // 2 for j in range(2)]
        return m[st[0]:st[0] + sz[0], st[1]:st[1] + sz[1]]

    def crop_field2RI(self, m):
        """:173-179 (crops to parameters.size(1:2), offset by RI_center, as the reference does;
        MATLAB stops with an index error when the window leaves the array)."""
        sz = self.parameters['size']
        rc = self.parameters['RI_center']
        st = [m.shape[j] // 2 - sz[j] // 2 + rc[j] for j in range(2)]
        if any(st[j] < 0 or st[j] + sz[j] > m.shape[j] for j in range(2)):
            raise IndexError('crop_field2RI: RI_center moves the crop outside the array (index exceeds dimensions)')
        return m[st[0]:st[0] + sz[0], st[1]:st[1] + sz[1]]

    def crop_conv2RI(self, m):
        """:169-172."""
        return self.crop_field2RI(self.crop_conv2field(m))

    # ---- potential set-up (:180-396)
    def set_RI(self, RI):
        """:180-198."""
        RI = np.asarray(RI)
        if list(RI.shape[:3]) != self.expected_RI_size:                         # :188-191
            raise ValueError('The refractive index does not have the expected size : %d %d %d'
                             % tuple(self.expected_RI_size))
        if RI.ndim > 3 and RI.shape[3] > 1:
            raise NotImplementedError('tensor RI is out of scope')
        self.RI = RI.reshape(self.expected_RI_size).astype(np.complex128)
        self.condition_RI()
        self.init()

    def condition_RI(self):
        """:199-222."""
        p = self.parameters
        pot = ri2potential(self.RI, p['wavelength'], p['RI_bg'])
        self.eps_imag = float(np.max(np.abs(pot)) * 1.01)                        # :205
        k2 = (2 * np.pi * p['RI_bg'] / p['wavelength']) ** 2
        if self.eps_imag <= 1e-12 * k2:
            # The reference divides by eps_imag (:207,326,596,599); a homogeneous RI
            # would give Inf/NaN there.  The oracle refuses it explicitly.
            raise ValueError('homogeneous RI (eps_imag = 0) is not solvable by CBS')
        step = abs(2 * (2 * np.pi * (p['RI_bg'] / p['wavelength'])) / self.eps_imag)   # :207
        self.pixel_step_size = matlab_round(step / np.array(p['resolution'], dtype=np.float64))
        self.ROI = self.create_boundary_RI()                                   # :210
        self.Vt = ri2potential(self.RI, p['wavelength'], p['RI_bg'])           # :212 (padded RI)
        self.V = self.Vt - 1j * self.eps_imag                                  # :215

    def create_boundary_RI(self):
        """:223-279.  Pads the potential with zeros and builds the attenuation mask."""
        p = self.parameters
        b = self.boundary_thickness_pixel
        old = self.RI.shape
        pot = ri2potential(self.RI, p['wavelength'], p['RI_bg'])                # :232
        pot = np.pad(pot, [(int(b[0]), int(b[0])), (int(b[1]), int(b[1])), (int(b[2]), int(b[2]))])  # :238-240
        ROI = [b[0] + 1, b[0] + old[0], b[1] + 1, b[1] + old[1], b[2] + 1, b[2] + old[2]]   # :243-246 (1-based)
        mask = np.ones((1, 1, 1))
        for j in range(3):                                                     # :252-271
            n = pot.shape[j]
            x = np.abs(np.arange(1, n + 1) - (np.floor(n / 2 + 1)) + 0.5) - 0.5
            x = np.roll(x, -(n // 2))
            with np.errstate(divide='ignore', invalid='ignore'):
                x = x / (b[j] - 0.5)
            val0 = x.copy()
            val0[np.abs(x) >= 1] = 1
            val0 = np.abs(val0)
            val0 = 1 - val0
            sh = p['boundary_sharpness']
            val0[val0 > sh] = sh
            val0 = val0 / sh
            if b[j] == 0:
                val0[:] = 0
            shp = [1, 1, 1]
            shp[j] = -1
            mask = mask * (1 - val0.reshape(shp))
        self.attenuation_mask = mask
        self.RI = potential2ri(pot, p['wavelength'], p['RI_bg'])               # :274
        return ROI

    def init(self):
        """:280-396 (scalar)."""
        p = self.parameters
        pb = dict(p)
        pb['size'] = list(self.RI.shape)
        ub = derive_optical_tool(pb)                                           # :284
        self.utility_border = ub
        fc = ub['fourier_coor']
        fr = ub['fourier_res']
        if p['acyclic'] and self.cyclic_boundary_xy:                           # :300-315
            sc = [fc[0], fc[1], fc[2] + fr[2] / 4]
        elif p['acyclic']:
            sc = [fc[0] + fr[0] / 4, fc[1] + fr[1] / 4, fc[2] + fr[2] / 4]
        else:
            sc = [fc[0], fc[1], fc[2]]
        self.shifted_coordinate = sc
        step = abs(2 * (2 * np.pi * ub['k0_nm']) / self.eps_imag)               # :326
        sz = np.array(self.RI.shape, dtype=np.float64)
        bornmax_opt = np.ceil(np.linalg.norm(sz * np.array(p['resolution'])) / step / 2 + 1) * 2  # :327
        it = p['iterations_number']
        if it == 0:                                                            # :330-336
            self.Bornmax = 0
        elif it < 0:
            self.Bornmax = int(bornmax_opt * abs(it))
        else:
            self.Bornmax = int(it)
        g = 1.0 / (4 * np.pi ** 2 * np.abs(sc[0] ** 2 + sc[1] ** 2 + sc[2] ** 2)
                   - (2 * np.pi * ub['k0_nm']) ** 2 - 1j * self.eps_imag)      # :347-351
        b = self.boundary_thickness_pixel
        if p['acyclic']:                                                       # :355-377
            if b[0] == 0 and b[1] == 0:
                n = self.V.shape[2]
                x = np.exp(-1j * np.pi * np.arange(n) / n / 2)
                x = x / x[n // 2]
                ramp = x.reshape(1, 1, -1)
            else:
                ramp = np.ones((1, 1, 1), dtype=np.complex128)
                for j in range(3):
                    n = self.V.shape[j]
                    x = np.exp(-1j * np.pi * np.arange(n) / n / 2)
                    x = x / x[n // 2]
                    shp = [1, 1, 1]
                    shp[j] = -1
                    ramp = ramp * x.reshape(shp)
        else:
            ramp = np.ones((1, 1, 1), dtype=np.complex128)
        self.phase_ramp = ramp
        self.Greenp = sfft.ifftshift(g, axes=(0, 1, 2))                         # :384
        # dyadic terms (:318-324): rads = k_shifted / k0_nm per axis, and the
        # absorption correction k^2 / (k^2 + i eps) of the longitudinal part
        full = self.Greenp.shape
        rads = np.stack([np.broadcast_to(sc[j] / ub['k0_nm'], full) for j in range(3)], axis=3)
        self.rads = sfft.ifftshift(rads, axes=(0, 1, 2))                         # :385
        k2 = (2 * np.pi * ub['k0_nm']) ** 2
        self.green_absorbtion_correction = k2 / (k2 + 1j * self.eps_imag)       # :324
        self.flip_count = 0

    # ---- 2D <-> 3D field transforms (FORWARD_SOLVER.m:38-129)
    def _na_abbe(self, F, inverse):
        u = self.utility
        na = u['NA_circle'][:, :, 0]
        shp = na.shape + (1,) * (F.ndim - 2)
        F = F * na.reshape(shp)
        if self.parameters['use_abbe_sine']:
            if inverse:                                                        # :43-47
                filt = na.astype(np.float64)
                filt[na] = filt[na] / np.sqrt(u['cos_theta'][:, :, 0][na])
            else:                                                              # :72-75
                filt = np.sqrt(u['cos_theta'][:, :, 0]) * na
            F = F * filt.reshape(shp)
        return F

    def transform_field_3D(self, F):
        """FORWARD_SOLVER.m:38-68: NA circle, 1/sqrt(cos theta) (Abbe sine) and, for two
        polarisations, the (radial, perpendicular) -> 3D Ewald-tangent basis change."""
        F = self._na_abbe(F, True)
        if F.shape[2] == 1:
            return F
        if F.shape[2] != 2:
            raise ValueError('Far field has two polarisation')                 # :49-51
        p = self.parameters
        R2, P2, T3, Km = polarisation_utility(F.shape, p['RI_bg'], p['wavelength'], p['resolution'])   # :53
        F = F * Km[:, :, None, None]                                           # :55
        b1 = np.sum(F * R2[:, :, :, None], axis=2)                             # :58
        b2 = np.sum(F * P2[:, :, :, None], axis=2)                             # :59
        out = b1[:, :, None, :] * T3[:, :, :, None]                            # :62
        out[:, :, :2, :] += b2[:, :, None, :] * P2[:, :, :, None]              # :63
        return out

    def _to_2d(self, F, reflection):
        F = self._na_abbe(F, False)
        if F.shape[2] == 1:
            return F
        if F.shape[2] != 3:
            raise ValueError('Near field has three polarisation')              # :78-80
        p = self.parameters
        R2, P2, T3, Km = polarisation_utility(F.shape, p['RI_bg'], p['wavelength'], p['resolution'])
        if reflection:
            T3 = T3.copy()
            T3[:, :, 2] = -T3[:, :, 2]                                         # :114
        F = F * Km[:, :, None, None]
        b1 = np.sum(F * T3[:, :, :, None], axis=2)                             # :87
        b2 = np.sum(F[:, :, :2, :] * P2[:, :, :, None], axis=2)                # :88
        return b1[:, :, None, :] * R2[:, :, :, None] + b2[:, :, None, :] * P2[:, :, :, None]   # :90-92

    def transform_field_2D(self, F):
        """FORWARD_SOLVER.m:69-97 (3D -> 2 polarisations for vector fields)."""
        return self._to_2d(F, False)

    def transform_field_2D_reflection(self, F):
        """FORWARD_SOLVER.m:98-129 (the Ewald tangent's z part is negated, :114)."""
        return self._to_2d(F, True)

    # ---- solve (:397-520)
    def solve(self, input_field):
        """:397-520.  input_field (x, y, P, A); P = 1 for scalar, 2 for vector runs."""
        p = self.parameters
        inp = np.asarray(input_field)
        if inp.ndim == 2:
            inp = inp[:, :, None, None]
        elif inp.ndim == 3:
            inp = inp[:, :, None, :] if self.pole_num == 1 else inp[:, :, :, None]
        if inp.shape[2] > 1 and self.pole_num == 1:                            # :403-407
            raise ValueError('the source is 2D but parameter indicate a vectorial simulation')
        if inp.shape[2] == 1 and self.pole_num == 3:
            raise ValueError('the source is 3D but parameter indicate a non-vectorial simulation')
        if inp.shape[2] > 2:                                                   # :411-413
            raise ValueError('Input field must be either a scalar or a 2D vector')
        inp = cfft2(inp.astype(self.dtype), workers=self.workers)              # :415
        inp = self.transform_field_3D(inp)                                     # :417
        A = inp.shape[3]
        npol = inp.shape[2]
        out_pol = 2 if self.pole_num == 3 else 1                               # :419-422
        R = self.ROI
        nx, ny, nz = R[1] - R[0] + 1, R[3] - R[2] + 1, R[5] - R[4] + 1
        trans = np.ones((nx, ny, out_pol, A), dtype=self.dtype) if p['return_transmission'] else None
        refl = np.ones((nx, ny, out_pol, A), dtype=self.dtype) if p['return_reflection'] else None
        f3d = np.ones((nx, ny, nz, npol, A), dtype=self.dtype) if p['return_3D'] else None
        roi = (slice(R[0] - 1, R[1]), slice(R[2] - 1, R[3]), slice(R[4] - 1, R[5]))
        for n in range(A):                                                     # :437
            Field = self.solve_raw(inp[:, :, 0, n] if npol == 1 else inp[:, :, :, n])   # :439
            Field = Field.reshape(Field.shape[:3] + (npol,))
            if p['return_3D']:
                f3d[:, :, :, :, n] = Field
            if p['return_transmission'] or p['return_reflection']:
                pot = ri2potential(self.RI[roi], p['wavelength'], p['RI_bg'])  # :445
                em = Field * pot[:, :, :, None] * self.utility_border['dV']     # :448 (scalar RI)
                if not self.cyclic_boundary_xy:
                    em = self.padd_RI2conv(em)                                  # :458
                em = cfft2(em, workers=self.workers)                           # :460
            if p['return_transmission']:
                ft = self.crop_conv2field(cifft2(np.sum(em * self.kernel_trans[:, :, :, None], axis=2)))  # :472
                ft = cfft2(ft) + inp[:, :, :, n]                               # :481-484
                ft = self.transform_field_2D(ft[:, :, :, None])[:, :, :, 0]    # :490
                trans[:, :, :, n] = cifft2(ft)                                 # :491-492
            if p['return_reflection']:
                fr = self.crop_conv2field(cifft2(np.sum(em * self.kernel_ref[:, :, :, None], axis=2)))  # :500
                fr = cfft2(fr)                                                 # :505
                fr = self.transform_field_2D_reflection(fr[:, :, :, None])[:, :, :, 0]   # :506
                refl[:, :, :, n] = cifft2(fr)                                  # :507-508
        return trans, refl, f3d

    def incident_field(self, source):
        """:543-554: refocus the filtered 2D spectrum (x, y[, P]) into the 3D incident
        field u_in (x, y, z[, P]) on the RI grid, each polarisation separately."""
        src = np.asarray(source)
        vec = src.ndim == 3
        if not vec:
            src = src[:, :, None]
        s = cifft2(src.astype(self.dtype), workers=self.workers)               # :543
        s = self.padd_field2conv(s)                                            # :544
        s = cfft2(s, workers=self.workers)                                     # :548
        s = s[:, :, None, :] * self.refocusing_util[:, :, :, None]             # :550
        s = cifft2(s, workers=self.workers)                                    # :553
        s = self.crop_conv2RI(s)                                               # :554
        return s if vec else s[:, :, :, 0]

    def solve_raw(self, source):
        """:521-654 from the 2D spectrum of one angle."""
        u_in = self.incident_field(source)
        return self.born_loop(u_in)

    def born_loop(self, incident_field, increments=None, tol=None, stats=None, snapshots=None):
        """:564-647: the two-chain, alternating-shift Born series.

        Scalar: incident_field (x, y, z).  Vector (pole_num = 3): (x, y, z, 3), with
        the dyadic Green's function G (I - c rads rads^T) of :608-615, c the
        absorption correction of :324.

        ``self.Greenp``, ``self.rads`` and ``self.phase_ramp`` are flipped/conjugated
        at the top of every iteration and the flipped state persists across calls,
        exactly as the reference's handle properties do (:579-585); it is reset by
        set_RI.  For odd Bornmax this makes consecutive angles start on opposite
        shifts, as in the reference.

        ``increments``: an optional list that receives, per iteration, the FP64 sum
        over the simulation grid (and the components) of |Field_n|^2 at the moment
        Field_n is added to Field (:638), i.e. the norm of each series term.

        ``tol``: the build's FP64 convergence test, which the reference does not have
        (its loop always runs Bornmax iterations): the loop stops after the first
        iteration n >= 3 at which |Field_n|^2 <= tol^2 |F_1 + F_2|^2 (the norm of the
        series term against the first-order field, both over the simulation grid).
        The flipped Greenp state still advances by Bornmax, as if every iteration had
        run.  ``stats`` (a dict) receives 'iterations' and 'fref' (|F_1 + F_2|^2).

        ``snapshots``: a dict whose keys are iteration counts n; each value is set to
        Field[ROI] + u_in after iteration n, the result a Bornmax = n series gives.
        """
        p = self.parameters
        dt = self.dtype
        R = self.ROI
        roi = (slice(R[0] - 1, R[1]), slice(R[2] - 1, R[3]), slice(R[4] - 1, R[5]))
        vec = self.pole_num == 3
        V = self.V.astype(dt)
        mask = self.attenuation_mask.astype(self.rdtype)
        eps = self.eps_imag
        inc = np.asarray(incident_field).astype(dt)
        if vec:
            V = V[:, :, :, None]
            mask = mask[:, :, :, None] if mask.ndim == 3 else mask.reshape(mask.shape + (1,))
        shape = self.V.shape + ((3,) if vec else ())
        psi = np.zeros(shape, dt)
        Field = np.zeros(shape, dt)
        Field_n = np.zeros(shape, dt)
        source = (V[roi] + 1j * eps) * inc                                     # :565-575 (scalar RI)
        Greenp = self.Greenp.astype(dt)
        rads = self.rads.astype(dt) if vec else None
        cabs = self.green_absorbtion_correction
        ramp = self.phase_ramp.astype(dt)
        ax = (0, 1, 2)
        w = self.workers
        temp = None
        fref = None
        ran = self.Bornmax
        for jj in range(1, self.Bornmax + 1):                                  # :577
            if p['acyclic']:                                                   # :579-585
                Greenp = fft_flip(Greenp, [1, 1, 1], False)
                if vec:
                    rads = fft_flip(rads, [1, 1, 1], False)
                ramp = np.conj(ramp)
            psi[...] = 0                                                       # :588-589
            if jj in (1, 2):
                Field_n[...] = 0                                               # :590-592
                psi[roi] = (1j / eps) * source / 2                             # :596
            else:
                psi = (1j / eps) * Field_n * V                                 # :599
            if vec:
                PSI = np.zeros(shape, dt)
                for j2 in range(3):                                            # :607-611
                    coeff = sfft.fftn(psi[..., j2] * ramp, axes=ax, workers=w)
                    e = np.zeros(3)
                    e[j2] = 1
                    PSI = PSI + (Greenp[..., None] * (e - cabs * rads * rads[..., j2:j2 + 1])) * coeff[..., None]
                for j1 in range(3):                                            # :616-618
                    PSI[..., j1] = sfft.ifftn(PSI[..., j1], axes=ax, workers=w) * np.conj(ramp)
            else:
                PSI = Greenp * sfft.fftn(psi * ramp, axes=ax, workers=w)       # :609,613
                PSI = sfft.ifftn(PSI, axes=ax, workers=w) * np.conj(ramp)      # :617
            if jj not in (1, 2):
                Field_n = Field_n - psi                                        # :619-621
            Field_n = Field_n + V * PSI                                        # :623
            Field_n = Field_n * mask                                           # :630
            if jj == 2:
                temp = Field.copy()                                            # :634-637
            Field = Field + Field_n                                            # :638
            nrm = None
            if increments is not None or tol is not None:
                nrm = float(np.sum(np.abs(Field_n.astype(np.complex128)) ** 2))
            if increments is not None:
                increments.append(nrm)
            if jj == 2 and tol is not None:
                fref = float(np.sum(np.abs(Field.astype(np.complex128)) ** 2))
            if jj == 3:
                Field_n = Field_n + temp                                       # :639-642
            if snapshots is not None and jj in snapshots:
                snapshots[jj] = Field[roi] + inc
            if tol is not None and jj >= 3 and nrm <= tol * tol * fref:
                ran = jj
                if (self.Bornmax - jj) % 2 and p['acyclic']:
                    # the flips of the iterations not run
                    Greenp = fft_flip(Greenp, [1, 1, 1], False)
                    if vec:
                        rads = fft_flip(rads, [1, 1, 1], False)
                    ramp = np.conj(ramp)
                break
        if stats is not None:
            stats['iterations'] = ran
            stats['fref'] = fref
        if p['acyclic'] and self.Bornmax % 2:
            # persist the flipped state like the reference's handle properties
            self.Greenp = Greenp.astype(np.complex128)
            self.phase_ramp = ramp.astype(np.complex128)
            if vec:
                self.rads = rads.astype(np.complex128)
        return Field[roi] + inc                                                # :646-647
