"""FIELD_EXPERIMENTAL_RETRIEVAL: complex fields from off-axis holograms.

Python mirror of Codes/FIELD_EXPERIMENTAL_RETRIEVAL.m (the experimental input
pipeline that feeds the inverse solvers): the background and sample hologram stacks
are Fourier filtered around the +1 sideband, re-centred, matched to the simulation
resolution, cropped to the NA, normalised, and their phase ratio is unwrapped on the
GPU (cbs_phase_unwrap, the reference's unwrapp2_gpu) and flattened (PhiShiftMS).
get_aberration_pattern / correct_aberration / refocus / remove_dirty_fields follow the
class's remaining methods.  Images come in as arrays (n1, n2, A) in MATLAB index
order; the reference's loadTIFF (:454-466) is the caller's file reader, and the
interactive ROI selection (:48-100, drawrectangle) is not offered (a numeric ROI is).
The 2D FFTs are host NumPy, as they are host (or gpuArray) MATLAB there.
"""
from __future__ import annotations

import numpy as np

from .params import basic_optical_parameter, update_struct, centered_fft2, centered_ifft2, round_half_away


def get_default_parameters(init=None):
    """FIELD_EXPERIMENTAL_RETRIEVAL.get_default_parameters (:8-23)."""
    p = basic_optical_parameter()
    p.update(resolution_image=[0.1, 0.1], use_abbe_correction=True, cutout_portion=1 / 3, other_corner=False,
             conjugate_field=False, use_GPU=True)
    return update_struct(p, init) if init is not None else p


def derive_optical_tool(p):
    """The fields of DERIVE_OPTICAL_TOOL (Codes/sub_code/DERIVE_OPTICAL_TOOL.m:1-81) this
    class reads: centred Fourier coordinates, the NA circle, k3 and the refocusing kernel."""
    n1, n2 = int(p['size'][0]), int(p['size'][1])
    r1, r2 = float(p['resolution'][0]), float(p['resolution'][1])
    c1 = (np.arange(1, n1 + 1) - (n1 // 2 + 1)) / (r1 * n1)
    c2 = (np.arange(1, n2 + 1) - (n2 // 2 + 1)) / (r2 * n2)
    coorxy = np.sqrt(c1[:, None] ** 2 + c2[None, :] ** 2)
    k0_nm = p['RI_bg'] / p['wavelength']
    kmax = p['NA'] / p['wavelength']
    k3 = np.sqrt(np.maximum(k0_nm ** 2 - coorxy ** 2, 0.0))
    return {'coor': (c1, c2), 'res': (1 / (r1 * n1), 1 / (r2 * n2)), 'coorxy': coorxy, 'kmax': kmax,
            'NA_circle': coorxy < kmax, 'k3': k3, 'refocusing_kernel': 1j * 2 * np.pi * k3}


def _matlab_prctile(x, q):
    """MATLAB prctile (sorted values at percentages 100 (i - 0.5) / n, linear in
    between, clamped at the ends): numpy's 'hazen' plotting positions."""
    return float(np.percentile(np.asarray(x, dtype=np.float64), q, method='hazen'))


def peak_subpixel_positioner(peak_img):
    """Codes/sub_code/peak_subpixel_positioner.m for a 2D image with the position grids of
    CPU_placement_finder_prepare (centre floor(n/2)+1): the integer peak plus the subpixel
    correction from the larger neighbour on each axis (the first of equal maxima, in
    MATLAB's column-major order)."""
    img = np.asarray(peak_img, dtype=np.float64)
    n1, n2 = img.shape
    flat = img.ravel(order='F')
    m = int(np.argmax(flat))
    i1, i2 = m % n1, m // n1
    pos = [i1 - n1 // 2, i2 - n2 // 2]
    if not (0 < i1 < n1 - 1 and 0 < i2 < n2 - 1):
        raise IndexError('peak on the image border: its neighbours are outside the image')
    p0 = img[i1, i2]
    d = []
    for axis in (0, 1):
        if axis == 0:
            a, b = img[i1 - 1, i2], img[i1 + 1, i2]
        else:
            a, b = img[i1, i2 - 1], img[i1, i2 + 1]
        j2 = -1 if a >= b else 1                 # max([a, b]) keeps the first on ties
        q = a if j2 < 0 else b
        dd = np.array([q / (q + p0), q / (q - p0)]) if (q - p0) != 0 else np.array([q / (q + p0), np.inf])
        dd = np.real(dd)
        dd[np.abs(dd) > 1] = 0
        d.append(np.sign(j2) * np.max(dd))
    return pos[0] + d[0], pos[1] + d[1]


def phi_shift_ms(p2, bsize=8):
    """PhiShiftMS(p2, 1, 1) (Codes/sub_code/PhiShiftMS.m:1-53,125-187): a plane a x + b y + c
    fitted by least squares to the border of width bsize - 1 (pixels with a nonzero value
    after the interquartile outlier cut) is subtracted.  x runs along the columns, y
    along the rows, both from 1."""
    p2 = np.asarray(p2, dtype=np.float64)
    imY, imX = p2.shape
    mask = np.ones((imY, imX))
    mask[bsize - 1:imY - bsize + 1, bsize - 1:imX - bsize + 1] = 0
    XX, YY = np.meshgrid(np.arange(1, imX + 1, dtype=np.float64), np.arange(1, imY + 1, dtype=np.float64))
    pm = p2 * mask
    vals = pm[pm != 0]
    p25, p75 = _matlab_prctile(vals, 25), _matlab_prctile(vals, 75)
    cmin, cmax = p25 - 1.5 * (p75 - p25), p75 + 1.5 * (p75 - p25)
    pm[pm < cmin] = 0
    pm[pm > cmax] = 0
    keep = pm.ravel(order='F') != 0
    AA = np.stack([XX.ravel(order='F')[keep], YY.ravel(order='F')[keep], np.ones(int(keep.sum()))], axis=1)
    rhs = pm.ravel(order='F')[keep]
    coeff = np.linalg.solve(AA.T @ AA, AA.T @ rhs)
    return p2 - coeff[2] - coeff[0] * XX - coeff[1] * YY


def _circshift2(a, s1, s2):
    return np.roll(a, (int(s1), int(s2)), axis=(0, 1))


def _mode_smallest(v):
    """MATLAB mode: the most frequent value, the smallest of equally frequent ones."""
    vals, counts = np.unique(np.asarray(v), return_counts=True)
    return vals[np.argmax(counts)]


def _isoutlier(x):
    """MATLAB isoutlier default: more than three scaled median absolute deviations
    (1.4826 MAD) from the median."""
    x = np.asarray(x, dtype=np.float64)
    med = np.median(x)
    mad = 1.4826 * np.median(np.abs(x - med))
    return np.abs(x - med) > 3 * mad


class FIELD_EXPERIMENTAL_RETRIEVAL:
    """Codes/FIELD_EXPERIMENTAL_RETRIEVAL.m.  unwrap: the phase unwrapper (default: the
    GPU Goldstein unwrap of this library, cbs_phase_unwrap)."""

    get_default_parameters = staticmethod(get_default_parameters)

    def __init__(self, params, device=0, unwrap=None):
        self.parameters = dict(params)
        self.utility = None
        self.device = device
        if unwrap is None:
            from .rytov import unwrap_phase
            unwrap = lambda w: unwrap_phase(w, device=self.device)       # noqa: E731
        self._unwrap = unwrap

    def get_fields(self, bg, sp, ROI=None):
        """:28-214 for hologram stacks bg, sp (n1, n2[, A]) and an optional numeric ROI
        [r0 r1 c0 c1] (1-based, inclusive).  Returns (input_field, output_field,
        updated_optical_parameters, k0s, is_overexposed, ROI); the fields are
        (n1 - 4', n2 - 4', 1, A) complex after the resolution matching."""
        p = self.parameters
        inp = np.asarray(bg, dtype=np.float64)
        out = np.asarray(sp, dtype=np.float64)
        if inp.ndim == 2:
            inp, out = inp[:, :, None], out[:, :, None]
        if inp.shape != out.shape:
            raise ValueError('Background and sample field must be of same size')
        if inp.shape[0] != inp.shape[1]:
            raise ValueError('the image must be a square')
        if p['resolution_image'][0] != p['resolution_image'][1]:
            raise ValueError('please enter an isotropic resolution for the image')
        if p['resolution'][0] != p['resolution'][1]:
            raise ValueError('please enter an isotropic resolution for the output image')
        if ROI is not None:
            if isinstance(ROI, str):
                raise ValueError('the interactive ROI selection is not available; give [r0 r1 c0 c1]')
            r = [int(v) for v in ROI]
            inp = inp[r[0] - 1:r[1], r[2] - 1:r[3]]
            out = out[r[0] - 1:r[1], r[2] - 1:r[3]]
        is_overexposed = np.max(np.maximum(inp, out), axis=(0, 1)) > 254                   # :106
        inp = centered_fft2(inp)                                                           # :108-109
        out = centered_fft2(out)
        n1, n2 = inp.shape[:2]
        c = p['cutout_portion']
        nb = inp[:, :, 0].copy()                                                           # :112-122
        nb[int(round_half_away(n1 * c)) - 1:, :] = 0
        if p['other_corner']:
            nb[:, :int(round_half_away(n2 * (1 - c)))] = 0
        else:
            nb[:, int(round_half_away(n2 * c)) - 1:] = 0
        a = np.abs(nb)
        hit = np.argwhere((a == a.max()).T)                                                # column-major order
        c2, c1 = hit[0]
        shift = lambda F: np.fft.fftshift(_circshift2(F, -c1, -c2), axes=(0, 1))          # noqa: E731  :126-127
        inp, out = shift(inp), shift(out)
        old = inp.shape[0]                                                                 # :131-145
        ratio = p['resolution'][0] / p['resolution_image'][0]
        if ratio >= 1:
            cs = int(round_half_away(0.5 * (old - old / ratio)))
            if cs > 0:
                inp, out = inp[cs:old - cs, cs:old - cs], out[cs:old - cs, cs:old - cs]
        else:
            pad = -int(round_half_away(0.5 * (old - old / ratio)))
            w = ((pad, pad), (pad, pad), (0, 0))
            inp, out = np.pad(inp, w), np.pad(out, w)
        res = list(p['resolution'])
        res[0] = p['resolution_image'][0] * old / inp.shape[0]
        res[1] = p['resolution_image'][1] * old / inp.shape[1]
        p['resolution'] = res
        size = list(p['size'])
        size[0], size[1] = inp.shape[0], inp.shape[1]
        p['size'] = size
        u = derive_optical_tool(p)                                                         # :151-155
        inp = inp * u['NA_circle'][:, :, None]
        out = out * u['NA_circle'][:, :, None]
        k0s = np.zeros((2, inp.shape[2]))                                                  # :157-169
        for j in range(inp.shape[2]):
            k0s[:, j] = peak_subpixel_positioner(np.abs(inp[:, :, j]))
        inp = centered_ifft2(inp)                                                          # :171-172
        out = centered_ifft2(out)
        inp, out = inp[2:-2, 2:-2], out[2:-2, 2:-2]                                        # :177-178
        size[0], size[1] = inp.shape[0], inp.shape[1]
        p['size'] = size
        a = np.sum(np.abs(inp), axis=(0, 1)) / (inp.shape[0] * inp.shape[1])               # :182-184
        inp, out = inp / a, out / a
        if p['conjugate_field']:
            inp, out = np.conj(inp), np.conj(out)
        ratio_f = out / inp
        phase = np.asarray(self._unwrap(np.angle(ratio_f).astype(np.float32)), dtype=np.float64)   # :190
        amp = np.abs(ratio_f)
        for j in range(phase.shape[2]):
            phase[:, :, j] = phi_shift_ms(phase[:, :, j])                                  # :194-198
        out = amp * inp * np.exp(1j * phase)                                               # :199
        inp = inp.reshape(inp.shape[0], inp.shape[1], 1, -1)
        out = out.reshape(out.shape[0], out.shape[1], 1, -1)
        self.utility = derive_optical_tool(p)
        base = basic_optical_parameter()                                                   # :208-211
        updated = {k: p[k] for k in base}
        return inp, out, updated, k0s, is_overexposed, ROI

    def get_aberration_pattern(self, input_field, output_field, correct_params):
        """:216-352: pupil stitching from the field ratios, relative phase matching in
        order of illumination angle, defocus pre-compensation over distList, a low-pass
        of the point-spread function, and the correction of both field stacks."""
        u = self.utility
        inp = np.asarray(input_field)
        out = np.asarray(output_field)
        ratio = np.squeeze(out / inp)
        if ratio.ndim == 2:
            ratio = ratio[:, :, None]
        fE = centered_fft2(ratio)                                                          # :218
        fMask = u['coorxy'] < u['coorxy'].max() * correct_params['overlap_factor']
        fdp = fE * np.conj(fE[:, :, :1])
        fdp = np.exp(1j * np.angle(fdp)) * fMask[:, :, None]
        k0s = np.asarray(correct_params['k0s'], dtype=np.float64)
        P = np.zeros_like(fE)
        for i in range(P.shape[2]):
            s = round_half_away(k0s[:, i])
            P[:, :, i] = _circshift2(fdp[:, :, i], s[0], s[1])
        P = P * u['NA_circle'][:, :, None]
        seq = np.argsort(np.sqrt(k0s[0] ** 2 + k0s[1] ** 2), kind='stable')               # :237
        sumT = 0
        for i in range(1, P.shape[2]):                                                     # :270-275
            sumT = sumT + P[:, :, seq[i - 1]]
            cur = P[:, :, seq[i]]
            both = (np.abs(sumT) > 0) & (np.abs(cur) > 0)
            b = np.angle(np.sum(sumT * np.conj(cur) * both))
            P[:, :, seq[i]] = cur * np.exp(1j * b)
        P2 = np.sum(P, axis=2)
        dist_list = np.asarray(correct_params['distList'], dtype=np.float64)               # :287-300
        cc = [abs(np.sum(np.exp(u['refocusing_kernel'] * d) * np.exp(-1j * np.angle(P2)) * u['NA_circle']))
              for d in dist_list]
        cmax = int(np.argmax(cc))
        approx = dist_list[cmax] * u['refocusing_kernel']
        n1, n2 = approx.shape
        approx = approx - approx[n1 // 2, n2 // 2]
        dist = _mode_smallest(np.diff(round_half_away(k0s[1])))                            # :310-312
        dist = abs(float(np.mean(dist))) * 2
        psf = centered_ifft2(P2 * np.exp(1j * approx))                                     # :316-320
        lpf = np.zeros(psf.shape)
        c1, c2 = n1 // 2, n2 // 2                                                          # floor(end/2+1), 0-based
        lpf[c1 - int(np.floor(n1 / dist)):c1 + int(np.ceil(n1 / dist)) + 1,
            c2 - int(np.floor(n2 / dist)):c2 + int(np.ceil(n2 / dist)) + 1] = 1
        psf = psf * lpf
        P3 = centered_fft2(psf) * np.exp(1j * approx)                                      # :340-342
        pattern = P3 * np.exp(-1j * np.angle(P3[c1, c2])) * u['NA_circle']
        params = {'pattern': pattern, 'utility': u}
        corr = np.exp(-1j * np.angle(pattern))
        shp = corr.shape + (1,) * (inp.ndim - 2)
        inp_f = centered_ifft2(centered_fft2(inp) * corr.reshape(shp))
        out_f = centered_ifft2(centered_fft2(out) * corr.reshape(shp))
        return inp_f, out_f, params

    def correct_aberration(self, input_field, field_trans, aberration_params):
        """:354-393: the pattern resampled to this field grid (Fourier-space crop or pad to
        the matching field of view), cut to the NA, and removed from both stacks."""
        pattern = aberration_params['pattern']
        au = aberration_params['utility']
        u = self.utility
        inp = np.asarray(input_field)
        n = inp.shape[0]
        target = int(round_half_away(n * au['res'][0] / u['res'][0] / 2) * 2)
        if target > n:
            pat = centered_fft2(pattern)
            cs = (target - pat.shape[0]) // 2
            pat0 = np.zeros((target, target), dtype=complex)
            pat0[cs:target - cs, cs:target - cs] = pat
            pattern = centered_ifft2(pat0)
        elif target < n:
            pat = centered_fft2(pattern)
            cs = (pat.shape[0] - target) // 2
            pattern = centered_ifft2(pat[cs:pat.shape[0] - cs, cs:pat.shape[0] - cs])
        if pattern.shape[0] > n:
            cs = (pattern.shape[0] - n) // 2
            pattern = centered_ifft2(pattern[cs:pattern.shape[0] - cs, cs:pattern.shape[0] - cs])
        elif pattern.shape[0] < n:
            pat0 = np.zeros((target, target), dtype=complex)
            cs = (target - pattern.shape[0]) // 2
            pat0[cs:target - cs, cs:target - cs] = pattern
            pattern = pat0
        c1 = au['coor'][0] * pattern.shape[0] / target
        c2 = au['coor'][1] * pattern.shape[1] / target
        na = np.sqrt(c1[:, None] ** 2 + c2[None, :] ** 2) < u['kmax']
        pattern = pattern * na
        corr = np.exp(-1j * np.angle(pattern))
        shp = corr.shape + (1,) * (inp.ndim - 2)
        return (centered_ifft2(centered_fft2(inp) * corr.reshape(shp)),
                centered_ifft2(centered_fft2(np.asarray(field_trans)) * corr.reshape(shp)))

    def refocus(self, field_trans, z):
        """:395-397: propagation by z with the angular spectrum exp(z i 2 pi k3)."""
        f = np.asarray(field_trans)
        k = np.exp(z * self.utility['refocusing_kernel'])
        return centered_ifft2(centered_fft2(f) * k.reshape(k.shape + (1,) * (f.ndim - 2)))

    def remove_dirty_fields(self, input_field, output_field, prctile=95):
        """:399-450 without the manual review: an angle is dropped when the spread of its
        amplitude-difference gradient is an outlier among the angles, or when more than
        prctile % of its unwrapped phase lies outside 1.5x the first angle's phase range;
        the first angle is always kept."""
        from .rytov import vector2scalarfield
        bg, sp = vector2scalarfield(np.asarray(input_field), np.asarray(output_field))
        bg = bg.reshape(bg.shape[0], bg.shape[1], -1)
        sp = sp.reshape(sp.shape[0], sp.shape[1], -1)
        amp = np.abs(sp / bg)
        phase = np.asarray(self._unwrap(np.angle(sp / bg).astype(np.float32)), dtype=np.float64)
        p0 = phase[:, :, 0]
        b1 = np.zeros(phase.shape[2])
        b2 = np.zeros(phase.shape[2], dtype=bool)
        for k in range(phase.shape[2]):
            f1 = amp[:, :, k] - amp[:, :, 0]
            gy, gx = np.gradient(f1)                   # MATLAB [XX, YY] = gradient: XX along columns
            b1[k] = np.std((gx ** 2 + gy ** 2).ravel(), ddof=1)
            f2 = (phase[:, :, k] > p0.max() * 1.5).astype(int) + (phase[:, :, k] < p0.min() * 1.5).astype(int)
            b2[k] = f2.sum() / p0.size > prctile / 100
        keep = np.flatnonzero(_isoutlier(b1).astype(int) + b2.astype(int) == 0)
        keep = np.unique(np.concatenate([[0], keep]))
        inp = np.asarray(input_field).reshape(bg.shape[0], bg.shape[1], -1)[:, :, keep]
        out = np.asarray(output_field).reshape(bg.shape[0], bg.shape[1], -1)[:, :, keep]
        return inp[:, :, None, :], out[:, :, None, :]
