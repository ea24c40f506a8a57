"""BACKWARD_SOLVER_MULTI gradient / FISTA loop and TV_FISTA_inner_v2 (FP64
restatement).  Test infrastructure only (see oracle/__init__.py)."""
from __future__ import annotations

import numpy as np


def get_gradiant_static(RI, input_field, output_field, forward_solver):
    """Codes/BACKWARD_SOLVER_MULTI.m:30-54.

    forward_solver: an oracle ForwardSolverConvergentBorn with return_transmission
    and return_3D on.  Returns (gradient on the RI grid, err)."""
    forward_solver.set_RI(RI)                                                   # :32
    trans, _, u3d = forward_solver.solve(input_field)                          # :34
    diff = trans - output_field                                                # :36
    err = float(np.sum(np.abs(diff) ** 2))                                     # :37
    forward_solver.set_RI(np.conj(np.flip(RI, axis=2)))                        # :48
    _, _, w3d = forward_solver.solve(np.conj(diff))                            # :49
    w3d = np.flip(np.conj(w3d), axis=2)                                        # :50
    grad = -(1j * np.mean(np.conj(u3d) * w3d, axis=(3, 4)))                    # :53
    return grad, err


def _tv_L(P):
    """TV_FISTA_inner_v2.m:99-113 (3D): divergence-like adjoint of the forward differences."""
    out = P[0] + P[1] + P[2]
    out = out - np.roll(P[0], 1, axis=0)
    out = out - np.roll(P[1], 1, axis=1)
    out = out - np.roll(P[2], 1, axis=2)
    return out


def _tv_L_trans(x):
    """TV_FISTA_inner_v2.m:114-123 (3D): periodic forward differences x - circshift(x, -1)."""
    return [x - np.roll(x, -1, axis=0), x - np.roll(x, -1, axis=1), x - np.roll(x, -1, axis=2)]


def _project_non_neg(x, Vmin, Vmax, Vimag_max, non_neg, dirichlet):
    """TV_FISTA_inner_v2.m:57-75."""
    x = np.array(x, copy=True)
    if non_neg:
        lo = np.real(x) < Vmin
        x[lo] = Vmin + 1j * np.imag(x[lo]) if np.iscomplexobj(x) else Vmin
        hi = np.real(x) > Vmax
        x[hi] = Vmax + 1j * np.imag(x[hi]) if np.iscomplexobj(x) else Vmax
    if np.iscomplexobj(x):
        neg = np.imag(x) < 0
        x[neg] = np.real(x[neg])
        if Vimag_max > 0:
            big = np.imag(x) > Vimag_max
            x[big] = np.real(x[big]) + 1j * Vimag_max
    if dirichlet:
        x[0, :, :] = 0
        x[-1, :, :] = 0
        x[:, 0, :] = 0
        x[:, -1, :] = 0
        x[:, :, 0] = 0
        x[:, :, -1] = 0
    return x


def _project_tv(P):
    """TV_FISTA_inner_v2.m:77-88 (3D)."""
    A = np.sqrt(np.maximum(np.abs(P[0]) ** 2 + np.abs(P[1]) ** 2 + np.abs(P[2]) ** 2, 1))
    return [P[0] / A, P[1] / A, P[2] / A]


def tv_fista_inner_v2(in_mat, lam, Vmin, Vmax, Vimag_max, non_neg, dirichlet, inner_itt):
    """Codes/sub_code/TV_FISTA_inner_v2.m:1-53 (3D input)."""
    in_mat = np.asarray(in_mat)
    if Vimag_max == 0:                                                          # :10-12
        in_mat = np.real(in_mat)
    dividend = 12                                                               # :16-20 (3D)
    P_np = [np.zeros_like(in_mat) for _ in range(3)]                            # :22-27
    R = [p.copy() for p in P_np]
    t_np = 1.0
    for _ in range(inner_itt):                                                  # :34
        P_n = P_np
        t_n = t_np
        P_np = _tv_L_trans(_project_non_neg(in_mat - lam * _tv_L(R), Vmin, Vmax, Vimag_max, non_neg, dirichlet))
        P_np = [R[k] + 1.0 / (dividend * lam) * P_np[k] for k in range(3)]     # :39-41
        P_np = _project_tv(P_np)                                                # :42
        t_np = (1 + np.sqrt(1 + 4 * t_n ** 2)) / 2                              # :44
        R = [P_np[k] + ((t_n - 1) / t_np) * (P_np[k] - P_n[k]) for k in range(3)]  # :45-47
    return _project_non_neg(in_mat - lam * _tv_L(P_np), Vmin, Vmax, Vimag_max, non_neg, dirichlet)   # :51


def tv_value(x):
    """Isotropic TV of TV_FISTA_inner_v2.m:90-97 (3D)."""
    P = _tv_L_trans(x)
    return float(np.sum(np.sqrt(np.abs(P[0]) ** 2 + np.abs(P[1]) ** 2 + np.abs(P[2]) ** 2)))


def backward_multi_solve(init_RI, input_field, output_field, forward_solver, p):
    """Codes/BACKWARD_SOLVER_MULTI.m:92-196 with the initial RI given (the Rytov
    initialiser of :94 is replaced by init_RI) and num_scan_per_iteration = 0.

    p: dict with RI_bg, step, tv_param, use_non_negativity, nmin, nmax, kappamax,
    inner_itt, itter_max.  Returns (RI, err_list)."""
    RI = np.asarray(init_RI).astype(np.complex128)
    n_m = p['RI_bg']
    alpha = 1.0 / p['step']                                                     # :107
    t_np = 1.0
    u_n = RI ** 2 / n_m ** 2 - 1                                                # :112
    x_n = RI ** 2 / n_m ** 2 - 1                                                # :113
    if RI.shape[2] % 2 == 0:                                                    # :117-119
        raise ValueError('need to be odd size')
    Vmin = (p['nmin'] ** 2 - p['kappamax'] ** 2) / n_m ** 2 - 1                 # :133
    Vmax = p['nmax'] ** 2 / n_m ** 2 - 1                                        # :134
    Vimag_max = 2 * p['kappamax'] * p['nmax'] / n_m ** 2                        # :135
    err_list = []
    for _ in range(p['itter_max']):                                             # :137
        grad, err = get_gradiant_static(RI, input_field, output_field, forward_solver)   # :143
        err_list.append(err)
        t_n = t_np                                                              # :147
        s_n = tv_fista_inner_v2(u_n - (1 / alpha) * grad, p['tv_param'] / alpha, Vmin, Vmax, Vimag_max,
                                p['use_non_negativity'], False, p['inner_itt'])  # :152-153
        t_np = (1 + np.sqrt(1 + 4 * t_n ** 2)) / 2                              # :156
        u_n = s_n + (t_n - 1) / t_np * (s_n - x_n)                              # :157
        x_n = s_n                                                               # :158
        RI = n_m * np.sqrt(u_n.astype(np.complex128) + 1)                       # :159-160
    return RI, err_list
