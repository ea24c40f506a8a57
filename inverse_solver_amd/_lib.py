"""ctypes binding of libcbs_hip.so (the C ABI declared in include/cbs_hip.h).

The library is the only compute path: importing this module fails loudly if
the built library is missing, and every call raises CbsError on a non-zero
status, with the library's own message.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libcbs_hip.so')

CBS_OK, CBS_ERR_ARG, CBS_ERR_HANDLE, CBS_ERR_HIP, CBS_ERR_UNSUPPORTED, CBS_ERR_STATE, CBS_ERR_MEMORY = range(7)
CBS_MEM_HOST, CBS_MEM_DEVICE = 0, 1
CBS_FORM_FUSED, CBS_FORM_PLAIN, CBS_FORM_ROWS, CBS_FORM_AXES = 0, 1, 2, 3

EXPORTED = [
    'cbs_create', 'cbs_destroy', 'cbs_set_options', 'cbs_get_options', 'cbs_set_potential', 'cbs_potential_max',
    'cbs_set_potential_ri', 'cbs_solve_3d',
    'cbs_set_io_kernels', 'cbs_set_io_kernels_octant', 'cbs_solve_2d', 'cbs_get_norms', 'cbs_get_kernel_times',
    'cbs_get_stream',
    'cbs_footprint', 'cbs_get_plan', 'cbs_autotune', 'cbs_tv_prox', 'cbs_tv_prox_complex', 'cbs_field_gradient',
    'cbs_phase_residues', 'cbs_phase_integrate', 'cbs_phase_branch_cuts', 'cbs_phase_unwrap',
    'cbs_set_convergence', 'cbs_get_iterations', 'cbs_solve_3d_noalloc', 'cbs_fista_prox_input',
    'cbs_fista_momentum', 'cbs_field_pupil', 'cbs_solve_field', 'cbs_last_error', 'cbs_abi_version',
]


class CbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'[cbs status {code}] {msg}')
        self.code = code


class Geometry(C.Structure):
    _fields_ = [('simulation_size', C.c_uint32 * 3), ('RI_size', C.c_uint32 * 3), ('field_size', C.c_uint32 * 3),
                ('conv_size', C.c_uint32 * 3), ('device', C.c_int32), ('RI_center', C.c_int32 * 2),
                ('pole_num', C.c_int32)]


class PotentialParams(C.Structure):
    _fields_ = [('k0_nm', C.c_double), ('eps_imag', C.c_double), ('fourier_res', C.c_double * 3),
                ('ROI', C.c_uint32 * 6), ('Bornmax', C.c_uint32), ('dV', C.c_double),
                ('boundary_sharpness', C.c_double), ('acyclic', C.c_uint32), ('shift_axes', C.c_uint32),
                ('ramp_axes', C.c_uint32)]


class Options(C.Structure):
    _fields_ = [('form', C.c_int32), ('split_a', C.c_int32), ('tile_t', C.c_int32), ('max_batch', C.c_int32),
                ('track_norms', C.c_int32), ('use_graph', C.c_int32), ('time_kernels', C.c_int32),
                ('threads_real', C.c_int32), ('threads_kspace', C.c_int32), ('block_order', C.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is missing: build it with __graft_entry__.build() '
                          f'(make -C inverse_solver_amd/csrc). There is no CPU fallback.')
    lib = C.CDLL(LIB_PATH)
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    sig = {
        'cbs_create': ([C.POINTER(Geometry), C.POINTER(vp)], i32),
        'cbs_destroy': ([vp], i32),
        'cbs_set_options': ([vp, C.POINTER(Options)], i32),
        'cbs_get_options': ([vp, C.POINTER(Options)], i32),
        'cbs_set_potential': ([vp, C.POINTER(PotentialParams), vp, i32], i32),
        'cbs_potential_max': ([vp, vp, i32, C.c_double, C.c_double, i32, C.POINTER(C.c_double)], i32),
        'cbs_set_potential_ri': ([vp, C.POINTER(PotentialParams), vp, i32, i32, C.c_double, C.c_double, i32], i32),
        'cbs_solve_3d': ([vp, vp, u32, vp, i32], i32),
        'cbs_set_io_kernels': ([vp, vp, vp, vp, i32], i32),
        'cbs_set_io_kernels_octant': ([vp, vp, vp, i32], i32),
        'cbs_solve_2d': ([vp, vp, u32, vp, vp, vp, i32], i32),
        'cbs_get_norms': ([vp, C.POINTER(C.c_double), u32, u32], i32),
        'cbs_get_kernel_times': ([vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64), u32], i32),
        'cbs_get_stream': ([vp, C.POINTER(vp)], i32),
        'cbs_footprint': ([vp, u32, C.POINTER(C.c_uint64)], i32),
        'cbs_get_plan': ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                          C.POINTER(C.c_int32)], i32),
        'cbs_autotune': ([vp, u32, u32, C.POINTER(C.c_double)], i32),
        'cbs_tv_prox': ([vp, vp, vp, C.POINTER(C.c_uint32), C.c_double, C.c_double, C.c_double, i32, i32, u32,
                         i32], i32),
        'cbs_tv_prox_complex': ([vp, vp, vp, C.POINTER(C.c_uint32), C.c_double, C.c_double, C.c_double,
                                 C.c_double, i32, i32, u32, i32], i32),
        'cbs_field_gradient': ([vp, vp, vp, u32, C.POINTER(C.c_uint32), vp, i32], i32),
        'cbs_phase_residues': ([i32, vp, vp, C.POINTER(C.c_uint32), i32], i32),
        'cbs_phase_integrate': ([i32, vp, vp, vp, vp, C.POINTER(C.c_uint32), i32], i32),
        'cbs_phase_branch_cuts': ([i32, vp, vp, vp, C.POINTER(C.c_uint32), i32], i32),
        'cbs_phase_unwrap': ([i32, vp, vp, C.POINTER(C.c_uint32), i32], i32),
        'cbs_set_convergence': ([vp, C.c_double], i32),
        'cbs_solve_3d_noalloc': ([vp, vp, u32, vp], i32),
        'cbs_get_iterations': ([vp, C.POINTER(C.c_uint32), u32], i32),
        'cbs_fista_prox_input': ([vp, vp, vp, C.c_double, C.c_uint64, i32, vp], i32),
        'cbs_fista_momentum': ([vp, vp, i32, vp, vp, vp, C.c_uint64, C.c_double, C.c_double], i32),
        'cbs_field_pupil': ([vp, vp, vp, u32, vp, vp, i32], i32),
        'cbs_solve_field': ([vp, vp, u32, vp, vp, vp, vp, vp, i32], i32),
        'cbs_last_error': ([], C.c_char_p),
        'cbs_abi_version': ([], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


lib = _load()


def check(status):
    if status != CBS_OK:
        msg = lib.cbs_last_error()
        raise CbsError(status, msg.decode() if msg else '')
    return status
