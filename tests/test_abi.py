"""The C ABI library loads and exports every entry point include/cbs_hip.h
declares; argument errors are reported without touching a GPU."""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, '..', 'include', 'cbs_hip.h')


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(cbs_\w+)\s*\(', txt, re.M)))


def test_header_declarations_are_exported():
    from inverse_solver_amd import _lib
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(_lib.lib, n), n
    assert set(names) == set(_lib.EXPORTED)
    assert _lib.lib.cbs_abi_version() == 1


def test_argument_errors_without_gpu():
    from inverse_solver_amd import _lib
    lib = _lib.lib
    h = C.c_void_p()
    assert lib.cbs_create(None, C.byref(h)) == _lib.CBS_ERR_ARG
    assert 'null' in lib.cbs_last_error().decode()
    assert lib.cbs_destroy(C.c_void_p(12345)) == _lib.CBS_ERR_HANDLE
    assert 'handle' in lib.cbs_last_error().decode()
    assert lib.cbs_solve_3d(None, None, 1, None, 0) == _lib.CBS_ERR_HANDLE
    # the scalar field entry points check the handle before anything else
    assert lib.cbs_solve_field(None, None, 1, None, None, None, None, None, 0) == _lib.CBS_ERR_HANDLE
    assert lib.cbs_field_pupil(C.c_void_p(12345), None, None, 1, None, None, 0) == _lib.CBS_ERR_HANDLE
    assert 'handle' in lib.cbs_last_error().decode()


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with status 2 and a JSON message."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert 'one GPU' in json.loads(r.stdout.strip().splitlines()[-1])['error']


def test_probe_library_validates_arguments_without_gpu():
    """libcbs_probe.so (the bench's bandwidth and access-pattern probes) loads, exports both
    entry points, and refuses bad arguments before touching a device (status 1)."""
    import ctypes as C
    lib = C.CDLL(os.path.join(HERE, '..', 'inverse_solver_amd', 'libcbs_probe.so'))
    ms = C.c_float(0.0)
    lib.cbs_probe_stream.restype = C.c_int
    lib.cbs_probe_pattern.restype = C.c_int
    fake = C.c_void_p(16)
    # unknown kind, missing arrays, bytes not a multiple of 16
    assert lib.cbs_probe_stream(7, fake, fake, fake, C.c_longlong(1024), 1, None, C.byref(ms)) == 1
    assert lib.cbs_probe_stream(1, fake, None, fake, C.c_longlong(1024), 1, None, C.byref(ms)) == 1
    assert lib.cbs_probe_stream(0, fake, None, fake, C.c_longlong(1000), 1, None, C.byref(ms)) == 1
    dims = (C.c_int * 6)(160, 160, 160, 40, 4, 49)
    bad_ny = (C.c_int * 6)(160, 150, 160, 40, 4, 49)      # ny != a * b
    bad_nx = (C.c_int * 6)(150, 160, 160, 40, 4, 49)      # k-space tiles of 16 need nx % 16 == 0
    assert lib.cbs_probe_pattern(5, fake, fake, fake, fake, dims, 1, None, C.byref(ms)) == 1
    assert lib.cbs_probe_pattern(1, fake, fake, fake, fake, bad_ny, 1, None, C.byref(ms)) == 1
    assert lib.cbs_probe_pattern(0, fake, fake, fake, fake, bad_nx, 1, None, C.byref(ms)) == 1
    assert lib.cbs_probe_pattern(2, fake, fake, fake, None, dims, 1, None, C.byref(ms)) == 1    # odd launches need Field
    assert lib.cbs_probe_pattern(4, fake, fake, fake, None, dims, 1, None, C.byref(ms)) == 1
    assert lib.cbs_probe_pattern(1, fake, fake, fake, fake, dims, 0, None, C.byref(ms)) == 1    # reps >= 1
