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
