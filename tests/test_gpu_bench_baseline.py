"""GPU: the reference-form baseline that bench.py times beside the product (the
reference's gpuArray loop of solve_raw restated op for op with torch/rocFFT) gives
the same Born-loop result as the product, so the speed ratio compares like with like.

It runs in a child process that creates torch's HIP context before the solver
library's, the order bench.py uses (torch finds no device when it comes second)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import torch
torch.zeros(1, device='cuda:0')
import bench
from helpers import params, sphere_ri, random_field, rel
from inverse_solver_amd import FORWARD_SOLVER_CONVERGENT_BORN
for M in (5, 8):
    ri = (16, 16, 16)
    s = FORWARD_SOLVER_CONVERGENT_BORN(params(ri, iterations=M))
    s.set_RI(sphere_ri(ri, 0.5, n_s=1.42))
    u = random_field(ri, 9)
    got = s.solve_incident(u)
    ms, ref_form = bench.reference_form_gpu(s, u, M, lambda *a: None)
    assert ref_form.shape == got.T.shape
    print('RESULT', M, rel(ref_form, got.T), flush=True)
    s.close()
'''


def test_reference_form_baseline_matches_product():
    code = CHILD.format(root=ROOT, tests=os.path.join(ROOT, 'tests'))
    r = subprocess.run([sys.executable, '-c', code], capture_output=True, text=True, timeout=600, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    res = [line.split() for line in r.stdout.splitlines() if line.startswith('RESULT')]
    assert len(res) == 2
    for _, M, err in res:
        assert float(err) < 1e-5, (M, err)
