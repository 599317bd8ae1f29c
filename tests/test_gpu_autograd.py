"""The adjoint kernel (pftcc_grad_batch) through the torch.autograd.Function wrapper.

Runs in a child process that loads the HIP library before torch, with LD_BIND_NOW=1, so that the
library's calls are bound to the system HIP runtime (torch is only used for CPU tensors there).
Checks: torch.autograd.gradcheck (central finite differences, h = 1e-3 px) on the shifts, and the
gradient against the oracle's gencorr_grad_for_rot_8 restatement (1e-9 relative, FP64 on both sides).
"""
import os
import subprocess
import sys
import textwrap

import pytest

from gpu_helpers import say

pytestmark = pytest.mark.gpu

CHILD = textwrap.dedent("""
    import os, sys
    root = sys.argv[1]
    for sub in ("tests", "oracle", "simple3.0_amd"):
        sys.path.insert(0, os.path.join(root, sub))
    import numpy as np
    import pftcc
    pftcc.load_library()
    from cases import make_case
    from gpu_helpers import to_lib
    pf, cps = make_case(seed=71, n=308, kfrom=2, kto=30, ctf=True, nrefs=4, nptcls=6, ldim=200, eo_mixed=True)
    lib = to_lib(pf, cps)
    import torch
    from pftcc.autograd import corr_for_rot
    rng = np.random.default_rng(72)
    npairs = 8
    irefs = rng.integers(1, 5, npairs)
    iptcls = rng.integers(1, 7, npairs)
    irots = rng.integers(1, 617, npairs)
    shifts = torch.tensor(rng.uniform(-4, 4, (npairs, 2)), dtype=torch.float64, requires_grad=True)
    fn = lambda s: corr_for_rot(lib, irefs, iptcls, s, irots)
    ok = torch.autograd.gradcheck(fn, (shifts,), eps=1e-3, atol=1e-6, rtol=1e-4)
    print("gradcheck", ok, flush=True)
    f = fn(shifts)
    f.sum().backward()
    worst = 0.0
    for t in range(npairs):
        fo, go = pf.gencorr_grad_for_rot_8(irefs[t] - 1, iptcls[t] - 1, shifts[t].detach().numpy(), irots[t])
        assert abs(f[t].item() - fo) <= 1e-9 * max(1.0, abs(fo))
        d = np.abs(shifts.grad[t].numpy() - go) / np.maximum(1.0, np.abs(go))
        worst = max(worst, float(np.max(d)))
    assert worst <= 1e-9, worst
    print("max relative gradient error vs oracle", worst, flush=True)
    lib.kill()
""")


def test_autograd_wrapper_with_gpu_backend(repo_root):
    say("autograd wrapper over pftcc_grad_batch (child process)")
    env = dict(os.environ, LD_BIND_NOW="1")
    r = subprocess.run([sys.executable, "-c", CHILD, repo_root], env=env, capture_output=True, text=True,
                       timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "gradcheck True" in r.stdout
