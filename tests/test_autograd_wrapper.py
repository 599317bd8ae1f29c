"""The torch.autograd.Function wrapper of the batched adjoint, on the FP64 oracle as its backend.

Pins the wrapper's plumbing (forward = objective only, backward = recomputed gradient times the
incoming gradient) with torch.autograd.gradcheck, i.e. central finite differences at h = 1e-3 px
on the shifts.  The GPU backend goes through the same wrapper in tests/test_gpu_autograd.py.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cases import make_case  # noqa: E402
from pftcc.autograd import CorrForRot, corr_for_rot  # noqa: E402


def oracle_backend(pf):
    def backend(irefs, iptcls, shvecs, irots, with_grad):
        f = np.zeros(len(irefs))
        g = np.zeros((len(irefs), 3))
        for t in range(len(irefs)):
            fo, go = pf.gencorr_grad_for_rot_8(irefs[t] - 1, iptcls[t] - 1, shvecs[t], irots[t])
            f[t] = fo
            g[t, :2] = go
            g[t, 2] = pf.grad_angle(irefs[t] - 1, iptcls[t] - 1, shvecs[t], (irots[t] - 1) * np.pi / pf.pftsz)
        return (f, g) if with_grad else f
    return backend


@pytest.mark.parametrize("match_filt", [False, True])
def test_gradcheck_shifts_on_oracle_backend(match_filt):
    pf, _ = make_case(seed=61, n=24, kto=9, ctf=True, nrefs=3, nptcls=3, match_filt=match_filt)
    rng = np.random.default_rng(62)
    npairs = 5
    irefs = rng.integers(1, 4, npairs)
    iptcls = rng.integers(1, 4, npairs)
    irots = rng.integers(1, 2 * pf.pftsz + 1, npairs)
    shifts = torch.tensor(rng.uniform(-4, 4, (npairs, 2)), dtype=torch.float64, requires_grad=True)
    backend = oracle_backend(pf)
    fn = lambda s: corr_for_rot(backend, irefs, iptcls, s, irots)  # noqa: E731
    assert torch.autograd.gradcheck(fn, (shifts,), eps=1e-3, atol=1e-6, rtol=1e-4)


def test_angle_gradient_and_chain_rule():
    pf, _ = make_case(seed=63, n=24, kto=9, ctf=False, nrefs=2, nptcls=2)
    backend = oracle_backend(pf)
    irefs, iptcls, irots = np.array([1, 2]), np.array([2, 1]), np.array([3, 17])
    shifts = torch.tensor([[0.5, -1.0], [2.0, 1.5]], dtype=torch.float64, requires_grad=True)
    dpsi = torch.zeros(2, dtype=torch.float64, requires_grad=True)
    f = CorrForRot.apply(shifts, dpsi, backend, irefs, iptcls, irots)
    (3.0 * f[0] - 2.0 * f[1]).backward()
    _, g = backend(irefs, iptcls, shifts.detach().numpy(), irots, True)
    assert np.allclose(shifts.grad.numpy(), np.array([3.0, -2.0])[:, None] * g[:, :2], rtol=1e-14, atol=0)
    assert np.allclose(dpsi.grad.numpy(), np.array([3.0, -2.0]) * g[:, 2], rtol=1e-14, atol=0)
    # the angle derivative is that of the continuous rotational interpolant (central FD)
    for t in range(2):
        psi = (irots[t] - 1) * np.pi / pf.pftsz
        h = 1e-5
        fd = (pf.corr_at_angle(irefs[t] - 1, iptcls[t] - 1, shifts[t].detach().numpy(), psi + h) -
              pf.corr_at_angle(irefs[t] - 1, iptcls[t] - 1, shifts[t].detach().numpy(), psi - h)) / (2 * h)
        assert abs(fd - g[t, 2]) <= 1e-7 * max(1.0, abs(fd))
    with pytest.raises(ValueError):
        CorrForRot.apply(shifts, torch.ones(2, dtype=torch.float64), backend, irefs, iptcls, irots)
