"""The oracle and the C++ restatements against the committed golden fixtures (tests/golden/).

The fixtures hold seeded inputs and the FP64 brute-force outputs (direct sums over every
rotation, direct single-rotation sums, closed-form CTF), written by tests/golden/make_golden.py.
Here the oracle's FFT route, its gradient routine, its CTF model and the C++ FP64 brute force
must reproduce them; a change to any of them that moves a result shows up as a fixture diff.
"""
import os

import numpy as np
import pytest

import cpu_port
import pftcc_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["small_n48", "box128_n192"]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))


def oracle_from(g):
    n, kfrom, kto, ldim = int(g["n"]), int(g["kfrom"]), int(g["kto"]), int(g["ldim"])
    grid = O.make_grid(n, kfrom, kto, ldim)
    return O.Pftcc(grid=grid, kstop=kto, refs_even=g["refs_even"], refs_odd=g["refs_odd"], ptcls=g["ptcls"],
                   ctfmats=g["ctf_fp32"], iseven=g["iseven"].astype(bool))


def ctf_list(g):
    return [O.CtfParams(smpd=c[0], kv=c[1], cs=c[2], fraca=c[3], dfx=c[4], dfy=c[5], angast=c[6], phshift=c[7],
                        phaseplate=bool(c[8])) for c in g["ctfparams"]]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_fft_route_reproduces_golden_gencorrs(name):
    g = load(name)
    pf = oracle_from(g)
    cc = g["gencorrs"]
    for i in range(cc.shape[0]):
        for r in range(cc.shape[1]):
            got = pf.gencorrs(r, i, None, method="fft")
            assert np.max(np.abs(got - cc[i, r, 0])) <= 1e-12 * np.max(np.abs(cc[i, r, 0]))
            for s, sh in enumerate(g["shifts"]):
                got = pf.gencorrs(r, i, tuple(sh), method="fft")
                assert np.max(np.abs(got - cc[i, r, s + 1])) <= 1e-12 * np.max(np.abs(cc[i, r, s + 1]))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_gradients_ctf_and_tables_reproduce_golden(name):
    g = load(name)
    pf = oracle_from(g)
    for t in range(len(g["q_f"])):
        f, gr = pf.gencorr_grad_for_rot_8(int(g["q_ref"][t]), int(g["q_ptcl"][t]), g["q_sh"][t], int(g["q_rot"][t]))
        assert f == pytest.approx(g["q_f"][t], rel=1e-13, abs=1e-15)
        assert np.allclose(gr, g["q_grad"][t], rtol=1e-12, atol=1e-15)
    for i, p in enumerate(ctf_list(g)):
        assert np.max(np.abs(O.ctf_abs_mat(pf.grid, p) - g["ctf64"][i])) <= 1e-14
    assert np.max(np.abs(O.shmat_closed(pf.grid, tuple(g["shifts"][0])) - g["shmat"])) <= 1e-14
    assert [O.magic_pftsz(int(r)) for r in g["magic_ring2"]] == list(g["magic_pftsz"])


@pytest.mark.parametrize("name", NAMES)
def test_cpp_brute_force_reproduces_golden(name):
    g = load(name)
    cc = g["gencorrs"]
    even = np.where(g["iseven"] != 0)[0]
    ks = int(g["kto"]) - int(g["kfrom"]) + 1
    out = cpu_port.gencorrs_direct64(g["refs_even"], g["ptcls"][even], g["ctf_fp32"][even], None, ks, False, 2)
    for a, i in enumerate(even):
        for r in range(cc.shape[1]):
            assert np.max(np.abs(out[r, a] - cc[i, r, 0])) <= 1e-12 * np.max(np.abs(cc[i, r, 0]))
