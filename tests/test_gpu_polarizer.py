"""GPU parity of the polarizer (pftcc_init_polarizer / pftcc_polarize[_dev]) against the oracle
restatement of init_polarizer / polarize (simple_polarizer.f90:41-151).

Tolerance: the kernel and the oracle use the same FP32 weights up to a few ulps and accumulate
the 9 products in FP32 (kernel) or FP64 (oracle), so the polar values agree within 2e-6 of the
largest one.  Particle slots must also carry refreshed sqsums (set_ptcl_pft, :438-446).
"""
import numpy as np
import pytest

import pftcc_oracle as O
import polarizer_oracle as P
from gpu_helpers import say
from pftcc import Params, PolarFTCorrCalc, PftccError

pytestmark = pytest.mark.gpu


def _images(rng, count, L):
    """Real images with a smooth spectrum plus noise, and their transforms as the image object
    holds them (cmat, Fortran order)."""
    out = []
    yy, xx = np.mgrid[0:L, 0:L].astype(np.float64)
    for _ in range(count):
        img = rng.standard_normal((L, L)) * 0.2
        for _b in range(4):
            cx, cy = rng.uniform(-L / 5, L / 5, 2)
            s = rng.uniform(1.5, 4.0)
            img += np.exp(-((xx - L / 2 - cx) ** 2 + (yy - L / 2 - cy) ** 2) / (2 * s * s))
        out.append(P.image_fft(img))
    return np.stack(out)


def _handle(L, n, kfrom, kto, nrefs, nptcls, kstop=None, pfrom=1):
    prm = Params(boxmatch=L, kfromto=(kfrom, kto), kstop=kstop or kto, pftsz=n, ctf=False)
    return PolarFTCorrCalc().new(nrefs, (pfrom, pfrom + nptcls - 1), prm)


@pytest.mark.parametrize("L,n,kfrom,kto,griddev", [(64, 48, 2, 30, False), (128, 192, 3, 40, False),
                                                   (64, 48, 2, 32, True)])
def test_polarize_particles_and_references_match_oracle(L, n, kfrom, kto, griddev):
    say(f"polarize parity: box {L}, pftsz {n}, k [{kfrom},{kto}], griddev {griddev}")
    rng = np.random.default_rng(L + n + kto)
    nptcls, nrefs = 6, 3
    lib = _handle(L, n, kfrom, kto, nrefs, nptcls, kstop=kto - 2, pfrom=11)
    lib.init_polarizer(griddev=griddev)
    g = O.make_grid(n, kfrom, kto, L)
    tab = P.polarizer_tables(g.polar_x, g.polar_y, n, L, griddev=griddev)
    imgs = _images(rng, nptcls + 2 * nrefs, L)
    # particles in scrambled order, references even and odd
    ids = rng.permutation(nptcls) + 11
    lib.polarize(ids, imgs[:nptcls], isptcl=True)
    lib.polarize(np.arange(1, nrefs + 1), imgs[nptcls:nptcls + nrefs], isptcl=False, iseven=True)
    lib.polarize(np.arange(1, nrefs + 1), imgs[nptcls + nrefs:], isptcl=False, iseven=False)
    for b, i in enumerate(ids):
        ref = P.polarize_vec(tab, imgs[b])
        got = lib.get_ptcl_pft(int(i))
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max(), (b, i)
        sq = float(np.sum(np.abs(ref[: kto - 2 - kfrom + 1]) ** 2))
        assert lib.get_sqsum_ptcl(int(i)) == pytest.approx(sq, rel=2e-5)
    for r in range(nrefs):
        for even, off in ((True, nptcls), (False, nptcls + nrefs)):
            ref = P.polarize_vec(tab, imgs[off + r])
            got = lib.get_ref_pft(r + 1, even)
            assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max(), (r, even)
    lib.kill()


def test_polarize_band_mask_and_device_source():
    say("polarize: band mask, device-resident transforms")
    from pftcc.synth import DeviceBuffer
    L, n, kfrom, kto = 64, 48, 2, 24
    rng = np.random.default_rng(7)
    lib = _handle(L, n, kfrom, kto, 2, 5)
    lib.init_polarizer()
    g = O.make_grid(n, kfrom, kto, L)
    tab = P.polarizer_tables(g.polar_x, g.polar_y, n, L)
    imgs = _images(rng, 5, L)
    mask = np.ones(40, np.int8)          # the reference's 1-based band mask, mask(k) for shell k
    mask[[4, 9, 17]] = 0                 # shells 5, 10, 18 out of band
    d = DeviceBuffer.from_array(imgs)
    lib.polarize_dev(np.arange(1, 6), d.ptr, isptcl=True, mask=mask)
    d.close()
    shells = np.arange(kfrom, kto + 1)
    keep = mask[shells - 1].astype(bool)
    for b in range(5):
        ref = P.polarize_vec(tab, imgs[b], mask=keep)
        got = lib.get_ptcl_pft(b + 1)
        assert np.all(got[~keep] == 0)
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
    lib.kill()


def test_polarize_errors():
    say("polarize: error paths")
    L, n = 64, 48
    lib = _handle(L, n, 2, 20, 2, 3)
    imgs = np.zeros((2, L, L // 2 + 1), np.complex64)
    with pytest.raises(PftccError, match="init_polarizer"):
        lib.polarize([1, 2], imgs, isptcl=True)
    lib.init_polarizer()
    with pytest.raises(PftccError, match="twice"):
        lib.polarize([1, 1], imgs, isptcl=True)
    with pytest.raises(PftccError, match="outside"):
        lib.polarize([1, 4], imgs, isptcl=True)
    with pytest.raises(PftccError, match="mask"):
        lib.polarize([1, 2], imgs, isptcl=True, mask=np.ones(10, np.int8))
    with pytest.raises(PftccError):
        lib.polarize([1, 3], imgs, isptcl=False)
    lib.kill()
    big = _handle(32, 48, 2, 20, 1, 1)            # kfromto(2) = 20 beyond Nyquist 16
    with pytest.raises(PftccError, match="Nyquist"):
        big.init_polarizer()
    big.kill()


def test_polarized_particles_search_back_to_their_reference():
    """End to end: projections of a Gaussian-blob phantom rotated in-plane by known angles are
    polarized; the exhaustive search must return each particle's source and rotation."""
    say("polarize -> search: rotated copies of the references")
    L, kfrom, kto = 64, 2, 20
    n = O.magic_pftsz(27)
    rng = np.random.default_rng(11)
    yy, xx = np.mgrid[0:L, 0:L].astype(np.float64)
    refs_img = []
    for _ in range(4):
        img = np.zeros((L, L))
        for _b in range(5):
            cx, cy = rng.uniform(-12, 12, 2)
            s = rng.uniform(1.5, 3.5)
            img += rng.uniform(0.5, 1.5) * np.exp(-((xx - L / 2 - cx) ** 2 + (yy - L / 2 - cy) ** 2) / (2 * s * s))
        refs_img.append(img)
    # particle p = reference p rotated in-plane by a whole number of angular steps, built from the
    # reference's polarized rings (p(j) = r(j - rho) on the full ring of 2 pftsz angles)
    lib = _handle(L, n, kfrom, kto, 4, 4)
    lib.init_polarizer()
    cm = np.stack([P.image_fft(im) for im in refs_img])
    lib.polarize(np.arange(1, 5), cm, isptcl=False, iseven=True)
    lib.polarize(np.arange(1, 5), cm, isptcl=False, iseven=False)
    rots = [1, 17, 40, 2 * n - 5]
    ptcls = []
    for p in range(4):
        full = O.full_ring(lib.get_ref_pft(p + 1, True).astype(np.complex128))
        rho = rots[p] - 1
        # particle(j) = ref(j - rho): the search's best rotation is then irot = rots[p]
        ptcls.append(np.roll(full, rho, axis=1)[:, :n].astype(np.complex64))
    lib.set_ptcl_pfts(1, np.stack(ptcls))
    lib.memoize_ffts()
    cc, ir, _ish, rot = lib.search(np.arange(1, 5))
    lib.kill()
    assert list(ir) == [1, 2, 3, 4]
    assert list(rot) == rots
    assert np.all(cc > 0.9999)
