"""pftcc_search_dev_noalloc (the search that allocates nothing) against pftcc_search_dev.

Both entry points run the same kernels, so the results must be identical: best cc, reference,
shift and rotation of every particle.  The no-allocation form keeps its working memory in one
static array per GPU that every call shares, so the test makes two calls in a row, the second with
fewer particles, references and shifts, and checks that nothing of the first call's contents leaks
into the second.  A handle that only ever runs the no-allocation form must not change the GPU's
free memory across its calls, and a request beyond the static array's size is refused.
"""
import ctypes

import numpy as np
import pytest

from cases import make_case
from gpu_helpers import say, to_lib
import pftcc

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def shift_grid(lim, step):
    g = np.arange(-lim, lim + 1e-6, step)
    return np.array([(x, y) for x in g for y in g], np.float32)


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _outputs(np_):
    return (torch.empty(np_, device="cuda", dtype=torch.float32),
            torch.empty(np_, device="cuda", dtype=torch.int32),
            torch.empty(np_, device="cuda", dtype=torch.int32),
            torch.empty(np_, device="cuda", dtype=torch.int32))


def _search(lib, fn, ids, refs, shifts, outs):
    torch.cuda.synchronize()
    d_ids = _dev(ids, torch.int32)
    d_refs = _dev(refs, torch.int32) if refs is not None else None
    d_sh = _dev(shifts, torch.float32) if shifts is not None and len(shifts) else None
    nr = len(refs) if refs is not None else lib.nrefs
    ns = 0 if shifts is None else len(shifts)
    cc, ir, ish, rot = outs
    fn(len(ids), d_ids.data_ptr(), nr, d_refs.data_ptr() if d_refs is not None else None, ns,
       d_sh.data_ptr() if d_sh is not None else None, cc.data_ptr(), ir.data_ptr(), ish.data_ptr(), rot.data_ptr())
    lib.sync()
    return tuple(t.cpu().numpy().copy() for t in outs)


def _same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y), (x, y)


def test_noalloc_search_equals_search_dev():
    say("no-alloc search: n 308, K 59, box 200, 169 shifts, then a smaller call")
    pf, cps = make_case(seed=611, n=308, kfrom=2, kto=60, ctf=True, nrefs=3, nptcls=20, ldim=200)
    lib_a = to_lib(pf, cps)          # only ever runs the no-allocation search
    lib_b = to_lib(pf, cps)          # the allocating pftcc_search_dev
    big = shift_grid(3.0, 0.5)       # 169 shifts: the staged-reference kernel
    small = shift_grid(1.0, 0.5)     # 25 shifts
    ids1 = np.arange(1, 21, dtype=np.int32)
    ids2 = np.array([7, 3, 19, 11, 2], np.int32)
    refs2 = np.array([3, 1], np.int32)
    outs1, outs2 = _outputs(20), _outputs(5)
    for lib in (lib_a, lib_b):
        tile, staged = ctypes.c_int32(), ctypes.c_int32()
        assert lib._lib.pftcc_get_corr_tile(lib._h, len(big), ctypes.byref(tile), ctypes.byref(staged)) == 0
        assert staged.value == 1
    a1 = _search(lib_a, lib_a.search_dev_noalloc, ids1, None, big, outs1)
    a2 = _search(lib_a, lib_a.search_dev_noalloc, ids2, refs2, small, outs2)
    b1 = _search(lib_b, lib_b.search_dev, ids1, None, big, outs1)
    b2 = _search(lib_b, lib_b.search_dev, ids2, refs2, small, outs2)
    _same(a1, b1)
    _same(a2, b2)
    # the winners of the second call come from its own lists
    assert set(a2[1].tolist()) <= {1, 3} and np.all(a2[2] < len(small))
    # host-list search, the same winners again
    cc, ir, ish, rot = lib_b.search(ids2, refs2, small)
    assert np.array_equal(ir, a2[1]) and np.array_equal(ish, a2[2]) and np.array_equal(rot, a2[3])
    # unshifted form (ns = 0) on the same handle, after the shifted calls
    a3 = _search(lib_a, lib_a.search_dev_noalloc, ids2, None, None, outs2)
    b3 = _search(lib_b, lib_b.search_dev, ids2, None, None, outs2)
    _same(a3, b3)
    # no device memory is allocated by the no-allocation form (after one call has loaded its kernels)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(2):
        _search(lib_a, lib_a.search_dev_noalloc, ids1, None, big, outs1)
        _search(lib_a, lib_a.search_dev_noalloc, ids2, refs2, small, outs2)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    lib_a.kill()
    lib_b.kill()


def test_noalloc_search_limits_and_errors():
    say("no-alloc search: scratch limit and bad indices")
    pf, cps = make_case(seed=612, n=308, kfrom=2, kto=60, ctf=True, nrefs=2, nptcls=4, ldim=200)
    lib = to_lib(pf, cps)
    outs = _outputs(4)
    too_many = shift_grid(4.5, 0.5)     # 361 shifts: 52 MB of shift table, over the 32 MiB scratch
    with pytest.raises(pftcc.PftccError) as e:
        _search(lib, lib.search_dev_noalloc, np.arange(1, 5, dtype=np.int32), None, too_many, outs)
    assert "scratch" in str(e.value)
    with pytest.raises(pftcc.PftccError):
        _search(lib, lib.search_dev_noalloc, np.array([1, 2, 9, 3], np.int32), None, None, outs)
    # the handle still works after both refusals
    ok = _search(lib, lib.search_dev_noalloc, np.arange(1, 5, dtype=np.int32), None, None, outs)
    ref = lib.search(np.arange(1, 5, dtype=np.int32), None, None)
    assert np.array_equal(ok[1], ref[1]) and np.array_equal(ok[3], ref[3])
    lib.kill()
