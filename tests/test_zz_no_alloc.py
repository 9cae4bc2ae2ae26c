"""The allocation-free entry point, ccpol3b_eval_device_noalloc, called through
the C ABI (ctypes) against ccpol3b_eval_device on the same inputs: two calls in
a row on one context, the second with a smaller box, in both precisions.  Its
scratch array keeps the previous call's contents, so the second call also shows
that everything it reads is rewritten or cleared.  Lists and list-derived counts
must be identical, and so must energies (fixed-order sums over the same chunk
layout) and forces (fixed-point sums), bit for bit."""
import ctypes as C

import numpy as np
import pytest

from ccpol23_plus_2025_mbmd_amd.geometry import lattice_box

pytestmark = pytest.mark.gpu


def device_inputs(box):
    import torch
    dev = torch.device("cuda:0")
    com = torch.tensor(box.com.T.copy(), dtype=torch.float64, device=dev).contiguous()
    site = torch.tensor(np.stack([box.xs.reshape(box.nm, box.ld), box.ys.reshape(box.nm, box.ld),
                                  box.zs.reshape(box.nm, box.ld)]), dtype=torch.float64, device=dev).contiguous()
    return com, site


def call_noalloc(tb, com, site, box_len):
    import torch
    from ccpol23_plus_2025_mbmd_amd import _lib
    L = _lib.load()
    nm, ld = com.shape[1], site.shape[2]
    fsite = torch.zeros((3, nm, ld), dtype=torch.float64, device=com.device)
    fcom2 = torch.zeros((3, nm), dtype=torch.float64, device=com.device)
    fcom3 = torch.zeros((3, nm), dtype=torch.float64, device=com.device)
    energies = torch.zeros(3, dtype=torch.float64, device=com.device)
    p = lambda t, *i: C.c_void_p(t[i].data_ptr()) if i else C.c_void_p(t.data_ptr())
    rc = L.ccpol3b_eval_device_noalloc(tb._ctx, nm, ld, p(com, 0), p(com, 1), p(com, 2), p(site, 0),
                                       p(site, 1), p(site, 2), float(box_len), p(fsite, 0), p(fsite, 1),
                                       p(fsite, 2), p(fcom2), p(fcom3), p(energies),
                                       C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, fsite, fcom2, fcom3, energies


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_eval_device(params_fs2fs3, precision):
    import torch
    from ccpol23_plus_2025_mbmd_amd import _lib
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params_fs2fs3, precision)
    for n, seed in ((512, 3), (200, 4)):  # the second call is smaller
        box = lattice_box(n, seed=seed)
        com, site = device_inputs(box)
        print(f"noalloc {precision}: {n} waters, seed {seed}", flush=True)
        ref = tb.eval_device(com, site, box.box)
        torch.cuda.synchronize()
        st_ref, l3_ref = tb.stats(), tb.fs3_list()
        rc, *got = call_noalloc(tb, com, site, box.box)
        torch.cuda.synchronize()
        _lib.check(rc, tb._ctx, "ccpol3b_eval_device_noalloc")
        st = tb.stats()
        for k in ("n_fs3", "n_fs3_sr", "n_fs3_switch", "n_fs3_lr", "n_fs2_items", "n_fs2_pairs"):
            assert st[k] == st_ref[k], k
        np.testing.assert_array_equal(tb.fs3_list(), l3_ref)
        assert torch.equal(got[3], ref[3])  # qfs2, qfs3, vir increment
        for a, b in zip(got[:3], ref[:3]):
            assert torch.equal(a, b)


def test_noalloc_rejects_oversized_input(params_fs2fs3):
    """More molecules than the scratch array holds: a clear error, no device
    work, outputs untouched."""
    import torch
    from ccpol23_plus_2025_mbmd_amd import _lib
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params_fs2fs3, "fp64")
    box = lattice_box(1100, seed=5)
    com, site = device_inputs(box)
    rc, fsite, fcom2, fcom3, energies = call_noalloc(tb, com, site, box.box)
    torch.cuda.synchronize()
    assert rc == _lib.ECAPACITY
    assert "1024" in _lib.load().ccpol3b_last_error(tb._ctx).decode()
    assert float(fsite.abs().max()) == 0.0 and float(energies.abs().max()) == 0.0
