"""Cell-grid neighbour search (periodic boxes of at least 4 cells a side, each
side at least the FS2 cutoff of 7.5 A; tuning neighbour_cells = 2 forces the
grid from 4 cells a side, the default takes it from 6): it must give the
lists of the sweep over the whole box, which are the reference's
(threebody_potential.f90:297-336, 859-892), so every list, energy and force
equals the sweep's bit for bit.  Cases: 1,024 waters (4 cells a side), 2,048
waters (5), 1,024 waters with the molecules moved by whole box lengths in and
out of [0, L) (cell assignment wraps positions), 256 waters at 1/4 density (4
cells a side) against the oracle's lists, and 4,096 waters (6 cells a side)
with the default tuning."""
import numpy as np
import pytest

from ccpol23_plus_2025_mbmd_amd.geometry import WaterBox, lattice_box

pytestmark = pytest.mark.gpu


def evaluate(params, box, cells):
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    tb = ThreeBody(params, "fp64")
    tb.set_tuning(neighbour_cells=cells)
    got = tb.evaluate(box)
    return got, tb.stats(), tb.fs3_list(), tb.fs2_list()


def cases():
    b1024 = lattice_box(1024, seed=41)
    rng = np.random.default_rng(42)
    shift = rng.integers(-3, 4, size=(1024, 3)) * b1024.box
    shifted = WaterBox(b1024.com + shift, b1024.quats, b1024.box)
    b256 = lattice_box(256, seed=43)
    s = 4.0 ** (1 / 3)
    dilute = WaterBox(b256.com * s, b256.quats, b256.box * s)
    return {"1024": (b1024, 4, 2), "2048": (lattice_box(2048, seed=44), 5, 2), "1024_shifted": (shifted, 4, 2),
            "256_dilute": (dilute, 4, 2), "4096_default": (lattice_box(4096, seed=46), 6, 1)}


CASES = cases()


@pytest.mark.parametrize("name", list(CASES))
def test_cells_match_sweep(params_fs2fs3, oracle, name):
    box, nc, tuning = CASES[name]
    print(f"cell lists: {name} ({box.nm} waters, L = {box.box:.2f} A, {nc} cells a side)", flush=True)
    got, st, l3, l2 = evaluate(params_fs2fs3, box, tuning)
    ref, st0, r3, r2 = evaluate(params_fs2fs3, box, 0)
    assert st["n_cells"] == nc ** 3 and st0["n_cells"] == 0
    print(f"lists: {st['ms_lists']:.3f} ms with cells, {st0['ms_lists']:.3f} ms sweeping; "
          f"{st['n_fs3']} triplets, {st['n_fs2_items']} FS2 items", flush=True)
    assert st["n_fs3"] == st0["n_fs3"] > 0 and st["n_fs2_items"] == st0["n_fs2_items"] > 0
    np.testing.assert_array_equal(l3, r3)
    np.testing.assert_array_equal(l2, r2)
    for k in ("qfs2", "qfs3", "vir"):
        assert got[k] == ref[k], k
    for k in ("fxs", "fys", "fzs", "fxfs2", "fyfs2", "fzfs2", "fxfs3", "fyfs3", "fzfs3"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    if box.nm <= 256:
        np.testing.assert_array_equal(l3, oracle.fs3_list(params_fs2fs3, box.x, box.y, box.z, box.box))
        np.testing.assert_array_equal(l2, oracle.fs2_list(params_fs2fs3, box.x, box.y, box.z, box.box))


def test_cells_small_box_sweeps(params_fs2fs3):
    """512 waters (L = 24.8 A, 3 cells a side): no grid even when forced; 1,024
    waters (4 cells a side): no grid by default."""
    _, st, _, _ = evaluate(params_fs2fs3, lattice_box(512, seed=45), 2)
    assert st["n_cells"] == 0 and st["n_fs3"] > 0
    _, st, _, _ = evaluate(params_fs2fs3, lattice_box(1024, seed=45), 1)
    assert st["n_cells"] == 0 and st["n_fs3"] > 0


def test_cells_in_a_batch_of_uneven_boxes(params_fs2fs3):
    """The 96 uneven boxes (64-1024 waters) in one batched call with the grid
    forced from 4 cells a side: the boxes of more than ~900 waters get a grid
    (each with its own range of cells), the others are swept; every box's
    energies, forces and list counts equal the all-sweep call bit for bit."""
    import torch
    from ccpol23_plus_2025_mbmd_amd.geometry import uneven_boxes
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    boxes = uneven_boxes(96, seed=2025)
    ms = np.cumsum([0] + [b.nm for b in boxes]).astype(np.int32)
    bl = np.array([b.box for b in boxes], float)
    com = torch.tensor(np.concatenate([b.com for b in boxes]).T.copy(), device=dev).contiguous()
    site = torch.tensor(np.concatenate([b.site for b in boxes]).transpose(2, 0, 1).copy(), device=dev).contiguous()
    out = {}
    for cells in (2, 0):
        tb = ThreeBody(params_fs2fs3, "fp64")
        tb.set_tuning(neighbour_cells=cells)
        res = [t.cpu().numpy() for t in tb.eval_batch_device(ms, bl, com, site)]
        out[cells] = (res, tb.stats())
    (r2, st2), (r0, st0) = out[2], out[0]
    n_grid = sum(1 for b in boxes if np.floor(b.box / (7.5 * (1 + 1e-9))) >= 4)
    print(f"batch: {n_grid} of 96 boxes with a grid, {st2['n_cells']} cells", flush=True)
    assert n_grid >= 3 and st2["n_cells"] >= 64 * n_grid and st0["n_cells"] == 0
    assert (st2["n_fs3"], st2["n_fs2_items"]) == (st0["n_fs3"], st0["n_fs2_items"])
    for a, b in zip(r2, r0):
        np.testing.assert_array_equal(a, b)
