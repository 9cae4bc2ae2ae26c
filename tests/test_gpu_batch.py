"""GPU tests of the batched path: HIP-graph replays, the capacity flag of a
replay, and the autotuner's launch shapes (results must not depend on them)."""
import numpy as np
import pytest
import torch

from ccpol23_plus_2025_mbmd_amd.geometry import uneven_boxes

pytestmark = pytest.mark.gpu


def pack(boxes, dev):
    com = np.concatenate([b.com for b in boxes])
    site = np.concatenate([b.site for b in boxes])
    ms = np.cumsum([0] + [b.nm for b in boxes]).astype(np.int32)
    bl = np.array([b.box for b in boxes], float)
    return (ms, bl, torch.tensor(com.T.copy(), device=dev),
            torch.tensor(site.transpose(2, 0, 1).copy(), device=dev))


@pytest.fixture(scope="module")
def batch():
    return uneven_boxes(6, seed=21, nmin=64, nmax=256)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager(params_fs2fs3, batch, precision):
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    ms, bl, com, site = pack(batch, dev)
    print(f"graph replay: {len(batch)} boxes, {ms[-1]} molecules, {precision}", flush=True)
    tb = ThreeBody(params_fs2fs3, precision)
    fs, c2, c3, en = tb.eval_batch_device(ms, bl, com, site)
    g = tb.graph(ms, bl, com, site)
    torch.cuda.synchronize()
    assert g.status() == 0
    for rep in range(3):
        g.fsite.zero_()
        g.energies.fill_(-1.0)
        g.launch()
        torch.cuda.synchronize()
        assert g.status() == 0
        # fixed-order energy sums and fixed-point force sums: the same bits
        assert torch.equal(g.energies, en), rep
        assert torch.equal(g.fsite, fs), rep
        assert torch.equal(g.fcom2, c2) and torch.equal(g.fcom3, c3), rep
        # eager work between replays
        tb.eval_batch_device(ms, bl, com, site)
        torch.cuda.synchronize()
    g.close()


def test_graph_flags_capacity_and_writes_nothing(params_fs2fs3, batch):
    """A replay whose geometry outgrew the recorded triplet capacity reports it
    and leaves every output untouched (forced from the host by compressing a box)."""
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    from ccpol23_plus_2025_mbmd_amd import _lib
    dev = torch.device("cuda:0")
    ms, bl, com, site = pack(batch, dev)
    tb = ThreeBody(params_fs2fs3, "fp64")
    g = tb.graph(ms, bl, com, site, headroom=0.0)
    torch.cuda.synchronize()
    # compress the largest box towards its centre: more triplets than recorded
    b = int(np.argmax(np.diff(ms)))
    sl = slice(int(ms[b]), int(ms[b + 1]))
    com[:, sl] *= 0.85
    g.energies.fill_(-7.0)
    g.fsite.fill_(3.0)
    g.launch()
    torch.cuda.synchronize()
    assert g.status() == _lib.ECAPACITY
    assert torch.all(g.energies == -7.0) and torch.all(g.fsite == 3.0)
    # the eager path grows its capacities and succeeds on the same geometry
    _, _, _, en = tb.eval_batch_device(ms, bl, com, site)
    assert torch.all(torch.isfinite(en))
    g.close()


def test_autotune_keeps_results(params_fs2fs3, batch):
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    ms, bl, com, site = pack(batch, dev)
    tb = ThreeBody(params_fs2fs3, "mixed")
    fs0, _, _, en0 = tb.eval_batch_device(ms, bl, com, site)
    best, t_ms = tb.autotune(ms, bl, com, site, reps=1)
    print(f"autotune: {best} {t_ms:.3f} ms", flush=True)
    assert tb.get_tuning() == best and t_ms > 0
    assert 1 <= best["fs2_parts"] <= 8
    fs1, _, _, en1 = tb.eval_batch_device(ms, bl, com, site)
    assert torch.allclose(en1, en0, rtol=1e-12, atol=1e-12)
    assert torch.allclose(fs1, fs0, rtol=1e-10, atol=1e-10 * fs0.abs().max().item())
    # every launch shape gives the same answer (to round-off: the FS2 block
    # split decides how a spectator's items are grouped before they are summed)
    for shape in ({"fs3_lds_coef": 1, "fs2_parts": 1, "fs2_pair_parts": 1},
                  {"fs3_lds_coef": 0, "fs2_parts": 8, "fs2_pair_parts": 8}):
        tb.set_tuning(**shape)
        fs2, _, _, en2 = tb.eval_batch_device(ms, bl, com, site)
        assert torch.allclose(en2, en0, rtol=1e-12, atol=1e-12), shape
        assert torch.allclose(fs2, fs0, rtol=1e-10, atol=1e-10 * fs0.abs().max().item()), shape
