"""Batched uneven boxes (the 96-box workload): many independent periodic water
boxes of uneven size in one batched evaluation, recorded into a HIP graph with
launch shapes from the autotuner.

* All 96 boxes (N log-uniform in 64-1024, 38,109 molecules), FS2+FS3: the graph
  replay equals the eager batched call bit for bit; every box equals its own
  single-box call bit for bit (energies: fixed-order sums over box-aligned
  chunks; forces: fixed-point sums); every box's FS3 triplet and FS2 item
  counts equal the triangle counts of its cutoff graphs where those are the
  reference's lists (twice the cutoff below half the box), and the oracle's
  list counts for the small boxes; every box's total force vanishes
  (translation invariance of E).  In FP64 and in mixed precision.
* Eight small uneven boxes (64-128 waters) in one batch, box by box against
  the oracle: the FP64 parity bar, and the mixed path's bar (site forces 1e-5)
  with its measured envelopes on energies (FS3 and FS2) and COM forces.
"""
import numpy as np
import pytest
import torch

from ccpol23_plus_2025_mbmd_amd.geometry import uneven_boxes

pytestmark = pytest.mark.gpu

B2A = 0.529177249


def pack(boxes, dev):
    com = np.concatenate([b.com for b in boxes])
    site = np.concatenate([b.site for b in boxes])
    ms = np.cumsum([0] + [b.nm for b in boxes]).astype(np.int32)
    bl = np.array([b.box for b in boxes], float)
    return (ms, bl, torch.tensor(com.T.copy(), device=dev).contiguous(),
            torch.tensor(site.transpose(2, 0, 1).copy(), device=dev).contiguous())


def triangles(com_a, box_a, cut_a):
    """Triangles of the graph 'COM distance < cut' with the list kernels' FP64
    minimum image in bohr."""
    x = com_a / B2A
    L = box_a / B2A
    d = x[:, None, :] - x[None, :, :]
    d = d - L * (np.sign(d / L) * np.floor(np.abs(d / L) + 0.5))
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    a = (r2 < (cut_a / B2A) ** 2).astype(np.float32)
    np.fill_diagonal(a, 0.0)
    return int(round(float(((a @ a) * a).sum(dtype=np.float64)) / 6))


def rel_rms(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def boxes96():
    return uneven_boxes(96, seed=2025)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_96_uneven_boxes_graph(params_fs2fs3, oracle, boxes96, precision):
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    ms, bl, com, site = pack(boxes96, dev)
    print(f"96 uneven boxes ({ms[-1]} molecules, {min(np.diff(ms))}-{max(np.diff(ms))} per box), {precision}",
          flush=True)
    tb = ThreeBody(params_fs2fs3, precision)
    tune, t_ms = tb.autotune(ms, bl, com, site, reps=1)
    print(f"autotune: {tune}, {t_ms:.2f} ms", flush=True)
    fs, c2, c3, en = tb.eval_batch_device(ms, bl, com, site)
    torch.cuda.synchronize()
    total = tb.stats()
    g = tb.graph(ms, bl, com, site)
    for _ in range(2):
        g.fsite.zero_()
        g.launch()
        torch.cuda.synchronize()
        assert g.status() == 0
        assert torch.equal(g.energies, en) and torch.equal(g.fsite, fs)
        assert torch.equal(g.fcom2, c2) and torch.equal(g.fcom3, c3)
    g.close()
    fs_h, c2_h, c3_h, en_h = (t.cpu().numpy() for t in (fs, c2, c3, en))
    n3 = n2 = 0
    checked = [0, 0, 0]
    for b, box in enumerate(boxes96):
        sl = slice(int(ms[b]), int(ms[b + 1]))
        cb, sb = com[:, sl].contiguous(), site[:, sl].contiguous()
        f1, c21, c31, e1 = tb.eval_device(cb, sb, box.box)
        torch.cuda.synchronize()
        st = tb.stats()
        n3 += st["n_fs3"]
        n2 += st["n_fs2_items"]
        assert np.array_equal(e1.cpu().numpy(), en_h[b]), b
        assert np.array_equal(f1.cpu().numpy(), fs_h[:, sl]), b
        assert np.array_equal(c21.cpu().numpy(), c2_h[:, sl]) and np.array_equal(c31.cpu().numpy(), c3_h[:, sl]), b
        # the derived third vector of a triplet (x_jk from the images of j and
        # k nearest i) is the minimum image when 2 cut < L / 2: then the
        # triplets are the triangles of the cutoff graph
        if box.box > 4 * 6.3:
            assert st["n_fs3"] == triangles(box.com, box.box, 6.3), b
            checked[0] += 1
        if box.box > 4 * params_fs2fs3.rcut3:
            assert st["n_fs2_items"] == 3 * triangles(box.com, box.box, params_fs2fs3.rcut3), b
            checked[1] += 1
        if box.nm <= 160:  # and the oracle's lists, cheap for small boxes
            assert st["n_fs3"] == len(oracle.fs3_list(params_fs2fs3, box.x, box.y, box.z, box.box)), b
            assert st["n_fs2_items"] == len(oracle.fs2_list(params_fs2fs3, box.x, box.y, box.z, box.box)), b
            checked[2] += 1
        fsb = fs_h[:, sl, :17]
        net = fsb.sum(axis=(1, 2)) + c2_h[:, sl].sum(1) + c3_h[:, sl].sum(1)
        assert np.all(np.abs(net) <= 1e-10 * np.abs(fsb).sum()), (b, net)
    assert (total["n_fs3"], total["n_fs2_items"]) == (n3, n2)
    print(f"list counts checked: FS3 triangles {checked[0]}, FS2 triangles {checked[1]}, oracle {checked[2]} boxes")
    assert min(checked) >= 5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_small_uneven_batch_against_oracle(params_fs2fs3, oracle, precision):
    from ccpol23_plus_2025_mbmd_amd.threebody import ThreeBody
    dev = torch.device("cuda:0")
    boxes = uneven_boxes(8, seed=77, nmin=64, nmax=128)
    ms, bl, com, site = pack(boxes, dev)
    print(f"8 small uneven boxes vs oracle: {list(np.diff(ms))}, {precision}", flush=True)
    tb = ThreeBody(params_fs2fs3, precision)
    fs, c2, c3, en = (t.cpu().numpy() for t in tb.eval_batch_device(ms, bl, com, site))
    for b, box in enumerate(boxes):
        sl = slice(int(ms[b]), int(ms[b + 1]))
        ref = oracle.threebody_term(params_fs2fs3, box, nthreads=16)
        f_ref = np.stack([ref[k][:, :17] for k in ("fxs", "fys", "fzs")])
        c2_ref = np.stack([ref[k] for k in ("fxfs2", "fyfs2", "fzfs2")])
        c3_ref = np.stack([ref[k] for k in ("fxfs3", "fyfs3", "fzfs3")])
        err_f = rel_rms(fs[:, sl, :17], f_ref)
        err_all = rel_rms(np.concatenate([fs[:, sl, :17].ravel(), c2[:, sl].ravel(), c3[:, sl].ravel()]),
                          np.concatenate([f_ref.ravel(), c2_ref.ravel(), c3_ref.ravel()]))
        err_c3 = rel_rms(c3[:, sl], c3_ref)
        err_e2 = abs(en[b, 0] - ref["qfs2"]) / abs(ref["qfs2"])
        err_e3 = abs(en[b, 1] - ref["qfs3"]) / abs(ref["qfs3"])
        print(f"box {b} ({box.nm} waters): site {err_f:.2e} all {err_all:.2e} FS3 COM {err_c3:.2e} "
              f"qfs2 {err_e2:.2e} qfs3 {err_e3:.2e}", flush=True)
        if precision == "fp64":
            assert err_e2 <= 1e-10 and err_e3 <= 1e-10
            assert err_f <= 1e-9 and err_all <= 1e-9
            assert err_c3 <= 1e-9
        else:
            assert err_f <= 1e-5 and err_all <= 1e-5
            assert err_e2 <= 2e-4  # FS2 with FP32 charge-site distances: measured <= 3.4e-5
            assert err_e3 <= 3e-3 and err_c3 <= 4e-3  # measured envelopes, not a bar
