"""Continuous-batching server: every request's result is the same alone (plain chain call) and in
the mix, admission respects the HBM budget, canonical labels and boxes match the labels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _tiles():
    from nscale_amd.synth import synth_he
    shapes = [(256, 256), (300, 520), (512, 384), (700, 700), (1024, 600), (333, 999), (640, 1280), (1024, 1024),
              (257, 263), (900, 450), (512, 1024), (768, 768)]
    return [synth_he(h, w, 100 + k) for k, (h, w) in enumerate(shapes)]


def _plain(bgr_dev):
    from nscale_amd import gpu
    st, lab, cc, _, _ = gpu.segment_nuclei(bgr_dev)
    torch.cuda.synchronize()
    return st, lab.cpu().numpy(), cc


def test_results_alone_equal_results_in_the_mix():
    from nscale_amd import gpu
    tiles = [torch.from_numpy(t).cuda() for t in _tiles()]
    alone = [_plain(t) for t in tiles]
    srv = gpu.Server(slots=6, hbm_budget=64 << 30, max_w=1024, max_h=1024)
    print("server mix of", len(tiles), flush=True)
    tickets = [srv.submit(t) for t in tiles for _ in range(2)]
    slots = set()
    for k, tk in enumerate(tickets):
        st, n, lab, _, _, res = srv.wait(tk)
        ast, alab, acc = alone[k // 2]
        assert res.error == 0 and st == ast and n == acc
        assert np.array_equal(lab.cpu().numpy(), alab), "tile %d differs in the mix" % (k // 2)
        slots.add(res.slot)
    srv.close()
    assert len(slots) > 1, "requests should spread over slots"


def test_small_tiles_match_the_oracle_through_the_server():
    from nscale_amd import gpu
    from tests._lib import o_segment
    tiles = _tiles()[:3]
    srv = gpu.Server(slots=3, hbm_budget=16 << 30, max_w=1024, max_h=1024)
    tks = [srv.submit(torch.from_numpy(t).cuda()) for t in tiles]
    tks.append(srv.submit(torch.from_numpy(tiles[1])))  # host input: copied by the worker
    for t, tk in zip(tiles + [tiles[1]], tks):
        st, n, lab, _, _, _ = srv.wait(tk)
        ost, olab, _ = o_segment(t)
        assert st == ost and np.array_equal(lab.cpu().numpy(), olab)
    srv.close()


def test_budget_bounds_the_slots_memory():
    """A budget with room for one 1024^2 workspace: requests run one at a time in whichever slot
    holds that memory, the slots never hold more than the budget, results stay exact."""
    from nscale_amd import gpu
    tiles = [torch.from_numpy(t).cuda() for t in _tiles()[:6]]
    alone = [_plain(t) for t in tiles]
    one = gpu.request_bytes(1024, 1024)
    srv = gpu.Server(slots=3, hbm_budget=one + one // 4, max_w=1024, max_h=1024)
    tks = [srv.submit(t) for t in tiles]
    spans = []
    for k, tk in enumerate(tks):
        st, n, lab, _, _, res = srv.wait(tk)
        assert res.error == 0 and st == alone[k][0] and np.array_equal(lab.cpu().numpy(), alone[k][1])
        spans.append((res.start_ms, res.end_ms, gpu.request_bytes(int(lab.shape[1]), int(lab.shape[0]))))
        assert srv.stats()[2] <= one + one // 4
    srv.close()
    # at any instant the running requests' workspaces fit in the budget
    for a in spans:
        live = sum(b[2] for b in spans if b[0] <= a[0] < b[1])
        assert live <= one + one // 4


def test_stage_requests_pipeline_s1_and_a4():
    """S1 (plFindNucleusCandidates) and A4 (plSeparateNuclei) requests of different tiles in one
    server: each equals the oracle's stage 11 / stage 21 of the same tile."""
    from nscale_amd import gpu
    from tests._lib import o_segment
    tiles = _tiles()[:4]
    exp = [o_segment(t, stages=(11, 13, 21))[2] for t in tiles]
    dev = [torch.from_numpy(t).cuda() for t in tiles]
    opened = [torch.from_numpy(e[13]).cuda() for e in exp]
    srv = gpu.Server(slots=4, hbm_budget=32 << 30, max_w=1024, max_h=1024)
    tks = [(k, 1, srv.submit(d, stage=1)) for k, d in enumerate(dev)]
    tks += [(k, 2, srv.submit(d, stage=2, seg_open=o)) for k, (d, o) in enumerate(zip(dev, opened))]
    for k, stage, tk in tks:
        st, _, mask, _, _, res = srv.wait(tk)
        assert res.error == 0 and st == 4
        assert np.array_equal(mask.cpu().numpy(), exp[k][11 if stage == 1 else 21]), (k, stage)
    srv.close()


def test_canonical_labels_boxes_and_areas():
    from nscale_amd import gpu
    t = torch.from_numpy(_tiles()[7]).cuda()
    ast, alab, acc = _plain(t)
    srv = gpu.Server(slots=2, hbm_budget=16 << 30, max_w=1024, max_h=1024)
    st, n, lab, bbox, area, _ = srv.wait(srv.submit(t, canonical=True, max_labels=20000))
    srv.close()
    lab = lab.cpu().numpy()
    assert st == ast and n == acc
    # canonical = rank of the reference label (smallest raster index) among the labels, bg 0
    roots = np.unique(alab[alab >= 0])
    exp = np.zeros_like(alab)
    exp[alab >= 0] = np.searchsorted(roots, alab[alab >= 0]) + 1
    assert np.array_equal(lab, exp)
    bbox = bbox.cpu().numpy()
    area = area.cpu().numpy()
    ys, xs = np.nonzero(lab)
    ids = lab[ys, xs]
    assert np.array_equal(bbox[0], np.arange(1, n + 1))
    assert np.array_equal(area, np.bincount(ids, minlength=n + 1)[1:])
    mnx = np.full(n + 1, 1 << 30); np.minimum.at(mnx, ids, xs)
    mxx = np.full(n + 1, -1); np.maximum.at(mxx, ids, xs)
    mny = np.full(n + 1, 1 << 30); np.minimum.at(mny, ids, ys)
    mxy = np.full(n + 1, -1); np.maximum.at(mxy, ids, ys)
    assert np.array_equal(bbox[1], mnx[1:]) and np.array_equal(bbox[2], mxx[1:])
    assert np.array_equal(bbox[3], mny[1:]) and np.array_equal(bbox[4], mxy[1:])
