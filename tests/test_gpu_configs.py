"""The benchmark configurations at their own sizes, through the C ABI:

* 4096^2 synthetic H&E tile (seed 2026) through the whole chain, stage by stage against the oracle,
  with every phase-1 LDS tier of the watershed in use and no sequential fallback;
* the reference's own chain parameters (nu-segmentTest.cpp:207: T1 4.0, T2 3.5, G1 44, 5..9000,
  G2 31, 20, 5..1150, connectivities 4/4/4) at 1024^2 and 4096^2 against the oracle;
* 256 mask/marker pairs, sides log-uniform in [512, 4096] x aspect U[1, 2], conn 4/8, batched ==
  single calls for uint8 reconstruction and float32 h-minima, 8 sampled pairs against the oracle;
* the server over 32 tiles of that size distribution, each result equal alone and in the mix;
* four 4096^2 whole-slide tiles with 64-px overlap, served with canonical labels, boxes and areas.
"""
import numpy as np
import pytest

from tests._lib import STAGES, boxes_numpy, o_imhmin, o_recon, o_segment, ref_test_params

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def chain_vs_oracle(bgr, params=None):
    from nscale_amd import gpu
    stages = tuple(sorted(STAGES))
    st, lab, cc, bufs, _ = gpu.segment_nuclei(torch.from_numpy(bgr).cuda(), params=params, stages=stages)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in bufs.items()}
    lab = lab.cpu().numpy()
    del bufs
    ost, olab, obufs = o_segment(bgr, params=params, stages=stages)
    assert st == ost
    if st == 0:
        for k in sorted(obufs):
            assert np.array_equal(got[k], obufs[k]), "stage %d differs" % k
        assert np.array_equal(lab, olab)
        assert cc == len(np.unique(olab[olab >= 0]))
    return st, olab


def test_c1_chain_4096_stage_by_stage():
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    bgr = synth_he(4096, 4096, 2026)
    f0, b0 = gpu.watershed_stats()
    t0 = gpu.watershed_tier_runs()
    print("C1 chain 4096^2 seed 2026", flush=True)
    st, olab = chain_vs_oracle(bgr)
    assert st == 0
    f1, b1 = gpu.watershed_stats()
    t1 = gpu.watershed_tier_runs()
    runs = [b - a for a, b in zip(t0, t1)]
    print("phase-1 runs per tier", runs, flush=True)
    assert f1 == f0 + 1 and b1 == b0, "the decomposed flood must serve the tile without fallback"
    # every LDS tier (frames up to 512, 1024, 2048, 4096, 6400 px) floods runs of this tile; the
    # global-memory bucket beyond them is exercised by test_gpu_watershed_props
    assert all(r > 0 for r in runs[:5]), runs
    assert len(np.unique(olab[olab >= 0])) > 5000


@pytest.mark.parametrize("size,seed", [(1024, 5), (4096, 2026)])
def test_reference_test_parameters(size, seed):
    from nscale_amd.synth import synth_he
    bgr = synth_he(size, size, seed)
    print("reference test parameters", size, seed, flush=True)
    st, _ = chain_vs_oracle(bgr, params=ref_test_params())
    assert st == 0


def _c2_pairs(n=256, seed=2026):
    """Crops of two 4096 x 8192 ellipse masks (the generator of the C0 tile), sides as the config."""
    from nscale_amd.synth import synth_mask
    from nscale_amd.synth import serve_sizes
    rng = np.random.default_rng(seed)
    bases = [synth_mask(4096, 8192, seed + k) for k in range(2)]
    pairs = []
    for k, (h, w) in enumerate(serve_sizes(n, seed)):
        b = bases[k % 2] if h <= w else bases[k % 2].T
        y0 = int(rng.integers(0, b.shape[0] - h + 1))
        x0 = int(rng.integers(0, b.shape[1] - w + 1))
        mask = np.ascontiguousarray(b[y0:y0 + h, x0:x0 + w])
        pairs.append((mask, int(rng.choice([4, 8]))))
    return pairs


def test_c2_batched_at_size():
    from nscale_amd import gpu
    pairs = _c2_pairs()
    sizes = [m.shape for m, _ in pairs]
    assert len(pairs) == 256 and min(min(s) for s in sizes) >= 512 and max(max(s) for s in sizes) <= 8192
    conns = [c for _, c in pairs]
    dev = []
    for m, _ in pairs:
        mk = torch.from_numpy(m).cuda()
        dev.append(((mk.to(torch.int16) - 10).clamp_(min=0).to(torch.uint8), mk))
    print("C2 batched u8: %d pairs, %.0f Mpx" % (len(pairs), sum(m.size for m, _ in pairs) / 1e6), flush=True)
    outs = gpu.imreconstruct_batched(dev, conns)
    torch.cuda.synchronize()
    for (mk, m), c, o in zip(dev, conns, outs):
        assert torch.equal(o, gpu.imreconstruct(mk, m, c))
    sample = np.random.default_rng(3).choice(len(pairs), 8, replace=False)
    for i in sample:
        m = pairs[i][0]
        mk = dev[i][0].cpu().numpy()
        assert np.array_equal(outs[i].cpu().numpy(), o_recon(mk, m, conns[i])), "pair %d vs oracle" % i
    del outs
    # float32 h-minima (h = 1) on the negated distance maps of the thresholded masks
    imgs = [-gpu.dist_transform((m < 150).to(torch.uint8) * 255) for _, m in dev]
    del dev
    print("C2 batched f32 h-minima", flush=True)
    hb = gpu.imhmin_batched(imgs, 1.0, conns)
    torch.cuda.synchronize()
    for im, c, o in zip(imgs, conns, hb):
        assert torch.equal(o, gpu.imhmin(im, 1.0, c))
    for i in sample[:4]:
        assert np.array_equal(hb[i].cpu().numpy(), o_imhmin(imgs[i].cpu().numpy(), 1.0, conns[i])), "pair %d" % i


def test_c3_server_mix_equals_alone():
    """32 requests of the served size distribution (device-rendered H&E tiles): each result equals
    the plain chain's result on the same tile."""
    from nscale_amd import gpu
    from nscale_amd.synth import serve_sizes
    sizes = serve_sizes(32, 77)
    tiles = [gpu.synth_slide_tile(1000 * k, 500 * k, w, h, 2026) for k, (h, w) in enumerate(sizes)]
    alone = []
    for t in tiles:
        st, lab, cc, _, _ = gpu.segment_nuclei(t)
        alone.append((st, lab.clone(), cc))
    torch.cuda.synchronize()
    mx = max(max(s) for s in sizes)
    srv = gpu.Server(slots=8, hbm_budget=200 << 30, max_w=4096, max_h=max(8192, mx))
    print("C3 server mix of %d tiles" % len(tiles), flush=True)
    tickets = [srv.submit(t) for t in tiles]
    for tk, (st, lab, cc) in zip(tickets, alone):
        st2, n2, lab2, _, _, res = srv.wait(tk)
        assert res.error == 0 and st2 == st
        if st == 0:
            assert n2 == cc and torch.equal(lab2, lab)
    srv.close()
    assert sum(st == 0 for st, _, _ in alone) >= 30, "the tiles hold nuclei: the chain should succeed"


def test_c4_slide_tiles_served_with_boxes():
    """Four 4096^2 tiles (64-px overlap, 2 x 2) of the procedural whole slide through the server with
    canonical labels, boxes and areas; equal to the plain chain relabelled, boxes equal to numpy."""
    from nscale_amd import gpu
    T, stride = 4096, 4096 - 64
    xy = [(0, 0), (stride, 0), (0, stride), (stride, stride)]
    tiles = [gpu.synth_slide_tile(x, y, T, T, 2026) for x, y in xy]
    srv = gpu.Server(slots=4, hbm_budget=200 << 30, max_w=T, max_h=T)
    print("C4 slide tiles", flush=True)
    tickets = [srv.submit(t, canonical=True, max_labels=65536) for t in tiles]
    for (x, y), t, tk in zip(xy, tiles, tickets):
        st, n, lab, bbox, area, res = srv.wait(tk)
        assert res.error == 0 and st == 0 and n > 1000
        _, plain, cc, _, _ = gpu.segment_nuclei(t)
        assert gpu.cc_relabel(plain) == n == cc
        assert torch.equal(plain, lab)
        eb, ea = boxes_numpy(lab.cpu().numpy(), bg=0)
        assert np.array_equal(bbox.cpu().numpy(), eb)
        assert np.array_equal(area.cpu().numpy(), ea)
    # the overlaps render the same pixels in neighbouring tiles
    assert torch.equal(tiles[0][:, stride:], tiles[1][:, :64])
    srv.close()
