"""Full-size checks of the HIP cv::watershed that the flooding itself fixes, independent of the
oracle's queue order: on the chain's own stage-20 inputs of a 4096^2 tile (nuclei image, markers =
bwlabel2(localMinima(distance2)), both padded by one pixel as watershed2 does,
MorphologicOperations.cpp:3717-3760):

* the outer frame is -1 and every seed pixel keeps its label;
* every interior pixel ends up a dam (-1), one of the seed labels, or 0 where it never touches a
  labelled pixel (a pocket closed off by dams is never queued);
* every 4-connected region of one label contains a seed pixel of that label;
* every interior dam pixel has two different labels among its 4 neighbours (a dam is written
  exactly where the popped pixel's labelled neighbours disagree).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def label_regions_contain_seeds(m, seeds):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    h, w = m.shape
    idx = np.arange(h * w, dtype=np.int64).reshape(h, w)
    pos = m > 0
    eh = pos[:, :-1] & (m[:, :-1] == m[:, 1:])
    ev = pos[:-1, :] & (m[:-1, :] == m[1:, :])
    a = np.concatenate([idx[:, :-1][eh], idx[:-1, :][ev]])
    b = np.concatenate([idx[:, 1:][eh], idx[1:, :][ev]])
    g = coo_matrix((np.ones(a.size, np.int8), (a, b)), shape=(h * w, h * w)).tocsr()
    _, comp = connected_components(g, directed=False)
    comp = comp.reshape(h, w)
    seeded = np.zeros(comp.max() + 1, bool)
    seeded[comp[seeds & pos]] = True
    return bool(seeded[comp[pos]].all())


@pytest.mark.parametrize("size,seed,conn", [(4096, 2026, 8), (2048, 9, 4)])
def test_watershed_properties_full_size(size, seed, conn):
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    p = gpu.default_params()
    p.watershedConnectivity = conn
    bgr = torch.from_numpy(synth_he(size, size, seed)).cuda()
    print("watershed properties", size, seed, conn, flush=True)
    st, _, _, bufs, _ = gpu.segment_nuclei(bgr, params=p, stages=(18, 19))
    assert st == 0
    markers = gpu.bwlabel2(gpu.local_minima(bufs[18], conn), conn)
    pm = torch.nn.functional.pad(markers, (1, 1, 1, 1), value=0)
    pb = torch.zeros(size + 2, size + 2, 3, dtype=torch.uint8, device="cuda")
    pb[1:-1, 1:-1] = bufs[19]
    f0, b0 = gpu.watershed_stats()
    out = gpu.cv_watershed(pb, pm)
    torch.cuda.synchronize()
    assert gpu.watershed_stats()[1] == b0, "decomposed flood expected"
    m = out.cpu().numpy()
    mk = pm.cpu().numpy()
    seeds = mk > 0
    frame = np.ones_like(seeds)
    frame[1:-1, 1:-1] = False
    assert (m[frame] == -1).all()
    inner = ~frame
    assert np.array_equal(m[seeds & inner], mk[seeds & inner])
    allowed = np.unique(mk[seeds])
    vals = np.unique(m[inner])
    assert set(vals.tolist()) <= set(allowed.tolist()) | {-1, 0}
    c0 = m[1:-1, 1:-1] == 0
    touch = np.zeros(c0.shape, bool)
    for nb in (m[:-2, 1:-1], m[2:, 1:-1], m[1:-1, :-2], m[1:-1, 2:]):
        touch |= nb > 0
    assert not (c0 & touch).any(), "an unlabelled pixel next to a labelled one would have been flooded"
    assert label_regions_contain_seeds(m, seeds)
    # dams: two different positive labels among the 4 neighbours
    c = m[1:-1, 1:-1]
    nb = [m[:-2, 1:-1], m[2:, 1:-1], m[1:-1, :-2], m[1:-1, 2:]]
    two = np.zeros(c.shape, bool)
    for i in range(4):
        for j in range(i + 1, 4):
            two |= (nb[i] > 0) & (nb[j] > 0) & (nb[i] != nb[j])
    dams = c == -1
    assert dams.sum() > 1000
    assert two[dams].all()


def test_watershed_runs_of_every_tier_including_the_global_bucket():
    """Unknown runs of every LDS tier and one beyond them (flooded from global memory): large noisy
    colour regions around a few seeds, against the oracle's cv::watershed."""
    from nscale_amd import gpu
    from tests._lib import o_cv_watershed
    rng = np.random.default_rng(5)
    h, w = 700, 900
    bgr = np.zeros((h, w, 3), np.uint8)
    markers = np.zeros((h, w), np.int32)
    lab = 1
    # squares whose frames (box + 2-px margin) fall in the 512, 1024, 2048, 4096, 6400 tiers and beyond
    spots = [(20, 20), (20, 300), (20, 600), (300, 20), (300, 300), (470, 650)]
    for (y0, x0), side in zip(spots, [14, 26, 40, 58, 74, 180]):
        assert y0 + side + 2 < h and x0 + side + 2 < w
        bgr[y0:y0 + side, x0:x0 + side] = rng.integers(40, 220, (side, side, 3))
        for _ in range(2):
            yy, xx = y0 + int(rng.integers(0, side)), x0 + int(rng.integers(0, side))
            markers[yy, xx] = lab
            lab += 1
    t0 = gpu.watershed_tier_runs()
    got = gpu.cv_watershed(torch.from_numpy(bgr).cuda(), torch.from_numpy(markers).cuda())
    torch.cuda.synchronize()
    runs = [b - a for a, b in zip(t0, gpu.watershed_tier_runs())]
    print("phase-1 runs per tier", runs, flush=True)
    assert all(r > 0 for r in runs), runs
    assert np.array_equal(got.cpu().numpy(), o_cv_watershed(bgr, markers))
