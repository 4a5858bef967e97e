"""GPU per-nucleus features (nsc_object_features) against the oracle's restatement of
ObjFeatures.cpp / Operators.cpp, on the chain's canonical labels and boxes: integer features
and the exact moment sums bit-exact, the perimeter within 1e-12 (its sum is re-associated), the
float histogram statistics within 1e-5 relative (log2 / pow of the device library)."""
import numpy as np
import pytest

from tests._lib import oracle, ptr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def o_features(lab, box, gray):
    h, w = lab.shape
    n = box.shape[1]
    o = oracle()
    lab = np.ascontiguousarray(lab)
    box = np.ascontiguousarray(box)
    g = np.ascontiguousarray(gray)
    area = np.zeros(n, np.int32)
    o.orc_feat_area(ptr(box), n, ptr(lab), w, h, ptr(area))
    ma, mi, ec, per, ext, circ = (np.zeros(n) for _ in range(6))
    o.orc_feat_ellipse(ptr(box), ptr(area), n, ptr(lab), w, h, ptr(ma), ptr(mi), ptr(ec))
    o.orc_feat_perimeter(ptr(box), n, ptr(lab), w, h, ptr(per))
    o.orc_feat_extent_circ(ptr(box), ptr(area), ptr(per), n, ptr(ext), ptr(circ))
    f8 = np.zeros(8 * n, np.float32)
    o.orc_feat_intensity(ptr(box), n, ptr(lab), ptr(g), w, h, ptr(f8))
    f6 = np.zeros(6 * n, np.float32)
    o.orc_feat_gradient(ptr(box), n, ptr(lab), ptr(g), w, h, ptr(f6))
    return dict(area=area, major_axis=ma, minor_axis=mi, eccentricity=ec, perimeter=per, extent_ratio=ext,
                circularity=circ, intensity=f8.reshape(n, 8), gradient=f6.reshape(n, 6))


@pytest.mark.parametrize("size,seed,canonical", [(1024, 3, True), (2048, 2026, False)])
def test_features_of_chain_nuclei(size, seed, canonical):
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    bgr_h = synth_he(size, size, seed)
    bgr = torch.from_numpy(bgr_h).cuda()
    print("features", size, seed, canonical, flush=True)
    st, lab, n, nuclei, box, area, _ = gpu.segment_nuclei_io(bgr, canonical=canonical, bbox_mode=gpu.BBOX_ALL,
                                                             max_boxes=1 << 16, with_area=True)
    assert st == 0 and n > 100
    for ch in (0, 2):
        f = gpu.object_features(lab, box, bgr, channel=ch)
        torch.cuda.synchronize()
        exp = o_features(lab.cpu().numpy(), box.cpu().numpy(), np.ascontiguousarray(bgr_h[:, :, ch]))
        assert np.array_equal(f["area"].cpu().numpy(), exp["area"])
        assert np.array_equal(f["area"].cpu().numpy(), area.cpu().numpy())
        for k in ("major_axis", "minor_axis", "eccentricity", "extent_ratio"):
            assert np.array_equal(f[k].cpu().numpy(), exp[k]), k
        for k in ("perimeter", "circularity"):
            assert np.allclose(f[k].cpu().numpy(), exp[k], rtol=1e-12, atol=0), k
        for k in ("intensity", "gradient"):
            assert np.allclose(f[k].cpu().numpy(), exp[k], rtol=1e-5, atol=1e-5), k
        # the integer-valued intensity statistics (max, min) are exact
        assert np.array_equal(f["intensity"].cpu().numpy()[:, 1:3], exp["intensity"][:, 1:3])
