"""GPU parity of HistologicalEntities::segmentNuclei, stage by stage, against the oracle chain."""
import numpy as np
import pytest

from tests._lib import STAGES, o_segment

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ALL_STAGES = tuple(sorted(STAGES))


def run_both(bgr, stages=ALL_STAGES):
    from nscale_amd import gpu
    st, lab, cc, bufs, _ = gpu.segment_nuclei(torch.from_numpy(bgr).cuda(), stages=stages)
    torch.cuda.synchronize()
    ost, olab, obufs = o_segment(bgr, stages=stages)
    return (st, lab.cpu().numpy(), cc, {k: v.cpu().numpy() for k, v in bufs.items()}), (ost, olab, obufs)


@pytest.mark.parametrize("size,seed", [((256, 256), 1), ((384, 520), 2), ((1024, 1024), 3)])
def test_chain_stages_synthetic(size, seed):
    from nscale_amd.synth import synth_he
    bgr = synth_he(size[0], size[1], seed)
    print("chain", size, seed, flush=True)
    (st, lab, cc, bufs), (ost, olab, obufs) = run_both(bgr)
    assert st == ost == 0
    for k in sorted(obufs):
        assert np.array_equal(bufs[k], obufs[k]), "stage %d differs" % k
    assert np.array_equal(lab, olab)
    assert cc == len(np.unique(olab[olab >= 0]))


def test_chain_reference_image():
    import os
    from PIL import Image
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "watershed.png")
    bgr = np.ascontiguousarray(np.asarray(Image.open(gold).convert("RGB"))[:, :, ::-1])
    print("chain reference image", flush=True)
    (st, lab, cc, bufs), (ost, olab, obufs) = run_both(bgr)
    assert st == ost
    if st == 0:
        for k in sorted(obufs):
            assert np.array_equal(bufs[k], obufs[k]), "stage %d differs" % k
        assert np.array_equal(lab, olab)


def test_chain_status_codes():
    bg = np.full((64, 64, 3), 240, np.uint8)
    from nscale_amd import gpu
    assert gpu.segment_nuclei(torch.from_numpy(bg).cuda())[0] == 1
    bg2 = bg.copy()
    bg2[:3] = 100
    assert gpu.segment_nuclei(torch.from_numpy(bg2).cuda())[0] == 2
    rng = np.random.default_rng(0)
    noise = rng.integers(150, 200, (64, 64, 3)).astype(np.uint8)
    assert gpu.segment_nuclei(torch.from_numpy(noise).cuda())[0] == 3


def test_chain_rbc_branch():
    # a red-blood-cell-like blob with R/G > 5 next to nuclei exercises getRBC's selection
    from nscale_amd.synth import synth_he
    bgr = synth_he(256, 256, 4).copy()
    bgr[100:110, 100:110] = (60, 30, 220)   # B, G, R: R/G = 7.3 > T1, R/B > 1
    bgr[110:114, 100:110] = (60, 50, 215)   # R/G = 4.3: in bw2 only, joined through bwselect
    (st, lab, cc, bufs), (ost, olab, obufs) = run_both(bgr)
    assert st == ost
    assert obufs[2].any(), "the RBC branch should fire"
    assert np.array_equal(bufs[2], obufs[2])
    if st == 0:
        assert np.array_equal(lab, olab)


@pytest.mark.parametrize("size,seed", [(1024, 11), (1024, 2026), (2048, 7)])
def test_watershed_decomposition_used_and_equal_to_sequential(monkeypatch, size, seed):
    """The chain's watershed runs on the decomposed parallel flood (no fallback), and the
    sequential per-component flood gives the same bytes on the same input."""
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    print("watershed decomposition", size, seed, flush=True)
    bgr = torch.from_numpy(synth_he(size, size, seed)).cuda()
    f0, b0 = gpu.watershed_stats()
    st, lab, cc, bufs, _ = gpu.segment_nuclei(bgr, stages=(20,))
    torch.cuda.synchronize()
    f1, b1 = gpu.watershed_stats()
    assert f1 == f0 + 1 and b1 == b0, "decomposed flood not used"
    monkeypatch.setenv("NSC_WATERSHED", "seq")
    st2, lab2, cc2, bufs2, _ = gpu.segment_nuclei(bgr, stages=(20,))
    torch.cuda.synchronize()
    assert st == st2 == 0
    assert torch.equal(bufs[20], bufs2[20]) and torch.equal(lab, lab2)


def test_watershed_fallback_and_latency_guard(monkeypatch):
    """A watershed input that the decomposed flood does not serve (forced from the host with
    NSC_WATERSHED_REJECT=1) goes to the exact sequential flood and gives the decomposed result's
    bytes; with NSC_WATERSHED=fast the chain returns NSC_ERR_LIMIT instead of waiting for it."""
    from nscale_amd import gpu
    from nscale_amd._lib import NscError
    from nscale_amd.synth import synth_he
    print("watershed fallback and latency guard 512 seed 5", flush=True)
    bgr = torch.from_numpy(synth_he(512, 512, 5)).cuda()
    st, lab, _, bufs, _ = gpu.segment_nuclei(bgr, stages=(20,))
    torch.cuda.synchronize()
    assert st == 0
    monkeypatch.setenv("NSC_WATERSHED_REJECT", "1")
    f0, b0 = gpu.watershed_stats()
    st2, lab2, _, bufs2, _ = gpu.segment_nuclei(bgr, stages=(20,))
    torch.cuda.synchronize()
    assert gpu.watershed_stats() == (f0, b0 + 1), "the sequential flood must serve the rejected input"
    assert st2 == 0 and torch.equal(bufs[20], bufs2[20]) and torch.equal(lab, lab2)
    monkeypatch.setenv("NSC_WATERSHED", "fast")
    with pytest.raises(NscError, match="decomposed flood does not apply"):
        gpu.segment_nuclei(bgr, stages=(20,))
    torch.cuda.synchronize()
    assert gpu.watershed_stats() == (f0, b0 + 2)
