"""GPU parity of the chain's boundary: pitched (row-stepped) images, the reference's compcount/bbox
outputs, the literal ConnComponents::boundingBox, and the stage-split entry points
plFindNucleusCandidates (S1) / plSeparateNuclei (A4), all through the C ABI."""
import ctypes

import numpy as np
import pytest

from tests._lib import boxes_numpy, o_cc_label, o_segment, o_segment_bbox, oracle, ptr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def pitched(arr, pad_cols, offset):
    """A device view of arr whose rows are spaced wider than its width (a GpuMat with a step)."""
    h, w = arr.shape[:2]
    shape = (h, w + pad_cols) + arr.shape[2:]
    big = torch.full(shape, 77, dtype=torch.from_numpy(arr[:1]).dtype, device="cuda")
    view = big[:, offset:offset + w]
    view.copy_(torch.from_numpy(np.ascontiguousarray(arr)).cuda())
    return big, view


def canon(lab):
    lab = lab.copy()
    oracle().orc_cc_relabel(lab.shape[1], lab.shape[0], ptr(lab), -1)
    return lab


@pytest.mark.parametrize("canonical", [False, True])
def test_pitched_chain_boxes_all(canonical):
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    h, w = 1000, 1224
    bgr = synth_he(h, w, 21)
    _, bview = pitched(bgr, 37, 13)  # row step 3*(w+37), start 39 bytes in: not 4-byte aligned
    lbig = torch.full((h, w + 11), -9, dtype=torch.int32, device="cuda")
    lview = lbig[:, 5:5 + w]
    print("pitched chain", canonical, flush=True)
    st, lab, n, nuclei, box, area, _ = gpu.segment_nuclei_io(bview, labels=lview, canonical=canonical,
                                                             bbox_mode=gpu.BBOX_ALL, max_boxes=1 << 15, with_area=True)
    torch.cuda.synchronize()
    ost, olab, _ = o_segment(bgr)
    assert st == ost == 0
    exp = canon(olab) if canonical else olab
    got = lbig.cpu().numpy()
    assert np.array_equal(got[:, 5:5 + w], exp)
    assert (got[:, :5] == -9).all() and (got[:, 5 + w:] == -9).all(), "writes outside the pitched view"
    eb, ea = boxes_numpy(exp, bg=0 if canonical else -1)
    assert n == nuclei == eb.shape[1] > 50
    assert np.array_equal(box.cpu().numpy(), eb)
    assert np.array_equal(area.cpu().numpy(), ea)


def test_reference_cpu_compcount_and_bbox():
    """bbox_mode REFERENCE_CPU reproduces ConnComponents::boundingBox(labels, 0) as the reference CPU
    chain calls it (HistologicalEntities.cpp:590-591): compcount and the box block equal the oracle's."""
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    for k, (h, w) in enumerate([(512, 640), (700, 333)]):
        bgr = synth_he(h, w, 40 + k).copy()
        if k == 1:
            bgr[:30, :30] = (140, 60, 90)  # a nucleus-coloured block on the first pixel
        st, lab, n, nuclei, box, _, _ = gpu.segment_nuclei_io(torch.from_numpy(bgr).cuda(),
                                                              bbox_mode=gpu.BBOX_REFERENCE_CPU, max_boxes=4096)
        torch.cuda.synchronize()
        ost, olab, on, obox = o_segment_bbox(bgr)
        assert st == ost == 0
        assert np.array_equal(lab.cpu().numpy(), olab)
        assert n == on
        assert np.array_equal(box.cpu().numpy(), obox)
        assert nuclei == len(np.unique(olab[olab >= 0]))


def _labelled(rng, h, w, levels):
    img = rng.integers(0, levels, (h, w)).astype(np.uint8)
    img[rng.random((h, w)) < 0.3] = 0
    return img


@pytest.mark.parametrize("bg", [0, -1])
def test_cc_bounding_box_literal_vs_oracle(bg):
    """nsc_cc_bounding_box against the oracle's literal restatement (itself pinned against the
    reference's ConnComponents.cpp): multi-valued labellings meet several labels before the first
    -1 or 0, binary ones usually none; no-background and foreground-first images too."""
    from nscale_amd._lib import lib
    rng = np.random.default_rng(7 + bg)
    cases = []
    for it in range(10):
        h, w = (int(v) for v in rng.integers(3, 90, 2))
        img = _labelled(rng, h, w, 2 if it % 3 == 0 else 5)
        if it == 4:
            img[:] = 3  # one component, no background at all
        if it == 5:
            img[0, :] = 2  # first row one component (label 0)
        cases.append(o_cc_label(img, 8 if it % 2 else 4, bgval=-1))
    cases.append(np.arange(40 * 30, dtype=np.int32).reshape(40, 30))  # every label distinct, 0 first
    lab = np.arange(50 * 20, dtype=np.int32).reshape(50, 20)
    lab[0, 0] = 17
    cases.append(lab)  # label 0 nowhere before many distinct labels
    for lab in cases:
        h, w = lab.shape
        cap = h * w
        exp = np.zeros(5 * cap, np.int32)
        n_exp = oracle().orc_cc_bounding_box_literal(w, h, ptr(lab), bg, ptr(exp), cap)
        d = torch.from_numpy(lab).cuda()
        out = torch.zeros(5 * cap, dtype=torch.int32, device="cuda")
        n = ctypes.c_int(-1)
        rc = lib().nsc_cc_bounding_box(ctypes.c_void_p(d.data_ptr()), w, h, bg, ctypes.c_void_p(out.data_ptr()), cap,
                                       ctypes.byref(n), None)
        assert rc == 0
        assert n.value == n_exp
        got = out.cpu().numpy()[:5 * n_exp].reshape(5, n_exp)
        assert np.array_equal(got, exp.reshape(5, cap)[:, :n_exp])


def test_stage_split_entry_points():
    """plFindNucleusCandidates (S1) gives the oracle's stage 11 (seg_norbc); plSeparateNuclei (A4) on
    the oracle's stage 13 (seg_open) gives its stage 21 (seg_nonoverlap); pitched in and out."""
    from nscale_amd import gpu
    from nscale_amd.synth import synth_he
    h, w = 768, 900
    bgr = synth_he(h, w, 33)
    ost, _, ob = o_segment(bgr, stages=(11, 13, 21))
    assert ost == 0
    _, bview = pitched(bgr, 20, 4)
    obig = torch.zeros((h, w + 9), dtype=torch.uint8, device="cuda")
    ws = gpu.Workspace(w, h)
    st, seg = gpu.pl_find_nucleus_candidates(bview, ws=ws, out=obig[:, 3:3 + w])
    torch.cuda.synchronize()
    assert st == 4  # CONTINUE
    assert np.array_equal(seg.cpu().numpy(), ob[11])
    _, oview = pitched(ob[13], 5, 2)
    st2, sep = gpu.pl_separate_nuclei(bview, oview, ws=ws)
    torch.cuda.synchronize()
    assert st2 == 4
    assert np.array_equal(sep.cpu().numpy(), ob[21])
    bg = torch.full((64, 64, 3), 240, dtype=torch.uint8, device="cuda")
    assert gpu.pl_find_nucleus_candidates(bg)[0] == 1  # BACKGROUND
