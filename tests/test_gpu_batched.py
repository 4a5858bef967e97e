"""Batched reconstruction (many pairs of uneven sizes and mixed connectivity per launch) gives
exactly the per-pair results; small pairs are checked against the oracle too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _pairs(n, seed, lo=64, hi=700):
    from nscale_amd.synth import synth_mask, synth_marker
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        a = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        b = int(a * rng.uniform(1.0, 2.0))
        h, w = (a, b) if rng.random() < 0.5 else (b, a)
        m = synth_mask(h, w, seed * 1000 + k)
        out.append((synth_marker(m), m, int(rng.choice([4, 8]))))
    return out


def test_batched_u8_equals_single_calls_and_oracle():
    from nscale_amd import gpu
    from tests._lib import o_recon
    pairs = _pairs(24, 5)
    dev = [(torch.from_numpy(mk).cuda(), torch.from_numpy(m).cuda()) for mk, m, _ in pairs]
    conns = [c for _, _, c in pairs]
    print("batched u8", len(pairs), flush=True)
    outs = gpu.imreconstruct_batched(dev, conns)
    torch.cuda.synchronize()
    for (mk, m, c), (dmk, dm), o in zip(pairs, dev, outs):
        single = gpu.imreconstruct(dmk, dm, c)
        torch.cuda.synchronize()
        assert torch.equal(o, single)
        if m.size <= 200000:
            assert np.array_equal(o.cpu().numpy(), o_recon(mk, m, c))


def test_batched_f32_and_hminima_equal_single_calls():
    from nscale_amd import gpu
    from tests._lib import o_edt
    pairs = _pairs(12, 9)
    imgs, conns = [], []
    for mk, m, c in pairs:
        bw = ((m < 150) * 255).astype(np.uint8)
        imgs.append(torch.from_numpy(-o_edt(bw)).cuda())  # distance maps of the thresholded masks, negated
        conns.append(c)
    print("batched f32", len(imgs), flush=True)
    outs = gpu.imhmin_batched(imgs, 1.0, conns)
    torch.cuda.synchronize()
    for im, c, o in zip(imgs, conns, outs):
        assert torch.equal(o, gpu.imhmin(im, 1.0, c))
    rec = gpu.imreconstruct_batched([(im - 1.0, im) for im in imgs], conns)
    torch.cuda.synchronize()
    for im, c, o in zip(imgs, conns, rec):
        assert torch.equal(o, gpu.imreconstruct(im - 1.0, im, c))


def test_batched_single_pair_and_tiny_images():
    from nscale_amd import gpu
    from tests._lib import o_recon
    rng = np.random.default_rng(3)
    pairs = []
    for h, w in [(1, 1), (1, 200), (130, 1), (63, 63), (64, 62), (125, 3)]:
        m = rng.integers(0, 256, (h, w), dtype=np.uint8)
        mk = np.minimum(m, rng.integers(0, 256, (h, w), dtype=np.uint8))
        pairs.append((mk, m))
    dev = [(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()) for a, b in pairs]
    outs = gpu.imreconstruct_batched(dev, [8] * len(dev))
    torch.cuda.synchronize()
    for (mk, m), o in zip(pairs, outs):
        assert np.array_equal(o.cpu().numpy(), o_recon(mk, m, 8))
    one = gpu.imreconstruct_batched(dev[3:4], [4])
    assert np.array_equal(one[0].cpu().numpy(), o_recon(pairs[3][0], pairs[3][1], 4))


def test_autotuner_choice_keeps_results():
    from nscale_amd import gpu
    pairs = _pairs(8, 21)
    dev = [(torch.from_numpy(mk).cuda(), torch.from_numpy(m).cuda()) for mk, m, _ in pairs]
    conns = [c for _, _, c in pairs]
    ref = gpu.imreconstruct_batched(dev, conns)
    t = gpu.autotune_recon(dev, conns)
    assert t["region"] in (1, 2) and t["passes"] in (0, 4, 8) and all(x > 0 for x in t["timings_ms"])
    assert len(t["timings_ms"]) == 5
    got = gpu.imreconstruct_batched(dev, conns)
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(ref, got))
