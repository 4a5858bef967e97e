"""nsc_imreconstruct_u8_noalloc (fixed device scratch, no allocation) against nsc_imreconstruct_u8,
through the C ABI: two calls in a row, the second on a smaller image, bit-exact; sizes beyond the
fixed scratch are rejected before any launch; device memory in use does not change."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _call(name, marker, mask, conn, out=None):
    from nscale_amd._lib import lib
    h, w = mask.shape
    out = torch.empty_like(mask) if out is None else out
    rc = getattr(lib(), name)(ctypes.c_void_p(marker.data_ptr()), ctypes.c_void_p(mask.data_ptr()),
                              ctypes.c_void_p(out.data_ptr()), w, h, conn,
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, out


def test_noalloc_matches_allocating_entry_point_twice():
    from nscale_amd.synth import synth_mask, synth_marker
    cases = [((2048, 1536), 8, 11), ((700, 900), 4, 12)]  # the second call is smaller
    inputs = []
    for (h, w), conn, seed in cases:
        m = synth_mask(h, w, seed)
        inputs.append((torch.from_numpy(synth_marker(m)).cuda(), torch.from_numpy(m).cuda(), conn))
    rc, _ = _call("nsc_imreconstruct_u8_noalloc", *inputs[1])  # first use on this device
    assert rc == 0
    got = [torch.empty_like(mask) for _, mask, _ in inputs]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for (marker, mask, conn), out in zip(inputs, got):
        print("noalloc recon", tuple(mask.shape), conn, flush=True)
        rc, _ = _call("nsc_imreconstruct_u8_noalloc", marker, mask, conn, out)
        assert rc == 0
        torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    for (marker, mask, conn), out in zip(inputs, got):
        rc, ref = _call("nsc_imreconstruct_u8", marker, mask, conn)
        assert rc == 0
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
    assert free1 == free0, "the no-allocation calls changed the device memory in use"


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from nscale_amd._lib import lib
    t = torch.zeros(16, dtype=torch.uint8, device="cuda")
    p = ctypes.c_void_p(t.data_ptr())
    w = h = 62 * 300  # 90000 tiles: beyond the fixed scratch, rejected on the host before any launch
    rc = lib().nsc_imreconstruct_u8_noalloc(p, p, p, w, h, 8, None)
    assert rc == -13
    assert b"fixed scratch" in lib().nsc_last_error()
