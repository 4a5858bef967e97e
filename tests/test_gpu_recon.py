"""GPU parity: nsc_imreconstruct_{u8,f32} (through the C ABI) against the oracle."""
import numpy as np
import pytest

from tests._lib import o_recon

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def gpu_recon(marker, mask, conn):
    from nscale_amd import gpu
    m = torch.from_numpy(np.ascontiguousarray(marker)).cuda()
    k = torch.from_numpy(np.ascontiguousarray(mask)).cuda()
    out = gpu.imreconstruct(m, k, conn)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", [(1, 1), (1, 70), (70, 1), (5, 5), (62, 62), (63, 64), (64, 63),
                                   (124, 125), (200, 131), (257, 513)])
def test_recon_u8_random_shapes(conn, shape):
    rng = np.random.default_rng(hash(shape) % 2**32 + conn)
    h, w = shape
    mask = rng.integers(0, 256, shape).astype(np.uint8)
    marker = (mask.astype(np.int32) - rng.integers(0, 120, shape)).clip(0, 255).astype(np.uint8)
    marker[rng.random(shape) < 0.05] = 255
    print("recon u8", shape, conn, flush=True)
    assert np.array_equal(gpu_recon(marker, mask, conn), o_recon(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_u8_long_range(conn):
    # one bright seed must travel across every tile of a plateau-and-maze mask
    rng = np.random.default_rng(conn)
    mask = (rng.random((700, 900)) < 0.7).astype(np.uint8) * 200 + 10
    marker = np.zeros_like(mask)
    marker[350, 450] = 255
    print("recon u8 long range", conn, flush=True)
    assert np.array_equal(gpu_recon(marker, mask, conn), o_recon(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_u8_synthetic_mask(conn):
    from nscale_amd.synth import synth_mask, synth_marker
    mask = synth_mask(1024, 1536, 7)
    marker = synth_marker(mask)
    print("recon u8 synthetic", conn, flush=True)
    assert np.array_equal(gpu_recon(marker, mask, conn), o_recon(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_f32_negative_and_border(conn):
    rng = np.random.default_rng(40 + conn)
    mask = rng.uniform(-4, 4, (300, 277)).astype(np.float32)
    marker = (mask - rng.uniform(0, 3, mask.shape)).astype(np.float32)
    print("recon f32", conn, flush=True)
    assert np.array_equal(gpu_recon(marker, mask, conn), o_recon(marker, mask, conn))


def test_recon_u8_full_tile_4096():
    from nscale_amd.synth import synth_mask, synth_marker
    mask = synth_mask(4096, 4096, 2026)
    marker = synth_marker(mask)
    print("recon u8 4096^2", flush=True)
    got = gpu_recon(marker, mask, 8)
    assert np.array_equal(got, o_recon(marker, mask, 8))
