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


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_i32_vs_oracle(conn):
    rng = np.random.default_rng(90 + conn)
    mask = rng.integers(-50000, 50000, (333, 517)).astype(np.int32)
    mask[rng.random(mask.shape) < 0.4] = 2 ** 31 - 1
    marker = (mask.astype(np.int64) - rng.integers(0, 100000, mask.shape)).clip(-2 ** 31, 2 ** 31 - 1).astype(np.int32)
    print("recon i32", conn, flush=True)
    assert np.array_equal(gpu_recon(marker, mask, conn), o_recon(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_serpentine_crosses_a_seam_per_segment(conn):
    """31 segments each crossing the seam between two tile rows: about 31 passes on 2 tiles."""
    from tests._lib import serpentine
    marker, mask, exp = serpentine()
    print("recon serpentine", conn, flush=True)
    got = gpu_recon(marker, mask, conn)
    assert np.array_equal(got, exp)
    assert np.array_equal(got, o_recon(marker, mask, conn))


@pytest.mark.parametrize("conn", [4, 8])
def test_recon_maze_1024(conn):
    """A 1024^2 one-pixel maze: thousands of passes, far above a per-tile pass budget."""
    from tests._lib import boustrophedon
    marker, mask, exp = boustrophedon(1024)
    print("recon maze 1024", conn, flush=True)
    got = gpu_recon(marker, mask, conn)
    assert np.array_equal(got, exp)
    assert np.array_equal(got, o_recon(marker, mask, conn))


@pytest.fixture
def work_queue():
    """The reconstruction's device-side work queue for the calling thread, reset afterwards."""
    from nscale_amd import gpu
    gpu.set_recon_schedule(False, 1, 0, 1)
    gpu.set_recon_schedule(True, 1, 0, 1)
    yield
    gpu.set_recon_schedule(False, 0, 0, -1)
    gpu.set_recon_schedule(True, 0, 0, -1)


@pytest.mark.parametrize("conn", [4, 8])
def test_work_queue_schedule_matches_oracle(work_queue, conn):
    """The device-side work queue (one launch drains the marked tiles) gives the oracle's bytes on
    the long-path fixtures, a 4096^2 ellipse tile and float data."""
    from nscale_amd.synth import synth_mask, synth_marker
    from tests._lib import boustrophedon, serpentine
    print("work queue", conn, flush=True)
    for marker, mask, exp in (serpentine(), boustrophedon(512)):
        assert np.array_equal(gpu_recon(marker, mask, conn), exp)
    mask = synth_mask(4096, 4096, 2026)
    marker = synth_marker(mask)
    assert np.array_equal(gpu_recon(marker, mask, conn), o_recon(marker, mask, conn))
    rng = np.random.default_rng(70 + conn)
    fm = rng.uniform(-4, 4, (300, 277)).astype(np.float32)
    fk = (fm - rng.uniform(0, 3, fm.shape)).astype(np.float32)
    assert np.array_equal(gpu_recon(fk, fm, conn), o_recon(fk, fm, conn))
