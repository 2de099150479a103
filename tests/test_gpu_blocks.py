"""Rectangular bootstrap blocks on the GPU (tpcf_rect_blocks_*) against the oracle.

* Labels: bit-exact against the oracle's rectangular_blocks labels
  (src/tpcf.f90:1029-1049, through the single-precision find_bin_index the
  oracle pins to the reference's own outputs) for FP32 and FP64 inputs, host
  and device arrays, several grids and offsets, with objects on the box's
  edges and maxima; 8 seeded random grids, boxes and offsets.
* Errors: fewer than 2 blocks, labels past the field's range (n_x < n_y),
  non-finite coordinates.
* End to end (FP64 catalogue): labels as lens patches, a correlation with
  patch-pair sums, block_sums and the bootstrap with given multiplicities
  equal the oracle's survey_marked_point_block_bootstrap built from per-lens
  marks (block sums in object order, empty blocks dropped, G = 0 where K = 0;
  src/tpcf.f90:1043-1069, :1165-1191) within 1e-10 of the sums of |G|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SHEAR = 0


def field(seed, n, dtype):
    r = np.random.default_rng(seed)
    x = r.uniform(-40.0, 25.0, n)
    y = r.uniform(100.0, 180.0, n)
    # the box's corners and edge values, repeated objects
    x[:6] = [-40.0, 25.0, -40.0, 25.0, 0.0, 0.0]
    y[:6] = [100.0, 180.0, 180.0, 100.0, 140.0, 140.0]
    return x.astype(dtype), y.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nx, ny, offset", [(4, 4, 0), (5, 3, 0), (7, 2, 30), (2, 1, 5), (10, 10, 100)])
def test_labels_equal_oracle(tp, orc, ctx, dtype, nx, ny, offset):
    x, y = field(nx * 100 + ny, 50000, dtype)
    print(f"rect labels {dtype.__name__} {nx}x{ny}+{offset}", flush=True)
    lab = tp.rect_blocks(x, y, nx, ny, offset=offset, ctx=ctx)
    want = orc.rect_block_labels(x.astype(np.float32), y.astype(np.float32), nx, ny, offset)
    np.testing.assert_array_equal(lab, want)
    assert lab.min() >= offset and lab.max() < offset + nx * ny


@pytest.mark.parametrize("seed", range(8))
def test_labels_random_configurations(tp, orc, ctx, seed):
    """Seeded random grids against the oracle, bit-exact: 6 to 200000
    objects in a box of random position and size (0.01' to 300' a side) with
    its corners among them, 1 <= n_y <= n_x < 16 blocks, offsets to 200,
    FP32 or FP64 inputs."""
    r = np.random.default_rng(8000 + seed)
    n = int(r.choice([6, 100, 5000, 200000]))
    ny = int(r.integers(1, 12))
    nx = int(r.integers(max(ny, 2), 16))
    offset = int(r.integers(0, 200))
    x0, y0 = r.uniform(-500.0, 500.0, 2)
    wx, wy = r.uniform(0.01, 300.0, 2)
    dt = np.float32 if r.random() < 0.5 else np.float64
    x, y = r.uniform(x0, x0 + wx, n), r.uniform(y0, y0 + wy, n)
    x[:4] = [x0, x0 + wx, x0, x0 + wx]
    y[:4] = [y0, y0 + wy, y0 + wy, y0]
    x, y = x.astype(dt), y.astype(dt)
    print(f"\nseed {seed}: {n} objects, {nx} x {ny} + {offset}, box {wx:.3f} x {wy:.3f}, {np.dtype(dt).name}",
          flush=True)
    lab = tp.rect_blocks(x, y, nx, ny, offset=offset, ctx=ctx)
    np.testing.assert_array_equal(lab, orc.rect_block_labels(x.astype(np.float32), y.astype(np.float32), nx, ny,
                                                             offset))


def test_labels_device_arrays(tp, orc, ctx):
    import torch
    x, y = field(7, 100000, np.float32)
    dx = torch.from_numpy(x).cuda()
    dy = torch.from_numpy(y).cuda()
    out = torch.full((len(x),), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    tp.rect_blocks(dx.data_ptr(), dy.data_ptr(), 6, 6, offset=3, device_inputs=True, labels=out.data_ptr(),
                   n=len(x), precision="f32", ctx=ctx)
    np.testing.assert_array_equal(out.cpu().numpy(), orc.rect_block_labels(x, y, 6, 6, 3))


def test_label_errors(tp, ctx):
    x, y = field(3, 1000, np.float32)
    with pytest.raises(tp.TpcfError):
        tp.rect_blocks(x, y, 1, 1, ctx=ctx)          # fewer than 2 blocks
    with pytest.raises(tp.TpcfError):
        tp.rect_blocks(x, y, 2, 5, ctx=ctx)          # the stride 5 runs past 2 x 5 labels
    bad = x.copy()
    bad[10] = np.nan
    with pytest.raises(tp.TpcfError):
        tp.rect_blocks(bad, y, 3, 3, ctx=ctx)
    # the context stays usable
    assert tp.rect_blocks(x, y, 3, 3, ctx=ctx).max() < 9


def test_block_bootstrap_end_to_end(tp, orc, ctx):
    r = np.random.default_rng(21)
    n_a, n_b, nx, ny = 3000, 4000, 4, 3
    P = nx * ny
    lx = r.uniform(0, 60, n_a)
    ly = r.uniform(0, 60, n_a)
    # no lens in the first block (x < 15, y < 20): it is dropped
    hole = (lx < 15) & (ly < 20)
    lx[hole] += 15
    lx[:2] = [0.0, 60.0]
    ly[:2] = [60.0, 0.0]
    lens = dict(x=lx, y=ly, z=np.zeros(n_a), val=r.normal(0, 0.02, n_a), w=r.uniform(0.5, 1.5, n_a))
    src = dict(x=r.uniform(0, 60, n_b), y=r.uniform(0, 60, n_b), z=np.ones(n_b), e1=r.normal(0, 0.27, n_b),
               e2=r.normal(0, 0.27, n_b), w=r.uniform(0.5, 1.5, n_b), m=r.normal(0, 0.05, n_b))
    lab = tp.rect_blocks(lx, ly, nx, ny, ctx=ctx)
    olab = orc.rect_block_labels(lx.astype(np.float32), ly.astype(np.float32), nx, ny)
    np.testing.assert_array_equal(lab, olab)
    assert not np.any(lab == 0)
    bins = tp.create_bins(1.0, 20.0, 8, tp.LOG)
    res = tp.correlate(SHEAR, lens, src, bins, patch_A=lab, patch_B=np.zeros(n_b, np.int32), n_patch_A=P,
                       n_patch_B=1, real_bytes=8, precision="f64", ctx=ctx)
    blocks = tp.block_sums(res, lab, lens["w"], P)
    o = orc.correlate(SHEAR, lens, src, bins.bin_edges, marks=True)
    ob = orc.survey_block_sums(olab, lens["w"], o["xi_e"], o["xi_b"], o["k_marks"], P)
    # the stride n_y = 3 < n_x = 4 merges rows: labels 0..9, label 0 empty
    assert sorted(set(lab.tolist())) == list(range(1, 10))
    assert blocks["g_t"].shape == ob["g_t"].shape == (9, 8)
    np.testing.assert_allclose(blocks["weight"], ob["weight"], rtol=1e-14)
    scale = o["gt_abs"] + o["gx_abs"]
    for key in ("g_t", "g_x"):
        assert np.all(np.abs(blocks[key] - ob[key]) <= 1e-10 * scale), key
    np.testing.assert_allclose(blocks["k"], ob["k"], rtol=1e-11)
    M = r.integers(0, 3, (50, 9)).astype(np.int32)
    g = tp.bootstrap(blocks, 50, multiplicity=M, ctx=ctx)
    og = orc.bootstrap(ob, M)
    Mk = M.astype(np.float64) @ ob["k"]
    for key, sc in (("g_t", o["gt_abs"]), ("g_x", o["gx_abs"])):
        tol = 1e-10 * (M.max() * sc) / np.where(Mk > 0, Mk, 1.0)
        assert np.all(np.abs(g[key] - og[key]) <= tol), key
    np.testing.assert_allclose(g["weights"], og["weights"], rtol=1e-13)
