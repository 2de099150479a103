"""The oracle's rectangular bootstrap blocks (the reference's default block
scheme, survey_marked_point_block_bootstrap, src/tpcf.f90:1017-1069) on the CPU.

The labels go through rp_find_bin, which test_oracle.py pins to the
reference's own find_bin_index on every golden case; here the index formula
and the block sums are checked against labels and sums worked out by hand
for small fields, chosen so that the plausible slips fail: a row stride of
n_x_blocks instead of the reference's n_y_blocks, a 1-based label, the
object at the box's maximum (the reference's last bin), an empty block kept,
and G left non-zero where K is 0.
"""
import numpy as np
import pytest


def test_labels_of_a_4x3_field(orc):
    # x in [0, 8] cut into 4 bins of 2, y in [0, 6] cut into 3 bins of 2
    x = np.array([0.0, 1.0, 2.0, 3.9, 5.0, 7.0, 8.0, 8.0, 0.0])
    y = np.array([0.0, 1.0, 2.5, 5.0, 3.0, 6.0, 0.0, 6.0, 6.0])
    # 1-based (ix, iy): (1,1) (1,1) (2,2) (2,3) (3,2) (4,3) (4,1) (4,3) (1,3)
    # label = (iy - 1) * n_y + ix - 1 with n_y = 3
    want = [0, 0, 4, 7, 5, 9, 3, 9, 6]
    np.testing.assert_array_equal(orc.rect_block_labels(x, y, 4, 3), want)
    np.testing.assert_array_equal(orc.rect_block_labels(x, y, 4, 3, offset=12), np.array(want) + 12)


def test_row_stride_is_n_y_blocks(orc):
    # n_x = 3 > n_y = 2: the reference's stride 2 makes (ix=3, iy=1) and
    # (ix=1, iy=2) share label 2 (a stride of n_x would give 2 and 3)
    x = np.array([0.0, 2.9, 3.0, 0.5])
    y = np.array([0.0, 0.0, 0.0, 2.0])
    lab = orc.rect_block_labels(x, y, 3, 2)
    np.testing.assert_array_equal(lab, [0, 2, 2, 2])


def test_single_precision_box(orc):
    # the box and the bins are single precision: a coordinate that rounds to
    # the float maximum is the last bin
    # (0.30000001 and 0.3 are the same float, the box's maximum), and y = 0.5
    # is (0.5 - 0) / 0.5 = 1 bin sizes up: bin int(1) + 1 = 2
    assert np.float32(0.30000001) == np.float32(0.3)
    x = np.array([0.1, 0.30000001, 0.3])
    y = np.array([0.0, 1.0, 0.5])
    np.testing.assert_array_equal(orc.rect_block_labels(x, y, 2, 2), [0, 3, 3])


def test_block_sums_rules(orc):
    # 3 objects, 2 bins, 4 blocks: block 1 is empty (dropped), block 3 has
    # K = 0 in bin 1 (its G set to 0 there) and a NaN mark there
    labels = np.array([0, 2, 3, 0])
    w = np.array([1.0, 2.0, 3.0, 0.5])
    xi_e = np.array([[1.0, 2.0], [3.0, 4.0], [5.0, np.nan], [0.5, 0.25]])
    xi_b = -xi_e
    k = np.array([[2.0, 2.0], [1.0, 1.0], [4.0, 0.0], [1.0, 1.0]])
    b = orc.survey_block_sums(labels, w, xi_e, xi_b, k, 4)
    np.testing.assert_array_equal(b["weight"], [1.5, 2.0, 3.0])
    np.testing.assert_array_equal(b["g_t"], [[1.5, 2.25], [3.0, 4.0], [5.0, 0.0]])
    np.testing.assert_array_equal(b["g_x"], [[-1.5, -2.25], [-3.0, -4.0], [-5.0, 0.0]])
    np.testing.assert_array_equal(b["k"], [[3.0, 3.0], [1.0, 1.0], [4.0, 0.0]])


@pytest.mark.parametrize("seed", [0, 1])
def test_labels_cover_the_grid(orc, seed):
    # with n_x == n_y every (ix, iy) has its own label; a uniform field of
    # many objects fills all of them, and each object sits in the cell its
    # coordinates say (away from the bin edges, where rounding decides)
    r = np.random.default_rng(seed)
    nx = ny = 5
    x = r.uniform(-3, 7, 20000).astype(np.float32)
    y = r.uniform(10, 30, 20000).astype(np.float32)
    lab = orc.rect_block_labels(x, y, nx, ny, offset=7)
    assert set(lab.tolist()) == set(range(7, 7 + nx * ny))
    fx = (x - x.min()) / (x.max() - x.min()) * nx
    fy = (y - y.min()) / (y.max() - y.min()) * ny
    ok = (np.abs(fx - np.round(fx)) > 1e-4) & (np.abs(fy - np.round(fy)) > 1e-4)
    ix = np.minimum(np.floor(fx), nx - 1)
    iy = np.minimum(np.floor(fy), ny - 1)
    np.testing.assert_array_equal(lab[ok], (iy * ny + ix + 7)[ok])
