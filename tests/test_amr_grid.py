"""The box bookkeeping of the AMR hierarchy (hyperflow_amd.amr_grid, host
only): Berger-Rigoutsos clustering and the regrid's box rules, pinned by the
properties AMReX's regrid guarantees (the reference delegates it to AMReX,
HyperFlow_V3/src/main.cpp:33-36, which is not available to compare bits
with): every buffered tag is covered, boxes are disjoint, aligned to the
blocking factor, at most max_grid_size per side, inside the domain and its
margins, and properly nested; each cluster is grid_eff full or one block."""
import numpy as np
import pytest

from hyperflow_amd import amr_grid as g


def covered(boxes, shape):
    return g.boxes_mask(boxes, shape)


def disjoint(boxes):
    for a in range(len(boxes)):
        for b in range(a + 1, len(boxes)):
            p, q = boxes[a], boxes[b]
            if not (p[2] < q[0] or q[2] < p[0] or p[3] < q[1] or q[3] < p[1]):
                return False
    return True


def test_cluster_known_cases():
    f = np.zeros((8, 16), dtype=bool)
    assert g.cluster(f, 0.7) == []
    f[3, 5] = True
    assert g.cluster(f, 0.7) == [(5, 3, 5, 3)]
    # two blobs separated by a hole are cut at the hole
    f = np.zeros((8, 16), dtype=bool)
    f[1:4, 1:4] = True
    f[4:7, 10:14] = True
    assert g.cluster(f, 0.7) == [(1, 1, 3, 3), (10, 4, 13, 6)]
    # a full rectangle stays one box
    f = np.zeros((8, 16), dtype=bool)
    f[2:6, 3:12] = True
    assert g.cluster(f, 0.7) == [(3, 2, 11, 5)]


def test_cluster_l_shape_is_cut_at_the_inflection():
    f = np.zeros((16, 16), dtype=bool)
    f[0:16, 0:3] = True          # a tall bar
    f[0:3, 0:16] = True          # and a wide one: an L, 37 % full as one box
    boxes = g.cluster(f, 0.7)
    assert covered(boxes, f.shape)[f].all() and disjoint(boxes)
    assert len(boxes) == 2
    for b in boxes:
        assert f[b[1]:b[3] + 1, b[0]:b[2] + 1].mean() >= 0.7


@pytest.mark.parametrize("seed", range(40))
def test_cluster_random_patterns(seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((24, 40), dtype=bool)
    for _ in range(rng.integers(1, 6)):
        i, j = rng.integers(0, 40), rng.integers(0, 24)
        w, h = rng.integers(1, 12), rng.integers(1, 8)
        f[j:j + h, i:i + w] = True
    f &= rng.random(f.shape) < 0.9
    boxes = g.cluster(f, 0.7)
    assert covered(boxes, f.shape)[f].all()
    assert disjoint(boxes)
    for b in boxes:
        sub = f[b[1]:b[3] + 1, b[0]:b[2] + 1]
        assert sub.any()
        assert sub.mean() >= 0.7 or b[0] == b[2] or b[1] == b[3] or (sub.shape[0] <= 2 and sub.shape[1] <= 2)


def test_chop_known_answers():
    assert g.chop((0, 0, 63, 63), 64, 16) == [(0, 0, 63, 63)]
    assert g.chop((0, 0, 95, 15), 64, 16) == [(0, 0, 47, 15), (48, 0, 95, 15)]
    assert g.chop((32, 16, 191, 31), 64, 16) == [(32, 16, 95, 31), (96, 16, 143, 31), (144, 16, 191, 31)]


def check_hierarchy(grids, tags, nx0, ny0, blocking, mgs, nbuf, nproper, margins=(0, 0, 0, 0)):
    masks = g.level_masks(grids, nx0, ny0)           # raises when a level is not nested
    for lev in range(1, len(masks)):
        boxes = grids[lev]
        shape = masks[lev].shape
        assert disjoint(boxes)
        for b in boxes:
            w, h = b[2] - b[0] + 1, b[3] - b[1] + 1
            assert b[0] % blocking == 0 and b[1] % blocking == 0 and w % blocking == 0 and h % blocking == 0
            assert w <= mgs and h <= mgs
            assert b[0] >= 0 and b[1] >= 0 and b[2] < shape[1] and b[3] < shape[0]
            s = 2 ** lev
            assert b[0] >= 2 * margins[0] * s // 2 and b[2] < shape[1] - margins[1] * s
        # proper nesting: the fine boxes, coarsened, lie n_proper cells inside level lev-1's
        # valid cells except on the domain sides
        if lev >= 2:
            coarse_valid = masks[lev - 1] > 0
            nest = g.shrink_mask(coarse_valid, nproper)
            fine_c = g.coarsen_mask(masks[lev] > 0, 2)
            assert not (fine_c & ~nest).any()
    # level 1 covers the buffered level-0 tags (inside the margins)
    t0 = g.grow_mask(tags[0], nbuf)
    inner = np.zeros_like(t0)
    inner[margins[2]:ny0 - margins[3], margins[0]:nx0 - margins[1]] = True
    if len(masks) > 1:
        cov0 = masks[0] == 2
        assert not (t0 & inner & ~cov0).any()


@pytest.mark.parametrize("seed", range(12))
def test_make_grids_two_levels_random(seed):
    rng = np.random.default_rng(100 + seed)
    nx0, ny0, blocking, mgs = 128, 64, 16, 64
    tags0 = np.zeros((ny0, nx0), dtype=bool)
    for _ in range(rng.integers(1, 5)):
        i, j = rng.integers(0, nx0), rng.integers(0, ny0)
        tags0[j:j + rng.integers(1, 20), i:i + rng.integers(1, 30)] = True
    valid0 = np.ones_like(tags0)
    grids1 = g.make_grids([tags0], [valid0], 2, blocking, mgs, 0.7)
    check_hierarchy(grids1, [tags0], nx0, ny0, blocking, mgs, 2, 2)
    # second level from tags on the new level 1 (inside its valid region)
    m = g.level_masks(grids1, nx0, ny0)
    if len(m) < 2:
        return
    valid1 = m[1] > 0
    tags1 = valid1 & (rng.random(valid1.shape) < 0.02)
    grids2 = g.make_grids([tags0, tags1], [valid0, valid1], 2, blocking, mgs, 0.7)
    check_hierarchy(grids2, [tags0, tags1], nx0, ny0, blocking, mgs, 2, 2)
    m2 = g.level_masks(grids2, nx0, ny0)
    if len(m2) == 3:
        # tags of level 1 inside its nesting region are covered by level 2
        nest = g.shrink_mask(valid1, 2)
        bf = blocking // 2
        allowed = nest.reshape(nest.shape[0] // bf, bf, nest.shape[1] // bf, bf).all(axis=(1, 3))
        allowed_cells = np.kron(allowed, np.ones((bf, bf), dtype=bool)).astype(bool)
        t1 = g.grow_mask(tags1, 2) & allowed_cells
        assert not (t1 & ~(m2[1] == 2)).any()


def test_make_grids_margins_keep_boxes_off_periodic_sides():
    nx0, ny0 = 64, 64
    tags0 = np.zeros((ny0, nx0), dtype=bool)
    tags0[:, 0:3] = True              # along the x-lo side
    grids = g.make_grids([tags0], [np.ones_like(tags0)], 1, 8, 64, 0.7, margins=(8, 8, 0, 0))
    for b in grids[1]:
        assert b[0] >= 16


def test_level_masks_values():
    grids = [[], [(16, 16, 47, 31)], [(48, 40, 79, 55)]]
    m = g.level_masks(grids, 32, 32)
    assert [a.shape for a in m] == [(32, 32), (64, 64), (128, 128)]
    assert (m[0][8:16, 8:24] == 2).all() and m[0].sum() == 32 * 32 + 8 * 16
    assert (m[1][20:28, 24:40] == 2).all()
    assert ((m[1] == 1).sum() + (m[1] == 2).sum()) == 32 * 16
    assert (m[2] == 1).sum() == 32 * 16
    with pytest.raises(ValueError):
        g.level_masks([[], [(16, 16, 31, 31)], [(0, 0, 15, 15)]], 32, 32)
