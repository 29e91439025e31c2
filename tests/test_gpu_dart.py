"""The reference's dart thrower on the GPU (lg_dart_throw, LightGraph.h:206-274)
against the oracle's sequential restatement (oracle/lgoracle.cpp
lgo_dart_throw): the same accepted points in the same order, bit for bit, and
the same final minimum distance.  The oracle's mt19937 / libstdc++ float draws
are pinned by the known-answer tests in test_oracle_core.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(lga, ctx):
    """C1 volume: 64^3 sphere r=24 x fBm4, seed 1."""
    spec = lga.scene_spec(lga.LG_SCENE_SPHERE_FBM, 64, seed=1, octaves=4, radius=24.0)
    grid, _ = lga.Grid.generate(ctx, spec)
    dense = grid.read_dense()
    yield grid, dense
    grid.destroy()


def _start_min_dist(lo, hi):
    """main.cc:514-518 in float arithmetic: MIN_DIST_MULT * mean bbox side."""
    size = [np.float32(float(hi[a]) - float(lo[a])) for a in range(3)]
    return float(np.float32(0.13) * (np.float32(np.float32(size[0] + size[1]) + size[2]) / np.float32(3.0)))


def test_active_bbox_matches_numpy(lga, ctx, c1):
    grid, dense = c1
    lo, hi = lga.active_bbox(ctx, grid)
    z, y, x = np.nonzero(dense)
    assert lo == (int(x.min()), int(y.min()), int(z.min()))
    assert hi == (int(x.max()), int(y.max()), int(z.max()))
    empty = lga.Grid.from_dense(ctx, np.zeros((16, 16, 16), np.float32))
    assert lga.active_bbox(ctx, empty) == ((0, 0, 0), (-1, -1, -1))
    empty.destroy()


@pytest.mark.parametrize("iters,scatter,target", [([0, 1, 2, 5], 100000, 750), ([3], 1000000, 750),
                                                  ([0, 7], 20000, 10), ([4], 100, 750)])
def test_dart_throw_equals_oracle(lga, ctx, oracle, c1, iters, scatter, target):
    grid, dense = c1
    lo, hi = lga.active_bbox(ctx, grid)
    md0 = _start_min_dist(lo, hi)
    pts, md = lga.dart_throw(ctx, grid, iters, scatter_count=scatter, target=target, min_dist=md0)
    for k, it in enumerate(iters):
        want, want_md = oracle.dart_throw(dense, it, scatter_count=scatter, target=target, min_dist=md0)
        assert len(pts[k]) == len(want), f"iteration {it}: {len(pts[k])} vs {len(want)} points"
        np.testing.assert_array_equal(pts[k], want, err_msg=f"iteration {it}")
        # scatter 100 gives batch 1: the first rescale is 0/0 and the distance NaN on both sides
        a, b = np.float32(md[k]), np.float32(want_md)
        assert a == b or (np.isnan(a) and np.isnan(b)), f"iteration {it}: {a} vs {b}"
        assert len(want) <= target + 1
        if target == 10:
            assert len(want) == target + 1  # early stop (LightGraph.h:232-238)


def test_dart_throw_empty_grid_and_bad_arguments(lga, ctx, c1):
    empty = lga.Grid.from_dense(ctx, np.zeros((16, 16, 16), np.float32))
    pts, md = lga.dart_throw(ctx, empty, [0], scatter_count=1000, target=750, min_dist=1.5,
                             bbox=([0.0, 0.0, 0.0], [15.0, 15.0, 15.0]))
    assert len(pts[0]) == 0 and md[0] == np.float32(1.5)
    grid, _ = c1
    with pytest.raises(lga.LGError):
        lga.dart_throw(ctx, grid, [0], scatter_count=99, target=750)
    with pytest.raises(lga.LGError):
        lga.dart_throw(ctx, grid, [0], scatter_count=1000, target=4000)
    empty.destroy()
