"""GPU BVH8 and software ray queries against the oracle.

The reference's traversal lives in the Vulkan driver (rayQueryProceed), so
its arithmetic is not available; the contract pinned here is the one stated
in traverse.hpp and restated by the oracle: closest hit on (tmin, tmax) with
ties to the lower triangle id, Moller-Trumbore, front face <=> det > 0.
  * Morton codes of the world triangles are bit-exact with the oracle's;
  * world triangles are bit-exact with the oracle's transform;
  * closest-hit results (t, triangle, u, v bits, front face) are bit-exact
    with the oracle's brute force over all triangles, any-hit likewise;
  * the BVH8 is well formed (every triangle once, child boxes inside the
    node's box) and the oracle's traversal of the same arrays visits the same
    nodes and triangles as the GPU's (so the traversal counts are pinned).
"""
import numpy as np
import pytest

import oracle as O
import scenes
import wsrt

pytestmark = pytest.mark.gpu


def random_rays(n, seed, lo=-0.95, hi=0.95, tmin=1e-4, tmax=100.0):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3], r[:, 7] = tmin, tmax
    return r


def caustic_pair(device):
    return scenes.build_caustic(wsrt.SceneBuilder()).build(device), \
        scenes.build_caustic(O.Scene()).build()


def test_world_and_morton_bit_exact(device, cornell_pair):
    g, o = cornell_pair
    gt, gi = g.world()
    ot, oi = o.world()
    np.testing.assert_array_equal(gt, ot)
    np.testing.assert_array_equal(gi, oi)
    codes, order = g.morton()
    np.testing.assert_array_equal(codes, o.morton())
    # build order: by (size class, code); a triangle is large when its box
    # spans more than 1/8 of the centroid bounds on some axis (exact in f32)
    v = gt.reshape(-1, 3, 3)
    lo, hi = v.min(axis=1), v.max(axis=1)
    cent = (lo + hi) * np.float32(0.5)
    ext = cent.max(axis=0) - cent.min(axis=0)
    large = ((ext > 0) & ((hi - lo) > np.float32(0.125) * ext)).any(axis=1)
    assert 0 < large.sum() < len(large)  # the Cornell box: walls and boxes large, the sphere small
    keys = codes.astype(np.int64) | (large.astype(np.int64) << 30)
    assert np.all(np.diff(keys[order]) >= 0)
    assert sorted(order.tolist()) == list(range(len(codes)))


@pytest.mark.parametrize("scene", ["cornell", "caustic"])
def test_closest_and_any_hit_match_brute_force(device, cornell_pair, scene):
    g, o = cornell_pair if scene == "cornell" else caustic_pair(device)
    print(f"closest/any hit {scene}", flush=True)
    lo, hi = (-0.95, 0.95) if scene == "cornell" else (-1.5, 2.0)
    rays = random_rays(40000, 1, lo, hi)
    gh, gf = g.closest_hit(rays)
    oh, of = o.closest_hit(rays, brute=True)
    np.testing.assert_array_equal(gf, of)
    np.testing.assert_array_equal(gh, oh)
    assert (gf & 1).mean() > 0.2
    rays[:, 3], rays[:, 7] = 1e-3, np.random.default_rng(2).uniform(0.01, 2.0, len(rays))
    np.testing.assert_array_equal(g.any_hit(rays), o.any_hit(rays, brute=True))


@pytest.mark.parametrize("scene", ["cornell", "caustic"])
def test_closest_hit_any_interval_matches_brute_force(device, cornell_pair, scene):
    """Closest hits on intervals the render kernels never use: tmin of 0 or
    0.25, tmax equal to tmin, tmax of 0, negative, infinite or NaN. The
    wave-shared triangle tests start each ray's search key from tmax, so
    these must agree with brute force too. (tmin stays >= 0: the padded box
    test assumes distances ahead of the origin, as the reference's ray
    queries do.)"""
    g, o = cornell_pair if scene == "cornell" else caustic_pair(device)
    lo, hi = (-0.95, 0.95) if scene == "cornell" else (-1.5, 2.0)
    rays = random_rays(40000, 4, lo, hi)
    rng = np.random.default_rng(5)
    n = len(rays)
    rays[:, 3] = rng.choice(np.array([0.0, 1e-4, 1e-3, 0.25], np.float32), n)
    tmax = rng.uniform(0.05, 3.0, n).astype(np.float32)
    pick = rng.integers(0, 8, n)
    tmax[pick == 0] = 0.0
    tmax[pick == 1] = -0.5
    tmax[pick == 2] = np.inf
    tmax[pick == 3] = np.nan
    tmax[pick == 4] = rays[pick == 4, 3]  # empty interval (tmin, tmin)
    rays[:, 7] = tmax
    gh, gf = g.closest_hit(rays)
    oh, of = o.closest_hit(rays, brute=True)
    np.testing.assert_array_equal(gf, of)
    np.testing.assert_array_equal(gh, oh)
    hit = (gf & 1) != 0
    assert hit.mean() > 0.05 and hit[pick == 2].any()
    assert not hit[pick <= 1].any() and not hit[pick == 3].any() and not hit[pick == 4].any()


def test_camera_rays_match_brute_force(device, cornell_pair):
    # rays from the camera position: the reference's first-bounce workload
    g, o = cornell_pair
    rng = np.random.default_rng(3)
    n = 20000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = (0, 0, 3)
    d = np.column_stack([rng.uniform(-0.4, 0.4, n), rng.uniform(-0.4, 0.4, n), -np.ones(n)])
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3], rays[:, 7] = 1e-4, 100
    gh, gf = g.closest_hit(rays)
    oh, of = o.closest_hit(rays, brute=True)
    np.testing.assert_array_equal(gh, oh)
    np.testing.assert_array_equal(gf, of)


NODE = np.dtype([("origin", np.float32, 3), ("exp", np.int8, 3), ("imask", np.uint8),
                 ("child_base", np.uint32), ("tri_base", np.uint32), ("meta", np.uint8, 8),
                 ("qlo", np.uint8, (3, 8)), ("qhi", np.uint8, (3, 8))])


@pytest.mark.parametrize("scene", ["cornell", "caustic"])
def test_bvh8_structure_and_traversal_counts(device, cornell_pair, scene):
    g, o = cornell_pair if scene == "cornell" else caustic_pair(device)
    info = g.info
    nodes_b, tris_b = g.bvh()
    nodes = nodes_b.view(NODE)
    tris = tris_b.view(np.float32).reshape(-1, 12)
    assert NODE.itemsize == 80 and len(nodes) == info.n_nodes
    ids = tris[:, 3].view(np.uint32)
    assert sorted(ids.tolist()) == list(range(info.n_triangles))
    # leaf records hold v0 and edges of the world triangle
    wt, _ = g.world()
    np.testing.assert_array_equal(tris[:, 0:3], wt[ids, 0:3])
    np.testing.assert_array_equal(tris[:, 4:7], wt[ids, 3:6] - wt[ids, 0:3])
    # children: every internal child index in range, each non-root node referenced once
    seen = np.zeros(len(nodes), int)
    for nd in nodes:
        for c in range(8):
            m = nd["meta"][c]
            if m & 0x80:
                seen[nd["child_base"] + (m & 0x7F)] += 1
            elif m:
                cnt, off = m >> 5, m & 0x1F
                assert 1 <= cnt <= 3
                # dequantized child box contains its triangles
                s = np.ldexp(np.float64(1), nd["exp"].astype(np.int64))
                lo = nd["origin"] + nd["qlo"][:, c] * s
                hi = nd["origin"] + nd["qhi"][:, c] * s
                for k in range(cnt):
                    t = ids[nd["tri_base"] + off + k]
                    v = wt[t].reshape(3, 3)
                    assert np.all(v >= lo - 1e-12) and np.all(v <= hi + 1e-12)
    assert seen[0] == 0 and np.all(seen[1:] == 1)
    # the oracle's traversal of these arrays: same hits as brute force, same counts as the GPU
    rays = random_rays(20000, 4, -0.95, 0.95)
    hits, flags, nv, nt = O.bvh8_closest(nodes_b, tris_b, rays)
    oh, of = o.closest_hit(rays, brute=True)
    np.testing.assert_array_equal(hits, oh)
    np.testing.assert_array_equal(flags, of)
    gv, gtc = g.traversal_counts(rays)
    np.testing.assert_array_equal(gv, nv)
    np.testing.assert_array_equal(gtc, nt)
    assert info.stack_size == 1 + 7 * info.max_depth
