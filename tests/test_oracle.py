"""Pins the CPU oracle (no GPU): known answers, closed forms, invariants and
brute force, each chosen so that a plausible mistake in the restatement (a
dropped term, a wrong sign, a transposed index) fails at least one of them."""
import json
import math
import pathlib

import numpy as np
import pytest
from scipy import stats

import oracle
from compound_ray_amd import scenes

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------ Philox RNG ----

# Published known-answer vectors of Philox4x32-10 (Random123 kat_vectors).
RANDOM123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_random123_kat(ctr, key, out):
    assert list(oracle.philox(ctr, key)) == out


def test_philox_matches_rocrand():
    vectors = json.loads((GOLDEN / "philox_rocrand.json").read_text())
    assert len(vectors) >= 5
    for v in vectors:
        assert list(oracle.philox(v["ctr"], v["key"])) == v["out"], v


# ---------------------------------------------------- deterministic math ----

def ulp_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    spacing = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    return np.abs(got - want) / spacing


def test_sincos_accuracy():
    x = np.concatenate([np.linspace(-30, 30, 200001), np.random.default_rng(0).normal(0, 0.1, 100000)])
    x = x.astype(np.float32)
    s, c = oracle.math_fn("sin", x), oracle.math_fn("cos", x)
    xs = x.astype(np.float64)
    # absolute error relative to the unit circle
    assert np.max(np.abs(s - np.sin(xs))) < 2.5e-7
    assert np.max(np.abs(c - np.cos(xs))) < 2.5e-7
    small = np.abs(xs) < 0.7
    assert np.max(ulp_err(s[small], np.sin(xs[small]))) <= 2.0


def test_log_exp_pow_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(2 ** -24, 1, 200000), rng.uniform(1, 1e6, 50000),
                        np.float32(2.0) ** np.arange(-140, 120)]).astype(np.float32)
    assert np.max(ulp_err(oracle.math_fn("log", x), np.log(x.astype(np.float64)))) <= 2.0
    e = rng.uniform(-120, 120, 200000).astype(np.float32)
    assert np.max(ulp_err(oracle.math_fn("exp2", e), np.exp2(e.astype(np.float64)))) <= 2.0
    c = rng.uniform(0, 1, 200000).astype(np.float32)
    p = oracle.math_fn("pow", c, np.full_like(c, 2.2))
    want = np.power(c.astype(np.float64), float(np.float32(2.2)))
    # 2^(y log2 x): the float32 exponent carries ~1 ulp of |y log2 x|, so the
    # relative error grows with it (about 2e-6 near x = 2^-8)
    bound = (np.abs(2.2 * np.log2(c.astype(np.float64))) + 4.0) * 2.0 ** -23
    assert np.all(np.abs(p - want) / np.maximum(want, 1e-30) <= bound)
    assert oracle.math_fn("pow", np.float32([0.0, 1.0, -0.5]), np.float32([2.2] * 3)).tolist() == [0.0, 1.0, 0.0]


def test_asin_atan2_accuracy():
    x = np.linspace(-1, 1, 400001).astype(np.float32)
    a = oracle.math_fn("asin", x)
    assert np.max(np.abs(a - np.arcsin(x.astype(np.float64)))) < 3e-7
    assert oracle.math_fn("asin", np.float32([1.0]))[0] == np.float32(math.pi / 2)
    rng = np.random.default_rng(2)
    y, z = rng.normal(size=(2, 200000)).astype(np.float32)
    t = oracle.math_fn("atan2", y, z)
    assert np.max(np.abs(t - np.arctan2(y.astype(np.float64), z.astype(np.float64)))) < 5e-7


# ------------------------------------------------------- ray generation ----

def one_omm(axis, acceptance, pos=(0, 0, 0), focal=0.0):
    o = np.zeros((1, 8), np.float32)
    o[0, 0:3] = pos
    o[0, 3:6] = axis
    o[0, 6] = acceptance
    o[0, 7] = focal
    return o


IDENT = np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1])


@pytest.mark.parametrize("axis", [(1, 0, 0), (0, 0, -1), (0.26726124, 0.5345225, 0.8017837)])
def test_cone_closed_forms(axis):
    """Angle between sample and axis is |splay| ~ |N(0, sigma)|, sigma =
    acceptance / 2.35482 (FWHM = acceptance), azimuth uniform."""
    acc = np.float32(0.2)
    n = 200000
    rays = oracle.generate_rays(one_omm(axis, acc), IDENT, spp=n)
    d = rays[:, 4:7].astype(np.float64)
    a = np.asarray(axis, np.float64)
    cosang = (d @ a) / np.linalg.norm(d, axis=1) / np.linalg.norm(a)
    ang = np.arccos(np.clip(cosang, -1, 1))
    sigma = float(acc) / 2.35482004503094938
    assert stats.kstest(ang, stats.halfnorm(scale=sigma).cdf).pvalue > 1e-3
    assert abs(np.mean(ang) - sigma * math.sqrt(2 / math.pi)) < 3e-3 * sigma * 10
    # azimuth around the axis is uniform over (0, 2 pi)
    u = np.cross(a, [0.3, 0.9, 0.1])
    u /= np.linalg.norm(u)
    v = np.cross(a, u)
    az = np.arctan2(d @ v, d @ u)
    assert stats.kstest((az + math.pi) / (2 * math.pi), "uniform").pvalue > 1e-3
    # directions are unit length when the axis is
    assert np.max(np.abs(np.linalg.norm(d, axis=1) - 1)) < 1e-5


def test_zero_acceptance_gives_axis_and_origin():
    rays = oracle.generate_rays(one_omm((0, 0, 1), 0.0, pos=(0.1, 0.2, 0.3), focal=0.05), IDENT, spp=16)
    assert np.allclose(rays[:, 4:7], [0, 0, 1], atol=1e-7)
    # origin moved back along the normalised axis by the focal offset, tmin = offset
    assert np.allclose(rays[:, 0:3], [0.1, 0.2, 0.25], atol=1e-7)
    assert np.all(rays[:, 3] == np.float32(0.05)) and np.all(rays[:, 7] == np.float32(1e16))


def test_degenerate_perpendicular_rule():
    """shaders.cu:657 tests the component sum of cross((0,1,0), a): axes with
    a.x == a.z fall back to (0,0,1).  Compare with a float64 restatement."""
    for axis in [(0, 1, 0), (0.6, 0.52915025, 0.6), (0.48, 0.6, 0.64)]:
        acc = 0.3
        rays = oracle.generate_rays(one_omm(axis, acc), IDENT, spp=64)
        a = np.asarray(axis, np.float64)
        perp = np.cross([0, 1, 0], a)
        perp = np.array([0, 0, 1.0]) if perp.sum() == 0 else perp / np.linalg.norm(perp)
        # recover splay and spin from the random stream as the oracle draws them
        for s in range(64):
            r = oracle.philox([0, s, 0, 0], [42, 0])
            u1 = ((int(r[0]) >> 8) + 1) * 2.0 ** -24
            u2 = (int(r[1]) >> 8) * 2.0 ** -24
            splay = math.sqrt(-2 * math.log(u1)) * math.cos(2 * math.pi * u2) * acc / 2.35482004503094938
            spin = ((int(r[2]) >> 8) + 1) * 2.0 ** -24 * math.pi

            def rot(p, ang, k):
                return math.cos(ang) * p + math.sin(ang) * np.cross(k, p) + (1 - math.cos(ang)) * np.dot(k, p) * k
            want = rot(rot(a, splay, perp), spin, a)
            assert np.allclose(rays[s, 4:7], want, atol=2e-6), (axis, s)


def test_pose_restatement():
    """setCameraPose: rotations about world X then Y then Z, then translation."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        p6 = np.concatenate([rng.normal(size=3), rng.uniform(-3, 3, size=3)]).astype(np.float32)
        got = oracle.pose12(p6)
        want = scenes.pose_matrix12(p6)
        assert np.allclose(got, want, atol=2e-6)


# ---------------------------------------------------------- intersection ----

def flat_scene(h=0.25, half=5.0):
    pos = np.float32([[-half, h, -half], [half, h, -half], [half, h, half], [-half, h, half]])
    tri = np.concatenate([pos[[0, 1, 2]].reshape(1, 9), pos[[0, 2, 3]].reshape(1, 9)])
    col = np.tile(np.float32([0.2, 0.4, 0.6] * 3), (2, 1))
    return oracle.OracleScene(tri, col, np.zeros(2, np.uint32), 1)


def test_flat_plane_closed_form():
    o = flat_scene()
    rng = np.random.default_rng(3)
    xz = rng.uniform(-4.9, 4.9, size=(2000, 2)).astype(np.float32)
    rays = np.zeros((2000, 8), np.float32)
    rays[:, 0] = xz[:, 0]
    rays[:, 1] = 1.25
    rays[:, 2] = xz[:, 1]
    rays[:, 5] = -1.0
    rays[:, 7] = 1e16
    ids, tbb = o.trace(rays)
    # triangle 0 holds points with x >= z (below the diagonal (0,0)-(2,2) of the quad)
    want = np.where(xz[:, 0] > xz[:, 1], 0, 1)
    off_diag = np.abs(xz[:, 0] - xz[:, 1]) > 1e-6
    assert np.array_equal(ids[off_diag], want[off_diag])
    assert np.allclose(tbb[:, 0], 1.0, atol=1e-6)
    # barycentrics reproduce the hit point
    v = o.pos.reshape(-1, 3, 3)[ids]
    p = (1 - tbb[:, 1:2] - tbb[:, 2:3]) * v[:, 0] + tbb[:, 1:2] * v[:, 1] + tbb[:, 2:3] * v[:, 2]
    assert np.allclose(p[:, [0, 2]], xz, atol=2e-6)
    # rays pointing up, or starting below tmin of the hit, miss
    rays[:, 5] = 1.0
    assert np.all(o.trace(rays)[0] == -1)


def icosphere(subdiv=2):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.asarray(v, np.float32), np.asarray(f, np.uint32)


def test_watertight_closed_mesh():
    """Rays from inside a closed mesh aimed at its vertices and edge midpoints
    (where non-watertight tests leak) always hit."""
    v, f = icosphere(2)
    tri = v[f].reshape(-1, 9)
    o = oracle.OracleScene(tri, np.ones_like(tri), np.zeros(len(f), np.uint32), 1)
    targets = np.concatenate([v, (v[f[:, 0]] + v[f[:, 1]]) / 2, (v[f[:, 1]] + v[f[:, 2]]) / 2])
    origins = np.float32([[0, 0, 0], [0.1, -0.2, 0.05], [0.3, 0.3, -0.3]])
    rays = []
    for org in origins:
        d = targets - org
        r = np.zeros((len(d), 8), np.float32)
        r[:, 0:3] = org
        r[:, 4:7] = d
        r[:, 7] = 1e16
        rays.append(r)
    rays = np.concatenate(rays)
    ids, _ = o.trace(rays, bvh=True)
    assert np.all(ids >= 0)
    bids, _ = o.trace(rays, bvh=False)
    assert np.array_equal(ids, bids)


def test_bvh_equals_brute_force_c1(config1, oracle_c1):
    pose = oracle.pose12(config1["pose"])
    rays = oracle.generate_rays(config1["eyes"][0][::5], pose, spp=32)
    ids, tbb = oracle_c1.trace(rays, bvh=True)
    bids, btbb = oracle_c1.trace(rays, bvh=False)
    assert np.array_equal(ids, bids)
    assert np.array_equal(tbb.view(np.uint32), btbb.view(np.uint32))
    assert 0.2 < np.mean(ids >= 0) < 0.9


# --------------------------------------------------------------- shading ----

def test_shading_barycentric_and_sky():
    """A ray straight down onto a vertex-coloured triangle returns
    pow(barycentric colour, 2.2); misses return the simple_sky gradient."""
    tri = np.float32([[0, 0, 0, 1, 0, 0, 0, 0, 1]])
    col = np.float32([[1, 0, 0, 0, 1, 0, 0, 0, 1]])
    o = oracle.OracleScene(tri, col, np.zeros(1, np.uint32), 1)
    omm = one_omm((0, -1, 0), 0.0, pos=(0.25, 1.0, 0.5))
    r = o.render(omm, IDENT, spp=1)["colour_gpu_order"][0, 0]
    # hit point (0.25, 0, 0.5): weights v0 = 0.25, v1 = 0.25, v2 = 0.5
    assert np.allclose(r, np.power([0.25, 0.25, 0.5], 2.2), rtol=2e-6)
    lower = np.float32([143, 179, 203]) / np.float32(255) * np.float32(0.8)
    upper = np.float32([1, 31, 117]) / np.float32(255)
    empty = oracle.OracleScene(np.zeros((0, 9), np.float32), np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), 1)
    horiz = empty.render(one_omm((1, 0, 0), 0.0), IDENT, spp=4)["colour_gpu_order"][0, 0]
    assert np.array_equal(horiz, lower)
    down = empty.render(one_omm((0, -1, 0), 0.0), IDENT, spp=4)["colour_gpu_order"][0, 0]
    assert np.array_equal(down, lower)
    up = empty.render(one_omm((0, 1, 0), 0.0), IDENT, spp=4)["colour_gpu_order"][0, 0]
    assert np.allclose(up, upper, atol=1e-7)


def test_default_background_closed_forms():
    """__miss__default_background (shaders.cu:740-747): r = (atan2(z, x) + pi)
    / 2pi and g = (asin(y) + pi/2) / pi of the normalised direction, b = 0,
    and black where any |component| < 0.01.  The diagonal has the closed form
    (5/8, 1/2 + asin(1/sqrt 3)/pi); directions in every octant (none symmetric
    in x and z, so swapped atan2 arguments show) match float64 numpy; the
    border bands are black on either side of each axis plane and colour just
    outside them."""
    empty = oracle.OracleScene(np.zeros((0, 9), np.float32), np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), 0)

    def miss(dirs):
        d = np.asarray(dirs, np.float32).reshape(-1, 3)
        rays = np.zeros((len(d), 8), np.float32)
        rays[:, 4:7] = d
        rays[:, 7] = 1e16
        return empty.radiance(rays)

    diag = miss([1, 1, 1])[0]
    assert abs(diag[0] - 0.625) <= 2e-7
    assert abs(diag[1] - (0.5 + math.asin(1 / math.sqrt(3)) / math.pi)) <= 2e-7
    assert diag[1] == pytest.approx(0.69591, abs=1e-5)
    assert diag[2] == 0.0
    dirs = np.float32([[2, 1, -3], [-1, 0.5, -2], [-3, -1, 0.5], [0.3, -2, 0.7], [-0.2, 3, 0.9],
                       [0.7, -0.4, -0.1], [-1, -1, -3], [5, 0.2, 0.1]])
    got = miss(dirs)
    d = dirs.astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    want = np.stack([(np.arctan2(d[:, 2], d[:, 0]) + np.pi) / (2 * np.pi),
                     (np.arcsin(d[:, 1]) + np.pi / 2) / np.pi, np.zeros(len(d))], 1)
    assert np.all(np.abs(got - want) <= 2e-6), np.abs(got - want).max()
    # the -x half plane: z -> +0 and z -> -0 give the two ends of the r range
    hi, lo = miss([[-1, 0.5, 0.02], [-1, 0.5, -0.02]])
    assert hi[0] > 0.99 and lo[0] < 0.01
    for axis in range(3):
        for v in (0.0, 0.009, -0.009, 0.0099):
            dd = np.ones(3, np.float32)
            dd[axis] = v
            dd[(axis + 1) % 3] = -1.0
            assert np.all(miss(dd)[0] == 0.0), (axis, v)
        dd = np.ones(3, np.float32)
        dd[axis] = 0.02   # 0.02 / |(0.02, 1, 1)| = 0.01414 > 0.01
        assert np.all(miss(dd)[0][:2] > 0.0), axis
    # the simple_sky mode of the same scene differs, so the mode switch is live
    simple = oracle.OracleScene(np.zeros((0, 9), np.float32), np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), 1)
    rays = np.zeros((1, 8), np.float32)
    rays[0, 4:7] = (1, 1, 1)
    rays[0, 7] = 1e16
    assert not np.array_equal(simple.radiance(rays)[0], diag)


def test_material_colour_exact_and_sum_order():
    """Base-colour triangles are returned unchanged (no linearize).  With 64
    samples the pre-scaled values c/64 are exact and the wave-order sum only
    doubles them, so the average is exactly c; the sequential order rounds at
    the partial sums 3c/64, 5c/64, ... and stays within 64 ulp."""
    v, f = icosphere(1)
    tri = v[f].reshape(-1, 9) * 3
    col = np.tile(np.float32([0.3, 0.55, 0.9] * 3), (len(f), 1))
    o = oracle.OracleScene(tri, col, np.ones(len(f), np.uint32), 1)
    eye = scenes.fibonacci_eye(50, 0.001, 0.1)
    r = o.render(eye, IDENT, spp=64)
    c = np.float32([0.3, 0.55, 0.9])
    assert np.all(r["colour_gpu_order"][0] == c)
    assert np.all(np.abs(r["colour_seq"][0] - c) <= 64 * np.spacing(c))


def test_sum_orders_agree(config1, oracle_c1):
    pose = oracle.pose12(config1["pose"])
    for spp in (32, 100, 256):
        r = oracle_c1.render(config1["eyes"][0][:200], pose, spp=spp, want_f64=True)
        g, s, f64 = r["colour_gpu_order"], r["colour_seq"], r["colour_f64"]
        # summation-order bound: |error| <= spp * eps * sum|x|  (sum|x| == result here)
        bound = spp * 2.0 ** -24 * np.abs(f64) + 1e-30
        assert np.all(np.abs(s - f64) <= bound)
        assert np.all(np.abs(g - f64) <= bound)
        assert np.all(np.abs(g.astype(np.float64) - s) <= 1e-5 * np.abs(s) + 1e-12)


def test_make_color_table():
    c = (np.arange(256, dtype=np.float32) / 255).repeat(3).reshape(-1, 3)
    got = oracle.make_color(c)
    want = np.floor(np.power(c.astype(np.float64), 1 / 2.2) * 255).astype(int)
    assert np.max(np.abs(got[:, :3].astype(int) - want)) <= 1
    assert np.all(got[:, 3] == 255)
    clamp = oracle.make_color(np.float32([[-1, 2, 0.5]]))
    assert clamp[0, 0] == 0 and clamp[0, 1] == 255


# ------------------------------------------------------------------- BVH ----

def test_lbvh_invariants(config1, oracle_c1):
    info = oracle_c1.info()
    # nodes of 128-byte halves: (lox hix loy hiy loz hiz) x 4 slots, 4 child words, 4 pad words
    nh = info["node_bytes"] // 128
    width = 4 * nh
    assert info["node_bytes"] == 128 * nh and width in (4, 8)
    halves = oracle_c1.nodes().view(np.uint8).reshape(-1, nh, 128)
    f = np.concatenate([halves[:, h, :96].copy().view(np.float32).reshape(-1, 6, 4) for h in range(nh)], axis=2)
    child = np.concatenate([halves[:, h, 96:112].copy().view(np.int32).reshape(-1, 4) for h in range(nh)], axis=1)
    assert np.all(halves[:, :, 112:] == 0)
    ids, codes = oracle_c1.order()
    tri = config1["scene"].tri_pos().reshape(-1, 3, 3)
    # sort order: permutation, sorted by (code, id), codes recomputed in float64
    assert np.array_equal(np.sort(ids), np.arange(len(ids)))
    key = codes.astype(np.uint64) << np.uint64(32) | ids.astype(np.uint64)
    assert np.all(np.diff(key.astype(np.int64)) > 0)
    lo, hi = tri.min(axis=1), tri.max(axis=1)
    cen = (lo + hi) * np.float32(0.5)
    clo, chi = cen.min(axis=0), cen.max(axis=0)
    q = np.clip(np.floor((cen[ids].astype(np.float64) - clo) / (chi.astype(np.float64) - clo) * 1024), 0,
                1023).astype(np.uint32)

    def spread(v):
        r = np.zeros_like(v)
        for b in range(10):
            r |= ((v >> b) & 1) << (3 * b)
        return r
    assert np.array_equal((spread(q[:, 0]) << 2) | (spread(q[:, 1]) << 1) | spread(q[:, 2]), codes)
    # every leaf's triangles lie inside the slot bounds of every ancestor slot
    covered = np.zeros(len(ids), int)
    stack = [(0, None)]
    while stack:
        w, anc = stack.pop()
        for s in range(width):
            c = int(child[w, s])
            if c == 0x7FFFFFFF:
                continue
            box = (f[w, 0::2, s], f[w, 1::2, s])
            boxes = (anc or []) + [box]
            if c < 0:
                first, cnt = c & 0x07FFFFFF, ((c >> 27) & 0xF) + 1
                covered[first:first + cnt] += 1
                t = tri[ids[first:first + cnt]]
                for blo, bhi in boxes:
                    assert np.all(t.min(axis=1) >= blo) and np.all(t.max(axis=1) <= bhi)
            else:
                assert c > w
                stack.append((c, boxes))
    assert np.all(covered == 1)
    # slots fill in order, empty slots hold the far box
    filled = child != 0x7FFFFFFF
    assert np.all(filled[:, :-1] | ~filled[:, 1:])
    assert np.all(f[:, :, :][~np.repeat(filled[:, None, :], 6, axis=1)] == np.float32(3.0e38))
    # a node pushes at most width - 1 children per level on the way down
    assert info["stack_bound"] <= (width - 1) * info["levels"]


# ---------------------------------------------------------- convergence ----

def test_monte_carlo_convergence(config1, oracle_c1):
    """RMS error against an FP64 4096-spp render falls as N^-1/2."""
    pose = oracle.pose12(config1["pose"])
    omm = config1["eyes"][0][:120]
    ref = oracle_c1.render(omm, pose, spp=4096, want_f64=True, frame=7)["colour_f64"]
    errs = []
    for spp in (16, 64, 256):
        c = oracle_c1.render(omm, pose, spp=spp)["colour_gpu_order"]
        errs.append(float(np.sqrt(np.mean((c - ref) ** 2))))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 1.5 < r1 < 2.7 and 1.5 < r2 < 2.7, errs


# ------------------------------------------------------------------ MLP ----

def mlp_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    sizes = [(oracle.MLP_IN * 64, oracle.MLP_IN), (64, 0), (64 * 64, 64), (64, 0), (64 * 64, 64), (64, 0), (64, 64), (1, 0)]
    parts = []
    for n, fan in sizes:
        lim = math.sqrt(6.0 / fan) if fan else 0.1
        parts.append(rng.uniform(-lim, lim, n) * scale)
    return np.concatenate(parts).astype(np.float32)


def test_mlp_backprop_matches_finite_differences():
    """The FP64 oracle MLP's analytic gradient equals central differences
    (which pins the backward pass the GPU gradients are compared with)."""
    rng = np.random.default_rng(1)
    p = mlp_params(1)
    x = rng.normal(size=(40, oracle.MLP_IN)).astype(np.float32)
    t = rng.normal(size=40).astype(np.float32)
    _, g, loss = oracle.mlp_loss_grad(p, x, t, mode=0)
    assert loss > 0
    idx = np.concatenate([rng.choice(768, 12, replace=False), 768 + rng.choice(64, 6, replace=False),
                          832 + rng.choice(4096, 12, replace=False), 4992 + rng.choice(4096, 12, replace=False),
                          9088 + rng.choice(64, 6, replace=False), 9152 + np.arange(0, 64, 9), [9216]])
    p64 = p.astype(np.float64)
    for k in idx:
        # small steps keep the ReLU kinks out of the difference quotient
        h = 1e-5 * max(1.0, abs(p64[k]))
        # parameters are float32 at the boundary: use perturbations exactly representable
        hp = np.float32(p64[k] + h) - np.float32(p64[k])
        pp, pm = p.copy(), p.copy()
        pp[k] += hp
        pm[k] -= hp
        lp = oracle.mlp_loss_grad(pp, x, t, mode=0)[2]
        lm = oracle.mlp_loss_grad(pm, x, t, mode=0)[2]
        fd = (lp - lm) / (2 * float(hp))
        assert abs(fd - g[k]) <= 2e-3 * max(1e-3, abs(g[k])) + 1e-6, (k, fd, g[k])


def test_mlp_bf16_mode_close_to_exact():
    rng = np.random.default_rng(2)
    p = mlp_params(2)
    x = rng.normal(size=(200, oracle.MLP_IN)).astype(np.float32)
    t = rng.normal(size=200).astype(np.float32)
    y0, g0, _ = oracle.mlp_loss_grad(p, x, t, mode=0)
    y1, g1, _ = oracle.mlp_loss_grad(p, x, t, mode=1)
    assert np.max(np.abs(y1 - y0)) < 0.05 * np.max(np.abs(y0))
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)


def test_spherical_ids_closed_forms():
    """Spherical projections (shaders.cu:412-640): the six axis ommatidia own
    the pixels looking along their axes; on random eyes the chosen ommatidium
    maximises the float64 cosine (up to near-ties); the split variant keeps
    each half of the window to ommatidia on its side of x = 0."""
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    omm = np.zeros((6, 8), np.float32)
    omm[:, 0:3] = axes * 0.001
    omm[:, 3:6] = axes
    w, h = 8, 4
    for variant in (0, 1):
        ids = oracle.spherical_ids(omm, w, h, variant)
        # d = 2 (x/w, y/h) - 1; lon = pi/2 - pi d.x; lat = pi/2 d.y
        assert ids[2, 4] == 4            # d = (0, 0): lon = pi/2 -> (0, 0, 1)
        assert ids[2, 6] == 0            # d.x = 1/2: lon = 0 -> (1, 0, 0)
        assert ids[2, 0] == 5            # d.x = -1: lon = 3 pi/2 -> (0, 0, -1)
        assert ids[0, 4] == 3            # d.y = -1: lat = -pi/2 -> (0, -1, 0)
    rng = np.random.default_rng(4)
    d = rng.normal(size=(300, 3))
    omm = np.zeros((300, 8), np.float32)
    omm[:, 3:6] = d
    omm[:, 0:3] = d * 0.001
    w, h = 40, 20
    ids = oracle.spherical_ids(omm, w, h, 1)
    ys, xs = np.mgrid[0:h, 0:w]
    lon = (2 * xs / w - 1) * -np.pi + np.pi / 2
    lat = (2 * ys / h - 1) * np.pi / 2
    s = np.stack([np.cos(lon) * np.cos(lat), np.sin(lat), np.sin(lon) * np.cos(lat)], -1)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    cos = s.reshape(-1, 3) @ dn.T
    want = cos.argmax(axis=1)
    got = ids.reshape(-1)
    agree = got == want
    assert agree.mean() > 0.995
    gap = cos[np.arange(len(got)), want] - cos[np.arange(len(got)), got]
    assert np.all(gap[~agree] < 1e-5)
    # split: left half only ommatidia with x < 0 (index 0 is the initial pick)
    sids = oracle.spherical_ids(omm, w, h, 2)
    for y in range(h):
        for x in range(w):
            i = sids[y, x]
            if i == 0:
                continue
            if x / w < 0.5:
                assert omm[i, 0] < 0
            elif x / w > 0.5:
                assert omm[i, 0] > 0


def test_camera_rays_closed_forms():
    """Regular cameras (shaders.cu:198-333) on an identity pose: the pinhole's
    centre pixel looks along z and its corners along z - x sx - y sy; the
    orthographic camera's rays are parallel to z from a grid spanning +-sx,
    +-sy; panoramic directions are unit vectors, the centre pixel along z and
    a quarter of the width to the left along -x."""
    pose = np.array([1, 2, 3, 1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    w, h = 8, 4
    sc = np.array([0.5, 0.25, 1.0], np.float32)
    pin = oracle.camera_rays(0, pose, sc, w, h).reshape(h, w, 8)
    assert np.array_equal(pin[2, 4, 4:7], [0, 0, 1])
    assert np.array_equal(pin[0, 0, 4:7], [-0.5, -0.25, 1])
    assert np.all(pin[..., 0:3] == [1, 2, 3]) and np.all(pin[..., 3] == np.float32(0.01))
    assert np.all(pin[..., 7] == np.float32(1e16))
    ortho = oracle.camera_rays(1, pose, np.array([2.0, 1.0, 0.0], np.float32), w, h).reshape(h, w, 8)
    assert np.all(ortho[..., 4:7] == [0, 0, 1])
    assert np.array_equal(ortho[0, 0, 0:3], [1 - 2, 2 - 1, 3])
    assert np.allclose(ortho[:, :, 0], 1 + 2 * (2 * np.arange(w) / w - 1)[None, :])
    pano = oracle.camera_rays(2, pose, sc, w, h).reshape(h, w, 8)
    assert np.allclose(np.linalg.norm(pano[..., 4:7], axis=-1), 1.0, atol=1e-6)
    assert np.allclose(pano[2, 4, 4:7], [0, 0, 1], atol=1e-6)
    assert np.allclose(pano[2, 6, 4:7], [1, 0, 0], atol=1e-6)    # lon = 0
    assert np.allclose(pano[2, 2, 4:7], [-1, 0, 0], atol=1e-6)   # lon = pi
    assert np.all(pano[..., 0:3] == [1, 2, 3])


def test_radiance_is_shaded_trace(oracle_c1, config1):
    """oracle radiance = shading of the closest hit, or the sky on a miss: rays
    straight down hit the terrain (vertex-colour shading); rays straight up see
    the top sky colour."""
    n = 64
    rng = np.random.default_rng(2)
    rays = np.zeros((2 * n, 8), np.float32)
    rays[:, 0] = rng.uniform(-4, 4, 2 * n)
    rays[:, 2] = rng.uniform(-4, 4, 2 * n)
    rays[:, 1] = 5.0
    rays[:n, 5] = -1.0
    rays[n:, 5] = 1.0
    rays[:, 3] = 0.0
    rays[:, 7] = 1e16
    col = oracle_c1.radiance(rays)
    ids, _ = oracle_c1.trace(rays)
    assert np.all(ids[:n] >= 0) and np.all(ids[n:] < 0)
    upper = np.float32([1, 31, 117]) / np.float32(255)
    assert np.allclose(col[n:], upper, atol=1e-6)
    assert np.all(col[:n] >= 0) and np.all(col[:n] <= 1)
    ref = oracle_c1.render(config1["eyes"][0][:4], oracle.pose12(config1["pose"]), spp=1, want_samples=True)
    assert ref["samples"].shape[0] == 4
