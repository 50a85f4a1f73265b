"""Pins of the oracle's BSDF sampling and bounce loop by what the
mathematics fixes (no GPU).

The GPU renders are compared with the oracle bit for bit, so these tests
are what stands between a plausible mistake shared by both sides and a
green suite.  Each one is chosen so that such a mistake fails it:

Samplers (rendering/taichi/scatter.py, through orc_scatter_dirs):
  * cosine_hemisphere: unit, in the hemisphere, P(cos t <= c) = c^2,
    E[cos t] = 2/3 (a uniform hemisphere gives c and 1/2), uniform azimuth,
    for normals on both helper-axis branches (:14-30)
  * random_unit_vector: unit, z uniform on [-1, 1] (Archimedes), E[x^2] =
    E[y^2] = E[z^2] = 1/3 (:6-11)
  * metal_scatter: the exact mirror at fuzz 0; E[d.m] = 1 - r^2/3 for
    fuzz r <= 1 (:50-54)
  * dielectric_scatter: reflects iff Schlick(cos, eta) > xi, with Schlick's
    r0 = ((1-eta)/(1+eta))^2 (0.04 at normal incidence), always reflects
    past the critical angle, and refracts by Snell's law, |perp| = eta sin
    (:40-49, :57-60, :71-81)
Bounce loop (rendering/taichi/trace.py:139-236, through the oracle render):
  * a diffuse floor under the sky gradient has expectation
    a * (7/12, 3/4, 1) (E[d_y] = 2/3 for a cosine-weighted direction)
  * a metal floor of fuzz r: a * sky(1 - r^2/3); a glossy floor of
    roughness r: a * sky(r 2/3 + (1 - r)(1 - r^2/3)) (the diffuse branch
    taken with probability r, :62-69)
  * a glass floor seen at 60 degrees reflects with probability
    0.04 + 0.96 * 0.5^5 and otherwise refracts by eta = 1/1.5 (an inverted
    eta reflects totally)
  * dielectric bookkeeping (:216-225): a glass slab entered at 60 degrees
    leaves by eta = 1.5 at its back face, and an internal reflection there
    resets current_ior to 1, so the ray then crosses the top face unbent;
    a face met from behind has eta = current_ior / 1 (a floor wound
    downward is crossed unbent from above); a camera inside a glass cube starts at current_ior 1 and crosses a face
    head-on without refraction
  * Russian roulette (:227-232): a mirror corridor left after k bounces
    returns exactly a^k for k = 3 (no roulette before depth 3); for k = 4, 6
    every sample is 0 or the throughput divided by its largest channel,
    with survival a_max^k, so the expectation a^k is unbiased
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_oracle_pins import _camera, _tri_scene

N_DIRS = 400_000


def _uniforms(seed, n=N_DIRS):
    rng = np.random.default_rng(seed)
    return rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32)


def _within(got, want, sd, k=5.0):
    assert abs(got - want) <= k * sd, (got, want, sd)


# ------------------------------------------------------------------ samplers
@pytest.mark.parametrize("normal", [(0, 1, 0), (0.95, 0.2, -0.2396), (-0.3, -0.5, 0.8124)])
def test_cosine_hemisphere_distribution(normal):
    n = np.asarray(normal, np.float64)
    n /= np.linalg.norm(n)
    u1, u2 = _uniforms(1)
    d = orc.scatter_dirs("cosine", u1, u2, n=n).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    c = d @ n
    assert c.min() >= -1e-6
    sd = math.sqrt(0.5 - 4.0 / 9.0) / math.sqrt(len(c))
    _within(c.mean(), 2.0 / 3.0, sd)
    # CDF of cos(theta) is c^2 (Kolmogorov-Smirnov bound at 5 sigma)
    for q in (0.1, 0.3, 0.5, 0.7, 0.9):
        p = q * q
        _within(np.mean(c <= q), p, math.sqrt(p * (1 - p) / len(c)))
    # azimuth uniform about n
    t = np.cross(n, [0.0, 0.0, 1.0] if abs(n[2]) < 0.9 else [1.0, 0.0, 0.0])
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    phi = np.arctan2(d @ b, d @ t)
    for lo in np.linspace(-math.pi, math.pi, 8, endpoint=False):
        _within(np.mean((phi >= lo) & (phi < lo + math.pi / 4)), 0.125, math.sqrt(0.125 * 0.875 / len(c)))


def test_random_unit_vector_distribution():
    u1, u2 = _uniforms(2)
    d = orc.scatter_dirs("unit", u1, u2).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    n = len(d)
    for k in range(3):
        _within(d[:, k].mean(), 0.0, math.sqrt(1 / 3) / math.sqrt(n))
        _within((d[:, k] ** 2).mean(), 1.0 / 3.0, math.sqrt(4 / 45) / math.sqrt(n))
    # z uniform on [-1, 1]
    for q in (-0.8, -0.4, 0.0, 0.4, 0.8):
        p = (q + 1) / 2
        _within(np.mean(d[:, 2] <= q), p, math.sqrt(p * (1 - p) / n))


def test_metal_fuzz_zero_is_the_mirror():
    d = np.float32([0.6, -0.8, 0.0])
    nrm = np.float32([0.0, 1.0, 0.0])
    u1, u2 = _uniforms(3, 1000)
    out = orc.scatter_dirs("metal", u1, u2, d=d, n=nrm, param=0.0)
    np.testing.assert_allclose(out, np.tile([0.6, 0.8, 0.0], (1000, 1)), atol=1e-6)


@pytest.mark.parametrize("fuzz", [0.2, 0.5, 0.9])
def test_metal_fuzz_closed_form(fuzz):
    """d' = normalize(m + r u), u uniform on the sphere: E[d'.m] = 1 - r^2/3."""
    d = np.float32([0.6, -0.8, 0.0])
    m = np.float64([0.6, 0.8, 0.0])
    u1, u2 = _uniforms(4)
    out = orc.scatter_dirs("metal", u1, u2, d=d, n=(0, 1, 0), param=fuzz).astype(np.float64)
    c = out @ m
    _within(c.mean(), 1.0 - fuzz * fuzz / 3.0, c.std() / math.sqrt(len(c)))


def _glass(d, n, eta, xi):
    xi = np.asarray(xi, np.float32)
    return orc.scatter_dirs("dielectric", xi, d=d, n=n, param=eta).astype(np.float64)


def _schlick(cos, eta):
    r0 = ((1 - eta) / (1 + eta)) ** 2
    return r0 + (1 - r0) * (1 - cos) ** 5


@pytest.mark.parametrize("theta_deg,eta", [(0.0, 1 / 1.5), (0.0, 1.5), (60.0, 1 / 1.5), (30.0, 1.5), (45.0, 1 / 1.33)])
def test_dielectric_schlick_threshold_and_snell(theta_deg, eta):
    th = math.radians(theta_deg)
    d = np.float32([math.sin(th), -math.cos(th), 0.0])
    n = (0.0, 1.0, 0.0)
    cos = float(-np.dot(d.astype(np.float64), n))
    p = _schlick(cos, eta)
    if theta_deg == 0.0:
        assert abs(p - 0.04) < 1e-12
    if theta_deg == 60.0:
        assert abs(p - (0.04 + 0.96 * 0.5 ** 5)) < 1e-7
    below, above = _glass(d, n, eta, [p * (1 - 1e-4), min(p * (1 + 1e-4) + 1e-7, 0.999)])
    mirror = np.float64([d[0], -d[1], 0.0])
    np.testing.assert_allclose(below, mirror, atol=1e-6)         # xi < Schlick: reflect
    assert above[1] < 0                                           # xi > Schlick: refract
    np.testing.assert_allclose(np.linalg.norm(above), 1.0, atol=1e-6)
    sin_t = math.hypot(above[0], above[2])
    assert abs(sin_t - eta * math.sin(th)) < 2e-6                 # Snell
    # the reflected fraction over uniform xi is the Schlick probability
    xi = np.random.default_rng(5).random(200_000, dtype=np.float32)
    refl = _glass(d, n, eta, xi)[:, 1] > 0
    _within(refl.mean(), p, math.sqrt(p * (1 - p) / len(xi)))


def test_dielectric_total_internal_reflection():
    """From inside (eta = 1.5) past the critical angle asin(1/1.5) = 41.8
    degrees every xi reflects; just below it refraction happens."""
    n = (0.0, 1.0, 0.0)
    xi = np.random.default_rng(6).random(10_000, dtype=np.float32)
    for deg in (42.0, 60.0, 85.0):
        th = math.radians(deg)
        d = np.float32([math.sin(th), -math.cos(th), 0.0])
        out = _glass(d, n, 1.5, xi)
        np.testing.assert_allclose(out, np.tile([d[0], -d[1], 0.0], (len(xi), 1)), atol=1e-6)
    th = math.radians(41.0)
    d = np.float32([math.sin(th), -math.cos(th), 0.0])
    assert (_glass(d, n, 1.5, xi)[:, 1] < 0).mean() > 0.5


# ------------------------------------------------------------------ bounce loop
FLOOR = [[[-100, 0, -100], [100, 0, -100], [100, 0, 100]], [[-100, 0, -100], [100, 0, 100], [-100, 0, 100]]]
# the same floor wound the other way: face normal +y, so a ray from above is
# a front-face hit (trace.py:161-163 decides inside/outside by the normal)
FLOOR_UP = [[t[0], t[2], t[1]] for t in FLOOR]


def _sky(dy):
    t = 0.5 * (dy + 1.0)
    return np.array([1 - t + 0.5 * t, 1 - t + 0.7 * t, 1.0])


def _render(sc, cam, prm, spp, first=0):
    acc = orc.Accum(cam.width * cam.height)
    orc.render(sc, cam, prm, first, spp, acc)
    return acc


def _down_camera(w=64, h=64):
    return _camera((0, 1.0, 0), (0, -1, 0), up=(0, 0, 1), fov_deg=1e-3, w=w, h=h)


def _pixel_mean_check(acc, want):
    c = acc.color.astype(np.float64)
    sd = c.std(axis=0) / math.sqrt(len(c)) + 2e-7  # floor: fp32 rounding of constant channels
    for k in range(3):
        _within(c[:, k].mean(), want[k], sd[k])


def test_diffuse_floor_under_sky_gradient():
    a = np.float64([0.3, 0.6, 0.9])
    sc = _tri_scene(FLOOR, mats=[(0, tuple(a), 0.0, 1.0, 0.0, -1)])
    acc = _render(sc, _down_camera(), orc.params(max_depth=2, use_sky=1, seed=11), 8)
    want = a * np.float64([7 / 12, 3 / 4, 1.0])
    _pixel_mean_check(acc, want)
    # a uniform hemisphere would give E[d_y] = 1/2: a * sky(1/2), far outside
    assert abs(acc.color[:, 0].mean() - a[0] * _sky(0.5)[0]) > 0.01


@pytest.mark.parametrize("fuzz", [0.3, 0.8])
def test_metal_floor_fuzz_under_sky_gradient(fuzz):
    a = np.float64([0.9, 0.8, 0.7])
    sc = _tri_scene(FLOOR, mats=[(1, tuple(a), fuzz, 1.0, 0.0, -1)])
    acc = _render(sc, _down_camera(), orc.params(max_depth=2, use_sky=1, seed=12), 8)
    _pixel_mean_check(acc, a * _sky(1 - fuzz * fuzz / 3))


@pytest.mark.parametrize("rough", [0.25, 0.6])
def test_glossy_floor_mixes_diffuse_with_probability_roughness(rough):
    a = np.float64([0.5, 0.7, 0.9])
    sc = _tri_scene(FLOOR, mats=[(4, tuple(a), rough, 1.0, 0.0, -1)])
    acc = _render(sc, _down_camera(), orc.params(max_depth=2, use_sky=1, seed=13), 8)
    dy = rough * 2 / 3 + (1 - rough) * (1 - rough * rough / 3)
    _pixel_mean_check(acc, a * _sky(dy))
    # the branch probabilities swapped (diffuse with 1 - r) would give this
    wrong = (1 - rough) * 2 / 3 + rough * (1 - rough * rough / 3)
    assert abs(acc.color[:, 1].mean() - a[1] * _sky(wrong)[1]) > 0.002


def _classes(values, centres, tol=2e-4):
    """Index of the class each sample's red channel falls in (all must)."""
    dist = np.abs(values[:, None] - np.asarray(centres)[None, :])
    assert (dist.min(axis=1) <= tol).all(), np.unique(np.round(values, 4))
    return dist.argmin(axis=1)


def _frac_check(cls, k, p):
    n = len(cls)
    _within(np.mean(cls == k), p, math.sqrt(p * (1 - p) / n) + 1e-9)


def test_glass_floor_at_60_degrees():
    sc = _tri_scene(FLOOR_UP, mats=[(2, (1.0, 1.0, 1.0), 0.0, 1.5, 0.0, -1)])
    cam = _camera((0, 1.0, 0), (math.sin(math.radians(60)), -0.5, 0), up=(0, 1, 0), fov_deg=1e-3, w=128, h=128)
    acc = _render(sc, cam, orc.params(max_depth=2, use_sky=1, seed=14), 1)
    refl = _sky(0.5)
    refr = _sky(-math.sqrt(1 - (math.sin(math.radians(60)) / 1.5) ** 2))
    cls = _classes(acc.color[:, 0].astype(np.float64), [refl[0], refr[0]])
    _frac_check(cls, 0, 0.04 + 0.96 * 0.5 ** 5)
    np.testing.assert_allclose(acc.color[cls == 1], np.tile(refr, ((cls == 1).sum(), 1)), atol=2e-4)


def test_glass_floor_seen_from_its_back_is_crossed_unbent():
    """The same floor wound downward: a ray from above meets a back face,
    so eta = current_ior / 1 = 1 (trace.py:221-223): Schlick with r0 = 0
    reflects with probability 0.5^5 and every other ray goes on unbent."""
    sc = _tri_scene(FLOOR, mats=[(2, (1.0, 1.0, 1.0), 0.0, 1.5, 0.0, -1)])
    cam = _camera((0, 1.0, 0), (math.sin(math.radians(60)), -0.5, 0), up=(0, 1, 0), fov_deg=1e-3, w=128, h=128)
    acc = _render(sc, cam, orc.params(max_depth=2, use_sky=1, seed=21), 1)
    cls = _classes(acc.color[:, 0].astype(np.float64), [_sky(0.5)[0], _sky(-0.5)[0]])
    _frac_check(cls, 0, 0.5 ** 5)


def test_glass_floor_at_normal_incidence():
    sc = _tri_scene(FLOOR_UP, mats=[(2, (1.0, 1.0, 1.0), 0.0, 1.5, 0.0, -1)])
    acc = _render(sc, _down_camera(128, 128), orc.params(max_depth=2, use_sky=1, seed=15), 1)
    cls = _classes(acc.color[:, 0].astype(np.float64), [_sky(1.0)[0], _sky(-1.0)[0]])
    _frac_check(cls, 0, 0.04)


def test_glass_slab_ior_bookkeeping():
    """Top face y = 0 (normal +y), bottom face y = -1 (normal -y), glass
    1.5, entered at 60 degrees.  Outcomes (red channel of the sky):
      reflect at the top                                      p 0.07
      refract in (eta 1/1.5), out at the bottom (eta 1.5)     p 0.93 (1 - s)
      refract in, reflect inside (current_ior -> 1), cross
      the top unbent (eta 1)                                  p 0.93 s
    with s = Schlick(cos, 1.5) at the bottom, cos = sqrt(1 - (sin60/1.5)^2)."""
    top = [[[-100, 0, -100], [100, 0, 100], [100, 0, -100]], [[-100, 0, -100], [-100, 0, 100], [100, 0, 100]]]
    bot = [[[-100, -1, -100], [100, -1, -100], [100, -1, 100]], [[-100, -1, -100], [100, -1, 100], [-100, -1, 100]]]
    tris = np.asarray(top + bot, np.float32)
    sc = _tri_scene(tris, mats=[(2, (1.0, 1.0, 1.0), 0.0, 1.5, 0.0, -1)])
    # the helper takes face normals from the winding: top +y, bottom -y
    assert sc.n0[:, 1].max() == 1.0 and sc.n0[:, 1].min() == -1.0
    s60 = math.sin(math.radians(60))
    cam = _camera((0, 1.0, 0), (s60, -0.5, 0), up=(0, 1, 0), fov_deg=1e-3, w=128, h=128)
    acc = _render(sc, cam, orc.params(max_depth=6, use_sky=1, seed=16), 1)
    cos_in = math.sqrt(1 - (s60 / 1.5) ** 2)
    s = _schlick(cos_in, 1.5)
    centres = [_sky(0.5)[0], _sky(-0.5)[0], _sky(cos_in)[0]]
    cls = _classes(acc.color[:, 0].astype(np.float64), centres)
    r_top = 0.04 + 0.96 * 0.5 ** 5
    _frac_check(cls, 0, r_top)
    _frac_check(cls, 1, (1 - r_top) * (1 - s))
    _frac_check(cls, 2, (1 - r_top) * s)


def _cube(half=1.0):
    """12 triangles of an axis-aligned cube, outward face normals."""
    tris, nrms = [], []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            u, v = [(a) for a in range(3) if a != axis]
            c = np.zeros((4, 3))
            for k, (su, sv) in enumerate([(-1, -1), (1, -1), (1, 1), (-1, 1)]):
                c[k, axis] = sign * half
                c[k, u] = su * half
                c[k, v] = sv * half
            for tri in ([c[0], c[1], c[2]], [c[0], c[2], c[3]]):
                tris.append(tri)
                nv = np.zeros(3)
                nv[axis] = sign
                nrms.append(nv)
    return np.asarray(tris, np.float32), np.asarray(nrms, np.float32)


def test_camera_inside_glass_cube_starts_at_ior_one():
    """current_ior starts at 1 (trace.py:143): from inside, the first face is
    a back face with eta = 1 / 1 and is crossed head-on without bending, so
    every sample is albedo * sky(d) exactly."""
    tris, nrms = _cube()
    alb = (0.9, 0.8, 0.7)
    sc = orc.OracleScene(tris[:, 0], tris[:, 1], tris[:, 2], nrms, nrms, nrms, np.zeros(len(tris), np.int32),
                         [(2, alb, 0.0, 1.5, 0.0, -1)])
    cam = _camera((0, 0, 0), (0, 1, 0), up=(0, 0, 1), fov_deg=1e-3, w=32, h=32)
    acc = _render(sc, cam, orc.params(max_depth=4, use_sky=1, seed=17), 4)
    np.testing.assert_allclose(acc.color, np.tile(np.float64(alb) * _sky(1.0), (32 * 32, 1)), atol=1e-5)


def _corridor(k):
    """Two facing mirror plates, y = 0 and y = 1, x in [0, k]; a ray from
    (0, 0.5, 0) along (1, -1, 0) hits them at x = 0.5, 1.5, ... and leaves
    after k bounces.  Plate diagonals avoid the hit points."""
    tris = []
    for y in (0.0, 1.0):
        tris.append([[0, y, -5], [k, y, 4], [k, y, -5]])
        tris.append([[0, y, -5], [0, y, 4], [k, y, 4]])
    return tris


ALB = (0.8, 0.5, 0.3)


def _corridor_render(k, seed):
    sc = _tri_scene(_corridor(k), mats=[(1, ALB, 0.0, 1.0, 0.0, -1)])
    cam = _camera((0, 0.5, 0), (1, -1, 0), up=(0, 1, 0), fov_deg=1e-3, w=128, h=128)
    return _render(sc, cam, orc.params(max_depth=k + 1, use_sky=0, bg=(1, 1, 1), seed=seed, sample_clamp=0.0), 1)


def test_corridor_three_bounces_has_no_roulette():
    acc = _corridor_render(3, 18)
    a = np.float32(ALB)
    want = (a * a) * a
    np.testing.assert_allclose(acc.color, np.tile(want, (len(acc.color), 1)), rtol=1e-6)


@pytest.mark.parametrize("k,seed", [(4, 19), (6, 20)])
def test_russian_roulette_is_unbiased(k, seed):
    acc = _corridor_render(k, seed)
    a = np.float64(ALB)
    survivor = (a / a.max()) ** k
    cls = _classes(acc.color[:, 0].astype(np.float64), [0.0, 1.0], tol=1e-6)
    _frac_check(cls, 1, a.max() ** k)
    alive = acc.color[cls == 1].astype(np.float64)
    np.testing.assert_allclose(alive, np.tile(survivor, (len(alive), 1)), rtol=1e-5)
    assert (acc.color[cls == 0] == 0).all()
    # unbiased: the mean is a^k per channel
    c = acc.color.astype(np.float64)
    for ch in range(3):
        _within(c[:, ch].mean(), a[ch] ** k, c[:, ch].std() / math.sqrt(len(c)) + 1e-12)
