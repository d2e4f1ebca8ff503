"""Pins the oracle's camera (src/camera.cpp:40-146) and projector
(src/render-engine.cpp:797-1042) by matrix properties, library routines and the
projected-grid invariants, and checks the library's own (double) host projector
against it."""
import math

import numpy as np
import pytest

import oracle as orc
import wavetool_amd as wt
from wavetool_amd import scenes


def cam(height=10.0, pitch=-15.0, yaw=-90.0, fov=60.0, aspect=16 / 9, near=0.1, far=1000.0, x=0.0, z=0.0):
    return scenes.camera_tuple(height=height, pitch_deg=pitch, yaw_deg=yaw, fov_deg=fov, aspect=aspect, near=near,
                               far=far, x=x, z=z)


def cross2(a, b):
    return a[0] * b[1] - a[1] * b[0]


def test_inverse_matches_numpy():
    rng = np.random.default_rng(0)
    for _ in range(20):
        m = rng.normal(size=(4, 4))
        assert np.allclose(orc.inverse(m), np.linalg.inv(m), atol=1e-10)


def test_camera_basis_and_view_matrix():
    for c in (cam(), cam(yaw=30.0, pitch=-40.0, height=3.0, x=5.0, z=-7.0), cam(yaw=200.0, pitch=60.0)):
        m = orc.camera(c)
        yaw, pitch = math.radians(c[7]), math.radians(c[8])
        fwd = np.array([math.cos(yaw) * math.cos(pitch), math.sin(pitch), math.sin(yaw) * math.cos(pitch)])
        assert np.allclose(m["forward"], fwd, atol=1e-12)
        V = m["view"]
        eye = np.array(c[4:7])
        # the view matrix is a rigid motion taking the eye to the origin and forward to -z
        R = V[:3, :3]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(R) - 1) < 1e-12
        assert np.allclose(V @ np.append(eye, 1), [0, 0, 0, 1], atol=1e-9)
        assert np.allclose(R @ fwd, [0, 0, -1], atol=1e-12)
        # up stays in the vertical plane through forward (no roll)
        assert abs(np.cross(m["right"], [0, 1, 0]) @ fwd) < 1e-12 or abs(m["right"][1]) < 1e-12


def test_perspective_maps_frustum_to_ndc():
    c = cam(fov=72.0, aspect=1.5, near=0.5, far=80.0)
    P = orc.camera(c)["proj"]
    n, f, t = 0.5, 80.0, math.tan(math.radians(72.0) / 2)
    for z, expect in ((-n, -1.0), (-f, 1.0)):
        q = P @ np.array([0, 0, z, 1])
        assert abs(q[2] / q[3] - expect) < 1e-12
    q = P @ np.array([0, t * 10.0, -10.0, 1])  # top edge of the frustum at depth 10
    assert abs(q[1] / q[3] - 1.0) < 1e-12
    q = P @ np.array([1.5 * t * 10.0, 0, -10.0, 1])  # right edge (aspect 1.5)
    assert abs(q[0] / q[3] - 1.0) < 1e-12


CAMERAS = [cam(), cam(fov=72.0, aspect=1.0, far=100.0, height=4.0, pitch=0.0, z=70.0),
           cam(height=2.0, pitch=-35.0, yaw=17.0), cam(height=50.0, pitch=-5.0, yaw=300.0),
           cam(height=0.5, pitch=-80.0), cam(height=-3.0, pitch=10.0), cam(height=10.0, pitch=0.0)]


@pytest.mark.parametrize("c", CAMERAS)
def test_projector_invariants(c):
    amp = 2.0
    st, corners, info = orc.projector(c, amp)
    assert st == 0
    # corners lie on y = 0
    assert np.abs(corners[:, 1]).max() < 1e-9
    # the projector sits outside the displaceable slab, above it
    assert info["projector_pos"][1] >= amp + 1.0 - 1e-9 or c[5] >= amp + 1.0
    # every frustum-slab point, dropped onto y = 0, lies inside the projector range
    # [xmin, xmax] x [ymin, ymax] (the range matrix maps [0,1]^2 onto it)
    ndc = info["ndc_points"]
    assert (ndc[:, 0] >= info["xmin"] - 1e-9).all() and (ndc[:, 0] <= info["xmax"] + 1e-9).all()
    assert (ndc[:, 1] >= info["ymin"] - 1e-9).all() and (ndc[:, 1] <= info["ymax"] + 1e-9).all()
    # the corner quad is convex with consistent orientation BL -> BR -> TR -> TL
    BL, TL, BR, TR = corners[:, [0, 2]]
    quad = [BL, BR, TR, TL]
    crosses = [cross2(quad[(i + 1) % 4] - quad[i], quad[(i + 2) % 4] - quad[(i + 1) % 4]) for i in range(4)]
    assert all(np.sign(x) == np.sign(crosses[0]) for x in crosses)


def test_symmetric_camera_gives_symmetric_corners():
    st, corners, _ = orc.projector(cam(yaw=-90.0), 2.0)  # looking along -z from x = 0
    BL, TL, BR, TR = corners
    assert abs(BL[0] + BR[0]) < 1e-9 and abs(TL[0] + TR[0]) < 1e-9
    assert abs(BL[2] - BR[2]) < 1e-9 and abs(TL[2] - TR[2]) < 1e-9


def test_aim_point_inside_grid_when_looking_down():
    c = cam(height=20.0, pitch=-50.0, yaw=40.0)
    st, corners, info = orc.projector(c, 2.0)
    fwd = orc.camera(c)["forward"]
    eye = np.array(c[4:7])
    hit = eye + fwd * (-eye[1] / fwd[1])  # forward ray on y = 0
    BL, TL, BR, TR = corners[:, [0, 2]]
    quad = [BL, BR, TR, TL]
    s = [cross2(quad[(i + 1) % 4] - quad[i], hit[[0, 2]] - quad[i]) for i in range(4)]
    assert all(np.sign(x) == np.sign(s[0]) for x in s)


def test_no_intersection_when_frustum_misses_slab():
    # high above the slab and looking up: the reference skips the water (render-engine.cpp:894)
    st, corners, info = orc.projector(cam(height=100.0, pitch=80.0, far=50.0), 2.0)
    assert st == orc.NO_INTERSECTION and info["npoints"] == 0


def test_underwater_and_horizon_branches():
    st, corners, info = orc.projector(cam(height=-3.0, pitch=10.0), 2.0)
    assert st == 0 and info["raised"]
    # underwater: lifted by MIN - 2 * y (render-engine.cpp:915)
    assert abs(info["projector_pos"][1] - (2.0 + 1.0 - 2 * -3.0 + -3.0)) < 1e-9
    st, corners, info = orc.projector(cam(height=10.0, pitch=0.0), 2.0)
    assert st == 0 and np.isfinite(corners).all()


@pytest.mark.parametrize("c", CAMERAS)
def test_library_projector_matches_oracle(c):
    """The product's host projector (double, Gauss-Jordan inverse) against the oracle
    (cofactor inverse): an independent implementation of the same steps."""
    st, ref, _ = orc.projector(c, 2.0)
    got = wt.projector_corners(c, 2.0)
    assert np.abs(got - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


def test_library_projector_many_random_cameras():
    rng = np.random.default_rng(11)
    for _ in range(200):
        c = cam(height=rng.uniform(-20, 60), pitch=rng.uniform(-89, 89), yaw=rng.uniform(0, 360),
                fov=rng.uniform(20, 100), aspect=rng.uniform(0.5, 2.5), far=rng.uniform(50, 2000))
        st, ref, _ = orc.projector(c, 2.0)
        if st != 0:
            with pytest.raises(wt.WtError) as e:
                wt.projector_corners(c, 2.0)
            assert e.value.status == st
            continue
        got = wt.projector_corners(c, 2.0)
        assert np.abs(got - ref).max() <= 1e-8 * max(1.0, np.abs(ref).max())


def test_library_camera_matrices_match_oracle():
    for c in CAMERAS:
        v, p, f = wt.camera_matrices(c)
        m = orc.camera(c)
        assert np.allclose(v, m["view"], atol=1e-12) and np.allclose(p, m["proj"], atol=1e-12)
        assert np.allclose(f, m["forward"], atol=1e-12)


# ------------------------------------------------------------------ closed-form pins
# These pin the two free constants of the projector, the frustum padding scalar
# (src/render-engine.cpp:828-832) and the aim-point blend with its horizon epsilon
# (:920-953), by facts that a wrong constant or a swapped mix violates. Each pin is
# also run against the oracle with that mistake planted, and must fail there.
AIM_EPS = 0.001  # render-engine.cpp:921
PAD_PROBE = 1.19  # NDC magnitude inside the padded frustum (1.2) and outside the unpadded one


def quad_margin(corners, pts):
    """Signed distance (m) of xz points inside the convex corner quad BL, BR, TR, TL:
    positive inside."""
    poly = corners[[0, 2, 3, 1]][:, [0, 2]]
    area = sum(cross2(poly[i], poly[(i + 1) % 4]) for i in range(4))
    sgn = 1.0 if area > 0 else -1.0
    margin = np.full(len(pts), np.inf)
    for i in range(4):
        a, b = poly[i], poly[(i + 1) % 4]
        e = b - a
        d = pts - a
        margin = np.minimum(margin, sgn * (e[0] * d[:, 1] - e[1] * d[:, 0]) / np.linalg.norm(e))
    return margin


def padded_frustum_hits(c, lim=PAD_PROBE, n=49):
    """Where camera rays through the NDC border |x| = lim or |y| = lim meet y = 0 between
    the near and the far plane (xz points)."""
    m = orc.camera(c)
    inv = np.linalg.inv(m["proj"] @ m["view"])
    ts = np.linspace(-lim, lim, n)
    ndc = [(sx * lim, t) for sx in (-1, 1) for t in ts] + [(t, sy * lim) for sy in (-1, 1) for t in ts]
    out = []
    for x, y in ndc:
        a = inv @ np.array([x, y, -1.0, 1.0])
        b = inv @ np.array([x, y, 1.0, 1.0])
        a, b = a[:3] / a[3], b[:3] / b[3]
        if (a[1] > 0) != (b[1] > 0):
            p = a + (a[1] / (a[1] - b[1])) * (b - a)
            out.append((p[0], p[2]))
    return np.array(out)


PIN_CAMERAS = [
    cam(),                                                                   # the benchmark camera
    cam(height=4.0, pitch=0.0, yaw=-90.0, fov=72.0, aspect=1.0, far=100.0, z=70.0),  # reference default (:52)
    cam(height=1.0, pitch=-10.0, yaw=30.0),                                  # projector raised above the slab
    cam(height=25.0, pitch=-60.0, yaw=135.0, fov=45.0),                      # steep
    cam(height=10.0, pitch=5.0, yaw=-20.0, far=500.0),                       # looking up
    cam(height=40.0, pitch=-30.0, yaw=250.0, fov=90.0, aspect=2.0, far=300.0),
]


@pytest.mark.parametrize("ci", range(len(PIN_CAMERAS)))
def test_padded_frustum_lies_inside_the_corner_quad(ci):
    """Every point of y = 0 that the padded frustum sees lies inside the corner quad: the
    range window bounds the padded frustum's slab points, so rays through NDC +-1.19 (inside
    the x1.2 padding) meet the plane inside the grid. Without the padding, or with 1.15,
    some of them fall outside."""
    c = PIN_CAMERAS[ci]
    pts = padded_frustum_hits(c)
    assert len(pts) >= 40
    st, corners, _ = orc.projector(c, 2.0)
    assert st == 0
    assert quad_margin(corners, pts).min() > 0.0
    for wrong in (1.0, 1.15):
        st, corners_w, _ = orc.projector(c, 2.0, pad=wrong)
        assert st == 0 and quad_margin(corners_w, pts).min() < 0.0, wrong


def aim_closed_form(c, amp):
    """Projector pitch and aim point in closed form for a camera above the slab's base
    plane: forward f = (cos p cos y, sin p, cos p sin y), a = max(|sin p|, eps).
    aim_1 meets y = 0 along f with its y clamped to +-a (mirrored when looking up), at
    horizontal distance h cos p / a; aim_2 is f's horizontal part (length cos p). The
    blend a aim_1 + (1 - a) aim_2 therefore sits (1 - a) cos p + h cos p ahead, and the
    projector, at height max(h, amp + 1), pitches down by atan(height / distance)."""
    h, yaw, p = c[5], math.radians(c[7]), math.radians(c[8])
    a = max(abs(math.sin(p)), AIM_EPS)
    dist = (1.0 - a) * math.cos(p) + h * math.cos(p)
    hp = max(h, amp + 1.0)
    aim = np.array([c[4] + dist * math.cos(yaw), 0.0, c[6] + dist * math.sin(yaw)])
    return -math.degrees(math.atan2(hp, dist)), aim


@pytest.mark.parametrize("h,pitch,yaw", [(10.0, 0.0, -90.0), (4.0, 0.0, 45.0), (10.0, -15.0, 30.0),
                                         (1.0, -10.0, 200.0), (25.0, -60.0, 10.0), (10.0, 5.0, 77.0),
                                         (6.0, 0.005, 0.0), (6.0, -0.005, 123.0)])
def test_projector_aim_closed_form(h, pitch, yaw):
    """At pitch 0 the aim point is 1 - eps + h ahead and the projector pitch is
    -atan(h_p / (h + 1 - eps)); other pitches, the raised projector and the clamped
    horizon case follow the same closed form. Swapping the mix's operands (the aim then
    sits h / eps ahead at the horizon) fails it."""
    amp = 2.0
    c = cam(height=h, pitch=pitch, yaw=yaw)
    want_pitch, want_aim = aim_closed_form(c, amp)
    st, _, info = orc.projector(c, amp)
    assert st == 0
    assert abs(info["projector_pitch"] - want_pitch) < 1e-9
    assert np.abs(info["aim"] - want_aim).max() < 1e-9 * max(1.0, np.abs(want_aim).max())
    assert abs(info["projector_pos"][1] - max(c[5], amp + 1.0)) < 1e-12
    st, _, bad = orc.projector(c, amp, swap_aim_mix=True)
    assert st == 0 and abs(bad["projector_pitch"] - want_pitch) > 1.0


def test_horizon_aim_distance_at_pitch_zero():
    """The pitch-0 case written out: h = 10 m, eps = 0.001: the aim point is 10.999 m
    ahead and the projector looks down atan(10 / 10.999)."""
    st, _, info = orc.projector(cam(height=10.0, pitch=0.0, yaw=-90.0), 2.0)
    assert st == 0
    assert abs(info["aim"][2] - (-(1.0 - AIM_EPS + 10.0))) < 1e-9 and abs(info["aim"][0]) < 1e-9
    assert abs(info["projector_pitch"] + math.degrees(math.atan(10.0 / 10.999))) < 1e-9
