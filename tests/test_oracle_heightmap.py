"""Pins the oracle's heightmap sampler (water-grid.vert:93-111, sampler state
src/texture.cpp:54-72) with GL wrap tables, symmetries and the textbook bilinear form."""
import numpy as np
import pytest

import oracle as orc


def test_repeat_wrap_table():
    for n in (1, 3, 4, 7):
        for i in range(-3 * n - 2, 3 * n + 2):
            assert orc.wrap_index(i, n, orc.WRAP_REPEAT) == i % n


def test_mirrored_repeat_wrap_table():
    # GL MIRRORED_REPEAT on a size-4 axis: 0 1 2 3 3 2 1 0 0 1 2 3 ... in both directions
    seq = [0, 1, 2, 3, 3, 2, 1, 0]
    for i in range(-16, 24):
        assert orc.wrap_index(i, 4, orc.WRAP_MIRRORED_REPEAT) == seq[i % 8], i


def brute_bilinear(tex, s, t, wrap):
    """Textbook GL_LINEAR: weights of the 2x2 texels around (s W - 1/2, t H - 1/2)."""
    H, W = tex.shape
    u, v = s * W - 0.5, t * H - 0.5
    i0, j0 = int(np.floor(u)), int(np.floor(v))
    a, b = u - i0, v - j0
    idx = lambda i, n: orc.wrap_index(i, n, wrap)  # noqa: E731
    val = 0.0
    for di, wi in ((0, 1 - a), (1, a)):
        for dj, wj in ((0, 1 - b), (1, b)):
            val += wi * wj * float(tex[idx(j0 + dj, H), idx(i0 + di, W)])
    return val


@pytest.mark.parametrize("wrap", [orc.WRAP_REPEAT, orc.WRAP_MIRRORED_REPEAT])
def test_bilinear_against_textbook(wrap):
    tex = np.random.default_rng(0).random((5, 7)).astype(np.float32)
    for s, t in np.random.default_rng(1).uniform(-3, 3, size=(300, 2)):
        v, _, _ = orc.sample_bilinear(tex, orc.TEX_R32F, wrap, s, t)
        assert abs(v - brute_bilinear(tex, s, t, wrap)) < 1e-12


def test_texel_centres_exact_and_constant_texture():
    tex = np.random.default_rng(2).random((4, 6)).astype(np.float32)
    for j in range(4):
        for i in range(6):
            v, _, _ = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_REPEAT, (i + 0.5) / 6, (j + 0.5) / 4)
            assert v == float(tex[j, i])
    c = np.full((3, 3), 0.625, np.float32)
    for s, t in np.random.default_rng(3).uniform(-5, 5, size=(50, 2)):
        v, ds, dt = orc.sample_bilinear(c, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, s, t)
        assert v == 0.625 and ds == 0.0 and dt == 0.0


def test_wrap_symmetries():
    tex = np.random.default_rng(4).random((8, 8)).astype(np.float32)
    for s, t in np.random.default_rng(5).uniform(-2, 2, size=(100, 2)):
        a = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, s, t)[0]
        b = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, -s, t)[0]
        c = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, 1 + s, t)[0]
        d = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, 1 - s, t)[0]
        assert abs(a - b) < 1e-12 and abs(c - d) < 1e-12
        r1 = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_REPEAT, s, t)[0]
        r2 = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_REPEAT, s + 3, t - 2)[0]
        assert abs(r1 - r2) < 1e-12


def test_r8_unorm_is_byte_over_255():
    tex = np.array([[0, 51], [204, 255]], np.uint8)
    v, _, _ = orc.sample_bilinear(tex, orc.TEX_R8_UNORM, orc.WRAP_REPEAT, 0.75, 0.25)
    assert abs(v - 51 / 255) < 1e-15


def test_derivatives_match_differences_inside_a_cell():
    tex = np.random.default_rng(6).random((9, 11)).astype(np.float32)
    for s, t in np.random.default_rng(7).uniform(-1, 2, size=(100, 2)):
        u, w = s * 11 - 0.5, t * 9 - 0.5
        if min(abs(u - round(u)), abs(w - round(w))) < 1e-4:
            continue
        _, ds, dt = orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, s, t)
        h = 1e-7
        fs = (orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, s + h, t)[0] -
              orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, s - h, t)[0]) / (2 * h)
        ft = (orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, s, t + h)[0] -
              orc.sample_bilinear(tex, orc.TEX_R32F, orc.WRAP_MIRRORED_REPEAT, s, t - h)[0]) / (2 * h)
        assert abs(ds - fs) < 1e-5 * max(1, abs(ds)) and abs(dt - ft) < 1e-5 * max(1, abs(dt))


def test_sample_point_is_scaled_x_minus_z():
    """water-grid.vert:103: texture(hm, s_h * (x, -z)) of the displaced point."""
    tex = np.random.default_rng(8).random((6, 6)).astype(np.float32)
    corners = np.array([[1.3, 0, 2.7, 1]] * 4)
    sd = orc.SceneData(corners, (0, 100, 0), 2, np.zeros((0, 8), np.float32), 0.0, sample_scale=0.11,
                       displacement_scale=0.5, texels=tex)
    y = orc.vertices(sd, count=1)["pos"][0, 1]
    r = brute_bilinear(tex, 0.11 * 1.3, 0.11 * -2.7, orc.WRAP_REPEAT)
    assert abs(y - 0.5 * 2 * (r - 0.5)) < 1e-12
