"""Seeded scene generators for the water-displacement workloads.

All randomness comes from SplitMix64 with u = (x >> 11) * 2^-53, so every draw is
reproducible across platforms (std / numpy distributions are not). The parameter
mapping onto the reference's wave struct (src/render-engine.h:130-167,
assets/shaders/water-grid.vert:45-51):
  frequency_w        w   = 2 pi / L                  (wavenumber)
  phaseConstant_phi  phi = sqrt(g 2 pi / L)          (deep-water angular speed,
                                                      the reference's phi = S w)
  phaseOffset_phi0   phi0 ~ U[0, 2 pi)
  steepness_Q_i      Q_i = q / (w A Count)           (render-engine.cpp:1073,
                                                      Count = normaliser)
so the kernel phase w D.x + phi t + phi0 is the reference's w D.x + phi t plus the
drawn offset.
"""
import math

import numpy as np

MASK = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15
G = 9.81


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK

    def next_u64(self):
        self.state = (self.state + GAMMA) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo

    def normal(self):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def splitmix64_array(seed, n):
    """The first n outputs of SplitMix64(seed), vectorised (uint64 wraps mod 2^64)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK) + i * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_array(seed, n):
    return (splitmix64_array(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def draw_wave_params(rng, W, wind_deg=0.0):
    """Per wave (L, A, angle_rad, phi0, q) drawn as: L ~ U[1,60] m, A = L / U[60,120],
    angle = wind + U[-45, 45] deg, phi0 ~ U[0, 2 pi), q ~ U[0, 1]."""
    out = []
    for _ in range(W):
        L = rng.uniform(1.0, 60.0)
        A = L / rng.uniform(60.0, 120.0)
        ang = math.radians(wind_deg + rng.uniform(-45.0, 45.0))
        phi0 = rng.uniform(0.0, 2.0 * math.pi)
        q = rng.uniform(0.0, 1.0)
        out.append((L, A, ang, phi0, q))
    return out


def waves_from_params(params, normaliser):
    """float32 array (W, 8) in wt_gerstner_wave layout {A, w, phi, Q_i, Dx, Dz, phi0, 0}."""
    W = len(params)
    arr = np.zeros((W, 8), dtype=np.float32)
    for i, (L, A, ang, phi0, q) in enumerate(params):
        w = np.float32(2.0 * math.pi / L)
        a = np.float32(A)
        arr[i, 0] = a
        arr[i, 1] = w
        arr[i, 2] = np.float32(math.sqrt(G * 2.0 * math.pi / L))
        wa = np.float32(w * a)
        # render-engine.cpp:1073, evaluated in FP32
        arr[i, 3] = np.float32(np.float32(q) / np.float32(wa * np.float32(normaliser))) if wa != 0 else np.float32(0)
        arr[i, 4] = np.float32(math.cos(ang))
        arr[i, 5] = np.float32(math.sin(ang))
        arr[i, 6] = np.float32(phi0)
    return arr


def gen_waves(seed, W, normaliser=None, wind_deg=0.0):
    rng = SplitMix64(seed)
    return waves_from_params(draw_wave_params(rng, W, wind_deg), W if normaliser is None else normaliser)


def value_noise(seed, size=512, octaves=8, base_cells=4):
    """Tiling value noise in [0, 1]: octave o has base_cells * 2^o lattice cells per
    side (period = size texels), lattice values U[0,1) from SplitMix64, smoothstep
    interpolation at texel centres, amplitudes 2^-o normalised to sum 1."""
    out = np.zeros((size, size), dtype=np.float64)
    total = 0.0
    centres = (np.arange(size) + 0.5) / size
    for o in range(octaves):
        cells = base_cells * (1 << o)
        lat = uniform_array(seed * 1000003 + o, cells * cells).reshape(cells, cells)
        f = centres * cells
        i0 = np.floor(f).astype(np.int64) % cells
        i1 = (i0 + 1) % cells
        t = f - np.floor(f)
        t = t * t * (3.0 - 2.0 * t)
        # rows (j) then columns (i)
        a = lat[np.ix_(i0, i0)] * (1 - t)[None, :] + lat[np.ix_(i0, i1)] * t[None, :]
        b = lat[np.ix_(i1, i0)] * (1 - t)[None, :] + lat[np.ix_(i1, i1)] * t[None, :]
        v = a * (1 - t)[:, None] + b * t[:, None]
        amp = 0.5 ** o
        out += amp * v
        total += amp
    return np.clip(out / total, 0.0, 1.0).astype(np.float32)


# --------------------------------------------------------------------- workloads
def camera_tuple(height=10.0, pitch_deg=-15.0, yaw_deg=-90.0, fov_deg=60.0, aspect=16.0 / 9.0, near=0.1, far=1000.0,
                 x=0.0, z=0.0):
    """(fov, aspect, near, far, px, py, pz, yaw, pitch), float32-representable values."""
    f = lambda v: float(np.float32(v))  # noqa: E731
    return tuple(f(v) for v in (fov_deg, aspect, near, far, x, height, z, yaw_deg, pitch_deg))


DISPLACEMENT_BOUND = 2.0  # projector slab half-height of the plain-grid scenes


def scene_256_w4():
    """256 x 256 projected grid, 4 seeded waves, one frame at t = 10 s, with tangents and
    bitangents (the size a CPU finishes in seconds)."""
    return dict(N=256, W=4, seed=0, frames=1, t0=10.0, dt=0.0, heightmap=None, camera=camera_tuple(),
                sample_scale=0.0, displacement_scale=0.0, scroll=(0.0, 0.0), amp=DISPLACEMENT_BOUND,
                with_tb=True)


def scene_1024_w64_heightmap():
    """1024 x 1024 projected grid, 64 seeded waves and a 512 x 512 value-noise heightmap,
    240 frames at t = 10 + k/60 s.

    Heightmap: R32F value noise x 0.05 m, scrolled 0.3 m/s along the wind (angle 0 -> +x),
    sampled with REPEAT at s_h = 0.02 (the reference default, render-engine.h:225): one
    texture tile spans 50 m."""
    return dict(N=1024, W=64, seed=1, frames=240, t0=10.0, dt=1.0 / 60.0, heightmap=dict(seed=1, size=512),
                camera=camera_tuple(), sample_scale=0.02, displacement_scale=0.05, scroll=(0.3, 0.0),
                amp=DISPLACEMENT_BOUND, with_tb=False)


def batched_patches(seed=2, count=512):
    """A batch of patches (512 by default): per patch grid side log-uniform in [32, 1024], W ~ U{4..128}, camera
    height U[2,50] m, pitch U[-35,-5] deg, yaw U[0,360) deg, t ~ U[0, 3600) s."""
    rng = SplitMix64(seed)
    out = []
    for p in range(count):
        N = int(round(math.exp(rng.uniform(math.log(32.0), math.log(1024.0)))))
        W = rng.randint(4, 128)
        h = rng.uniform(2.0, 50.0)
        pitch = rng.uniform(-35.0, -5.0)
        yaw = rng.uniform(0.0, 360.0)
        t = rng.uniform(0.0, 3600.0)
        out.append(dict(N=N, W=W, camera=camera_tuple(height=h, pitch_deg=pitch, yaw_deg=yaw), time=t,
                        wave_seed=seed * 100003 + p))
    return out


def fit_problem(seed, n, S, W_hidden, W_fit, spacing=0.5, dt=0.25, noise=0.01):
    """One fit problem: hidden waves (drawn as for the plain scenes, normalised by their count), target
    displacement on an n x n world grid (centred on the origin) at S snapshots
    t = s * dt, plus N(0, noise) on every component; a perturbed initial guess
    (every parameter x U[0.8, 1.2], direction + U[-15, 15] deg).
    Parameters per wave: (A, L, alpha, Q, phi0)."""
    rng = SplitMix64(seed)
    hid = draw_wave_params(rng, W_hidden)
    hw = waves_from_params(hid, W_hidden)
    hidden = np.array([(float(hw[i, 0]), L, ang, float(hw[i, 3]), phi0)
                       for i, (L, A, ang, phi0, q) in enumerate(hid)], dtype=np.float64)
    fit_src = hidden[:W_fit] if W_fit <= W_hidden else np.concatenate(
        [hidden, hidden[: W_fit - W_hidden]], axis=0)
    init = fit_src.copy()
    for i in range(init.shape[0]):
        for k in (0, 1, 3, 4):
            init[i, k] *= rng.uniform(0.8, 1.2)
        init[i, 2] += math.radians(rng.uniform(-15.0, 15.0))
    x0 = -0.5 * (n - 1) * spacing
    return dict(n=n, S=S, spacing=spacing, x0=x0, z0=x0, t0=0.0, dt=dt, noise=noise, hidden=hidden,
                init=init, noise_seed=seed * 7919 + 1)


def fit_batch_specs(seed=4, count=256):
    """A batch of fit problems (256 by default): per problem grid side log-uniform in [64, 512] (0.5 m spacing), hidden and
    fitted wave count U{8..64}, snapshots U{4..32} at 0.25 s, noise 0.01 m."""
    rng = SplitMix64(seed)
    out = []
    for p in range(count):
        n = int(round(math.exp(rng.uniform(math.log(64.0), math.log(512.0)))))
        W = rng.randint(8, 64)
        S = rng.randint(4, 32)
        out.append(dict(n=n, W=W, S=S, seed=seed * 100003 + p))
    return out


def fit_target(prob, chunk=1 << 16):
    """Target displacement [S][n*n][3] float32 of the hidden waves plus noise, in FP64."""
    n, S = prob["n"], prob["S"]
    V = n * n
    h = prob["hidden"]
    A, L, al, Q, ph0 = (h[:, k] for k in range(5))
    w = 2.0 * math.pi / L
    om = np.sqrt(G * w)
    dx, dz = np.cos(al), np.sin(al)
    idx = np.arange(V)
    xs = prob["x0"] + (idx % n) * prob["spacing"]
    zs = prob["z0"] + (idx // n) * prob["spacing"]
    out = np.zeros((S, V, 3), dtype=np.float32)
    for s in range(S):
        t = prob["t0"] + s * prob["dt"]
        for lo in range(0, V, chunk):
            x = xs[lo:lo + chunk, None]
            z = zs[lo:lo + chunk, None]
            th = w[None, :] * (dx[None, :] * x + dz[None, :] * z) + om[None, :] * t + ph0[None, :]
            c, sn = np.cos(th), np.sin(th)
            out[s, lo:lo + chunk, 0] = (c * (Q * A * dx)[None, :]).sum(1)
            out[s, lo:lo + chunk, 1] = (sn * A[None, :]).sum(1)
            out[s, lo:lo + chunk, 2] = (c * (Q * A * dz)[None, :]).sum(1)
    if prob["noise"] > 0:
        u1 = np.maximum(uniform_array(prob["noise_seed"], S * V * 3), 1e-300)
        u2 = uniform_array(prob["noise_seed"] + 1, S * V * 3)
        nz = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
        out += (prob["noise"] * nz).reshape(S, V, 3).astype(np.float32)
    return out


# --------------------------------------------------------------------- fragment stage
SHADE_TIME_OF_DAY_H = 9.0  # timeOfDayInHours default (src/render-engine.h:233)


def shade_uniforms(width, height, time_of_day_h=SHADE_TIME_OF_DAY_H):
    """The fragment shader's uniforms with the reference's defaults
    (src/render-engine.h:213-235): sun from the time of day as at
    src/render-engine.cpp:289-291, fog colour at :295 (all in float32)."""
    f32 = np.float32
    theta = f32(f32(time_of_day_h) / f32(24.0)) * f32(2.0 * math.pi) - f32(0.5 * math.pi)
    sun = [float(f32(np.cos(theta))), float(f32(np.sin(theta))), 0.0]
    k = float(f32(min(max(sun[1], 0.0), 1.0)))
    return {"fogColourFarAtCurrentTime": [k, k, k, 0.1], "fogDepthRadiusFar": 1.0, "fogDepthRadiusNear": 0.0,
            "softEdgesDeltaDepthThreshold": 0.05, "sunPosition": sun, "sunShininess": 50.0, "sunStrength": 1.0,
            "tintDeltaDepthThreshold": 0.1, "viewportWidthHeight": [float(width), float(height)],
            "waterClarity": 0.3, "zFar": 100.0, "zNear": 0.1}


def cube_texel_directions(length, face=None):
    """(6, L, L, 3) directions through the texel centres of a cube map's faces in GL
    target order (or (L, L, 3) for one face), inverting the face table of the GL spec
    (table 8.19): +X (1, -tc, -sc), -X (-1, -tc, sc), +Y (sc, 1, tc), -Y (sc, -1, -tc),
    +Z (sc, -tc, 1), -Z (-sc, -tc, -1), with sc = 2 s - 1 along a row, tc = 2 t - 1."""
    c = (np.arange(length) + 0.5) / length * 2.0 - 1.0
    sc = np.broadcast_to(c[None, :], (length, length))
    tc = np.broadcast_to(c[:, None], (length, length))
    one = np.ones((length, length))
    table = [lambda: (one, -tc, -sc), lambda: (-one, -tc, sc), lambda: (sc, one, tc), lambda: (sc, -one, -tc),
             lambda: (sc, -tc, one), lambda: (-sc, -tc, -one)]
    if face is not None:
        return np.stack(table[face](), axis=-1)
    return np.stack([np.stack(f(), axis=-1) for f in table])


def _noise_rgba(seed, h, w, alpha):
    """Smooth RGBA8 image: value noise per channel (bilinearly resampled), alpha from
    `alpha` (a constant, or 'mask' for a half-covered pattern of 0 and 255 regions)."""
    out = np.empty((h, w, 4), dtype=np.uint8)
    size = 64
    for ch in range(4 if alpha == "noise" else 3):
        base = value_noise(seed * 7 + ch, size=size, octaves=4, base_cells=2).astype(np.float64)
        ys = (np.arange(h) + 0.5) / h * size - 0.5
        xs = (np.arange(w) + 0.5) / w * size - 0.5
        y0 = np.floor(ys).astype(np.int64)
        x0 = np.floor(xs).astype(np.int64)
        fy, fx = ys - y0, xs - x0
        y0m, y1m, x0m, x1m = y0 % size, (y0 + 1) % size, x0 % size, (x0 + 1) % size
        top = base[np.ix_(y0m, x0m)] * (1 - fx) + base[np.ix_(y0m, x1m)] * fx
        bot = base[np.ix_(y1m, x0m)] * (1 - fx) + base[np.ix_(y1m, x1m)] * fx
        v = top * (1 - fy)[:, None] + bot * fy[:, None]
        out[..., ch] = np.clip(np.rint(v * 255.0), 0, 255).astype(np.uint8)
    if alpha == "mask":
        out[..., 3] = np.where(out[..., 0] > 128, 255, 0).astype(np.uint8)
    elif alpha != "noise":
        out[..., 3] = int(alpha)
    return out


def shading_scene(width, height, seed=0, sky_length=256, local_scale=1.0, eye_height=6.0, pitch_deg=-12.0,
                  yaw_deg=30.0, fov_deg=60.0):
    """A synthetic water frame for the fragment stage: the fragment buffer a rasteriser
    would produce for a rippled sea seen from a camera `eye_height` m up, plus the
    textures the shader samples.

    Per pixel the camera ray meets the plane y = 0 at P; the surface height and normal
    come from a sum of sines; `normal` is shortened by up to 2% (interpolation);
    viewVecRaw = eye - P; the yaw-only view normal is the normal rotated by -yaw about y;
    xyPositionNDCSpaceHeight0 projects (P.x, 0, P.z); gl_FragCoord.z is P's window depth.
    Pixels whose ray misses the sea or runs past zFar are not covered. The depth texture
    holds a sea bed 1-6 m down (1.0 = sky where no bed is hit), reflections / refractions
    are smooth RGBA8 noise at local_scale x the viewport, the skybox a smooth gradient of
    the direction plus noise."""
    u = shade_uniforms(width, height)
    zf, zn = u["zFar"], u["zNear"]
    aspect = width / height
    th = math.tan(math.radians(fov_deg) / 2.0)
    yaw, pitch = math.radians(yaw_deg), math.radians(pitch_deg)
    fwd = np.array([math.cos(pitch) * math.sin(yaw), math.sin(pitch), -math.cos(pitch) * math.cos(yaw)])
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    eye = np.array([0.0, eye_height, 0.0])
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = right, up, -fwd
    view[:3, 3] = -view[:3, :3] @ eye
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1] = 1.0 / (aspect * th), 1.0 / th
    proj[2, 2], proj[2, 3], proj[3, 2] = -(zf + zn) / (zf - zn), -2.0 * zf * zn / (zf - zn), -1.0
    vp = proj @ view

    def window_depth(dist_along_fwd):
        z_ndc = (zf + zn) / (zf - zn) - 2.0 * zf * zn / ((zf - zn) * dist_along_fwd)
        return 0.5 * z_ndc + 0.5

    rng = SplitMix64(seed)
    ks = [(rng.uniform(0.2, 1.5), rng.uniform(0.0, 2.0 * math.pi), rng.uniform(0.0, 2.0 * math.pi),
           rng.uniform(0.02, 0.08)) for _ in range(6)]
    frags = np.zeros((height, width, 12), dtype=np.float32)
    cover = np.zeros((height, width), dtype=np.uint8)
    depth = np.ones((height, width), dtype=np.float32)
    xs = ((np.arange(width) + 0.5) / width) * 2.0 - 1.0
    shorten = uniform_array(seed * 31 + 5, width * height).reshape(height, width)
    for row in range(height):
        y = (row + 0.5) / height * 2.0 - 1.0
        d = fwd[None, :] + (xs[:, None] * th * aspect) * right[None, :] + (y * th) * up[None, :]
        down = d[:, 1] < -1e-6
        t = np.where(down, eye_height / np.where(down, -d[:, 1], 1.0), np.inf)
        P = eye[None, :] + t[:, None] * d
        along = t * (d @ fwd)
        ok = down & (along < zf * 0.999) & (along > zn)
        px, pz = np.where(ok, P[:, 0], 0.0), np.where(ok, P[:, 2], 0.0)
        hx = np.zeros(width)
        hz = np.zeros(width)
        hh = np.zeros(width)
        for k, ang, ph, amp in ks:
            arg = k * (math.cos(ang) * px + math.sin(ang) * pz) + ph
            hh += amp * np.sin(arg)
            hx += amp * k * math.cos(ang) * np.cos(arg)
            hz += amp * k * math.sin(ang) * np.cos(arg)
        n = np.stack([-hx, np.ones(width), -hz], axis=1)
        n /= np.linalg.norm(n, axis=1)[:, None]
        n *= (1.0 - 0.02 * shorten[row])[:, None]
        Pw = np.stack([px, hh, pz], axis=1)
        vv = eye[None, :] - Pw
        cy, sy = math.cos(yaw), math.sin(yaw)
        ny = np.stack([cy * n[:, 0] - sy * n[:, 2], n[:, 1], sy * n[:, 0] + cy * n[:, 2]], axis=1)
        ny /= np.linalg.norm(ny, axis=1)[:, None]
        h0 = np.stack([px, np.zeros(width), pz, np.ones(width)], axis=1) @ vp.T
        ndc0 = h0[:, :2] / np.where(ok, h0[:, 3], 1.0)[:, None]
        fz = window_depth(np.where(ok, along, 1.0))
        rec = np.concatenate([n, vv, ny, ndc0, fz[:, None]], axis=1)
        frags[row] = np.where(ok[:, None], rec, 0.0).astype(np.float32)
        cover[row] = ok
        bed = 1.0 + 5.0 * (0.5 + 0.5 * np.sin(0.05 * px + 0.3) * np.cos(0.07 * pz))
        tb = np.where(down, (eye_height + bed) / np.where(down, -d[:, 1], 1.0), np.inf)
        bed_along = tb * (d @ fwd)
        hit = down & (bed_along < zf)
        depth[row] = np.where(hit, window_depth(np.where(hit, bed_along, 1.0)), 1.0).astype(np.float32)
    lw, lh = max(1, int(round(width * local_scale))), max(1, int(round(height * local_scale)))
    refl = _noise_rgba(seed * 3 + 1, lh, lw, "mask")
    refr = _noise_rgba(seed * 3 + 2, lh, lw, 255)
    sky = np.empty((6, sky_length, sky_length, 4), dtype=np.uint8)
    for f in range(6):
        dirs = cube_texel_directions(sky_length, face=f)
        dirs = dirs / np.linalg.norm(dirs, axis=-1, keepdims=True)
        grain = uniform_array(seed * 3 + 3 + 1000 * f, sky_length * sky_length).reshape(sky_length, sky_length)
        sky[f, ..., 0] = np.clip(np.rint((0.35 + 0.3 * dirs[..., 0] + 0.02 * grain) * 255), 0, 255)
        sky[f, ..., 1] = np.clip(np.rint((0.55 + 0.3 * dirs[..., 1] + 0.02 * grain) * 255), 0, 255)
        sky[f, ..., 2] = np.clip(np.rint((0.75 + 0.2 * dirs[..., 2] + 0.02 * grain) * 255), 0, 255)
    sky[..., 3] = 255
    return dict(uniforms=u, fragments=frags, coverage=cover, depth=depth, reflections=refl, refractions=refr,
                skybox=sky, view_projection=vp)
