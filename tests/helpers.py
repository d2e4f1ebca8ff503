"""Shared helpers for the parity tests: scene set-up and the tolerance checks."""
import math

import numpy as np

import oracle as orc
import wavetool_amd as wt

POS_TOL = 1e-4   # metres, within RADIUS of the camera
NRM_TOL = 1e-4   # radians
RADIUS = 1000.0  # metres
EDGE_EPS = 2e-3  # texel units: heightmap samples this close to a texel edge are left out of the normal check
FLIP_EPS = 1e-5  # |cos| between the oracle normal and the view vector below which the
                 # camera-facing flip (water-grid.vert:190) is ill-conditioned


def oracle_scene(corners, cam_pos, N, waves, time, sample_scale=0.0, displacement_scale=0.0, bounce=0.0,
                 scroll=(0.0, 0.0), texels=None, tex_format=orc.TEX_R32F, tex_wrap=orc.WRAP_REPEAT):
    return orc.SceneData(corners, cam_pos, N, waves, time, sample_scale, displacement_scale, bounce, scroll,
                         texels, tex_format, tex_wrap)


def angle_between(a, b):
    a = a / np.linalg.norm(a, axis=1, keepdims=True)
    b = b / np.linalg.norm(b, axis=1, keepdims=True)
    c = np.clip((a * b).sum(1), -1.0, 1.0)
    # atan2 form is accurate for tiny angles
    cr = np.linalg.norm(np.cross(a, b), axis=1)
    return np.arctan2(cr, c)


def near_texel_edge(pos_ref, sample_scale, scroll, time, tex_w, tex_h, eps=EDGE_EPS):
    s = sample_scale * (pos_ref[:, 0] - scroll[0] * time)
    t = sample_scale * -(pos_ref[:, 2] - scroll[1] * time)
    u = s * tex_w - 0.5
    v = t * tex_h - 0.5
    du = np.abs(u - np.round(u))
    dv = np.abs(v - np.round(v))
    return (du < eps) | (dv < eps)


def flip_ambiguous(ref, cam_pos, eps=FLIP_EPS):
    """Vertices whose normal is within eps of perpendicular to the view vector: the sign
    chosen by the camera-facing flip depends on the last bits there."""
    view = np.asarray(cam_pos, dtype=np.float64)[None, :] - ref["pos"]
    view /= np.linalg.norm(view, axis=1, keepdims=True)
    n = ref["nrm"] / np.linalg.norm(ref["nrm"], axis=1, keepdims=True)
    return np.abs((n * view).sum(1)) < eps


def compare(out, ref, cam_pos, edge_mask=None, radius=RADIUS):
    """Max position error (m) and normal angle error (rad) within radius of the camera.
    Normals at heightmap texel edges (edge_mask) and at flip-ambiguous vertices are
    excluded from the normal check and counted in n_excluded / n_flip."""
    d = np.linalg.norm(ref["pos"] - np.asarray(cam_pos)[None, :], axis=1)
    near = d <= radius
    perr = np.linalg.norm(out["pos"].astype(np.float64) - ref["pos"], axis=1)
    ang = angle_between(out["nrm"].astype(np.float64), ref["nrm"])
    flip = flip_ambiguous(ref, cam_pos)
    nmask = near & ~flip
    if edge_mask is not None:
        nmask &= ~edge_mask
    res = dict(pos_max=float(perr[near].max()) if near.any() else 0.0,
               nrm_max=float(ang[nmask].max()) if nmask.any() else 0.0,
               n_near=int(near.sum()), n_excluded=int((near & edge_mask).sum()) if edge_mask is not None else 0,
               n_flip=int((near & flip).sum()))
    for k in ("tan", "bit"):
        if out.get(k) is not None and ref.get(k) is not None:
            res[k + "_max"] = float(angle_between(out[k].astype(np.float64), ref[k])[nmask].max())
    return res


def reference_indices(n):
    """src/program.cpp:686-717 restated: ids row-major from the bottom-left, per cell
    (r,c), (r+1,c+1), (r+1,c) then (r,c), (r,c+1), (r+1,c+1)."""
    ids = np.arange(n * n).reshape(n, n)
    out = []
    for r in range(n - 1):
        for c in range(n - 1):
            out += [ids[r, c], ids[r + 1, c + 1], ids[r + 1, c], ids[r, c], ids[r, c + 1], ids[r + 1, c + 1]]
    return np.array(out, dtype=np.int64)


def reference_indices_vectorized(n):
    """The same list built with array operations, for large n (pinned against the loop
    form on small grids)."""
    r, c = np.meshgrid(np.arange(n - 1, dtype=np.int64), np.arange(n - 1, dtype=np.int64), indexing="ij")
    bl = r * n + c
    tl = bl + n
    return np.stack([bl, tl + 1, tl, bl, bl + 1, tl + 1], axis=-1).reshape(-1)
