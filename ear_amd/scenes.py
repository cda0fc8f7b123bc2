"""Seeded procedural scenes for the benchmark and tests.

Every generator returns a plain dict (numpy arrays + python values) that both
the device path (ear_amd.api) and the CPU oracle can consume:
    tris     float32 [n, 3, 3]  (the merged MESH blocks, in order)
    mat      uint16  [n]
    materials  [(refl[3], refr[3] | None, spec[3] | None)]  -- .EAR MAT rows
    src      source position, recorders  list of recorder dicts
    air      per-band absorbtion_factor = 1 - SET absorption (exporter
             defaults 0.01/0.02/0.05, blender/render_EAR/__init__.py:132-134)
Absorption alpha maps to refl = 1 - alpha, no refraction, as a .EAR MAT block
with refl only (src/Material.cpp:26-72); specularity is the optional third
triple, written with a zero refraction triple in front of it.
"""
from __future__ import annotations

import math

import numpy as np

EXPORTER_AIR = (1.0 - 0.01, 1.0 - 0.02, 1.0 - 0.05)
# StereoRecorder ctor stores max(0, (1 - h)^4) (src/StereoRecorder.cpp:53-57);
# exporter defaults h = 0.1 / 0.3 / 0.9 (blender/render_EAR/__init__.py:88-90)
EXPORTER_HEAD_ABS = tuple(max(0.0, (1.0 - h) ** 4) for h in (0.1, 0.3, 0.9))
MASTER_SEED = 0xEA2


def _quad(a, b, c, d):
    """Two triangles (a,b,c), (a,c,d) of the planar quad a-b-c-d (CCW)."""
    return [(a, b, c), (a, c, d)]


def _grid_quad(p0, u, v, nu, nv):
    """Quad p0 + s*u + t*v, s,t in [0,1], tessellated nu x nv."""
    p0, u, v = (np.asarray(x, np.float64) for x in (p0, u, v))
    out = []
    for i in range(nu):
        for j in range(nv):
            a = p0 + u * (i / nu) + v * (j / nv)
            b = p0 + u * ((i + 1) / nu) + v * (j / nv)
            c = p0 + u * ((i + 1) / nu) + v * ((j + 1) / nv)
            d = p0 + u * (i / nu) + v * ((j + 1) / nv)
            out += _quad(a, b, c, d)
    return out


def box_walls(lx, ly, lz, origin=(0.0, 0.0, 0.0), tess=(1, 1, 1)):
    """Six walls of an axis-aligned box as lists of triangles, outward CCW.
    Order: floor, ceiling, x=0, x=lx, y=0, y=ly."""
    ox, oy, oz = origin
    tx, ty, tz = tess
    P = lambda x, y, z: (ox + x, oy + y, oz + z)  # noqa: E731
    walls = [
        _grid_quad(P(0, 0, 0), (0, ly, 0), (lx, 0, 0), ty, tx),      # floor, normal -z
        _grid_quad(P(0, 0, lz), (lx, 0, 0), (0, ly, 0), tx, ty),     # ceiling, +z
        _grid_quad(P(0, 0, 0), (0, 0, lz), (0, ly, 0), tz, ty),      # x = 0, -x
        _grid_quad(P(lx, 0, 0), (0, ly, 0), (0, 0, lz), ty, tz),     # x = lx, +x
        _grid_quad(P(0, 0, 0), (lx, 0, 0), (0, 0, lz), tx, tz),      # y = 0, -y
        _grid_quad(P(0, ly, 0), (0, 0, lz), (lx, 0, 0), tz, tx),     # y = ly, +y
    ]
    return walls


def _pack(groups):
    tris, mat = [], []
    for m, g in groups:
        tris += g
        mat += [m] * len(g)
    t = np.asarray(tris, np.float64).reshape(-1, 3, 3).astype(np.float32)
    return t, np.asarray(mat, np.uint16)


def absorption_materials(alpha, spec=None):
    """alpha [n_mat][3] -> MAT rows refl = 1 - alpha (+ zero refr and spec)."""
    mats = []
    for i, a in enumerate(alpha):
        refl = [float(1.0 - x) for x in a]
        if spec is None:
            mats.append((refl, None, None))
        else:
            s = spec[i]
            s = [float(s)] * 3 if np.ndim(s) == 0 else [float(x) for x in s]
            mats.append((refl, [0.0, 0.0, 0.0], s))
    return mats


def shoebox(lx=10.0, ly=8.0, lz=4.0, *, seed=MASTER_SEED, alpha_range=(0.05, 0.4), spec=0.5,
            src=(2.0, 3.0, 1.5), rec=(7.0, 5.0, 1.7), tess=(1, 1, 1), stereo=False):
    """The shoebox room: 10x8x4 m box, 12 triangles, one material per wall,
    alpha ~ U(0.05, 0.4) per wall and band, specularity 0.5."""
    rng = np.random.default_rng(seed)
    walls = box_walls(lx, ly, lz, tess=tess)
    tris, mat = _pack(list(enumerate(walls)))
    alpha = rng.uniform(alpha_range[0], alpha_range[1], size=(6, 3))
    mats = absorption_materials(alpha, spec=[spec] * 6 if spec is not None else None)
    recs = [stereo_recorder(rec) if stereo else dict(kind=0, pos=tuple(rec))]
    return dict(name=f"shoebox {lx}x{ly}x{lz}", tris=tris, mat=mat, materials=mats, alpha=alpha, src=tuple(src),
                recorders=recs, air=EXPORTER_AIR, volume=lx * ly * lz, area=2 * (lx * ly + lx * lz + ly * lz))


def stereo_recorder(pos, right_ear=(0.0, 1.0, 0.0), head_size=0.17, head_abs=EXPORTER_HEAD_ABS):
    return dict(kind=1, pos=tuple(pos), right_ear=tuple(right_ear), head_size=head_size, head_abs=tuple(head_abs))


def hall(*, seed=MASTER_SEED, target_tris=200_000, W=40.0, D=28.0, H=14.0, amp=1.5, n_mat=8,
         alpha_range=(0.02, 0.6), spec=None):
    """Procedural concert hall of ~target_tris triangles (200k by default).

    Box shell W x D x H; the ceiling is z = H + amp sin(3*2pi x/W) sin(2*2pi y/D)
    (zero at the walls, so the shell is closed), tessellated into a grid;
    the audience area holds seat blocks: axis-aligned extruded quads of random
    height U(0.4, 1.2) m (top + 4 sides, 10 triangles each).  8 materials:
    floor, 4 walls, ceiling, seat tops, seat sides; alpha ~ U(0.02, 0.6).
    Source on the stage, stereo recorder (head 0.17 m) above the seats."""
    rng = np.random.default_rng(seed)
    # ceiling gets ~2/3 of the budget, seats ~1/3
    ceil_tris = int(target_tris * 0.66)
    nx = int(round(math.sqrt(ceil_tris / 2 * W / D)))
    ny = max(2, int(ceil_tris / 2 / nx))
    xs = np.linspace(0.0, W, nx + 1)
    ys = np.linspace(0.0, D, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    Z = H + amp * np.sin(3 * 2 * math.pi * X / W) * np.sin(2 * 2 * math.pi * Y / D)
    Z[0, :] = H
    Z[-1, :] = H
    Z[:, 0] = H
    Z[:, -1] = H
    P = np.stack([X, Y, Z], -1)
    a, b, c, d = P[:-1, :-1], P[1:, :-1], P[1:, 1:], P[:-1, 1:]
    # downward-facing (into the room) is irrelevant: surfaces are double sided;
    # keep outward (+z) orientation for the signed volume
    ceil = np.concatenate([np.stack([a, b, c], -2).reshape(-1, 3, 3), np.stack([a, c, d], -2).reshape(-1, 3, 3)])
    walls = box_walls(W, D, H)
    groups = [(0, walls[0]), (1, walls[2]), (2, walls[3]), (3, walls[4]), (4, walls[5])]
    shell, shell_mat = _pack(groups)
    seat_budget = target_tris - len(ceil) - len(shell)
    n_blocks = max(1, seat_budget // 10)
    ax0, ax1, ay0, ay1 = 0.3 * W, 0.95 * W, 0.06 * D, 0.94 * D
    aspect = (ax1 - ax0) / (ay1 - ay0)
    bx = max(1, int(round(math.sqrt(n_blocks * aspect))))
    by = max(1, n_blocks // bx)
    px, py = (ax1 - ax0) / bx, (ay1 - ay0) / by
    gap = 0.12
    heights = rng.uniform(0.4, 1.2, size=(bx, by))
    tops, sides = [], []
    for i in range(bx):
        x0 = ax0 + i * px + gap / 2
        x1 = x0 + px - gap
        for j in range(by):
            y0 = ay0 + j * py + gap / 2
            y1 = y0 + py - gap
            h = float(heights[i, j])
            q = [(x0, y0, h), (x1, y0, h), (x1, y1, h), (x0, y1, h)]
            tops += _quad(q[0], q[1], q[2], q[3])
            sides += _quad((x0, y0, 0), (x1, y0, 0), (x1, y0, h), (x0, y0, h))
            sides += _quad((x1, y0, 0), (x1, y1, 0), (x1, y1, h), (x1, y0, h))
            sides += _quad((x1, y1, 0), (x0, y1, 0), (x0, y1, h), (x1, y1, h))
            sides += _quad((x0, y1, 0), (x0, y0, 0), (x0, y0, h), (x0, y1, h))
    st, smat = _pack([(6, tops), (7, sides)])
    tris = np.concatenate([shell, ceil.astype(np.float32), st])
    mat = np.concatenate([shell_mat, np.full(len(ceil), 5, np.uint16), smat])
    alpha = rng.uniform(alpha_range[0], alpha_range[1], size=(n_mat, 3))
    mats = absorption_materials(alpha, spec=spec)
    src = (0.12 * W, 0.5 * D, 1.5)
    rec = stereo_recorder((0.62 * W, 0.5 * D + 0.003, 1.7))  # above the tallest seat block (1.2 m)
    return dict(name=f"hall {len(tris)} tris", tris=tris, mat=mat, materials=mats, alpha=alpha, src=src,
                recorders=[rec], air=EXPORTER_AIR)


def stage_path(n_keys=64, *, W=40.0, D=28.0, z=1.5):
    """A moving source for the hall: n_keys discrete positions on a closed
    loop over the hall's stage (x < 0.3 W, clear of the seat blocks).  The
    reference has no spline: keyframe i is position frames[i]
    (src/Animated.h:89-91), as the exporter samples a curve once per frame."""
    th = 2.0 * math.pi * np.arange(n_keys) / n_keys
    x = 0.12 * W + 0.0625 * W * np.cos(th)
    y = 0.5 * D + 0.32 * D * np.sin(th)
    zz = z + 0.3 * np.sin(3.0 * th)
    return [(float(a), float(b), float(c)) for a, b, c in zip(x, y, zz)]


def room(*, seed, target_tris=12, kind=None, alpha_range=(0.02, 0.6)):
    """One room of the batching workload: a box or L-shaped room whose walls are
    tessellated to reach ~target_tris triangles; seeded source/recorder."""
    rng = np.random.default_rng(seed)
    if kind is None:
        kind = "box" if rng.random() < 0.5 else "L"
    lx, ly, lz = rng.uniform(5, 20), rng.uniform(4, 15), rng.uniform(2.5, 6)
    if kind == "box":
        per_wall_quads = max(1, target_tris // 12)
        t = max(1, int(math.sqrt(per_wall_quads)))
        walls = box_walls(lx, ly, lz, tess=(t, t, t))
        tris, mat = _pack(list(enumerate(walls)))
        src = (rng.uniform(0.2, 0.8) * lx, rng.uniform(0.2, 0.8) * ly, rng.uniform(0.3, 0.7) * lz)
        rec = (rng.uniform(0.2, 0.8) * lx, rng.uniform(0.2, 0.8) * ly, rng.uniform(0.3, 0.7) * lz)
    else:
        # L shape: box [0,lx]x[0,ly] minus the corner [cx,lx]x[cy,ly]
        cx, cy = rng.uniform(0.4, 0.7) * lx, rng.uniform(0.4, 0.7) * ly
        quads_total = max(1, target_tris // 2)
        t = max(1, int(math.sqrt(quads_total / 16)))
        floor = _grid_quad((0, 0, 0), (0, cy, 0), (lx, 0, 0), t, t) + _grid_quad((0, cy, 0), (0, ly - cy, 0),
                                                                                  (cx, 0, 0), t, t)
        ceil = _grid_quad((0, 0, lz), (lx, 0, 0), (0, cy, 0), t, t) + _grid_quad((0, cy, lz), (cx, 0, 0),
                                                                                 (0, ly - cy, 0), t, t)
        w = []
        w += _grid_quad((0, 0, 0), (lx, 0, 0), (0, 0, lz), t, t)            # y = 0
        w += _grid_quad((lx, 0, 0), (0, cy, 0), (0, 0, lz), t, t)           # x = lx (lower part)
        w += _grid_quad((lx, cy, 0), (cx - lx, 0, 0), (0, 0, lz), t, t)     # y = cy step
        w += _grid_quad((cx, cy, 0), (0, ly - cy, 0), (0, 0, lz), t, t)     # x = cx
        w += _grid_quad((cx, ly, 0), (-cx, 0, 0), (0, 0, lz), t, t)         # y = ly
        w += _grid_quad((0, ly, 0), (0, -ly, 0), (0, 0, lz), t, t)          # x = 0
        tris, mat = _pack([(0, floor), (1, ceil), (2, w)])
        src = (rng.uniform(0.1, 0.9) * cx, rng.uniform(0.1, 0.9) * ly, rng.uniform(0.3, 0.7) * lz)
        rec = (rng.uniform(0.1, 0.9) * lx, rng.uniform(0.1, 0.9) * cy, rng.uniform(0.3, 0.7) * lz)
    n_mat = int(mat.max()) + 1
    alpha = rng.uniform(alpha_range[0], alpha_range[1], size=(n_mat, 3))
    return dict(name=f"{kind} room {len(tris)} tris", tris=tris, mat=mat, materials=absorption_materials(alpha),
                alpha=alpha, src=src, recorders=[dict(kind=0, pos=rec)], air=EXPORTER_AIR)


def batch_rooms(n_rooms=256, *, seed=3, tris=(12, 50_000), rays=(1e4, 2e6)):
    """The batching workload: n_rooms box or L rooms (room(seed=1000 + i))
    of log-uniform triangle counts in `tris` and log-uniform ray counts in
    `rays`, drawn from one seeded stream.  Returns [(room, n_rays)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_rooms):
        nt = int(np.exp(rng.uniform(np.log(tris[0]), np.log(tris[1]))))
        nr = int(np.exp(rng.uniform(np.log(rays[0]), np.log(rays[1]))))
        out.append((room(seed=1000 + i, target_tris=nt), nr))
    return out


def signed_volume(tris) -> float:
    """Mesh::Volume (src/Mesh.cpp:160-167) via Triangle::SignedVolume, in float64."""
    t = np.asarray(tris, np.float64)
    return float(np.sum(np.einsum("ij,ij->i", t[:, 0], np.cross(t[:, 1], t[:, 2]))) / 6.0)


def surface_area(tris) -> float:
    t = np.asarray(tris, np.float64)
    return float(np.sum(np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 1]), axis=1)) / 2.0)
