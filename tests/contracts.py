"""Raycast, integration-band and marching-cubes fixtures re-expressed from the reference's
own CPU contracts, with the same numbers and the same hand-derived expectations.

Each check takes a callable that runs the operation under test (the CPU oracle or the
GPU path), so one fixture pins the oracle and checks the product. Colour checks of the
reference contracts are dropped: no colour is stored on this path.
"""
import math
import os
import re

import numpy as np

f32 = np.float32

# ---------------------------------------------------------------------------
# tests/tsdf_raycast_contract.cpp:63-151: 32^3 lattice at 4 cm, origin (-0.64, -0.64, 0),
# an 8x8 image with fx = fy = 4, cx = cy = 3 (pixel (3, 3) on the optical axis),
# identity pose, every voxel observed (weight 4) and holding a z-only field.
#   wall: f(z) = clamp((0.38 - z) / 0.04) -> zero crossing at z = 0.38
#   slab: solid for z in [0.39, 0.41], tau 0.02 -> front zero 0.36 + 0.04 / 1.5
# ---------------------------------------------------------------------------
RC_RES, RC_VS, RC_TRUNC = 32, 0.04, 0.025
RC_ORIGIN = (-0.64, -0.64, 0.0)
RC_K = (4.0, 4.0, 3.0, 3.0)
RC_IMG = 8
RC_NEAR, RC_FAR = 0.12, 0.90
WALL_Z = 0.38
SLAB_FRONT_Z = 0.36 + 0.04 * (1.0 / 1.5)


def _clamp1(v):
    return np.minimum(np.maximum(v, f32(-1.0)), f32(1.0))


def wall_field(z):
    return _clamp1((f32(0.38) - z) / f32(RC_VS))


def slab_field(z):
    a, b, tau = f32(0.39), f32(0.41), f32(0.02)
    inside = (z >= a) & (z <= b)
    sdf = np.where(inside, -np.minimum(z - a, b - z), np.where(z < a, a - z, z - b)).astype(f32)
    return _clamp1(sdf / tau)


def layered_volume(field, res=RC_RES, vs=RC_VS, weight=4.0):
    """(tsdf, weight) arrays [z, y, x] with layer k holding field(k * vs) (float32)."""
    z = (np.arange(res, dtype=f32) * f32(vs)).astype(f32)
    col = field(z).astype(f32)
    t = np.broadcast_to(col[:, None, None], (res, res, res)).copy()
    return t, np.full((res, res, res), f32(weight))


def check_raycast_contract(cast, tol=1e-5):
    """cast(tsdf, weight, min_depth, max_depth) -> (vertices (8, 8, 3), normals (8, 8, 3)) of
    an identity-pose raycast with RC_K. Gates A-F of tests/tsdf_raycast_contract.cpp."""
    wall = layered_volume(wall_field)

    def empty(v, n, px, py):
        return not v[py, px].any() and not n[py, px].any()

    # A: on-axis planar hit at the derived crossing, outward unit normal
    v, n = cast(*wall, RC_NEAR, RC_FAR)
    h, hn = v[3, 3].astype(np.float64), n[3, 3].astype(np.float64)
    assert np.isfinite(h).all() and np.isfinite(hn).all()
    assert abs(h[0]) <= tol and abs(h[1]) <= tol and abs(h[2] - WALL_Z) <= tol, h
    assert abs(hn[0]) <= tol and abs(hn[1]) <= tol and abs(hn[2] + 1.0) <= tol, hn
    assert abs(np.linalg.norm(hn) - 1.0) <= tol
    control_v, control_n = v[3, 3].copy(), n[3, 3].copy()

    # B: the band is a Z-depth band: pixel (7, 3) has |ray| = sqrt(2), hit t = 0.537 > 0.42
    v, n = cast(*wall, RC_NEAR, 0.42)
    h = v[3, 7].astype(np.float64)
    assert not empty(v, n, 7, 3)
    assert abs(h[2] - WALL_Z) <= tol and abs(h[0] - WALL_Z) <= tol, h
    assert np.linalg.norm(np.cross(h, [1.0, 0.0, 1.0])) <= tol

    # C: the configured near / far bounds gate
    v, n = cast(*wall, 0.39, RC_FAR)
    assert empty(v, n, 3, 3)
    v, n = cast(*wall, RC_NEAR, 0.30)
    assert empty(v, n, 3, 3)

    # D: exterior start resolves the slab's entry zero; an interior start is no surface
    slab = layered_volume(slab_field)
    v, n = cast(*slab, RC_NEAR, RC_FAR)
    assert abs(float(v[3, 3, 2]) - SLAB_FRONT_Z) <= tol, v[3, 3]
    v, n = cast(*slab, 0.402, RC_FAR)
    assert empty(v, n, 3, 3)

    # E: a non-finite sample at the crossing is rejected, one behind the surface is not
    for bad in (np.nan, np.inf, -np.inf):
        t, w = layered_volume(wall_field)
        t[10] = f32(bad)
        v, n = cast(t, w, RC_NEAR, RC_FAR)
        assert empty(v, n, 3, 3) and np.isfinite(v).all() and np.isfinite(n).all()
    t, w = layered_volume(wall_field)
    t[20] = f32(np.nan)
    v, n = cast(t, w, RC_NEAR, RC_FAR)
    assert np.array_equal(v[3, 3], control_v) and np.array_equal(n[3, 3], control_n)

    # F: every pixel hits the wall in band with a unit normal; repeats are bit-identical
    v1, n1 = cast(*wall, RC_NEAR, RC_FAR)
    assert v1.any(axis=2).all()
    assert (np.abs(v1[..., 2].astype(np.float64) - WALL_Z) <= tol).all()
    assert (np.abs(np.linalg.norm(n1.astype(np.float64), axis=2) - 1.0) <= 1e-4).all()
    v2, n2 = cast(*wall, RC_NEAR, RC_FAR)
    assert np.array_equal(v1, v2) and np.array_equal(n1, n2)


# ---------------------------------------------------------------------------
# tests/tsdf_subvoxel_thin_feature_contract.cpp:56-315: the same 32^3 / 4 cm lattice with
# slabs of tau = 2 cm, weight 1, one on-axis ray (1x1 image, fx = fy = 1, cx = cy = 0).
# The expected depth is the first zero of the piecewise-linear interpolation through
# the lattice samples, z* = (k-1) vs + vs f(k-1) / (f(k-1) - f(k)).
# ---------------------------------------------------------------------------
TF_TAU = 0.02
TF_K = (1.0, 1.0, 0.0, 0.0)


def slabs_field(slabs):
    def field(z):
        out = np.empty_like(z)
        for i, zz in enumerate(z):
            inside, best_in, best_out = False, f32(0.0), f32(np.inf)
            for a, b in slabs:
                a, b = f32(a), f32(b)
                if a <= zz <= b:
                    inside = True
                    best_in = max(best_in, min(zz - a, b - zz))
                else:
                    best_out = min(best_out, a - zz if zz < a else zz - b)
            sd = -best_in if inside else best_out
            out[i] = f32(max(f32(-1.0), min(f32(1.0), f32(sd) / f32(TF_TAU))))
        return out
    return field


def lattice_zero(slabs):
    z = (np.arange(RC_RES, dtype=f32) * f32(RC_VS)).astype(f32)
    f = slabs_field(slabs)(z)
    for k in range(1, RC_RES):
        if f[k] < 0:
            if f[k - 1] <= 0:
                return float(f32(k - 1) * f32(RC_VS))
            return float(f32(k - 1) * f32(RC_VS) + f32(RC_VS) * f[k - 1] / (f[k - 1] - f[k]))
    return math.nan


def check_thin_feature_contract(cast1, tol=1e-5):
    """cast1(tsdf, weight, min_depth, max_depth) -> (vertex (3,), normal (3,)) of the single
    on-axis ray. Gates A-E of tests/tsdf_subvoxel_thin_feature_contract.cpp."""

    def run(slabs, lo=RC_NEAR, hi=RC_FAR):
        t, w = layered_volume(slabs_field(slabs), weight=1.0)
        return cast1(t, w, lo, hi)

    def hit(slabs, want, lo=RC_NEAR, hi=RC_FAR):
        v, n = run(slabs, lo, hi)
        assert v.any(), (slabs, want)
        assert abs(float(v[2]) - want) < tol and abs(float(v[0])) < tol and abs(float(v[1])) < tol, (v, want)
        return v, n

    def miss(slabs, lo=RC_NEAR, hi=RC_FAR):
        v, n = run(slabs, lo, hi)
        assert not v.any() and not n.any(), (slabs, lo, hi, v)

    # A: a plate closer than 0.3 m; near bound inside / past the plate
    want = lattice_zero([(0.19, 0.21)])
    assert abs(want - 0.18666667) < 1e-6
    v, n = hit([(0.19, 0.21)], want)
    assert abs(float(n[2]) + 1.0) < 1e-4 and abs(float(n[0])) < 1e-4 and abs(float(n[1])) < 1e-4
    miss([(0.19, 0.21)], lo=0.20)
    miss([(0.19, 0.21)], lo=0.22)
    # B: sub-voxel plates resolve off the lattice planes
    for a, b in ((0.19, 0.21), (0.39, 0.41), (0.395, 0.405)):
        want = lattice_zero([(a, b)])
        assert not math.isnan(want)
        v, _ = hit([(a, b)], want)
        frac = float(v[2]) / RC_VS - math.floor(float(v[2]) / RC_VS)
        assert 1e-3 < frac < 1.0 - 1e-3
    # C: of two plates one voxel apart the first crossing wins
    assert abs(lattice_zero([(0.47, 0.49)]) - 0.46666667) < 1e-6
    want = lattice_zero([(0.39, 0.41), (0.47, 0.49)])
    assert abs(want - 0.38666667) < 1e-6
    hit([(0.39, 0.41), (0.47, 0.49)], want)
    # D: the configured far bound truncates the march
    want = lattice_zero([(0.91, 0.93)])
    assert abs(want - 0.90666667) < 1e-6
    miss([(0.91, 0.93)], hi=0.90)
    hit([(0.91, 0.93)], want, hi=0.95)
    # E: phase sweep; unrepresentable phases report no surface
    hits = 0
    for d in (0.0, 0.005, 0.0075, 0.015, 0.02, 0.025):
        s = [(0.19 + d, 0.21 + d)]
        want = lattice_zero(s)
        if math.isnan(want):
            miss(s)
        else:
            hit(s, want)
            hits += 1
    assert hits == 3


# ---------------------------------------------------------------------------
# tests/tsdf_integration_min_depth_contract.cpp:59-236: 64^3 at 1 cm, tau 2.5 cm, origin
# (-0.32, -0.32, 0), a 4x4 frame with one valid on-axis pixel (2, 2), fx = fy = 500,
# cx = cy = 2, identity pose. Axis voxel k takes sdf = D - k vs.
# ---------------------------------------------------------------------------
BAND_PARAMS = dict(resolution=64, voxel_size=0.01, truncation=0.025, max_weight=128.0, origin=(-0.32, -0.32, 0.0),
                   min_depth=0.3, max_depth=5.0)
BAND_K = (500.0, 500.0, 2.0, 2.0)
BAND_D = 0.50125


def band_frame(value):
    d = np.zeros((4, 4), f32)
    d[2, 2] = f32(value)
    return d


def check_band_contract(integrate, tol=1e-5):
    """integrate(depth (4, 4), min_depth, max_depth) -> (tsdf, weight) [z, y, x] of a fresh
    volume after one frame. tol bounds the stored surface value's error (1e-5 in the
    reference; a binary16 store of 0.05 is off by up to half its 2^-15 spacing)."""

    def empty(t, w):
        return bool((w == 0).all() and (t == 1).all())

    # A: min_depth is a per-pixel gate
    assert empty(*integrate(band_frame(BAND_D), 0.60, 3.0))
    t, w = integrate(band_frame(BAND_D), 0.10, 3.0)
    assert (w > 0).sum() > 0
    # B: an accepted pixel folds the hand values whatever min_depth is
    t2, w2 = integrate(band_frame(BAND_D), 0.50, 3.0)
    for tt, ww in ((t, w), (t2, w2)):
        assert ww[50, 32, 32] == 1.0 and abs(float(tt[50, 32, 32]) - (BAND_D - 0.50) / 0.025) < tol
        assert ww[47, 32, 32] == 1.0 and tt[47, 32, 32] == 1.0
    assert np.array_equal(t.view(np.uint32), t2.view(np.uint32)) and np.array_equal(w, w2)
    # C: the far bound and non-finite depths gate
    assert empty(*integrate(band_frame(BAND_D), 0.1, 0.40))
    for bad in (np.nan, np.inf, -np.inf):
        assert empty(*integrate(band_frame(bad), 0.1, 3.0))
    assert (integrate(band_frame(BAND_D), 0.1, 0.51)[1] > 0).sum() > 0


# ---------------------------------------------------------------------------
# Marching cubes: tests/marching_cubes_table_contract.cpp (crossing edges from corner
# signs; the three rows the reference repaired), marching_cubes_sphere_contract.cpp
# (48^3 at 2 cm, r = 0.30 m centred on the volume: 8588 triangles, closed, manifold,
# outward), marching_cubes_winding_contract.cpp (0.99 V_sphere < V_signed < V_sphere),
# mesh_welding_contract.cpp (welded vertex count = distinct crossing edges used).
# ---------------------------------------------------------------------------
CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def crossing_mask(cfg):
    m = 0
    for e, (a, b) in enumerate(EDGES):
        if ((cfg >> a) & 1) != ((cfg >> b) & 1):
            m |= 1 << e
    return m


def tri_rows():
    """The triangle rows of the shared table header, as lists of edge ids."""
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "azu_amd", "csrc",
                            "mc_tables.h")).read()
    body = src[src.index("kTriRows[256]"):]
    rows = re.findall(r'"([0-9a-f]*)"', body)[:256]
    return [[int(c, 16) for c in r] for r in rows]


def sphere_volume(res, vs, radius, trunc, weight):
    """tsdf = clamp((|x - c| - r) / trunc) on the lattice, c the volume centre, origin 0."""
    half = f32(0.5) * f32(res - 1) * f32(vs)
    idx = np.arange(res, dtype=f32) * f32(vs)
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    d = np.sqrt((x - half) ** 2 + ((y - half) ** 2 + (z - half) ** 2)).astype(f32)
    t = _clamp1(((d - f32(radius)) / f32(trunc)).astype(f32)).astype(f32)
    return t, np.full_like(t, f32(weight)), np.array([half, half, half], np.float64)


def mesh_topology(idx):
    tris = np.asarray(idx, np.int64).reshape(-1, 3)
    directed = {}
    for t in tris:
        for i in range(3):
            e = (int(t[i]), int(t[(i + 1) % 3]))
            directed[e] = directed.get(e, 0) + 1
    boundary = sum(1 for (a, b) in directed if (b, a) not in directed)
    nonmanifold = sum(1 for c in directed.values() if c > 1)
    return boundary, nonmanifold


def signed_volume(pos, idx):
    p = np.asarray(pos, np.float64)[np.asarray(idx, np.int64).reshape(-1, 3)]
    return float(np.einsum("ij,ij->i", p[:, 0], np.cross(p[:, 1], p[:, 2])).sum() / 6.0)


def welded_keys(t, w, rows):
    """Distinct canonical edge keys (lower voxel, axis) used by the triangles the
    reference's per-cube rule emits (supported corners: finite tsdf, weight > 1e-3)."""
    R = t.shape[0]
    sup = np.isfinite(t) & (w > 1e-3)
    val = np.where(sup, t, f32(1.0))
    keys = set()
    for z in range(R - 1):
        for y in range(R - 1):
            for x in range(R - 1):
                cfg = 0
                ok = []
                for c, (dx, dy, dz) in enumerate(CORNERS):
                    if val[z + dz, y + dy, x + dx] < 0:
                        cfg |= 1 << c
                    ok.append(sup[z + dz, y + dy, x + dx])
                mask = crossing_mask(cfg)
                if not mask:
                    continue
                r = rows[cfg]
                for i in range(0, len(r), 3):
                    tri = r[i:i + 3]
                    if not all((mask >> e) & 1 and ok[EDGES[e][0]] and ok[EDGES[e][1]] for e in tri):
                        continue
                    for e in tri:
                        a, b = EDGES[e]
                        pa = (x + CORNERS[a][0], y + CORNERS[a][1], z + CORNERS[a][2])
                        pb = (x + CORNERS[b][0], y + CORNERS[b][1], z + CORNERS[b][2])
                        axis = [i2 for i2 in range(3) if pa[i2] != pb[i2]][0]
                        keys.add((min(pa, pb), axis))
    return keys


def check_mc_contracts(extract):
    """extract(tsdf, weight, voxel_size) -> (positions, normals, indices, truncated)."""
    rows = tri_rows()
    # table contract: every triangle edge crosses; the reference's repaired rows
    assert crossing_mask(213) == 0x83F and crossing_mask(214) == 0xB35 and crossing_mask(215) == 0xA3C
    for cfg in range(256):
        m = crossing_mask(cfg)
        assert len(rows[cfg]) % 3 == 0 and all((m >> e) & 1 for e in rows[cfg]), cfg
        assert (len(rows[cfg]) == 0) == (m == 0), cfg
    # sphere contract (48^3, 2 cm, r 0.30, trunc 2.5 voxels, weight 64)
    t, w, c = sphere_volume(48, 0.02, 0.30, 2.5 * f32(0.02), 64.0)
    pos, nrm, idx, tr = extract(t, w, 0.02)
    assert not tr and len(idx) // 3 == 8588
    assert np.isfinite(pos).all() and np.isfinite(nrm).all()
    boundary, nonmanifold = mesh_topology(idx)
    assert boundary == 0 and nonmanifold == 0
    p = pos.astype(np.float64)[idx.astype(np.int64).reshape(-1, 3)]
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    area = np.linalg.norm(fn, axis=1)
    assert (area > 1e-9).all()
    radial = p.mean(axis=1) - c
    cosang = np.einsum("ij,ij->i", fn, radial) / (area * np.linalg.norm(radial, axis=1))
    assert (cosang > 0.5).all()
    V, Vs = signed_volume(pos, idx), 4.0 / 3.0 * math.pi * 0.30 ** 3
    assert 0.99 * Vs < V < Vs, (V, Vs)
    pos2, nrm2, idx2, _ = extract(t, w, 0.02)
    assert np.array_equal(pos, pos2) and np.array_equal(nrm, nrm2) and np.array_equal(idx, idx2)
    # welding contract (28^3, 3 cm, r 0.30, trunc 3 voxels, weight 8)
    t, w, _ = sphere_volume(28, 0.03, 0.30, 3.0 * f32(0.03), 8.0)
    pos, nrm, idx, tr = extract(t, w, 0.03)
    assert len(pos) == len(welded_keys(t, w, rows))
    assert mesh_topology(idx)[0] == 0
