"""Mesh generators (no reference counterpart beyond jst-py/tools/genmesh.py's ellipse family):
closed forms of the NACA 4-digit section, the requested wall spacing, and the geometry the
reference computes on the generated meshes (jst-rs geometry.rs:270-296 checks: positive
volumes, wall normals into the fluid, far-field normals outward, metric closure), with the
closed-ring variant exercising Mesh::parse's closing-duplicate trim (mesh.rs:169-189)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import config_dict
from purinecfd_amd import meshgen


def test_naca4_closed_forms():
    # closed trailing edge: the thickness polynomial vanishes at x = 1
    xu, yu, xl, yl = meshgen.naca4(0.0, 0.4, 0.12, 1.0)
    assert abs(yu) < 1e-15 and abs(yl) < 1e-15 and xu == xl == 1.0
    # NACA 0012: half thickness 0.6 * (0.2969 sqrt(.3) - .126 .3 - .3516 .09 + .2843 .027
    # - .1036 .0081) at x = 0.3
    _, yu, _, yl = meshgen.naca4(0.0, 0.4, 0.12, 0.3)
    want = 0.6 * (0.2969 * math.sqrt(0.3) - 0.126 * 0.3 - 0.3516 * 0.09 + 0.2843 * 0.027 - 0.1036 * 0.0081)
    assert abs(yu - want) < 1e-15 and yl == -yu and abs(want - 0.0600078) < 1e-6
    # camber line at x = p is y = m, with zero slope there: upper and lower straddle it
    xu, yu, xl, yl = meshgen.naca4(0.04, 0.4, 0.12, 0.4)
    assert xu == xl == 0.4 and abs(0.5 * (yu + yl) - 0.04) < 1e-15


def _nodes(po, text):
    n_rings, n_theta, xy = po.parse_mesh(text)
    return n_rings, n_theta, xy.reshape(n_rings, n_theta, 2)


def test_cylinder_first_spacing_and_far_field(oracle_lib):
    text, q = meshgen.cylinder_mesh(40, 17, 0.5, 20.0)
    n_rings, n_theta, xy = _nodes(oracle_lib, text)
    assert (n_rings, n_theta) == (17, 40)
    r = np.hypot(xy[..., 0], xy[..., 1])
    assert np.max(np.abs(r[0] - 0.5)) < 1e-15 and np.max(np.abs(r[-1] - 20.0)) < 1e-13
    h1 = 2.0 * math.pi * 0.5 / 40  # default: first spacing = wall circumferential spacing
    assert np.max(np.abs((r[1] - r[0]) - h1)) < 1e-12
    steps = np.diff(r[:, 0])
    assert np.allclose(steps[1:] / steps[:-1], q, rtol=1e-10)


def test_closed_ring_is_trimmed_to_the_open_one(oracle_lib):
    for open_, closed in ((meshgen.cylinder_mesh(32, 9, 0.5, 20.0)[0],
                           meshgen.cylinder_mesh(32, 9, 0.5, 20.0, closed=True)[0]),
                          (meshgen.naca_mesh(64, 17, 0.02, 0.4, 0.12, 1e-3, 30.0)[0],
                           meshgen.naca_mesh(64, 17, 0.02, 0.4, 0.12, 1e-3, 30.0, closed=True)[0])):
        assert closed.split("\n", 1)[0].split()[1] == str(int(open_.split("\n", 1)[0].split()[1]) + 1)
        a, b = _nodes(oracle_lib, open_), _nodes(oracle_lib, closed)
        assert a[:2] == b[:2] and np.array_equal(a[2], b[2])


def _geometry_checks(po, text):
    s = po.OracleSolver(config_dict(), text)
    ni, nj = s.ni, s.nj
    vol = s.get("vol").reshape(ni, nj)
    assert np.all(vol > 0.0)
    cx, cy = s.get("cx").reshape(ni, nj), s.get("cy").reshape(ni, nj)
    tnx, tny = s.get("tau_nx").reshape(ni + 1, nj), s.get("tau_ny").reshape(ni + 1, nj)
    tmx, tmy = s.get("tau_mx").reshape(ni + 1, nj), s.get("tau_my").reshape(ni + 1, nj)
    nnx, nny = s.get("n_nx").reshape(ni, nj), s.get("n_ny").reshape(ni, nj)
    # wall faces point into the fluid (towards the first cell), far-field faces outward
    assert np.all(tnx[0] * (cx[0] - tmx[0]) + tny[0] * (cy[0] - tmy[0]) > 0.0)
    assert np.all(tnx[ni] * (tmx[ni] - cx[ni - 1]) + tny[ni] * (tmy[ni] - cy[ni - 1]) > 0.0)
    # metric closure per cell: tau(i+1) - tau(i) + n(j+1) - n(j) = 0 (geometry.rs:17-18)
    clx = tnx[1:] - tnx[:-1] + np.roll(nnx, -1, axis=1) - nnx
    cly = tny[1:] - tny[:-1] + np.roll(nny, -1, axis=1) - nny
    scale = np.max(np.hypot(tnx, tny))
    assert np.max(np.abs(clx)) < 1e-12 * scale and np.max(np.abs(cly)) < 1e-12 * scale
    return s


@pytest.mark.parametrize("kind", ["cylinder", "ellipse", "naca0012", "naca4412"])
def test_generated_meshes_have_valid_geometry(oracle_lib, kind):
    if kind == "cylinder":
        text = meshgen.cylinder_mesh(64, 33, 0.5, 20.0)[0]
    elif kind == "ellipse":
        text = meshgen.ellipse_mesh(17, 48, 1.0, 0.5, 5.0, 1.1)
    elif kind == "naca0012":
        text = meshgen.naca_mesh(128, 33, 0.0, 0.4, 0.12, 1e-4, 30.0)[0]
    else:
        text = meshgen.naca_mesh(128, 33, 0.04, 0.4, 0.12, 1e-4, 30.0)[0]
    _geometry_checks(oracle_lib, text)


def test_naca_wall_ring(oracle_lib):
    text, _ = meshgen.naca_mesh(256, 33, 0.0, 0.4, 0.12, 1e-4, 30.0)
    _, _, xy = _nodes(oracle_lib, text)
    x, y = xy[0, :, 0], xy[0, :, 1]
    # counter-clockwise wall ring (positive shoelace area) from the trailing edge
    area = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
    assert area > 0.0 and (x[0], y[0]) == (1.0, 0.0)
    # section area = 2 * integral of the half thickness over the chord, in closed form for the
    # closed-trailing-edge polynomial: 10 t (2/3 .2969 - .126/2 - .3516/3 + .2843/4 - .1036/5)
    exact = 10 * 0.12 * (2.0 / 3.0 * 0.2969 - 0.126 / 2 - 0.3516 / 3 + 0.2843 / 4 - 0.1036 / 5)
    assert abs(area - exact) < 1e-3 * exact  # polygon of 256 cosine-spaced points
    # symmetric section: the lower surface mirrors the upper one
    assert np.max(np.abs(y[1:] + y[1:][::-1])) < 1e-15
    # first wall spacing at the leading edge, where the wall normal and the far-field ray
    # are both -x and |far - wall| = 29.5: (1 - s_1) s_1 30 + s_1^2 29.5 with s_1 30 = 1e-4
    s1 = 1e-4 / 30.0
    d = np.hypot(xy[1, 128, 0] - x[128], xy[1, 128, 1] - y[128])
    assert abs(d - ((1.0 - s1) * s1 * 30.0 + s1 * s1 * 29.5)) < 1e-15


def _cell_areas(xy):
    """Signed quad areas of an O-mesh node array (rings, n_theta, 2); positive for the
    orientation the reference assumes (counter-clockwise rings, i outward)."""
    a, b = xy[:-1], np.roll(xy[:-1], -1, axis=1)
    c, d = np.roll(xy[1:], -1, axis=1), xy[1:]

    def cross(u, v):
        return u[..., 0] * v[..., 1] - u[..., 1] * v[..., 0]
    return -0.5 * (cross(a, b) + cross(b, c) + cross(c, d) + cross(d, a))


def _wall_angle(xy):
    """Degrees by which the first grid line leaves the wall away from its normal (the wall
    tangent is the central difference of the wall ring)."""
    t = np.roll(xy[0], -1, axis=0) - np.roll(xy[0], 1, axis=0)
    r = xy[1] - xy[0]
    cos = np.abs(np.sum(t * r, axis=1)) / (np.linalg.norm(t, axis=1) * np.linalg.norm(r, axis=1))
    return np.degrees(np.arcsin(np.clip(cos, 0.0, 1.0)))


@pytest.mark.parametrize("n_theta,rings,m,t,h", [(256, 129, 0.0, 0.12, 1e-5),
                                                  (128, 65, 0.04, 0.08, 1e-3),
                                                  (512, 257, 0.04, 0.18, 1e-3)])
def test_naca_grid_lines_leave_the_wall_along_its_normal(oracle_lib, n_theta, rings, m, t, h):
    """Away from the trailing edge the first grid line is the wall normal (to 0.01 degree);
    next to the sharp trailing edge the lines fan out (at most 40 degrees off the normal, where
    straight lines to far-field points evenly spaced in angle leave at 80)."""
    _, _, xy = _nodes(oracle_lib, meshgen.naca_mesh(n_theta, rings, m, 0.4, t, h, 30.0)[0])
    ang = _wall_angle(xy)
    far_from_te = np.abs(((np.arange(n_theta) + n_theta // 2) % n_theta) - n_theta // 2) > 16
    assert ang[far_from_te].max() < 0.01 and ang.max() < 40.0
    assert np.all(_cell_areas(xy) > 0.0)


def test_every_c4_mesh_has_positive_cells():
    """The whole 64-case distribution (cambered sections included, where normals from the
    concave lower surface converge): no folded cell."""
    from purinecfd_amd import workloads
    for w in workloads.c4_batch():
        lines = w["mesh"].split("\n")
        rings, per = map(int, lines[0].split())
        xy = np.array([list(map(float, ln.split())) for ln in lines[1:1 + rings * per]])
        xy = xy.reshape(rings, per, 2)[:, :-1]  # closed rings: drop the closing duplicate
        assert np.all(_cell_areas(xy) > 0.0), w["name"]
