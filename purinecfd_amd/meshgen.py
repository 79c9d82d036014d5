"""O-mesh generators writing the reference mesh text format (jst-rs mesh.rs:3-10).

Header ``i_total j_total`` then ``x y`` per node, ring by ring (i = 0 is the wall ring,
i increasing outward), points counter-clockwise within a ring -- the orientation the
reference normals assume (jst-rs geometry.rs:149-183).  ``closed=True`` repeats the first
point of every ring at its end, which Mesh::parse trims (mesh.rs:169-189).

* ``ellipse_mesh`` is the family of jst-py/tools/genmesh.py:26-73 (ellipse wall -> circular
  far field, geometric radial stretch); a_wall == b_wall gives a circular cylinder.
* ``naca_mesh`` is an algebraic O-mesh around a NACA 4-digit section with a closed trailing
  edge: cosine clustering along the chord, grid lines that leave the wall along its normal
  and bend over to a far-field circle centred at mid-chord, geometric radial spacing set by
  the first wall spacing.
"""
from __future__ import annotations

import math

import numpy as np


def _svals_geometric(n_rings: int, stretch: float) -> list[float]:
    if abs(stretch - 1.0) < 1e-12:
        return [i / (n_rings - 1) for i in range(n_rings)]
    w = [stretch ** k for k in range(n_rings - 1)]
    tot = sum(w)
    out, acc = [0.0], 0.0
    for x in w:
        acc += x
        out.append(acc / tot)
    return out


def stretch_for_first_spacing(n_rings: int, length: float, h1: float) -> float:
    """Geometric ratio q with h1 * (q^(n-1) - 1)/(q - 1) == length (bisection)."""
    n = n_rings - 1
    if h1 * n >= length:
        return 1.0
    lo, hi = 1.0 + 1e-12, 2.0
    while h1 * (hi ** n - 1.0) / (hi - 1.0) < length:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if h1 * (mid ** n - 1.0) / (mid - 1.0) < length:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def _fmt(points, n_rings, per, fmt):
    out = [f"{n_rings} {per}\n"]
    out.extend(fmt.format(x, y) for x, y in points)
    return "".join(out)


def ellipse_mesh(n_rings: int, n_theta: int, a_wall: float = 1.0, b_wall: float = 0.5,
                 r_far: float = 5.0, stretch: float = 1.0, closed: bool = False,
                 fmt: str = "{:.17g} {:.17g}\n") -> str:
    """genmesh.py:26-73 family; ``fmt="{:.10f} {:.10f}\\n"`` reproduces its files."""
    if n_rings < 4 or n_theta < 8:
        raise ValueError("need at least 4 rings and 8 circumferential points")
    per = n_theta + 1 if closed else n_theta
    pts = []
    for s in _svals_geometric(n_rings, stretch):
        a = a_wall + s * (r_far - a_wall)
        b = b_wall + s * (r_far - b_wall)
        for j in range(per):
            th = 2.0 * math.pi * (j % n_theta) / n_theta
            pts.append((a * math.cos(th), b * math.sin(th)))
    return _fmt(pts, n_rings, per, fmt)


def cylinder_mesh(n_theta: int, n_rings: int, radius: float = 0.5, r_far: float = 20.0,
                  first_spacing: float | None = None, closed: bool = False) -> tuple[str, float]:
    """Circular cylinder O-mesh; default first radial spacing = circumferential spacing at
    the wall.  Returns (mesh text, geometric stretch ratio used)."""
    h1 = first_spacing if first_spacing else 2.0 * math.pi * radius / n_theta
    q = stretch_for_first_spacing(n_rings, r_far - radius, h1)
    return ellipse_mesh(n_rings, n_theta, radius, radius, r_far, q, closed), q


def naca4(m: float, p: float, t: float, x: float) -> tuple[float, float, float, float]:
    """NACA 4-digit upper/lower surface points at chord station x (closed trailing edge)."""
    yt = 5.0 * t * (0.2969 * math.sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x ** 3
                    - 0.1036 * x ** 4)
    if m > 0.0 and p > 0.0:
        if x < p:
            yc = m / (p * p) * (2.0 * p * x - x * x)
            dyc = 2.0 * m / (p * p) * (p - x)
        else:
            yc = m / ((1.0 - p) ** 2) * ((1.0 - 2.0 * p) + 2.0 * p * x - x * x)
            dyc = 2.0 * m / ((1.0 - p) ** 2) * (p - x)
    else:
        yc, dyc = 0.0, 0.0
    th = math.atan(dyc)
    return (x - yt * math.sin(th), yc + yt * math.cos(th),
            x + yt * math.sin(th), yc - yt * math.cos(th))


def _smooth_directions(v: np.ndarray, sigma: float) -> np.ndarray:
    """Unit vectors v (n x 2, periodic in the first index) convolved with a periodic Gaussian
    of standard deviation ``sigma`` points, renormalised."""
    n = v.shape[0]
    k = np.fft.fftfreq(n)
    g = np.exp(-2.0 * (math.pi * sigma * k) ** 2)
    out = np.fft.ifft(np.fft.fft(v, axis=0) * g[:, None], axis=0).real
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def naca_mesh(n_theta: int, n_rings: int, m: float = 0.0, p: float = 0.4, t: float = 0.12,
              first_spacing: float = 1e-5, r_far: float = 30.0, closed: bool = False,
              wall_sigma: float = 3.0, far_sigma: float = 0.125) -> tuple[str, float]:
    """Algebraic O-mesh around a NACA 4-digit section (chord 1, LE at x=0, TE at x=1).

    Wall points j = 0..n_theta-1 follow theta_j = 2 pi j / n_theta: x = (1 + cos theta)/2,
    upper surface for theta < pi (counter-clockwise from the trailing edge), lower surface
    after.  N_j is the outward wall normal from the central difference of the wall ring (at
    the trailing-edge point the bisector), smoothed over ``wall_sigma`` points so that the
    grid lines fan out around the sharp trailing edge; F_j is where the ray from W_j along
    N_j smoothed over ``far_sigma`` of the ring meets the far-field circle (centre (0.5, 0),
    radius r_far; the wide smoothing keeps these rays from crossing under a concave lower
    surface).  Ring i, with s_i geometric and s_1 r_far = first_spacing, is

        P_ij = W_j + (1 - s_i) s_i r_far N_j + s_i^2 (F_j - W_j),

    i.e. normal to the wall at the wall (first spacing first_spacing (1 - s_1) + O(s_1^2)),
    the straight line W_j -> F_j at the far field.  Grid lines that meet the wall at a
    shallow angle (straight lines to far-field points spaced evenly in angle meet it at 10
    degrees next to the trailing edge) make the explicit scheme unstable at the trailing-edge
    wall cells.  Returns (mesh text, stretch ratio)."""
    q = stretch_for_first_spacing(n_rings, r_far, first_spacing)
    svals = np.asarray(_svals_geometric(n_rings, q))
    wall = []
    for j in range(n_theta):
        th = 2.0 * math.pi * j / n_theta
        x = 0.5 * (1.0 + math.cos(th))
        xu, yu, xl, yl = naca4(m, p, t, x)
        if j == 0:
            wall.append((1.0, 0.0))
        elif th < math.pi:
            wall.append((xu, yu))
        else:
            wall.append((xl, yl))
    w = np.array(wall)
    tang = np.roll(w, -1, axis=0) - np.roll(w, 1, axis=0)
    nrm = np.stack([tang[:, 1], -tang[:, 0]], axis=1)  # outward for a counter-clockwise ring
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    # the smoothed normals blend into the exact ones away from the trailing edge, so that the
    # curved leading edge keeps grid lines normal to the wall
    d_te = np.minimum(np.arange(n_theta), n_theta - np.arange(n_theta))
    wgt = np.exp(-0.5 * (d_te / (2.0 * wall_sigma)) ** 2)[:, None]
    n_wall = wgt * _smooth_directions(nrm, wall_sigma) + (1.0 - wgt) * nrm
    n_wall /= np.linalg.norm(n_wall, axis=1, keepdims=True)
    n_far = _smooth_directions(nrm, far_sigma * n_theta)
    rel = w - np.array([0.5, 0.0])
    b = np.sum(rel * n_far, axis=1)
    dist = -b + np.sqrt(b * b - (np.sum(rel * rel, axis=1) - r_far * r_far))
    far = w + dist[:, None] * n_far
    pts = (w[None] + ((1.0 - svals) * svals * r_far)[:, None, None] * n_wall[None]
           + (svals * svals)[:, None, None] * (far - w)[None])
    pts[0] = w
    pts[-1] = far
    if closed:
        pts = np.concatenate([pts, pts[:, :1]], axis=1)
    per = pts.shape[1]
    return _fmt(pts.reshape(-1, 2).tolist(), n_rings, per, "{:.17g} {:.17g}\n"), q
