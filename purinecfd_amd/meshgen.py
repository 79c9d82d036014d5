"""O-mesh generators writing the reference mesh text format (jst-rs mesh.rs:3-10).

Header ``i_total j_total`` then ``x y`` per node, ring by ring (i = 0 is the wall ring,
i increasing outward), points counter-clockwise within a ring -- the orientation the
reference normals assume (jst-rs geometry.rs:149-183).  ``closed=True`` repeats the first
point of every ring at its end, which Mesh::parse trims (mesh.rs:169-189).

* ``ellipse_mesh`` is the family of jst-py/tools/genmesh.py:26-73 (ellipse wall -> circular
  far field, geometric radial stretch); a_wall == b_wall gives a circular cylinder.
* ``naca_mesh`` is an algebraic O-mesh around a NACA 4-digit section with a closed trailing
  edge: cosine clustering along the chord, straight radial lines to a far-field circle
  centred at mid-chord, geometric radial spacing set by the first wall spacing.
"""
from __future__ import annotations

import math


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


def naca_mesh(n_theta: int, n_rings: int, m: float = 0.0, p: float = 0.4, t: float = 0.12,
              first_spacing: float = 1e-5, r_far: float = 30.0,
              closed: bool = False) -> tuple[str, float]:
    """Algebraic O-mesh around a NACA 4-digit section (chord 1, LE at x=0, TE at x=1).

    Wall points j = 0..n_theta-1 follow theta_j = 2 pi j / n_theta: x = (1 + cos theta)/2,
    upper surface for theta < pi (counter-clockwise from the trailing edge), lower surface
    after.  Ring i is wall + s_i (far - wall) with the far point at mid-chord + r_far
    (cos theta, sin theta) and s_i geometric with s_1 * r_far == first_spacing.
    Returns (mesh text, stretch ratio)."""
    q = stretch_for_first_spacing(n_rings, r_far, first_spacing)
    svals = _svals_geometric(n_rings, q)
    per = n_theta + 1 if closed else n_theta
    wall, far = [], []
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
        far.append((0.5 + r_far * math.cos(th), r_far * math.sin(th)))
    pts = []
    for s in svals:
        for j in range(per):
            (wx, wy), (fx, fy) = wall[j % n_theta], far[j % n_theta]
            pts.append((wx + s * (fx - wx), wy + s * (fy - wy)))
    return _fmt(pts, n_rings, per, "{:.17g} {:.17g}\n"), q
