"""Seeded scene generators and the named render configurations.

Every generator emits the reference's preorder text format
(README.md:37-61 of the reference), so generated scenes go through the same
parser as hand-written ones.  Numbers are written with 9 significant digits,
which round-trips float32 exactly.
"""
from __future__ import annotations

import math

import numpy as np


def _num(v: float) -> str:
    s = f"{float(v):.9g}"
    return "0" if s == "-0" else s


def sphere(c, colour, r):
    return f"Sphere {_num(c[0])} {_num(c[1])} {_num(c[2])} {colour} {_num(r)}"


def cube(c, colour, edge):
    return f"Cube {_num(c[0])} {_num(c[1])} {_num(c[2])} {colour} {_num(edge)}"


def cylinder(c, colour, r, h, rot):
    return (f"Cylinder {_num(c[0])} {_num(c[1])} {_num(c[2])} {colour} {_num(r)} {_num(h)} "
            f"{_num(rot[0])} {_num(rot[1])} {_num(rot[2])}")


def balanced(op: str, leaves: list[str]) -> list[str]:
    """Preorder lines of a balanced binary `op` tree over `leaves` (left half gets the extra leaf)."""
    out: list[str] = []

    def rec(lo, hi):
        if hi - lo == 1:
            out.append(leaves[lo])
            return
        mid = lo + (hi - lo + 1) // 2
        out.append(op)
        rec(lo, mid)
        rec(mid, hi)

    rec(0, len(leaves))
    return out


def random_ops_balanced(rng, leaves: list[str]) -> list[str]:
    """Balanced tree with operators drawn uniformly from Union/Intersection/Difference."""
    out: list[str] = []
    ops = ("Union", "Intersection", "Difference")

    def rec(lo, hi):
        if hi - lo == 1:
            out.append(leaves[lo])
            return
        mid = lo + (hi - lo + 1) // 2
        out.append(ops[int(rng.integers(0, 3))])
        rec(lo, mid)
        rec(mid, hi)

    rec(0, len(leaves))
    return out


def cheese(n_spheres=512, seed=512, r_lo=0.05, r_hi=0.2, edge=2.0) -> str:
    """Cube of edge 2 at the origin minus a balanced Union of spheres with centres
    U[-1,1]^3 and radii U[r_lo, r_hi] (the plain-kernel benchmark scene)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, size=(n_spheres, 3))
    r = rng.uniform(r_lo, r_hi, size=n_spheres)
    leaves = [sphere(c[i], "F5F500", r[i]) for i in range(n_spheres)]
    lines = ["Difference", cube((0, 0, 0), "FFFF00", edge)] + balanced("Union", leaves)
    return "\n".join(lines) + "\n"


def stress(n_prims=4095, seed=4095) -> str:
    """Cube of edge 2 minus a balanced Union of mixed small spheres and rotated
    cylinders (8191 nodes for 4095 primitives)."""
    rng = np.random.default_rng(seed)
    leaves = []
    for _ in range(n_prims):
        c = rng.uniform(-1.0, 1.0, size=3)
        if rng.random() < 0.5:
            leaves.append(sphere(c, "F5F500", rng.uniform(0.02, 0.1)))
        else:
            rot = rng.uniform(0.0, 180.0, size=3)
            leaves.append(cylinder(c, "00C0F0", rng.uniform(0.01, 0.05), rng.uniform(0.2, 1.0), rot))
    return "\n".join(["Difference", cube((0, 0, 0), "FFFF00", 2.0)] + balanced("Union", leaves)) + "\n"


def random_tree(rng) -> str:
    """One tree of the batched mix: leaf count round(exp(U[ln 2, ln 1024])),
    leaves uniform over sphere / cylinder / cube with the mix's size ranges,
    centres U[-1,1]^3, balanced shape, operators uniform."""
    n = int(round(math.exp(rng.uniform(math.log(2), math.log(1024)))))
    n = max(2, min(1024, n))
    leaves = []
    for _ in range(n):
        c = rng.uniform(-1.0, 1.0, size=3)
        kind = int(rng.integers(0, 3))
        col = "%02X%02X%02X" % tuple(int(v) for v in rng.integers(40, 256, size=3))
        if kind == 0:
            leaves.append(sphere(c, col, rng.uniform(0.1, 0.8)))
        elif kind == 1:
            leaves.append(cylinder(c, col, rng.uniform(0.1, 0.6), rng.uniform(0.5, 2.0),
                                   rng.uniform(0.0, 180.0, size=3)))
        else:
            leaves.append(cube(c, col, rng.uniform(0.2, 1.5)))
    return "\n".join(random_ops_balanced(rng, leaves)) + "\n"


def orbit_eye(rng, radius=4.0, max_elev_deg=80.0):
    az = rng.uniform(0.0, 2 * math.pi)
    el = math.radians(rng.uniform(-max_elev_deg, max_elev_deg))
    return (radius * math.cos(el) * math.sin(az), radius * math.sin(el), radius * math.cos(el) * math.cos(az))


def batch_mix(n=256, seed=256, resolutions=(128, 256, 512, 1024)):
    """The batched mix: n random trees (see random_tree), each with a random
    orbit camera at radius 4 looking at the origin and a square resolution
    drawn uniformly from `resolutions`.  Returns [(text, eye, size)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        text = random_tree(rng)
        eye = orbit_eye(rng, 4.0, 85.0)
        size = int(resolutions[int(rng.integers(0, len(resolutions)))])
        out.append((text, eye, size))
    return out


# ---- named configurations --------------------------------------------------

# Shared camera / light / shading of the Wikipedia and cheese configurations:
# eye (3, 2.5, 4) looking at the origin with a 60 degree vertical fov; light
# travelling along (-1,-1,-1), i.e. toward-light direction (1,1,1); Phong
# ka 0.1, kd 0.7, ks 0.2, n 32.
VIEW_EYE = (3.0, 2.5, 4.0)
VIEW_TARGET = (0.0, 0.0, 0.0)
VIEW_VFOV_DEG = 60.0
LIGHT_TOWARD = (1.0, 1.0, 1.0)
PHONG = (0.1, 0.7, 0.2, 32.0)
BACKGROUND = (0.08, 0.08, 0.11, 1.0)

WIKIPEDIA = """Difference
	Intersection
		Cube 0 0 0 FF0000 2
		Sphere 0 0 0 0000FF 1.35
	Union
		Union
			Cylinder 0 0 0 00FF00 0.7 2 90 0 0
			Cylinder 0 0 0 00FF00 0.7 2 0 0 0
		Cylinder 0 0 0 00FF00 0.7 2 0 0 90
"""
