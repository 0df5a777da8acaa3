"""Scene texts with deep operator chains, for the traversal stack's depth range.

The reference's three stacks hold 32 entries (RaycastingKernels.cuh:19,
CudaStack.cuh:14-31), which keeps it exact up to 15 operator levels; this
library holds one slot per operator level up to 64 and rejects deeper trees.
These chains put 16 to 65 operators on one root-to-leaf path.
"""
import numpy as np


def _sphere(rng, r_lo=0.15, r_hi=0.35, span=1.0):
    c = rng.uniform(-span, span, 3)
    col = "".join(f"{int(v):02X}" for v in rng.integers(40, 256, 3))
    return f"Sphere {c[0]:.6g} {c[1]:.6g} {c[2]:.6g} {col} {rng.uniform(r_lo, r_hi):.6g}"


def _cube(rng, lo=0.3, hi=0.8):
    c = rng.uniform(-1, 1, 3)
    return f"Cube {c[0]:.6g} {c[1]:.6g} {c[2]:.6g} 80C0FF {rng.uniform(lo, hi):.6g}"


def _cylinder(rng):
    c = rng.uniform(-1, 1, 3)
    rot = rng.uniform(0, 180, 3)
    return (f"Cylinder {c[0]:.6g} {c[1]:.6g} {c[2]:.6g} FFC080 {rng.uniform(0.1, 0.3):.6g} "
            f"{rng.uniform(0.5, 1.5):.6g} {rot[0]:.6g} {rot[1]:.6g} {rot[2]:.6g}")


def chain(depth, ops, sides, leaves):
    """Preorder text of a chain of `depth` operators: operator i (from the
    root) has ops[i] as its type and continues the chain on sides[i] ('L' or
    'R'), its other child a leaf; the bottom operator has two leaves.
    `leaves` yields leaf lines, the bottom-most first."""
    it = iter(leaves)

    def build(level):
        if level == depth:
            return [next(it)]
        sub = build(level + 1)
        leaf = [next(it)]
        return [ops[level]] + (sub + leaf if sides[level] == "L" else leaf + sub)

    return "\n".join(build(0)) + "\n"


def left_deep_difference(depth, seed=0):
    """A cube of edge 2 with `depth` spheres subtracted one after another:
    Difference(Difference(... Difference(cube, s1) ..., s_{d-1}), s_d)."""
    rng = np.random.default_rng(seed)
    leaves = ["Cube 0 0 0 E0E0E0 2"] + [_sphere(rng) for _ in range(depth)]
    return chain(depth, ["Difference"] * depth, ["L"] * depth, leaves)


def right_deep_union(depth, seed=0):
    """Union(s0, Union(s1, ... Union(s_{d-1}, s_d))): spheres, the chain on the right."""
    rng = np.random.default_rng(seed)
    leaves = [_sphere(rng, 0.15, 0.4) for _ in range(depth + 1)]
    return chain(depth, ["Union"] * depth, ["R"] * depth, leaves)


def mixed_unbalanced(depth, seed=0, cylinders=True):
    """Operators among Union (half), Difference and Intersection (a quarter
    each), the chain turning left or right at random (a Difference keeps it
    on the left, so that it subtracts its leaf); an Intersection's leaf
    is a large central sphere (so the solid does not vanish), the other
    leaves spheres, cubes and (optionally) tilted cylinders."""
    rng = np.random.default_rng(seed)
    ops = [str(rng.choice(["Union", "Union", "Intersection", "Difference"])) for _ in range(depth)]
    # a Difference subtracts its leaf from the chain (the chain on its left);
    # Unions and Intersections take the chain on either side
    sides = ["L" if op == "Difference" else str(rng.choice(["L", "R"])) for op in ops]
    kinds = [_sphere, _cube] + ([_cylinder] if cylinders else [])
    # leaves come bottom-most first: the bottom operator's two, then one per
    # operator upwards (operator depth - 1 - i for the i-th of those)
    leaves = ["Cube 0 0 0 F0F0F0 1.8", kinds[int(rng.integers(0, len(kinds)))](rng)]
    for level in range(depth - 2, -1, -1):
        if ops[level] == "Intersection":
            leaves.append(f"Sphere 0 0 0 C0C0C0 {1.6 + 0.013 * level:.4g}")  # distinct, no coincident surfaces
        else:
            leaves.append(kinds[int(rng.integers(0, len(kinds)))](rng))
    return chain(depth, ops, sides, leaves)


SHAPES = {"left-deep difference": left_deep_difference, "right-deep union": right_deep_union,
          "mixed unbalanced": mixed_unbalanced}
