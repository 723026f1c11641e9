"""Procedural closed triangle meshes (outward-facing) for scene setup and tests: an axis-aligned
box and a cup (a box with a box-shaped cavity open at the top, like the bucket obstacles of the
reference's scenes)."""
import numpy as np


def box(lo, hi):
    """12 triangles of the box [lo, hi], outward normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = [(0, 3, 2), (0, 2, 1),  # z0 (normal -z)
         (4, 5, 6), (4, 6, 7),  # z1
         (0, 1, 5), (0, 5, 4),  # y0
         (3, 7, 6), (3, 6, 2),  # y1
         (0, 4, 7), (0, 7, 3),  # x0
         (1, 2, 6), (1, 6, 5)]  # x1
    return v[np.asarray(f)]


def cup(lo, hi, wall):
    """Closed surface of a cup: the box [lo, hi] minus the cavity [lo + wall, hi - wall] extended
    through the top face (y = hi[1]).  The rim ring and the cavity walls complete the surface."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    a0, b0 = x0 + wall, z0 + wall
    a1, b1 = x1 - wall, z1 - wall
    yb = y0 + wall  # cavity floor
    P = lambda x, y, z: np.array([x, y, z], np.float32)
    quads = []
    # outer sides and bottom (outward)
    quads += [(P(x0, y0, z0), P(x1, y0, z0), P(x1, y0, z1), P(x0, y0, z1))]  # bottom, -y
    quads += [(P(x0, y0, z0), P(x0, y1, z0), P(x1, y1, z0), P(x1, y0, z0))]  # -z
    quads += [(P(x0, y0, z1), P(x1, y0, z1), P(x1, y1, z1), P(x0, y1, z1))]  # +z
    quads += [(P(x0, y0, z0), P(x0, y0, z1), P(x0, y1, z1), P(x0, y1, z0))]  # -x
    quads += [(P(x1, y0, z0), P(x1, y1, z0), P(x1, y1, z1), P(x1, y0, z1))]  # +x
    # rim (top ring, +y) as four quads
    quads += [(P(x0, y1, z0), P(x0, y1, z1), P(a0, y1, z1), P(a0, y1, z0))]
    quads += [(P(a1, y1, z0), P(a1, y1, z1), P(x1, y1, z1), P(x1, y1, z0))]
    quads += [(P(a0, y1, z0), P(a0, y1, b0), P(a1, y1, b0), P(a1, y1, z0))]
    quads += [(P(a0, y1, b1), P(a0, y1, z1), P(a1, y1, z1), P(a1, y1, b1))]
    # cavity floor (+y, facing into the cavity) and walls (facing the cavity)
    quads += [(P(a0, yb, b0), P(a0, yb, b1), P(a1, yb, b1), P(a1, yb, b0))]
    quads += [(P(a0, yb, b0), P(a1, yb, b0), P(a1, y1, b0), P(a0, y1, b0))]  # wall z = b0, normal +z
    quads += [(P(a0, yb, b1), P(a0, y1, b1), P(a1, y1, b1), P(a1, yb, b1))]  # wall z = b1, normal -z
    quads += [(P(a0, yb, b0), P(a0, y1, b0), P(a0, y1, b1), P(a0, yb, b1))]  # wall x = a0, normal +x
    quads += [(P(a1, yb, b0), P(a1, yb, b1), P(a1, y1, b1), P(a1, y1, b0))]  # wall x = a1, normal -x
    tris = []
    for q in quads:
        tris.append([q[0], q[1], q[2]])
        tris.append([q[0], q[2], q[3]])
    return np.asarray(tris, np.float32)


def inside_cup(p, lo, hi, wall):
    """Analytic inside test of cup(lo, hi, wall) for points (n, 3)."""
    p = np.asarray(p, np.float64)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    in_box = np.all((p > lo) & (p < hi), axis=1)
    cav_lo = np.array([lo[0] + wall, lo[1] + wall, lo[2] + wall])
    cav_hi = np.array([hi[0] - wall, hi[1] + 1.0, hi[2] - wall])
    in_cav = np.all((p > cav_lo) & (p < cav_hi), axis=1)
    return in_box & ~in_cav


def to_obj(tris, path):
    """Writes triangles as a Wavefront OBJ (one vertex record per corner)."""
    with open(path, "w") as f:
        for t in tris:
            for v in t:
                f.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for i in range(len(tris)):
            f.write(f"f {3 * i + 1} {3 * i + 2} {3 * i + 3}\n")
