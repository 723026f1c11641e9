"""CPU oracle (test infrastructure) of the mesh voxelisation used for scene setup.

numpy FP64 restatement of assets/shaders/sdf.comp:53-115 (the editor's SDF compute pass that
Editor::initFluid / initBoundaries dispatch, src/core/Editor.cpp:90-258): a cell centre
domain_min + cell_size * (i + 0.5) (cell_size = domain_size / (W, H, D)) is inside when the
generalized winding number, the sum of the triangles' solid angles
2 atan2(a . (b x c), 1 + a.b + b.c + c.a) with a, b, c the unit vectors to the vertices, exceeds
pi in magnitude; cells of the last layer of each axis are left outside.  Returns the winding
number as well, so that tests can exclude the cells whose decision is within rounding of the
threshold.

Pinned in tests/test_scene_json.py against the analytic inside test of closed meshes (a cube
and a closed cup) and the winding number of points inside/outside (4 pi / 0).
"""
import numpy as np


def winding(points, tris):
    """points (n, 3), tris (m, 3, 3) -> winding number sum (n,) in FP64."""
    p = np.asarray(points, np.float64)
    w = np.zeros(len(p))
    for t in np.asarray(tris, np.float64):
        a = t[0] - p
        b = t[1] - p
        c = t[2] - p
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        c /= np.linalg.norm(c, axis=1, keepdims=True)
        num = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = 1.0 + np.einsum("ij,ij->i", a, b) + np.einsum("ij,ij->i", b, c) + np.einsum("ij,ij->i", c, a)
        w += 2.0 * np.arctan2(num, den)
    return w


def voxelise(dims, tris, domain_center, domain_size):
    """dims (W, H, D) -> (inside bool (D, H, W), winding (D, H, W))."""
    W, H, D = dims
    cs = np.asarray(domain_size, np.float64) / np.array([W, H, D], np.float64)
    dmin = np.asarray(domain_center, np.float64) - 0.5 * np.asarray(domain_size, np.float64)
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    pts = np.stack([dmin[0] + cs[0] * x + cs[0] * 0.5, dmin[1] + cs[1] * y + cs[1] * 0.5, dmin[2] + cs[2] * z + cs[2] * 0.5],
                   axis=-1).reshape(-1, 3)
    w = winding(pts, tris).reshape(D, H, W)
    inside = np.abs(w) > np.pi
    inside[-1, :, :] = False
    inside[:, -1, :] = False
    inside[:, :, -1] = False
    return inside, w
