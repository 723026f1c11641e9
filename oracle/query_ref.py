"""CPU oracle (test infrastructure) of the query-box statistics.

numpy restatement of QueryNode::update's box loops (src/simulation/scene/nodes/QueryNode.cpp:
45-84): over z, y, x in [max(0, lo), min(n - 1, hi)), cells with rho < 0.5 skipped when
only_fluid, the running max / min of the cell coordinates, the count, and the sums of the
coordinates, of the velocity (u, v, w at the cell) and of the pressure; means when the count is
not 0, else max = min = 0.  With dtype float32 the sums run sequentially in the loop order, as
the reference's glm::vec3 / float accumulators do; with float64 they are exact to rounding.

Pinned in tests/test_query.py by closed forms: count = box volume, min = lo, max = hi - 1,
mean position = (lo + hi - 1) / 2 on a full box, constant fields averaging to the constant,
an empty box giving zeros, and the fluid-only filter on a half-filled box.
"""
import numpy as np


def query_box(rho, u, v, w, p, lo, hi, only_fluid, dtype=np.float64):
    """Fields (D, H, W) (w may be None in 2D, where D = 1) -> dict of count, max_pos, min_pos,
    avg_pos, vel, pressure (float arrays of 3, pressure a float)."""
    D, H, W = rho.shape
    n = (W, H, D)
    a = [max(0, int(lo[k])) for k in range(3)]
    b = [max(a[k], min(n[k] - 1, int(hi[k]))) for k in range(3)]
    if D == 1:
        a[2], b[2] = 0, 1
    sl = (slice(a[2], b[2]), slice(a[1], b[1]), slice(a[0], b[0]))
    z, y, x = np.meshgrid(np.arange(a[2], b[2]), np.arange(a[1], b[1]), np.arange(a[0], b[0]), indexing="ij")
    keep = np.ones(z.shape, bool)
    if only_fluid:
        keep = rho[sl] >= np.float32(0.5)
    xs, ys, zs = x[keep], y[keep], z[keep]  # C order = the loop order z, y, x
    cnt = int(keep.sum())
    out = {"count": cnt, "max_pos": np.zeros(3), "min_pos": np.zeros(3), "avg_pos": np.zeros(3), "vel": np.zeros(3),
           "pressure": 0.0}
    if cnt == 0:
        return out

    def total(vals):
        vals = np.asarray(vals).astype(dtype)
        return np.cumsum(vals, dtype=dtype)[-1] if dtype == np.float32 else vals.sum(dtype=np.float64)

    zero = np.zeros(rho.shape, np.float32)
    out["max_pos"] = np.array([xs.max(), ys.max(), zs.max()], np.float64)
    out["min_pos"] = np.array([xs.min(), ys.min(), zs.min()], np.float64)
    out["avg_pos"] = np.array([total(xs), total(ys), total(zs)], np.float64) / cnt
    out["vel"] = np.array([total(u[sl][keep]), total(v[sl][keep]), total((w if w is not None else zero)[sl][keep])],
                          np.float64) / cnt
    out["pressure"] = float(total(p[sl][keep])) / cnt
    return out
