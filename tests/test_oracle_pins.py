"""Closed-form pins of the oracle's per-cell stages (oracle/efs_oracle.cpp).

The GPU is tested bit-exact against these oracle functions, so each one is checked here against
what the reference's code fixes by hand computation: the cells a stage touches, the axis and sign
of each term, the neighbour rule, the number of passes and the half-cell conventions.  Each
expected value is a literal or a closed form (an affine field, an integer shift), chosen so that a
plausible mistake (a wrong threshold, axis, sign, offset or pass count, a dropped guard) fails.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle2D, Oracle3D, default_params


def interior3(shape):
    D, H, W = shape
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


# ---------------------------------------------------------------- vel_decay (CPUSimulator3D.cpp:54-59)
@pytest.mark.parametrize("dbl", [False, True])
def test_vel_decay_scales_interior_air_cells_only(dbl):
    """vel *= VEL_DECAY on interior cells with rho < 0.5 (solid cells included: the reference
    does not test the boundary); rho = 0.5 exactly is not decayed; border cells are untouched."""
    W, H, D = 9, 8, 7
    o = Oracle3D(W, H, D, params=default_params(vel_decay=0.5), dbl=dbl)
    rng = np.random.default_rng(10)
    levels = np.array([0.0, 0.25, 0.4999, 0.5, 0.75, 1.0])
    rho = levels[rng.integers(0, len(levels), (D, H, W))]
    rho[3, 3, 3:5] = (0.0, 0.75)  # interior solids: one below and one above the threshold
    o.field("rho")[...] = rho
    o.field("boundary")[3, 3, 3:5] = 1.0
    v0 = rng.integers(-64, 64, (D, H, W, 3)).astype(np.float64)  # exact halves
    o.field("vel")[...] = v0
    o.stage(0)
    decayed = interior3((D, H, W)) & (rho < 0.5)
    expect = np.where(decayed[..., None], v0 * 0.5, v0)
    np.testing.assert_array_equal(o.field("vel"), expect)
    np.testing.assert_array_equal(o.field("vel")[3, 3, 3], v0[3, 3, 3] * 0.5)  # solid, rho 0: decayed
    np.testing.assert_array_equal(o.field("vel")[3, 3, 4], v0[3, 3, 4])  # solid, rho 0.75: kept
    # 2D (CPUSimulator2D.cpp:65-70)
    o2 = Oracle2D(W, H, params=default_params(vel_decay=0.5), dbl=dbl)
    rho2 = rho[0]
    o2.field("rho")[...] = rho2
    v2 = v0[0, ..., :2].copy()
    o2.field("vel")[...] = v2
    o2.stage(0)
    m2 = np.zeros((H, W), bool)
    m2[1:-1, 1:-1] = True
    m2 &= rho2 < 0.5
    np.testing.assert_array_equal(o2.field("vel"), np.where(m2[..., None], v2 * 0.5, v2))


# ---------------------------------------------------------------- vel_extrapolate (:61-111)
SENT = np.array([7.0, 7.0, 7.0])


def _line3d(W):
    """A 1-cell-thick line of interior cells along x (H = D = 3): every y/z neighbour is border."""
    o = Oracle3D(W, 3, 3, dbl=True)
    o.field("vel")[...] = SENT
    return o


def test_vel_extrapolate_six_passes_one_ring_each():
    """From one valid cell the velocity spreads exactly one cell per pass, 6 passes in all
    (3 ping-pong double passes, VEL_EXTRAPOLATE_ITER is ignored)."""
    o = _line3d(20)
    a = np.array([1.0, -2.0, 0.5])
    o.field("rho")[1, 1, 2] = 1.0
    o.field("vel")[1, 1, 2] = a
    o.stage(1)
    v = o.field("vel")[1, 1]
    for x in range(1, 9):  # x = 1 (1 pass) and x = 3..8 (1..6 passes)
        np.testing.assert_array_equal(v[x], a, err_msg=f"x={x}")
    for x in range(9, 19):
        np.testing.assert_array_equal(v[x], SENT, err_msg=f"x={x}")
    np.testing.assert_array_equal(v[0], SENT)  # border untouched
    np.testing.assert_array_equal(v[19], SENT)


def test_vel_extrapolate_vec3_mean_of_valid_neighbours():
    """Two sources a (x=5) and b (x=9): x=7 becomes valid in pass 2 with both neighbours valid,
    so it takes the vec3 mean (a + b) / 2 = (2, 0, -1.5); validity is rho >= 0.5 (0.5 counts),
    a solid cell (x=12) is never filled and blocks the spread."""
    o = _line3d(16)
    a = np.array([1.0, -2.0, 0.5])
    b = np.array([3.0, 2.0, -3.5])
    o.field("rho")[1, 1, 5] = 0.5  # exactly the threshold: valid
    o.field("rho")[1, 1, 9] = 1.0
    o.field("rho")[1, 1, 11] = 0.4999  # not valid: gets b from x=10 in pass 2
    o.field("vel")[1, 1, 5] = a
    o.field("vel")[1, 1, 9] = b
    o.field("boundary")[1, 1, 12] = 1.0
    o.stage(1)
    v = o.field("vel")[1, 1]
    np.testing.assert_array_equal(v[7], [2.0, 0.0, -1.5])
    for x in (1, 2, 3, 4, 6):
        np.testing.assert_array_equal(v[x], a, err_msg=f"x={x}")
    for x in (8, 10, 11):
        np.testing.assert_array_equal(v[x], b, err_msg=f"x={x}")
    np.testing.assert_array_equal(v[12], SENT)  # solid: velocity kept
    np.testing.assert_array_equal(v[13], SENT)  # behind the solid: unreachable
    np.testing.assert_array_equal(v[14], SENT)


def test_vel_extrapolate_3d_manhattan_ball():
    """In 3D the valid region after 6 passes is the Manhattan ball of radius 6."""
    n = 17
    o = Oracle3D(n, n, n, dbl=True)
    o.field("vel")[...] = SENT
    a = np.array([0.25, -1.0, 2.0])
    o.field("rho")[8, 8, 8] = 1.0
    o.field("vel")[8, 8, 8] = a
    o.stage(1)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    dist = np.abs(i - 8) + np.abs(j - 8) + np.abs(k - 8)
    inner = interior3((n, n, n))
    v = o.field("vel")
    assert np.all(v[inner & (dist <= 6)] == a)
    assert np.all(v[inner & (dist > 6)] == SENT)


def test_vel_extrapolate_2d_line():
    o = Oracle2D(16, 3, dbl=True)
    o.field("vel")[...] = 7.0
    a, b = np.array([1.0, -2.0]), np.array([3.0, 4.0])
    o.field("rho")[1, 5] = 1.0
    o.field("rho")[1, 9] = 1.0
    o.field("vel")[1, 5] = a
    o.field("vel")[1, 9] = b
    o.stage(1)
    v = o.field("vel")[1]
    np.testing.assert_array_equal(v[7], [2.0, 1.0])
    np.testing.assert_array_equal(v[1], a)
    np.testing.assert_array_equal(v[14], b)


# ---------------------------------------------------------------- vel_advect (:478-504)
def test_vel_advect_uniform_field_is_identity():
    W, H, D = 16, 14, 12
    for dbl, tol in ((True, 1e-13), (False, 2e-6)):
        o = Oracle3D(W, H, D, params=default_params(dt=0.4), dbl=dbl)
        o.field("vel")[...] = [1.5, -2.25, 0.75]
        o.stage(4)
        v = o.field("vel")[interior3((D, H, W))]
        np.testing.assert_allclose(v, np.broadcast_to([1.5, -2.25, 0.75], v.shape), rtol=0, atol=tol)


def test_vel_advect_linear_field_shifts_exactly():
    """u = U everywhere, v = 0, w stored at x-coordinate (i - 0.5) as c * (i - 0.5): the trace goes
    back by U dt along x (the midpoint rule integrates the linear-in-time w exactly), the sample at
    prev - 0.5 reads c * (i - U dt - 0.5), and the z-face average of two equal cells keeps it.  So
    w_new = c * (i - 0.5 - U dt): the field shifts by U dt.  Dropping the -0.5 sample shift moves
    every value by c / 2."""
    W, H, D = 24, 12, 24
    U, c, dt = 1.5, 0.1, 0.5
    o = Oracle3D(W, H, D, params=default_params(dt=dt), dbl=True)
    v = o.field("vel")
    i = np.arange(W)[None, None, :]
    v[..., 0] = U
    v[..., 1] = 0.0
    v[..., 2] = c * (i - 0.5)
    o.stage(4)
    v = o.field("vel")
    sl = (slice(5, D - 5), slice(3, H - 3), slice(4, W - 3))
    expect_w = np.broadcast_to(c * (np.arange(W)[4:W - 3] - 0.5 - U * dt), v[sl][..., 2].shape)
    np.testing.assert_allclose(v[sl][..., 2], expect_w, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v[sl][..., 0], U, rtol=0, atol=1e-12)
    np.testing.assert_allclose(v[sl][..., 1], 0.0, rtol=0, atol=1e-12)


def test_vel_advect_2d_uniform_identity():
    o = Oracle2D(14, 12, params=default_params(dt=0.4), dbl=True)
    o.field("vel")[...] = [1.5, -2.25]
    o.stage(4)
    np.testing.assert_allclose(o.field("vel")[1:-1, 1:-1], np.broadcast_to([1.5, -2.25], (10, 12, 2)), atol=1e-13)


# ---------------------------------------------------------------- add_force (:458-476)
@pytest.mark.parametrize("dbl", [False, True])
def test_add_force_gravity_is_minus_y_on_every_non_solid_interior_cell(dbl):
    W, H, D = 8, 7, 6
    o = Oracle3D(W, H, D, params=default_params(gravity_force=10.0, dt=0.0625), dbl=dbl)
    rng = np.random.default_rng(11)
    v0 = rng.integers(-32, 32, (D, H, W, 3)).astype(np.float64)
    o.field("vel")[...] = v0
    o.field("rho")[2:4, 2:4, 2:4] = 1.0  # fluid and air both receive gravity
    o.field("boundary")[2, 4, 4] = 1.0
    o.stage(5)
    hit = interior3((D, H, W)) & (o.field("boundary") != 1)
    expect = v0.copy()
    expect[hit, 1] -= 0.625  # 10 * 0.0625, along -y only
    np.testing.assert_array_equal(o.field("vel"), expect)


def test_add_force_adds_forces_times_dt_to_all_components():
    W, H, D = 8, 7, 6
    o = Oracle3D(W, H, D, params=default_params(gravity_enabled=0, dt=0.25), dbl=True)
    F = np.zeros((D, H, W, 3))
    F[3, 3, 3] = [4.0, -8.0, 2.0]
    F[0, 3, 3] = [4.0, 4.0, 4.0]  # border: ignored
    o.field("forces")[...] = F
    o.stage(5)
    v = o.field("vel")
    np.testing.assert_array_equal(v[3, 3, 3], [1.0, -2.0, 0.5])
    assert np.count_nonzero(v) == 3


def test_add_force_2d_is_face_wise():
    """2D (CPUSimulator2D.cpp:404-425): a face between two solids is zeroed; every other face gets
    F dt, and y-faces also -g dt."""
    W, H = 7, 6
    o = Oracle2D(W, H, params=default_params(gravity_force=10.0, dt=0.0625), dbl=True)
    o.field("vel")[...] = 1.0
    o.stage(5)
    v = o.field("vel")
    # interior face between two non-solid cells
    np.testing.assert_array_equal(v[3, 3], [1.0, 1.0 - 0.625])
    # x = 1 x-face sits between the border cell x = 0 and a fluid cell: not both solid, kept
    np.testing.assert_array_equal(v[3, 1], [1.0, 1.0 - 0.625])
    # y = 1 y-face: between border y = 0 and an interior cell: gets gravity
    assert v[1, 3, 1] == 1.0 - 0.625
    o2 = Oracle2D(W, H, params=default_params(gravity_force=10.0, dt=0.0625), dbl=True)
    o2.field("boundary")[3, 2:4] = 1.0  # solids (y=3, x=2) and (y=3, x=3): adjacent along x
    o2.field("boundary")[2, 5] = 1.0    # solids (y=2, x=5) and (y=3, x=5): adjacent along y
    o2.field("boundary")[3, 5] = 1.0
    o2.field("vel")[...] = 1.0
    o2.stage(5)
    v2 = o2.field("vel")
    assert v2[3, 3, 0] == 0.0          # x-face between solids (3,2) and (3,3)
    assert v2[3, 5, 1] == 0.0          # y-face between solids (2,5) and (3,5)
    np.testing.assert_array_equal(v2[3, 4], [1.0, 1.0 - 0.625])  # (3,4) fluid, its -x neighbour solid: kept


# ---------------------------------------------------------------- pressure gradient (:550-563)
def test_gradient_linear_pressure_constant_decrement_with_guards():
    """p = 3x - 2y + 5z: non-solid interior cells lose (3, -2, 5), except that x == 1 keeps its
    x-component (the x > 1 guard), likewise y == 1 and z == 1; solid cells are skipped; a cell
    next to a solid still uses the solid's pressure (no neighbour test)."""
    W, H, D = 8, 7, 6
    o = Oracle3D(W, H, D, dbl=True)
    k, j, i = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    p = 3.0 * i - 2.0 * j + 5.0 * k
    p[3, 3, 4] = 100.0
    o.field("boundary")[3, 3, 4] = 1.0
    o.field("pressure")[...] = p
    o.gradient()
    v = o.field("vel")
    assert list(v[3, 3, 3]) == [-3.0, 2.0, -5.0]
    assert list(v[2, 2, 1]) == [0.0, 2.0, -5.0]
    assert list(v[2, 1, 2]) == [-3.0, 0.0, -5.0]
    assert list(v[1, 2, 2]) == [-3.0, 2.0, 0.0]
    assert list(v[1, 1, 1]) == [0.0, 0.0, 0.0]
    assert list(v[3, 3, 4]) == [0.0, 0.0, 0.0]  # the solid itself
    # (x=5, y=3, z=3): its -x neighbour is the solid with p = 100; p(5) = 15 - 6 + 15 = 24
    assert list(v[3, 3, 5]) == [-(24.0 - 100.0), 2.0, -5.0]
    border = ~interior3((D, H, W))
    assert not v[border].any()


def test_gradient_2d_face_wise_zeroes_faces_touching_solids():
    W, H = 8, 7
    o = Oracle2D(W, H, dbl=True)
    j, i = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    o.field("pressure")[...] = 3.0 * i - 2.0 * j
    o.field("boundary")[3, 4] = 1.0
    o.field("vel")[...] = 10.0
    o.gradient()
    v = o.field("vel")
    assert list(v[2, 2]) == [7.0, 12.0]
    assert list(v[2, 1]) == [0.0, 12.0]   # x-face at x = 1 touches the border solid
    assert list(v[1, 2]) == [7.0, 0.0]
    assert list(v[3, 4]) == [0.0, 0.0]    # the solid cell's faces
    assert list(v[3, 5]) == [0.0, 12.0]   # -x neighbour is solid
    assert list(v[4, 4]) == [7.0, 0.0]    # -y neighbour is solid


# ---------------------------------------------------------------- wall enforcement
def test_walls_3d_one_sided_2d_two_sided():
    """3D (CPUSimulator3D.cpp:657-679) zeroes only velocity pointing into a solid; 2D
    (CPUSimulator2D.cpp:597-607) zeroes every face touching one."""
    o = Oracle3D(6, 6, 6, dbl=True)
    v = o.field("vel")
    v[2, 2, 1] = [-1.0, 1.0, 1.0]   # u at the -x wall points into it: zeroed
    v[2, 2, 2] = [-1.0, 1.0, 1.0]   # interior: kept
    v[2, 2, 5] = [1.0, 1.0, 1.0]    # cell x = 5 is border (solid): u > 0 zeroed
    v[2, 3, 3] = [1.0, 1.0, 1.0]
    o.enforce_walls()
    assert list(v[2, 2, 1]) == [0.0, 1.0, 1.0]
    assert list(v[2, 2, 2]) == [-1.0, 1.0, 1.0]
    assert list(v[2, 2, 5]) == [0.0, 0.0, 0.0]
    assert list(v[2, 3, 3]) == [1.0, 1.0, 1.0]
    o2 = Oracle2D(6, 6, dbl=True)
    v2 = o2.field("vel")
    v2[...] = 1.0
    o2.enforce_walls()
    assert list(v2[2, 1]) == [0.0, 1.0]  # x-face touching the -x border, pointing away
    assert list(v2[2, 2]) == [1.0, 1.0]
    assert list(v2[1, 2]) == [1.0, 0.0]


# ---------------------------------------------------------------- rho_sharpen (:366-456)
def _sharpen_line(dbl, dt, profile, H=7, D=7):
    """rho constant over a 5 x 5 interior cross-section, varying along x."""
    W = len(profile) + 2
    o = Oracle3D(W, H, D, params=default_params(dt=dt), dbl=dbl)
    r = np.zeros((D, H, W))
    r[1:-1, 1:-1, 1:-1] = np.array(profile)[None, None, :]
    o.field("rho")[...] = r
    return o


def test_sharpen_delta_hand_computed_profile():
    """dt = 0.25 (delta t = 3 dt = 0.75).  At the middle row (y = z = 3) only x varies:
      x=1 rho 1.0  > 0.5                                  -> 0
      x=2 rho 0.46 max |diff| 0.54 >= 0.2, weight 0       -> 0
      x=3 rho 0.45 nbrs 0.46/0.40: w = (-0.05)^3 (1 - 0.05/0.2) = -9.375e-5,
                   max_dec = 0.75 * 0.05 = 0.0375         -> -3.515625e-6
      x=4 rho 0.40 nbrs 0.45/0.35: w = -7.5e-4, max_dec 0.0375 -> -2.8125e-5
      x=5 rho 0.35 nbrs 0.40/0.20: w = (-0.15)^3 * 0.25, max_dec 0.75 * 0.15 -> -9.4921875e-5
      x=6 rho 0.20 nbrs 0.35/2e-5: w = (-0.3)^3 (1 - 0.19998/0.2f), max_dec 0.75 * 0.19998
                   where 0.2f is the reference's float literal 0.20000000298 (the weight
                   cancels, so it shows: with the double 0.2 the delta is -4.049595e-7)
                                                          -> -4.0501984e-7
      x=8 rho 5e-6 < 1e-5                                 -> -rho = -5e-6
    """
    prof = [1.0, 0.46, 0.45, 0.40, 0.35, 0.20, 2e-5, 5e-6, 0.0]
    o = _sharpen_line(True, 0.25, prof)
    d = o.sharpen_delta()[3, 3]
    expect = {1: 0.0, 2: 0.0, 3: -3.515625e-6, 4: -2.8125e-5, 5: -9.4921875e-5, 6: -4.0501984e-7, 8: -5e-6}
    for x, e in expect.items():
        assert d[x] == pytest.approx(e, rel=1e-6, abs=1e-15), (x, d[x], e)
    # FP32 instance: same values to float rounding (the weight of x=6 cancels catastrophically)
    d32 = _sharpen_line(False, 0.25, prof).sharpen_delta()[3, 3]
    for x in (3, 4, 5):
        assert d32[x] == pytest.approx(expect[x], rel=2e-4)


def test_sharpen_delta_isolated_cell_and_negative_clamp():
    """A lone cell rho = 1e-4 among zeros: every axis contributes (0.75 * 1e-4)^2 to max_dec, so
    max_dec = sqrt(3) * 7.5e-5 and delta = (1e-4 - 0.5)^3 (1 - 1e-4 / 0.2) max_dec = -1.62201e-5.
    With dt = 2 (delta t 6) the decrease exceeds rho and is clamped to -rho."""
    for dt, e in ((0.25, -1.6220121e-5), (2.0, -1e-4)):
        o = Oracle3D(9, 9, 9, params=default_params(dt=dt), dbl=True)
        o.field("rho")[4, 4, 4] = 1e-4
        d = o.sharpen_delta()
        assert d[4, 4, 4] == pytest.approx(e, rel=1e-6)
        d[4, 4, 4] = 0
        assert not d.any()
    # 2D: two axes, sqrt(2)
    o2 = Oracle2D(9, 9, params=default_params(dt=0.25), dbl=True)
    o2.field("rho")[4, 4] = 1e-4
    assert o2.sharpen_delta()[4, 4] == pytest.approx(-1.3243674e-5, rel=1e-6)


def test_trace_along_field_climbs_gradient_until_half():
    """rho = 0.1 (10 - x) along x (constant across y, z); rho_d = 3 gives steps of 0.3 along -x
    (the gradient direction).  ScalField3D::interpolate puts cell i at i + 0.5, so the
    interpolated rho at x is 0.1 (10.5 - x), which reaches 0.5 at x <= 5.5: from x = 8 the trace
    stops after 9 steps at 8 - 2.7 = 5.3."""
    prof = [max(0.0, 0.1 * (10 - x)) for x in range(1, 13)]
    o = _sharpen_line(True, 0.25, prof)
    o.params.rho_d = 3.0
    o.set_params(o.params)
    end = o.trace_field((8.0, 3.0, 3.0))
    np.testing.assert_allclose(end, [5.3, 3.0, 3.0], atol=1e-12)
    # a start already at rho >= 0.5 does not move
    np.testing.assert_array_equal(o.trace_field((4.0, 3.0, 3.0)), [4.0, 3.0, 3.0])


def test_scatter_values_trilinear_weights_with_solid_renormalisation():
    o = Oracle3D(9, 9, 9, dbl=True)
    o.scatter((4.25, 3.5, 3.0), 1.0)
    r = o.field("rho")
    # corners at floor(pos) and +1; fx 0.25, fy 0.5, fz 0
    assert r[3, 3, 4] == 0.375 and r[3, 3, 5] == 0.125 and r[3, 4, 4] == 0.375 and r[3, 4, 5] == 0.125
    assert r.sum() == 1.0
    o2 = Oracle3D(9, 9, 9, dbl=True)
    o2.field("boundary")[3, 4, 5] = 1.0
    o2.scatter((4.25, 3.5, 3.0), 1.0)
    r2 = o2.field("rho")
    assert r2[3, 4, 5] == 0.0
    assert r2[3, 3, 4] == pytest.approx(0.375 / 0.875, rel=1e-15)
    assert r2[3, 3, 5] == pytest.approx(0.125 / 0.875, rel=1e-15)
    assert r2.sum() == pytest.approx(1.0, rel=1e-15)


# ---------------------------------------------------------------- 2D interpolation (ScalField2D.cpp:21)
def test_scalar_interpolation_2d_has_plus_one_shift():
    """ScalField2D::interpolate adds 1 to the position before sampling (ScalField2D.cpp:21), so an
    affine field with cell i at i + 0.5 reads f(pos + 1)."""
    W, H = 12, 10
    o = Oracle2D(W, H, dbl=True)
    j, i = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    o.field("gamma")[...] = 2.0 + 0.5 * (i + 0.5) - 0.25 * (j + 0.5)
    rng = np.random.default_rng(12)
    for _ in range(40):
        p = rng.uniform([0.0, 0.0], [W - 3.0, H - 3.0])
        assert o.sinterp("gamma", p) == pytest.approx(2.0 + 0.5 * (p[0] + 1) - 0.25 * (p[1] + 1), abs=1e-12)


def test_trace_2d_uniform_velocity():
    o = Oracle2D(16, 16, dbl=True)
    o.field("vel")[..., 0] = 3.0
    o.field("vel")[..., 1] = -2.0
    np.testing.assert_allclose(o.trace((7.0, 6.5), -0.5), [5.5, 7.5], atol=1e-12)
    r = o.trace((2.0, 8.0), -2.0)
    assert 1.0 <= r[0] < 2.0
