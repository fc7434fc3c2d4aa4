"""Pins of the oracle's level transfers (oracle/amr.py) and of its inflow /
wedge ghost fill (oracle/driver.py fill_ghosts):

* ghost fill: the known answers of the reference's own boundary test
  (HyperFlow_V3/tests/boundary_check.cpp:15-36: an 8 x 4 domain, value
  1000 n + 10 j + i + 1, fixed inflow left and top, outflow right, a wedge
  bottom whose tip lies on the x = 0 face, inlet 20000 + n);
* interpolation: a uniform field is reproduced; a linear field is reproduced
  exactly at the fine centres; every parent's four children average back to
  it; on trace-species fields whose neighbours push minmod slopes past the
  parent's density, theta is the closed form max{2^-k : q - 2^-k (sx + sy)/4
  >= 0}; an inadmissible parent is reported;
* average-down: exact sums of integer-valued fields, mass conservation, and
  the round trip average_down(interp(C)) == C to roundoff.
"""
import numpy as np
import pytest

from oracle.amr import average_down, interp
from oracle.driver import INFLOW, PERIODIC, REFLECTING, TRANSMISSIVE, WEDGE, fill_ghosts

R = 8314.462618


def test_boundary_check_known_answers():
    ns, nx, ny = 9, 8, 4
    nv = ns + 3
    value = lambda i, j, n: 1000.0 * n + 10.0 * j + i + 1.0
    G = np.full((nv, ny + 4, nx + 4), -1e100)
    for n in range(nv):
        for j in range(ny):
            for i in range(nx):
                G[n, j + 2, i + 2] = value(i, j, n)
    inlet = 20000.0 + np.arange(nv)
    # prob lo x = -0.002, dx = 0.001, wedge start 0: the wall starts at column 2
    fill_ghosts(G, ns, (INFLOW, TRANSMISSIVE, WEDGE, INFLOW), False, inlet, wedge_i=2)
    a = lambda i, j, n: G[n, j + 2, i + 2]
    my = ns + 1
    for n in range(nv):
        assert a(-1, 1, n) == inlet[n] and a(3, 4, n) == inlet[n] and a(-1, -1, n) == inlet[n]
        assert a(8, 2, n) == value(7, 2, n)
        assert a(0, -1, n) == value(0, 0, n) and a(0, -2, n) == value(0, 0, n)
        sign = -1.0 if n == my else 1.0
        assert a(3, -1, n) == sign * value(3, 0, n) and a(3, -2, n) == sign * value(3, 1, n)
    # every ghost cell is written
    assert np.all(G > -1e99)


def test_flow_fill_corners_and_clamp():
    """Corners beyond an inflow side are inlet; a wedge-bottom corner beyond a
    non-inflow side reads the clamped column; plain corners follow filcc."""
    rng = np.random.default_rng(4)
    ns, nx, ny = 9, 5, 4
    nv = ns + 3
    G = rng.standard_normal((nv, ny + 4, nx + 4))
    inlet = rng.standard_normal(nv)
    fill_ghosts(G, ns, (REFLECTING, INFLOW, WEDGE, TRANSMISSIVE), False, inlet, wedge_i=1)
    a = lambda i, j: G[:, j + 2, i + 2]
    my = ns + 1
    for g in (-1, -2):
        # x-lo reflecting + wedge bottom (column -1 < wedge_i: outflow from row 0, column clamped to 0)
        assert np.array_equal(a(-1, g), a(0, 0)) and np.array_equal(a(-2, g), a(0, 0))
        # x-hi inflow corners
        assert np.array_equal(a(nx, g), inlet) and np.array_equal(a(nx + 1, ny - 1 - g), inlet)
        # wall columns mirror
        w = a(3, g).copy()
        w[my] *= -1
        assert np.array_equal(w, a(3, -g - 1))
    # x-lo reflecting + y-hi transmissive corner: the y rule on the filled x-ghost column
    assert np.array_equal(a(-1, ny), a(-1, ny - 1)) and np.array_equal(a(-2, ny + 1), a(-2, ny - 1))


def state(o, T, p, Y, u=0.0, v=0.0):
    """Conserved state of an ideal gas at (T, p, Y, u, v) (BlockAMR.cpp:20-24)."""
    Y = np.asarray(Y, dtype=float)
    rho = p / (R * T * np.sum(Y / o.m["W"]))
    e, _ = o.e_cv(T, Y)
    return np.concatenate([rho * Y, [rho * u, rho * v, rho * (e + 0.5 * (u * u + v * v))]])


Y_FRESH = np.array([0.028, 0.226, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.746])
Y_BURNT = np.array([0.002, 0.01, 0.22, 2e-4, 1.5e-3, 1.5e-2, 2e-5, 2e-6, 0.7515])


def smooth_field(o, nx, ny, seed):
    """An admissible H2/air field with steep but smooth T, p, Y and velocity variation."""
    rng = np.random.default_rng(seed)
    C = np.zeros((o.ns + 3, ny, nx))
    for j in range(ny):
        for i in range(nx):
            f = 0.5 + 0.5 * np.sin(0.9 * i + 0.6 * j + rng.uniform(-0.3, 0.3))
            Y = (1 - f) * Y_FRESH + f * Y_BURNT
            C[:, j, i] = state(o, 900.0 + 1400.0 * f, 1e5 * (1 + 0.5 * f), Y / Y.sum(),
                               80.0 * np.cos(0.5 * i), 40.0 * np.sin(0.7 * j))
    return C


def test_interp_uniform_and_linear(oracle_h2):
    o = oracle_h2
    q = state(o, 1500.0, 2e5, [0.01, 0.2, 0.1, 1e-4, 1e-3, 1e-2, 1e-5, 1e-6, 0.67889], 100.0, -40.0)
    C = np.repeat(np.repeat(q[:, None, None], 6, 1), 7, 2)
    F, theta, failed = interp(o, C, (0, 0), (2, 2, 9, 7))
    assert failed.size == 0 and np.all(theta == 1.0)
    assert np.array_equal(F, np.broadcast_to(q[:, None, None], F.shape))
    # a linear field in every component: minmod keeps the slope, the children
    # are the field at the fine centres ((I + 1/2)/2 - 1/2 in coarse units)
    ii, jj = np.meshgrid(np.arange(7), np.arange(6))
    L = q[:, None, None] * (1 + 0.01 * ii[None] + 0.02 * jj[None])
    F, theta, failed = interp(o, L, (0, 0), (2, 2, 11, 9))
    assert failed.size == 0 and np.all(theta == 1.0)
    fi, fj = np.meshgrid(np.arange(2, 12), np.arange(2, 10))
    exact = q[:, None, None] * (1 + 0.01 * ((fi + 0.5) / 2 - 0.5)[None] + 0.02 * ((fj + 0.5) / 2 - 0.5)[None])
    assert np.allclose(F, exact, rtol=1e-14, atol=0)


def test_interp_children_average_to_parent(oracle_h2):
    o = oracle_h2
    C = smooth_field(o, 12, 10, 1)
    F, theta, failed = interp(o, C, (0, 0), (2, 2, 21, 17))
    assert failed.size == 0
    mean = average_down(F, (2, 2), (1, 1, 10, 8))
    assert np.allclose(mean, C[:, 1:9, 1:11], rtol=2e-16 * 4, atol=1e-300)
    # odd region bounds: only the children inside are written, same values
    F2, theta2, _ = interp(o, C, (0, 0), (3, 2, 20, 16))
    assert np.array_equal(F2, F[:, 0:15, 1:19]) and np.array_equal(theta2, theta)


def limited_field(o, seed=2):
    """Trace H (index 3) densities whose left / lower neighbours are negative:
    minmod slopes s > 2 q push the (-,-) child negative, everything else stays
    admissible; returns the field and the closed-form theta of each parent."""
    rng = np.random.default_rng(seed)
    base = state(o, 1400.0, 1.5e5, [0.02, 0.2, 0.05, 0.0, 1e-3, 1e-2, 1e-5, 1e-6, 0.71889], 20.0, 10.0)
    nx = ny = 15
    C = np.repeat(np.repeat(base[:, None, None], ny, 1), nx, 2)
    h = 1e-6 * base[:o.ns].sum()
    theta = {}
    for j in range(1, ny - 1, 3):                       # parents three apart: no shared neighbours
        for i in range(1, nx - 1, 3):
            q = h * rng.uniform(0.5, 2.0)
            kx, ky = np.exp(rng.uniform(np.log(0.2), np.log(60.0), size=2))   # slopes kx q, ky q after minmod
            C[3, j, i] = q
            C[3, j, i - 1] = q - kx * q
            C[3, j, i + 1] = q + kx * q * 1.5
            C[3, j - 1, i] = q - ky * q
            C[3, j + 1, i] = q + ky * q * 1.5
            sx, sy = min(kx * q, 1.5 * kx * q), min(ky * q, 1.5 * ky * q)
            t = 1.0
            while q - t * (0.25 * sx + 0.25 * sy) < 0:
                t *= 0.5
            theta[(i, j)] = t
    return C, theta


def test_interp_theta_closed_form(oracle_h2):
    o = oracle_h2
    C, expect = limited_field(o)
    # region = the parent's own children
    for (i, j), t in expect.items():
        F, theta, failed = interp(o, C, (0, 0), (2 * i, 2 * j, 2 * i + 1, 2 * j + 1))
        assert failed.size == 0
        assert theta[0, 0] == t, (i, j, theta[0, 0], t)
    assert min(expect.values()) <= 1 / 8 and any(t == 1 for t in expect.values())


def test_interp_inadmissible_parent(oracle_h2):
    o = oracle_h2
    C = smooth_field(o, 8, 8, 3)
    C[4, 3, 5] = -1e-3 * C[:o.ns, 3, 5].sum()             # parent (5, 3) holds negative O
    F, theta, failed = interp(o, C, (0, 0), (2, 2, 13, 13))
    # parents of region: i, j in 1..6; (5, 3) is row 2, column 4
    assert list(failed) == [2 * 6 + 4]
    assert np.isnan(F[:, 4:6, 8:10]).all() and not np.isnan(np.delete(F, np.s_[8:10], axis=2)).any()


def test_average_down_exact_and_conservative():
    rng = np.random.default_rng(7)
    F = rng.integers(-1000, 1000, size=(12, 10, 14)).astype(float)
    Cd = average_down(F, (4, 6), (2, 3, 8, 7))
    for n in (0, 5, 11):
        for j in range(3, 8):
            for i in range(2, 9):
                s = sum(F[n, 2 * j + b - 6, 2 * i + a - 4] for b in (0, 1) for a in (0, 1))
                assert Cd[n, j - 3, i - 2] == 0.25 * s
    assert np.array_equal(4 * Cd.sum(axis=(1, 2)), F.sum(axis=(1, 2)))


def composite_stage(o, seed=5, nxc=16, nyc=12, fine_box=(8, 6, 15, 13), dxc=1e-4, cfl=0.3):
    """One Euler stage of a two-level composite (the same h on both levels,
    BlockAMR.cpp:202-256): a periodic coarse level nxc x nyc and one interior
    fine box whose valid cells and two ghost layers come from the coarse level
    by conservative interpolation. Returns (coarse before, coarse after the
    stage, fine after the stage, coarse fluxes, fine fluxes, h, dxc, dyc)."""
    from oracle.driver import PERIODIC, fill_ghosts
    C = smooth_field(o, nxc, nyc, seed)
    ns, nv = o.ns, o.ns + 3
    dyc = dxc
    # coarse stage on the periodic level
    G = np.zeros((nv, nyc + 4, nxc + 4))
    G[:, 2:-2, 2:-2] = C
    fill_ghosts(G, ns, (PERIODIC,) * 4, False)
    rho = C[:ns].sum(0)
    _, a, _, _ = o.properties(C.reshape(nv, -1).T.copy())
    speed = (np.abs(C[ns] / rho).max() + np.abs(C[ns + 1] / rho).max() + 2 * a.max())
    h = cfl * 0.5 * dxc / speed
    r, C1, _, _, cfx, cfy = o.flux_stage(G, h / dxc, h / dyc, fluxes=True)
    assert r == 0
    # fine box: interpolation of the coarse level over the box grown by 2 (its ghosts)
    FI0, FJ0, FI1, FJ1 = fine_box
    Fg, _, failed = interp(o, G, (-2, -2), (FI0 - 2, FJ0 - 2, FI1 + 2, FJ1 + 2))
    assert failed.size == 0
    r, F1, _, _, ffx, ffy = o.flux_stage(Fg, 2 * h / dxc, 2 * h / dyc, fluxes=True)
    assert r == 0
    return C, C1, F1, (cfx, cfy, (0, 0)), (ffx, ffy, (FI0, FJ0)), h, dxc, dyc


def _composite_totals(C1, F1, fine_box, corrected, o, coarse, fine, h, dxc, dyc):
    from oracle.amr import reflux
    C2 = C1.copy()
    if corrected:
        reflux(C2, (0, 0), coarse, fine, fine_box, 15, h, dxc, dyc)
    FI0, FJ0, FI1, FJ1 = fine_box
    C2[:, FJ0 // 2:FJ1 // 2 + 1, FI0 // 2:FI1 // 2 + 1] = average_down(F1, (FI0, FJ0),
                                                                      (FI0 // 2, FJ0 // 2, FI1 // 2, FJ1 // 2))
    return C2.sum(axis=(1, 2))


def test_reflux_composite_conservation(oracle_h2):
    o = oracle_h2
    box = (8, 6, 15, 13)
    C, C1, F1, coarse, fine, h, dxc, dyc = composite_stage(o, fine_box=box)
    before = C.sum(axis=(1, 2))
    scale = np.abs(C).sum(axis=(1, 2))
    with_reflux = _composite_totals(C1, F1, box, True, o, coarse, fine, h, dxc, dyc)
    without = _composite_totals(C1, F1, box, False, o, coarse, fine, h, dxc, dyc)
    err = np.abs(with_reflux - before) / scale
    miss = np.abs(without - before) / scale
    assert err.max() < 1e-14, err
    # the correction is material: without it the composite loses conservation
    assert miss.max() > 1e-9, miss
    # a flipped sign (the correction applied to the wrong side) is caught too
    ffx, ffy, fl = fine
    flipped = _composite_totals(C1, F1, box, True, o, coarse, (-ffx, -ffy, fl), h, dxc, dyc)
    assert (np.abs(flipped - before) / scale).max() > 1e-9


def _ghosted(Q, ns):
    from oracle.driver import TRANSMISSIVE, fill_ghosts
    G = np.zeros((Q.shape[0], Q.shape[1] + 4, Q.shape[2] + 4))
    G[:, 2:-2, 2:-2] = Q
    fill_ghosts(G, ns, (TRANSMISSIVE,) * 4, False)
    return G


def step_field(o, nx, ny, left, right, k):
    """Columns i < k hold state left, the others state right."""
    Q = np.zeros((o.ns + 3, ny, nx))
    Q[:, :, :k] = left[:, None, None]
    Q[:, :, k:] = right[:, None, None]
    return Q


def fresh_Y(o):
    W = o.m["W"]
    X = np.zeros(o.ns)
    X[0], X[1], X[8] = 2.0, 1.0, 3.76                    # H2, O2, N2 (H2_Li order)
    Y = X * W
    return Y / Y.sum()


def test_tag_cells_closed_forms(oracle_h2):
    from oracle.amr import tag_cells
    o = oracle_h2
    Y = fresh_Y(o)
    nx, ny, k = 10, 6, 4
    a = state(o, 300.0, 101325.0, Y)
    # uniform gas: nothing
    assert not tag_cells(o, _ghosted(step_field(o, nx, ny, a, a, k), o.ns), 1e-3, 1e-3).any()
    # 10 % pressure step at fixed T: exactly the two columns beside it; 5 %: none
    for ratio, cols in ((1.10, {k - 1, k}), (1.05, set())):
        b = state(o, 300.0, 101325.0 * ratio, Y)
        t = tag_cells(o, _ghosted(step_field(o, nx, ny, a, b, k), o.ns), 1e-3, 1e-3)
        assert set(np.flatnonzero(t.any(axis=0))) == cols and (t.all(axis=0) == t.any(axis=0)).all()
    # temperature step at fixed p: 10 % tags the two columns, 5 % none
    for ratio, cols in ((1.10, {k - 1, k}), (1.05, set())):
        b = state(o, 300.0 * ratio, 101325.0, Y)
        t = tag_cells(o, _ghosted(step_field(o, nx, ny, a, b, k), o.ns), 1e-3, 1e-3)
        assert set(np.flatnonzero(t.any(axis=0))) == cols
    # geometry only: box x in [2, 4], y in [1, 1.6], moving at (1, 0), at t = 0.5 with unit cells:
    # the moved centres are (i, j + 0.5), so cells i = 2, 3, 4 of row 1
    t = tag_cells(o, _ghosted(step_field(o, nx, ny, a, a, k), o.ns), 1.0, 1.0, time=0.5, physics=False,
                  box=(2.0, 4.0, 1.0, 1.6), velocity=(1.0, 0.0))
    assert sorted(map(tuple, np.argwhere(t))) == [(1, 2), (1, 3), (1, 4)]


def test_tag_cells_fuel_rate(oracle_h2):
    """A threshold halfway between a cold and a hot state's fuel consumption
    rate (the oracle's kinetics) tags exactly the hot cells."""
    from oracle.amr import tag_cells
    o = oracle_h2
    W = o.m["W"]
    cold = state(o, 300.0, 101325.0, fresh_Y(o))
    Yh = np.array([0.02, 0.18, 0.06, 1e-4, 1e-3, 4e-3, 1e-5, 1e-6, 0.0])
    Yh[8] = 1.0 - Yh.sum()
    hot = state(o, 1500.0, 101325.0, Yh)
    rate = lambda q, T: -o.wdot(T, q[:o.ns] / W)[0] * W[0] / q[:o.ns].sum()
    r_cold, r_hot = rate(cold, 300.0), rate(hot, 1500.0)
    assert r_hot > 100 * max(r_cold, 1e-30)
    Q = step_field(o, 8, 4, cold, hot, 5)
    t = tag_cells(o, _ghosted(Q, o.ns), 1e-3, 1e-3, pressure_tag=1e9, temperature_tag=1e9,
                  fuel_rate_tag=0.5 * r_hot)
    assert set(np.flatnonzero(t.any(axis=0))) == {5, 6, 7} and (t.all(axis=0) == t.any(axis=0)).all()


def two_level_case(o, nx=40, ny=24, dx=1e-4):
    """Cold H2/air with a hot, high-pressure band that crosses the fine box's
    x-hi side (coarse cells 14..28 x 10..14); fine box = coarse 8..23 x 6..17."""
    Y = fresh_Y(o)
    cold = state(o, 300.0, 101325.0, Y)
    hot = state(o, 1800.0, 5 * 101325.0, Y)
    Q = np.repeat(np.repeat(cold[:, None, None], ny, 1), nx, 2)
    Q[:, 10:15, 14:29] = hot[:, None, None]
    return Q, dx, (16, 12, 47, 35)


def test_two_level_step_conserves(oracle_h2):
    """A periodic two-level Strang step (flow, chemistry, flow) keeps the
    composite's mass, momenta and energy to roundoff."""
    from oracle.driver import PERIODIC, OracleTwoLevelDriver
    o = oracle_h2
    Q, dx, box = two_level_case(o)
    d = OracleTwoLevelDriver(o, Q.shape[2], Q.shape[1], dx, dx, (PERIODIC,) * 4, box, cfl=0.5)
    d.set_state(Q)
    ns = o.ns
    tot = lambda: np.concatenate([[d.Qc[:ns].sum()], d.Qc[ns:].sum(axis=(1, 2))])
    before = tot()
    # momenta against the total mass moving at 1000 m/s (the gas starts at rest)
    mass = np.abs(Q[:ns]).sum()
    scale = np.array([mass, 1000.0 * mass, 1000.0 * mass, np.abs(Q[ns + 2]).sum()])
    d.step()
    d.step()
    after = tot()
    assert d.time > 0 and (np.abs(after - before) / scale).max() < 1e-13
    # the fine box carries its own solution: it differs from the interpolated start
    assert not np.allclose(d.Qf[ns + 2], Q[ns + 2, 6:18, 8:24].repeat(2, 0).repeat(2, 1))


def test_fine_box_for_tags_closed_forms():
    from oracle.driver import fine_box_for_tags
    t = np.zeros((24, 40), dtype=bool)
    assert fine_box_for_tags(t, 2, 4, 40, 24) is None
    t[10, 17] = True
    # grown to 15..19 x 8..12, blocks of 4: 12..19 x 8..15 -> fine 24..39 x 16..31
    assert fine_box_for_tags(t, 2, 4, 40, 24) == (24, 16, 39, 31)
    t[1, 39] = True                                        # near the corner: clamped 2 cells inside
    assert fine_box_for_tags(t, 2, 4, 40, 24) == (24, 4, 75, 31)


def test_two_level_regrid_conserves(oracle_h2):
    """Steps and regrids of the oracle's two-level driver keep the conserved
    integrals to roundoff (the remap defect check) and follow the hot band."""
    from oracle.driver import PERIODIC, OracleTwoLevelDriver
    o = oracle_h2
    Q, dx, box = two_level_case(o)
    d = OracleTwoLevelDriver(o, Q.shape[2], Q.shape[1], dx, dx, (PERIODIC,) * 4, box, cfl=0.5)
    d.set_state(Q)
    d.step()
    new_box, defect = d.regrid()
    assert new_box != box and defect < 1e-13
    d.step()


def test_keep_sides_closed_forms():
    """HF_BC_KEEP sides keep their (interpolated) ghosts; a ghost beyond a
    physical x side beside a kept y side takes the x rule; a physical y side
    applies its rule on every column, kept x ghosts included."""
    from oracle.driver import KEEP
    rng = np.random.default_rng(9)
    ns, nx, ny = 9, 6, 5
    G = rng.standard_normal((ns + 3, ny + 4, nx + 4))
    G0 = G.copy()
    fill_ghosts(G, ns, (KEEP, REFLECTING, KEEP, TRANSMISSIVE), False)
    a = lambda A, i, j: A[:, j + 2, i + 2]
    for j in range(-2, ny):                                # x-lo kept on every row below y-hi
        for i in (-2, -1):
            assert np.array_equal(a(G, i, j), a(G0, i, j))
    for j in (-2, -1):                                     # y-lo kept inside x; reflecting beyond x-hi
        for i in range(-2, nx):
            assert np.array_equal(a(G, i, j), a(G0, i, j))
        w = a(G0, nx - 1, j).copy()
        w[ns] *= -1
        assert np.array_equal(a(G, nx, j), w)
    for i in range(-2, nx + 2):                            # y-hi transmissive on every column
        assert np.array_equal(a(G, i, ny), a(G, i, ny - 1)) and np.array_equal(a(G, i, ny + 1), a(G, i, ny - 1))


def test_two_level_box_at_a_wall_conserves(oracle_h2):
    """A fine box on the x-lo wall of a closed (all-reflecting) box: two
    two-level Strang steps keep mass and energy to roundoff."""
    from oracle.driver import OracleTwoLevelDriver
    o = oracle_h2
    Q, dx, _ = two_level_case(o)
    Q[:, 10:15, 0:6] = Q[:, 10:15, 14:15]                 # hot gas against the wall too
    d = OracleTwoLevelDriver(o, Q.shape[2], Q.shape[1], dx, dx, (REFLECTING,) * 4, (0, 12, 47, 35), cfl=0.5)
    assert d.reflux_sides == 14 and d.fine_walls == 1
    d.set_state(Q)
    ns = o.ns
    before = np.array([d.Qc[:ns].sum(), d.Qc[ns + 2].sum()])
    d.step()
    d.step()
    after = np.array([d.Qc[:ns].sum(), d.Qc[ns + 2].sum()])
    assert (np.abs(after - before) / np.abs(before)).max() < 1e-13


def test_box_choice_matches_the_oracle_restatement():
    """hyperflow_amd.amr.choose_fine_box against oracle.driver.fine_box_for_tags
    on random tag patterns, buffers, block sizes and per-side margins."""
    from hyperflow_amd.amr import choose_fine_box
    from oracle.driver import fine_box_for_tags
    rng = np.random.default_rng(13)
    for _ in range(300):
        nx, ny = rng.integers(8, 60, size=2)
        t = rng.random((ny, nx)) < rng.choice([0.0, 0.002, 0.02, 0.2])
        buf, blk = int(rng.integers(0, 4)), int(rng.choice([1, 2, 4, 8]))
        m = tuple(int(v) for v in rng.integers(0, 3, size=4))
        assert choose_fine_box(t, buf, blk, nx, ny, m) == fine_box_for_tags(t, buf, blk, nx, ny, m)


def repair_case(o, seed=7):
    """Two-level arrays with roundoff-size negatives: in leaf coarse cells,
    in covered coarse cells (which the repair must leave to the
    average-down) and in fine cells; species 2 only in fine cells, so its
    deficit moves to the coarse level when the largest mass sits there."""
    Q, dx, box = two_level_case(o)
    rng = np.random.default_rng(seed)
    ns = o.ns
    oc = np.abs(Q) * (1 + 0.1 * rng.random(Q.shape))
    FI0, FJ0, FI1, FJ1 = box
    of = np.abs(rng.random((Q.shape[0], FJ1 - FJ0 + 1, FI1 - FI0 + 1))) * oc[:, 0:1, 0:1].mean()
    eps = np.finfo(float).eps
    for n, lev, j, i in ((1, 0, 0, 0), (1, 0, 3, 30), (2, 1, 4, 5), (2, 1, 9, 17), (5, 1, 0, 0), (5, 0, 23, 39)):
        A = oc if lev == 0 else of
        A[n, j, i] = -(2 + j % 5) * eps * max(oc[n].max(), of[n].max())
    # a covered coarse cell with a negative value: not a leaf, not repaired
    oc[3, FJ0 // 2 + 1, FI0 // 2 + 1] = -10 * eps * oc[3].max()
    return oc, of, dx, box


def test_two_level_repair_over_leaves(oracle_h2):
    """repairRoundoff over the leaf cells of both levels (BlockAMR.cpp:
    299-334): every negative leaf value is zeroed, the covered coarse cells
    are left alone, each species' integral over the leaves (value times cell
    volume) is unchanged to roundoff, and the deficit goes to the one cell
    of largest mass over both levels."""
    from oracle.driver import PERIODIC, OracleTwoLevelDriver
    o = oracle_h2
    oc, of, dx, box = repair_case(o)
    d = OracleTwoLevelDriver(o, oc.shape[2], oc.shape[1], dx, dx, (PERIODIC,) * 4, box, cfl=0.5)
    ns = o.ns
    FI0, FJ0, FI1, FJ1 = box
    leaf = np.ones(oc.shape[1:], dtype=bool)
    leaf[FJ0 // 2:FJ1 // 2 + 1, FI0 // 2:FI1 // 2 + 1] = False
    vc, vf = dx * dx, 0.25 * dx * dx
    integral = lambda C, F: (np.where(leaf, C[:ns], 0.0).sum(axis=(1, 2)) * vc + F[:ns].sum(axis=(1, 2)) * vf)
    before = integral(oc, of)
    covered = oc[:, ~leaf].copy()
    C, F = oc.copy(), of.copy()
    n_rep = d._repair(C, F)
    assert n_rep == 6
    assert (np.where(leaf, C[:ns], 0.0) >= 0).all() and (F[:ns] >= 0).all()
    np.testing.assert_array_equal(C[:, ~leaf], covered)
    after = integral(C, F)
    np.testing.assert_allclose(after, before, rtol=8 * np.finfo(float).eps, atol=0)
    # species 2's deficit (fine cells only) went to the largest mass over both levels
    changed_c = np.argwhere(C[2] != oc[2])
    changed_f = np.argwhere((F[2] != of[2]) & (of[2] > 0))
    assert len(changed_c) + len(changed_f) == 1
