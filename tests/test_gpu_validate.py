"""The reference's integration checks, re-expressed on the GPU hierarchy
driver (AmrDriver): validate.py's groups (HyperFlow_V3/tests/validate.py:
90-202) and check_cases.py's smoke invariants for sun_cj and teng_case1
(HyperFlow_V3/tests/check_cases.py:39-91), with the reference's
configurations, initial states (BlockAMR.cpp:59-118) and thresholds.

Differences from the reference runs, each forced by this driver:
  - the comparison across chemistry thread counts (8 vs 1 OpenMP workers)
    becomes plain vs batched chemistry, the two scheduling forms of the same
    per-cell integration (the thresholds are the reference's);
  - fine levels are filled from the coarser level at set-up (initial_grids)
    rather than sampled from the initial condition;
  - sun_cj's base level is 6252 x 4 cells over 0.50016 m x 2 m (the smoke
    case has 6250 x 2 over 0.5 m x 1 m with blocking factor 2; this driver
    needs a blocking factor of at least 4 that divides the base cells): the
    same 80 um x 0.5 m cells, two more columns at the transmissive end and
    twice the transverse extent, so the hot-spot volume doubles.
The zero-dimensional check against a converged reactor is in
test_gpu_chem.py and the thread-invariance of a single batch in
test_gpu_chem_forms.py."""
import numpy as np
import pytest

import hyperflow_amd as hf
from hyperflow_amd import HF_BC_PERIODIC, HF_BC_REFLECTING, HF_BC_TRANSMISSIVE
from hyperflow_amd.hierarchy import AmrDriver

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

AIR = {"H2": 2.0, "O2": 1.0, "N2": 3.76}


def mixture_Y(mech):
    X = np.array([AIR.get(n, 0.0) for n in mech.species])
    Y = X * mech.W
    return Y / Y.sum()


def primitive(mech, T, p, u=0.0, v=0.0):
    """BlockAMR.cpp:20 primitive() for the H2:2,O2:1,N2:3.76 mole fractions."""
    return mech.state_from_TPY(T, p, mixture_Y(mech), u, v)


def initial_field(mech, problem, nx, ny, dx, dy, x0=0.0, y0=0.0, quadrature=4, **c):
    """BlockAMR::initial on an nx x ny level: (nv, ny, nx)."""
    fresh = primitive(mech, c.get("T0", 300.0), c.get("P0", 101325.0), c.get("u0", 0.0), c.get("v0", 0.0))
    hot = primitive(mech, c.get("hot_T", 1500.0), c.get("hot_P", 10 * 101325.0))
    Q = np.zeros((mech.nv, ny, nx))
    xc = x0 + (np.arange(nx) + 0.5) * dx
    yc = y0 + (np.arange(ny) + 0.5) * dy
    if problem == "uniform":
        Q[:] = fresh[:, None, None]
        return Q
    if problem == "sun_cj":
        hot = primitive(mech, 3000.0, 90 * 101325.0)
        f = np.clip((c.get("hotspot_width", 0.002) - (xc - 0.5 * dx)) / dx, 0.0, 1.0)
        Q[:] = (f[None, :] * hot[:, None] + (1 - f[None, :]) * fresh[:, None])[:, None, :]
        return Q
    cx, cy = c["center"]
    for j in range(ny):
        for i in range(nx):
            acc = np.zeros(mech.nv)
            for b in range(quadrature):
                for a in range(quadrature):
                    xx = xc[i] + dx * ((a + 0.5) / quadrature - 0.5)
                    yy = yc[j] + dy * ((b + 0.5) / quadrature - 0.5)
                    if problem == "entropy_wave":
                        r2 = (xx - cx) ** 2 + (yy - cy) ** 2
                        bump = 1 + c.get("amplitude", 0.2) * np.exp(-r2 / c["sigma"] ** 2)
                        acc += primitive(mech, c.get("T0", 300.0) / bump, c.get("P0", 101325.0), c.get("u0", 0.0),
                                         c.get("v0", 0.0))
                    else:   # blast2d
                        acc += hot if np.hypot(xx - cx, yy - cy) < c["radius"] else fresh
            Q[:, j, i] = acc / quadrature ** 2
    return Q


def leaf_rows(d):
    """The leaf cells of a hierarchy as the reference's state CSV rows."""
    mech = d.mech
    out = {k: [] for k in ("level", "i", "j", "x", "y", "dx", "dy")}
    cols = []
    for lev in range(d.finest + 1):
        Q, mask = d.level_state(lev)
        jj, ii = np.nonzero(mask == 1)
        L = d.levels[lev]
        out["level"].append(np.full(ii.size, lev))
        out["i"].append(ii)
        out["j"].append(jj)
        out["x"].append(d.x0 + (ii + 0.5) * L.dx)
        out["y"].append(d.y0 + (jj + 0.5) * L.dy)
        out["dx"].append(np.full(ii.size, L.dx))
        out["dy"].append(np.full(ii.size, L.dy))
        cols.append(Q[:, jj, ii])
    r = {k: np.concatenate(v) for k, v in out.items()}
    U = np.ascontiguousarray(np.concatenate(cols, axis=1))
    p, _, T = (x.cpu().numpy() for x in mech.properties(torch.from_numpy(U).cuda()))
    rho = U[:mech.ns].sum(0)
    r.update(rho=rho, u=U[mech.ns] / rho, v=U[mech.ns + 1] / rho, p=p, T=T, rhoE=U[mech.ns + 2], Y=U[:mech.ns] / rho)
    return r


def check_geometry(r, d, length):
    """validate.py check_geometry: cell sizes per level, the area, and the
    leaf cells covering the finest raster exactly once."""
    assert np.allclose(r["dx"], d.dx0 / 2.0 ** r["level"], rtol=1e-13, atol=0)
    assert np.allclose(r["dy"], d.dy0 / 2.0 ** r["level"], rtol=1e-13, atol=0)
    assert abs(np.sum(r["dx"] * r["dy"]) / np.prod(length) - 1) < 1e-13
    full = np.zeros((d.ny0 << d.max_level, d.nx0 << d.max_level), dtype=np.int16)
    for lev, i, j in zip(r["level"], r["i"], r["j"]):
        k = 1 << (d.max_level - lev)
        full[j * k:(j + 1) * k, i * k:(i + 1) * k] += 1
    assert np.all(full == 1), "leaf cells overlap or leave a hole"
    return full.size


def raster(r, d, field):
    img = np.zeros((d.ny0 << d.max_level, d.nx0 << d.max_level))
    for lev, i, j, v in zip(r["level"], r["i"], r["j"], r[field]):
        k = 1 << (d.max_level - lev)
        img[j * k:(j + 1) * k, i * k:(i + 1) * k] = v
    return img


def conserved_error(a, b, mech, remap=0.0):
    mass = a[:mech.ns].sum()
    return dict(mass=abs(b[:mech.ns].sum() - mass) / mass,
                energy=abs(b[mech.ns + 2] - a[mech.ns + 2]) / max(abs(a[mech.ns + 2]), 1e-100), remap=remap)


def run(d, Q, end, outputs=None):
    """BlockAMR::run: steps to each output time (the last is end); returns
    the leaf rows at t = 0 and at every output."""
    d.initial_grids(Q)
    d._average("U")
    frames = [leaf_rows(d)]
    for target in (outputs or [end]):
        while d.time < target:
            d.step(target=target)
        frames.append(leaf_rows(d))
    return frames


def normalized_difference(a, b, fields):
    assert len(a["x"]) == len(b["x"])
    oa, ob = np.lexsort((a["y"], a["x"])), np.lexsort((b["y"], b["x"]))
    errs = {}
    for n in fields:
        va, vb = (a[n][oa], b[n][ob]) if n != "Y" else (a["Y"][:, oa], b["Y"][:, ob])
        errs[n] = float(np.max(np.abs(va - vb)) / max(float(np.max(np.abs(va))), 1.0))
    return errs


BASE = dict(nx0=16, ny0=16, length=(0.01, 0.01), max_level=2, blocking=4, max_grid_size=16, n_error_buf=1,
            cfl=0.6, regrid_interval=2)


def driver(mech, cfg, bc, tag, chemistry, chem_mode=hf.HF_CHEM_BATCHED, **kw):
    nx0, ny0 = cfg["nx0"], cfg["ny0"]
    return AmrDriver(mech, nx0, ny0, cfg["length"][0] / nx0, cfg["length"][1] / ny0, bc,
                     max_level=cfg["max_level"], cfl=cfg["cfl"], chem_mode=chem_mode, blocking=cfg["blocking"],
                     max_grid_size=cfg["max_grid_size"], n_error_buf=cfg["n_error_buf"],
                     regrid_interval=cfg["regrid_interval"], tag=tag, chemistry=chemistry, **kw)


def test_uniform_local_amr():
    """validate.py:92-103: a uniform periodic state stays uniform to 1e-12
    across a two-level refined box (reflux, interpolation and average-down
    exact for uniform data), with no retry."""
    print("[gpu] validate uniform_local_amr", flush=True)
    mech = hf.Mechanism(hf.H2_LI)
    d = driver(mech, BASE, (HF_BC_PERIODIC,) * 4, dict(mode="box", box=(0.003, 0.007, 0.003, 0.007)), False)
    Q = initial_field(mech, "uniform", 16, 16, d.dx0, d.dy0)
    d.initial_grids(Q)
    start = d.totals()
    while d.time < 2e-7:
        d.step(target=2e-7)
    r = leaf_rows(d)
    full = check_geometry(r, d, BASE["length"])
    deviation = max(float(np.ptp(r[n]) / max(abs(r[n][0]), 1.0)) for n in ("rho", "u", "v", "p", "T", "rhoE"))
    assert deviation < 1e-12
    assert d.finest == 2 and len(r["x"]) < full
    assert d.retries == 0
    err = conserved_error(start, d.totals(), mech)
    assert err["mass"] < 1e-12 and err["energy"] < 1e-12
    print(f"[gpu]   leaves {len(r['x'])} of {full}, deviation {deviation:.2e}, conservation {err}", flush=True)


def test_moving_refine_coarsen():
    """validate.py:105-122: an entropy wave under a tag box moving with the
    flow; the refined region moves (cells refined and cells coarsened),
    mass and energy are conserved to 1e-11, the regrid remap defect stays
    below 1e-9, no retry."""
    print("[gpu] validate moving_refine_coarsen", flush=True)
    mech = hf.Mechanism(hf.H2_LI)
    cfg = dict(BASE, regrid_interval=4)
    d = driver(mech, cfg, (HF_BC_PERIODIC,) * 4,
               dict(mode="box", box=(0.002, 0.004, 0.003, 0.005), velocity=(300.0, 200.0)), False)
    Q = initial_field(mech, "entropy_wave", 16, 16, d.dx0, d.dy0, center=(0.003, 0.004), sigma=0.0007, u0=300.0,
                      v0=200.0)
    d.initial_grids(Q)
    before_rows = leaf_rows(d)
    a = d.totals()
    for target in (2e-6, 4e-6):
        while d.time < target:
            d.step(target=target)
    r = leaf_rows(d)
    check_geometry(r, d, cfg["length"])
    before, after = raster(before_rows, d, "level"), raster(r, d, "level")
    refined, coarsened = int(np.sum(after > before)), int(np.sum(after < before))
    err = conserved_error(a, d.totals(), mech, d.max_remap_defect)
    assert refined > 0 and coarsened > 0
    assert err["mass"] < 1e-11 and err["energy"] < 1e-11 and err["remap"] < 1e-9
    assert d.retries == 0
    print(f"[gpu]   refined {refined}, coarsened {coarsened} finest pixels, {d.steps} steps, {err}", flush=True)


def test_blast2d_inert():
    """validate.py:124-140: an inert blast in a reflecting box with physics
    tagging; mass and energy conserved to 1e-11, no retry. The pressure
    image's transpose asymmetry is reported, as the reference does."""
    print("[gpu] validate blast2d_inert", flush=True)
    mech = hf.Mechanism(hf.H2_LI)
    cfg = dict(BASE, nx0=24, ny0=24, length=(0.02, 0.02), regrid_interval=4)
    d = driver(mech, cfg, (HF_BC_REFLECTING,) * 4, dict(pressure_tag=0.06, temperature_tag=0.06), False)
    Q = initial_field(mech, "blast2d", 24, 24, d.dx0, d.dy0, center=(0.01, 0.01), radius=0.002)
    d.initial_grids(Q)
    a = d.totals()
    for target in (1e-6, 2e-6):
        while d.time < target:
            d.step(target=target)
    r = leaf_rows(d)
    check_geometry(r, d, cfg["length"])
    err = conserved_error(a, d.totals(), mech)
    assert err["mass"] < 1e-11 and err["energy"] < 1e-11
    assert d.retries == 0
    img = raster(r, d, "p")
    symmetry = float(np.mean(np.abs(img - img.T)) / np.mean(img))
    print(f"[gpu]   {d.steps} steps, finest {d.finest}, {err}, transpose pressure error {symmetry:.2e}", flush=True)


def _blast_reacting(chem_mode):
    mech = hf.Mechanism(hf.H2_LI)
    cfg = dict(BASE, max_level=1)
    bc = (HF_BC_REFLECTING,) * 4        # boundary_lo's default and boundary_hi reflecting
    d = driver(mech, cfg, bc, {}, True, chem_mode=chem_mode)
    Q = initial_field(mech, "blast2d", 16, 16, d.dx0, d.dy0, center=(0.005, 0.005), radius=0.0015)
    d.initial_grids(Q)
    a = d.totals()
    for target in (1e-7, 2e-7):
        while d.time < target:
            d.step(target=target)
    return d, a, leaf_rows(d)


def test_blast2d_reacting_and_chemistry_form_invariance():
    """validate.py:175-202: a reacting blast with physics and fuel-rate
    tagging: conservation to 1e-11, Y >= 0 and sum(Y) = 1 to 1e-12, T
    finite and positive, radicals grown, no retry; the same run with the
    other chemistry form agrees to 1e-10 (normalised maximum difference)."""
    print("[gpu] validate blast2d_reacting (batched, then plain chemistry)", flush=True)
    d, a, r = _blast_reacting(hf.HF_CHEM_BATCHED)
    mech = d.mech
    err = conserved_error(a, d.totals(), mech)
    assert err["mass"] < 1e-11 and err["energy"] < 1e-11
    Y = r["Y"]
    assert np.min(Y) >= 0 and np.max(np.abs(Y.sum(0) - 1)) < 1e-12
    assert np.isfinite(r["T"]).all() and np.min(r["T"]) > 0
    sp = {n: k for k, n in enumerate(mech.species)}
    radicals = float(np.max(Y[sp["H"]] + Y[sp["O"]] + Y[sp["OH"]]))
    assert radicals > 1e-12
    assert d.retries == 0
    d1, _, r1 = _blast_reacting(hf.HF_CHEM_PLAIN)
    errs = normalized_difference(r, r1, ("rho", "u", "v", "p", "T", "rhoE", "Y"))
    assert max(errs.values()) < 1e-10, errs
    print(f"[gpu]   {d.steps} steps, radicals {radicals:.3e}, {err}, forms differ by {max(errs.values()):.1e}",
          flush=True)


def _case_checks(d, r, end):
    """check_cases.py:43-50: completed at the end time with no retry, every
    field finite, Y >= 0 and sum(Y) = 1 to 1e-12, rho, p, T positive."""
    assert d.retries == 0 and abs(d.time - end) < 1e-20
    for k, v in r.items():
        assert np.isfinite(v).all(), k
    assert np.min(r["Y"]) >= 0 and np.max(np.abs(r["Y"].sum(0) - 1)) < 1e-12
    assert all(np.min(r[n]) > 0 for n in ("rho", "p", "T"))


def _sun_cj(chem_mode):
    mech = hf.Mechanism(hf.H2_LI)
    cfg = dict(nx0=6252, ny0=4, length=(6252 * 8e-5, 2.0), max_level=2, blocking=4, max_grid_size=32, n_error_buf=2, cfl=0.5,
               regrid_interval=16)
    bc = (HF_BC_REFLECTING, HF_BC_TRANSMISSIVE, HF_BC_REFLECTING, HF_BC_REFLECTING)
    d = driver(mech, cfg, bc, dict(pressure_tag=0.08, temperature_tag=0.08, fuel_rate_tag=50.0,
                                   box=(0.0, 0.002, 0.0, 2.0)), True, chem_mode=chem_mode)
    Q = initial_field(mech, "sun_cj", 6252, 4, d.dx0, d.dy0)
    frames = run(d, Q, 1e-7)
    return d, frames


def test_sun_cj_smoke_invariants():
    """check_cases.py:39-83 for sun_cj: the hot spot's volume and the
    initial integrals against the fresh and hot states, mass conserved to
    1e-11 and energy to 1e-9, transverse uniformity (every column of leaf
    cells equal to 1e-8 of each field's scale), and the other chemistry
    form within 1e-10."""
    print("[gpu] check_cases sun_cj (batched, then plain chemistry)", flush=True)
    d, (initial, final) = _sun_cj(hf.HF_CHEM_BATCHED)
    mech = d.mech
    _case_checks(d, final, 1e-7)
    assert abs(d.dx0 / 4 - 20e-6) < 1e-16
    assert np.max(np.abs(initial["u"])) == 0 and np.max(np.abs(initial["v"])) == 0
    volume = initial["dx"] * initial["dy"]
    hot = np.isclose(initial["T"], 3000.0, rtol=1e-10)
    total = 6252 * 8e-5 * 2.0
    assert abs(volume.sum() - total) < 1e-13 and abs(volume[hot].sum() - 0.004) < 1e-14
    fresh, spot = primitive(mech, 300.0, 101325.0), primitive(mech, 3000.0, 90 * 101325.0)
    rho = lambda q: q[:mech.ns].sum()  # noqa: E731
    # the integrals at t = 0 from the leaf rows (the hierarchy before its first step)
    a0m = float(np.sum(initial["rho"] * volume))
    a0e = float(np.sum(initial["rhoE"] * volume))
    expected_mass = rho(fresh) * (total - 0.004) + rho(spot) * 0.004
    expected_energy = fresh[mech.ns + 2] * (total - 0.004) + spot[mech.ns + 2] * 0.004
    assert abs(a0m / expected_mass - 1) < 1e-12 and abs(a0e / expected_energy - 1) < 1e-12
    vol_f = final["dx"] * final["dy"]
    mass_error = abs(float(np.sum(final["rho"] * vol_f)) / a0m - 1)
    energy_error = abs(float(np.sum(final["rhoE"] * vol_f)) / a0e - 1)
    assert mass_error < 1e-11 and energy_error < 1e-9
    fields = ("rho", "u", "v", "p", "T", "rhoE")
    scales = {n: max(float(np.max(np.abs(final[n]))), 1.0) for n in fields}
    deviation = 0.0
    for lev in np.unique(final["level"]):
        for i in np.unique(final["i"][final["level"] == lev]):
            g = (final["level"] == lev) & (final["i"] == i)
            for n in fields:
                deviation = max(deviation, float(np.ptp(final[n][g])) / scales[n])
            deviation = max(deviation, float(np.max(np.ptp(final["Y"][:, g], axis=1))))
    assert deviation < 1e-8
    _, (_, one) = _sun_cj(hf.HF_CHEM_PLAIN)
    errs = normalized_difference(final, one, fields + ("Y",))
    assert max(errs.values()) < 1e-10, errs
    print(f"[gpu]   {d.steps} steps, leaves {len(initial['x'])} -> {len(final['x'])}, mass {mass_error:.1e}, "
          f"energy {energy_error:.1e}, transverse {deviation:.1e}, forms {max(errs.values()):.1e}", flush=True)


def _teng(chem_mode):
    from hyperflow_amd.driver import teng_osw_base
    mech = hf.Mechanism(hf.H2_LI)
    w = teng_osw_base(mech)
    cfg = dict(nx0=144, ny0=64, length=(0.002304, 0.001024), max_level=2, blocking=4, max_grid_size=32,
               n_error_buf=2, cfl=0.6, regrid_interval=8)
    d = driver(mech, cfg, w["bc"], dict(pressure_tag=0.08, temperature_tag=0.06, fuel_rate_tag=50.0,
                                        box=(-3.2e-5, 0.000128, 0.0, 6.4e-5)), True, chem_mode=chem_mode,
               inlet=w["inlet"], wedge_i=w["wedge_i"], x0=-0.000128)
    frames = run(d, w["Q"], 2e-9)
    return d, frames


def test_teng_case1_smoke_invariants():
    """check_cases.py:39-52,84-87 for teng_case1 (the oblique-shock wedge
    inflow): completed with no retry, finite and admissible leaves, finest
    cells 4 um, and the other chemistry form within 1e-10."""
    print("[gpu] check_cases teng_case1 (batched, then plain chemistry)", flush=True)
    d, (_, final) = _teng(hf.HF_CHEM_BATCHED)
    _case_checks(d, final, 2e-9)
    assert abs(d.dx0 / (1 << d.max_level) - 4e-6) < 1e-16 and abs(d.dy0 / (1 << d.max_level) - 4e-6) < 1e-16
    _, (_, one) = _teng(hf.HF_CHEM_PLAIN)
    errs = normalized_difference(final, one, ("rho", "u", "v", "p", "T", "rhoE", "Y"))
    assert max(errs.values()) < 1e-10, errs
    print(f"[gpu]   {d.steps} steps, finest {d.finest}, leaves {len(final['x'])}, forms {max(errs.values()):.1e}",
          flush=True)
