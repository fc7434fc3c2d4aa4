"""GPU parity of the flux sweep and EOS kernels against the CPU oracle.

Bar: every updated conserved component within 1e-12 of the oracle,
normalised per component by the field's global scale max|U_c|
(the comparison metric of check_cases.py:35)."""
import numpy as np
import pytest

import hyperflow_amd as hf
from hyperflow_amd import _lib
from hyperflow_amd.physics import WALL_XHI, WALL_XLO, WALL_YHI, WALL_YLO, FluxSweep, box_for_arrays
from tests.conftest import H2
from tests.test_oracle_flux import random_states

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
TOL = 1e-12


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def smooth_box(o, nx, ny, seed, amp=0.3, umax=300.0):
    """Ghosted (nv, ny+4, nx+4) field with smooth variations plus a few jumps."""
    ns = o.ns
    rng = np.random.default_rng(seed)
    gx, gy = nx + 4, ny + 4
    base = random_states(o, 4, seed, umax=umax)
    x = np.linspace(0, 1, gx)[None, :]
    y = np.linspace(0, 1, gy)[:, None]
    w = 0.5 + 0.5 * np.sin(2 * np.pi * (x * rng.uniform(0.5, 2) + y * rng.uniform(0.5, 2)))
    G = base[0][:, None, None] * (1 - w) + base[1][:, None, None] * w
    jump = (x > 0.6) & (y < 0.4)
    G = np.where(jump[None], base[2][:, None, None] * (1 + amp * w), G)
    return np.ascontiguousarray(G)


def normalised_error(a, b):
    scale = np.abs(b).reshape(b.shape[0], -1).max(1).clip(min=1e-300)
    return (np.abs(a - b).reshape(b.shape[0], -1).max(1) / scale).max()


@pytest.mark.parametrize("nx,ny,walls,first_order", [(40, 33, 0, False), (31, 17, WALL_XLO | WALL_YHI, False),
                                                     (16, 47, WALL_XHI | WALL_YLO, True), (15, 15, 15, False)])
def test_flux_stage_matches_oracle(oracle_h2, mech_h2, nx, ny, walls, first_order):
    print(f"[gpu] flux parity {nx}x{ny} walls={walls} first_order={first_order}", flush=True)
    G = smooth_box(oracle_h2, nx, ny, seed=nx * 100 + ny)
    dtdx, dtdy = 2e-4, 1.5e-4
    r, ref, fb, ff = oracle_h2.flux_stage(G, dtdx, dtdy, first_order=first_order, walls=walls)
    assert r == 0
    Gd = dev(G)
    out = torch.zeros((G.shape[0], ny, nx), dtype=torch.float64, device="cuda")
    sweep = FluxSweep(mech_h2)
    st = sweep.stage([box_for_arrays(G, out, dtdx, dtdy, walls)], Gd, out, first_order)
    err = normalised_error(out.cpu().numpy(), ref)
    print(f"[gpu] max normalised error {err:.3e} stats {st}", flush=True)
    assert err < TOL
    assert st["first_order_sides"] == fb


def test_flux_stage_one_dimensional(oracle_h2, mech_h2):
    nx = 600
    q = random_states(oracle_h2, 2, 21, umax=50.0)
    G = np.empty((12, 5, nx + 4))
    G[:, :, : (nx + 4) // 3] = q[0][:, None, None]
    G[:, :, (nx + 4) // 3:] = q[1][:, None, None]
    r, ref, fb, ff = oracle_h2.flux_stage(G, 3e-4, 0.0, walls=WALL_XLO, dims=1)
    assert r == 0
    out = torch.zeros((12, 1, nx), dtype=torch.float64, device="cuda")
    FluxSweep(mech_h2).stage([box_for_arrays(G, out, 3e-4, 0.0, WALL_XLO)], dev(G), out)
    assert normalised_error(out.cpu().numpy(), ref) < TOL


def test_multiple_boxes_in_one_call(oracle_h2, mech_h2):
    shapes = [(20, 18), (64, 16), (17, 30)]
    Gs = [smooth_box(oracle_h2, nx, ny, seed=i + 40) for i, (nx, ny) in enumerate(shapes)]
    flat_in = np.concatenate([g.reshape(-1) for g in Gs])
    outs = sum(12 * nx * ny for nx, ny in shapes)
    din, dout = dev(flat_in), torch.zeros(outs, dtype=torch.float64, device="cuda")
    boxes, ioff, ooff = [], 0, 0
    for (nx, ny), g in zip(shapes, Gs):
        b = _lib.hf_box(nx, ny, 0, 0, ioff, nx + 4, (nx + 4) * (ny + 4), ooff, nx, nx * ny, 1e-4, 2e-4)
        boxes.append(b)
        ioff += g.size
        ooff += 12 * nx * ny
    FluxSweep(mech_h2).stage(boxes, din, dout)
    res = dout.cpu().numpy()
    ooff = 0
    for (nx, ny), g in zip(shapes, Gs):
        r, ref, fb, ff = oracle_h2.flux_stage(g, 1e-4, 2e-4)
        got = res[ooff: ooff + 12 * nx * ny].reshape(12, ny, nx)
        assert normalised_error(got, ref) < TOL
        ooff += 12 * nx * ny


def test_flux_stage_reports_inadmissible_cells(oracle_h2, mech_h2):
    G = smooth_box(oracle_h2, 20, 20, seed=5)
    G[0, 12, 9] = -5.0                      # cell (i=7, j=10) of the valid region
    out = torch.zeros((12, 20, 20), dtype=torch.float64, device="cuda")
    with pytest.raises(hf.HyperflowError) as e:
        FluxSweep(mech_h2).stage([box_for_arrays(G, out, 1e-4, 1e-4)], dev(G), out)
    assert e.value.code == _lib.HF_EDOMAIN
    assert (e.value.first_failed >> 32) == 0


def test_properties_and_stable_dt_match_oracle(oracle_h2, mech_h2):
    U = random_states(oracle_h2, 5000, 31).T.copy()
    p, a, T = mech_h2.properties(dev(U))
    pr, ar, Tr, st = oracle_h2.properties(U.T.copy())
    for x, y in ((p, pr), (a, ar), (T, Tr)):
        assert np.max(np.abs(x.cpu().numpy() - y) / np.abs(y)) < 1e-13
    dt = mech_h2.stable_dt(dev(U), 1e-4, 2e-4, 0.5)
    rho = U[:9].sum(0)
    rate = (np.abs(U[9] / rho) + ar) / 1e-4 + (np.abs(U[10] / rho) + ar) / 2e-4
    assert dt == pytest.approx(0.5 / rate.max(), rel=1e-13)


@pytest.mark.parametrize("nx,ny,walls,first_order", [(23, 19, WALL_XLO | WALL_YHI, False), (16, 9, 0, True)])
def test_dme_flux_stage_matches_oracle(oracle_dme, nx, ny, walls, first_order):
    """The 39-species DME mechanism (NV = 42): 7 x 7 tiles of one wave."""
    print(f"[gpu] DME flux parity {nx}x{ny} walls={walls} first_order={first_order}", flush=True)
    mech = hf.Mechanism(hf.DME_SK39)
    G = smooth_box(oracle_dme, nx, ny, seed=nx * 100 + ny + 7)
    dtdx, dtdy = 2e-4, 1.5e-4
    r, ref, fb, ff = oracle_dme.flux_stage(G, dtdx, dtdy, first_order=first_order, walls=walls)
    assert r == 0
    out = torch.zeros((G.shape[0], ny, nx), dtype=torch.float64, device="cuda")
    st = FluxSweep(mech).stage([box_for_arrays(G, out, dtdx, dtdy, walls)], dev(G), out, first_order)
    err = normalised_error(out.cpu().numpy(), ref)
    print(f"[gpu] DME max normalised error {err:.3e} stats {st}", flush=True)
    assert err < TOL
    assert st["first_order_sides"] == fb


def test_dme_flux_stage_one_dimensional(oracle_dme):
    mech = hf.Mechanism(hf.DME_SK39)
    nx = 150
    q = random_states(oracle_dme, 2, 23, umax=50.0)
    G = np.empty((42, 5, nx + 4))
    G[:, :, : (nx + 4) // 3] = q[0][:, None, None]
    G[:, :, (nx + 4) // 3:] = q[1][:, None, None]
    r, ref, fb, ff = oracle_dme.flux_stage(G, 3e-4, 0.0, walls=WALL_XLO, dims=1)
    assert r == 0
    out = torch.zeros((42, 1, nx), dtype=torch.float64, device="cuda")
    FluxSweep(mech).stage([box_for_arrays(G, out, 3e-4, 0.0, WALL_XLO)], dev(G), out)
    assert normalised_error(out.cpu().numpy(), ref) < TOL


def test_tile_shapes_give_the_same_stage(oracle_h2, mech_h2):
    """The flux stage's online tile-shape choice (15 x 15, 31 x 7, 7 x 31 on
    the first calls of a size class) never changes the result."""
    print("[gpu] flux tile shapes 70x45", flush=True)
    G = smooth_box(oracle_h2, 70, 45, seed=77)
    Gd = dev(G)
    sweep = FluxSweep(hf.Mechanism(H2))
    ref, shapes = None, []
    for _ in range(4):
        out = torch.zeros((G.shape[0], 45, 70), dtype=torch.float64, device="cuda")
        st = sweep.stage([box_for_arrays(G, out, 2e-4, 1.5e-4)], Gd, out)
        shapes.append((st["tile_x"], st["tile_y"]))
        o = out.cpu().numpy()
        ref = o if ref is None else ref
        assert np.array_equal(o, ref)
    print(f"[gpu]   shapes {shapes}", flush=True)
    assert shapes[:3] == [(15, 15), (31, 7), (7, 31)] and shapes[3] in shapes[:3]
