"""Every chemistry form on the GPU against the CPU oracle, at the stated
floors, and the forms' independence from scheduling.

Forms: H2 plain FP64 (one cell per lane, the drop-in), H2 batched FP64 and
H2 batched mixed (FP32 Newton matrix and LU, FP64 residuals, state and
error control: the form the headline throughput is measured on); DME plain
FP64, batched FP64 and batched mixed (one cell per wave).

Parity is stated at two species floors: Y_k within tol * max(|Y_k|, floor)
with floor 1e-6 and floor 1e-10 (the stated floor of the parity target). T within
1e-6 relative, element masses within 1e-12.

Reference tolerance (rtol 1e-10, atol 1e-14): the reference algorithm (the
oracle at the same setting) is itself away from the converged solution by
its truncation error, which for species near 1e-10 is set by atol (an
absolute error of order 1e-14 per step is 1e-4 relative at Y = 1e-10). At
that setting the bar is therefore the reference algorithm's own distance:
the GPU must be at most twice as far from the converged solution as the
oracle at the same tolerances (and within 1e-6 wherever the oracle is).
At rtol 1e-12 / atol 1e-20 every H2 form meets 1e-6 at the 1e-10 floor.

Scheduling independence (chemistry.cpp:57-59, chemistry_batch.cpp:14-18;
recorded 0.0 in docs/amr_checks.json:66-94): each cell's trajectory
depends on that cell only, so the batched form gives the same bits for any
refill group, bin count, resident waves and Jacobian grouping, and for
repeated launches.
"""
import threading
import time

import numpy as np
import pytest

import hyperflow_amd as hf
from hyperflow_amd import _lib
from tests.conftest import H2

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CONVERGED = dict(rtol=1e-12, atol=1e-20)
H2_FORMS = [(hf.HF_CHEM_PLAIN, hf.HF_PREC_FP64), (hf.HF_CHEM_BATCHED, hf.HF_PREC_FP64),
            (hf.HF_CHEM_BATCHED, hf.HF_PREC_MIXED)]
FORM_IDS = ["plain-fp64", "batched-fp64", "batched-mixed"]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def errors(o, U, R, floor):
    """(T relative, Y relative at floor) of U against R, SoA host arrays."""
    ns = len(o.m["W"])
    Yu, Yr = U[:ns] / U[:ns].sum(0), R[:ns] / R[:ns].sum(0)
    _, _, Tu, su = o.properties(U.T.copy())
    _, _, Tr, sr = o.properties(R.T.copy())
    assert (su == 0).all() and (sr == 0).all()
    return float(np.max(np.abs(Tu - Tr) / Tr)), float(np.max(np.abs(Yu - Yr) / np.maximum(np.abs(Yr), floor)))


def element_drift(m, U0, U1):
    ns = len(m["W"])
    el0 = (U0[:ns] / m["W"][:, None]).T @ m["comp"]
    el1 = (U1[:ns] / m["W"][:, None]).T @ m["comp"]
    return float(np.max(np.abs(el1 - el0) / np.abs(el0)))


def run_gpu(path, U0, dt, mode, precision, **opt):
    mech = hf.Mechanism(path, precision=precision, **opt)
    Ud = dev(U0)
    st = hf.ChemistryBatch(mech, mode=mode).advance(Ud, dev(dt))
    return Ud.cpu().numpy(), st


@pytest.fixture(scope="module")
def h2_cells(oracle_h2):
    """512 seed-2 cells of the batched workload, the converged oracle
    solution and the oracle at the reference tolerance."""
    mech = hf.Mechanism(H2)
    U0, dt = mech.synth_cells(0, 2, 512)
    conv = U0.copy()
    r, _, _, msg = oracle_h2.chem_advance(conv, dt, **CONVERGED)
    assert r == 0, msg
    ref10 = U0.copy()
    r, _, _, msg = oracle_h2.chem_advance(ref10, dt, rtol=1e-10, atol=1e-14)
    assert r == 0, msg
    return U0, dt, conv, ref10


@pytest.mark.parametrize("mode,precision", H2_FORMS, ids=FORM_IDS)
def test_h2_form_at_default_tolerance(oracle_h2, h2_cells, mode, precision):
    """Each H2 form at the product default (the reference's CVODE setting,
    rtol 1e-8 / atol 1e-15) against the converged oracle: T and Y within 1e-6
    at the 1e-6 floor, element masses within 1e-12, momenta and energy
    untouched. The 1e-10-floor figure is reported."""
    U0, dt, conv, _ = h2_cells
    print(f"[gpu] H2 form mode={mode} precision={precision} n=512 seed=2 rtol 1e-8", flush=True)
    out, st = run_gpu(H2, U0, dt, mode, precision)
    eT, eY6 = errors(oracle_h2, out, conv, 1e-6)
    _, eY10 = errors(oracle_h2, out, conv, 1e-10)
    drift = element_drift(oracle_h2.m, U0, out)
    print(f"[gpu]   T {eT:.3e}  Y(floor 1e-6) {eY6:.3e}  Y(floor 1e-10) {eY10:.3e}  elements {drift:.3e}; "
          f"steps {st['steps']} rhs {st['rhs']} jac {st['jac']}", flush=True)
    assert eT < 1e-6 and eY6 < 1e-6
    assert drift < 1e-12
    assert np.array_equal(out[9:], U0[9:])


@pytest.mark.parametrize("n", [1, 63, 65, 1000])
def test_h2_mixed_ragged_sizes(oracle_h2, n):
    """The headline form on batches that do not fill whole waves, at rtol
    1e-10 / atol 1e-20, against the converged oracle (1e-6 at the 1e-10
    floor)."""
    print(f"[gpu] H2 batched mixed ragged n={n} seed=9", flush=True)
    mech = hf.Mechanism(H2)
    U0, dt = mech.synth_cells(0, 9, n)
    conv = U0.copy()
    assert oracle_h2.chem_advance(conv, dt, **CONVERGED)[0] == 0
    out, st = run_gpu(H2, U0, dt, hf.HF_CHEM_BATCHED, hf.HF_PREC_MIXED, rtol=1e-10, atol=1e-20)
    assert st["cells"] == n
    eT, eY = errors(oracle_h2, out, conv, 1e-10)
    print(f"[gpu]   T {eT:.3e}  Y(floor 1e-10) {eY:.3e}", flush=True)
    assert eT < 1e-6 and eY < 1e-6


@pytest.mark.parametrize("mode,precision", H2_FORMS, ids=FORM_IDS)
def test_h2_form_at_reference_tolerance(oracle_h2, h2_cells, mode, precision):
    """rtol 1e-10 / atol 1e-14 (the stated CVODE setting): against the
    converged oracle at both floors, the GPU is within 1e-6 or within twice
    the reference algorithm's own distance at the same setting; T within
    1e-6."""
    U0, dt, conv, ref10 = h2_cells
    print(f"[gpu] H2 form mode={mode} precision={precision} n=512 seed=2 rtol 1e-10 atol 1e-14", flush=True)
    out, _ = run_gpu(H2, U0, dt, mode, precision, rtol=1e-10, atol=1e-14)
    for floor in (1e-6, 1e-10):
        eT, eY = errors(oracle_h2, out, conv, floor)
        oT, oY = errors(oracle_h2, ref10, conv, floor)
        sT, sY = errors(oracle_h2, out, ref10, floor)
        print(f"[gpu]   floor {floor:g}: GPU vs converged T {eT:.3e} Y {eY:.3e}; reference algorithm vs converged "
              f"T {oT:.3e} Y {oY:.3e}; GPU vs reference algorithm T {sT:.3e} Y {sY:.3e}", flush=True)
        assert eT < 1e-6
        assert eY < max(1e-6, 2.0 * oY)
    assert element_drift(oracle_h2.m, U0, out) < 1e-12


@pytest.mark.parametrize("mode,precision", H2_FORMS, ids=FORM_IDS)
def test_h2_form_tight_meets_floor_1e10(oracle_h2, h2_cells, mode, precision):
    """At rtol 1e-12 / atol 1e-20 every H2 form is within 1e-6 of the
    converged oracle at the 1e-10 floor."""
    U0, dt, conv, _ = h2_cells
    print(f"[gpu] H2 form mode={mode} precision={precision} n=512 seed=2 rtol 1e-12 atol 1e-20", flush=True)
    out, _ = run_gpu(H2, U0, dt, mode, precision, rtol=1e-12, atol=1e-20)
    eT, eY = errors(oracle_h2, out, conv, 1e-10)
    print(f"[gpu]   T {eT:.3e}  Y(floor 1e-10) {eY:.3e}", flush=True)
    assert eT < 1e-6 and eY < 1e-6


DME_FORMS = [(hf.HF_CHEM_PLAIN, hf.HF_PREC_FP64), (hf.HF_CHEM_BATCHED, hf.HF_PREC_FP64),
             (hf.HF_CHEM_BATCHED, hf.HF_PREC_MIXED)]


def _accepted(run, n):
    """Indices of the cells a batch call accepts: the call is repeated
    without the lowest failing cell until it succeeds (each failure must be
    the reference's accepted-species check, HF_EDOMAIN). Returns (kept
    indices, rejected indices, output of the final call)."""
    keep, rejected = list(range(n)), []
    while True:
        r, ff, out = run(keep)
        if r == 0:
            return keep, rejected, out
        assert r == _lib.HF_EDOMAIN, r
        rejected.append(keep.pop(ff))


@pytest.mark.parametrize("mode,precision", DME_FORMS, ids=FORM_IDS)
def test_dme_form_at_reference_tolerance(oracle_dme, mode, precision):
    """DME at rtol 1e-10 / atol 1e-14 on 96 seed-4 cells. At this atol the
    reference algorithm rejects cells: a trace species ends below -1e-13
    (chemistry.cpp:69-73), which atol 1e-14 does not resolve. Which cells
    sit on that threshold differs between two integrations at the roundoff
    level, so the forms may reject a few cells the oracle accepts and the
    other way round. Asserted: every rejection is that check (HF_EDOMAIN);
    the form and the oracle agree on at least 90 % of the cells; on the
    cells both accept, T within 1e-6 and Y at the 1e-6 floor within 1e-6 or
    twice the reference algorithm's own distance (measured: GPU 6.6e-5,
    oracle 5.9e-5). At the 1e-10 floor the figure is reported only: atol
    1e-14 leaves species below about 1e-8 unresolved in both (measured on
    this batch: the oracle 2.5e-2, the plain GPU form 0.66, i.e. absolute
    errors of 2.5e-12 and 6.6e-11 on species near 1e-10), so two
    integrations at this setting differ there by their own truncation;
    test_dme_form_tight_floor_1e10 holds the 1e-10 floor at a tolerance
    that resolves them."""
    print(f"[gpu] DME form mode={mode} precision={precision} rtol 1e-10 atol 1e-14 n=96 seed=4", flush=True)
    mech = hf.Mechanism(hf.DME_SK39)
    n = 96
    U0, dt = mech.synth_cells(1, 4, n)

    def oracle_run(idx):
        X = np.ascontiguousarray(U0[:, idx])
        r, _, ff, _ = oracle_dme.chem_advance(X, np.ascontiguousarray(dt[idx]), rtol=1e-10, atol=1e-14)
        return r, ff, X

    gmech = hf.Mechanism(hf.DME_SK39, precision=precision, rtol=1e-10, atol=1e-14)
    batch = hf.ChemistryBatch(gmech, mode=mode)

    def gpu_run(idx):
        Ud = dev(U0[:, idx])
        try:
            batch.advance(Ud, dev(dt[idx]))
        except hf.HyperflowError as e:
            return e.code, e.first_failed, None
        return 0, -1, Ud.cpu().numpy()

    ok_o, rej_o, _ = _accepted(oracle_run, n)
    ok_g, rej_g, _ = _accepted(gpu_run, n)
    both = sorted(set(ok_o) & set(ok_g))
    agree = n - len(set(rej_o) ^ set(rej_g))
    print(f"[gpu]   rejected: oracle {sorted(rej_o)}, form {sorted(rej_g)}; verdicts agree on {agree}/{n}",
          flush=True)
    assert agree >= 0.9 * n
    U0, dt = np.ascontiguousarray(U0[:, both]), np.ascontiguousarray(dt[both])
    conv = U0.copy()
    assert oracle_dme.chem_advance(conv, dt, **CONVERGED)[0] == 0
    ref10 = U0.copy()
    assert oracle_dme.chem_advance(ref10, dt, rtol=1e-10, atol=1e-14)[0] == 0
    out, _ = run_gpu(hf.DME_SK39, U0, dt, mode, precision, rtol=1e-10, atol=1e-14)
    for floor in (1e-6, 1e-10):
        eT, eY = errors(oracle_dme, out, conv, floor)
        oT, oY = errors(oracle_dme, ref10, conv, floor)
        print(f"[gpu]   floor {floor:g}: GPU vs converged T {eT:.3e} Y {eY:.3e}; reference algorithm vs converged "
              f"T {oT:.3e} Y {oY:.3e}", flush=True)
        assert eT < 1e-6
        if floor == 1e-6:
            assert eY < max(1e-6, 2.0 * oY)


@pytest.mark.parametrize("mode,precision", DME_FORMS, ids=FORM_IDS)
def test_dme_form_tight_floor_1e10(oracle_dme, mode, precision):
    """At rtol 1e-12 / atol 1e-20 every DME form is within 1e-6 of the
    converged oracle at the 1e-10 floor: the species that atol 1e-14 leaves
    unresolved are resolved when the tolerance asks for them."""
    print(f"[gpu] DME form mode={mode} precision={precision} n=64 seed=5 rtol 1e-12 atol 1e-20", flush=True)
    mech = hf.Mechanism(hf.DME_SK39)
    U0, dt = mech.synth_cells(1, 5, 64)
    conv = U0.copy()
    assert oracle_dme.chem_advance(conv, dt, **CONVERGED)[0] == 0
    out, _ = run_gpu(hf.DME_SK39, U0, dt, mode, precision, rtol=1e-12, atol=1e-20)
    eT, eY = errors(oracle_dme, out, conv, 1e-10)
    print(f"[gpu]   T {eT:.3e}  Y(floor 1e-10) {eY:.3e}", flush=True)
    assert eT < 1e-6 and eY < 1e-6


# ------------------------------------------------------------------ scheduling independence
def _batched(U0, dt, precision=hf.HF_PREC_FP64, **opt):
    out, st = run_gpu(H2, U0, dt, hf.HF_CHEM_BATCHED, precision, eager=True, **opt)
    return out, st


@pytest.fixture(scope="module")
def cells_8k():
    mech = hf.Mechanism(H2)
    return mech.synth_cells(0, 2, 8192)


@pytest.mark.parametrize("precision", [hf.HF_PREC_FP64, hf.HF_PREC_MIXED], ids=["fp64", "mixed"])
def test_batched_independent_of_tuning(cells_8k, precision):
    """Same bits for every refill group, bin count, resident-wave count and
    Jacobian grouping (the reference asserts < 1e-12 across thread counts,
    validate.py:163-171; here the difference is exactly 0)."""
    U0, dt = cells_8k
    print(f"[gpu] batched tuning invariance precision={precision} n=8192", flush=True)
    ref, st0 = _batched(U0, dt, precision, chunk=8, nbins=32, waves_per_cu=4, jac_wait=-1, setup_wait=-1)
    for opt in (dict(chunk=1, nbins=8, waves_per_cu=2, jac_wait=0, setup_wait=0), dict(chunk=16, nbins=64, waves_per_cu=8,
                                                                        jac_wait=8),
                dict(chunk=4, nbins=128, waves_per_cu=4, jac_wait=32), dict(chunk=64, nbins=1, waves_per_cu=1,
                                                                           jac_wait=4)):
        out, st = _batched(U0, dt, precision, **opt)
        print(f"[gpu]   {opt}: max |diff| {np.max(np.abs(out - ref)):.3e}; steps {st['steps']} "
              f"(reference run {st0['steps']})", flush=True)
        assert np.array_equal(out, ref)
        for k in ("steps", "rhs", "jac", "lu", "newton", "conv_failures", "err_failures", "max_cell_steps"):
            assert st[k] == st0[k], k


def test_batched_fp64_matches_plain(cells_8k):
    """The batched FP64 form integrates each cell exactly as the plain
    one-cell-per-lane form does (normalised difference below 1e-12, as
    validate.py:163-171 asks across worker counts; measured 0)."""
    U0, dt = cells_8k
    print("[gpu] batched FP64 vs plain FP64 n=8192", flush=True)
    plain, sp = run_gpu(H2, U0, dt, hf.HF_CHEM_PLAIN, hf.HF_PREC_FP64, eager=True)
    batched, sb = _batched(U0, dt)
    scale = np.maximum(np.abs(plain).max(axis=1, keepdims=True), 1e-300)   # momenta are 0
    d = float(np.max(np.abs(batched - plain) / scale))
    print(f"[gpu]   normalised max difference {d:.3e}; steps plain {sp['steps']} batched {sb['steps']}", flush=True)
    assert d < 1e-12
    assert sp["steps"] == sb["steps"] and sp["rhs"] == sb["rhs"]


def test_repeated_launches_identical():
    """Repeated graph replays of the same 2^16-cell batch give identical
    states and identical step counts."""
    print("[gpu] repeated batched launches n=65536", flush=True)
    mech = hf.Mechanism(H2, precision=hf.HF_PREC_MIXED)
    U0, dt = mech.synth_cells(0, 2, 65536)
    U0d, dtd = dev(U0), dev(dt)
    Ud = U0d.clone()
    b = hf.ChemistryBatch(mech, mode=hf.HF_CHEM_BATCHED)
    first, steps = None, set()
    for rep in range(3):
        Ud.copy_(U0d)
        st = b.advance(Ud, dtd)
        steps.add((st["steps"], st["rhs"], st["jac"]))
        out = Ud.cpu().numpy()
        if first is None:
            first = out
        assert np.array_equal(out, first)
    print(f"[gpu]   counters {steps}", flush=True)
    assert len(steps) == 1


def test_concurrent_call_is_ebusy():
    """A second call on a context while one is running returns HF_EBUSY
    (chemistry_batch.cpp:22 throws for a nested batch) and leaves its field
    unchanged; the running call completes normally."""
    print("[gpu] HF_EBUSY on a concurrent call", flush=True)
    mech = hf.Mechanism(H2, eager=True)
    U0, dt = mech.synth_cells(0, 2, 1 << 20)
    big, bdt = dev(U0), dev(dt)
    small0, sdt = mech.synth_cells(0, 3, 64)
    small = dev(small0)
    result = {}

    def long_call():
        try:
            hf.ChemistryBatch(mech, mode=hf.HF_CHEM_PLAIN).advance(big, bdt)
            result["ok"] = True
        except Exception as e:            # noqa: BLE001 - reported below
            result["err"] = e

    t = threading.Thread(target=long_call)
    t.start()
    seen = None
    deadline = time.time() + 30
    while t.is_alive() and time.time() < deadline:
        try:
            hf.ChemistryBatch(mech, mode=hf.HF_CHEM_PLAIN).advance(small, sdt)
            small.copy_(dev(small0))          # ran before or after the long call: try again
        except hf.HyperflowError as e:
            seen = e
            break
        time.sleep(0.002)
    t.join()
    assert result.get("ok"), result
    assert seen is not None and seen.code == _lib.HF_EBUSY, seen
    assert "nested or concurrent" in seen.message
    assert np.array_equal(small.cpu().numpy(), small0)


def test_online_tuner_keeps_every_result():
    """With every tuning value left to the library, the first calls of a size
    class try candidate configurations (refill group, resident waves, bins,
    Jacobian and setup waits) one at a time; every call gives the same bits,
    and the tuner settles on one configuration."""
    print("[gpu] online tuner, 12 calls of 16384 cells", flush=True)
    mech = hf.Mechanism(H2, precision=hf.HF_PREC_MIXED)
    U0, dt = mech.synth_cells(0, 2, 16384)
    U0d, dtd = dev(U0), dev(dt)
    b = hf.ChemistryBatch(mech, mode=hf.HF_CHEM_BATCHED)
    first, seen, tuned = None, set(), []
    for _ in range(12):
        Ud = U0d.clone()
        st = b.advance(Ud, dtd)
        cfg = (st["chunk"], st["waves_per_cu"], st["nbins"], st["jac_wait"], st["setup_wait"])
        seen.add(cfg)
        tuned.append(st["tuned"])
        out = Ud.cpu().numpy()
        first = out if first is None else first
        assert np.array_equal(out, first)
    print(f"[gpu]   {len(seen)} configurations tried, states {tuned}, final {cfg}", flush=True)
    assert len(seen) >= 6 and tuned[0] == 1 and tuned[-1] == 2


@pytest.mark.parametrize("precision", [hf.HF_PREC_FP64, hf.HF_PREC_MIXED], ids=["fp64", "mixed"])
def test_dme_analytic_jacobian_newton_behaviour(precision):
    """The one-cell-per-wave DME form builds its Jacobian analytically; the
    plain form uses CVODE's difference quotients (chemistry.cpp:15-25).
    The Jacobian only steers the Newton iteration, so a correct one keeps
    the step sequence of the difference-quotient form (both approximate the
    same derivative: within a few per cent in steps, Newton iterations per
    step and setups per cell) while making no rate evaluations of its own
    (about one evaluation per step instead of two). A wrong entry (a sign, a
    missing third-body or temperature term) shows as slower Newton
    convergence: more iterations and failures per step."""
    print(f"[gpu] DME analytic vs difference-quotient Jacobian, precision={precision}, 192 cells seed 4",
          flush=True)
    n = 192
    U0, dt = hf.Mechanism(hf.DME_SK39).synth_cells(1, 4, n)
    stats = {}
    for name, mode, prec in (("plain", hf.HF_CHEM_PLAIN, hf.HF_PREC_FP64), ("wave", hf.HF_CHEM_BATCHED, precision)):
        mech = hf.Mechanism(hf.DME_SK39, rtol=1e-8, atol=1e-20, precision=prec)
        Ud = dev(U0)
        stats[name] = hf.ChemistryBatch(mech, mode=mode).advance(Ud, dev(dt))
    p, w = stats["plain"], stats["wave"]
    print(f"[gpu]   plain: steps {p['steps']} rhs {p['rhs']} newton {p['newton']} lu {p['lu']} jac {p['jac']} "
          f"ncf {p['conv_failures']}; wave: steps {w['steps']} rhs {w['rhs']} newton {w['newton']} lu {w['lu']} "
          f"jac {w['jac']} ncf {w['conv_failures']}", flush=True)
    assert abs(w["steps"] - p["steps"]) <= 0.05 * p["steps"]
    assert w["newton"] / w["steps"] <= 1.05 * p["newton"] / p["steps"]
    assert w["lu"] <= 1.05 * p["lu"]
    assert w["conv_failures"] <= p["conv_failures"] + n // 64
    # no rate evaluations for the Jacobian: about one per Newton iteration
    assert w["rhs"] <= 1.02 * w["newton"] + w["steps"] * 0.2
