"""The C ABI library: it loads, exports every declared entry point, and its
host-side logic (mechanism loading, argument checks, set-up helpers) behaves
without a GPU. No kernel is launched here."""
import re

import numpy as np
import pytest

import hyperflow_amd as hf
from hyperflow_amd import _lib
from tests.conftest import DME, GOLDEN, H2, ROOT


def declared_functions():
    text = (ROOT / "include" / "hyperflow.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hf_[a-z_A-Z0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.EXPORTED) == names


def test_mechanism_tables_match_independent_loader(oracle_h2, oracle_dme):
    for path, o in ((H2, oracle_h2), (DME, oracle_dme)):
        m = hf.Mechanism(path)
        assert m.species == o.m["species"]
        assert m.nr == o.m["nr"]
        assert np.array_equal(m.W, o.m["W"])


def test_unknown_topology_and_bad_file_are_rejected(tmp_path):
    text = H2.read_text()
    # drop the last reaction: a valid mechanism whose topology has no kernels
    cut = text.rstrip().rfind("- equation:")
    p = tmp_path / "h2_minus_one.yaml"
    p.write_text(text[:cut])
    with pytest.raises(hf.HyperflowError) as e:
        hf.Mechanism(p)
    assert e.value.code == _lib.HF_EINVAL and "compiled kernels" in e.value.message
    with pytest.raises(hf.HyperflowError) as e:
        hf.Mechanism(tmp_path / "missing.yaml")
    assert e.value.code == _lib.HF_EINVAL


def test_argument_validation_before_any_device_work(mech_h2):
    batch = hf.ChemistryBatch(mech_h2)
    U = np.zeros((12, 4))
    with pytest.raises(hf.HyperflowError) as e:
        batch.advance(U, -1.0)
    assert e.value.code == _lib.HF_EINVAL
    with pytest.raises(ValueError):
        batch.advance(U, 1e-7, count=5)
    # count 0 is a successful no-op (chemistry_batch.cpp:25)
    batch.advance(U, 1e-7, count=0)
    assert batch.lastTeamSize() == 0


def test_synthetic_cells_are_deterministic(mech_h2):
    U1, dt1 = mech_h2.synth_cells(0, 2, 64)
    g = np.load(GOLDEN / "synth_h2_seed2.npz")
    assert np.array_equal(U1, g["U"]) and np.array_equal(dt1, g["dt"])
    U2, _ = mech_h2.synth_cells(0, 3, 64)
    assert not np.array_equal(U1, U2)
    assert (dt1 >= 1e-8).all() and (dt1 <= 1e-6).all()


def test_synthetic_cells_sample_the_stated_ranges(oracle_h2, mech_h2):
    U, dt = mech_h2.synth_cells(0, 11, 2000)
    p, a, T, st = oracle_h2.properties(U.T.copy())
    assert (st == 0).all()
    assert T.min() >= 1100 - 1e-6 and T.max() <= 2400 + 1e-6
    assert p.min() >= 0.5 * 101325 * (1 - 1e-9) and p.max() <= 50 * 101325 * (1 + 1e-9)
    # H/O element ratio is 2 phi: phi in [0.5, 2]
    m = oracle_h2.m
    el = (U[:9] / m["W"][:, None]).T @ m["comp"]
    phi = el[:, 0] / el[:, 1] / 2 * 1.0
    assert phi.min() > 0.5 - 1e-9 and phi.max() < 2.0 + 1e-9


def test_equilibrium_conserves_elements(oracle_h2, mech_h2):
    m = oracle_h2.m
    X = np.zeros(9)
    X[0], X[1], X[8] = 3.0, 1.0, 3.76
    X /= X.sum()
    Y = mech_h2.equilibrate_tp(2400.0, 2 * 101325.0, X)
    b_in = X @ m["comp"]
    Xe = (Y / m["W"]) / np.sum(Y / m["W"])
    b_out = Xe @ m["comp"]
    assert np.allclose(b_in / b_in.sum(), b_out / b_out.sum(), rtol=1e-12)


def test_state_from_TPY_round_trip(oracle_h2, mech_h2):
    Y = np.array([0.02, 0.2, 0.05, 1e-5, 1e-4, 1e-3, 1e-6, 1e-7, 0.0])
    Y[8] = 1 - Y.sum()
    q = mech_h2.state_from_TPY(1234.5, 3.3e5, Y, 120.0, -40.0)
    p, a, T, st = oracle_h2.properties(q)
    assert st[0] == 0 and abs(T[0] - 1234.5) < 1e-9 and abs(p[0] - 3.3e5) / 3.3e5 < 1e-12


def test_no_gpu_means_device_error(mech_h2):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    U, dt = mech_h2.synth_cells(0, 2, 8)
    with pytest.raises(hf.HyperflowError) as e:
        hf.ChemistryBatch(mech_h2).advance(U, dt)
    assert e.value.code == _lib.HF_EDEVICE


def test_level_transfer_entry_points_reject_bad_arguments(mech_h2):
    """The ghost-fill, interpolation, reflux and tagging entry points check
    their arguments before any device work (HF_EINVAL, no kernel launched)."""
    from ctypes import byref, c_double, c_int32, c_int64
    L, h = _lib.load(), mech_h2.handle
    box = _lib.hf_box(8, 4, 0, 0, 0, 12, 96, 0, 8, 32, 0.0, 0.0)
    inlet = (c_double * mech_h2.nv)()
    for bc, q in (((_lib.HF_BC_WEDGE, 0, 0, 0), inlet),        # wedge on x-lo
                  ((0, 0, 0, _lib.HF_BC_WEDGE), inlet),        # wedge on y-hi
                  ((0, 0, _lib.HF_BC_WEDGE, -1), inlet),       # wedge on a 1-D box
                  ((_lib.HF_BC_INFLOW, 0, 0, 0), None),        # inflow without an inlet state
                  ((0, 0, 7, 0), inlet)):                      # unknown code
        assert L.hf_fill_ghosts_flow(h, None, byref(box), (c_int32 * 4)(*bc), q) == _lib.HF_EINVAL
    # the plain fill refuses the flow codes
    assert L.hf_fill_ghosts(h, None, byref(box), (c_int32 * 4)(_lib.HF_BC_INFLOW, 0, 0, 0)) == _lib.HF_EINVAL
    p = _lib.hf_patch(0, 8, 32)
    ff = c_int64(0)
    assert L.hf_interp_coarse_to_fine(h, None, byref(p), None, byref(p), (c_int32 * 4)(4, 4, 3, 7), byref(ff),
                                      byref(ff)) == _lib.HF_EINVAL           # empty region
    assert L.hf_average_down(h, None, byref(p), None, byref(p), (c_int32 * 4)(0, 0, 1, 1)) == _lib.HF_EINVAL
    fl = _lib.hf_face_fluxes()
    for fb, sides, step in (((1, 0, 7, 7), 15, 1e-9),                 # not aligned to ratio 2
                            ((0, 0, 7, 7), 16, 1e-9),                 # invalid side mask
                            ((0, 0, 7, 7), 15, -1.0)):                # invalid step
        assert L.hf_reflux(h, None, byref(p), byref(fl), byref(fl), (c_int32 * 4)(*fb), sides, step, 1e-3,
                           1e-3) == _lib.HF_EINVAL
    prm = _lib.hf_tag_params(0.08, 0.08, 50.0, 1, 1, mech_h2.ns, 0)     # fuel index out of range
    prm.dx = prm.dy = 1e-3
    from ctypes import c_uint8
    tags = (c_uint8 * 32)()                      # reaches the fuel-range check, which comes first
    assert L.hf_tag_cells(h, None, byref(box), byref(prm), tags, byref(ff), byref(ff)) == _lib.HF_EINVAL
    assert "fuel" in L.hf_last_error(h).decode()
    assert L.hf_flux_stage_fluxes(h, byref(box), 1, None, None, 0, None, None, None) == _lib.HF_EINVAL


def test_chemistry_interval_argument_checks():
    """One interval for every cell from a number, a 0-d or a one-element
    array or tensor; per-cell intervals must be 1-D, contiguous and hold at
    least as many values as cells (a shorter buffer would be read past its
    end by the library)."""
    import torch

    from hyperflow_amd.chemistry import _interval
    a, per = _interval(2e-7, 10)
    assert per == 0 and a.shape == (1,) and a[0] == 2e-7
    a, per = _interval(np.array(3e-7), 10)
    assert per == 0 and a[0] == 3e-7
    a, per = _interval(np.array([4e-7]), 10)
    assert per == 0 and a[0] == 4e-7
    a, per = _interval(torch.tensor(5e-7, dtype=torch.float64), 10)
    assert per == 0 and a.numel() == 1 and float(a[0]) == 5e-7
    a, per = _interval(np.full(10, 1e-7), 10)
    assert per == 1
    for bad in (np.full(9, 1e-7), torch.full((9,), 1e-7, dtype=torch.float64), np.full((2, 5), 1e-7),
                np.full(20, 1e-7)[::2]):
        with pytest.raises(ValueError):
            _interval(bad, 10)


def test_cxx_consumer_builds_against_the_header(tmp_path):
    """tests/abi_consumer.cpp (the integration guide's ChemistryBatch wrapper
    and the AMReX-style per-box table) compiles with the host compiler
    against include/hyperflow.h, links the library and runs its no-device
    check: a missing mechanism throws std::runtime_error."""
    import subprocess
    from hyperflow_amd import build as B
    exe = tmp_path / "abi_consumer"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", str(B.INCLUDE), "-I",
           "/opt/rocm/include", str(B.CONSUMER_SRC), "-o", str(exe), "-L", str(B.PKG), "-lhyperflow_hip", "-L",
           "/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{B.PKG}:/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(H2), "--compile-only"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "PASS" in r.stdout and "FAIL" not in r.stdout, r.stdout + r.stderr
