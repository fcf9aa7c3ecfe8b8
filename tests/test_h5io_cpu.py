"""Observable and checkpoint files in the reference's HDF5 layout (pffrg/h5io.py over
libpffrg_h5.so).  No file written by the reference ships with it, so the layout is pinned by
what HDF5.jl's conventions fix: h5dump of the written file shows every dataset under the name
the reference composes (checkpoint_su2.jl:2-39, Action/disk.jl, Observable/disk.jl:2-29,
Lattice/disk.jl:2-19), arrays with the Julia dimensions reversed and the Julia (column-major)
bytes in storage order, scalars as scalar dataspaces, strings as fixed-length UTF-8 and Bool as
an 8-bit bitfield.  Λ keys follow Julia's Float64 printing (known answers of Base.Ryu's rule)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pffrg import action as A
from pffrg import h5io as H
from pffrg.mesh import launch_mesh

ROOT = Path(__file__).resolve().parents[1]
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if Path("/opt/conda/bin/h5dump").exists() else None)

pytestmark = pytest.mark.skipif(not H.LIB_PATH.exists(), reason="libpffrg_h5.so not built (no HDF5 headers)")


def _action(ns=3, seed=4):
    m = launch_mesh(1.0, num_sigma=7, num_Omega=5, num_nu=4, num_chi=3)
    a = A.synthetic_action(ns, m, np.array([[0.5, -0.25, 0.1], [0.0, 0.0, 0.0]])[:, :ns], seed=seed)
    return m, a


def _dump(path, *args):
    return subprocess.run([H5DUMP, *args, str(path)], capture_output=True, text=True, check=True).stdout


def test_header_and_binding_agree():
    import ctypes
    import re
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "pffrg_h5.h").read_text(), flags=re.S)
    declared = sorted(set(re.findall(r"\b(pffrg_h5_[a-z0-9_]+)\s*\(", text)))
    lib = ctypes.CDLL(str(H.LIB_PATH))
    assert [s for s in declared if not hasattr(lib, s)] == []
    assert H.exported_symbols() == declared


@pytest.mark.parametrize("x,s", [(50.0, "50.0"), (0.3, "0.3"), (45.0, "45.0"), (4.05, "4.05"), (1e-4, "0.0001"),
                                 (1e-5, "1.0e-5"), (2.5e-7, "2.5e-7"), (100000.0, "100000.0"), (1e6, "1.0e6"),
                                 (123456.7, "123456.7"), (1234567.8, "1.2345678e6"), (-0.5, "-0.5"),
                                 (0.19796360741190397, "0.19796360741190397"), (1e16, "1.0e16")])
def test_julia_float_names(x, s):
    assert H.julia_float(x) == s


def test_checkpoint_round_trip_and_nearest_cutoff(tmp_path):
    m, a = _action()
    p = tmp_path / "cp.h5"
    with H.H5File(p, "cw") as f:
        for L, dL, seed in ((2.0, 0.2, 1), (0.5, 0.05, 2), (0.3, 0.03, 3)):
            aa = A.synthetic_action(a.ns, m, a.bare, seed=seed)
            H.checkpoint(f, L, dL, m, aa)
    with H.H5File(p, "r") as f:
        assert f.read("symmetry") == "su2" and f.read("S") == 0.5
        assert sorted(f.keys("a")) == ["0.3", "0.5", "2.0"]
        Lc, dL, m2, b = H.read_checkpoint(f, 0.45)
    assert (Lc, dL) == (0.5, 0.05)
    ref = A.synthetic_action(a.ns, m, a.bare, seed=2)
    for name in ("sigma", "Os", "ns", "Ot", "nt", "chi"):
        np.testing.assert_array_equal(getattr(m2, name), getattr(m, name))
    np.testing.assert_array_equal(b.gamma, ref.gamma)
    np.testing.assert_array_equal(b.Sigma, ref.Sigma)
    np.testing.assert_array_equal(b.bare, ref.bare)


def test_u1_dm_checkpoint_and_observables(tmp_path):
    # checkpoint_u1_dm.jl:2-78: six vertices Γxx .. Γdz, symmetry "u1-dm", no spin length;
    # save_χ! (Observable/disk.jl:19-24): χ/Λ/xx, zz, xy
    from pffrg import flow as F
    prob = F.Problem.build("triangular", 2, [[1.0, 0.6, 0.4]], initial=1.0, num_sigma=6, num_Omega=4, num_nu=4,
                           normalize=False, model="triangular-dm-c3")
    m = prob.mesh
    a = A.synthetic_action(prob.r.n_sites, m, prob.bare, seed=5)
    assert a.nc == 6
    p, q = tmp_path / "cp_u1.h5", tmp_path / "obs_u1.h5"
    chi = np.arange(3 * prob.r.n_sites * m.num_chi, dtype=np.float64).reshape(3, prob.r.n_sites, m.num_chi)
    with H.H5File(p, "cw") as f:
        H.checkpoint(f, 0.7, 0.07, m, a)
    with H.H5File(q, "cw") as f:
        H.save_chi(f, 0.7, "u1-dm", m, chi)
    with H.H5File(q, "r") as f:
        assert f.read("symmetry") == "u1-dm"
        for k, label in enumerate(("xx", "zz", "xy")):
            np.testing.assert_array_equal(H.read_chi(f, 0.7, label)[1], chi[k])
    with H.H5File(p, "r") as f:
        assert f.read("symmetry") == "u1-dm" and not f.haskey("S")
        assert sorted(f.keys("a/0.7/Γ")) == sorted(["Γxx", "Γzz", "ΓDM", "Γdd", "Γzd", "Γdz"])
        Lc, dL, m2, b = H.read_checkpoint(f, 0.6)
    assert (Lc, dL, b.nc) == (0.7, 0.07, 6)
    np.testing.assert_array_equal(b.gamma, a.gamma)
    np.testing.assert_array_equal(b.bare, a.bare)
    np.testing.assert_array_equal(b.Sigma, a.Sigma)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")
def test_layout_as_hdf5_jl_writes_it(tmp_path):
    m, a = _action()
    ns, nO, nn = a.ns, m.num_Omega, m.num_nu
    p = tmp_path / "cp.h5"
    with H.H5File(p, "cw") as f:
        H.checkpoint(f, 0.5, 0.05, m, a)
        H.save_lattice(f, "square", 3, "heisenberg", False, [np.array([1.0]), np.array([0.5])])
    head = _dump(p, "-H")
    # every dataset name the reference composes
    names = ["dΛ/0.5", "σ/0.5", "Ωs/0.5", "νs/0.5", "Ωt/0.5", "νt/0.5", "χ/0.5", "a/0.5/Σ", "S", "symmetry"]
    for comp in ("spin", "dens"):
        names.append(f"a/0.5/Γ/{comp}/bare")
        names += [f"a/0.5/Γ/{comp}/{ch}/{q}" for ch in ("ch_s", "ch_t", "ch_u") for q in ("q1", "q2_1", "q2_2", "q3")]
    names += ["reduced_lattice/name", "reduced_lattice/size", "reduced_lattice/model", "reduced_lattice/euclidean",
              "reduced_lattice/J/1", "reduced_lattice/J/2"]
    for n in names:
        d = _dump(p, "-H", "-d", "/" + n)
        assert "DATASET" in d, n
    q3 = _dump(p, "-H", "-d", "/a/0.5/Γ/spin/ch_t/q3")
    assert f"SIMPLE {{ ( {nn}, {nn}, {nO}, {ns} ) / ( {nn}, {nn}, {nO}, {ns} ) }}" in q3 and "H5T_IEEE_F64LE" in q3
    q21 = _dump(p, "-H", "-d", "/a/0.5/Γ/dens/ch_u/q2_1")
    assert f"( {nn}, {nO}, {ns} )" in q21
    s = _dump(p, "-H", "-d", "/S")
    assert "SCALAR" in s and "H5T_IEEE_F64LE" in s
    sym = _dump(p, "-H", "-d", "/symmetry")
    assert "STRSIZE 3;" in sym and "H5T_STR_NULLTERM" in sym and "H5T_CSET_UTF8" in sym
    assert "H5T_STD_B8LE" in _dump(p, "-H", "-d", "/reduced_lattice/euclidean")
    size = _dump(p, "-H", "-d", "/reduced_lattice/size")
    assert "H5T_STD_I64LE" in size and "SCALAR" in size
    # storage order = the Julia (column-major) array's bytes
    raw = tmp_path / "q2_2.bin"
    subprocess.run([H5DUMP, "-d", "/a/0.5/Γ/spin/ch_s/q2_2", "-b", "LE", "-o", str(raw), str(p)], capture_output=True,
                   check=True)
    np.testing.assert_array_equal(np.fromfile(raw, dtype="<f8"), a.view(0, 0, "q2_2").ravel(order="F"))


def test_observables_and_measurement(tmp_path):
    m, a = _action()
    obs, cp = tmp_path / "obs.h5", tmp_path / "cp.h5"
    meas = H.Measurement(obs, cp, wt=24.0, ct=0.0, lattice=("square", 3, "heisenberg", False, [np.array([1.0])]))
    chi = np.arange(a.ns * m.num_chi, dtype=np.float64).reshape(a.ns, m.num_chi)
    calls = []

    def get_action():
        calls.append(1)
        return a

    for L in (1.0, 0.9, 0.9):
        meas(L, 0.1, m, chi * L, a.Sigma, get_action)
    assert meas.checkpoints == [1.0, 0.9] and len(calls) == 2  # no second checkpoint of a cutoff
    with H.H5File(obs, "r") as f:
        assert f.keys("χ") == ["0.9", "1.0"] and f.read("symmetry") == "su2"
        Lc, c = H.read_chi(f, 0.95)
        np.testing.assert_array_equal(f.read("χ/1.0/mesh"), m.chi)
        np.testing.assert_array_equal(f.read("Σ/0.9"), a.Sigma)
        assert H.read_lattice(f)[:4] == ("square", 3, "heisenberg", False)
    assert Lc in (0.9, 1.0) and c.shape == (a.ns, m.num_chi)
    # a long checkpoint interval writes observables only
    meas2 = H.Measurement(tmp_path / "o2.h5", tmp_path / "c2.h5", wt=24.0, ct=1.0)
    meas2(0.5, 0.05, m, chi, a.Sigma, get_action)
    assert meas2.checkpoints == []


def test_errors(tmp_path):
    with pytest.raises(H.H5Error):
        H.H5File(tmp_path / "missing.h5", "r")
    with H.H5File(tmp_path / "x.h5", "cw") as f:
        f.write("a/b", np.zeros(3))
        with pytest.raises(H.H5Error):
            f.write("a/b", np.zeros(3))  # file[path] = v on an existing dataset fails, as in HDF5.jl
        with pytest.raises(H.H5Error):
            f.read("a/c")
        assert f.haskey("a/b") and not f.haskey("a/c") and not f.haskey("z/y/x")


def test_flow_measurement_on_the_oracle(tmp_path):
    # the flow driver's measurement hook, run on the FP64 oracle: every accepted step's χ and Σ
    # land in the observable file and a checkpoint per step (ct = 0) holds that step's action
    from pffrg import flow as F
    from helpers import small_problem
    from oracle_ctx import OracleContext
    prob = small_problem("square", 2, 6, 5.0)
    meas = H.Measurement(tmp_path / "obs.h5", tmp_path / "cp.h5", ct=0.0)
    octx = OracleContext(prob)
    res = F.launch_flow(octx, prob, prob.bare_action(), 5.0, 1.0, max_steps=2, measurement=meas)
    assert res.accepted >= 1 and meas.checkpoints == res.Lambdas
    with H.H5File(tmp_path / "obs.h5", "r") as f:
        for L, c in zip(res.Lambdas, res.chi):
            np.testing.assert_array_equal(f.read(f"χ/{H.julia_float(L)}/diag"), c)
    with H.H5File(tmp_path / "cp.h5", "r") as f:
        Lc, dL, m, a = H.read_checkpoint(f, res.Lambdas[0])
    assert Lc == res.Lambdas[0]
    # the first step is measured on the initial meshes; its checkpoint reproduces its χ
    for name in ("sigma", "Os", "ns", "Ot", "nt", "chi"):
        np.testing.assert_array_equal(getattr(m, name), getattr(prob.mesh, name))
    octx.set_mesh(m)
    octx.upload(F.A_INTER, a)
    np.testing.assert_array_equal(octx.compute_chi(Lc, F.A_INTER, (1e-8, 1e-4)), res.chi[0])


def test_flow_forces_checkpoints_at_required_cutoffs(tmp_path):
    # launcher_1l.jl:59-69, 143-172, 201-209: with a checkpoint interval that never elapses
    # (ct = 24 h), checkpoints are still written at the user's cutoffs and at Λ_f (measure with
    # ct = 0), the steps land on those cutoffs, and both files get the finished flag
    from pffrg import flow as F
    from helpers import small_problem
    from oracle_ctx import OracleContext
    prob = small_problem("square", 2, 6, 5.0)
    meas = H.Measurement(tmp_path / "obs.h5", tmp_path / "cp.h5", wt=24.0, ct=24.0)
    octx = OracleContext(prob)
    res = F.launch_flow(octx, prob, prob.bare_action(), 5.0, 3.0, measurement=meas, cps=[4.0])
    assert res.status == "finished" and res.Lambdas[-1] == 3.0
    assert 4.0 in res.Lambdas
    assert meas.checkpoints == [4.0, 3.0]
    with H.H5File(tmp_path / "cp.h5", "r") as f:
        assert f.read("finished") is True
        assert f.keys("a") == ["3.0", "4.0"]
        Lc, dL, m, a = H.read_checkpoint(f, 3.0)
        assert Lc == 3.0
    with H.H5File(tmp_path / "obs.h5", "r") as f:
        assert f.read("finished") is True
        assert len(f.keys("χ")) == res.accepted
    # the same flow without a measurement hook takes the same steps
    res2 = F.launch_flow(OracleContext(prob), prob, prob.bare_action(), 5.0, 3.0, cps=[4.0])
    assert res2.Lambdas == res.Lambdas
