"""Parity of the HIP path (through the C ABI) with the CPU restatement and the
reference fixtures.  Tolerances: FP64 relative L2 per variable <= 1e-11 after
the stated step counts; single RHS evaluations and
operator calls <= 1e-12 (differences come only from FMA contraction)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import _oracle as orc
from conftest import GOLDEN, ROOT, build_deck
import piafs_amd
from piafs_amd import decks

pytestmark = pytest.mark.gpu

TOL_STEPS = 1e-11
TOL_RHS = 1e-12


def _rel(a, b):
    """per-variable relative L2, guarded by the variable's scale"""
    num = np.sqrt(((a - b) ** 2).sum(axis=0))
    den = np.sqrt((b ** 2).sum(axis=0))
    scale = np.maximum(den, 1e-300)
    return np.where(den > 0, num / scale, num)


@pytest.mark.parametrize("n,ns", [(257, 300), (6000, 7)])
def test_tridiag_matches_oracle(n, ns):
    """Team solves (n <= 4097) and the sequential solver of longer systems
    against the oracle's tridiagLU."""
    rng = np.random.default_rng(11)
    a = rng.uniform(-1, 1, (n, ns)); c = rng.uniform(-1, 1, (n, ns))
    b = 2.5 + rng.uniform(0, 1, (n, ns)); x = rng.uniform(-1, 1, (n, ns))
    a[0] = 0; c[-1] = 0
    rc, ref = orc.tridiag(a, b, c, x)
    got = piafs_amd.tridiag(a, b, c, x)
    assert rc == 0
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_weno_weights_match_oracle(kind):
    rng = np.random.default_rng(kind)
    st = rng.uniform(-1, 1, (500, 5))
    st[:100] = np.cumsum(rng.uniform(0, 0.01, (100, 5)), axis=1)  # smooth
    for copt in ((0.1, 0.6, 0.3), (0.2, 0.5, 0.3)):
        got = piafs_amd.weno_weights(kind, copt, st)
        ref = np.array([orc.weno_weights(kind, copt, s) for s in st])
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15)


def _wide_stencils(rng, n, lo, hi):
    mag = 10.0 ** rng.uniform(lo, hi, (n, 1))
    st = rng.uniform(-1, 1, (n, 5)) * mag
    q = n // 4
    st[:q] = np.cumsum(rng.uniform(0, 1e-3, (q, 5)), axis=1) * mag[:q]  # smooth
    st[q:q + q // 3] = np.repeat(mag[q:q + q // 3], 5, axis=1)            # constant: every beta zero
    st[2 * q:2 * q + q // 3, 2:] *= 1e6                                    # jumps
    return st


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_weno_weights_bitwise_over_wide_ranges(kind):
    """The weights (piafs_weno_weights_host, include/interpolation.h:255-508),
    whose divisions run the Newton-Raphson core of the target's IEEE double
    division without its range scaling and fix-up steps, against the
    oracle's C divisions.  Inside the envelope where those steps are
    identities -- stencils of magnitude 1e-30 .. 1e30 of both signs (random,
    smooth, constant, with jumps of 1e6), epsilon 1e-6 and 1e-40, both sets of
    optimal weights -- every weight has the same bits.  Far outside it
    (magnitudes up to 1e70 with epsilon 1e-100, where weights fall within
    2^53 of the underflow limit and Z / YC ratios overflow) a weight may
    differ in its last bits below 1e-290, and a row whose reference weights
    are not finite is not finite here either: the interpolated value is then
    the same (NaN, or the same sum)."""
    rng = np.random.default_rng(100 + kind)
    st = _wide_stencils(rng, 6000, -30, 30)
    for eps in (1e-6, 1e-40):
        for copt in ((0.1, 0.6, 0.3), (0.2, 0.5, 0.3)):
            got = piafs_amd.weno_weights(kind, copt, st, eps)
            ref = np.array([orc.weno_weights(kind, copt, s, eps) for s in st])
            assert np.all(np.isfinite(ref))
            same = got.view(np.int64) == ref.view(np.int64)
            assert same.all(), (eps, copt, np.argwhere(~same)[:5], got[~same][:5], ref[~same][:5])
    st = _wide_stencils(rng, 6000, -70, 70)
    for copt in ((0.1, 0.6, 0.3), (0.2, 0.5, 0.3)):
        got = piafs_amd.weno_weights(kind, copt, st, 1e-100)
        ref = np.array([orc.weno_weights(kind, copt, s, 1e-100) for s in st])
        row_fin = np.isfinite(ref).all(axis=1)
        assert not np.isfinite(got[~row_fin]).all(axis=1).any()
        g, r = got[row_fin], ref[row_fin]
        close = (np.abs(g - r) <= 4 * np.spacing(np.abs(r))) | ((np.abs(r) < 1e-290) & (np.abs(g) < 1e-290))
        assert close.all(), (copt, g[~close][:5], r[~close][:5])


RHS_DECKS = {
    "sod_char": lambda: decks.sod_1d(N=64, n_iter=1),
    "sod_rfchar": lambda: decks.sod_1d(N=64, n_iter=1, upwinding="rf-char", mapped=1),
    "sine_crweno": lambda: decks.density_sine_1d(N=64),
    "ns2d_chem_crweno": lambda: decks.ns2d_photochem(nx=32, ny=24, n_iter=1),
    "ns2d_chem_weno_z": lambda: decks.ns2d_photochem(nx=24, ny=32, scheme="weno5", weights="z"),
    "ns2d_chem_yc_par2": lambda: decks.ns2d_photochem(nx=20, ny=20, weights="yc", par="2"),
    "ns2d_chem_mapped": lambda: decks.ns2d_photochem(nx=20, ny=20, weights="mapped"),
    "ns2d_chem_2layer": lambda: decks.ns2d_photochem(nx=16, ny=20, z_mm=1.5),
    "ns2d_nochem": lambda: decks.ns2d_photochem(nx=20, ny=16, include_chemistry=False),
    "vortex": lambda: decks.isentropic_vortex(N=32),
}


@pytest.mark.parametrize("name", sorted(RHS_DECKS))
def test_rhs_matches_oracle(name):
    dk = RHS_DECKS[name]()
    rng = np.random.default_rng(3)
    u = dk.u * (1.0 + 1e-3 * rng.uniform(-1, 1, dk.u.shape))
    t = 0.45  # inside the photolysis pulse
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        ref = orc.OCase(d).rhs(u, t)
        with piafs_amd.Case(d) as c:
            got = c.rhs(u, t)
    scale = np.abs(ref).max(axis=0) + 1e-300
    err = (np.abs(got - ref).max(axis=0) / scale)
    assert np.all(err <= TOL_RHS), err


FIXTURES = [
    "sod_char_roe_N128", "sod_char_mapped_N96", "sod_rfchar_mapped_rk44_N101", "sine_crweno_comp_N128",
    "sine_weno_comp_N100", "ns2d_chem_crweno_js_24x20", "ns2d_chem_weno_z_16x24", "ns2d_chem_par2_yc_16x16",
    "ns2d_chem_mapped_nolimit_16x16", "ns2d_chem_pulse_24x20", "ns2d_chem_2layer_16x16",
    "ns2d_nochem_mapped_20x16", "ns2d_chem_weno_mixed_17x31", "ns2d_chem_crweno_48x40",
    "sod_char_roe_ssprk3_N1024", "ns2d_chem_char_rf_24x20", "ns2d_chem_char_roe_mapped_20x16",
    "ns2d_chem_comp_rf_crweno_20x24", "ns2d_nochem_char_rf_z_16x20", "riemann2d_char_rf_41_40",
    "riemann2d_comp_roe_crweno_33_30", "ns2d_walls_inout_re200_24x20", "ns2d_walls_overlap_crweno_llf_20x24",
    "ns2d_chem_supersonic_jet_char_rf_16x20", "ns2d_chem_subsonic_jet_char_llf_16x20",
    "ns2d_dirichlet_ambivalent_sponge_rusanov_24x20", "riemann2d_char_rusanov_33_30",
    "riemann2d_comp_llf_crweno_33_30", "sod_slipwall_char_llf_N64", "sod_dirichlet_crweno_rusanov_N80",
    "sod_sponge_comp_llf_N64", "sod_char_rusanov_N64", "sod_crweno_char_roe_N64", "sod_crweno_char_rf_mapped_N48",
    "ns2d_crweno_char_flow_20x16", "ns2d_chem_crweno_char_rf_flow_16x12", "ns2d_crweno_char_llf_mapped_inout_20x16",
    "sod_upw5_char_roe_N64", "sod_upw5_comp_rf_N64", "ns2d_chem_upw5_comp_roe_20x16",
    "ns2d_upw5_char_llf_flow_20x16", "ns2d_upw5_cavity_re400_20x20",
    "ns2d_chem_crweno_16x4", "ns2d_chem_char_rf_24x4", "ns2d_crweno_char_flow_16x4",
    "sod_char_roe_rk1fe_N64", "sine_crweno_comp_rk22_N64", "ns2d_chem_crweno_rk33_16x12",
    "ns2d_chem_weno_rk22_12x16",
    "sod_crweno_comp_N8193", "sod_crweno_comp_rf_N5000", "ns2d_chem_crweno_long_4200x8",
]


@pytest.mark.parametrize("name", FIXTURES)
def test_steps_match_reference_fixture(ref_runs, name):
    data, meta = ref_runs
    dk = build_deck(meta[name])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            c.step(int(dk.solver["n_iter"]))
            got = c.solution()
            t = c.time
    ref = data[name]
    err = _rel(got, ref)
    assert np.all(err <= TOL_STEPS), err
    assert abs(t - dk.solver["dt"] * dk.solver["n_iter"]) < 1e-12


def _sum_err(ours, ref, absref):
    """per-variable error of line sums, relative to the sums of absolute
    values (the condition of a sum whose terms cancel)"""
    num = np.sqrt(((ours - ref) ** 2).sum(axis=0))
    den = np.sqrt((absref ** 2).sum(axis=0))
    return np.where(den > 0, num / np.maximum(den, 1e-300), num)


def _close_in_last_digit(ours, printed, fmt):
    a = float(fmt % ours)
    ulp = 10.0 ** (np.floor(np.log10(abs(printed))) - (3 if "3E" in fmt else 4))
    return abs(a - printed) <= ulp * 1.01


@pytest.mark.parametrize("example", ["SodShockTube", "ShuOsherProblem", "LaxShockTube"])
def test_shipped_logs(examples, example):
    rows = examples[example]["rows"]
    with piafs_amd.Case(os.path.join(GOLDEN, "examples", example)) as c:
        cfl, norm = c.step_diag(rows[-1][0])
    for it, t, cf, nm in rows:
        assert _close_in_last_digit(cfl[it - 1], cf, "%1.3E"), (it, cfl[it - 1], cf)
        assert _close_in_last_digit(norm[it - 1], nm, "%1.4E"), (it, norm[it - 1], nm)


SHIPPED = ["Riemann2DCase4", "RadialExpansionWave", "LidDrivenCavity", "Uniform_Supersonic_Jet",
           "Uniform_Subsonic_Jet", "FlatPlateLaminar", "BeamInducedDensityMod_1D_x", "PressurePerturbation_1D_x",
           "DensitySineWave_2D", "InviscidVortexConvection", "BeamInducedDensityMod", "PressurePerturbation",
           "DensitySineWave_1D"]


def _run_shipped(example, d, ni, nj, solver_set=None):
    """Copy the stored deck of a shipped example into d (optionally changing
    solver.inp keywords), write its initial state, run the command-line
    driver there; returns the driver's output and the files it made."""
    src = os.path.join(GOLDEN, "examples", example)
    decks.copy_example_deck(src, d, sorted(os.listdir(src)), solver_set)
    decks.shipped_example_initial(example, d, ni, nj)
    before = set(os.listdir(d))
    r = subprocess.run([os.path.join(ROOT, "piafs_amd", "bin", "piafs_amd"), d], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout, sorted(set(os.listdir(d)) - before)


def _screen_rows(out):
    rows = {}
    for line in out.splitlines():
        t = line.split()
        if t and t[0].startswith("iter="):
            rows[int(line.split("iter=")[1].split()[0])] = (float(line.split("CFL=")[1].split()[0]),
                                                            float(line.split("norm=")[1].split()[0]))
    return rows


@pytest.mark.parametrize("example", SHIPPED)
def test_shipped_examples_through_the_driver(examples, example):
    """A shipped example as a user runs it: its own decks and the initial
    state (and exact solution or beam map) its aux programs write, through the
    command-line driver.
    Riemann2DCase4 / RadialExpansionWave: characteristic WENO5, mapped
    weights, rf-char, SSPRK3, mirror BCs, conservation check, Tecplot output.
    LidDrivenCavity: upw5, no-slip walls with a moving lid, Re 3200.  The
    jets: O2/O3 photochemistry, characteristic WENO5, rf-char, supersonic
    inflow / outflow and subsonic inflow / extrapolation in y, periodic in x,
    numbered binary output and species files (the subsonic jet through its
    photolysis pulse).  FlatPlateLaminar: WENO5 components, Roe, slip and
    no-slip walls, subsonic inflow / outflow, Re 1e5.  The beam-induced
    density modulation with photochemistry (2D, and 1D Euler with three
    z-layers), the pressure perturbations (conservation check), the density
    sine waves (text output, errors against exact.inp) and the isentropic
    vortex (CRWENO5, SSPRK3, Tecplot output, errors and conservation).
    Against the
    reference's run of the same directory (a shortened n_iter): every screen
    line to its printed digits, the names of the files written, the
    conservation errors, and the final solution (norms, a subsample, row and
    column sums)."""
    ref = examples[example]
    ni, nj = ref["size"]
    z = np.load(os.path.join(GOLDEN, "examples2d_ref.npz"))
    with tempfile.TemporaryDirectory() as d:
        out, made = _run_shipped(example, d, ni, nj)
        assert made == ref["files"]
        ops = [f for f in made if f.startswith("op") and not f.startswith("op_species")]
        nv = len(z[example + "_norm"])
        U = decks.read_solution_file(os.path.join(d, ops[-1]), nv).reshape(nj, ni, nv)
        cons = ([float(v) for v in open(os.path.join(d, "conservation.dat")).read().split()[5:]]
                if ref["conservation"] is not None else None)
        nd = 1 if nj == 1 else 2
        errs = [float(v) for v in open(os.path.join(d, "errors.dat")).read().split()][2 * nd + 1:2 * nd + 4]
    rows = _screen_rows(out)
    assert sorted(rows) == [row[0] for row in ref["rows"]]
    for it, _, cf, nm in ref["rows"]:
        assert _close_in_last_digit(rows[it][0], cf, "%1.3E"), (it, rows[it][0], cf)
        assert _close_in_last_digit(rows[it][1], nm, "%1.4E"), (it, rows[it][1], nm)
    for it, _, cf, nm in ref.get("log_rows", []):  # the example's own shipped output.log, where it has one
        assert _close_in_last_digit(rows[it][0], cf, "%1.3E"), (it, rows[it][0], cf)
        assert _close_in_last_digit(rows[it][1], nm, "%1.4E"), (it, rows[it][1], nm)
    if cons is not None:
        np.testing.assert_allclose(cons, ref["conservation"], rtol=0, atol=1e-11)
    shipped = examples.get(example + "_conservation_dat")
    if shipped is not None:  # the example's own conservation.dat (2x2 ranks): roundoff-level errors as well
        np.testing.assert_allclose(cons, shipped["errors"], rtol=0, atol=1e-12)
    if "errors" in ref:  # L1, L2, Linf against the example's exact.inp (zeros without one)
        np.testing.assert_allclose(errs, ref["errors"], rtol=1e-9, atol=1e-15)
    norm = np.sqrt((U.reshape(-1, nv) ** 2).sum(0))
    assert np.all(np.abs(norm - z[example + "_norm"]) <= 1e-12 * z[example + "_norm"])
    st = ref["stride"]
    sub = U[::st, ::st, :].reshape(-1, nv)
    assert np.all(_rel(sub, z[example + "_sub"].reshape(-1, nv)) <= TOL_STEPS)
    assert np.all(_sum_err(U.sum(1), z[example + "_rowsum"], z[example + "_rowabs"]) <= TOL_STEPS)
    assert np.all(_sum_err(U.sum(0), z[example + "_colsum"], z[example + "_colabs"]) <= TOL_STEPS)


@pytest.mark.parametrize("example", ["ShuOsherProblem", "LaxShockTube"])
def test_shipped_1d_conservation_dat(examples, example):
    """Examples/1D_Euler/{ShuOsherProblem,LaxShockTube}/conservation.dat: the
    conservation error of each variable after the full run, through the
    command-line driver (conservation_check yes in the shipped decks), bit
    for bit equal to the serial reference's run of the same deck.  The
    ShuOsherProblem file was written on one rank and equals that run; the
    LaxShockTube file was written on two ranks, whose sums run in another
    order, and agrees with it at roundoff level (2e-15)."""
    d0 = os.path.join(GOLDEN, "examples", example)
    with tempfile.TemporaryDirectory() as d:
        decks.copy_example_deck(d0, d, sorted(os.listdir(d0)))
        r = subprocess.run([os.path.join(ROOT, "piafs_amd", "bin", "piafs_amd"), d], capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        got = [float(v) for v in open(os.path.join(d, "conservation.dat")).read().split()[3:]]
    assert got == examples[example]["conservation_dat_serial"], got
    np.testing.assert_allclose(got, examples[example]["conservation_dat"], rtol=0, atol=2e-15)


COMPARE = os.path.join(ROOT, "oracle", "_ref", "CompareSolutions")


def _write_binary_op(path, dk, u):
    dims = list(dk.solver["size"]) if dk.solver["ndims"] == 2 else [dk.solver["size"]]
    decks.write_binary_op(path, dk.solver["ndims"], dims, dk.x, u)


@pytest.mark.skipif(not os.path.exists(COMPARE), reason="the reference's CompareSolutions is not built")
@pytest.mark.parametrize("name", ["ns2d_chem_crweno_48x40", "riemann2d_char_rf_41_40", "sod_char_roe_N128",
                                  "ns2d_walls_inout_re200_24x20", "ns2d_crweno_char_flow_20x16"])
def test_reference_compare_solutions_tool(ref_runs, name):
    """The reference's own comparison tool (Extras/CompareSolutions.c, built
    from source into oracle/_ref) on the driver's solution file against the
    reference's run of the same deck: relative L1, L2 and Linf norms of the
    difference, over all variables, within the parity bar."""
    data, meta = ref_runs
    dk = build_deck(meta[name])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            c.step(int(dk.solver["n_iter"]))
            c.write_binary(os.path.join(d, "gpu.bin"))
        _write_binary_op(os.path.join(d, "ref.bin"), dk, data[name])
        with open(os.path.join(d, "compare.inp"), "w") as f:
            f.write("gpu.bin ref.bin\n")
        r = subprocess.run([COMPARE], cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        l1, l2, linf = (float(v) for v in open(os.path.join(d, "compare.out")).read().split())
    assert l1 <= TOL_STEPS and l2 <= TOL_STEPS and linf <= TOL_STEPS, (l1, l2, linf)


def test_vortex_errors_match_errors_dat(examples):
    ref = examples["InviscidVortexConvection_errors_dat"]["errors"]
    dk = decks.isentropic_vortex(N=60, n_iter=400, dt=0.05)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            c.step(400)
            u = c.solution()
    np.testing.assert_allclose(decks.error_norms(u, dk.u), ref, rtol=1e-8)


def test_write_binary_layout_and_cli(ref_runs):
    data, meta = ref_runs
    name = "ns2d_chem_crweno_js_24x20"
    dk = build_deck(meta[name])
    dk.solver["op_file_format"] = "binary"
    dk.solver["op_overwrite"] = "yes"
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        cli = os.path.join(ROOT, "piafs_amd", "bin", "piafs_amd")
        r = subprocess.run([cli], cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        nd, nv, dims, x, u = decks.read_binary_op(os.path.join(d, "op.bin"))
    assert (nd, nv, dims) == (2, 11, [24, 20])
    np.testing.assert_array_equal(x, np.concatenate(dk.x))
    assert np.all(_rel(u, data[name]) <= TOL_STEPS)


def test_sod_1024_full_run(ref_runs):
    """1D Sod, N = 1024, WENO5 characteristic + Roe, SSPRK3, 901 steps to t = 0.2."""
    data, meta = ref_runs
    dk = build_deck(meta["sod_char_roe_ssprk3_N1024"])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            c.step(dk.solver["n_iter"])
            got = c.solution()
    assert np.all(_rel(got, data["sod_char_roe_ssprk3_N1024"]) <= TOL_STEPS)


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "ns2d_chem_crweno_1024_200steps_ref.npz")),
                    reason="1024^2 reference fixture not generated")
def test_ns2d_chem_1024_full_run():
    """2D NS + photochemistry, 1024^2, CRWENO5, RK4, 200 steps vs the reference run."""
    z = np.load(os.path.join(GOLDEN, "ns2d_chem_crweno_1024_200steps_ref.npz"))
    dk = decks.ns2d_photochem(nx=1024, n_iter=200)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            c.step(199)
            cfl, norm = c.step_diag(1)
            u = c.solution()
    # the reference's screen line at iter 200 (printed %1.3E / %1.4E)
    assert _close_in_last_digit(cfl[0], float(z["cfl_200"]), "%1.3E")
    assert _close_in_last_digit(norm[0], float(z["norm_200"]), "%1.4E")
    U = u.reshape(1024, 1024, -1)
    np.testing.assert_allclose(np.sqrt((U * U).sum(axis=(0, 1))), z["l2"], rtol=1e-12)
    assert np.all(_rel(U[::32, ::32, :].reshape(-1, U.shape[2]), z["sub"].reshape(-1, U.shape[2])) <= TOL_STEPS)
    assert np.all(_sum_err(U.sum(axis=0), z["colsum"], z["colabs"]) <= TOL_STEPS)
    assert np.all(_sum_err(U.sum(axis=1), z["rowsum"], z["rowabs"]) <= TOL_STEPS)


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "ns2d_nochem_crweno_4096_10steps_ref.npz")),
                    reason="4096^2 reference fixture not generated")
def test_4096_lines_nochem():
    """4096^2 (lines of 4097 interfaces, 256-lane teams spanning four waves),
    CRWENO5 + viscous, chemistry off, RK4 dt 0.0025, 10 steps vs the reference
    run."""
    z = np.load(os.path.join(GOLDEN, "ns2d_nochem_crweno_4096_10steps_ref.npz"))
    dk = decks.ns2d_photochem(nx=4096, n_iter=10, dt=0.0025, include_chemistry=False, Re=1000.0)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            c.step(10)
            u = c.solution()
    U = u.reshape(4096, 4096, -1)
    np.testing.assert_allclose(np.sqrt((U * U).sum(axis=(0, 1))), z["l2"], rtol=1e-12)
    assert np.all(_rel(U[::128, ::128, :].reshape(-1, U.shape[2]), z["sub"].reshape(-1, U.shape[2])) <= TOL_STEPS)
    assert np.all(_sum_err(U.sum(axis=0), z["colsum"], z["colabs"]) <= TOL_STEPS)
    assert np.all(_sum_err(U.sum(axis=1), z["rowsum"], z["rowabs"]) <= TOL_STEPS)


CRWENO_FIXTURES = [f for f in FIXTURES if "crweno" in f]


@pytest.mark.parametrize("name", CRWENO_FIXTURES)
def test_line_solve_relay_from_plain_start(ref_runs, name, monkeypatch):
    """The line solves started from trivial segment estimates (b' = 1, x' = 0
    instead of the scanned ones) need many relay passes and must still land
    on the same bits."""
    data, meta = ref_runs
    dk = build_deck(meta[name])
    n = int(dk.solver["n_iter"])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            c.step(n)
            fast = c.solution()
        monkeypatch.setenv("PIAFS_AMD_LINE_SOLVE_PLAIN_START", "1")
        with piafs_amd.Case(d) as c:
            c.step(n)
            slow = c.solution()
    assert np.array_equal(fast, slow)
    assert np.all(_rel(slow, data[name]) <= TOL_STEPS)


@pytest.mark.parametrize("N", [4098, 4099, 8194])
def test_long_line_window_edges_match_oracle(N):
    """Lines just past one 4096-row window (4097 and 4098 interior rows: a
    last window of one and of two rows, held by lane 0) and one row past two
    windows, solved by the windowed team relay, against the oracle's
    sequential sweep."""
    dk = decks.sod_1d(N=N, n_iter=3, t_final=3 * 0.25 / N, scheme="crweno5", interp="components")
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        oc = orc.OCase(d)
        oc.step(3)
        ref = oc.solution()
        with piafs_amd.Case(d) as c:
            c.step(3)
            got = c.solution()
    assert np.all(_rel(got, ref) <= 1e-13), _rel(got, ref)


_PERIODIC_2D = [("periodic", 0, 1, 0, 0, -1e99, 1e99), ("periodic", 0, -1, 0, 0, -1e99, 1e99),
                ("periodic", 1, 1, -1e99, 1e99, 0, 0), ("periodic", 1, -1, -1e99, 1e99, 0, 0)]
BLOCK_RUN_DECKS = {
    # 1D Sod: 4 runs of 256 rows; 2 ragged runs of 150; the shock and plateaus make reruns
    "sod_crweno_char_N1024": lambda: decks.sod_1d(N=1024, n_iter=12, t_final=12 * 0.2 / 901, scheme="crweno5",
                                                  interp="characteristic"),
    "sod_crweno_char_rf_N300": lambda: decks.sod_1d(N=300, n_iter=12, t_final=12 * 0.2 / 264, scheme="crweno5",
                                                    interp="characteristic", upwinding="rf-char"),
    # x lines of 641 rows in 4 runs with chemistry; 2049 rows in 16 runs of 128
    "ns2d_chem_crweno_char_640x6": lambda: decks.ns2d_open_box(
        nx=640, ny=6, n_iter=3, dt=1e-4, flow=(1.0, 0.3, 0.2, None), scheme="crweno5", interp="characteristic",
        upwinding="roe", include_chemistry=True, Re=1000.0, boundary=_PERIODIC_2D),
    "ns2d_crweno_char_2048x4": lambda: decks.ns2d_open_box(
        nx=2048, ny=4, n_iter=3, dt=2e-5, flow=(1.0, 0.3, 0.2, None), scheme="crweno5", interp="characteristic",
        upwinding="roe", Re=1000.0, boundary=_PERIODIC_2D),
}


@pytest.mark.parametrize("name", sorted(BLOCK_RUN_DECKS))
def test_block_chain_runs_match_oracle(name):
    """Characteristic CRWENO5 lines long enough that k_cw_solve splits each
    block chain into runs relayed bit for bit (runs of at least 128 rows),
    against the oracle's sequential block sweep (blocktridiagLU.c:159-301)."""
    dk = BLOCK_RUN_DECKS[name]()
    n = int(dk.solver["n_iter"])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        oc = orc.OCase(d)
        oc.step(n)
        ref = oc.solution()
        with piafs_amd.Case(d) as c:
            c.step(n)
            got = c.solution()
    assert np.all(np.isfinite(ref))
    assert np.all(_rel(got, ref) <= 1e-11), _rel(got, ref)


def _rel_max(got, ref):
    scale = np.abs(ref).max(axis=tuple(range(ref.ndim - 1))) + 1e-300
    return np.abs(got - ref).max(axis=tuple(range(ref.ndim - 1))) / scale


@pytest.mark.parametrize("nx,ny", [(520, 530), (300, 400)])
def test_species_fluxes_from_the_reconstruction_match_oracle(nx, ny):
    """NS2D with chemistry and paired Roe on lines whose teams have 64 lanes
    (513 .. 1025 rows: the pair's two teams are two waves and meet through an
    LDS counter) or 32 lanes (258 .. 512 rows: one wave) in both sweeps, where
    the reconstruction writes the species' Roe interface fluxes itself (|v_n|
    per interface from k_flux) and k_upwind the flow variables: the interface
    flux of every line (piafs_interface_flux) against the oracle's upwind
    output (orc_rhs_parts, NavierStokes2DUpwind.c:26-92) to 1e-12, and one step
    against the oracle's to 1e-11."""
    dk = decks.ns2d_open_box(nx=nx, ny=ny, n_iter=1, dt=5e-4, flow=(1.0, 0.3, 0.2, None), scheme="crweno5",
                             upwinding="roe", include_chemistry=True, Re=1000.0, boundary=_PERIODIC_2D)
    rng = np.random.default_rng(23)
    u = dk.u * (1.0 + 1e-3 * rng.uniform(-1, 1, dk.u.shape))
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        oc = orc.OCase(d)
        P = oc.rhs_parts(u, 0.45)
        with piafs_amd.Case(d) as c:
            fi = [c.interface_flux(u, 0.45, k) for k in range(2)]
            c.step(1)
            got = c.solution()
        oc.step(1)
        ref = oc.solution()
    for k in range(2):
        assert fi[k].shape == P["fi%d" % k].shape
        assert np.all(_rel_max(fi[k], P["fi%d" % k]) <= 1e-12), (k, _rel_max(fi[k], P["fi%d" % k]))
    assert np.all(_rel(got, ref) <= 1e-11), _rel(got, ref)


def _eps(dk, eps):
    dk.weno["epsilon"] = eps
    return dk


SMALL_EPS_DECKS = {
    # the weights' quotients c_k / (beta_k + eps)^2 reach 1e80 where the data are smooth
    "sod_crweno_comp_eps1e-40": lambda: _eps(decks.sod_1d(N=256, n_iter=20, t_final=20 * 0.2 / 225,
                                                          scheme="crweno5", interp="components"), 1e-40),
    "sod_weno_char_mapped_eps1e-40": lambda: _eps(decks.sod_1d(N=200, n_iter=20, t_final=20 * 0.2 / 176,
                                                               scheme="weno5", interp="characteristic", mapped=1),
                                                  1e-40),
    "ns2d_chem_crweno_mapped_eps1e-40": lambda: _eps(decks.ns2d_open_box(
        nx=48, ny=40, n_iter=4, dt=5e-4, flow=(1.0, 0.3, 0.2, None), scheme="crweno5", weights="mapped",
        include_chemistry=True, Re=1000.0, boundary=_PERIODIC_2D), 1e-40),
    "ns2d_crweno_z_eps1e-30": lambda: _eps(decks.ns2d_open_box(
        nx=40, ny=48, n_iter=4, dt=5e-4, flow=(1.0, 0.3, 0.2, None), scheme="crweno5", weights="z",
        Re=1000.0, boundary=_PERIODIC_2D), 1e-30),
}


@pytest.mark.parametrize("name", sorted(SMALL_EPS_DECKS))
def test_small_weno_epsilon_matches_oracle(name):
    """WENO weights with a small epsilon (weno.inp `epsilon`, WENOInitialize.c),
    where the weights' divisions take operands far from the default's, against
    the oracle after the deck's steps: 1e-13 for Euler, 1e-11 with chemistry."""
    dk = SMALL_EPS_DECKS[name]()
    n = int(dk.solver["n_iter"])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        oc = orc.OCase(d)
        oc.step(n)
        ref = oc.solution()
        with piafs_amd.Case(d) as c:
            c.step(n)
            got = c.solution()
    assert np.all(np.isfinite(ref))
    tol = 1e-13 if name.startswith("sod") else 1e-11
    assert np.all(_rel(got, ref) <= tol), _rel(got, ref)


def test_derivative_exchange_translation_invariant_and_matches_oracle():
    """The derivative-ghost exchange of a run decomposed along the periodic
    dimensions (piafs_set_derivative_exchange): the right-hand side of a
    shifted state is the shifted right-hand side bit for bit, it matches the
    oracle's exchange to 1e-12, and steps match the oracle's to 1e-11; switched
    off, the serial semantics come back."""
    from test_oracle_pinning import _roll, translation_deck
    dk = translation_deck()
    rng = np.random.default_rng(23)
    u = dk.u * (1.0 + 1e-2 * rng.uniform(-1, 1, dk.u.shape))
    us = _roll(u, dk, 3, 5)
    n = int(dk.solver["n_iter"])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        o = orc.OCase(d)
        serial_ref = o.rhs(u, 0.0)
        o.set_derivative_exchange(True)
        ref = o.rhs(u, 0.0)
        o.step(n)
        ref_steps = o.solution()
        with piafs_amd.Case(d) as c:
            c.set_derivative_exchange(True)
            got, got_s = c.rhs(u, 0.0), c.rhs(us, 0.0)
            c.step(n)
            steps = c.solution()
            c.set_derivative_exchange(False)
            serial = c.rhs(u, 0.0)
    assert np.array_equal(_roll(got, dk, 3, 5), got_s)
    scale = np.abs(ref).max(axis=0) + 1e-300
    assert np.all(np.abs(got - ref).max(axis=0) / scale <= TOL_RHS)
    assert np.all(np.abs(serial - serial_ref).max(axis=0) / scale <= TOL_RHS)
    assert not np.array_equal(serial, got)
    assert np.all(_rel(steps, ref_steps) <= TOL_STEPS)
