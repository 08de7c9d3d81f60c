"""Pin the CPU restatement (oracle/) before trusting it as the checker.

* bit-for-bit against reference runs (tests/golden/ref_runs.npz, produced by
  the reference binary built from source; see tests/golden/make_golden.py);
* against the fixtures the reference tree ships: per-iteration CFL and norm of
  Examples/1D_Euler/{SodShockTube,ShuOsherProblem,LaxShockTube}/output.log and
  the error norms of Examples/2D_NavierStokes/InviscidVortexConvection/errors.dat;
* against closed forms: Thomas vs a dense solve, WENO weights on linear data,
  fourth-order derivatives of quartics, free-stream preservation and design
  order of accuracy of the no-limiting schemes.
"""
import os
import tempfile

import numpy as np
import pytest

import _oracle as orc
from conftest import GOLDEN, build_deck
from piafs_amd import decks


def _run_oracle(dk, n=None):
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        c = orc.OCase(d)
        c.step(int(dk.solver["n_iter"]) if n is None else n)
        return c.solution()


@pytest.mark.parametrize("name", [
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
])
def test_oracle_matches_reference_bitwise(ref_runs, name):
    data, meta = ref_runs
    u = _run_oracle(build_deck(meta[name]))
    ref = data[name]
    assert u.shape == ref.shape
    assert np.array_equal(u, ref), f"max diff {np.abs(u - ref).max()}"


@pytest.mark.parametrize("example", ["SodShockTube", "ShuOsherProblem", "LaxShockTube"])
def test_oracle_matches_shipped_logs(examples, example):
    """CFL (%1.3E) and step norm (%1.4E) rows printed by the reference run."""
    rows = examples[example]["rows"]
    d = os.path.join(GOLDEN, "examples", example)
    c = orc.OCase(d)
    last = rows[-1][0]
    cfl, norm = c.step(last, diag=True)
    for it, t, cf, nm in rows:
        assert float("%1.3E" % cfl[it - 1]) == cf, (it, cfl[it - 1], cf)
        assert float("%1.4E" % norm[it - 1]) == nm, (it, norm[it - 1], nm)


def test_oracle_vortex_errors_match_errors_dat(examples):
    """CRWENO5 + tridiagonal + SSPRK3 + Roe + mapped weights, 60^2, t = 20."""
    ref = examples["InviscidVortexConvection_errors_dat"]["errors"]
    dk = decks.isentropic_vortex(N=60, n_iter=400, dt=0.05)
    u = _run_oracle(dk)
    e = decks.error_norms(u, dk.u)
    # the shipped run used 2x2 MPI ranks with an approximate (Jacobi) reduced
    # solve, the serial oracle an exact one: agreement to ~10 digits
    np.testing.assert_allclose(e, ref, rtol=1e-8)


def test_tridiag_matches_dense_solve():
    rng = np.random.default_rng(7)
    n, ns = 37, 9
    a = rng.uniform(-1, 1, (n, ns))
    c = rng.uniform(-1, 1, (n, ns))
    b = 3.0 + rng.uniform(0, 1, (n, ns))
    a[0] = 0.0
    c[-1] = 0.0
    x = rng.uniform(-1, 1, (n, ns))
    rc, sol = orc.tridiag(a, b, c, x)
    assert rc == 0
    for s in range(ns):
        M = np.diag(b[:, s]) + np.diag(a[1:, s], -1) + np.diag(c[:-1, s], 1)
        np.testing.assert_allclose(sol[:, s], np.linalg.solve(M, x[:, s]), rtol=1e-12, atol=1e-13)


def test_tridiag_zero_pivot_flag():
    a = np.zeros((3, 1)); b = np.ones((3, 1)); c = np.zeros((3, 1)); x = np.ones((3, 1))
    b[0, 0] = 0.0
    rc, _ = orc.tridiag(a, b, c, x)
    assert rc == -1


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("copt", [(0.1, 0.6, 0.3), (0.2, 0.5, 0.3)])
def test_weights_optimal_on_linear_data(kind, copt):
    """All three smoothness indicators equal h^2 on linear data, so every
    weight formulation returns the optimal weights."""
    h = 0.37
    st = [1.0 + h * k for k in (-2, -1, 0, 1, 2)]
    w = orc.weno_weights(kind, copt, st)
    np.testing.assert_allclose(w, copt, rtol=1e-12)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_weights_avoid_discontinuity(kind):
    """A jump between m1 and p1 drives the weights of the stencils crossing it
    to ~0 (the left-biased stencil 1 stays smooth)."""
    st = [1.0, 1.0, 1.0, 0.0, 0.0]
    w = orc.weno_weights(kind, (0.1, 0.6, 0.3), st)
    assert abs(w.sum() - 1.0) < 1e-14
    assert w[0] > 0.99


def test_fd4_exact_on_quartic():
    N, g = 20, 3
    x = np.arange(-g, N + g, dtype=np.float64) * 0.1
    f = 3 * x ** 4 - 2 * x ** 3 + x - 5
    df = orc.fd4_line(f, N, g) / 0.1
    np.testing.assert_allclose(df, 12 * x ** 3 - 6 * x ** 2 + 1, rtol=1e-9, atol=1e-9)


def test_free_stream_preserved():
    """A uniform moving state is an exact steady solution: RHS == 0."""
    dk = decks.ns2d_photochem(nx=12, ny=10, include_chemistry=False, Re=500.0)
    dk.u[:, 0] = 1.2
    dk.u[:, 1] = 1.2 * 0.3
    dk.u[:, 2] = 1.2 * -0.2
    dk.u[:, 3] = 2.0
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        c = orc.OCase(d)
        r = c.rhs(dk.u, 0.0)
    assert np.abs(r).max() < 1e-13


@pytest.mark.parametrize("scheme", ["weno5", "crweno5", "upw5"])
def test_linear_schemes_fifth_order(scheme):
    """With no_limiting the reconstructions are the linear fifth-order schemes
    (upw5 is one by construction):
    the advection RHS of a density wave converges at order ~5 in the middle half
    of the line (CRWENO's lines end in explicit WENO5 rows even on a periodic
    grid, which is fourth order at the two ends and decays geometrically)."""
    errs = []
    for N in (32, 64):
        dk = decks.density_sine_1d(N=N, n_iter=1, scheme=scheme, amp=1e-3)
        dk.weno["no_limiting"] = 1
        with tempfile.TemporaryDirectory() as d:
            dk.write(d)
            r = orc.OCase(d).rhs(dk.u, 0.0)
        x = dk.x[0]
        exact = -1.0 * 1e-3 * 2 * np.pi * np.cos(2 * np.pi * x)  # -(rho u)_x with u = 1
        errs.append(np.abs(r[N // 4:3 * N // 4, 0] - exact[N // 4:3 * N // 4]).max())
    order = np.log2(errs[0] / errs[1])
    assert order > 4.5, (errs, order)


@pytest.mark.parametrize("interp", ["components", "characteristic"])
def test_upw5_is_weno5_at_optimal_weights(interp):
    """The optimal weights (0.1, 0.6, 0.3) combine the three WENO5 candidate
    stencils into the single five-point upwind stencil (2, -13, 47, 27, -3)/60,
    so upw5 and WENO5 with no_limiting give the same right-hand side up to
    rounding, on any data: a wrong coefficient or a reversed bias in either
    fails this."""
    dk = decks.sod_1d(N=48, n_iter=1, scheme="upw5", interp=interp)
    rng = np.random.default_rng(11)
    u = dk.u * (1.0 + 0.05 * rng.uniform(-1, 1, dk.u.shape))
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        r_upw = orc.OCase(d).rhs(u, 0.0)
    dk.solver["hyp_space_scheme"] = "weno5"
    dk.weno["no_limiting"] = 1
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        r_weno = orc.OCase(d).rhs(u, 0.0)
    scale = np.abs(r_weno).max(axis=0)
    assert np.all(np.abs(r_upw - r_weno).max(axis=0) <= 1e-12 * scale)
    assert not np.array_equal(r_upw, r_weno)  # different arithmetic, same polynomial


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref",
                                                   "CompareSolutions")),
                    reason="the reference's CompareSolutions is not built")
def test_reference_compare_tool_reports_a_planted_difference(ref_runs):
    """Extras/CompareSolutions (built from source) reads solution files written
    in WriteBinary's layout and reports zero for identical files, and for one
    value changed by a relative 1e-9 an Linf of exactly |change| / max|u|."""
    import json
    import subprocess
    data, meta = ref_runs
    name = "ns2d_chem_crweno_48x40"
    dk = build_deck(meta[name])
    u = data[name]
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "CompareSolutions")
    up = u.copy()
    up[100, 0] *= 1 + 1e-9
    out = []
    with tempfile.TemporaryDirectory() as d:
        for fn, arr in (("a.bin", u), ("b.bin", up)):
            decks.write_binary_op(os.path.join(d, fn), 2, list(dk.solver["size"]), dk.x, arr)
        for pair in ("a.bin a.bin", "b.bin a.bin"):
            open(os.path.join(d, "compare.inp"), "w").write(pair + "\n")
            subprocess.run([exe], cwd=d, capture_output=True, check=True)
            out.append([float(v) for v in open(os.path.join(d, "compare.out")).read().split()])
    assert out[0] == [0.0, 0.0, 0.0]
    assert out[1][2] == abs(up[100, 0] - u[100, 0]) / np.abs(u).max()


@pytest.mark.parametrize("name", ["ns2d_chem_crweno_js_24x20", "ns2d_walls_inout_re200_24x20",
                                  "ns2d_dirichlet_ambivalent_sponge_rusanov_24x20", "sod_char_roe_N128"])
def test_oracle_rhs_parts_recombine(ref_runs, name):
    """orc_rhs_parts, the oracle's pieces for the per-operator entry points:
    the terms recombine bit for bit into orc_rhs (the reference's sum order,
    TimeRHSFunctionExplicit.c:29-77), and the captured interface fluxes give
    the hyperbolic term through hyp = sum_d dxinv_d (fI[k+1] - fI[k])
    (HyperbolicFunction.c:34-127), which pins their [line][interface][var]
    layout."""
    _, meta = ref_runs
    dk = build_deck(meta[name])
    rng = np.random.default_rng(5)
    u = dk.u * (1.0 + 1e-3 * rng.uniform(-1, 1, dk.u.shape))
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        c = orc.OCase(d)
        P = c.rhs_parts(u, 0.45)
        rhs = c.rhs(u, 0.45)
    assert np.array_equal(((0.0 + -1.0 * P["hyp"]) + 1.0 * P["par"]) + 1.0 * P["src"], rhs)
    nd = dk.solver["ndims"]
    dims = list(dk.solver["size"]) if nd == 2 else [dk.solver["size"]]
    nv = dk.u.shape[1]
    # dxinv of the uniform grids these decks use, from the grid coordinates
    inv = [1.0 / (dk.x[k][1] - dk.x[k][0]) for k in range(nd)]
    hyp = np.zeros((dims[1] if nd == 2 else 1, dims[0], nv))
    fi0 = P["fi0"]                                    # [y line][x interface][var]
    hyp = hyp + inv[0] * (fi0[:, 1:, :] - fi0[:, :-1, :])
    if nd == 2:
        fi1 = P["fi1"]                                # [x line][y interface][var]
        hyp = hyp + inv[1] * np.transpose(fi1[:, 1:, :] - fi1[:, :-1, :], (1, 0, 2))
    assert np.allclose(hyp.reshape(-1, nv), P["hyp"], rtol=1e-12, atol=1e-12 * np.abs(P["hyp"]).max())


def test_chemistry_constants_from_formulas():
    """Re and Pr of the 1024^2 chemistry deck follow the reference-scale formulas
    (ChemistryInitialize.c:289-314, NavierStokes2DInitialize.c:204-210)."""
    dk = decks.ns2d_photochem(nx=8, ny=8, n_iter=1)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        c = orc.OCase(d)
        Re, Pr, gam = decks.chemistry_reynolds(dk.chemistry["Ptot"])
        np.testing.assert_allclose(c.param("Re"), Re, rtol=1e-12)
        np.testing.assert_allclose(c.param("Re"), 1000.0, rtol=1e-12)
        np.testing.assert_allclose(c.param("Pr"), Pr, rtol=1e-12)
        np.testing.assert_allclose(c.param("gamma"), 918.45 / 650.0, rtol=1e-15)


def test_riemann2dcase4_conservation_fixture_against_shipped_file(examples):
    """Examples/2D_NavierStokes/Riemann2DCase4/conservation.dat (written on 2x2
    ranks) against the reference's serial run of the same example, the fixture
    the driver test compares with: both are roundoff-level conservation errors
    of each variable."""
    np.testing.assert_allclose(examples["Riemann2DCase4"]["conservation"],
                               examples["Riemann2DCase4_conservation_dat"]["errors"], rtol=0, atol=1e-12)


def reflect_decks():
    """Decks whose every face is a reflect zone (BCReflect.c:17-50): a 2D box
    with a moving gas (two zones on one face, split across it) and a 1D tube.
    A step with them is meaningless (the ghost density is negative, which the
    upwinders take the square root of); the boundary fill alone is pinned."""
    two = decks.ns2d_open_box(nx=14, ny=11, n_iter=1, flow=(1.0, 0.3, 0.2, None), boundary=[
        ("reflect", 0, 1, 0, 0, -1e99, 0.4), ("reflect", 0, 1, 0, 0, 0.4, 1e99), ("reflect", 0, -1, 0, 0, -1e99, 1e99),
        ("reflect", 1, 1, -1e99, 1e99, 0, 0), ("reflect", 1, -1, -1e99, 1e99, 0, 0)])
    one = decks.sod_1d(N=24, n_iter=1, boundary=[("reflect", 0, 1, 0, 0), ("reflect", 0, -1, 0, 0)])
    return {"2d": two, "1d": one}


@pytest.mark.parametrize("which", ["2d", "1d"])
def test_reflect_ghosts_are_negated_mirror(which):
    """The oracle's reflect fill equals the closed form of BCReflect.c:40-46:
    ghost point g-1-b (low face) or N+b (high face) along the zone's dimension
    holds -1 times interior point b (low) or N-1-b (high), every variable; the
    corners, which no zone reaches, stay zero."""
    dk = reflect_decks()[which]
    rng = np.random.default_rng(17)
    u = dk.u * (1.0 + 1e-2 * rng.uniform(-1, 1, dk.u.shape))
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        P = orc.OCase(d).rhs_parts(u, 0.0)
    g, nv = 3, dk.u.shape[1]
    if which == "1d":
        N = dk.solver["size"]
        ug = P["u_ghosted"].reshape(N + 2 * g, nv)
        assert np.array_equal(ug[g:g + N], u)
        for b in range(g):
            assert np.array_equal(ug[g - 1 - b], -1.0 * u[b])
            assert np.array_equal(ug[g + N + b], -1.0 * u[N - 1 - b])
        return
    nx, ny = dk.solver["size"]
    ug = P["u_ghosted"].reshape(ny + 2 * g, nx + 2 * g, nv)
    ui = u.reshape(ny, nx, nv)
    assert np.array_equal(ug[g:g + ny, g:g + nx], ui)
    for b in range(g):
        assert np.array_equal(ug[g:g + ny, g - 1 - b], -1.0 * ui[:, b])
        assert np.array_equal(ug[g:g + ny, g + nx + b], -1.0 * ui[:, nx - 1 - b])
        assert np.array_equal(ug[g - 1 - b, g:g + nx], -1.0 * ui[b, :])
        assert np.array_equal(ug[g + ny + b, g:g + nx], -1.0 * ui[ny - 1 - b, :])
    for cj in (slice(0, g), slice(g + ny, None)):
        for ci in (slice(0, g), slice(g + nx, None)):
            assert not ug[cj, ci].any()


PERIODIC_2D = [("periodic", 0, 1, 0, 0, -1e99, 1e99), ("periodic", 0, -1, 0, 0, -1e99, 1e99),
               ("periodic", 1, 1, -1e99, 1e99, 0, 0), ("periodic", 1, -1, -1e99, 1e99, 0, 0)]


def translation_deck():
    """A viscous periodic 2D box on a grid of spacing 1/8 (every dxinv, ghosts
    included, is exactly 8, so a translation of the state is a symmetry of the
    discrete operator), WENO5 components (its interface values are pointwise
    in the stencil values; the compact scheme's end rows are not), no source."""
    dk = decks.ns2d_open_box(nx=20, ny=16, n_iter=6, dt=0.002, L=19.0 / 8.0, flow=(1.0, 0.3, 0.2, None),
                             Re=200.0, scheme="weno5", boundary=PERIODIC_2D)
    assert all(np.all(np.diff(xd) == 0.125) for xd in dk.x)
    return dk


def _roll(u, dk, sy, sx):
    nx, ny = dk.solver["size"]
    return np.roll(u.reshape(ny, nx, -1), (sy, sx), axis=(0, 1)).reshape(u.shape)


def test_derivative_exchange_is_translation_invariant():
    """With the derivative ghosts exchanged as a run decomposed along the
    periodic dimensions does (MPIExchangeBoundariesnD.c:45-176 at
    NavierStokes2DParabolicFunction.c:91-94), the right-hand side of a
    periodically shifted state is the shifted right-hand side, bit for bit.
    With the serial semantics (one-sided derivatives at the ghost points, the
    other derivative zero there) it is not, near the edges only: the two
    differ only within two points of the boundary, where the viscous stencil
    reaches the ghost layers."""
    dk = translation_deck()
    rng = np.random.default_rng(23)
    u = dk.u * (1.0 + 1e-2 * rng.uniform(-1, 1, dk.u.shape))
    us = _roll(u, dk, 3, 5)
    nx, ny = dk.solver["size"]
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        o = orc.OCase(d)
        serial, serial_s = o.rhs(u, 0.0), o.rhs(us, 0.0)
        o.set_derivative_exchange(True)
        xchg, xchg_s = o.rhs(u, 0.0), o.rhs(us, 0.0)
    assert np.array_equal(_roll(xchg, dk, 3, 5), xchg_s)
    assert not np.array_equal(_roll(serial, dk, 3, 5), serial_s)
    a, b = serial.reshape(ny, nx, -1), xchg.reshape(ny, nx, -1)
    assert np.array_equal(a[2:-2, 2:-2], b[2:-2, 2:-2])
    assert not np.array_equal(a, b)
    # density has no viscous term: the exchange leaves it alone everywhere
    assert np.array_equal(a[..., 0], b[..., 0])


def _ghosted(dk, u):
    """The boundary-filled ghosted state of u (the oracle's rhs_parts)."""
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        return orc.OCase(d).rhs_parts(u, 0.0)["u_ghosted"]


def _interfaces(dk, u, dir, t=0.0):
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        c = orc.OCase(d)
        P = c.rhs_parts(u, t)
        iv, w = c.rhs_interfaces(u, t, dir)
    return P, iv, w


@pytest.mark.parametrize("name,dir", [("sine_weno_comp_N100", 0), ("ns2d_chem_weno_z_16x24", 0),
                                      ("ns2d_chem_weno_z_16x24", 1)])
def test_captured_weights_make_the_interface_values(ref_runs, name, dir):
    """orc_rhs_interfaces: WENO5 (Interp1PrimFifthOrderWENO.c:124-166)
    assembled from the captured weights and the stencils of the boundary-filled
    state and flux reproduces the captured uL, uR, fL, fR bit for bit.  The
    interface values come out of the reconstruction the reference runs pin;
    this pins the weights' set order (LF, LU, RF, RU), their interface-point
    order and the [line][interface][var] layout of both."""
    _, meta = ref_runs
    dk = build_deck(meta[name])
    rng = np.random.default_rng(3)
    u = dk.u * (1.0 + 1e-2 * rng.uniform(-1, 1, dk.u.shape))
    P, iv, w = _interfaces(dk, u, dir)
    nd = dk.solver["ndims"]
    g, nv = 3, dk.u.shape[1]
    dims = list(dk.solver["size"]) if nd == 2 else [dk.solver["size"]]
    N = dims[dir]
    shape = [n + 2 * g for n in dims][::-1] + [nv]
    uG = P["u_ghosted"].reshape(shape)
    fG = P["flux%d" % dir].reshape(shape)
    s6 = 1.0 / 6.0
    for src, (lname, rname), (sl, sr) in ((uG, ("uL", "uR"), (1, 3)), (fG, ("fL", "fR"), (0, 2))):
        if nd == 2:   # lines along dir: [line][position][var]
            lines = src[g:-g, :, :] if dir == 0 else np.transpose(src[:, g:-g, :], (1, 0, 2))
        else:
            lines = src[None, :, :]
        k = np.arange(N + 1)
        pos = lambda o: lines[:, g + k + o, :]                     # noqa: E731
        for side, name_, st in ((0, lname, sl), (1, rname, sr)):
            if side == 0:
                m3, m2, m1, p1, p2 = pos(-3), pos(-2), pos(-1), pos(0), pos(1)
            else:
                m3, m2, m1, p1, p2 = pos(2), pos(1), pos(0), pos(-1), pos(-2)
            f1 = (2 * s6) * m3 + (-7 * s6) * m2 + (11 * s6) * m1
            f2 = (-s6) * m2 + (5 * s6) * m1 + (2 * s6) * p1
            f3 = (2 * s6) * m1 + (5 * s6) * p1 + (-s6) * p2
            nl = lines.shape[0]
            ws = [x[st].reshape((nl, N + 1, nv) if (nd == 1 or dir == 0) else (N + 1, nl, nv)) for x in w]
            if nd == 2 and dir == 1:
                ws = [np.transpose(x, (1, 0, 2)) for x in ws]
            got = ws[0] * f1 + ws[1] * f2 + ws[2] * f3
            assert np.array_equal(got, iv[name_]), (name_, np.abs(got - iv[name_]).max())


@pytest.mark.parametrize("name", ["sod_char_roe_N128", "sod_crweno_char_roe_N64", "ns2d_chem_char_rf_24x20",
                                  "ns2d_crweno_char_flow_20x16", "ns2d_chem_crweno_js_24x20"])
def test_captured_weights_are_normalised(ref_runs, name):
    """Every captured weight set (component-wise and characteristic, WENO5
    and CRWENO5) lies in [0, 1] and sums to one to roundoff."""
    _, meta = ref_runs
    dk = build_deck(meta[name])
    for dir in range(dk.solver["ndims"]):
        _, _, w = _interfaces(dk, dk.u, dir)
        assert min(x.min() for x in w) >= 0.0 and max(x.max() for x in w) <= 1.0
        assert np.allclose(w[0] + w[1] + w[2], 1.0, rtol=0, atol=1e-14)


@pytest.mark.parametrize("interp", ["characteristic", "components"])
def test_interface_values_and_weights_are_biased_the_right_way(interp):
    """On the Sod step (density 1 | 0.125 between cells k0-1 and k0): the
    left-biased value at interface k0 is the left state and the right-biased
    one the right state (ENO stencils); at interface k0+1 the left-biased
    solution weights put everything on the third stencil (cells k0..k0+2, the
    only one without the jump), while the right-biased ones give the third
    stencil (cells k0+1..k0-1, the only one with it) nothing and the other two
    their optimal ratio 0.1 : 0.6."""
    dk = decks.sod_1d(N=64, n_iter=10, interp=interp)
    u = dk.u
    k0 = int(np.argmax(u[:, 0] < 0.5))
    _, iv, w = _interfaces(dk, u, 0)
    assert abs(iv["uL"][0, k0, 0] - 1.0) < 1e-6 and abs(iv["uR"][0, k0, 0] - 0.125) < 1e-6
    LU, RU = 1, 3
    nv = u.shape[1]
    # the variables that jump (the momentum is zero on both sides, so its
    # component-wise weights stay optimal)
    jv = np.abs(u[k0] - u[k0 - 1]) > 0 if interp == "components" else np.ones(nv, bool)
    w3_lu = w[2][LU].reshape(-1, nv)[k0 + 1][jv]
    w_ru = [x[RU].reshape(-1, nv)[k0 + 1][jv] for x in w]
    assert np.all(w3_lu > 0.999), w3_lu
    assert np.all(w_ru[2] < 1e-6) and np.allclose(w_ru[0], 1.0 / 7.0, atol=1e-6), w_ru


def test_interface_layout_across_y():
    """2D, dimension 1: on the Riemann quadrants (a density jump along y in
    every column), the captured [x line][y interface] values of a column sit
    at that column's jump: uL is the state below it and uR the state above."""
    dk = decks.riemann_2d(N=33, n_iter=4, dt=0.004)
    n = 33
    u = dk.u
    _, iv, _ = _interfaces(dk, u, 1)
    U = u.reshape(n, n, -1)  # [y][x][var]
    for i in (3, n - 4):
        col = U[:, i, 0]
        k0 = int(np.argmax(np.abs(np.diff(col)) > 1e-3)) + 1   # first cell above the jump
        # within 1e-4 of the states either side of a jump of about 0.6 (mapped
        # weights leave the stencils across the jump a little weight)
        assert abs(iv["uL"][i, k0, 0] - col[k0 - 1]) < 1e-4 and abs(iv["uR"][i, k0, 0] - col[k0]) < 1e-4


@pytest.mark.parametrize("par", ["4", "2"])
def test_first_derivative_exact_on_polynomials(par):
    """orc_first_derivative: with par_space_scheme 4 the central and one-sided
    fourth-order stencils (FirstDerivativeFourthOrder.c:39-132) are exact for
    quartics at every position of every interior line, ghosts included; with
    "2" (FirstDerivativeFirstOrder.c:20-95, bias +1 and -1) the differences are
    exact for linear data.  Both dimensions; the points off the interior lines
    are not written."""
    dk = decks.ns2d_photochem(nx=12, ny=9, n_iter=2, par=par)
    g, nv, nx, ny = 3, dk.u.shape[1], 12, 9
    X, Y = np.meshgrid(np.arange(-g, nx + g, dtype=float), np.arange(-g, ny + g, dtype=float))
    deg = 4 if par == "4" else 1
    coef = np.arange(1, nv + 1, dtype=float)
    f = np.stack([c * (0.5 + 0.1 * X) ** deg + (1.5 - 0.2 * Y) ** deg for c in coef], axis=-1)
    dfx = np.stack([c * deg * 0.1 * (0.5 + 0.1 * X) ** (deg - 1) for c in coef], axis=-1)
    dfy = np.stack([deg * (-0.2) * (1.5 - 0.2 * Y) ** (deg - 1) + 0 * c for c in coef], axis=-1)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        c = orc.OCase(d)
        for bias in ((1,) if par == "4" else (1, -1)):
            gx = c.first_derivative(f.reshape(-1, nv), 0, bias).reshape(f.shape)
            gy = c.first_derivative(f.reshape(-1, nv), 1, bias).reshape(f.shape)
            assert np.allclose(gx[g:-g, :], dfx[g:-g, :], rtol=0, atol=1e-11)
            assert np.allclose(gy[:, g:-g], dfy[:, g:-g], rtol=0, atol=1e-11)
            assert not gx[:g].any() and not gx[-g:].any() and not gy[:, :g].any() and not gy[:, -g:].any()
