"""Per-operator parity through the C ABI: the boundary fill, the physical
flux, the upwind interface flux and the three right-hand-side terms of one
evaluation, each against the oracle's piece of the same evaluation
(orc_rhs_parts).  Tolerances: relative to each variable's largest magnitude,
1e-14 for the boundary fill and the flux (copies and single expressions),
1e-12 for the interface flux and the terms (reconstruction weights, line
solves, device exp/log in the viscous and chemistry terms)."""
import tempfile

import numpy as np
import pytest

import _oracle as orc
from conftest import build_deck
import piafs_amd

pytestmark = pytest.mark.gpu

# reference fixtures whose decks cover every boundary type family, both
# models, component-wise / characteristic / compact / upw5 reconstruction and
# all four upwinders
DECKS = [
    "ns2d_chem_crweno_js_24x20",                       # CRWENO5 paired, Roe, viscous, photochemistry
    "ns2d_walls_inout_re200_24x20",                    # slip / no-slip walls, subsonic inflow / outflow
    "ns2d_dirichlet_ambivalent_sponge_rusanov_24x20",  # dirichlet, ambivalent, supersonic outflow, sponge
    "ns2d_chem_supersonic_jet_char_rf_16x20",          # characteristic WENO5, rf-char, supersonic inflow
    "ns2d_crweno_char_flow_20x16",                     # characteristic CRWENO5 (block line solves)
    "ns2d_upw5_cavity_re400_20x20",                    # upw5, lid-driven walls
    "sod_char_roe_N128",                               # euler1d, characteristic WENO5, extrapolate
    "sod_sponge_comp_llf_N64",                         # euler1d, llf-char, sponge
]

TOL_COPY = 1e-14
TOL_TERMS = 1e-12


def _rel_max(got, ref):
    """per-variable max |got - ref| over the variable's largest |ref|"""
    got, ref = got.reshape(-1, got.shape[-1]), ref.reshape(-1, ref.shape[-1])
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref).max(axis=0)
    return np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), err)


@pytest.mark.parametrize("name", DECKS)
def test_operator_pieces_match_oracle(ref_runs, name):
    _, meta = ref_runs
    dk = build_deck(meta[name])
    rng = np.random.default_rng(7)
    u = dk.u * (1.0 + 1e-3 * rng.uniform(-1, 1, dk.u.shape))
    t = 0.45  # inside the photolysis pulse of the chemistry decks
    nd = dk.solver["ndims"]
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        oc = orc.OCase(d)
        P = oc.rhs_parts(u, t)
        rhs_ref = oc.rhs(u, t)
        with piafs_amd.Case(d) as c:
            ug = c.boundary_fill(u, t)
            flux = [c.flux(u, t, k) for k in range(nd)]
            fi = [c.interface_flux(u, t, k) for k in range(nd)]
            hyp, par, src = c.rhs_terms(u, t)
            rhs = c.rhs(u, t)
            after = c.solution()
    # the boundary fill: every ghost point the reference fills, corners zero in both
    assert np.all(_rel_max(ug, P["u_ghosted"]) <= TOL_COPY)
    # the flux, on the positions of the interior lines along each dimension
    g = dk.solver["ghost"]
    dims = list(dk.solver["size"]) if nd == 2 else [dk.solver["size"]]
    shape = [n + 2 * g for n in dims][::-1] + [dk.u.shape[1]]
    for k in range(nd):
        got, ref = flux[k].reshape(shape), P["flux%d" % k].reshape(shape)
        if nd == 2:
            sl = (slice(g, g + dims[1]), slice(None)) if k == 0 else (slice(None), slice(g, g + dims[0]))
            got, ref = got[sl], ref[sl]
        assert np.all(_rel_max(got, ref) <= TOL_COPY), k
    # the upwind interface flux of every line
    for k in range(nd):
        assert fi[k].shape == P["fi%d" % k].shape
        assert np.all(_rel_max(fi[k], P["fi%d" % k]) <= TOL_TERMS), k
    # the three terms, and their sum is the right-hand side the step uses
    for got, key in ((hyp, "hyp"), (par, "par"), (src, "src")):
        assert np.all(_rel_max(got, P[key]) <= TOL_TERMS), key
    assert np.array_equal(((-hyp) + par) + src, rhs)
    assert np.all(_rel_max(rhs, rhs_ref) <= TOL_TERMS)
    # the case's own state is left as it was
    assert np.array_equal(after, dk.u)


@pytest.mark.parametrize("which", ["2d", "1d"])
def test_reflect_boundary_fill_matches_oracle(which):
    """The reflect rule (BCReflect.c:17-50) through piafs_boundary_fill: the
    ghost points equal the oracle's fill bit for bit (a negated copy), which
    test_oracle_pinning pins against the closed form."""
    from test_oracle_pinning import reflect_decks
    dk = reflect_decks()[which]
    rng = np.random.default_rng(17)
    u = dk.u * (1.0 + 1e-2 * rng.uniform(-1, 1, dk.u.shape))
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        ref = orc.OCase(d).rhs_parts(u, 0.0)["u_ghosted"]
        with piafs_amd.Case(d) as c:
            got = c.boundary_fill(u, 0.0)
    assert np.array_equal(got, ref)
