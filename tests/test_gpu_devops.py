"""Device-buffer per-operator entry points (piafs_dev_*) through the C ABI,
on arrays in the reference's layouts held in device memory and a stream of
the caller's, each against the oracle's piece of the same evaluation:

* boundary fill and flux: the ghosted state and the flux at every ghosted
  point of orc_rhs_parts (1e-14 of each variable's largest magnitude);
* InterpolateInterfacesHyp: uL, uR, fL, fR of orc_rhs_interfaces (1e-12);
* Upwind on the oracle's own interface values: the interface flux of
  orc_rhs_parts (1e-12);
* FirstDerivativePar: orc_first_derivative (1e-14);
* the WENO weights: the weights orc_rhs_interfaces captured inside the
  reconstruction (1e-12)."""
import tempfile

import numpy as np
import pytest

import _oracle as orc
from _hip import DevArray, Stream
from conftest import build_deck
import piafs_amd

pytestmark = pytest.mark.gpu

DECKS = [
    "ns2d_chem_crweno_js_24x20",                       # CRWENO5 components, Roe (paired in a step), chemistry
    "ns2d_chem_weno_z_16x24",                          # WENO5, Z weights
    "ns2d_walls_inout_re200_24x20",                    # walls, subsonic inflow / outflow
    "ns2d_chem_char_rf_24x20",                         # characteristic WENO5, rf-char
    "ns2d_crweno_char_flow_20x16",                     # characteristic CRWENO5 (block line solves)
    "ns2d_upw5_cavity_re400_20x20",                    # upw5 (no weights), lid-driven walls
    "ns2d_dirichlet_ambivalent_sponge_rusanov_24x20",  # dirichlet, ambivalent, rusanov
    "sod_char_roe_N128",                               # euler1d, characteristic WENO5
    "sine_crweno_comp_N128",                           # euler1d, CRWENO5 components
]

TOL_COPY = 1e-14
TOL_TERMS = 1e-12


def _rel_max(got, ref):
    got, ref = got.reshape(-1, got.shape[-1]), ref.reshape(-1, ref.shape[-1])
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref).max(axis=0)
    return np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), err)


def _to_ref_iface(a, nd, dir):
    """[line][interface][var] -> the reference's interface-array order"""
    return np.transpose(a, (1, 0, 2)) if (nd == 2 and dir == 1) else a


@pytest.mark.parametrize("name", DECKS)
def test_device_operators_match_oracle(ref_runs, name):
    _, meta = ref_runs
    dk = build_deck(meta[name])
    rng = np.random.default_rng(11)
    u = dk.u * (1.0 + 1e-3 * rng.uniform(-1, 1, dk.u.shape))
    t = 0.45
    nd = dk.solver["ndims"]
    nv = dk.u.shape[1]
    g = dk.solver["ghost"]
    dims = list(dk.solver["size"]) if nd == 2 else [dk.solver["size"]]
    gshape = [n + 2 * g for n in dims][::-1] + [nv]
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        oc = orc.OCase(d)
        P = oc.rhs_parts(u, t)
        IV = [oc.rhs_interfaces(u, t, k) for k in range(nd)]
        fdf = rng.uniform(-1, 1, P["u_ghosted"].shape)
        DF = [oc.first_derivative(fdf, k, b) for k in range(nd) for b in (1, -1)]
        st = Stream()
        bufs = []

        def dev(shape, host=None):
            bufs.append(DevArray(shape, host))
            return bufs[-1]

        try:
            with piafs_amd.Case(d) as c:
                before = c.solution()
                # boundary fill: interior = u, ghosts zero on entry
                ug0 = np.zeros(gshape)
                sl = tuple(slice(g, g + n) for n in dims[::-1])
                ug0[sl] = u.reshape(dims[::-1] + [nv])
                dug = dev(gshape, ug0)
                c.dev_boundary_fill(dug.ptr, t, st.handle)
                duG = dev(gshape, P["u_ghosted"])
                for k in range(nd):
                    N = dims[k]
                    nl = dims[1 - k] if nd == 2 else 1
                    ishape = (nl, N + 1, nv) if (nd == 1 or k == 0) else (N + 1, nl, nv)
                    # flux at every ghosted point
                    df = dev(gshape)
                    c.dev_flux(duG.ptr, k, df.ptr, st.handle)
                    # interface values: uflag 1 interpolates the state, 0 the flux
                    dfC = dev(gshape, P["flux%d" % k])
                    out = {}
                    for key, upw, uflag in (("uL", 1, 1), ("uR", -1, 1), ("fL", 1, 0), ("fR", -1, 0)):
                        out[key] = dev(ishape)
                        c.dev_interp_interfaces(out[key].ptr, duG.ptr if uflag else dfC.ptr, duG.ptr, upw, k, uflag,
                                                st.handle)
                    # upwinding of the oracle's own interface values
                    iv, w = IV[k]
                    din = {key: dev(ishape, _to_ref_iface(iv[key], nd, k)) for key in ("uL", "uR", "fL", "fR")}
                    dfi = dev(ishape)
                    c.dev_upwind(dfi.ptr, din["fL"].ptr, din["fR"].ptr, din["uL"].ptr, din["uR"].ptr, duG.ptr, k, t,
                                 st.handle)
                    # first derivative, both biases
                    dff = dev(gshape, fdf)
                    ddf = [dev(gshape), dev(gshape)]
                    for j, b in enumerate((1, -1)):
                        c.dev_first_derivative_par(ddf[j].ptr, dff.ptr, k, b, st.handle)
                    # weights
                    size = (N + 1) * nl * nv
                    dw = [dev((4, size)) for _ in range(3)]
                    has_w = dk.solver["hyp_space_scheme"] != "upw5"
                    if has_w:
                        c.dev_weno_weights(dfC.ptr, duG.ptr, k, dw[0].ptr, dw[1].ptr, dw[2].ptr, size, st.handle)
                    st.sync()
                    assert np.all(_rel_max(df.download(), P["flux%d" % k]) <= TOL_COPY), ("flux", k)
                    for key in ("uL", "uR", "fL", "fR"):
                        ref = _to_ref_iface(iv[key], nd, k)
                        assert np.all(_rel_max(out[key].download(), ref) <= TOL_TERMS), (key, k)
                    fi_ref = _to_ref_iface(P["fi%d" % k], nd, k)
                    assert np.all(_rel_max(dfi.download(), fi_ref) <= TOL_TERMS), ("upwind", k)
                    for j in range(2):
                        assert np.all(_rel_max(ddf[j].download(), DF[2 * k + j]) <= TOL_COPY), ("fd", k, j)
                    if has_w:
                        for q in range(3):
                            got = dw[q].download().reshape(4, -1, nv)
                            assert np.all(_rel_max(got, w[q]) <= TOL_TERMS), ("weights", k, q)
                st.sync()
                assert np.all(_rel_max(dug.download(), P["u_ghosted"]) <= TOL_COPY), "boundary fill"
                # the case itself is left as it was, and still steps like an untouched one
                assert np.array_equal(c.solution(), before)
                c.step(2)
                stepped = c.solution()
            with piafs_amd.Case(d) as c2:
                c2.step(2)
                assert np.array_equal(stepped, c2.solution())
        finally:
            for b in bufs:
                b.free()
            st.destroy()


def test_device_weights_refuse_upw5(ref_runs):
    _, meta = ref_runs
    dk = build_deck(meta["ns2d_upw5_cavity_re400_20x20"])
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with piafs_amd.Case(d) as c:
            a = DevArray((8,))
            try:
                with pytest.raises(piafs_amd.PiafsError):
                    c.dev_weno_weights(a.ptr, a.ptr, 0, a.ptr, a.ptr, a.ptr, 10 ** 6)
            finally:
                a.free()
