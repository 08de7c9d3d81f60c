"""C-ABI library checks that need no GPU: it loads, exports every entry point
declared in include/piafs_amd.h, and its input-deck reader accepts the
reference's decks and rejects what the reference rejects (parsing runs on the
host only; no device call is made here)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from conftest import GOLDEN, ROOT
import piafs_amd
from piafs_amd import decks


def _declared():
    txt = open(os.path.join(ROOT, "include", "piafs_amd.h")).read()
    return sorted(set(re.findall(r"\b(piafs_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = piafs_amd.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(piafs_amd.EXPORTS)


def test_header_is_plain_c():
    txt = open(os.path.join(ROOT, "include", "piafs_amd.h")).read()
    assert 'extern "C"' in txt
    for bad in ("torch", "std::", "hipStream_t", "namespace", "template"):
        assert bad not in txt


@pytest.mark.parametrize("builder,kw", [
    ("sod_1d", dict(N=64)),
    ("ns2d_photochem", dict(nx=32, ny=16, n_iter=2)),
    ("ns2d_photochem", dict(nx=16, ny=16, z_mm=1.5)),
    ("isentropic_vortex", dict(N=20)),
])
def test_deck_parses_and_sizes(builder, kw):
    dk = getattr(decks, builder)(**kw)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        n = piafs_amd.estimate_bytes(d)
    pts = 1
    for s in dk.dims:
        pts *= s + 6
    assert n >= 8 * pts * dk.nvars * 4


@pytest.mark.parametrize("example", ["SodShockTube", "ShuOsherProblem", "LaxShockTube"])
def test_reference_example_decks_parse(example):
    assert piafs_amd.estimate_bytes(os.path.join(GOLDEN, "examples", example)) > 0


@pytest.mark.parametrize("mutate,msg", [
    (lambda dk: dk.solver.__setitem__("hyp_space_scheme", "muscl3"), "not a supported spatial interpolation"),
    (lambda dk: dk.solver.__setitem__("nvars", 7), "nvars has to be 11"),
    (lambda dk: dk.solver.__setitem__("time_scheme_type", "rk5"), "not a supported multi-stage"),
    (lambda dk: dk.physics.__setitem__("upwinding", "hllc"), "not a valid upwinding"),
    (lambda dk: dk.solver.__setitem__("par_space_scheme", "3"), "parabolic"),
    (lambda dk: dk.solver.__setitem__("hyp_interp_type", "upwind-char"), "not a supported interpolation"),
    (lambda dk: dk.boundary.__setitem__(0, ("thermal-wall", 0, 1, 0, 0, 0, 0)), "not supported"),
    (lambda dk: (dk.solver.__setitem__("size", (16, 3)), setattr(dk, "x", [dk.x[0], dk.x[1][:3]]),
                 setattr(dk, "u", dk.u[:48]), setattr(dk, "imap", dk.imap[:48])), "grid too small"),
])
def test_deck_errors_are_reported(mutate, msg):
    dk = decks.ns2d_photochem(nx=16, ny=16, n_iter=1)
    mutate(dk)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        with pytest.raises(piafs_amd.PiafsError):
            piafs_amd.estimate_bytes(d)
        err = piafs_amd.lib().piafs_last_error(None).decode()
    assert msg in err, err


def test_every_reference_fixture_deck_is_accepted(ref_runs):
    """Every configuration the shipped reference runs cover (both models,
    component-wise and characteristic interpolation, Roe and rf-char) is on
    this path: the library parses and sizes each deck."""
    from conftest import build_deck
    _, meta = ref_runs
    for name, m in meta.items():
        dk = build_deck(m)
        with tempfile.TemporaryDirectory() as d:
            dk.write(d)
            assert piafs_amd.estimate_bytes(d) > 0, name


def test_shipped_example_decks_are_accepted(examples):
    """The shipped example directories kept under tests/golden/examples (1D
    Euler, and 2D Navier-Stokes with the initial state of their aux/init
    programs: upw5 and walls, photochemistry jets, the flat plate) parse and
    size on this path."""
    import shutil
    base = os.path.join(os.path.dirname(__file__), "golden", "examples")
    for ex in sorted(os.listdir(base)):
        with tempfile.TemporaryDirectory() as d:
            for fn in os.listdir(os.path.join(base, ex)):
                shutil.copy(os.path.join(base, ex, fn), os.path.join(d, fn))
            if not os.path.exists(os.path.join(d, "initial.inp")):
                ni, nj = examples[ex]["size"]
                decks.shipped_example_initial(ex, d, ni, nj)
            assert piafs_amd.estimate_bytes(d) > 0, ex


def test_null_handles_are_input_errors():
    """Every entry point that takes a handle returns 1 (an input error) for a
    NULL handle instead of dereferencing it; no GPU is touched."""
    import ctypes
    L = piafs_amd.lib()
    d0 = ctypes.c_double(0.0)
    calls = [
        L.piafs_step(None, 1), L.piafs_sync(None), L.piafs_step_diag(None, 1, None, None),
        L.piafs_upload_state(None, None), L.piafs_download_state(None, None), L.piafs_set_time(None, d0),
        L.piafs_rhs(None, None, d0, None), L.piafs_rhs_terms(None, None, d0, None, None, None),
        L.piafs_boundary_fill(None, None, d0, None), L.piafs_flux(None, None, d0, 0, None),
        L.piafs_interface_flux(None, None, d0, 0, None), L.piafs_write_binary(None, None),
        L.piafs_solve(None, None, 0, None), L.piafs_autotune(None, None, None),
        L.piafs_batch_step(None, 1), L.piafs_batch_sync(None), L.piafs_batch_download(None, 0, None),
        L.piafs_batch_add(None, None, None), L.piafs_batch_remove(None, 0),
        L.piafs_server_submit(None, None, None), L.piafs_server_poll(None, 0, None, None),
        L.piafs_server_wait_all(None), L.piafs_server_stats(None, None, None, None),
    ]
    assert calls == [1] * len(calls)
    assert L.piafs_destroy(None) == 0


def test_missing_initial_solution_is_an_error():
    dk = decks.sod_1d(N=32)
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        os.remove(os.path.join(d, "initial.inp"))
        with pytest.raises(piafs_amd.PiafsError):
            piafs_amd.estimate_bytes(d)


def _c_blocks(md):
    return re.findall(r"```c\n(.*?)```", md, re.S)


def test_integration_binding_compiles(tmp_path):
    """INTEGRATION.md's reference-side bindings (the RHS binding with its
    interior gather/scatter and the device-buffer function-table wrappers)
    compile with a C compiler against include/piafs_amd.h and link against the
    library; the gather/scatter helpers are run on a 2D and a 1D ghosted
    array.  The four HyPar fields they read (include/hypar.h of the reference)
    are declared in a stand-in struct."""
    md = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = [b for b in _c_blocks(md) if "GpuRHS" in b or "GpuFFunction" in b]
    assert len(blocks) == 2
    src = tmp_path / "binding.c"
    main = r'''
#include <stdio.h>
#include <stdlib.h>
static int check(int ndims, int nx, int ny) {
  int dims[2] = {nx, ny}, g = 3, nv = 5;
  HyPar s = {ndims, nv, g, dims};
  long ng = (long)(nx + 2 * g) * (ndims > 1 ? ny + 2 * g : 1), ni = interior_points(&s);
  double *u = calloc(ng * nv, sizeof(double)), *w = calloc(ng * nv, sizeof(double)), *in = malloc(ni * nv * sizeof(double));
  for (long k = 0; k < ng * nv; k++) u[k] = (double)k;
  gather_interior(u, in, &s);
  scatter_interior(in, w, &s);
  int bad = 0;
  for (long k = 0; k < ng * nv; k++) {
    long p = k / nv; int i = (int)(p % (nx + 2 * g)) - g, j = ndims > 1 ? (int)(p / (nx + 2 * g)) - g : 0;
    int inside = i >= 0 && i < nx && j >= 0 && j < (ndims > 1 ? ny : 1);
    bad += inside ? (w[k] != u[k]) : (w[k] != 0.0);
  }
  bad += in[0] != u[(ndims > 1 ? (long)g * (nx + 2 * g) + g : g) * nv];
  free(u); free(w); free(in);
  return bad;
}
int main(void) {
  int bad = check(2, 7, 5) + check(1, 9, 1);
  printf("%s\n", bad ? "MISMATCH" : "OK");
  return bad != 0;
}
'''
    stand_in = "typedef struct { int ndims, nvars, ghosts; int *dim_local; } HyPar;\n"
    src.write_text(stand_in + blocks[0] + blocks[1] + main)
    exe = tmp_path / "binding"
    lib_dir = os.path.join(ROOT, "piafs_amd")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-Wno-unused-function", "-Wno-unused-variable",
                        "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", lib_dir,
                        "-lpiafs_amd", "-Wl,-rpath," + lib_dir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "OK", (r.stdout, r.stderr)


# --------------------------------------------------------------------------
# ensemble directories (simulation.inp), parsed on the host
# --------------------------------------------------------------------------
def _ens_members():
    return [decks.sod_1d(N=64, n_iter=10, t_final=0.02, scheme="crweno5", interp="components"),
            decks.sod_1d(N=96, n_iter=10, t_final=0.02, scheme="crweno5", interp="components",
                         boundary=[("periodic", 0, 1, 0, 0), ("periodic", 0, -1, 0, 0)]),
            decks.sod_1d(N=80, n_iter=10, t_final=0.02, scheme="crweno5", interp="components")]


def test_ensemble_directory_layout():
    """write_ensemble lays out the reference's ensemble files
    (EnsembleSimulationsDefine.cpp:31-119, ReadInputs.c:272-290,
    InitialSolution.c:32-41, InitializeBoundaries.c:51-75)."""
    ms = _ens_members()
    with tempfile.TemporaryDirectory() as d:
        decks.write_ensemble(d, ms)
        assert sorted(os.listdir(d)) == ["boundary.inp", "boundary_1.inp", "initial_0.inp", "initial_1.inp",
                                         "initial_2.inp", "physics.inp", "simulation.inp", "solver.inp",
                                         "weno.inp"]
        sim = open(os.path.join(d, "simulation.inp")).read().split()
        assert sim == ["begin", "nsims", "3", "end"]
        size = [ln.split()[1:] for ln in open(os.path.join(d, "solver.inp")) if ln.split()[:1] == ["size"]]
        assert size == [["64", "96", "80"]]
    assert [decks.member_tag(k, 12) for k in (0, 7, 11)] == ["_00", "_07", "_11"]
    assert decks.member_tag(0, 1) == ""


def test_ensemble_members_parse_as_their_own_decks():
    """Each member of an ensemble directory reads its own size, initial
    solution and zones: the ensemble's device memory is the sum of the
    members' own decks' (piafs_estimate_bytes parses both on the host)."""
    ms = _ens_members()
    with tempfile.TemporaryDirectory() as d:
        decks.write_ensemble(d, ms)
        total = piafs_amd.estimate_bytes(d)
    own = 0
    for dk in ms:
        with tempfile.TemporaryDirectory() as d:
            dk.write(d)
            own += piafs_amd.estimate_bytes(d)
    assert total == own


@pytest.mark.parametrize("mutate,msg", [
    (lambda d: os.remove(os.path.join(d, "initial_1.inp")), "domain 1: initial solution file not found"),
    (lambda d: os.remove(os.path.join(d, "boundary.inp")), "domain 0: boundary condition file boundary_0.inp or boundary.inp not found"),
    (lambda d: open(os.path.join(d, "simulation.inp"), "w").write("begin\n nsims 0\nend\n"), "invalid value for nsims (0)"),
    (lambda d: open(os.path.join(d, "simulation.inp"), "w").write("nsims 3\n"), "illegal format in"),
])
def test_ensemble_errors_are_reported(mutate, msg):
    with tempfile.TemporaryDirectory() as d:
        decks.write_ensemble(d, _ens_members())
        mutate(d)
        with pytest.raises(piafs_amd.PiafsError) as e:
            piafs_amd.estimate_bytes(d)
        assert msg in str(e.value)


def test_ensemble_create_without_simulation_inp_is_an_error():
    """piafs_ensemble_create reads simulation.inp before touching a device."""
    L = piafs_amd.lib()
    with tempfile.TemporaryDirectory() as d:
        decks.sod_1d(N=64, n_iter=2).write(d)
        h = ctypes.c_void_p()
        assert L.piafs_ensemble_create(d.encode(), 0, ctypes.byref(h)) == 1
        assert L.piafs_batch_last_error(h).decode() == "simulation.inp not found"
        assert L.piafs_batch_destroy(h) == 0
