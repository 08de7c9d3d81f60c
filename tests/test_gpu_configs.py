"""The benchmark workloads on the HIP path at their own sizes, through the C ABI.

* 1D Sod, N = 1024, 901 SSPRK3 steps: ``test_gpu_parity.test_sod_1024_full_run``
  (against the reference run).
* 2D NS + photochemistry 1024^2, 200 RK4 steps:
  ``test_gpu_parity.test_ns2d_chem_1024_full_run`` (against the reference run).
* The 256-case ensemble (sides 128-512, Re 1e2-1e4, WENO5/CRWENO5 mixed):
  ``test_ensemble_256_cases`` below.
* The served Poisson stream (30 % Sod with N 512-8192, 70 % 2D sides
  128-1024): ``test_served_stream_config_sizes`` below.
* 4096^2 NS + photochemistry, nvars 11, 50 RK4 steps:
  ``test_4096_chem_matches_oracle_then_conserves`` below.

Tolerances: FP64 relative L2 per variable <= 1e-11 against the oracle;
batched and served results equal the same deck run alone bit for bit (the same
device code runs on the same data)."""
import math
import os
import tempfile
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import _oracle as orc
from conftest import GOLDEN
import piafs_amd
from piafs_amd import decks
from test_gpu_parity import TOL_STEPS, _rel, _sum_err

pytestmark = pytest.mark.gpu


def _oracle_solution(d, steps):
    """The CPU restatement advanced `steps` steps (the ctypes call releases the
    GIL, so several of these run in parallel threads)."""
    o = orc.OCase(d)
    o.step(steps)
    return o.solution()


def _stratified(members, per_scheme=4):
    """Members spread over the size range of each scheme: the member at the
    middle of each size quartile, per scheme."""
    out = []
    for scheme in ("weno5", "crweno5"):
        idx = sorted((k for k, m in enumerate(members) if m.deck.solver["hyp_space_scheme"] == scheme),
                     key=lambda k: members[k].cells)
        for q in range(per_scheme):
            out.append(idx[int((q + 0.5) * len(idx) / per_scheme)])
    return out


ENSEMBLE_STEPS = 2


def test_ensemble_256_cases():
    """The 256-case ensemble (sides log-uniform in [128, 512], non-square, Re
    log-uniform in [1e2, 1e4], beam amplitude in [0.5, 2], WENO5 / CRWENO5
    50/50) advanced by the batched launches replayed from a HIP graph.  Every
    member equals its solo run bitwise; eight members spread over sizes and
    schemes match the oracle."""
    members = decks.ensemble(256, n_iter=ENSEMBLE_STEPS)
    assert {m.deck.solver["hyp_space_scheme"] for m in members} == {"weno5", "crweno5"}
    sides = np.array([m.deck.dims for m in members])
    assert sides.min() >= 128 and sides.max() <= 512
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for m in members:
            dirs.append(os.path.join(tmp, m.name))
            m.deck.write(dirs[-1])
        sample = _stratified(members)
        with ThreadPoolExecutor(len(sample)) as ex:  # the oracle on the CPU while the GPU works
            refs = {k: ex.submit(_oracle_solution, dirs[k], ENSEMBLE_STEPS) for k in sample}
            with piafs_amd.Batch(dirs) as b:
                b.step(ENSEMBLE_STEPS)
                b.sync()
                got = [b.solution(k) for k in range(len(dirs))]
                for k in range(len(dirs)):
                    assert b.steps_done(k) == ENSEMBLE_STEPS
            for k, d in enumerate(dirs):
                with piafs_amd.Case(d) as c:
                    c.step(ENSEMBLE_STEPS)
                    assert np.array_equal(got[k], c.solution()), members[k].name
            for k in sample:
                err = _rel(got[k], refs[k].result())
                assert np.all(err <= TOL_STEPS), (members[k].name, err)


def test_served_stream_config_sizes():
    """A Poisson stream of the serving mix at its own sizes (30 % Sod, N
    log-uniform in [512, 8192]; 70 % 2D NS + photochemistry, sides log-uniform
    in [128, 1024]), each request shortened to 120 (1D) or 8 (2D) steps of its
    own dt, served with continuous batching while later requests arrive.  Every
    result equals the same deck run alone bitwise; the smallest 1D and 2D
    requests match the oracle."""
    n = 16
    reqs = decks.request_stream(n, rate=60.0, steps_1d=120, steps_2d=8)
    is1d = [r.member.deck.solver["ndims"] == 1 for r in reqs]
    assert any(is1d) and not all(is1d)
    with tempfile.TemporaryDirectory() as tmp:
        dirs = []
        for r in reqs:
            dirs.append(os.path.join(tmp, r.member.name))
            r.member.deck.write(dirs[-1])
        small = [min((k for k in range(n) if is1d[k] == one), key=lambda k: reqs[k].member.cells)
                 for one in (True, False)]
        with ThreadPoolExecutor(2) as ex:
            refs = {k: ex.submit(_oracle_solution, dirs[k], reqs[k].member.steps) for k in small}
            with piafs_amd.Server() as srv:
                ids = [None] * n
                t0 = time.perf_counter()

                def feed():
                    for k, r in enumerate(reqs):
                        wait = t0 + r.arrival - time.perf_counter()
                        if wait > 0:
                            time.sleep(wait)
                        ids[k] = srv.submit(dirs[k])

                th = threading.Thread(target=feed)
                th.start()
                th.join()
                srv.wait_all()
                got = []
                for k in range(n):
                    st, lat = srv.poll(ids[k])
                    assert st == srv.DONE, (reqs[k].member.name, st, srv._err())
                    assert lat > 0.0
                    got.append(srv.result(ids[k]))
            for k, d in enumerate(dirs):
                with piafs_amd.Case(d) as c:
                    assert c.info.n_iter == reqs[k].member.steps
                    c.step(reqs[k].member.steps)
                    assert np.array_equal(got[k], c.solution()), reqs[k].member.name
            for k in small:
                err = _rel(got[k], refs[k].result())
                assert np.all(err <= TOL_STEPS), (reqs[k].member.name, err)


C5_FIXTURE = os.path.join(GOLDEN, "ns2d_chem_crweno_4096_2steps_oracle.npz")


@pytest.mark.skipif(not os.path.exists(C5_FIXTURE), reason="4096^2 chemistry fixture not generated")
def test_4096_chem_matches_oracle_then_conserves():
    """The 4096^2 photochemistry deck (nvars 11, CRWENO5 + fourth-order viscous
    terms, Roe, RK4, dt 0.0025): after the fixture's steps the state matches
    the oracle's run of the same deck (norms to 1e-12; stride-128 subsample,
    row and column sums to 1e-11); after all 50 steps the state is finite and
    the variables without a source or viscous term -- density, O2 and CO2 --
    are conserved (periodic box: the boundary fluxes of each line cancel).
    Conservation is measured on exact sums of the states (math.fsum), 1e-13
    relative.  The reference's own measure (CalculateConservationError.c:19,
    piafs_conservation_errors) sums 1.7e7 values in point order, and that
    rounding alone is about 1e-10 for O2, which starts at exactly 1 at every
    point; it is held to 1e-8."""
    z = np.load(C5_FIXTURE)
    steps = int(z["steps"])
    dk = decks.ns2d_photochem(nx=4096, n_iter=50, dt=0.0025)
    dk.solver["conservation_check"] = "yes"
    with tempfile.TemporaryDirectory() as d:
        dk.write(d)
        del dk
        cons = (0, 4, 10)
        with piafs_amd.Case(d) as c:
            assert (c.info.nvars, c.info.dim[0], c.info.dim[1]) == (11, 4096, 4096)
            u0 = c.solution()
            s0 = [math.fsum(u0[:, v]) for v in cons]
            del u0
            c.step(steps)
            U = c.solution().reshape(4096, 4096, -1)
            assert abs(c.time - float(z["t"])) < 1e-15
            np.testing.assert_allclose(np.sqrt((U * U).sum(axis=(0, 1))), z["l2"], rtol=1e-12)
            sub = U[::128, ::128, :].reshape(-1, U.shape[2])
            assert np.all(_rel(sub, z["sub"].reshape(-1, U.shape[2])) <= TOL_STEPS)
            assert np.all(_sum_err(U.sum(axis=0), z["colsum"], z["colabs"]) <= TOL_STEPS)
            assert np.all(_sum_err(U.sum(axis=1), z["rowsum"], z["rowabs"]) <= TOL_STEPS)
            del U, sub
            c.step(50 - steps)
            ce = c.conservation_errors()
            u = c.solution()
    assert np.isfinite(u).all()
    for v, a0 in zip(cons, s0):
        assert abs(math.fsum(u[:, v]) - a0) <= 1e-13 * max(1.0, abs(a0)), (v, math.fsum(u[:, v]), a0)
        assert 0.0 <= ce[v] <= 1e-8, (v, ce)
