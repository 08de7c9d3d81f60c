"""MI355X-native PIAFS right-hand-side / time-step path.

Python host mirror of the C ABI in ``include/piafs_amd.h``.  The compute path
is ``libpiafs_amd.so`` (hand-written HIP kernels for gfx950); this module only
marshals arguments.  There is no CPU fallback: if the library is missing or no
GPU is present, creating a case raises.

The object model follows the reference driver (src/main.cpp:133-235): a
:class:`Case` is created from a directory of PIAFS input decks (the
reference's ``Simulation::define`` + ``ReadInputs`` ... ``InitializePhysics``),
advanced with :meth:`Case.step` (``TimeStep`` = ``TimeRK``) or solved end to
end with :meth:`Case.solve` (``Solve``), and written with
:meth:`Case.write_binary` (``WriteBinary``).
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpiafs_amd.so")
_lib = None
# every open handle, so that the ones a program leaves open are closed by an
# exit hook while the interpreter and the HIP runtime are both fully alive,
# not by a finalizer during interpreter teardown
_live = weakref.WeakSet()


@atexit.register
def _close_live_handles():
    for h in list(_live):
        try:
            h.close()
        except Exception:
            pass


class PiafsError(RuntimeError):
    pass


class Info(ctypes.Structure):
    _fields_ = [("ndims", ctypes.c_int), ("nvars", ctypes.c_int), ("ghosts", ctypes.c_int),
                ("dim", ctypes.c_int * 3), ("n_iter", ctypes.c_int), ("restart_iter", ctypes.c_int),
                ("nstages", ctypes.c_int), ("dt", ctypes.c_double), ("t", ctypes.c_double),
                ("gamma", ctypes.c_double), ("Re", ctypes.c_double), ("Pr", ctypes.c_double),
                ("model", ctypes.c_int), ("scheme", ctypes.c_int), ("characteristic", ctypes.c_int),
                ("include_chem", ctypes.c_int), ("device_bytes", ctypes.c_size_t)]


EXPORTS = [
    "piafs_create", "piafs_destroy", "piafs_get_info", "piafs_last_error", "piafs_upload_state",
    "piafs_download_state", "piafs_set_time", "piafs_step", "piafs_step_diag", "piafs_sync", "piafs_rhs",
    "piafs_write_binary", "piafs_solve", "piafs_kernel_times", "piafs_tridiag_host",
    "piafs_weno_weights_host", "piafs_batch_create", "piafs_batch_destroy", "piafs_batch_step",
    "piafs_batch_sync", "piafs_batch_use_graph", "piafs_batch_add", "piafs_batch_remove",
    "piafs_batch_steps_done", "piafs_batch_download", "piafs_batch_info", "piafs_batch_last_error",
    "piafs_estimate_bytes", "piafs_autotune", "piafs_batch_autotune", "piafs_server_create", "piafs_server_destroy", "piafs_server_submit",
    "piafs_server_poll", "piafs_server_result", "piafs_server_release", "piafs_server_wait_all",
    "piafs_server_stats", "piafs_server_last_error", "piafs_boundary_fill", "piafs_flux",
    "piafs_interface_flux", "piafs_rhs_terms", "piafs_conservation_errors",
    "piafs_set_derivative_exchange", "piafs_dev_boundary_fill", "piafs_dev_flux", "piafs_dev_interp_interfaces",
    "piafs_dev_upwind", "piafs_dev_first_derivative_par", "piafs_dev_weno_weights", "piafs_tridiag_host_noalloc",
    "piafs_ensemble_create", "piafs_ensemble_solve",
]
# doubles in the device array of piafs_tridiag_host_noalloc (include/piafs_amd.h)
NOALLOC_SCRATCH_DOUBLES = 1 << 22


def lib():
    """Load libpiafs_amd.so (built in place by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PiafsError(f"{LIB_PATH} not built; run __graft_entry__.build() (make -C piafs_amd)")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip, cp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_char_p
    L.piafs_create.argtypes = [cp, ip, ctypes.POINTER(vp)]
    L.piafs_destroy.argtypes = [vp]
    L.piafs_get_info.argtypes = [vp, ctypes.POINTER(Info)]
    L.piafs_last_error.argtypes = [vp]
    L.piafs_last_error.restype = cp
    L.piafs_upload_state.argtypes = [vp, vp]
    L.piafs_download_state.argtypes = [vp, vp]
    L.piafs_set_time.argtypes = [vp, ctypes.c_double]
    L.piafs_step.argtypes = [vp, ip]
    L.piafs_step_diag.argtypes = [vp, ip, vp, vp]
    L.piafs_sync.argtypes = [vp]
    L.piafs_rhs.argtypes = [vp, vp, ctypes.c_double, vp]
    L.piafs_boundary_fill.argtypes = [vp, vp, ctypes.c_double, vp]
    L.piafs_flux.argtypes = [vp, vp, ctypes.c_double, ip, vp]
    L.piafs_interface_flux.argtypes = [vp, vp, ctypes.c_double, ip, vp]
    L.piafs_rhs_terms.argtypes = [vp, vp, ctypes.c_double, vp, vp, vp]
    L.piafs_write_binary.argtypes = [vp, cp]
    L.piafs_conservation_errors.argtypes = [vp, vp]
    L.piafs_set_derivative_exchange.argtypes = [vp, ip]
    L.piafs_dev_boundary_fill.argtypes = [vp, vp, ctypes.c_double, vp]
    L.piafs_dev_flux.argtypes = [vp, vp, ip, vp, vp]
    L.piafs_dev_interp_interfaces.argtypes = [vp, vp, vp, vp, ip, ip, ip, vp]
    L.piafs_dev_upwind.argtypes = [vp, vp, vp, vp, vp, vp, vp, ip, ctypes.c_double, vp]
    L.piafs_dev_first_derivative_par.argtypes = [vp, vp, vp, ip, ip, vp]
    L.piafs_dev_weno_weights.argtypes = [vp, vp, vp, ip, vp, vp, vp, ctypes.c_long, vp]
    L.piafs_solve.argtypes = [vp, cp, ip, dp]
    L.piafs_kernel_times.argtypes = [vp, ip, ctypes.POINTER(cp), dp, ctypes.POINTER(ctypes.c_int)]
    L.piafs_tridiag_host.argtypes = [vp, vp, vp, vp, ip, ip]
    L.piafs_tridiag_host_noalloc.argtypes = [vp, vp, vp, vp, ip, ip]
    L.piafs_weno_weights_host.argtypes = [ip, vp, vp, ip, ctypes.c_double, vp]
    L.piafs_batch_create.argtypes = [ctypes.POINTER(cp), ip, ip, ctypes.POINTER(vp)]
    L.piafs_batch_destroy.argtypes = [vp]
    L.piafs_batch_step.argtypes = [vp, ip]
    L.piafs_batch_sync.argtypes = [vp]
    L.piafs_batch_use_graph.argtypes = [vp, ip]
    L.piafs_batch_add.argtypes = [vp, cp, ctypes.POINTER(ctypes.c_int)]
    L.piafs_batch_remove.argtypes = [vp, ip]
    L.piafs_batch_steps_done.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int)]
    L.piafs_batch_download.argtypes = [vp, ip, vp]
    L.piafs_batch_info.argtypes = [vp, ip, ctypes.POINTER(Info)]
    L.piafs_batch_last_error.argtypes = [vp]
    L.piafs_batch_last_error.restype = cp
    L.piafs_ensemble_create.argtypes = [cp, ip, ctypes.POINTER(vp)]
    L.piafs_ensemble_solve.argtypes = [vp, cp, ip, dp]
    L.piafs_estimate_bytes.argtypes = [cp, ctypes.POINTER(ctypes.c_size_t)]
    L.piafs_autotune.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp]
    L.piafs_batch_autotune.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp]
    L.piafs_server_create.argtypes = [ip, ctypes.c_size_t, ctypes.POINTER(vp)]
    L.piafs_server_destroy.argtypes = [vp]
    L.piafs_server_submit.argtypes = [vp, cp, ctypes.POINTER(ctypes.c_int)]
    L.piafs_server_poll.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_int), dp]
    L.piafs_server_result.argtypes = [vp, ip, vp, ctypes.POINTER(Info)]
    L.piafs_server_release.argtypes = [vp, ip]
    L.piafs_server_wait_all.argtypes = [vp]
    L.piafs_server_wait_all.restype = ctypes.c_int
    L.piafs_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)]
    L.piafs_server_last_error.argtypes = [vp]
    L.piafs_server_last_error.restype = cp
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Case:
    """One PIAFS simulation resident on a GPU."""

    def __init__(self, directory: str | None = None, device: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.piafs_create(directory.encode() if directory else None, int(device), ctypes.byref(h))
        if rc:
            raise PiafsError(f"piafs_create failed ({rc}): {L.piafs_last_error(None).decode()}")
        self._h = h
        self._L = L
        self.directory = directory
        _live.add(self)

    # -- queries --
    @property
    def info(self) -> Info:
        inf = Info()
        self._check(self._L.piafs_get_info(self._h, ctypes.byref(inf)))
        return inf

    @property
    def shape(self):
        inf = self.info
        n = 1
        for d in range(inf.ndims):
            n *= inf.dim[d]
        return (n, inf.nvars)

    @property
    def time(self) -> float:
        return self.info.t

    def _check(self, rc):
        if rc:
            raise PiafsError(f"piafs error {rc}: {self._L.piafs_last_error(self._h).decode()}")

    # -- state --
    def solution(self) -> np.ndarray:
        out = np.empty(self.shape)
        self._check(self._L.piafs_download_state(self._h, _ptr(out)))
        return out

    def set_solution(self, u) -> None:
        u = _f64(u).reshape(self.shape)
        self._check(self._L.piafs_upload_state(self._h, _ptr(u)))

    def set_time(self, t: float) -> None:
        self._check(self._L.piafs_set_time(self._h, float(t)))

    # -- time stepping --
    def step(self, n: int = 1) -> None:
        self._check(self._L.piafs_step(self._h, int(n)))

    def step_diag(self, n: int = 1):
        cfl = np.zeros(n)
        norm = np.zeros(n)
        self._check(self._L.piafs_step_diag(self._h, int(n), _ptr(cfl), _ptr(norm)))
        return cfl, norm

    def sync(self) -> None:
        self._check(self._L.piafs_sync(self._h))

    def rhs(self, u, t: float) -> np.ndarray:
        u = _f64(u).reshape(self.shape)
        out = np.empty(self.shape)
        self._check(self._L.piafs_rhs(self._h, _ptr(u), float(t), _ptr(out)))
        return out

    def _ghosted_shape(self):
        inf = self.info
        g = inf.ghosts
        npts = 1
        for d in range(inf.ndims):
            npts *= inf.dim[d] + 2 * g
        return (npts, inf.nvars)

    def boundary_fill(self, u, t: float) -> np.ndarray:
        """The ghosted state after the boundary fill (piafs_boundary_fill)."""
        u = _f64(u).reshape(self.shape)
        out = np.empty(self._ghosted_shape())
        self._check(self._L.piafs_boundary_fill(self._h, _ptr(u), float(t), _ptr(out)))
        return out

    def flux(self, u, t: float, dim: int) -> np.ndarray:
        """The flux along dimension dim on the ghosted grid (piafs_flux)."""
        u = _f64(u).reshape(self.shape)
        out = np.empty(self._ghosted_shape())
        self._check(self._L.piafs_flux(self._h, _ptr(u), float(t), int(dim), _ptr(out)))
        return out

    def interface_flux(self, u, t: float, dim: int) -> np.ndarray:
        """The upwind interface flux of every line along dim, [line][interface][var]
        (piafs_interface_flux)."""
        u = _f64(u).reshape(self.shape)
        inf = self.info
        n = inf.dim[dim]
        lines = inf.dim[1 - dim] if inf.ndims == 2 else 1
        out = np.empty((lines, n + 1, inf.nvars))
        self._check(self._L.piafs_interface_flux(self._h, _ptr(u), float(t), int(dim), _ptr(out)))
        return out

    def rhs_terms(self, u, t: float):
        """hyp, par, src with rhs = -hyp + par + src (piafs_rhs_terms)."""
        u = _f64(u).reshape(self.shape)
        hyp, par, src = (np.empty(self.shape) for _ in range(3))
        self._check(self._L.piafs_rhs_terms(self._h, _ptr(u), float(t), _ptr(hyp), _ptr(par), _ptr(src)))
        return hyp, par, src

    def write_binary(self, path: str) -> None:
        self._check(self._L.piafs_write_binary(self._h, path.encode()))

    def set_derivative_exchange(self, on: bool) -> None:
        """Derivative ghosts of the viscous term as a run decomposed along the
        periodic dimensions computes them (on) or as a serial run (off, the
        default); piafs_set_derivative_exchange."""
        self._check(self._L.piafs_set_derivative_exchange(self._h, int(bool(on))))

    # -- device-buffer operators: arguments are device addresses (ints) of
    # arrays in the reference's layouts; stream is a hipStream_t (0: the
    # case's own stream); the work is enqueued, not waited for --
    def dev_boundary_fill(self, u_ghosted: int, t: float = 0.0, stream: int = 0) -> None:
        self._check(self._L.piafs_dev_boundary_fill(self._h, u_ghosted, float(t), stream or None))

    def dev_flux(self, u_ghosted: int, dir: int, f_ghosted: int, stream: int = 0) -> None:
        self._check(self._L.piafs_dev_flux(self._h, u_ghosted, int(dir), f_ghosted, stream or None))

    def dev_interp_interfaces(self, fI: int, fC: int, u_ghosted: int, upw: int, dir: int, uflag: int,
                              stream: int = 0) -> None:
        self._check(self._L.piafs_dev_interp_interfaces(self._h, fI, fC, u_ghosted, int(upw), int(dir), int(uflag),
                                                        stream or None))

    def dev_upwind(self, fI: int, fL: int, fR: int, uL: int, uR: int, u_ghosted: int, dir: int, t: float = 0.0,
                   stream: int = 0) -> None:
        self._check(self._L.piafs_dev_upwind(self._h, fI, fL, fR, uL, uR, u_ghosted, int(dir), float(t),
                                             stream or None))

    def dev_first_derivative_par(self, Df: int, f: int, dir: int, bias: int = 1, stream: int = 0) -> None:
        self._check(self._L.piafs_dev_first_derivative_par(self._h, Df, f, int(dir), int(bias), stream or None))

    def dev_weno_weights(self, fC: int, u_ghosted: int, dir: int, w1: int, w2: int, w3: int, size: int,
                         stream: int = 0) -> None:
        self._check(self._L.piafs_dev_weno_weights(self._h, fC, u_ghosted, int(dir), w1, w2, w3, int(size),
                                                   stream or None))

    def conservation_errors(self) -> np.ndarray:
        """Per-variable conservation error after the steps taken so far
        (piafs_conservation_errors; the deck must set conservation_check yes)."""
        out = np.empty(self.info.nvars)
        self._check(self._L.piafs_conservation_errors(self._h, _ptr(out)))
        return out

    def solve(self, directory: str | None = None, verbose: bool = False) -> float:
        secs = ctypes.c_double(0)
        d = directory if directory is not None else self.directory
        self._check(self._L.piafs_solve(self._h, d.encode() if d else None, int(verbose), ctypes.byref(secs)))
        return secs.value

    def profile(self, enable: bool) -> None:
        self._L.piafs_kernel_times(self._h, 1 if enable else 0, None, None, None)

    def autotune(self):
        """Pick the line-solve segment length by timing (piafs_autotune); returns
        (rows per lane, reconstruction ms)."""
        r, ms = ctypes.c_int(0), ctypes.c_double(0)
        self._check(self._L.piafs_autotune(self._h, ctypes.byref(r), ctypes.byref(ms)))
        return r.value, ms.value

    def kernel_times(self) -> dict:
        cap = 16
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        n = ctypes.c_int(cap)
        self._check(self._L.piafs_kernel_times(self._h, -1, names, ms, ctypes.byref(n)))
        return {names[k].decode(): ms[k] for k in range(n.value)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.piafs_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent cases advanced together (piafs_batch_* in the C ABI).

    Every step is one launch per kernel class and RK stage covering all member
    cases, replayed from a captured HIP graph unless ``use_graph(False)``.
    Members are addressed by slot numbers; :meth:`add` and :meth:`remove`
    change the member set between steps (continuous batching)."""

    def __init__(self, directories=(), device: int = 0):
        L = lib()
        directories = list(directories)
        arr = (ctypes.c_char_p * max(1, len(directories)))(*[d.encode() for d in directories])
        h = ctypes.c_void_p()
        rc = L.piafs_batch_create(arr, len(directories), int(device), ctypes.byref(h))
        self._h, self._L = h, L
        if rc:
            msg = L.piafs_batch_last_error(h).decode() if h else ""
            self.close()
            raise PiafsError(f"piafs_batch_create failed ({rc}): {msg}")
        self.slots = list(range(len(directories)))
        _live.add(self)

    @classmethod
    def ensemble(cls, directory: str, device: int = 0) -> "Batch":
        """The members of an ensemble directory (simulation.inp with nsims;
        piafs_ensemble_create), slot i holding member i."""
        L = lib()
        self = cls.__new__(cls)
        h = ctypes.c_void_p()
        rc = L.piafs_ensemble_create(str(directory).encode(), int(device), ctypes.byref(h))
        self._h, self._L = h, L
        if rc:
            msg = L.piafs_batch_last_error(h).decode() if h else ""
            self.close()
            raise PiafsError(f"piafs_ensemble_create failed ({rc}): {msg}")
        self.slots = []
        k = 0
        while True:
            inf = Info()
            if L.piafs_batch_info(h, k, ctypes.byref(inf)):
                break
            self.slots.append(k)
            k += 1
        _live.add(self)
        return self

    def solve(self, directory: str = None, verbose: bool = False) -> float:
        """piafs_ensemble_solve: the reference's run of an ensemble created by
        :meth:`ensemble`, files written into ``directory``; returns the
        solver's wall seconds."""
        secs = ctypes.c_double(0)
        d = None if directory is None else str(directory).encode()
        self._check(self._L.piafs_ensemble_solve(self._h, d, int(bool(verbose)), ctypes.byref(secs)),
                    "ensemble solve")
        return secs.value

    def _check(self, rc, what):
        if rc:
            raise PiafsError(f"{what} failed ({rc}): {self._L.piafs_batch_last_error(self._h).decode()}")

    def step(self, n: int = 1):
        self._check(self._L.piafs_batch_step(self._h, int(n)), "batch step")

    def sync(self):
        self._check(self._L.piafs_batch_sync(self._h), "batch sync")

    def use_graph(self, on: bool = True):
        self._check(self._L.piafs_batch_use_graph(self._h, int(bool(on))), "batch use_graph")

    def add(self, directory: str) -> int:
        k = ctypes.c_int(-1)
        self._check(self._L.piafs_batch_add(self._h, directory.encode(), ctypes.byref(k)), "batch add")
        self.slots.append(k.value)
        return k.value

    def remove(self, k: int):
        self._check(self._L.piafs_batch_remove(self._h, int(k)), "batch remove")
        self.slots.remove(k)

    def autotune(self):
        """piafs_batch_autotune: (rows per lane, reconstruction ms)."""
        r, ms = ctypes.c_int(0), ctypes.c_double(0)
        self._check(self._L.piafs_batch_autotune(self._h, ctypes.byref(r), ctypes.byref(ms)), "batch autotune")
        return r.value, ms.value

    def steps_done(self, k: int) -> int:
        n = ctypes.c_int(0)
        self._check(self._L.piafs_batch_steps_done(self._h, int(k), ctypes.byref(n)), "batch steps_done")
        return n.value

    def info(self, k: int) -> Info:
        inf = Info()
        self._check(self._L.piafs_batch_info(self._h, int(k), ctypes.byref(inf)), "batch info")
        return inf

    def solution(self, k: int) -> np.ndarray:
        inf = self.info(k)
        n = 1
        for d in range(inf.ndims):
            n *= inf.dim[d]
        out = np.empty((n, inf.nvars))
        self._check(self._L.piafs_batch_download(self._h, int(k), _ptr(out)), "batch download")
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._L.piafs_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (piafs_server_* in the C ABI).

    ``submit(dir)`` queues the solve of a deck directory to its n_iter steps
    and returns a request id; requests are admitted while their device memory
    fits ``hbm_budget`` (bytes; 0 = 80% of the free device memory), stay
    resident, and advance together with the other requests of their class.
    ``poll(id)`` gives (state, latency seconds): state 0 queued, 1 running,
    2 done, < 0 failed."""

    QUEUED, RUNNING, DONE = 0, 1, 2

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.piafs_server_create(int(device), int(hbm_budget), ctypes.byref(h))
        self._h, self._L = h, L
        if rc:
            msg = L.piafs_server_last_error(h).decode() if h else ""
            self.close()
            raise PiafsError(f"piafs_server_create failed ({rc}): {msg}")
        _live.add(self)

    def _err(self):
        return self._L.piafs_server_last_error(self._h).decode()

    def submit(self, directory: str) -> int:
        i = ctypes.c_int(-1)
        rc = self._L.piafs_server_submit(self._h, directory.encode(), ctypes.byref(i))
        if rc:
            raise PiafsError(f"submit failed ({rc}): {self._err()}")
        return i.value

    def poll(self, rid: int):
        st, lat = ctypes.c_int(0), ctypes.c_double(0)
        if self._L.piafs_server_poll(self._h, int(rid), ctypes.byref(st), ctypes.byref(lat)):
            raise PiafsError(f"unknown request {rid}")
        return st.value, lat.value

    def result(self, rid: int) -> np.ndarray:
        inf = Info()
        st, _ = self.poll(rid)
        if st < 0:
            raise PiafsError(f"request {rid} failed: {self._err()}")
        if st != self.DONE:
            raise PiafsError(f"request {rid} not finished (state {st})")
        rc = self._L.piafs_server_result(self._h, int(rid), None, ctypes.byref(inf))
        if rc:
            raise PiafsError(f"result failed ({rc}): {self._err()}")
        n = 1
        for d in range(inf.ndims):
            n *= inf.dim[d]
        out = np.empty((n, inf.nvars))
        rc = self._L.piafs_server_result(self._h, int(rid), _ptr(out), None)
        if rc:
            raise PiafsError(f"result failed ({rc}): {self._err()}")
        return out

    def release(self, rid: int):
        self._L.piafs_server_release(self._h, int(rid))

    def wait_all(self):
        self._L.piafs_server_wait_all(self._h)

    def stats(self):
        b, r, q = ctypes.c_size_t(0), ctypes.c_int(0), ctypes.c_int(0)
        self._L.piafs_server_stats(self._h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(q))
        return {"bytes_in_use": b.value, "running": r.value, "queued": q.value}

    def close(self):
        if getattr(self, "_h", None):
            self._L.piafs_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tridiag(a, b, c, x):
    """Batched tridiagonal solve on the GPU; arrays shaped [n, ns] (tridiagLU layout)."""
    a, b, c, x = (_f64(t).copy() for t in (a, b, c, x))
    n, ns = x.shape
    rc = lib().piafs_tridiag_host(_ptr(a), _ptr(b), _ptr(c), _ptr(x), n, ns)
    if rc:
        raise PiafsError(f"tridiag failed ({rc})")
    return x


def weno_weights(kind: int, copt, stencils, eps: float = 1e-6) -> np.ndarray:
    st = _f64(stencils).reshape(-1, 5)
    co = _f64(copt).reshape(3)
    w = np.empty((st.shape[0], 3))
    rc = lib().piafs_weno_weights_host(int(kind), _ptr(co), _ptr(st), st.shape[0], float(eps), _ptr(w))
    if rc:
        raise PiafsError(f"weno_weights failed ({rc})")
    return w


def estimate_bytes(directory: str) -> int:
    n = ctypes.c_size_t(0)
    L = lib()
    rc = L.piafs_estimate_bytes(directory.encode(), ctypes.byref(n))
    if rc:
        raise PiafsError(f"estimate_bytes failed ({rc}): {L.piafs_last_error(None).decode()}")
    return int(n.value)
