"""ctypes binding of oracle/liboracle.so -- the CPU restatement used as the
checker.  Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.orc_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        L.orc_free.argtypes = [vp]
        for f in ("orc_ndims", "orc_nvars", "orc_ghosts", "orc_n_iter"):
            getattr(L, f).argtypes = [vp]
        L.orc_dim.argtypes = [vp, ctypes.c_int]
        L.orc_time.argtypes = [vp]
        L.orc_time.restype = ctypes.c_double
        L.orc_param.argtypes = [vp, ctypes.c_char_p]
        L.orc_param.restype = ctypes.c_double
        L.orc_step.argtypes = [vp, ctypes.c_int, vp, vp]
        L.orc_get_solution.argtypes = [vp, vp]
        L.orc_set_solution.argtypes = [vp, vp]
        L.orc_rhs.argtypes = [vp, vp, ctypes.c_double, vp]
        L.orc_rhs_parts.argtypes = [vp, vp, ctypes.c_double] + [vp] * 8
        L.orc_tridiag.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int]
        L.orc_weno_weights.argtypes = [ctypes.c_int] + [ctypes.c_double] * 9 + [vp]
        L.orc_fd4_line.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
        L.orc_set_derivative_exchange.argtypes = [vp, ctypes.c_int]
        L.orc_rhs_interfaces.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int] + [vp] * 7
        L.orc_first_derivative.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class OCase:
    def __init__(self, d):
        L = lib()
        self.h = ctypes.c_void_p()
        rc = L.orc_load(d.encode(), ctypes.byref(self.h))
        if rc:
            raise RuntimeError(f"oracle load failed: {rc}")
        self.L = L
        nd = L.orc_ndims(self.h)
        n = 1
        for k in range(nd):
            n *= L.orc_dim(self.h, k)
        self.shape = (n, L.orc_nvars(self.h))

    def step(self, n, diag=False):
        if diag:
            cfl, norm = np.zeros(n), np.zeros(n)
            self.L.orc_step(self.h, n, _p(cfl), _p(norm))
            return cfl, norm
        self.L.orc_step(self.h, n, None, None)

    def solution(self):
        out = np.zeros(self.shape)
        self.L.orc_get_solution(self.h, _p(out))
        return out

    def set_solution(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        self.L.orc_set_solution(self.h, _p(u))

    def rhs(self, u, t):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        out = np.zeros(self.shape)
        self.L.orc_rhs(self.h, _p(u), float(t), _p(out))
        return out

    def rhs_parts(self, u, t):
        """Boundary-filled ghosted state, fluxes, interface fluxes and the three
        terms of one right-hand-side evaluation (orc_rhs_parts)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        nd, nv = self.L.orc_ndims(self.h), self.shape[1]
        g = self.L.orc_ghosts(self.h)
        dims = [self.L.orc_dim(self.h, d) for d in range(nd)]
        ng = int(np.prod([n + 2 * g for n in dims]))
        other = lambda d: dims[1 - d] if nd == 2 else 1
        out = {"u_ghosted": np.zeros((ng, nv)), "flux0": np.zeros((ng, nv)),
               "flux1": np.zeros((ng, nv)) if nd == 2 else None,
               "fi0": np.zeros((other(0), dims[0] + 1, nv)),
               "fi1": np.zeros((other(1), dims[1] + 1, nv)) if nd == 2 else None,
               "hyp": np.zeros(self.shape), "par": np.zeros(self.shape), "src": np.zeros(self.shape)}
        ptr = lambda a: _p(a) if a is not None else None
        self.L.orc_rhs_parts(self.h, _p(u), float(t), *[ptr(out[k]) for k in
                             ("u_ghosted", "flux0", "flux1", "fi0", "fi1", "hyp", "par", "src")])
        return out

    def _dims(self):
        nd = self.L.orc_ndims(self.h)
        return nd, [self.L.orc_dim(self.h, d) for d in range(nd)], self.L.orc_ghosts(self.h)

    def rhs_interfaces(self, u, t, dir):
        """uL, uR, fL, fR of every interior line along dir ([line][k][var]) and
        the weights w1, w2, w3 (each [4 sets][interface point][var], sets LF,
        LU, RF, RU, interface points in the reference's order) of one
        right-hand-side evaluation (orc_rhs_interfaces)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.shape)
        nd, dims, g = self._dims()
        nv = self.shape[1]
        N = dims[dir]
        nl = dims[1 - dir] if nd == 2 else 1
        iv = {k: np.zeros((nl, N + 1, nv)) for k in ("uL", "uR", "fL", "fR")}
        w = [np.zeros((4, (N + 1) * nl, nv)) for _ in range(3)]
        rc = self.L.orc_rhs_interfaces(self.h, _p(u), float(t), int(dir), *[_p(iv[k]) for k in ("uL", "uR", "fL", "fR")],
                                       *[_p(x) for x in w])
        assert rc == 0
        return iv, w

    def first_derivative(self, f, dir, bias=1):
        """FirstDerivativePar on a ghosted array (orc_first_derivative); the
        points it does not compute are zero."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        df = np.zeros_like(f)
        self.L.orc_first_derivative(self.h, _p(f), int(dir), int(bias), _p(df))
        return df

    def set_derivative_exchange(self, on):
        self.L.orc_set_derivative_exchange(self.h, int(bool(on)))

    def param(self, name):
        return self.L.orc_param(self.h, name.encode())

    @property
    def time(self):
        return self.L.orc_time(self.h)

    def __del__(self):
        try:
            self.L.orc_free(self.h)
        except Exception:
            pass


def tridiag(a, b, c, x):
    a, b, c, x = (np.ascontiguousarray(t, dtype=np.float64).copy() for t in (a, b, c, x))
    n, ns = x.shape
    rc = lib().orc_tridiag(_p(a), _p(b), _p(c), _p(x), n, ns)
    return rc, x


def weno_weights(kind, copt, st, eps=1e-6):
    w = np.zeros(3)
    lib().orc_weno_weights(int(kind), *[float(c) for c in copt], *[float(s) for s in st], float(eps), _p(w))
    return w


def fd4_line(f, N, g):
    f = np.ascontiguousarray(f, dtype=np.float64)
    df = np.zeros_like(f)
    lib().orc_fd4_line(_p(f), _p(df), N, g)
    return df
