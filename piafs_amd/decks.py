"""Input-deck writer for PIAFS cases.

Writes the reference solver's own input files (``solver.inp``,
``boundary.inp``, ``physics.inp``, ``weno.inp``, ``chemistry.inp``,
``initial.inp``, ``imap.inp``) so that the reference binary, the CPU
restatement and the MI355X path all read byte-identical inputs.

File formats follow the reference: keyword files ``begin ... end``
(src/Simulation/ReadInputs.c:244-429, doc/Input_Files.md:29-52), boundary
zones (src/Simulation/InitializeBoundaries.c:60-200), binary arrays with the
grid first and then the point-major solution (src/IOFunctions/ReadArray.c:
235-262).  Random perturbations use SplitMix64 so the sequence is fixed by
the seed, independent of numpy's generators.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0,1) from SplitMix64 (Steele, Lea, Flood 2014)."""
    out = np.empty(n, dtype=np.float64)
    s = seed & MASK64
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & MASK64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def splitmix64_fast(seed: int, n: int) -> np.ndarray:
    """Vectorised SplitMix64: same stream as :func:`splitmix64`."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        s = np.uint64(seed & MASK64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def _kw_file(path: str, items: dict) -> None:
    with open(path, "w") as f:
        f.write("begin\n")
        for k, v in items.items():
            if isinstance(v, (list, tuple)):
                v = " ".join(str(x) for x in v)
            elif isinstance(v, float):
                v = repr(v)
            f.write(f"  {k:<20s} {v}\n")
        f.write("end\n")


@dataclass
class Deck:
    """A complete case: solver keywords, boundary zones, physics/weno/chemistry
    keywords and the binary initial solution (grid + point-major u)."""
    solver: dict
    boundary: list
    physics: dict
    x: list                      # one coordinate array per dimension
    u: np.ndarray                # interior solution, shape (prod(dim), nvars)
    weno: dict = field(default_factory=dict)
    chemistry: dict | None = None
    imap: np.ndarray | None = None

    def write(self, d: str) -> str:
        os.makedirs(d, exist_ok=True)
        _kw_file(os.path.join(d, "solver.inp"), self.solver)
        with open(os.path.join(d, "boundary.inp"), "w") as f:
            f.write(f"{len(self.boundary)}\n")
            for b in self.boundary:
                f.write(" ".join(str(t) for t in b) + "\n")
        _kw_file(os.path.join(d, "physics.inp"), self.physics)
        if self.weno:
            _kw_file(os.path.join(d, "weno.inp"), self.weno)
        if self.chemistry is not None:
            _kw_file(os.path.join(d, "chemistry.inp"), self.chemistry)
        with open(os.path.join(d, "initial.inp"), "wb") as f:
            for xd in self.x:
                np.ascontiguousarray(xd, dtype=np.float64).tofile(f)
            np.ascontiguousarray(self.u, dtype=np.float64).tofile(f)
        if self.imap is not None:
            with open(os.path.join(d, "imap.inp"), "wb") as f:
                for xd in self.x:
                    np.ascontiguousarray(xd, dtype=np.float64).tofile(f)
                np.ascontiguousarray(self.imap, dtype=np.float64).tofile(f)
        return d

    @property
    def dims(self):
        return [len(xd) for xd in self.x]

    @property
    def nvars(self):
        return int(self.solver["nvars"])


# --------------------------------------------------------------------------
# physical constants needed to pick Ptot for a target Reynolds number
# (src/PhysicalModels/Chemistry/ChemistryInitialize.c:85-135,289-314 and
#  src/PhysicalModels/NavierStokes2D/NavierStokes2DInitialize.c:188-216)
# --------------------------------------------------------------------------

def chemistry_reynolds(Ptot: float, Ti: float = 288.0, f_CO2: float = 0.0,
                       lambda_UV: float = 2.48e-7, theta: float = 0.17 * math.pi / 180,
                       T0: float = 275.0, TS: float = 110.4, TA: float = 245.4, TB: float = 27.6):
    NA, kB = 6.02214076e23, 1.380649e-23
    f_O2 = 1.0 - f_CO2
    R = NA * kB / (f_O2 * 0.032 + f_CO2 * 0.044)
    Cp = f_O2 * 918.45 + f_CO2 * 842.86
    Cv = f_O2 * 650.0 + f_CO2 * 654.5
    gam = Cp / Cv
    mu0 = f_O2 * 1.99e-5 + f_CO2 * 1.45e-5
    kap0 = f_O2 * 2.70e-2 + f_CO2 * 1.58e-2
    kg = 2 * (2 * math.pi / lambda_UV) * math.sin(theta)
    cs = math.sqrt(gam * R * Ti)
    L_ref = 1.0 / kg
    rho_ref = Ptot / (R * Ti)
    mu_ref = mu0 * math.exp(1.5 * math.log(Ti / T0)) * ((T0 + TS) / (Ti + TS))
    kap_ref = kap0 * math.exp(1.5 * math.log(Ti / T0)) * (
        (T0 + TA * math.exp(-TB / T0)) / (Ti + TA * math.exp(-TB / Ti)))
    Re = rho_ref * cs * L_ref / mu_ref
    Pr = gam * R / (gam - 1) * mu_ref / kap_ref
    return Re, Pr, gam


def ptot_for_reynolds(Re_target: float) -> float:
    """Re is proportional to Ptot (rho_ref = Ptot/(R Ti)); scale from 1 atm."""
    Re1, _, _ = chemistry_reynolds(101325.0)
    return 101325.0 * Re_target / Re1


# --------------------------------------------------------------------------
# case builders
# --------------------------------------------------------------------------

def sod_1d(N: int = 1024, n_iter: int | None = None, t_final: float = 0.2, scheme: str = "weno5",
           interp: str = "characteristic", upwinding: str = "roe", rk: str = "ssprk3",
           mapped: int = 0, gamma: float = 1.4, x_from_generator: bool = False,
           op_file_format: str = "none", boundary: list | None = None) -> Deck:
    """Sod shock tube on [0,1]: (rho,u,p) = (1,0,1) | (0.125,0,0.1), mirror
    ("extrapolate") BCs, fixed dt = t_final/n_iter with n_iter = ceil(0.88 (N-1))
    (acoustic CFL about 0.5 for max|u|+c = 2.19 of the exact solution)."""
    if n_iter is None:
        n_iter = int(math.ceil(0.88 * (N - 1)))
    dt = t_final / n_iter
    x = np.arange(N, dtype=np.float64) / (N - 1)
    rho = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    u = np.zeros((N, 3))
    u[:, 0] = rho
    u[:, 2] = p / (gamma - 1.0)
    solver = dict(ndims=1, nvars=3, size=N, ghost=3, n_iter=n_iter, time_scheme="rk",
                  time_scheme_type=rk, hyp_space_scheme=scheme, hyp_interp_type=interp,
                  dt=dt, screen_op_iter=max(n_iter, 1), file_op_iter=n_iter * 10,
                  op_file_format=op_file_format, ip_file_type="binary", op_overwrite="yes",
                  model="euler1d")
    if boundary is None:
        boundary = [("extrapolate", 0, 1, 0, 0), ("extrapolate", 0, -1, 0, 0)]
    physics = dict(gamma=gamma, upwinding=upwinding)
    weno = dict(mapped=mapped, borges=0, yc=0, no_limiting=0, epsilon=1e-6)
    return Deck(solver, boundary, physics, [x], u, weno)


def ns2d_photochem(nx: int = 1024, ny: int | None = None, n_iter: int = 200, dt: float = 0.01,
                   seed: int = 20261202, scheme: str = "crweno5", interp: str = "components",
                   Re: float | None = 1000.0, amplitude: float = 1.0, dx: float | None = None,
                   L: float = 10.0 * math.pi, beam_width: float = 0.1, pert: float = 1e-3,
                   weights: str = "js", par: str = "4", rk: str = "44", z_mm: float = 0.0,
                   upwinding: str = "roe", op_file_format: str = "none", with_imap: bool = True,
                   include_chemistry: bool = True) -> Deck:
    """2D Navier-Stokes + photochemistry on a periodic box.

    Quiescent gas rho = 1 + pert*U(-1,1) (SplitMix64, `seed`), u = v = 0,
    p = 1/gamma_chem; species O2 = 1, O3 = 0.01, others 0; Gaussian beam map
    imap = amplitude*exp(-r^2/(2 (beam_width*L)^2)) centred; Re set through
    Ptot (Re scales with Ptot when chemistry is on)."""
    ny = nx if ny is None else ny
    if dx is None:
        dx = L / nx
    Lx, Ly = nx * dx, ny * dx
    x = np.arange(nx, dtype=np.float64) * dx
    y = np.arange(ny, dtype=np.float64) * dx
    _, _, gam = chemistry_reynolds(101325.0)
    nz = int(math.ceil(z_mm * 0.001 * 20 / 0.03)) + 1 if include_chemistry else 0
    nvars = 4 + 7 * nz if include_chemistry else 4
    g = gam if include_chemistry else 1.4
    npts = nx * ny
    r = splitmix64_fast(seed, npts)
    rho = 1.0 + pert * (2.0 * r - 1.0)
    u = np.zeros((npts, nvars))
    u[:, 0] = rho
    u[:, 3] = (1.0 / g) / (g - 1.0)
    if include_chemistry:
        for iz in range(nz):
            u[:, 4 + 7 * iz + 0] = 1.0
            u[:, 4 + 7 * iz + 1] = 0.01
    solver = dict(ndims=2, nvars=nvars, size=(nx, ny), ghost=3, n_iter=n_iter, time_scheme="rk",
                  time_scheme_type=rk, hyp_space_scheme=scheme, hyp_interp_type=interp,
                  par_space_scheme=par, dt=dt, screen_op_iter=max(n_iter, 1),
                  file_op_iter=n_iter * 10, op_file_format=op_file_format, ip_file_type="binary",
                  op_overwrite="yes", model="navierstokes2d")
    boundary = [("periodic", 0, 1, 0, 0, -1e99, 1e99), ("periodic", 0, -1, 0, 0, -1e99, 1e99),
                ("periodic", 1, 1, -1e99, 1e99, 0, 0), ("periodic", 1, -1, -1e99, 1e99, 0, 0)]
    physics = dict(gamma=1.4, upwinding=upwinding, include_chemistry="yes" if include_chemistry else "no")
    chem = None
    if include_chemistry:
        physics["Re"] = 1000.0 if Re is None else float(Re)
        chem = dict(f_CO2=0.0, f_O3=0.01, z_mm=z_mm, nz=20,
                    Ptot=ptot_for_reynolds(Re) if Re is not None else 101325.0)
    elif Re is not None:
        physics["Re"] = float(Re)
        physics["Pr"] = 0.72
    weno = dict(mapped=int(weights == "mapped"), borges=int(weights == "z"), yc=int(weights == "yc"),
                no_limiting=int(weights == "none"), epsilon=1e-6)
    imap = None
    if include_chemistry and with_imap:
        X, Y = np.meshgrid(x, y, indexing="xy")          # point-major, x fastest
        r2 = (X - 0.5 * Lx) ** 2 + (Y - 0.5 * Ly) ** 2
        w = beam_width * L
        imap = (amplitude * np.exp(-r2 / (2.0 * w * w))).reshape(-1)
    return Deck(solver, boundary, physics, [x, y], u, weno, chem, imap)


def ns2d_open_box(nx: int = 24, ny: int = 20, n_iter: int = 20, dt: float = 0.002, boundary: list | None = None,
                  flow=(1.0, 0.3, 0.0, None), bump: float = 0.05, pert: float = 1e-3, seed: int = 20261204,
                  scheme: str = "weno5", interp: str = "components", upwinding: str = "roe",
                  Re: float | None = None, rk: str = "44", weights: str = "js", par: str = "4",
                  include_chemistry: bool = False, L: float = 1.0) -> Deck:
    """2D Navier-Stokes on [0, L] x [0, L ny/nx] (grid points x_i = i h,
    h = L/(nx-1)) with open or wall boundaries: a uniform flow (rho, u, v, p) =
    `flow` (p defaults to 1/gamma) carrying a Gaussian bump of relative size
    `bump` in density and pressure at (0.35 L, 0.5 L ny/nx), density noise of
    size `pert` (SplitMix64, `seed`).  `boundary` lists the zones as written
    to boundary.inp: (type, dim, face, xmin, xmax, ymin, ymax, values...).
    With chemistry the species start at O2 = 1, O3 = 0.01 and Re is set
    through Ptot."""
    h = L / (nx - 1)
    x = np.arange(nx, dtype=np.float64) * h
    y = np.arange(ny, dtype=np.float64) * h
    _, _, gchem = chemistry_reynolds(101325.0)
    g = gchem if include_chemistry else 1.4
    nvars = 11 if include_chemistry else 4
    rho0, u0, v0, p0 = flow
    p0 = (1.0 / g) if p0 is None else p0
    X, Y = np.meshgrid(x, y, indexing="xy")
    r2 = ((X - 0.35 * L) ** 2 + (Y - 0.5 * y[-1]) ** 2) / (0.1 * L) ** 2
    b = bump * np.exp(-r2)
    noise = pert * (2.0 * splitmix64_fast(seed, nx * ny).reshape(ny, nx) - 1.0)
    rho = rho0 * (1.0 + b + noise)
    P = p0 * (1.0 + b)
    u = np.zeros((nx * ny, nvars))
    u[:, 0] = rho.reshape(-1)
    u[:, 1] = (rho * u0).reshape(-1)
    u[:, 2] = (rho * v0).reshape(-1)
    u[:, 3] = (P / (g - 1.0) + 0.5 * rho * (u0 * u0 + v0 * v0)).reshape(-1)
    if include_chemistry:
        u[:, 4] = 1.0
        u[:, 5] = 0.01
    solver = dict(ndims=2, nvars=nvars, size=(nx, ny), ghost=3, n_iter=n_iter, time_scheme="rk",
                  time_scheme_type=rk, hyp_space_scheme=scheme, hyp_interp_type=interp,
                  par_space_scheme=par, dt=dt, screen_op_iter=max(n_iter, 1), file_op_iter=n_iter * 10,
                  op_file_format="none", ip_file_type="binary", op_overwrite="yes", model="navierstokes2d")
    if boundary is None:
        boundary = [("extrapolate", 0, 1, 0, 0, -1e99, 1e99), ("extrapolate", 0, -1, 0, 0, -1e99, 1e99),
                    ("extrapolate", 1, 1, -1e99, 1e99, 0, 0), ("extrapolate", 1, -1, -1e99, 1e99, 0, 0)]
    physics = dict(gamma=1.4, upwinding=upwinding, include_chemistry="yes" if include_chemistry else "no")
    chem, imap = None, None
    if include_chemistry:
        physics["Re"] = 1000.0 if Re is None else float(Re)
        chem = dict(f_CO2=0.0, f_O3=0.01, z_mm=0.0, nz=20,
                    Ptot=ptot_for_reynolds(Re) if Re is not None else 101325.0)
        imap = (np.exp(-r2 / 4.0)).reshape(-1)
    elif Re is not None:
        physics["Re"] = float(Re)
        physics["Pr"] = 0.72
    weno = dict(mapped=int(weights == "mapped"), borges=int(weights == "z"), yc=int(weights == "yc"),
                no_limiting=int(weights == "none"), epsilon=1e-6)
    return Deck(solver, boundary, physics, [x, y], u, weno, chem, imap)


def example_2d_initial(name: str, d: str, NI: int, NJ: int, gamma: float = 1.4) -> None:
    """initial.inp of Examples/2D_NavierStokes/<name> (Riemann2DCase4 or
    RadialExpansionWave) as the example's aux/init.c writes it: ASCII with
    "%lf " (six decimals), x then y, then each variable with x fastest."""
    X = np.zeros((NJ, NI))
    Y = np.zeros((NJ, NI))
    if name == "Riemann2DCase4":
        dx, dy = 1.0 / (NI - 1), 1.0 / (NJ - 1)
        x = np.array([-0.5 + i * dx for i in range(NI)])
        y = np.array([-0.5 + j * dy for j in range(NJ)])
    elif name == "RadialExpansionWave":
        dx, dy = 8.0 / (NI - 1), 8.0 / (NJ - 1)
        x = np.array([-4.0 + i * dx for i in range(NI)])
        y = np.array([-4.0 + j * dy for j in range(NJ)])
    else:
        raise ValueError(name)
    u = np.zeros((4, NJ, NI))
    for j in range(NJ):
        for i in range(NI):
            xi, yj = x[i], y[j]
            if name == "Riemann2DCase4":
                if xi < 0 and yj < 0:
                    rho, P, uu, vv = 1.1, 1.1, 0.8939, 0.8939
                elif xi >= 0 and yj < 0:
                    rho, P, uu, vv = 0.5065, 0.35, 0.0, 0.8939
                elif xi < 0 and yj >= 0:
                    rho, P, uu, vv = 0.5065, 0.35, 0.8939, 0.0
                else:
                    rho, P, uu, vv = 1.1, 1.1, 0.0, 0.0
            else:
                r = math.sqrt(xi * xi + yj * yj)
                if r < 0.5:
                    q = 0.0
                elif r < 1.5:
                    with np.errstate(divide="ignore"):  # r = 0.5 divides by zero, as C does (tanh(-inf))
                        arg = np.float64(r - 1) / np.float64(0.25 - (r - 1) * (r - 1))
                    q = (1.0 / gamma) * (1.0 + math.tanh(arg))
                else:
                    q = 2.0 / gamma
                uu, vv = (0.0, 0.0) if r == 0 else (xi * q / r, yj * q / r)
                a = 1.0 - (gamma - 1.0) * q / 2.0
                rho = gamma * math.exp((2 / (gamma - 1.0)) * math.log(a))
                P = rho * a * a / gamma
            u[0, j, i] = rho
            u[1, j, i] = rho * uu
            u[2, j, i] = rho * vv
            u[3, j, i] = P / (gamma - 1.0) + 0.5 * rho * (uu * uu + vv * vv)
    with open(os.path.join(d, "initial.inp"), "w") as f:
        f.write("".join("%f " % v for v in x) + "\n")
        f.write("".join("%f " % v for v in y) + "\n")
        for k in range(4):
            f.write("".join("%f " % v for v in u[k].reshape(-1)) + "\n")


def _write_binary_initial(d: str, x, y, U) -> None:
    """initial.inp in the binary layout the examples' aux/init.c write
    (ReadArray.c): f64 x[NI], f64 y[NJ], then U point-major (x fastest, the
    nvars components of a point contiguous)."""
    with open(os.path.join(d, "initial.inp"), "wb") as f:
        np.asarray(x, dtype=np.float64).tofile(f)
        np.asarray(y, dtype=np.float64).tofile(f)
        np.ascontiguousarray(U, dtype=np.float64).tofile(f)


def _read_chem_keys(path: str) -> dict:
    """The begin/end keyword file chemistry.inp as the jets' aux/init.C read
    it: a number after each known keyword, one skipped token after any other."""
    known = {"lambda_UV", "theta", "f_CO2", "f_O3", "Ptot", "Ti", "Lz", "z_mm", "nz"}
    toks = open(path).read().split()
    out, k = {}, 1
    while k < len(toks) and toks[k] != "end":
        w = toks[k]
        if w in known:
            out[w] = float(toks[k + 1])
        k += 2
    return out


def example_jet_initial(name: str, d: str, NI: int, NJ: int) -> None:
    """initial.inp of Examples/2D_NavierStokes/Uniform_{Supersonic,Subsonic}_Jet
    as their aux/init.C writes it (binary): the reference scales of the O2/O3
    photochemistry (L_ref = 1/k_g, v_ref the sound speed, P_ref = rho_ref
    v_ref^2), a uniform gas of 1e7 Pa moving at 530 m/s along y on
    [-5, 5] mm x [0, 2] mm (supersonic) or [0, 10] mm (subsonic), and the
    species ratios O2 = 1, O3 = n_O3/n_O2, CO2 = n_CO2/n_O2 in each of the
    z_i + 1 layers.  The chemistry.inp of `d` supplies lambda_UV, theta, the
    fractions, Lz, z_mm and nz, as it does for the C program."""
    if name not in ("Uniform_Supersonic_Jet", "Uniform_Subsonic_Jet"):
        raise ValueError(name)
    sup = name == "Uniform_Supersonic_Jet"
    pi = 4.0 * math.atan(1.0)
    NA, kB = 6.02214076e23, 1.380649e-23
    M_O2, M_CO2 = 0.032, 0.044
    Cp_O2, Cv_O2, Cp_CO2, Cv_CO2 = 918.45, 650.0, 842.86, 654.5
    ck = _read_chem_keys(os.path.join(d, "chemistry.inp"))
    lambda_UV = ck.get("lambda_UV", 2.48e-7)
    theta = ck.get("theta", 0.17 * pi / 180)
    f_CO2, f_O3 = ck.get("f_CO2", 0.0), ck.get("f_O3", 0.05)
    if sup:
        Ptot, Ti = 101325.0, 288.0
    else:
        Ptot = 1.0e7
        Ti = Ptot / (kB * 1.0e26)
    Lz, z_mm, nz = ck.get("Lz", 0.03), ck.get("z_mm", 0.0), int(ck.get("nz", 22))
    kUV = 2 * pi / lambda_UV
    kg = 2 * kUV * math.sin(theta)
    f_O2 = 1.0 - f_CO2
    R = NA * kB / (f_O2 * M_O2 + f_CO2 * M_CO2)
    gamma = (f_O2 * Cp_O2 + f_CO2 * Cp_CO2) / (f_O2 * Cv_O2 + f_CO2 * Cv_CO2)
    n_O2 = f_O2 * Ptot / (kB * Ti)
    n_O3 = f_O3 * Ptot / (kB * Ti)
    n_CO2 = f_CO2 * Ptot / (kB * Ti)
    L_ref = 1.0 / kg
    rho_ref = Ptot / (R * Ti)
    v_ref = math.sqrt(gamma * R * Ti)
    P_ref = rho_ref * v_ref * v_ref
    nz = int(math.ceil(z_mm * 0.001 * nz / Lz)) + 1
    xmin, xmax = -0.005 / L_ref, 0.005 / L_ref
    ymin, ymax = 0.000 / L_ref, (0.002 if sup else 0.010) / L_ref
    if sup:
        rho0 = (f_O2 * 1.0e26 * M_O2 / NA + f_CO2 * 1.0e26 * M_CO2 / NA) / rho_ref
    else:
        rho0 = (n_O2 * M_O2 / NA + n_CO2 * M_CO2 / NA) / rho_ref
    p0 = 1.0e7 / P_ref
    u0, v0 = 0.0, 530 / v_ref
    dx = (xmax - xmin) / (NI - 1.0)
    dy = (ymax - ymin) / (NJ - 1.0)
    x = np.array([xmin + i * dx for i in range(NI)])
    y = np.array([j * dy for j in range(NJ)])
    nvars = 4 + 7 * nz
    U = np.zeros((NJ * NI, nvars))
    U[:, 0] = rho0
    U[:, 1] = rho0 * u0
    U[:, 2] = rho0 * v0
    U[:, 3] = p0 / (gamma - 1.0) + 0.5 * rho0 * (u0 * u0 + v0 * v0)
    for iz in range(nz):
        U[:, 4 + 7 * iz + 0] = 1.0
        U[:, 4 + 7 * iz + 1] = n_O3 / n_O2
        U[:, 4 + 7 * iz + 6] = n_CO2 / n_O2
    _write_binary_initial(d, x, y, U)


def example_wall_initial(name: str, d: str, NI: int, NJ: int) -> None:
    """initial.inp (binary) of Examples/2D_NavierStokes/LidDrivenCavity (unit
    square, rho = 1, p = 1/gamma, the rotating start u = Minf (y - 1/2),
    v = -Minf (x - 1/2) with Minf = 0.1) or FlatPlateLaminar/UniformGrid
    ([-0.25, 1] x [0, 0.25], rho = 1, p = 1/1.4, u = 0.3), as their aux/init.c
    write it."""
    gamma = 1.4
    if name == "LidDrivenCavity":
        dx, dy = 1.0 / (NI - 1.0), 1.0 / (NJ - 1.0)
        x = np.array([i * dx for i in range(NI)])
        y = np.array([j * dy for j in range(NJ)])
        X, Y = np.meshgrid(x, y, indexing="xy")
        Minf = 0.1
        u = Minf * (Y - 0.5)
        v = -(Minf * (X - 0.5))
        P = 1.0 / gamma
    elif name == "FlatPlateLaminar":
        xmin, xmax, ymin, ymax = -0.25, 1.0, 0.0, 0.25
        dx, dy = (xmax - xmin) / (NI - 1), (ymax - ymin) / (NJ - 1)
        x = np.array([xmin + i * dx for i in range(NI)])
        y = np.array([ymin + j * dy for j in range(NJ)])
        u = np.full((NJ, NI), 0.3)
        v = np.zeros((NJ, NI))
        P = 1.0 / 1.4
    else:
        raise ValueError(name)
    rho = np.ones((NJ, NI))
    U = np.stack([rho, rho * u, rho * v, P / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)], axis=-1)
    _write_binary_initial(d, x, y, U.reshape(NJ * NI, 4))


def _write_ascii_rows(path: str, rows, fmt: str) -> None:
    """Rows of numbers as the examples' aux/init programs print them: each
    value followed by a space, one row per line."""
    with open(path, "w") as f:
        for r in rows:
            f.write("".join(fmt % v for v in np.asarray(r).reshape(-1)) + "\n")


def _photochem_scales(d: str, Ptot: float = 101325.0, Ti: float = 288.0):
    """The reference scales the photochemistry examples' aux/init programs
    derive (O2/CO2 gas, k_g = 2 k_UV sin(theta)), with lambda_UV, theta, the
    fractions, Ptot, Ti, Lz, z_mm and nz taken from chemistry.inp in d."""
    pi = 4.0 * math.atan(1.0)
    NA, kB = 6.02214076e23, 1.380649e-23
    M_O2, M_CO2 = 0.032, 0.044
    ck = _read_chem_keys(os.path.join(d, "chemistry.inp"))
    Ptot, Ti = ck.get("Ptot", Ptot), ck.get("Ti", Ti)
    f_CO2, f_O3 = ck.get("f_CO2", 0.0), ck.get("f_O3", 0.05)
    Lz, z_mm, nz = ck.get("Lz", 0.03), ck.get("z_mm", 0.0), int(ck.get("nz", 22))
    f_O2 = 1.0 - f_CO2
    R = NA * kB / (f_O2 * M_O2 + f_CO2 * M_CO2)
    gamma = (f_O2 * 918.45 + f_CO2 * 842.86) / (f_O2 * 650.0 + f_CO2 * 654.5)
    n_O2 = f_O2 * Ptot / (kB * Ti)
    n_O3 = f_O3 * Ptot / (kB * Ti)
    n_CO2 = f_CO2 * Ptot / (kB * Ti)
    rho_ref = Ptot / (R * Ti)
    v_ref = math.sqrt(gamma * R * Ti)
    P_ref = rho_ref * v_ref * v_ref
    rho0 = (n_O2 * M_O2 / NA + n_CO2 * M_CO2 / NA) / rho_ref
    nz = int(math.ceil(z_mm * 0.001 * nz / Lz)) + 1
    return dict(pi=pi, gamma=gamma, rho0=rho0, p0=Ptot / P_ref, nz=nz, O3=n_O3 / n_O2, CO2=n_CO2 / n_O2)


def example_beam_initial(name: str, d: str, NI: int, NJ: int = 1) -> None:
    """Examples/1D_Euler/BeamInducedDensityMod (initial.inp in ASCII "%lf")
    and Examples/2D_NavierStokes/BeamInducedDensityMod_1D_x (binary, with the
    beam map imap.inp of its aux/imap.C: 1 + cos(x - xmin)), as their aux
    programs write them: gas at rest at Ptot on five acoustic periods
    [-5 pi, 5 pi], species O2 = 1, O3 = n_O3/n_O2, CO2 = n_CO2/n_O2 in each of
    the z_i + 1 layers."""
    c = _photochem_scales(d)
    pi, g, rho0, p0, nz = c["pi"], c["gamma"], c["rho0"], c["p0"], c["nz"]
    l_aw = 2 * pi / 1
    N_period_x = 5.0
    xmin, xmax = -N_period_x / 2 * l_aw, N_period_x / 2 * l_aw
    if name == "BeamInducedDensityMod":
        dx = (xmax - xmin) / (NI - 1)
        x = np.array([xmin + i * dx for i in range(NI)])
        rows = [x, np.full(NI, rho0), np.full(NI, rho0 * 0.0), np.full(NI, p0 / (g - 1.0) + 0.5 * rho0 * 0.0 * 0.0)]
        for _ in range(nz):
            rows += [np.full(NI, 1.0), np.full(NI, c["O3"])] + [np.zeros(NI)] * 4 + [np.full(NI, c["CO2"])]
        _write_ascii_rows(os.path.join(d, "initial.inp"), rows, "%f ")
        return
    if name != "BeamInducedDensityMod_1D_x":
        raise ValueError(name)
    dx = (xmax - xmin) / (NI - 1.0)
    x = np.array([xmin + i * dx for i in range(NI)])
    y = np.array([j * dx for j in range(NJ)])
    nvars = 4 + 7 * nz
    U = np.zeros((NJ * NI, nvars))
    U[:, 0] = rho0
    U[:, 3] = p0 / (g - 1.0) + 0.5 * rho0 * (0.0 * 0.0 + 0.0 * 0.0)
    for iz in range(nz):
        U[:, 4 + 7 * iz] = 1.0
        U[:, 4 + 7 * iz + 1] = c["O3"]
        U[:, 4 + 7 * iz + 6] = c["CO2"]
    _write_binary_initial(d, x, y, U)
    ival = np.tile(np.array([1.0 + 1.0 * math.cos((xi - xmin) * (1 - 0.0 * (xi - xmin))) for xi in x]), NJ)
    with open(os.path.join(d, "imap.inp"), "wb") as f:
        x.tofile(f)
        y.tofile(f)
        ival.tofile(f)


def example_pressure_initial(name: str, d: str, NI: int, NJ: int = 1) -> None:
    """Examples/1D_Euler/PressurePerturbation (ASCII "%lf") and
    Examples/2D_NavierStokes/PressurePerturbation_1D_x (binary), as their
    aux/init programs write them: rho = 1, p = 1 + exp(-x^2 / (2 sigma^2))
    with sigma = 5 / (2 sqrt(2 ln 2)), at rest, on ten periods [-10 pi, 10 pi]."""
    pi, gamma = 4.0 * math.atan(1.0), 1.4
    l_aw = 2 * pi / 1
    N_period_x = 10.0
    xmin, xmax = -N_period_x / 2 * l_aw, N_period_x / 2 * l_aw
    sigma = 5.0 / (2.0 * math.sqrt(2.0 * math.log(2.0)))
    if name == "PressurePerturbation":
        dx = (xmax - xmin) / (NI - 1)
    elif name == "PressurePerturbation_1D_x":
        dx = (xmax - xmin) / (NI - 1.0)
    else:
        raise ValueError(name)
    x = np.array([xmin + i * dx for i in range(NI)])
    pres = np.array([1.0 + 1.0 * 1.0 * math.exp(-0.5 * (xi / sigma) * (xi / sigma)) for xi in x])
    if name == "PressurePerturbation":
        e = pres / (gamma - 1.0) + 0.5 * 1.0 * 0.0 * 0.0
        _write_ascii_rows(os.path.join(d, "initial.inp"), [x, np.ones(NI), np.zeros(NI), e], "%f ")
        return
    y = np.array([j * dx for j in range(NJ)])
    U = np.zeros((NJ, NI, 4))
    U[:, :, 0] = 1.0
    U[:, :, 3] = pres / (gamma - 1.0) + 0.5 * 1.0 * (0.0 * 0.0 + 0.0 * 0.0)
    _write_binary_initial(d, x, y, U.reshape(-1, 4))


def _solver_keys(d: str) -> dict:
    toks = open(os.path.join(d, "solver.inp")).read().split()
    return {toks[k]: toks[k + 1] for k in range(len(toks) - 1)}


def example_sine_initial(name: str, d: str, NI: int, NJ: int = 1) -> None:
    """Examples/1D_Euler/DensitySineWave (`DensitySineWave_1D`) and
    Examples/2D_NavierStokes/DensitySineWave (`DensitySineWave_2D`): exact.inp at
    t = n_iter dt and initial.inp, both ASCII "%1.16E", as their aux/init
    programs write them (rho = 1 + 0.1 sin(2 pi (x - t)) [cos(2 pi (y - t))],
    u = v = 1, p = 1/gamma)."""
    pi, gamma = 4.0 * math.atan(1.0), 1.4
    sk = _solver_keys(d)
    tf = float(int(sk["n_iter"])) * float(sk["dt"])
    if name == "DensitySineWave_1D":
        dx = 1.0 / NI
        x = np.array([i * dx for i in range(NI)])
        for fn, t in (("exact.inp", tf), ("initial.inp", None)):
            rho = np.array([1.0 + (0.1 * math.sin(2 * pi * (xi - 1.0 * t)) if t is not None
                                   else 0.1 * math.sin(2 * pi * xi)) for xi in x])
            e = 1.0 / gamma / (gamma - 1.0) + 0.5 * rho * 1.0 * 1.0
            _write_ascii_rows(os.path.join(d, fn), [x, rho, rho * 1.0, e], "%1.16E ")
        return
    if name != "DensitySineWave_2D":
        raise ValueError(name)
    dx, dy = 1.0 / NI, 1.0 / NJ
    x = np.array([i * dx for i in range(NI)])
    y = np.array([j * dy for j in range(NJ)])
    for fn, t in (("exact.inp", tf), ("initial.inp", 0.0)):
        rho = np.array([[1.0 + 0.1 * math.sin(2 * pi * (x[i] - 1.0 * t)) * math.cos(2 * pi * (y[j] - 1.0 * t))
                         for i in range(NI)] for j in range(NJ)])
        P = 1.0 / gamma
        e = P / (gamma - 1.0) + 0.5 * rho * (1.0 * 1.0 + 1.0 * 1.0)
        _write_ascii_rows(os.path.join(d, fn), [x, y, rho, rho * 1.0, rho * 1.0, e], "%1.16E ")


def example_vortex_initial(name: str, d: str, NI: int, NJ: int) -> None:
    """Examples/2D_NavierStokes/InviscidVortexConvection: initial.inp and
    exact.inp (the vortex advected by u_inf t, t = n_iter dt folded into one
    period of 20) in binary, as its aux/init.c writes them."""
    if name != "InviscidVortexConvection":
        raise ValueError(name)
    pi, G = 4.0 * math.atan(1.0), 1.4
    sk = _solver_keys(d)
    tf = float(int(sk["n_iter"])) * float(sk["dt"])
    while tf > 20:
        tf -= 20
    dx, dy = 10.0 / NI, 10.0 / NJ
    x = np.array([i * dx for i in range(NI)])
    y = np.array([j * dy for j in range(NJ)])
    b = 0.5

    def state(x0, y0):
        U = np.zeros((NJ, NI, 4))
        for i in range(NI):
            for j in range(NJ):
                rx, ry = x[i] - x0, y[j] - y0
                if rx < -5:
                    rx += 10
                elif rx > 5:
                    rx -= 10
                rsq = rx * rx + ry * ry
                rho = math.exp((1.0 / (G - 1.0)) * math.log(1.0 - ((G - 1.0) * b * b) / (8.0 * G * pi * pi)
                                                               * math.exp(1.0 - rsq)))
                P = math.exp(G * math.log(rho))
                du = -b / (2.0 * pi) * math.exp(0.5 * (1.0 - rsq)) * ry
                dv = b / (2.0 * pi) * math.exp(0.5 * (1.0 - rsq)) * rx
                u, v = 0.5 + du, 0.0 + dv
                U[j, i] = (rho, rho * u, rho * v, P / (G - 1.0) + 0.5 * rho * (u * u + v * v))
        return U.reshape(-1, 4)

    _write_binary_initial(d, x, y, state(5.0, 5.0))
    x0 = 5.0 + tf * 0.5
    if x0 > 10:
        x0 -= 10
    with open(os.path.join(d, "exact.inp"), "wb") as f:
        x.tofile(f)
        y.tofile(f)
        state(x0, 5.0).tofile(f)


def shipped_example_initial(name: str, d: str, NI: int, NJ: int) -> None:
    """initial.inp of a shipped 2D example, written as its aux/init program
    writes it (the deck files, chemistry.inp included, already in `d`)."""
    if name in ("Riemann2DCase4", "RadialExpansionWave"):
        example_2d_initial(name, d, NI, NJ)
    elif name in ("Uniform_Supersonic_Jet", "Uniform_Subsonic_Jet"):
        example_jet_initial(name, d, NI, NJ)
    elif name.startswith("BeamInducedDensityMod"):
        example_beam_initial(name, d, NI, NJ)
    elif name.startswith("PressurePerturbation"):
        example_pressure_initial(name, d, NI, NJ)
    elif name.startswith("DensitySineWave"):
        example_sine_initial(name, d, NI, NJ)
    elif name == "InviscidVortexConvection":
        example_vortex_initial(name, d, NI, NJ)
    else:
        example_wall_initial(name, d, NI, NJ)


def copy_example_deck(src: str, dst: str, files, solver_set: dict | None = None,
                      solver_drop=()) -> None:
    """Copy an example's input files, changing solver.inp keywords (e.g. a
    shorter n_iter) and dropping the lines of `solver_drop`."""
    os.makedirs(dst, exist_ok=True)
    for fn in files:
        text = open(os.path.join(src, fn)).read()
        if fn == "solver.inp" and (solver_set or solver_drop):
            out = []
            for line in text.splitlines():
                t = line.split()
                if t and t[0] in solver_drop:
                    continue
                if t and solver_set and t[0] in solver_set:
                    line = "  %s %s" % (t[0], solver_set[t[0]])
                out.append(line)
            text = "\n".join(out) + "\n"
        with open(os.path.join(dst, fn), "w") as f:
            f.write(text)


def write_binary_op(path: str, ndims: int, dims, x_list, u) -> None:
    """A solution file in WriteBinary's layout (src/IOFunctions/WriteBinary.c:37-93):
    int32 ndims, int32 nvars, int32 dim[ndims], f64 x[sum dim], f64 u[prod dim * nvars]."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    with open(path, "wb") as f:
        np.array([ndims, u.shape[1]] + list(dims), dtype=np.int32).tofile(f)
        np.concatenate(x_list).astype(np.float64).tofile(f)
        u.tofile(f)


def read_binary_op(path: str):
    """Read a WriteBinary file (src/IOFunctions/WriteBinary.c:37-93):
    int32 ndims, int32 nvars, int32 dim[ndims], f64 x[sum dim], f64 u[prod dim * nvars]."""
    with open(path, "rb") as f:
        ndims, nvars = np.fromfile(f, dtype=np.int32, count=2)
        dims = np.fromfile(f, dtype=np.int32, count=int(ndims))
        x = np.fromfile(f, dtype=np.float64, count=int(dims.sum()))
        u = np.fromfile(f, dtype=np.float64, count=int(np.prod(dims)) * int(nvars))
    return int(ndims), int(nvars), [int(d) for d in dims], x, u.reshape(-1, int(nvars))


def read_solution_file(path: str, nv: int) -> np.ndarray:
    """A solution file the reference writes, as (npoints, nvars): binary
    (WriteBinary.c:37-93), or text / Tecplot (WriteText.c, WriteTecplot2D.c:
    one row per point, x fastest, the nvars values last; header lines are
    skipped)."""
    if path.endswith(".bin"):
        return read_binary_op(path)[4].reshape(-1, nv)
    rows = []
    for line in open(path):
        try:
            vals = [float(v) for v in line.split()]
        except ValueError:
            continue
        if len(vals) >= nv:
            rows.append(vals[-nv:])
    return np.array(rows)


def isentropic_vortex(N: int = 60, n_iter: int = 400, dt: float = 0.05, scheme: str = "crweno5",
                      rk: str = "ssprk3", mapped: int = 1, gamma: float = 1.4) -> Deck:
    """Isentropic vortex convected by u_inf = 0.5 across the periodic box
    [0,10)^2 (strength b = u_inf, centre (5,5)); after t = 20 it returns to its
    initial position, so the exact solution equals the initial one.  Same
    problem as Examples/2D_NavierStokes/InviscidVortexConvection (whose shipped
    errors.dat was produced with n_iter 400, dt 0.05); inputs written in binary
    so the grid is exact."""
    L = 10.0
    dx = L / N
    x = np.arange(N) * dx
    y = np.arange(N) * dx
    X, Y = np.meshgrid(x, y, indexing="xy")
    rx, ry = X - 5.0, Y - 5.0
    rx = np.where(rx < -5, rx + 10, np.where(rx > 5, rx - 10, rx))
    rsq = rx * rx + ry * ry
    b, pi = 0.5, 4.0 * math.atan(1.0)
    rho = np.exp((1.0 / (gamma - 1.0)) * np.log(1.0 - ((gamma - 1.0) * b * b) / (8.0 * gamma * pi * pi) * np.exp(1.0 - rsq)))
    P = np.exp(gamma * np.log(rho))
    du = -b / (2.0 * pi) * np.exp(0.5 * (1.0 - rsq)) * ry
    dv = b / (2.0 * pi) * np.exp(0.5 * (1.0 - rsq)) * rx
    uu, vv = 0.5 + du, 0.0 + dv
    u = np.stack([rho, rho * uu, rho * vv, P / (gamma - 1.0) + 0.5 * rho * (uu * uu + vv * vv)], axis=-1)
    u = u.reshape(-1, 4)
    solver = dict(ndims=2, nvars=4, size=(N, N), ghost=3, n_iter=n_iter, time_scheme="rk",
                  time_scheme_type=rk, hyp_space_scheme=scheme, hyp_interp_type="components",
                  par_space_scheme="4", dt=dt, screen_op_iter=n_iter, file_op_iter=n_iter * 10,
                  op_file_format="none", ip_file_type="binary", op_overwrite="yes", model="navierstokes2d")
    boundary = [("periodic", 0, 1, 0, 0, 0, 10.0), ("periodic", 0, -1, 0, 0, 0, 10.0),
                ("periodic", 1, 1, 0, 10.0, 0, 0), ("periodic", 1, -1, 0, 10.0, 0, 0)]
    physics = dict(gamma=gamma, upwinding="roe")
    weno = dict(mapped=mapped, borges=0, yc=0, no_limiting=0, epsilon=1e-6)
    return Deck(solver, boundary, physics, [x, y], u, weno)


def riemann_2d(N: int = 201, n_iter: int = 250, dt: float = 0.001, interp: str = "characteristic",
               upwinding: str = "rf-char", rk: str = "ssprk3", mapped: int = 1, scheme: str = "weno5",
               gamma: float = 1.4) -> Deck:
    """Two-dimensional Riemann problem on [-0.5, 0.5]^2 (grid points x_i =
    -0.5 + i/(N-1)), inviscid, mirror ("extrapolate") BCs on all faces: the
    four quadrant states of Examples/2D_NavierStokes/Riemann2DCase4 (rho, u, v,
    p) = (1.1, 0.8939, 0.8939, 1.1) for x < 0, y < 0; (0.5065, 0, 0.8939,
    0.35) for x >= 0, y < 0; (0.5065, 0.8939, 0, 0.35) for x < 0, y >= 0;
    (1.1, 0, 0, 1.1) otherwise."""
    h = 1.0 / (N - 1)
    x = -0.5 + np.arange(N) * h
    y = -0.5 + np.arange(N) * h
    X, Y = np.meshgrid(x, y, indexing="xy")
    rho = np.where(X < 0, np.where(Y < 0, 1.1, 0.5065), np.where(Y < 0, 0.5065, 1.1))
    P = np.where(X < 0, np.where(Y < 0, 1.1, 0.35), np.where(Y < 0, 0.35, 1.1))
    uu = np.where((X < 0) & (Y >= 0), 0.8939, np.where((X < 0) & (Y < 0), 0.8939, 0.0))
    vv = np.where((X >= 0) & (Y < 0), 0.8939, np.where((X < 0) & (Y < 0), 0.8939, 0.0))
    u = np.stack([rho, rho * uu, rho * vv, P / (gamma - 1.0) + 0.5 * rho * (uu * uu + vv * vv)], axis=-1)
    u = u.reshape(-1, 4)
    solver = dict(ndims=2, nvars=4, size=(N, N), ghost=3, n_iter=n_iter, time_scheme="rk",
                  time_scheme_type=rk, hyp_space_scheme=scheme, hyp_interp_type=interp,
                  dt=dt, screen_op_iter=n_iter, file_op_iter=n_iter * 10,
                  op_file_format="none", ip_file_type="binary", op_overwrite="yes", model="navierstokes2d")
    boundary = [("extrapolate", 0, 1, 0, 0, -0.5, 0.5), ("extrapolate", 0, -1, 0, 0, -0.5, 0.5),
                ("extrapolate", 1, 1, -0.5, 0.5, 0, 0), ("extrapolate", 1, -1, -0.5, 0.5, 0, 0)]
    physics = dict(gamma=gamma, upwinding=upwinding)
    weno = dict(mapped=mapped, borges=0, yc=0, no_limiting=0, epsilon=1e-6)
    return Deck(solver, boundary, physics, [x, y], u, weno)


def density_sine_1d(N: int = 128, n_iter: int = 50, dt: float = 0.002, scheme: str = "crweno5",
                    interp: str = "components", rk: str = "44", gamma: float = 1.4, amp: float = 0.1) -> Deck:
    """Smooth 1D density wave rho = 1 + amp sin(2 pi x), u = 1, p = 1/gamma on a
    periodic [0,1) grid."""
    x = np.arange(N, dtype=np.float64) / N
    rho = 1.0 + amp * np.sin(2 * math.pi * x)
    u = np.zeros((N, 3))
    u[:, 0] = rho
    u[:, 1] = rho * 1.0
    u[:, 2] = (1.0 / gamma) / (gamma - 1.0) + 0.5 * rho
    solver = dict(ndims=1, nvars=3, size=N, ghost=3, n_iter=n_iter, time_scheme="rk",
                  time_scheme_type=rk, hyp_space_scheme=scheme, hyp_interp_type=interp,
                  dt=dt, screen_op_iter=n_iter, file_op_iter=n_iter * 10, op_file_format="none",
                  ip_file_type="binary", op_overwrite="yes", model="euler1d")
    boundary = [("periodic", 0, 1, 0, 0), ("periodic", 0, -1, 0, 0)]
    physics = dict(gamma=gamma, upwinding="roe")
    weno = dict(mapped=0, borges=0, yc=0, no_limiting=0, epsilon=1e-6)
    return Deck(solver, boundary, physics, [x], u, weno)


def error_norms(u: np.ndarray, uex: np.ndarray):
    """L1/L2/Linf of u - uex relative to the norms of uex, aggregated over all
    variables as src/HyParFunctions/CalculateError.c:67-116 does."""
    npts = u.shape[0]
    d = u - uex
    n = (np.abs(uex).sum() / npts, math.sqrt((uex * uex).sum() / npts), np.abs(uex).max())
    e = [np.abs(d).sum() / npts, math.sqrt((d * d).sum() / npts), np.abs(d).max()]
    if min(n) > 1e-15:
        e = [e[k] / n[k] for k in range(3)]
    return tuple(e)


def rel_l2_per_var(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-variable relative L2 difference ||a-b|| / ||b|| (absolute where ||b|| = 0)."""
    num = np.sqrt(((a - b) ** 2).sum(axis=0))
    den = np.sqrt((b ** 2).sum(axis=0))
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)


# ---------------------------------------------------------------------------
# Ensembles and request streams
# ---------------------------------------------------------------------------
CHEM_RATE_PER_RE = 0.145   # fastest chemical rate of the O2/O3 set, per unit Re (normalised time)


def stable_dt(Re: float, dt_max: float = 0.01, safety: float = 1.5) -> float:
    """Largest fixed step for classical RK4 on the photochemistry source: the
    O(1D) quenching rate grows with Ptot, i.e. with Re (rate about
    CHEM_RATE_PER_RE*Re), and RK4 is stable for dt*rate < 2.78; `safety`
    keeps dt*rate at 1.5.  dt_max is the acoustic limit of the decks."""
    return min(dt_max, safety / (CHEM_RATE_PER_RE * Re))


@dataclass
class Member:
    name: str
    deck: Deck
    cells: int          # interior points
    steps: int


def member_tag(k: int, nsims: int) -> str:
    """File-name tag of member k of an ensemble of nsims ("" for one
    simulation): "_" and k zero-padded to floor(log10 nsims) + 1 digits
    (GetStringFromInteger as InitialSolution.c:36-41 calls it)."""
    if nsims <= 1:
        return ""
    return "_" + str(k).zfill(int(math.log10(nsims)) + 1)


def write_ensemble(d: str, members: list, shared_boundary: bool = True) -> str:
    """The reference's ensemble directory (EnsembleSimulationsDefine.cpp:31-119)
    for decks that differ only in grid size, initial solution and boundary
    zones: simulation.inp (nsims), one solver.inp whose `size` lists every
    member's sizes in turn (ReadInputs.c:272-290), the shared physics, weno
    and chemistry files, initial_<i>.inp, and boundary_<i>.inp for each member
    whose zones differ from member 0's (member 0's zones go to the shared
    boundary.inp when `shared_boundary`, read as the fallback by
    InitializeBoundaries.c:51-75)."""
    if not members:
        raise ValueError("an ensemble needs at least one member")
    base = members[0]
    n = len(members)
    for m in members[1:]:
        a = {k: v for k, v in m.solver.items() if k not in ("size", "size_exact")}
        b = {k: v for k, v in base.solver.items() if k not in ("size", "size_exact")}
        if a != b or m.physics != base.physics or m.weno != base.weno or m.chemistry != base.chemistry:
            raise ValueError("ensemble members share solver.inp (but for size), physics, weno and chemistry")
        if m.imap is not None:
            raise ValueError("the reference reads one imap.inp for every member; build members without it")
    os.makedirs(d, exist_ok=True)
    _kw_file(os.path.join(d, "simulation.inp"), {"nsims": n})
    solver = dict(base.solver)
    solver["size"] = [v for m in members for v in m.dims]
    _kw_file(os.path.join(d, "solver.inp"), solver)
    _kw_file(os.path.join(d, "physics.inp"), base.physics)
    if base.weno:
        _kw_file(os.path.join(d, "weno.inp"), base.weno)
    if base.chemistry is not None:
        _kw_file(os.path.join(d, "chemistry.inp"), base.chemistry)

    def zones(path, bl):
        with open(path, "w") as f:
            f.write(f"{len(bl)}\n")
            for b in bl:
                f.write(" ".join(str(t) for t in b) + "\n")

    if shared_boundary:
        zones(os.path.join(d, "boundary.inp"), base.boundary)
    for k, m in enumerate(members):
        t = member_tag(k, n)
        if not shared_boundary or m.boundary != base.boundary:
            zones(os.path.join(d, f"boundary{t}.inp"), m.boundary)
        with open(os.path.join(d, f"initial{t}.inp"), "wb") as f:
            for xd in m.x:
                np.ascontiguousarray(xd, dtype=np.float64).tofile(f)
            np.ascontiguousarray(m.u, dtype=np.float64).tofile(f)
    return d


def ensemble(n: int = 256, seed: int = 20261202, n_iter: int = 200, side=(128, 512),
             Re_range=(1e2, 1e4), amp_range=(0.5, 2.0)) -> list:
    """Independent 2D NS + photochemistry cases (case k drawn from SplitMix64
    stream seed + k): sides log-uniform in `side` (x and y independently),
    beam amplitude uniform in `amp_range`, Re log-uniform in `Re_range` (set
    through Ptot), WENO5 / CRWENO5 with equal probability; fixed dx = 10 pi /
    1024; dt = stable_dt(Re); seeded density perturbation as the single-case
    deck."""
    dx = 10.0 * math.pi / 1024
    out = []
    lo, hi = math.log(side[0]), math.log(side[1])
    for k in range(n):
        r = splitmix64_fast(seed + k, 5)
        nx = int(round(math.exp(lo + (hi - lo) * r[0])))
        ny = int(round(math.exp(lo + (hi - lo) * r[1])))
        amp = amp_range[0] + (amp_range[1] - amp_range[0]) * r[2]
        Re = math.exp(math.log(Re_range[0]) + (math.log(Re_range[1]) - math.log(Re_range[0])) * r[3])
        scheme = "crweno5" if r[4] < 0.5 else "weno5"
        dk = ns2d_photochem(nx=nx, ny=ny, n_iter=n_iter, dt=stable_dt(Re), seed=seed + k, scheme=scheme,
                            Re=Re, amplitude=amp, dx=dx, L=nx * dx)
        out.append(Member(f"c{k:03d}_{nx}x{ny}_{scheme}_Re{Re:.0f}", dk, nx * ny, n_iter))
    return out


@dataclass
class Request:
    arrival: float      # seconds after the start of the stream
    member: Member


def request_stream(n: int, rate: float, seed: int = 20261203, frac_1d: float = 0.3,
                   n1d=(512, 8192), side2d=(128, 1024), steps_1d: int | None = None,
                   steps_2d: int = 200) -> list:
    """Poisson arrivals (rate per second) of solve requests to a fixed final
    time: a fraction frac_1d are Sod shock tubes (N log-uniform in n1d,
    t_final 0.2, the single-case 1D deck), the rest 2D NS + photochemistry
    (side log-uniform in side2d, the single-case 2D deck with dx = 10 pi /
    1024, Re 1000, dt 0.01, t_final 2.0).  steps_1d / steps_2d shorten the
    requests to that many steps of the same dt (the same stream, stopped
    early)."""
    dx = 10.0 * math.pi / 1024
    out = []
    t = 0.0
    for k in range(n):
        r = splitmix64_fast(seed + 7919 * k, 3)
        t += -math.log(1.0 - r[0]) / rate
        if r[1] < frac_1d:
            N = int(round(math.exp(math.log(n1d[0]) + (math.log(n1d[1]) - math.log(n1d[0])) * r[2])))
            full = int(math.ceil(0.88 * (N - 1)))
            ns = full if steps_1d is None else min(steps_1d, full)
            dk = sod_1d(N=N, n_iter=ns, t_final=0.2 if ns == full else 0.2 * ns / full)
            m = Member(f"r{k:03d}_sod{N}", dk, N, ns)
        else:
            a = math.log(side2d[0])
            side = int(round(math.exp(a + (math.log(side2d[1]) - a) * r[2])))
            dk = ns2d_photochem(nx=side, n_iter=steps_2d, dt=0.01, seed=seed + k, dx=dx, L=side * dx)
            m = Member(f"r{k:03d}_ns{side}", dk, side * side, steps_2d)
        out.append(Request(t, m))
    return out
