"""Host-side problem generators (Init.h) and the benchmarked workload ICs.

These run once per run on the host and return a primitive array in the
reference's host layout (Ntz, Nty, Ntx, 5); ghost cells are left at zero
(the device builds ghosts from the boundary maps).  The reference's own
problems keep its expressions (Init.h:18-209), so Sod and blast are
bit-identical to the reference's initial conditions.

Random perturbations (H84, C91, the workload ICs) come from a counter-based
generator keyed by (seed, cell) -- splitmix64 (Steele, Lea, Flood 2014) --
so they do not depend on the thread mapping, unlike Kokkos'
Random_XorShift64_Pool in the reference (Init.h:114-116).
"""
from __future__ import annotations

import numpy as np

BC_REFLECTING, BC_PERIODIC = 1, 2

PROBLEMS = {"sod_x": 0, "sod_y": 1, "sod_z": 2, "blast": 3, "rayleigh-taylor": 4,
            "diffusion": 5, "H84": 6, "C91": 7}

_M64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_M64)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_M64)
    return x ^ (x >> np.uint64(31))


def cell_uniform(seed: int, cells: np.ndarray, lo: float, hi: float, stream: int = 0) -> np.ndarray:
    """U[lo, hi) per cell id, keyed by (seed, stream, cell)."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed & 0xFFFFFF) << np.uint64(40)) ^ np.uint64(stream << 36) ^ \
            cells.astype(np.uint64)
        r = splitmix64(key)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def _grid(p):
    """Cell-centre coordinates of the domain (getPos, SimInfo.h:162-167) and cell ids."""
    i = np.arange(p.ibeg, p.iend)
    j = np.arange(p.jbeg, p.jend)
    k = np.arange(p.kbeg, p.kend)
    x = p.xmin + ((i - p.ibeg) + 0.5) * p.dx
    y = p.ymin + ((j - p.jbeg) + 0.5) * p.dy
    z = p.zmin + ((k - p.kbeg) + 0.5) * p.dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    cells = (K.astype(np.int64) * p.Nty + J) * p.Ntx + I
    return X, Y, Z, cells


def _empty(p):
    return np.zeros((p.Ntz, p.Nty, p.Ntx, 5))


def _dom(Q, p):
    return Q[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend]


def sod(p, axis: int = 0):
    """initSodX/Y/Z (Init.h:18-67)."""
    Q = _empty(p)
    c = _grid(p)[axis]
    left = c <= 0.5
    d = _dom(Q, p)
    d[..., 0] = np.where(left, 1.0, 0.125)
    d[..., 4] = np.where(left, 1.0, 0.1)
    return Q


def blast(p, radius=0.2, rho_in=1.0, p_in=10.0, rho_out=1.2, p_out=0.1):
    """initBlast (Init.h:72-99); the defaults are the reference's constants."""
    Q = _empty(p)
    X, Y, Z, _ = _grid(p)
    xr = 0.5 * (p.xmin + p.xmax) - X
    yr = 0.5 * (p.ymin + p.ymax) - Y
    zr = 0.5 * (p.zmin + p.zmax) - Z
    r = np.sqrt(xr * xr + yr * yr + zr * zr)
    inside = r < radius
    d = _dom(Q, p)
    d[..., 0] = np.where(inside, rho_in, rho_out)
    d[..., 4] = np.where(inside, p_in, p_out)
    return Q


def diffusion(p):
    """initDiffusion (Init.h:155-178)."""
    Q = _empty(p)
    X, Y, Z, _ = _grid(p)
    x0 = X - 0.5 * (p.xmin + p.xmax)
    y0 = Y - 0.5 * (p.ymin + p.ymax)
    z0 = Z - 0.5 * (p.zmin + p.zmax)
    r = np.sqrt(x0 * x0 + y0 * y0 + z0 * z0)
    d = _dom(Q, p)
    d[..., 0] = np.where(r < 0.2, 1.0, 0.1)
    d[..., 1:5] = 1.0
    return Q


def rayleigh_taylor(p):
    """initRayleighTaylor (Init.h:183-209)."""
    Q = _empty(p)
    X, Y, Z, _ = _grid(p)
    zmid = 0.5 * (p.zmin + p.zmax)
    d = _dom(Q, p)
    d[..., 0] = np.where(Z < zmid, 1.0, 2.0)
    d[..., 4] = 2.5 + 0.1 * p.g * Z
    band = (Z > -1.0 / 3.0) & (Z < 1.0 / 3.0)
    w = 0.01 * (1.0 + np.cos(4 * np.pi * X)) * (1.0 + np.cos(4.0 * np.pi * Y)) * \
        (1 + np.cos(3.0 * np.pi * Z)) / 4.0
    d[..., 3] = np.where(band, w, 0.0)
    return Q


def h84(p):
    """initH84 (Init.h:104-123) with the counter-based perturbation."""
    Q = _empty(p)
    X, Y, Z, cells = _grid(p)
    d = _dom(Q, p)
    d[..., 0] = np.power(Z, p.m1)
    d[..., 3] = p.h84_pert * cell_uniform(p.seed, cells, -0.5, 0.5)
    d[..., 4] = np.power(Z, p.m1 + 1.0)
    return Q


def c91(p):
    """initC91 (Init.h:128-150) with the counter-based perturbation."""
    Q = _empty(p)
    X, Y, Z, cells = _grid(p)
    T = 1.0 + p.theta1 * Z
    d = _dom(Q, p)
    d[..., 0] = np.power(T, p.m1)
    prs = np.power(T, p.m1 + 1.0)
    d[..., 4] = prs * (1.0 + p.c91_pert * cell_uniform(p.seed, cells, -0.5, 0.5))
    return Q


def polytrope_convection(p, m=1.0, amplitude=1.0e-3):
    """Stratified polytropic layer with velocity noise (the 256x256x128
    conduction/viscosity workload).  z is depth: T = 1 + theta*z with
    theta = g/(m+1) so that dp/dz = rho*g (source +rho*g, Update.h:149);
    rho = T^m, p = T^(m+1); u, v, w += amplitude*U(-1/2, 1/2), seeded."""
    Q = _empty(p)
    X, Y, Z, cells = _grid(p)
    theta = p.g / (m + 1.0)
    T = 1.0 + theta * Z
    d = _dom(Q, p)
    d[..., 0] = np.power(T, m)
    d[..., 4] = np.power(T, m + 1.0)
    for s, slot in enumerate((1, 2, 3)):
        d[..., slot] = amplitude * cell_uniform(p.seed, cells, -0.5, 0.5, stream=s + 1)
    return Q


def _box_uniform(seed: int, box: int, q: int) -> float:
    """U[0, 1) keyed by (seed, box, draw q) -- per-box parameters of an ensemble."""
    return float(cell_uniform(seed, np.array([box * 64 + q]), 0.0, 1.0, stream=9)[0])


def rt_box_params(base, box: int, seed: int = 12345, sides=(32, 48, 64, 96)):
    """Parameters of box `box` of a Rayleigh-Taylor ensemble: each side drawn
    from `sides`, gamma ~ U[1.3, 1.7], on [-1/2, 1/2]^3 (cells need not be
    cubic), periodic x and y, reflecting z, gravity g = -0.1 (the reference's
    sign convention: the w source is +rho*g, Update.h:149-151).  `base`
    supplies the numerics (reconstruction, solver, time stepping, CFL)."""
    n = [sides[min(int(_box_uniform(seed, box, a) * len(sides)), len(sides) - 1)] for a in range(3)]
    gamma = 1.3 + 0.4 * _box_uniform(seed, box, 3)
    return base.replace(Nx=n[0], Ny=n[1], Nz=n[2], xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5,
                        zmin=-0.5, zmax=0.5, gamma0=gamma, gravity=1, g=-0.1,
                        boundary_x=BC_PERIODIC, boundary_y=BC_PERIODIC,
                        boundary_z=BC_REFLECTING, seed=seed + box)


def rt_box(p, box: int, seed: int = 12345, amplitude: float = 1.0e-2, modes: int = 6):
    """Rayleigh-Taylor box: heavy fluid (rho=2) over light (rho=1) across
    z=0, in hydrostatic balance p = 2.5 + rho*g*z (continuous at z=0; the
    reference's own RT problem uses p = 2.5 + 0.1*g*z for both fluids,
    Init.h:192-200).  w carries the reference's single mode
    (Init.h:205-208) plus `modes` random Fourier modes of integer
    wavenumbers 1..4 in x and y with random phases, under the same z
    envelope, scaled so that |w| <= amplitude."""
    Q = _empty(p)
    X, Y, Z, _ = _grid(p)
    d = _dom(Q, p)
    rho = np.where(Z > 0.0, 2.0, 1.0)
    d[..., 0] = rho
    d[..., 4] = 2.5 + rho * p.g * Z
    single = (1.0 + np.cos(4.0 * np.pi * X)) * (1.0 + np.cos(4.0 * np.pi * Y)) / 4.0
    rnd = np.zeros_like(X)
    wsum = 0.0
    for m in range(modes):
        kx = 1 + int(4 * _box_uniform(seed, box, 8 + 4 * m))
        ky = 1 + int(4 * _box_uniform(seed, box, 9 + 4 * m))
        ph = 2.0 * np.pi * _box_uniform(seed, box, 10 + 4 * m)
        a = _box_uniform(seed, box, 11 + 4 * m)
        rnd += a * np.cos(2.0 * np.pi * (kx * X + ky * Y) + ph)
        wsum += a
    env = np.where(np.abs(Z) < 1.0 / 3.0, (1.0 + np.cos(3.0 * np.pi * Z)) / 2.0, 0.0)
    d[..., 3] = amplitude * env * 0.5 * (single + rnd / max(wsum, 1e-300))
    return Q


def rt_ensemble(base, nbox: int, seed: int = 12345, sides=(32, 48, 64, 96)):
    """[(Params, Q)] for an ensemble of `nbox` Rayleigh-Taylor boxes."""
    out = []
    for b in range(nbox):
        p = rt_box_params(base, b, seed, sides)
        out.append((p, rt_box(p, b, seed)))
    return out


def spectral_field(rng, shape, slope):
    """Real Gaussian random field with power spectrum P(k) ~ k^slope on a
    periodic grid (k in units of the box), zero mean, unit rms."""
    white = rng.standard_normal(shape)
    f = np.fft.rfftn(white)
    kz = np.fft.fftfreq(shape[0]) * shape[0]
    ky = np.fft.fftfreq(shape[1]) * shape[1]
    kx = np.fft.rfftfreq(shape[2]) * shape[2]
    K = np.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)
    amp = np.zeros_like(K)
    amp[K > 0] = K[K > 0] ** (slope / 2.0)
    g = np.fft.irfftn(f * amp, s=shape, axes=(0, 1, 2))
    g -= g.mean()
    return g / np.sqrt(np.mean(g * g))


def _solenoidal_velocity(rng, shape, slope):
    """Divergence-free random velocity (u, v, w) on a periodic grid with
    spectrum ~ k^slope, projected in Fourier space, unit total rms."""
    kz = np.fft.fftfreq(shape[0]) * shape[0]
    ky = np.fft.fftfreq(shape[1]) * shape[1]
    kx = np.fft.rfftfreq(shape[2]) * shape[2]
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    amp = np.zeros_like(K2)
    amp[K2 > 0] = K2[K2 > 0] ** (slope / 4.0)
    comps = [np.fft.rfftn(rng.standard_normal(shape)) * amp for _ in range(3)]  # x, y, z
    div = (KX * comps[0] + KY * comps[1] + KZ * comps[2])
    with np.errstate(invalid="ignore", divide="ignore"):
        scale = np.where(K2 > 0, div / K2, 0.0)
    comps = [comps[0] - KX * scale, comps[1] - KY * scale, comps[2] - KZ * scale]
    vel = [np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) for c in comps]
    rms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    return [v / rms for v in vel]


def turbulent_box(p, seed: int = 12345, rho_amp: float = 0.05, v_rms: float = 0.1,
                  slope: float = -2.0, p_amp: float = 0.0):
    """Periodic box for the adjoint workload: rho = 1 + rho_amp * GRF
    (P(k) ~ k^slope, unit rms), p = 1 (+ p_amp * an independent GRF),
    solenoidal random velocity of rms v_rms, all drawn from numpy's PCG64
    seeded with `seed`."""
    rng = np.random.default_rng(seed)
    shape = (p.Nz, p.Ny, p.Nx)
    g = spectral_field(rng, shape, slope)
    u, v, w = _solenoidal_velocity(rng, shape, slope)
    gp = spectral_field(rng, shape, slope) if p_amp else 0.0
    Q = _empty(p)
    d = _dom(Q, p)
    d[..., 0] = 1.0 + rho_amp * g
    d[..., 1] = v_rms * u
    d[..., 2] = v_rms * v
    d[..., 3] = v_rms * w
    d[..., 4] = 1.0 + p_amp * gp
    return Q


def initial_condition(cfg) -> np.ndarray:
    """Q for a RunConfig (read_inifile) -- InitFunctor::init (Init.h:253-278)."""
    p = cfg.params
    name = cfg.problem
    if name == "sod_x":
        return sod(p, 0)
    if name == "sod_y":
        return sod(p, 1)
    if name == "sod_z":
        return sod(p, 2)
    if name == "blast":
        return blast(p, cfg.blast_radius, cfg.blast_rho_in, cfg.blast_p_in, cfg.blast_rho_out,
                     cfg.blast_p_out)
    if name == "diffusion":
        return diffusion(p)
    if name == "rayleigh-taylor":
        return rayleigh_taylor(p)
    if name == "H84":
        return h84(p)
    if name == "C91":
        return c91(p)
    if name == "polytrope":
        return polytrope_convection(p)
    raise RuntimeError("Error unknown problem " + name)  # Init.h:246-247
