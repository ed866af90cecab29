"""Seeded initial conditions and unit mapping (host side, FP64, numpy).

The reference builds its initial conditions in ``src/initial.f90`` (Thomas-Fermi
profile ``setup_thomas_fermi`` :637-652, vortex imprint ``single_vortex`` :109-121, noise
``add_noise`` :741-757) with a clock-seeded RNG (``src/mpi_var.f90:64-90``).  The
workloads here use a counter-based generator instead, so every input is reproducible
from (seed, stream, index); the same generator is written in Fortran in the host
driver (``trapbec_amd/driver/tbec_ic.f90``).

Arrays are indexed [x, y, z] (C order, z fastest) = the reference's Psi(k, j, i).

Unit map (the stepper integrates dpsi/dt = (i/2) rhs with
rhs = lap psi + (c0 - gnl*|psi|^2 - pot) psi + om_r*i*(x d_y - y d_x) psi,
``src/finite_diff.f90:23-132``): the physical equation
i dpsi/dt = -1/2 lap psi + V psi + g|psi|^2 psi - Omega L_z psi with
V = 1/2 sum w_a^2 x_a^2 maps to pot = 2V (harm_w_a = w_a^2 L_a^2), gnl = 2g, om_r = -2 Omega
and c0 = 0.  The reference's own equation is the same with psi_ref = sqrt(2g) psi,
gnl = 1 and c0 = 1 (a constant energy shift).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

M32 = 0xFFFFFFFF


def hash32(x: np.ndarray) -> np.ndarray:
    """32-bit integer mixer (xorshift-multiply); every intermediate fits in a signed
    64-bit integer so the Fortran driver can restate it with integer(8)."""
    x = np.asarray(x, dtype=np.int64) & M32
    x = x ^ (x >> 16)
    x = (x * 0x21F0AAAD) & M32
    x = x ^ (x >> 15)
    x = (x * 0x735A2D97) & M32
    x = x ^ (x >> 15)
    return x


def uniform(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Uniform variates in (0, 1) from (seed, stream, idx); idx < 2**32."""
    key = hash32(np.int64(seed & M32) ^ hash32(np.int64(stream)))
    bits = hash32(key ^ hash32(np.asarray(idx, dtype=np.int64)))
    return (bits.astype(np.float64) + 0.5) / 4294967296.0


def gaussian_pair(seed: int, stream: int, idx: np.ndarray):
    """Box-Muller pair of standard normals from streams (2*stream, 2*stream+1)."""
    u1 = uniform(seed, 2 * stream, idx)
    u2 = uniform(seed, 2 * stream + 1, idx)
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)


def coords(n: int, L: float) -> np.ndarray:
    """Cell-centred coordinates x_i = (i - 1/2) h - L/2, i = 1..n (src/initial.f90:34-42)."""
    i = np.arange(1, n + 1, dtype=np.float64)
    return ((2 * i - 1) / (2.0 * n)) * L - L / 2.0


def mesh(shape, box):
    nx, ny, nz = shape
    Lx, Ly, Lz = box
    return np.meshgrid(coords(nx, Lx), coords(ny, Ly), coords(nz, Lz), indexing="ij")


def trap_potential(shape, box, omega):
    X, Y, Z = mesh(shape, box)
    wx, wy, wz = omega
    return 0.5 * (wx * wx * X * X + wy * wy * Y * Y + wz * wz * Z * Z)


def tf_profile(shape, box, omega, g, mu=None, norm=None):
    """sqrt(max(mu - V, 0)/g); with ``norm`` the chemical potential is solved so that
    sum |psi|^2 dV = norm (bisection in FP64).  Returns (psi, mu)."""
    V = trap_potential(shape, box, omega)
    dV = np.prod([L / n for L, n in zip(box, shape)])
    if mu is None:
        if norm is None:
            raise ValueError("give mu or norm")

        def n_of(m):
            return np.maximum(m - V, 0.0).sum() * dV / g

        lo, hi = 0.0, 1.0
        while n_of(hi) < norm:
            hi *= 2.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if n_of(mid) < norm:
                lo = mid
            else:
                hi = mid
            if hi - lo <= 1e-15 * hi:
                break
        mu = 0.5 * (lo + hi)
    return np.sqrt(np.maximum(mu - V, 0.0) / g).astype(np.complex128), mu


def vortex_line_factor(shape, box, x0, y0, charge, xi):
    """tanh(r/(sqrt(2) xi)) * exp(i*charge*atan2(y-y0, x-x0)) for a straight z-line."""
    X, Y, _ = mesh(shape, box)
    dx, dy = X - x0, Y - y0
    r = np.sqrt(dx * dx + dy * dy)
    return np.tanh(r / (math.sqrt(2.0) * xi)) * np.exp(1j * charge * np.arctan2(dy, dx))


def vortex_ring_factor(shape, box, radius, z0, xi):
    """A ring of the given radius in the plane z=z0 about the z axis: the phase winds
    once around the core circle in each (rho, z) half-plane."""
    X, Y, Z = mesh(shape, box)
    rho = np.sqrt(X * X + Y * Y)
    a, b = rho - radius, Z - z0
    d = np.sqrt(a * a + b * b)
    return np.tanh(d / (math.sqrt(2.0) * xi)) * np.exp(1j * np.arctan2(b, a))


def noise(shape, seed, amp, stream=0):
    """amp*(g1 + i g2), g1, g2 standard normals indexed by the flat point index."""
    idx = np.arange(int(np.prod(shape)), dtype=np.int64)
    g1, g2 = gaussian_pair(seed, stream, idx)
    return (amp * (g1 + 1j * g2)).reshape(shape)


def phase_noise(shape, seed, amp, stream=0):
    """exp(i theta), theta uniform in [-amp, amp]."""
    idx = np.arange(int(np.prod(shape)), dtype=np.int64)
    th = amp * (2.0 * uniform(seed, stream, idx) - 1.0)
    return np.exp(1j * th).reshape(shape)


@dataclass
class Problem:
    """One condensate in physical units plus the stepper's unit choice."""
    shape: tuple
    box: tuple
    omega: tuple = (1.0, 1.0, 1.0)
    g: float = 500.0
    Omega: float = 0.0
    dt: complex = 1e-3
    reference_units: bool = False   # psi scaled by sqrt(2g), c0 = 1, gnl = 1
    meta: dict = field(default_factory=dict)

    def kernel_params(self) -> dict:
        """Stepper coefficients (reference units when reference_units)."""
        harm = tuple((w * L) ** 2 for w, L in zip(self.omega, self.box))
        if self.reference_units:
            c0, gnl = 1.0, 1.0
        else:
            c0, gnl = 0.0, 2.0 * self.g
        return dict(box=self.box, harm_w=harm, omega_r=-2.0 * self.Omega, c0=c0, gnl=gnl,
                    dt=complex(self.dt))

    def psi_scale(self) -> float:
        return math.sqrt(2.0 * self.g) if self.reference_units else 1.0


def config_c1(seed: int = 20261203, n: int = 64, reference_units: bool = True):
    """64^3 real time, box [-8,8)^3, V = r^2/2, g = 500, TF mu = 10 times a z-line
    vortex, plus complex Gaussian noise of amplitude 1e-3; dt = 1e-3."""
    shape, box = (n, n, n), (16.0, 16.0, 16.0)
    pr = Problem(shape, box, (1.0, 1.0, 1.0), 500.0, 0.0, 1e-3, reference_units,
                 {"name": "c1", "seed": seed, "steps": 1000})
    psi, _ = tf_profile(shape, box, pr.omega, pr.g, mu=10.0)
    psi = psi * vortex_line_factor(shape, box, 0.0, 0.0, 1, 1.0 / math.sqrt(10.0))
    psi = psi + noise(shape, seed, 1e-3)
    return pr, psi * pr.psi_scale()


def config_m(n: int = 512, seed: int = 20261203, Omega: float = 0.0,
             reference_units: bool = False):
    """The throughput workload: n^3 on [-16,16)^3 (h = 1/16 at n = 512), V = r^2/2,
    g = 500, TF mu = 10 times a z-line vortex, noise 1e-3, dt = 5e-4."""
    shape, box = (n, n, n), (32.0, 32.0, 32.0)
    pr = Problem(shape, box, (1.0, 1.0, 1.0), 500.0, Omega, 5e-4, reference_units,
                 {"name": f"m{n}", "seed": seed})
    psi, _ = tf_profile(shape, box, pr.omega, pr.g, mu=10.0)
    psi = psi * vortex_line_factor(shape, box, 0.0, 0.0, 1, 1.0 / math.sqrt(10.0))
    psi = psi + noise(shape, seed, 1e-3)
    return pr, psi * pr.psi_scale()


def config_c2(seed: int = 20261203, n: int = 256):
    """256^3 on [-12,12)^3, V = (x^2+y^2+0.25 z^2)/2, g = 2000, TF with norm 1, one ring
    of radius 4 in z=0 and 4 z-lines at seeded positions r<5 with seeded +-1 charges;
    dt = 5e-4."""
    shape, box = (n, n, n), (24.0, 24.0, 24.0)
    pr = Problem(shape, box, (1.0, 1.0, 0.5), 2000.0, 0.0, 5e-4, False,
                 {"name": "c2", "seed": seed, "steps": 2000})
    psi, mu = tf_profile(shape, box, pr.omega, pr.g, norm=1.0)
    xi = 1.0 / math.sqrt(2.0 * mu)
    psi = psi * vortex_ring_factor(shape, box, 4.0, 0.0, xi)
    u = uniform(seed, 101, np.arange(12))
    lines = []
    for q in range(4):
        r = 5.0 * math.sqrt(u[3 * q])
        th = 2 * math.pi * u[3 * q + 1]
        ch = 1 if u[3 * q + 2] < 0.5 else -1
        lines.append((r * math.cos(th), r * math.sin(th), ch))
        psi = psi * vortex_line_factor(shape, box, lines[-1][0], lines[-1][1], ch, xi)
    pr.meta.update(mu=mu, lines=lines)
    return pr, psi


def config_c3(seed: int = 20261203, n: int = 256):
    """Imaginary-time ground state in a rotating trap: 256^3 on [-10,10)^3,
    w = (1,1,0.8), Omega = 0.7, g = 3000, TF (norm 1) times exp(i theta) with seeded
    uniform phase noise of amplitude 0.1; dtau = 1e-3."""
    shape, box = (n, n, n), (20.0, 20.0, 20.0)
    pr = Problem(shape, box, (1.0, 1.0, 0.8), 3000.0, 0.7, -1e-3j, False,
                 {"name": "c3", "seed": seed})
    psi, mu = tf_profile(shape, box, pr.omega, pr.g, norm=1.0)
    psi = psi * phase_noise(shape, seed, 0.1)
    pr.meta.update(mu=mu)
    return pr, psi


def reference_soliton(shape, box, harm_w):
    """The reference's own 'soliton' initial state in reference units: the
    setup_thomas_fermi profile (src/initial.f90:637-652), TF = (1 - pot)/sqrt(2)
    clipped at 0, psi = (TF, TF), then |psi| times exp(i*0) for x < 0 and exp(i*pi) for
    x >= 0 with the reference's 12-digit pi (src/initial.f90:65-77, src/cdata.f90:13)."""
    X, Y, Z = mesh(shape, box)
    Lx, Ly, Lz = box
    tf = 1.0 - harm_w[0] * (X / Lx) ** 2 - harm_w[1] * (Y / Ly) ** 2 - harm_w[2] * (Z / Lz) ** 2
    tf = tf / math.sqrt(2.0)
    tf = np.where(tf < 0.0, 0.0, tf)
    ld = np.hypot(tf, tf)
    pi12 = 3.14159265359
    return np.where(X < 0.0, ld + 0j, ld * complex(math.cos(pi12), math.sin(pi12)))


C4_GRIDS = ((64, 64, 64), (96, 96, 96), (128, 128, 128), (160, 160, 160), (96, 96, 192))


def config_c4(seed: int = 20261203, n_inst: int = 48, dx: float = 0.2, dtau: float = 1e-3,
              grids=C4_GRIDS):
    """Ensemble of independent imaginary-time problems: grids drawn by seed from `grids`
    with spacing dx (box = shape*dx); w_perp in [0.8, 1.2], w_z in [0.3, 1.5],
    g in [200, 4000], Omega in [0, 0.8], seeded uniform; start = TF (norm 1) times seeded
    phase noise of amplitude 0.1.  Returns [(Problem, psi0)]."""
    u = uniform(seed, 301, np.arange(6 * n_inst)).reshape(n_inst, 6)
    out = []
    for i in range(n_inst):
        shape = tuple(grids[min(int(u[i, 0] * len(grids)), len(grids) - 1)])
        box = tuple(n * dx for n in shape)
        wp = 0.8 + 0.4 * u[i, 1]
        wz = 0.3 + 1.2 * u[i, 2]
        g = 200.0 + 3800.0 * u[i, 3]
        Om = 0.8 * u[i, 4]
        pr = Problem(shape, box, (wp, wp, wz), g, Om, -dtau * 1j, False,
                     {"name": f"c4_{i}", "seed": seed})
        psi, mu = tf_profile(shape, box, pr.omega, g, norm=1.0)
        psi = psi * phase_noise(shape, seed + i, 0.1)
        pr.meta.update(mu=mu)
        out.append((pr, psi))
    return out


def ic_spec(name: str, n: int = None, seed: int = 20261203) -> tuple:
    """(Problem, spec) of a configuration for the device generator (tbec_init_state);
    the spec reproduces config_<name>'s numpy state."""
    if name in ("m", "c1"):
        n = n or (512 if name == "m" else 64)
        box = (32.0,) * 3 if name == "m" else (16.0,) * 3
        dt = 5e-4 if name == "m" else 1e-3
        ref = name == "c1"
        pr = Problem((n, n, n), box, (1.0, 1.0, 1.0), 500.0, 0.0, dt, ref,
                     {"name": f"{name}{n}", "seed": seed})
        spec = dict(omega=(1.0, 1.0, 1.0), g=500.0, mu=10.0, scale=pr.psi_scale(),
                    xi=1.0 / math.sqrt(10.0), lines=[(0.0, 0.0, 1)], noise_amp=1e-3,
                    seed=seed)
        return pr, spec
    if name == "c2":
        pr, _ = config_c2(seed, n or 256)
        spec = dict(omega=pr.omega, g=pr.g, mu=pr.meta["mu"], xi=1.0 / math.sqrt(2.0 * pr.meta["mu"]),
                    ring_radius=4.0, ring_z0=0.0, lines=pr.meta["lines"], seed=seed)
        return pr, spec
    if name == "c3":
        pr, _ = config_c3(seed, n or 256)
        spec = dict(omega=pr.omega, g=pr.g, mu=pr.meta["mu"], xi=1.0, phase_noise_amp=0.1,
                    seed=seed)
        return pr, spec
    raise ValueError(name)
