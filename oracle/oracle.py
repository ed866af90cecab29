"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  It wraps liboracle.so (oracle/Makefile), the C
restatement of the reference stepper; see tbec_oracle.c for what each function follows.

Arrays: psi is a numpy complex array indexed [x, y, z] (C order, z fastest), the
reference's Psi(k, j, i) memory order.  Padding with 3 zero ghost cells per side
happens here.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dims_x", C.c_int), ("dims_y", C.c_int),
                ("coord_x", C.c_int), ("coord_y", C.c_int), ("two_dim", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("harm_w", C.c_double * 3), ("omega_z", C.c_double),
                ("c0", C.c_double), ("gnl", C.c_double),
                ("dt_re", C.c_double), ("dt_im", C.c_double),
                ("potential_off", C.c_double), ("t_spindown", C.c_double),
                ("boundaries", C.c_int), ("periodicz", C.c_int),
                ("potential", C.c_int), ("lbeam", C.c_int), ("wbeam", C.c_double),
                ("advective_vx", C.c_double), ("trap", C.c_void_p)]

BOUNDARIES = {"zero": 0, "solid": 1, "periodic": 2}
# the reference's init_cond names (src/initial.f90:58-156) and insertions (src/run.f90:26,37-40)
IC_KINDS = {"constant": 0, "soliton": 1, "random_phase": 2, "crow": 3, "single_vortex": 4,
            "sausage": 5, "random_points": 6, "leap_frog": 7, "vortex_rings": 8,
            "line_of_lines": 9, "linked_rings": 10, "lattice": 11, "lattice2": 12,
            "bundle": 13}
INSERTS = {"pcurrent": 0, "ring": 1, "dipole": 2, "phase_imprint": 3, "noise": 4}
# the reference's `potential` names (src/finite_diff.f90:30-117) plus a user table
POTENTIALS = {"zero": 0, "harmonic": 1, "harmonic_ring": 2, "harmonic_ring2": 3,
              "barrier": 4, "box_trap": 5, "channel": 6, "quartic": 7, "sphere": 8,
              "mobius": 9, "table": 10}


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OrcParams)
        vp = C.c_void_p
        for suf in ("_f64", "_f32"):
            getattr(L, "orc_get_rhs" + suf).argtypes = [P, vp, vp, C.c_double]
            getattr(L, "orc_rk3_step" + suf).argtypes = [P, vp, vp, vp, C.c_double]
            getattr(L, "orc_rk3_step" + suf).restype = C.c_double
            getattr(L, "orc_euler_step" + suf).argtypes = [P, vp, vp, C.c_double]
            getattr(L, "orc_psi_info" + suf).argtypes = [P, vp, vp]
            getattr(L, "orc_velocity" + suf).argtypes = [P, vp, vp]
            getattr(L, "orc_qpressure" + suf).argtypes = [P, vp, vp]
            getattr(L, "orc_potential" + suf).argtypes = [P, C.c_double, vp]
            getattr(L, "orc_mobius_trap" + suf).argtypes = [P, vp]
            getattr(L, "orc_imprint_vortex" + suf).argtypes = [P, vp, C.c_double,
                                                                C.c_double, C.c_double]
            getattr(L, "orc_reference_state" + suf).argtypes = [P, C.c_int, C.c_int,
                                                                 C.c_double, vp, C.c_int, vp]
            getattr(L, "orc_reference_insert" + suf).argtypes = [P, C.c_int, C.c_int,
                                                                  C.c_double, vp, vp]
        L.orc_sums_f64.argtypes = [P, vp, vp, vp]
        L.orc_energy_f64.argtypes = [P, vp, vp, C.c_double, vp]
        L.orc_imag_rk3_step_f64.argtypes = [P, vp, vp, vp, C.c_double, vp, vp]
        L.orc_vortex_plaquettes_f64.argtypes = [P, vp, C.c_double, C.c_int, vp, C.c_long]
        L.orc_vortex_plaquettes_f64.restype = C.c_long
        _lib = L
    return _lib


def params(shape, L, harm_w, omega_z=0.0, c0=1.0, gnl=1.0, dt=1e-3, gamma=0.0,
           dims=(1, 1), coords=(0, 0), two_dim=False, potential_off=None,
           boundaries="zero", periodicz=False, potential="harmonic", lbeam=False,
           wbeam=4.0, advective_vx=0.0, t_spindown=None):
    """Reference-unit parameters.  dt may be complex; the reference multiplies it by
    (1 - i*gamma) before the real-time loop (src/run.f90:42).  For the "mobius" and
    "table" potentials attach the Trap array with Stepper.set_trap."""
    nx, ny, nz = shape
    Lx, Ly, Lz = (L, L, L) if np.isscalar(L) else L
    dtc = complex(dt) * complex(1.0, -gamma)
    p = OrcParams(nx, ny, nz, dims[0], dims[1], coords[0], coords[1], int(two_dim),
                  Lx, Ly, Lz, (C.c_double * 3)(*harm_w), omega_z, c0, gnl,
                  dtc.real, dtc.imag,
                  np.finfo(np.float64).max if potential_off is None else potential_off,
                  np.finfo(np.float64).max if t_spindown is None else t_spindown,
                  BOUNDARIES[boundaries], int(periodicz), POTENTIALS[potential], int(lbeam),
                  wbeam, advective_vx, None)
    return p


def pad(psi: np.ndarray, prec: int = 64) -> np.ndarray:
    cdt = np.complex128 if prec == 64 else np.complex64
    nx, ny, nz = psi.shape
    out = np.zeros((nx + 6, ny + 6, nz + 6), dtype=cdt)
    out[3:-3, 3:-3, 3:-3] = psi
    return out


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


class Stepper:
    """Holds padded psi, w and rhs of one domain at one precision."""

    def __init__(self, p: OrcParams, psi0: np.ndarray, prec: int = 64, t: float = 0.0):
        self.p, self.prec, self.t = p, prec, t
        self.suf = "_f64" if prec == 64 else "_f32"
        self.cdt = np.complex128 if prec == 64 else np.complex64
        self.psi = pad(psi0, prec)
        self.w = np.zeros(psi0.shape, dtype=self.cdt)
        self.rhs = np.zeros(psi0.shape, dtype=self.cdt)

    @property
    def interior(self) -> np.ndarray:
        return self.psi[3:-3, 3:-3, 3:-3].copy()

    def get_rhs(self) -> np.ndarray:
        getattr(lib(), "orc_get_rhs" + self.suf)(C.byref(self.p), _ptr(self.psi),
                                                _ptr(self.rhs), self.t)
        return self.rhs.copy()

    def rk3(self, nsteps: int) -> None:
        f = getattr(lib(), "orc_rk3_step" + self.suf)
        for _ in range(nsteps):
            self.t = f(C.byref(self.p), _ptr(self.psi), _ptr(self.w), _ptr(self.rhs), self.t)

    def euler(self, nsteps: int) -> None:
        f = getattr(lib(), "orc_euler_step" + self.suf)
        for _ in range(nsteps):
            f(C.byref(self.p), _ptr(self.psi), _ptr(self.rhs), self.t)

    def info(self) -> np.ndarray:
        out = np.zeros(4)
        getattr(lib(), "orc_psi_info" + self.suf)(C.byref(self.p), _ptr(self.psi), _ptr(out))
        return out

    def velocity(self) -> np.ndarray:
        """get_phase's velocity field, indexed [component, x, y, z]."""
        out = np.zeros((3,) + self.rhs.shape, dtype=self.psi.real.dtype)
        getattr(lib(), "orc_velocity" + self.suf)(C.byref(self.p), _ptr(self.psi), _ptr(out))
        return out

    def qpressure(self) -> np.ndarray:
        """get_qpressure's field, indexed [component, x, y, z]."""
        out = np.zeros((3,) + self.rhs.shape, dtype=self.psi.real.dtype)
        getattr(lib(), "orc_qpressure" + self.suf)(C.byref(self.p), _ptr(self.psi), _ptr(out))
        return out

    def set_trap(self, trap: np.ndarray) -> None:
        """Attach Trap(k,j,i) ([x, y, z], working precision) for the mobius/table kinds."""
        self._trap = np.ascontiguousarray(trap, dtype=self.psi.real.dtype)
        assert self._trap.shape == self.rhs.shape
        self.p.trap = self._trap.ctypes.data

    def mobius_trap(self) -> np.ndarray:
        """The reference's mobius Trap table for this grid (src/initial.f90:160-196)."""
        out = np.zeros(self.rhs.shape, dtype=self.psi.real.dtype)
        getattr(lib(), "orc_mobius_trap" + self.suf)(C.byref(self.p), _ptr(out))
        return out

    def potential(self, t: float | None = None) -> np.ndarray:
        """pot over the interior at time t (default: the stepper's t)."""
        out = np.zeros(self.rhs.shape, dtype=self.psi.real.dtype)
        getattr(lib(), "orc_potential" + self.suf)(C.byref(self.p),
                                                   self.t if t is None else t, _ptr(out))
        return out

    def imprint_vortex(self, x0: float, y0: float, direction: float) -> None:
        """force_random_points' imprint at a given centre and sense (src/force.f90:32-43)."""
        getattr(lib(), "orc_imprint_vortex" + self.suf)(C.byref(self.p), _ptr(self.psi),
                                                        x0, y0, direction)

    def reference_state(self, kind: str, winding: int = 1, ring_rad: float = 4.0,
                        prm=None) -> None:
        """Replace psi by one of the reference's init_cond states; prm holds the drawn or
        precomputed values the C function documents (working-precision values)."""
        a = np.ascontiguousarray(np.zeros(1) if prm is None else prm, dtype=np.float64)
        getattr(lib(), "orc_reference_state" + self.suf)(
            C.byref(self.p), IC_KINDS[kind], winding, ring_rad, _ptr(a),
            0 if prm is None else a.size, _ptr(self.psi))

    def reference_insert(self, what: str, winding: int = 1, ring_rad: float = 4.0,
                         prm=None) -> None:
        """Apply one of the reference's insertions to psi."""
        a = np.ascontiguousarray(np.zeros(1) if prm is None else prm, dtype=np.float64)
        getattr(lib(), "orc_reference_insert" + self.suf)(
            C.byref(self.p), INSERTS[what], winding, ring_rad, _ptr(a), _ptr(self.psi))

    # FP64-only quantities with no reference counterpart
    def energy(self) -> dict:
        assert self.prec == 64
        out = np.zeros(4)
        lib().orc_energy_f64(C.byref(self.p), _ptr(self.psi), _ptr(self.rhs), self.t,
                             _ptr(out))
        return {"norm": out[0], "energy": out[1], "mu": out[2], "residual": out[3]}

    def sums(self) -> np.ndarray:
        assert self.prec == 64
        self.get_rhs()
        out = np.zeros(4)
        lib().orc_sums_f64(C.byref(self.p), _ptr(self.psi), _ptr(self.rhs), _ptr(out))
        return out

    def imag_rk3(self, nsteps: int, norm_target: float, tol: float = 0.0,
                 hits: int = 8, rgate: float = 0.0) -> list:
        """Renormalised imaginary-time RK3 (dt must be (0, -dtau)).  Stops early once the
        relative energy change between consecutive steps has stayed below tol for `hits`
        steps in a row; with rgate > 0 a step counts only if the GP residual of its input
        is at or below rgate as well.  Returns the per-step diagnostics
        [N, E, mu, residual] of the state entering each step."""
        assert self.prec == 64
        if not hasattr(self, "c0_eff"):
            self.c0_eff = np.array([self.p.c0])
        hist = []
        diag = np.zeros(4)
        run = 0
        for _ in range(nsteps):
            lib().orc_imag_rk3_step_f64(C.byref(self.p), _ptr(self.psi), _ptr(self.w),
                                        _ptr(self.rhs), norm_target, _ptr(self.c0_eff),
                                        _ptr(diag))
            hist.append(diag.copy())
            if tol > 0 and len(hist) > 1:
                ok = abs(hist[-1][1] - hist[-2][1]) < tol * abs(hist[-1][1]) and \
                    (rgate <= 0 or hist[-1][3] <= rgate)
                run = run + 1 if ok else 0
                if run >= hits:
                    break
        return hist

    def vortices(self, dens_min: float, orientation: int = 0, cap: int = 1 << 20) -> np.ndarray:
        assert self.prec == 64
        out = np.zeros((cap, 4), dtype=np.int32)
        n = lib().orc_vortex_plaquettes_f64(C.byref(self.p), _ptr(self.psi), dens_min,
                                            orientation, _ptr(out), cap)
        return out[:min(n, cap)].copy()
