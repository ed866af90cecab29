"""Python host interface to libtbec (ctypes over the C ABI in include/tbec.h).

This mirrors the reference's driver role (src/run.f90): parameters as in cparam.local
(src/cparam.local:1-44), psi in/out as in reload_data / var_print
(src/initial.f90:796-818, src/output.f90:71-108), the real-time RK3 loop
(src/run.f90:42-55), the imaginary-time loop (src/run.f90:19-33) and the
diagnostics (src/diagnostic.f90:7-25).  All compute runs in libtbec's HIP kernels;
there is no CPU fallback -- a missing library or GPU raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtbec.so")

TBEC_F32, TBEC_F64 = 32, 64
SUM_REFERENCE, SUM_COMPENSATED, SUM_CENTRED = 0, 1, 2
NUM_TILES = 5                   # tile shape ids 0..4 per precision
BC_ZERO, BC_SOLID, BC_PERIODIC = 0, 1, 2
BOUNDARIES = {"zero": BC_ZERO, "solid": BC_SOLID, "periodic": BC_PERIODIC}
REAL_RK3, IMAG_RK3, IMAG_EULER_REF = 0, 1, 2
ERRORS = {-1: "TBEC_ERR_ARG", -2: "TBEC_ERR_OOM", -3: "TBEC_ERR_HIP",
          -4: "TBEC_ERR_NONFINITE", -5: "TBEC_ERR_STATE"}


class TbecParams(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("prec", C.c_int32), ("sum_mode", C.c_int32), ("tile", C.c_int32),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("harm_wx", C.c_double), ("harm_wy", C.c_double), ("harm_wz", C.c_double),
                ("omega_z", C.c_double), ("c0", C.c_double), ("gnl", C.c_double),
                ("dt_re", C.c_double), ("dt_im", C.c_double), ("gamma", C.c_double)]


class TbecDiag(C.Structure):
    _fields_ = [("norm", C.c_double), ("energy", C.c_double), ("mu", C.c_double),
                ("residual", C.c_double), ("max_abs", C.c_double), ("min_abs", C.c_double),
                ("mean_abs", C.c_double), ("int_psi2", C.c_double)]


class TbecSolveResult(C.Structure):
    _fields_ = [("steps", C.c_int64), ("converged", C.c_int32), ("pad_", C.c_int32),
                ("energy", C.c_double), ("mu", C.c_double), ("residual", C.c_double),
                ("norm", C.c_double), ("rel_change", C.c_double)]


class TbecIc(C.Structure):
    _fields_ = [("omega", C.c_double * 3), ("g", C.c_double), ("mu", C.c_double),
                ("scale", C.c_double), ("xi", C.c_double), ("ring_radius", C.c_double),
                ("ring_z0", C.c_double), ("noise_amp", C.c_double),
                ("phase_noise_amp", C.c_double), ("line_x", C.c_double * 8),
                ("line_y", C.c_double * 8), ("line_q", C.c_int32 * 8),
                ("n_lines", C.c_int32), ("seed", C.c_uint32)]


class TbecPhysics(C.Structure):
    _fields_ = [("potential", C.c_int32), ("lbeam", C.c_int32), ("two_dim", C.c_int32),
                ("pad_", C.c_int32), ("wbeam", C.c_double),
                ("potential_off", C.c_double), ("t_spindown", C.c_double),
                ("advective_vx", C.c_double)]


# the reference's `potential` names (src/finite_diff.f90:30-117) plus a caller's table
POTENTIALS = {"zero": 0, "harmonic": 1, "harmonic_ring": 2, "harmonic_ring2": 3,
              "barrier": 4, "box_trap": 5, "channel": 6, "quartic": 7, "sphere": 8,
              "mobius": 9, "table": 10}


class TbecReferenceIc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("winding_number", C.c_int32),
                ("vor_ring_count", C.c_int32), ("seed", C.c_uint32), ("ring_rad", C.c_double)]


# the reference's init_cond names (src/initial.f90:58-156) and insertions (src/run.f90:26,37-40)
IC_KINDS = {"constant": 0, "soliton": 1, "random_phase": 2, "crow": 3, "single_vortex": 4,
            "sausage": 5, "random_points": 6, "leap_frog": 7, "vortex_rings": 8,
            "line_of_lines": 9, "linked_rings": 10, "lattice": 11, "lattice2": 12,
            "bundle": 13}
INSERTS = {"pcurrent": 0, "ring": 1, "dipole": 2, "phase_imprint": 3, "noise": 4}


class TbecError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


EXPORTS = ["tbec_create", "tbec_destroy", "tbec_upload_psi", "tbec_download_psi",
           "tbec_upload_psi_device", "tbec_download_psi_device", "tbec_step", "tbec_rhs",
           "tbec_diagnose", "tbec_diagnose_noalloc", "tbec_imag_solve",
           "tbec_set_residual_gate", "tbec_batch_set_residual_gate",
           "tbec_batch_set_tail_segment", "tbec_set_norm",
           "tbec_sync", "tbec_stream", "tbec_stream_rates",
           "tbec_time_stages", "tbec_set_tuning", "tbec_set_first_stage_tuning",
           "tbec_last_error", "tbec_version",
           "tbec_batch_create", "tbec_batch_destroy", "tbec_batch_upload_psi",
           "tbec_batch_download_psi", "tbec_batch_imag_solve", "tbec_batch_set_tuning",
           "tbec_batch_stream", "tbec_vortex_cores", "tbec_mlp_create", "tbec_mlp_destroy",
           "tbec_mlp_num_params", "tbec_mlp_get_params", "tbec_mlp_set_params",
           "tbec_mlp_forward", "tbec_mlp_loss_grad", "tbec_mlp_set_data", "tbec_mlp_train",
           "tbec_upload_coarse", "tbec_prolong", "tbec_batch_prolong",
           "tbec_set_xcd_remap", "tbec_init_state",
           "tbec_set_w_storage", "tbec_velocity", "tbec_qpressure", "tbec_set_boundaries",
           "tbec_set_physics", "tbec_set_potential_table", "tbec_potential", "tbec_set_time",
           "tbec_imprint_vortex", "tbec_reference_state", "tbec_reference_insert"]

_lib = None


def load(path: str = LIB_PATH):
    """Load libtbec.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    H = C.c_void_p
    vp = C.c_void_p
    L.tbec_create.argtypes = [C.POINTER(TbecParams), C.POINTER(H)]
    L.tbec_destroy.argtypes = [H]
    for f in ("tbec_upload_psi", "tbec_download_psi", "tbec_upload_psi_device",
              "tbec_download_psi_device", "tbec_rhs", "tbec_velocity", "tbec_qpressure"):
        getattr(L, f).argtypes = [H, vp]
    L.tbec_step.argtypes = [H, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    L.tbec_diagnose.argtypes = [H, C.POINTER(TbecDiag)]
    L.tbec_diagnose_noalloc.argtypes = [H, C.POINTER(TbecDiag)]
    L.tbec_imag_solve.argtypes = [H, C.c_int64, C.c_double, C.c_double, C.c_int,
                                  C.POINTER(TbecSolveResult)]
    L.tbec_set_norm.argtypes = [H, C.c_double]
    L.tbec_stream_rates.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
    L.tbec_set_residual_gate.argtypes = [H, C.c_double]
    L.tbec_batch_set_residual_gate.argtypes = [H, C.c_double]
    L.tbec_batch_set_tail_segment.argtypes = [H, C.c_int]
    L.tbec_set_tuning.argtypes = [H, C.c_int, C.c_int]
    L.tbec_set_first_stage_tuning.argtypes = [H, C.c_int, C.c_int]
    L.tbec_set_xcd_remap.argtypes = [H, C.c_int]
    L.tbec_set_w_storage.argtypes = [H, C.c_int]
    L.tbec_set_boundaries.argtypes = [H, C.c_int, C.c_int]
    L.tbec_init_state.argtypes = [H, C.POINTER(TbecIc)]
    L.tbec_set_physics.argtypes = [H, C.POINTER(TbecPhysics)]
    L.tbec_set_potential_table.argtypes = [H, vp]
    L.tbec_potential.argtypes = [H, C.c_double, vp]
    L.tbec_set_time.argtypes = [H, C.c_double]
    L.tbec_imprint_vortex.argtypes = [H, C.c_double, C.c_double, C.c_double]
    L.tbec_reference_state.argtypes = [H, C.POINTER(TbecReferenceIc)]
    L.tbec_reference_insert.argtypes = [H, C.c_int, C.POINTER(TbecReferenceIc)]
    L.tbec_sync.argtypes = [H]
    L.tbec_stream.argtypes = [H]
    L.tbec_stream.restype = C.c_void_p
    L.tbec_time_stages.argtypes = [H, C.c_int, C.POINTER(C.c_double)]
    L.tbec_last_error.restype = C.c_char_p
    L.tbec_version.restype = C.c_char_p
    L.tbec_vortex_cores.argtypes = [H, C.c_double, C.c_int, vp, C.c_int64,
                                     C.POINTER(C.c_int64)]
    L.tbec_batch_destroy.argtypes = [H]
    L.tbec_batch_upload_psi.argtypes = [H, C.c_int32, vp]
    L.tbec_batch_download_psi.argtypes = [H, C.c_int32, vp]
    L.tbec_batch_imag_solve.argtypes = [H, C.c_int64, C.c_double, C.c_double, C.c_int,
                                        C.POINTER(TbecSolveResult)]
    L.tbec_batch_set_tuning.argtypes = [H, C.c_int, C.c_int]
    L.tbec_batch_stream.argtypes = [H]
    L.tbec_batch_stream.restype = C.c_void_p
    for f in EXPORTS:
        if f not in ("tbec_stream", "tbec_last_error", "tbec_version", "tbec_batch_stream",
                     "tbec_mlp_num_params"):
            getattr(L, f).restype = C.c_int
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise TbecError(rc, load().tbec_last_error().decode())


class Condensate:
    """One condensate on the GPU (a libtbec handle).

    psi arrays are numpy complex64/complex128 (or torch device tensors for the *_device
    calls) indexed [x, y, z], C order -- the reference's Psi(k, j, i) memory order."""

    def __init__(self, shape, box, harm_w, *, omega_z=0.0, c0=1.0, gnl=1.0, dt=1e-3,
                 gamma=0.0, prec=64, sum_mode=None, tile=-1):
        L = load()
        if sum_mode is None:
            sum_mode = SUM_REFERENCE if prec == 64 else SUM_CENTRED
        box = (box, box, box) if np.isscalar(box) else tuple(box)
        dt = complex(dt)
        self.shape, self.box, self.prec = tuple(shape), box, prec
        self.cdt = np.complex128 if prec == 64 else np.complex64
        self.params = TbecParams(shape[0], shape[1], shape[2], prec, sum_mode, tile,
                                 box[0], box[1], box[2], harm_w[0], harm_w[1], harm_w[2],
                                 omega_z, c0, gnl, dt.real, dt.imag, gamma)
        self._h = C.c_void_p()
        _check(L.tbec_create(C.byref(self.params), C.byref(self._h)))
        self.t = 0.0

    @classmethod
    def from_problem(cls, pr, prec=64, **kw):
        kp = pr.kernel_params()
        return cls(pr.shape, pr.box, kp["harm_w"], omega_z=kp["omega_r"], c0=kp["c0"],
                   gnl=kp["gnl"], dt=kp["dt"], prec=prec, **kw)

    def close(self):
        if self._h:
            load().tbec_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # data movement
    def upload(self, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=self.cdt)
        assert a.shape == self.shape
        _check(load().tbec_upload_psi(self._h, a.ctypes.data_as(C.c_void_p)))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.cdt)
        _check(load().tbec_download_psi(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def init_state(self, spec: dict):
        """Generate a seeded initial state on the device (see initial.ic_spec)."""
        lines = spec.get("lines", [])
        ic = TbecIc((C.c_double * 3)(*spec["omega"]), spec["g"], spec["mu"],
                    spec.get("scale", 1.0), spec["xi"], spec.get("ring_radius", 0.0),
                    spec.get("ring_z0", 0.0), spec.get("noise_amp", 0.0),
                    spec.get("phase_noise_amp", 0.0),
                    (C.c_double * 8)(*[l[0] for l in lines]),
                    (C.c_double * 8)(*[l[1] for l in lines]),
                    (C.c_int32 * 8)(*[int(l[2]) for l in lines]), len(lines),
                    spec.get("seed", 0) & 0xFFFFFFFF)
        _check(load().tbec_init_state(self._h, C.byref(ic)))

    def upload_device(self, ptr: int):
        _check(load().tbec_upload_psi_device(self._h, C.c_void_p(ptr)))

    def download_device(self, ptr: int):
        _check(load().tbec_download_psi_device(self._h, C.c_void_p(ptr)))

    # stepping
    def step(self, nsteps: int, mode: int = REAL_RK3):
        t = C.c_double(self.t)
        _check(load().tbec_step(self._h, nsteps, mode, C.byref(t)))
        self.t = t.value

    def rhs(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.cdt)
        _check(load().tbec_rhs(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def diagnose(self) -> dict:
        d = TbecDiag()
        _check(load().tbec_diagnose(self._h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in TbecDiag._fields_}

    def diagnose_noalloc(self) -> dict:
        """tbec_diagnose_noalloc: the same diagnostics with no device allocation."""
        d = TbecDiag()
        _check(load().tbec_diagnose_noalloc(self._h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in TbecDiag._fields_}

    def imag_solve(self, max_steps: int, tol: float, norm: float = 1.0,
                   check_every: int = 50) -> dict:
        r = TbecSolveResult()
        _check(load().tbec_imag_solve(self._h, max_steps, tol, norm, check_every,
                                      C.byref(r)))
        return {k: getattr(r, k) for k, _ in TbecSolveResult._fields_ if k != "pad_"}

    def _rdt(self):
        return np.float64 if self.cdt == np.complex128 else np.float32

    def velocity(self) -> np.ndarray:
        """The reference's velocity output field (get_phase), [component, x, y, z]."""
        out = np.empty((3,) + tuple(self.shape), dtype=self._rdt())
        _check(load().tbec_velocity(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def qpressure(self) -> np.ndarray:
        """The reference's quantum-pressure output field, [component, x, y, z]."""
        out = np.empty((3,) + tuple(self.shape), dtype=self._rdt())
        _check(load().tbec_qpressure(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def vortex_cores(self, dens_min: float, orientation: int = 0, cap: int = 1 << 20):
        """(i, j, k, charge) of every plaquette with nonzero phase winding."""
        out = np.zeros((cap, 4), dtype=np.int32)
        cnt = C.c_int64()
        _check(load().tbec_vortex_cores(self._h, dens_min, orientation,
                                        out.ctypes.data_as(C.c_void_p), cap, C.byref(cnt)))
        return out[:min(cnt.value, cap)].copy()

    def set_residual_gate(self, rmax: float):
        """GP-residual gate of the imaginary-time stop rule (0: none)."""
        _check(load().tbec_set_residual_gate(self._h, rmax))

    def set_norm(self, norm: float):
        _check(load().tbec_set_norm(self._h, norm))

    def set_tuning(self, tile: int, seg: int):
        _check(load().tbec_set_tuning(self._h, tile, seg))

    def set_first_stage_tuning(self, tile: int, seg: int = 0):
        """Tile and segment of the first RK stage of real-time steps (tile -1: as
        set_tuning)."""
        _check(load().tbec_set_first_stage_tuning(self._h, tile, seg))

    def set_xcd_remap(self, on: int):
        """0 dispatch order, 1 XCD-contiguous tile order for every stage, 2 for the
        first RK stage only (TBEC_REMAP_FIRST_STAGE)."""
        _check(load().tbec_set_xcd_remap(self._h, int(on)))

    def set_boundaries(self, boundaries="zero", periodicz: bool = False):
        """Boundary faces as the reference's ghost_comm fills them: "zero", "solid"
        (mirror about the edge point) or "periodic"; periodicz makes z periodic."""
        bc = BOUNDARIES[boundaries] if isinstance(boundaries, str) else int(boundaries)
        _check(load().tbec_set_boundaries(self._h, bc, int(bool(periodicz))))

    def set_physics(self, potential="harmonic", lbeam: bool = False, wbeam: float = 4.0,
                    potential_off=None, t_spindown=None, advective_vx: float = 0.0,
                    two_dim: bool = False):
        """The reference's potential kind (a name of POTENTIALS), laser beam, switch-off
        times (None: never), advective velocity and two_dim (src/finite_diff.f90:23-130)."""
        kind = POTENTIALS[potential] if isinstance(potential, str) else int(potential)
        ph = TbecPhysics(kind, int(bool(lbeam)), int(bool(two_dim)), 0, wbeam,
                         math.inf if potential_off is None else potential_off,
                         math.inf if t_spindown is None else t_spindown, advective_vx)
        _check(load().tbec_set_physics(self._h, C.byref(ph)))

    def set_potential_table(self, trap: np.ndarray):
        """Trap(k,j,i) for potential="table", indexed [x, y, z]."""
        a = np.ascontiguousarray(trap, dtype=self._rdt())
        assert a.shape == tuple(self.shape)
        _check(load().tbec_set_potential_table(self._h, a.ctypes.data_as(C.c_void_p)))

    def potential(self, t: float | None = None) -> np.ndarray:
        """pot over the interior at time t (default: the condensate's t), [x, y, z]."""
        out = np.empty(tuple(self.shape), dtype=self._rdt())
        _check(load().tbec_potential(self._h, self.t if t is None else t,
                                     out.ctypes.data_as(C.c_void_p)))
        return out

    def set_time(self, t: float):
        _check(load().tbec_set_time(self._h, t))
        self.t = t

    def imprint_vortex(self, x0: float, y0: float, direction: float):
        """force_random_points' vortex imprint at (x0, y0) with sense +1 or -1."""
        _check(load().tbec_imprint_vortex(self._h, x0, y0, direction))

    def reference_state(self, kind: str, winding_number: int = 1, ring_rad: float = 4.0,
                        vor_ring_count: int = 1, seed: int = 0):
        """Replace psi by one of the reference's init_cond states, generated on the device."""
        ic = TbecReferenceIc(IC_KINDS[kind], winding_number, vor_ring_count,
                             seed & 0xFFFFFFFF, ring_rad)
        _check(load().tbec_reference_state(self._h, C.byref(ic)))

    def reference_insert(self, what: str, winding_number: int = 1, ring_rad: float = 4.0,
                         seed: int = 0):
        """Apply one of the reference's insertions (pcurrent, ring, dipole, phase_imprint,
        noise) to psi on the device."""
        ic = TbecReferenceIc(0, winding_number, 1, seed & 0xFFFFFFFF, ring_rad)
        _check(load().tbec_reference_insert(self._h, INSERTS[what], C.byref(ic)))

    def set_w_storage(self, stored: bool):
        """True: RK3 keeps w in its own array (the reference's scheme, 10 transfers per
        point-step); False (default): w is rebuilt from the two psi buffers (8)."""
        _check(load().tbec_set_w_storage(self._h, int(bool(stored))))

    def sync(self):
        _check(load().tbec_sync(self._h))

    def stream_ptr(self) -> int:
        return load().tbec_stream(self._h)

    def time_stages(self, nstages: int) -> float:
        ms = C.c_double()
        _check(load().tbec_time_stages(self._h, nstages, C.byref(ms)))
        return ms.value


class Batch:
    """Many independent condensates solved together in imaginary time (libtbec batch).

    `problems` is a list of trapbec_amd.initial.Problem with imaginary dt = (0, -dtau);
    all share one precision."""

    def __init__(self, problems, prec=32, sum_mode=None, tile=-1):
        L = load()
        if sum_mode is None:
            sum_mode = SUM_REFERENCE if prec == 64 else SUM_CENTRED
        self.n = len(problems)
        self.prec = prec
        self.cdt = np.complex128 if prec == 64 else np.complex64
        self.shapes = [tuple(pr.shape) for pr in problems]
        arr = (TbecParams * self.n)()
        for i, pr in enumerate(problems):
            kp = pr.kernel_params()
            dt = complex(kp["dt"])
            arr[i] = TbecParams(pr.shape[0], pr.shape[1], pr.shape[2], prec, sum_mode, tile,
                                pr.box[0], pr.box[1], pr.box[2], *kp["harm_w"], kp["omega_r"],
                                kp["c0"], kp["gnl"], dt.real, dt.imag, 0.0)
        self._params = arr
        self._h = C.c_void_p()
        L.tbec_batch_create.argtypes = [C.c_int32, C.POINTER(TbecParams), C.POINTER(C.c_void_p)]
        _check(L.tbec_batch_create(self.n, arr, C.byref(self._h)))

    def close(self):
        if self._h:
            load().tbec_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, i, psi):
        a = np.ascontiguousarray(psi, dtype=self.cdt)
        assert a.shape == self.shapes[i]
        _check(load().tbec_batch_upload_psi(self._h, i, a.ctypes.data_as(C.c_void_p)))

    def download(self, i):
        out = np.empty(self.shapes[i], dtype=self.cdt)
        _check(load().tbec_batch_download_psi(self._h, i, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_tuning(self, tile, seg):
        _check(load().tbec_batch_set_tuning(self._h, tile, seg))

    def set_residual_gate(self, rmax):
        _check(load().tbec_batch_set_residual_gate(self._h, rmax))

    def set_tail_segment(self, on: bool):
        """Shorter segment for a tail of few unfinished instances during solves (results
        then equal single-handle solves to rounding, not bit for bit)."""
        _check(load().tbec_batch_set_tail_segment(self._h, int(bool(on))))

    def imag_solve(self, max_steps, tol, norm=1.0, check_every=50):
        res = (TbecSolveResult * self.n)()
        _check(load().tbec_batch_imag_solve(self._h, max_steps, tol, norm, check_every, res))
        return [{k: getattr(r, k) for k, _ in TbecSolveResult._fields_ if k != "pad_"}
                for r in res]

    def stream_ptr(self):
        return load().tbec_batch_stream(self._h)


class MLP:
    """The warm-start network on the GPU (libtbec BF16-MFMA kernels)."""

    def __init__(self, n_in, width, depth, n_out, seed=1):
        L = load()
        for f, a in (("tbec_mlp_create", [C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_uint64, C.POINTER(C.c_void_p)]),
                     ("tbec_mlp_destroy", [C.c_void_p]),
                     ("tbec_mlp_get_params", [C.c_void_p, C.c_void_p]),
                     ("tbec_mlp_set_params", [C.c_void_p, C.c_void_p]),
                     ("tbec_mlp_forward", [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
                     ("tbec_mlp_loss_grad", [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                             C.POINTER(C.c_double), C.c_void_p]),
                     ("tbec_mlp_set_data", [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
                     ("tbec_mlp_train", [C.c_void_p, C.c_int32, C.c_int32, C.c_float,
                                         C.c_uint64, C.c_void_p]),
                     ("tbec_mlp_num_params", [C.c_void_p]),
                     ("tbec_upload_coarse", [C.c_void_p, C.c_void_p, C.c_int32])):
            getattr(L, f).argtypes = a
        L.tbec_mlp_num_params.restype = C.c_int64
        self.dims = [n_in] + [width] * depth + [n_out]
        self._h = C.c_void_p()
        _check(L.tbec_mlp_create(n_in, width, depth, n_out, seed, C.byref(self._h)))
        self.nparam = L.tbec_mlp_num_params(self._h)

    def close(self):
        if self._h:
            load().tbec_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get_params(self):
        out = np.empty(self.nparam, np.float32)
        _check(load().tbec_mlp_get_params(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam
        _check(load().tbec_mlp_set_params(self._h, p.ctypes.data_as(C.c_void_p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], self.dims[-1]), np.float32)
        _check(load().tbec_mlp_forward(self._h, x.ctypes.data_as(C.c_void_p), x.shape[0],
                                       y.ctypes.data_as(C.c_void_p)))
        return y

    def loss_grad(self, x, t, with_grads=True):
        """Mean squared error and (with_grads) every parameter's gradient; without, the
        gradients stay on the device (no host copy)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.nparam, np.float32) if with_grads else None
        loss = C.c_double()
        _check(load().tbec_mlp_loss_grad(self._h, x.ctypes.data_as(C.c_void_p),
                                         t.ctypes.data_as(C.c_void_p), x.shape[0],
                                         C.byref(loss),
                                         g.ctypes.data_as(C.c_void_p) if with_grads else None))
        return (loss.value, g) if with_grads else loss.value

    def set_data(self, x, t):
        self._x = np.ascontiguousarray(x, np.float32)
        self._t = np.ascontiguousarray(t, np.float32)
        _check(load().tbec_mlp_set_data(self._h, self._x.ctypes.data_as(C.c_void_p),
                                        self._t.ctypes.data_as(C.c_void_p), self._x.shape[0]))

    def train(self, epochs, batch, lr, seed=1):
        hist = np.zeros(epochs)
        _check(load().tbec_mlp_train(self._h, epochs, batch, lr, seed,
                                     hist.ctypes.data_as(C.c_void_p)))
        return hist


def stream_rates(nbytes: int, reps: int = 20):
    """(copy_us, add_us): plain streaming kernels over arrays of nbytes bytes."""
    cu, au = C.c_double(), C.c_double()
    _check(load().tbec_stream_rates(int(nbytes), int(reps), C.byref(cu), C.byref(au)))
    return cu.value, au.value


def upload_coarse(cond: Condensate, coarse: np.ndarray):
    """Start `cond` from the trilinear upsampling of a (3, n, n, n) coarse field."""
    c = np.ascontiguousarray(coarse, np.float32)
    assert c.ndim == 4 and c.shape[0] == 3
    L = load()
    L.tbec_upload_coarse.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    _check(L.tbec_upload_coarse(cond._h, c.ctypes.data_as(C.c_void_p), c.shape[1]))


def prolong(dst: Condensate, src: Condensate):
    """Start `dst` from the trilinear interpolation of `src`'s current psi at dst's cell
    centres (same box, precision and device); synchronises."""
    L = load()
    L.tbec_prolong.argtypes = [C.c_void_p, C.c_void_p]
    _check(L.tbec_prolong(dst._h, src._h))


def batch_prolong(dst: "Batch", src: "Batch"):
    """tbec_prolong for every instance: instance i of `dst` starts from instance i of
    `src` interpolated onto its grid; synchronises."""
    L = load()
    L.tbec_batch_prolong.argtypes = [C.c_void_p, C.c_void_p]
    _check(L.tbec_batch_prolong(dst._h, src._h))
