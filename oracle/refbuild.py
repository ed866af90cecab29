"""Build and run the reference TrapBEC Fortran (test infrastructure only).

Recipe: for one parameter set, make ``oracle/_ref/<name>/`` holding symlinks to the
reference's own ``src/*.f90`` (nothing is copied), a generated ``cparam.local``
(the reference's compile-time parameter file, ``src/cparam.local:1-44``), and the
program compiled with ``amdflang -O3`` against MPICH (``/opt/conda``), in the module
order of ``src/Makefile:25-26``.  Fortran ``include`` resolves ``cparam.local`` from the
directory of the including file first, so the symlinked ``cdata.f90`` picks up the
generated parameters.

FP64 builds add ``-fdefault-real-8`` and remap ``MPI_REAL``/``MPI_COMPLEX`` so multi-rank
halo exchanges (``src/ghost.f90:34-140``) move whole doubles.

Runs inject a seeded psi0 through the restart file ``data/procNN/var.dat``
(``src/initial.f90:44-57,796-818``: two-digit directory, records [itime] [t] [Psi]) and
read the snapshot ``data/procNNN/var0001.dat`` (``src/output.f90:71-78``: one record
``coords(2) int4, Lx, Ly, Lz, t, Psi``).  The reference never leaves this container;
only fixtures written from its outputs are committed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import numpy as np

REF_SRC = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
REF_OUT = os.path.join(HERE, "_ref")
MODULES = ["cdata", "mpi_var", "statistics", "initial", "ghost", "deriv",
           "finite_diff", "output", "diagnostic", "force"]
MPI_INC = "/opt/conda/include"
MPI_LIB = "/opt/conda/lib"
MPIEXEC = "/opt/conda/bin/mpiexec"

# Defaults: every switch that no configuration uses is off (zero BC, harmonic trap,
# no laser, no dissipation, no forcing, 3D).
DEFAULTS = dict(
    Lx=16.0, Ly=16.0, Lz=16.0, Nx=16, Ny=16, Nz=16, dt_re=1e-3, dt_im=0.0,
    shots=10, nsteps=10, boundaries="zero", periodicz=False, init_cond="constant",
    potential="harmonic", harm_wx=256.0, harm_wy=256.0, harm_wz=256.0,
    potential_off=None, gamma=0.0, two_dim=False, omega_z=0.0, advective_vx=0.0,
    lbeam=False, wbeam=4.0, t_spindown=None, winding_number=1, ring_rad=4.0,
    add_ring=False, add_dipole=False, add_pcurrent=False, phase_imprint_on=False,
    imaginary_time=False, imag_nsteps=1, full_binary_output=True,
    imaginary_output=False, vel_print=False, qpress_print=False,
)


def _fl(x: float) -> str:
    """A Fortran default-real literal that round-trips the double value."""
    r = repr(float(x))
    if "e" in r or "E" in r:
        m, e = r.lower().split("e")
        if "." not in m:
            m += "."
        return f"{m}e{e}"
    if "." not in r:
        r += "."
    return r


def _lg(b: bool) -> str:
    return ".true." if b else ".false."


def cparam_text(p: dict) -> str:
    off = "huge(1.)" if p["potential_off"] is None else _fl(p["potential_off"])
    spin = "huge(1.)" if p["t_spindown"] is None else _fl(p["t_spindown"])
    return "\n".join([
        f"real  :: Lx={_fl(p['Lx'])}, Ly={_fl(p['Ly'])}, Lz={_fl(p['Lz'])}",
        f"integer, parameter :: Nx={p['Nx']}, Ny={p['Ny']}, Nz={p['Nz']}",
        f"complex :: dt=({_fl(p['dt_re'])},{_fl(p['dt_im'])})",
        f"integer, parameter :: shots={p['shots']}",
        f"integer, parameter :: nsteps={p['nsteps']}",
        f"character(len=*), parameter :: boundaries='{p['boundaries']}'",
        f"logical, parameter :: periodicz={_lg(p['periodicz'])}",
        f"character(len=*), parameter :: init_cond='{p['init_cond']}'",
        f"character(len=*), parameter :: potential='{p['potential']}'",
        f"real, parameter :: harm_wx={_fl(p['harm_wx'])}, harm_wy={_fl(p['harm_wy'])}, "
        f"harm_wz={_fl(p['harm_wz'])}",
        f"real, parameter :: potential_off={off}",
        f"real, parameter :: gamma={_fl(p['gamma'])}",
        f"logical, parameter :: two_dim={_lg(p['two_dim'])}",
        f"logical, parameter :: phase_imprint_on={_lg(p['phase_imprint_on'])}",
        f"logical, parameter :: lbeam={_lg(p['lbeam'])}",
        f"logical, parameter :: add_pcurrent={_lg(p['add_pcurrent'])}",
        f"real, parameter :: wbeam={_fl(p['wbeam'])}",
        f"real, parameter :: advective_vx={_fl(p['advective_vx'])}",
        f"real, parameter :: omega_z={_fl(p['omega_z'])}",
        f"real, parameter :: t_spindown={spin}",
        f"integer, parameter :: winding_number={int(p['winding_number'])}",
        "integer, parameter :: vor_ring_count=1",
        f"real,parameter :: ring_rad={_fl(p['ring_rad'])}",
        f"logical, parameter :: add_ring={_lg(p['add_ring'])}",
        f"logical, parameter :: add_dipole={_lg(p['add_dipole'])}",
        "integer, parameter :: vor_line_count=1",
        f"logical, parameter :: imaginary_time={_lg(p['imaginary_time'])}",
        f"integer, parameter :: imag_nsteps={p['imag_nsteps']}",
        "logical, parameter :: forcing=.false.",
        "integer, parameter :: force_freq=10",
        f"logical, parameter :: full_binary_output={_lg(p['full_binary_output'])}",
        f"logical, parameter :: imaginary_output={_lg(p['imaginary_output'])}",
        "logical, parameter :: gnuplot_2d_print=.false.",
        f"logical, parameter :: qpress_print={_lg(p['qpress_print'])}",
        f"logical, parameter :: vel_print={_lg(p['vel_print'])}",
        "",
    ])


def build(name: str, prec: int = 64, contract: str | None = "off", **params) -> str:
    """Compile one reference variant into oracle/_ref/<name>/run.x; return its directory.
    contract: the -ffp-contract mode (fixtures use "off", so the C restatement can match
    bit for bit); None leaves the compiler's default, i.e. the reference Makefile's plain
    -O3 (src/Makefile:2), which the CPU timing builds use."""
    if not os.path.isdir(REF_SRC):
        raise FileNotFoundError(f"reference sources not found at {REF_SRC}")
    p = dict(DEFAULTS)
    unknown = set(params) - set(DEFAULTS)
    if unknown:
        raise KeyError(f"unknown cparam keys {sorted(unknown)}")
    p.update(params)
    d = os.path.join(REF_OUT, name)
    shutil.rmtree(d, ignore_errors=True)
    os.makedirs(d)
    for m in MODULES + ["run"]:
        os.symlink(os.path.join(REF_SRC, m + ".f90"), os.path.join(d, m + ".f90"))
    with open(os.path.join(d, "cparam.local"), "w") as f:
        f.write(cparam_text(p))
    flags = ["-O3", f"-I{MPI_INC}"]
    if contract is not None:
        flags.insert(1, f"-ffp-contract={contract}")
    if prec == 64:
        flags += ["-fdefault-real-8", "-cpp", "-DMPI_REAL=MPI_DOUBLE_PRECISION",
                  "-DMPI_COMPLEX=MPI_DOUBLE_COMPLEX"]
    objs = []
    for m in MODULES:
        subprocess.run(["amdflang", *flags, "-c", m + ".f90", "-o", m + ".o"], cwd=d,
                       check=True)
        objs.append(m + ".o")
    subprocess.run(["amdflang", *flags, "run.f90", *objs, "-o", "run.x", f"-L{MPI_LIB}",
                    f"-Wl,-rpath,{MPI_LIB}", "-lmpifort", "-lmpi"], cwd=d, check=True)
    with open(os.path.join(d, "build.txt"), "w") as f:
        f.write(f"prec={prec} contract={contract} params={p}\n")
    return d


def _dims_create(nprocs: int) -> tuple[int, int]:
    """MPI_DIMS_CREATE(nprocs, 2): the most balanced factorisation, larger first."""
    best = (nprocs, 1)
    for a in range(1, nprocs + 1):
        if nprocs % a == 0:
            b = nprocs // a
            if a >= b and a - b < best[0] - best[1]:
                best = (a, b)
    return best


def _write_record(f, payload: bytes) -> None:
    n = np.array([len(payload)], dtype=np.int32).tobytes()
    f.write(n + payload + n)


def _read_records(path: str) -> list[bytes]:
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        n = int(np.frombuffer(data[pos:pos + 4], dtype=np.int32)[0])
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def run(d: str, psi0, prec: int = 64, nprocs: int = 1, itime0: int = 1,
        t0: float = 0.0, timeout: float = 3600.0, shape=None, snapshot: str = "var0001.dat"
        ) -> dict:
    """Run the reference from psi0 (array indexed [x, y, z]); return the final state.

    The reference stores Psi(k, j, i) with z fastest; a C-ordered numpy array indexed
    [x, y, z] has the same memory order.
    """
    cdt = np.complex128 if prec == 64 else np.complex64
    rdt = np.float64 if prec == 64 else np.float32
    nx, ny, nz = psi0.shape if psi0 is not None else shape
    dx, dy = _dims_create(nprocs)
    mx, my = nx // dx, ny // dy
    data = os.path.join(d, "data")
    shutil.rmtree(data, ignore_errors=True)
    os.makedirs(data)
    for r in range(nprocs):
        cx, cy = r // dy, r % dy
        os.makedirs(os.path.join(data, f"proc{r:02d}"), exist_ok=True)
        os.makedirs(os.path.join(data, f"proc{r:03d}"), exist_ok=True)
        if psi0 is None:   # no restart file: the reference builds its own initial state
            os.rmdir(os.path.join(data, f"proc{r:02d}"))
            continue
        block = np.ascontiguousarray(psi0[cx * mx:(cx + 1) * mx, cy * my:(cy + 1) * my, :],
                                     dtype=cdt)
        with open(os.path.join(data, f"proc{r:02d}", "var.dat"), "wb") as f:
            _write_record(f, np.array([itime0], dtype=np.int32).tobytes())
            _write_record(f, np.array([t0], dtype=rdt).tobytes())
            _write_record(f, block.tobytes())
    for leftover in ("ERROR",):
        if os.path.exists(os.path.join(d, leftover)):
            os.remove(os.path.join(d, leftover))
    env = dict(os.environ)
    cmd = [MPIEXEC, "-n", str(nprocs), "./run.x"]
    res = subprocess.run(cmd, cwd=d, env=env, capture_output=True, text=True, errors="replace",
                         timeout=timeout)
    if res.returncode != 0 or os.path.exists(os.path.join(d, "ERROR")):
        raise RuntimeError(f"reference run failed ({res.returncode}):\n{res.stdout}\n"
                           f"{res.stderr}")
    out = np.zeros((nx, ny, nz), dtype=cdt)
    t_out = None
    for r in range(nprocs):
        recs = _read_records(os.path.join(data, f"proc{r:03d}", snapshot))
        rec = recs[0]
        coords = np.frombuffer(rec[:8], dtype=np.int32)
        off = 8
        hdr = np.frombuffer(rec[off:off + 4 * rdt().itemsize], dtype=rdt)
        off += 4 * rdt().itemsize
        blk = np.frombuffer(rec[off:], dtype=cdt).reshape(mx, my, nz)
        cx, cy = int(coords[0]), int(coords[1])
        out[cx * mx:(cx + 1) * mx, cy * my:(cy + 1) * my, :] = blk
        t_out = float(hdr[3])
    return {"psi": out, "t": t_out, "stdout": res.stdout}


def read_field(d: str, prefix: str, shape, nprocs: int = 1, prec: int = 64, shot: int = 1
               ) -> np.ndarray:
    """A vector field written by var_print (``src/output.f90:80-91``: ``vel%02d.dat`` /
    ``qpress%02d.dat``, one record ``coords(2) int4, field(nz, ny, nx, 3)``), assembled
    over ranks as an array indexed [component, x, y, z]."""
    rdt = np.float64 if prec == 64 else np.float32
    nx, ny, nz = shape
    dx, dy = _dims_create(nprocs)
    mx, my = nx // dx, ny // dy
    out = np.zeros((3, nx, ny, nz), dtype=rdt)
    for r in range(nprocs):
        rec = _read_records(os.path.join(d, "data", f"proc{r:03d}",
                                         f"{prefix}{shot:02d}.dat"))[0]
        coords = np.frombuffer(rec[:8], dtype=np.int32)
        blk = np.frombuffer(rec[8:], dtype=rdt).reshape(3, mx, my, nz)
        cx, cy = int(coords[0]), int(coords[1])
        out[:, cx * mx:(cx + 1) * mx, cy * my:(cy + 1) * my, :] = blk
    return out


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else "probe"))
