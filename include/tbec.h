/* libtbec: MI355X (gfx950) GPE time stepper -- C ABI.
 *
 * The reference has no plugin interface: the hot path is the argument-less module
 * procedure get_rhs() (src/finite_diff.f90:8-136) operating on module globals
 * (src/cdata.f90:5-17) and the inline RK3 / imaginary-time array updates of
 * program run (src/run.f90:19-55), configured at compile time by cparam.local
 * (src/cparam.local:1-44).  libtbec replaces that path with a handle that owns the
 * device state; the host program (trapbec_amd/driver, Fortran via iso_c_binding, or
 * Python via ctypes) keeps the reference's driver role.
 *
 * Conventions: every function returns 0 on success and a negative TBEC_ERR_* code on
 * failure; tbec_last_error() gives the message (thread-local).  Host psi buffers are
 * interior-only complex arrays (interleaved re, im) of the handle's precision, with z
 * fastest, then y, then x -- the reference's Psi(1:nz,1:ny,1:nx) (src/initial.f90:16).
 * A handle is used by one host thread at a time; work is ordered on its own stream.
 */
#ifndef TBEC_H
#define TBEC_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  TBEC_OK = 0,
  TBEC_ERR_ARG = -1,       /* invalid argument or parameter */
  TBEC_ERR_OOM = -2,       /* device allocation failed */
  TBEC_ERR_HIP = -3,       /* HIP runtime error */
  TBEC_ERR_NONFINITE = -4, /* NaN/Inf detected in a reduction */
  TBEC_ERR_STATE = -5      /* call not valid in the handle's current state */
};

enum { TBEC_F32 = 32, TBEC_F64 = 64 };

/* Stencil summation (FP32 only; FP64 always uses the reference order):
 *   TBEC_SUM_REFERENCE   the reference's left-to-right 7-tap sum (src/deriv.f90:24)
 *   TBEC_SUM_COMPENSATED each symmetric pair f[-k]+f[k] by an error-free TwoSum (s, e),
 *                        d_k = (s - 2 f0) + e, then 270 d1 - 27 d2 + 2 d3
 *   TBEC_SUM_CENTRED     d_k = (f[-k] - f0) + (f[k] - f0), both differences exact
 *                        (Sterbenz) for smooth data; same rounding as COMPENSATED,
 *                        fewer instructions. */
enum { TBEC_SUM_REFERENCE = 0, TBEC_SUM_COMPENSATED = 1, TBEC_SUM_CENTRED = 2 };

/* Step modes. */
enum {
  TBEC_REAL_RK3 = 0,       /* src/run.f90:42-55 (dt complex, times (1 - i*gamma)) */
  TBEC_IMAG_RK3 = 1,       /* RK3 with dt = (0,-dtau), FP64 renormalisation each step,
                              mu-shifted right-hand side (ground state is a fixed point) */
  TBEC_IMAG_EULER_REF = 2  /* src/run.f90:20-22: psi += dt*0.5*rhs, no renormalisation */
};

/* Run parameters: the cparam.local entries the stepper reads (src/cparam.local:1-44),
 * plus two coefficients that generalise the nonlinearity
 *   rhs += (c0 - gnl*|psi|^2 - pot)*psi          (src/finite_diff.f90:132; c0=gnl=1 there)
 * so the physical equation i psi_t = -lap/2 + V + g|psi|^2 - Omega L_z maps to
 * c0 = 0, gnl = 2g, harm_w = w^2 L^2, omega_z = -2 Omega. */
typedef struct {
  int32_t nx, ny, nz;       /* Nx, Ny, Nz */
  int32_t prec;             /* TBEC_F32 or TBEC_F64 */
  int32_t sum_mode;         /* TBEC_SUM_* */
  int32_t tile;             /* tile configuration id, -1 = library choice */
  double Lx, Ly, Lz;        /* box; h = L/N, x_i = (i - 1/2) h - L/2 */
  double harm_wx, harm_wy, harm_wz; /* pot = sum harm_w*(x/L)^2 (src/finite_diff.f90:34) */
  double omega_z;           /* frame term omega_z*i*(x D1y - y D1x) (src/finite_diff.f90:128) */
  double c0, gnl;
  double dt_re, dt_im;      /* complex dt (src/cparam.local:4) */
  double gamma;             /* dissipation: real-time dt -> dt*(1 - i*gamma) (src/run.f90:42) */
} tbec_params;

/* FP64 diagnostics of the current psi (dV = hx*hy*hz). */
typedef struct {
  double norm;       /* sum |psi|^2 dV */
  double energy;     /* sum [psi* H psi - (gnl/4)|psi|^4] dV, H from the stepper's rhs */
  double mu;         /* <psi|H|psi> / norm */
  double residual;   /* ||H psi - mu psi|| / ||psi|| */
  double max_abs;    /* maxPsi   (src/diagnostic.f90:10) */
  double min_abs;    /* minPsi   (src/diagnostic.f90:12) */
  double mean_abs;   /* meanPsi  (src/diagnostic.f90:14-16) */
  double int_psi2;   /* intPsi2 = sum |psi|^2 without dV (src/diagnostic.f90:22) */
} tbec_diag;

typedef struct {
  int64_t steps;        /* steps taken */
  int32_t converged;    /* 1 if the relative energy change stayed below tol for 8 steps */
  int32_t pad_;
  double energy, mu, residual, norm;
  double rel_change;    /* |E_n - E_{n-1}| / |E_n| at the last step */
} tbec_solve_result;

typedef struct tbec_handle tbec_handle;

/* Allocate device state for one condensate (src/initial.f90:16-25).  psi = 0. */
int tbec_create(const tbec_params *p, tbec_handle **out);
int tbec_destroy(tbec_handle *h);

/* Copy psi in/out (replaces reload_data / var_print's Psi record,
 * src/initial.f90:796-818, src/output.f90:103-107).  Host or device pointer variants;
 * the device variant takes an interior-only contiguous device buffer. */
int tbec_upload_psi(tbec_handle *h, const void *psi_host);
int tbec_download_psi(tbec_handle *h, void *psi_host);
int tbec_upload_psi_device(tbec_handle *h, const void *psi_dev);
int tbec_download_psi_device(tbec_handle *h, void *psi_dev);

/* Advance nsteps steps of the given mode (enqueued; returns before completion, except
 * TBEC_IMAG_RK3, which synchronises to read the finalisers' flag and returns
 * TBEC_ERR_NONFINITE for a non-finite norm or energy).
 * Real time advances the handle's time by Re(dt) per step in the working precision
 * (src/run.f90:55); t_inout, when not null, sets that time first and receives it after.
 * Each step evaluates the potential, the frame-term switch and the switch-off at the
 * time it starts (get_rhs reads t, src/finite_diff.f90:124,128). */
int tbec_step(tbec_handle *h, int64_t nsteps, int mode, double *t_inout);

/* One get_rhs evaluation of the current psi (src/finite_diff.f90:8-136) into a host
 * buffer (interior layout).  Test hook. */
int tbec_rhs(tbec_handle *h, void *rhs_host);

/* Diagnostics of the current psi; synchronises.  norm, energy and mu come from one
 * rhs pass; the GP residual ||(H - mu) psi|| / ||psi|| from a second pass with c0 = 2 mu,
 * which makes it the norm of that pass's rhs (no cancellation near a ground state). */
int tbec_diagnose(tbec_handle *h, tbec_diag *out);

/* tbec_diagnose that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The reduction partials live in a static
 * device array of 32768 four-double slots (1 MiB per GPU, held until the process ends)
 * instead of the handle's on-demand buffer; results equal tbec_diagnose's bit for bit.
 * TBEC_ERR_ARG when the grid, tile and segment need more slots (one per workgroup of a
 * stage pass, at most 8192 for Psi_info); TBEC_ERR_STATE when a tabulated potential of
 * the current time is not built yet.  Every call on one GPU shares the array: calls that
 * overlap on one GPU (other handles, other threads) race and must be serialised. */
int tbec_diagnose_noalloc(tbec_handle *h, tbec_diag *out);

/* Imaginary-time ground state: TBEC_IMAG_RK3 steps with dtau = -dt_im until
 * |E_n - E_{n-1}| < tol*|E_n| has held for 8 consecutive steps, or max_steps;
 * renormalises to `norm` every step.  The host checks convergence every check_every
 * steps (a captured graph in between).  With TBEC_PROGRESS=<seconds> in the
 * environment, a progress line goes to stderr at that interval. */
int tbec_imag_solve(tbec_handle *h, int64_t max_steps, double tol, double norm,
                    int check_every, tbec_solve_result *out);

/* Residual gate of the imaginary-time stop rule: with rmax > 0 a step counts toward
 * convergence (tbec_imag_solve, tbec_batch_imag_solve) only if its input's GP residual
 * ||(H - mu) psi|| / ||psi|| is at or below rmax as well, so a state that drifts slowly
 * far from any stationary state (the vortex-free saddle of a rotating trap) cannot meet
 * the rule.  0 (the default) disables the gate.  Takes effect at the next solve.  No
 * counterpart in the reference, whose imaginary time has no stop rule
 * (src/run.f90:19-35). */
int tbec_set_residual_gate(tbec_handle *h, double rmax);

/* Output-time diagnostics of the current psi, as the reference's var_print writes
 * them when vel_print / qpress_print are set (src/run.f90:64-66,
 * src/output.f90:80-91).  Both fill 3*nx*ny*nz reals of the working precision,
 * component slowest, then x, then y, z fastest (the memory order of the reference's
 * vel(k,j,i,a) / qpressure(k,j,i,a)); host buffers; synchronise.
 * tbec_velocity: get_phase (src/diagnostic.f90:27-46), the 6th-order gradient of
 *   atan2(Im psi, Re psi), not unwrapped.
 * tbec_qpressure: get_qpressure (src/diagnostic.f90:50-83), the 2nd-order gradient of
 *   the 4th-order Laplacian of |psi|.
 * Both use the handle's w array as scratch (w carries nothing across steps). */
int tbec_velocity(tbec_handle *h, void *vel_host);
int tbec_qpressure(tbec_handle *h, void *qp_host);

/* Vortex cores of the current psi by phase winding around grid plaquettes, counting
 * plaquettes whose corners all have |psi|^2 >= dens_min.  orientation 0: z-normal
 * (x,y) plaquettes, 1: x-normal (y,z), 2: y-normal (z,x).  Writes up to cap entries
 * (i, j, k, charge) -- the plaquette's lowest corner -- in (i, j, k) order and the total
 * number found in *count.  No reference counterpart in the solver; replaces the MATLAB
 * post-processing detectors (bin/vortex_detect.m, bin/gpeget2dvort_homg.m). */
int tbec_vortex_cores(tbec_handle *h, double dens_min, int orientation, int32_t *out,
                      int64_t cap, int64_t *count);

/* Seeded initial state generated on the device (no reference counterpart: the
 * reference builds states on the host, src/initial.f90, with a clock-seeded RNG):
 *   psi = scale * [ sqrt(max(mu - V, 0)/g) * ring * prod(lines) * exp(i theta) + noise ]
 * V = sum omega_a^2 x_a^2 / 2; ring and lines are tanh(d/(sqrt(2) xi)) exp(i q phi) vortex
 * factors; theta uniform in [-phase_noise_amp, phase_noise_amp]; noise complex Gaussian
 * of amplitude noise_amp.  Random numbers are a counter-based function of
 * (seed, stream, flat point index), identical to trapbec_amd/initial.py. */
typedef struct {
  double omega[3];
  double g, mu, scale, xi;
  double ring_radius, ring_z0;      /* no ring if ring_radius <= 0 */
  double noise_amp, phase_noise_amp;
  double line_x[8], line_y[8];
  int32_t line_q[8];
  int32_t n_lines;
  uint32_t seed;
} tbec_ic;
int tbec_init_state(tbec_handle *h, const tbec_ic *ic);

/* ---- the reference's own initial states and insertions --------------------------
 * The init_cond select of mesh_init (src/initial.f90:58-156, names in that order) and the
 * insertions applied before the real-time loop (src/run.f90:26,37-40), evaluated on the
 * device in the handle's precision, operand for operand as the reference writes them.
 * Values the reference draws from its clock-seeded generator (random_phase's waves and
 * amplitudes, random_points' centres, vortex_rings' angles and offsets, add_noise) are
 * drawn here from the counter-based generator: stream 201 random_phase (5 draws per
 * wave), 202 random_points (3 per line), 203 vortex_rings (6 per ring), 204 add_noise
 * (2 per point at the flat interior index), each mapped as runif(a, b) = a + u*(b - a)
 * in the working precision. */
enum {
  TBEC_IC_CONSTANT = 0, TBEC_IC_SOLITON = 1, TBEC_IC_RANDOM_PHASE = 2, TBEC_IC_CROW = 3,
  TBEC_IC_SINGLE_VORTEX = 4, TBEC_IC_SAUSAGE = 5, TBEC_IC_RANDOM_POINTS = 6,
  TBEC_IC_LEAP_FROG = 7, TBEC_IC_VORTEX_RINGS = 8, TBEC_IC_LINE_OF_LINES = 9,
  TBEC_IC_LINKED_RINGS = 10, TBEC_IC_LATTICE = 11, TBEC_IC_LATTICE2 = 12, TBEC_IC_BUNDLE = 13
};
enum {
  TBEC_INSERT_PCURRENT = 0,       /* setup_pcurrent (src/initial.f90:669-739), by the
                                     handle's potential kind */
  TBEC_INSERT_RING = 1,           /* setup_add_ring (:759-794) */
  TBEC_INSERT_DIPOLE = 2,         /* setup_add_dipole (:839-870) */
  TBEC_INSERT_PHASE_IMPRINT = 3,  /* phase_imprint (:820-836) */
  TBEC_INSERT_NOISE = 4           /* add_noise (:741-757) */
};
typedef struct {
  int32_t kind;             /* TBEC_IC_* (tbec_reference_state) */
  int32_t winding_number;   /* single_vortex, phase_imprint */
  int32_t vor_ring_count;   /* vortex_rings, 1..8 */
  uint32_t seed;            /* drawn values, see above */
  double ring_rad;          /* line_of_lines, linked_rings, vortex_rings, add_ring */
} tbec_reference_ic;
/* Replace psi by a reference state (harm_w of the params feeds the Thomas-Fermi ones;
 * two_dim of tbec_set_physics drops random_phase's z wave). */
int tbec_reference_state(tbec_handle *h, const tbec_reference_ic *ic);
/* Apply one insertion (TBEC_INSERT_*) to psi; winding_number, ring_rad and seed are read
 * from ic, the potential kind from tbec_set_physics. */
int tbec_reference_insert(tbec_handle *h, int what, const tbec_reference_ic *ic);

/* Set the norm target used by TBEC_IMAG_RK3 steps (default 1). */
int tbec_set_norm(tbec_handle *h, double norm);

/* Tile shape id (0..4, per precision) and x-segment length used by the stage kernel;
 * the autotuner (trapbec_amd/autotune.py) chooses them.  seg = 0 picks the longest
 * segment (256 down to 8 planes) that still gives >= 1024 workgroups, which is also
 * the default at creation.  Drops captured graphs. */
int tbec_set_tuning(tbec_handle *h, int tile, int seg);
/* A tile shape and segment of its own for the first RK stage of real-time steps
 * (TBEC_REAL_RK3), the stage that reads one array: tile = -1 (the default) runs it with
 * the tbec_set_tuning choice.  seg = 0: automatic, as for tbec_set_tuning.  Every other
 * stage, imaginary time and the diagnostics keep the handle's tuning.  No reference
 * counterpart (a knob of the stage kernel, chosen by the autotuner).  Drops captured
 * graphs. */
int tbec_set_first_stage_tuning(tbec_handle *h, int tile, int seg);
/* Boundary faces, as ghost_comm fills them (src/ghost.f90:27-192; `boundaries` and
 * `periodicz` of src/cparam.local): TBEC_BC_ZERO (default: ghosts 0), TBEC_BC_SOLID
 * (mirror about the edge point: ghost 0 = psi 2, -1 = 3, -2 = 4), TBEC_BC_PERIODIC
 * (x, y and z periodic, as the reference's periodic process topology).  periodicz = 1
 * makes z periodic whatever the boundaries.  Solid and periodic faces are refilled
 * before every stage (a pass over the 6 faces).  Drops captured graphs. */
enum { TBEC_BC_ZERO = 0, TBEC_BC_SOLID = 1, TBEC_BC_PERIODIC = 2 };
int tbec_set_boundaries(tbec_handle *h, int boundaries, int periodicz);

/* ---- physics beyond the harmonic trap -------------------------------------------
 * The reference selects these in cparam.local and evaluates them inside get_rhs
 * (src/finite_diff.f90:30-130).  The potential kinds are the reference's `potential`
 * names in order, plus TBEC_POT_TABLE (a caller-supplied Trap table, no reference
 * counterpart; the reference's own Trap is the mobius one). */
enum {
  TBEC_POT_ZERO = 0,            /* 'zero'           :31-32 */
  TBEC_POT_HARMONIC = 1,        /* 'harmonic'       :33-34 (default; harm_w of the params) */
  TBEC_POT_HARMONIC_RING = 2,   /* 'harmonic_ring'  :35-40, opening max(0, 5 - 0.1 t) */
  TBEC_POT_HARMONIC_RING2 = 3,  /* 'harmonic_ring2' :41-65 */
  TBEC_POT_BARRIER = 4,         /* 'barrier'        :66-71 */
  TBEC_POT_BOX_TRAP = 5,        /* 'box_trap'       :72-78, tilt 5 sin(0.1 t) z/Lz */
  TBEC_POT_CHANNEL = 6,         /* 'channel'        :79-96 */
  TBEC_POT_QUARTIC = 7,         /* 'quartic'        :97-99 */
  TBEC_POT_SPHERE = 8,          /* 'sphere'         :100-111 */
  TBEC_POT_MOBIUS = 9,          /* 'mobius'         :112-116, Trap of src/initial.f90:160-196 */
  TBEC_POT_TABLE = 10           /* pot = the table given to tbec_set_potential_table */
};
typedef struct {
  int32_t potential;        /* TBEC_POT_* */
  int32_t lbeam;            /* add the laser beam 20 exp(-|(x,y) - (70,0)|^2 / wbeam^2)
                               (src/finite_diff.f90:118-123) */
  int32_t two_dim;          /* drop the z second derivative (two_dim, :25-27) */
  int32_t pad_;
  double wbeam;             /* beam width */
  double potential_off;     /* pot = 0 while t > potential_off (:124); HUGE_VAL: never */
  double t_spindown;        /* frame term only while t < t_spindown (:128); HUGE_VAL: never */
  double advective_vx;      /* |vx| > epsilon: 2 i vx D1x psi replaces the frame term
                               (:126-127) */
} tbec_physics;
/* Select the physics.  Times are compared in the working precision, as the reference
 * compares its `real` t.  TBEC_POT_MOBIUS builds the Trap table on the device here.
 * Harmonic or zero potentials without laser or advective term keep the analytic
 * stage kernel; every other choice reads a tabulated potential, refreshed when the
 * potential changes with t (one extra pass per step for box_trap, none while the
 * table is constant).  Drops captured graphs. */
int tbec_set_physics(tbec_handle *h, const tbec_physics *ph);
/* The Trap table for TBEC_POT_TABLE (nx*ny*nz reals of the working precision, z
 * fastest, host buffer). */
int tbec_set_potential_table(tbec_handle *h, const void *trap_host);
/* pot over the interior at time t (nx*ny*nz reals, z fastest, host buffer); synchronises.
 * Uses the handle's w array as scratch (w carries nothing across steps). */
int tbec_potential(tbec_handle *h, double t, void *pot_host);
/* The handle's time t: real-time steps start from it and advance it; imaginary-time
 * steps and the diagnostics evaluate the potential at it (0 after creation, as the
 * reference's t, src/cdata.f90:17).  tbec_step's t_inout, when given, sets it first. */
int tbec_set_time(tbec_handle *h, double t);
/* Multiply psi by a straight vortex line along z at (x0, y0) with sense dir (+1 or -1):
 * psi *= pade(r) exp(i dir theta), the imprint of force_random_points
 * (src/force.f90:32-53; the reference draws x0, y0 uniform in [-L/4, L/4] and dir
 * every force_freq steps, src/run.f90:70-74). */
int tbec_imprint_vortex(tbec_handle *h, double x0, double y0, double dir);

/* How the low-storage RK3 carries w between stages (src/run.f90:46-54).
 * stored = 1: the reference's scheme, w in its own array (10 HBM transfers of a complex
 * value per point and step).  stored = 0 (default): w is rebuilt pointwise from the two
 * psi buffers, w_{s-1} = (psi_s - psi_{s-1}) / B_{s-1} (8 transfers); the same RK3
 * with one more rounding of psi per stage.  Real and imaginary time; batches always
 * rebuild.  Drops captured graphs. */
int tbec_set_w_storage(tbec_handle *h, int stored);
/* Tile order: 1 (default) gives each of the 8 XCDs a contiguous block of tiles so that
 * neighbouring tiles share that XCD's L2; 0 keeps the dispatch order; 2
 * (TBEC_REMAP_FIRST_STAGE) uses the contiguous order only for the stages that read one
 * array (the first RK stage, the reference's Euler step, the rhs pass) and the dispatch
 * order for the RK stages that also read a pointwise array. */
enum { TBEC_REMAP_FIRST_STAGE = 2 };
int tbec_set_xcd_remap(tbec_handle *h, int on);

/* Plain streaming rates over arrays of `bytes` bytes (allocated and freed here): the mean
 * time of a copy (one array read, one written) and of an add (two read, one written),
 * 16-byte accesses with non-temporal writes, over `reps` launches each.  The streaming
 * ceiling of the stepper's read/write mix.  No counterpart in the reference. */
int tbec_stream_rates(int64_t bytes, int reps, double *copy_us, double *add_us);

int tbec_sync(tbec_handle *h);
/* The handle's HIP stream (hipStream_t), for event timing by the caller. */
void *tbec_stream(tbec_handle *h);
/* Average duration (ms) of one stage kernel measured with HIP events over n stages
 * launched back to back on the handle's stream: real-time RK3 stages, or, when the
 * handle's dt is (0, -dtau), the imaginary-time stages with their reductions and
 * finalisers (this advances psi).  For the autotuner. */
int tbec_time_stages(tbec_handle *h, int nstages, double *ms_per_stage);

/* ---- batched ensembles: many independent condensates advanced together ---------
 * No reference counterpart (the reference runs one condensate per program).  All
 * instances share precision, summation mode and tile shape; grids, boxes, physics
 * parameters and dt may differ.  Each stage of every instance runs in one launch over
 * (instance, tile) work items; steps are captured into graphs; converged instances are
 * skipped by a per-instance flag. */
typedef struct tbec_batch tbec_batch;
int tbec_batch_create(int32_t n, const tbec_params *params, tbec_batch **out);
int tbec_batch_destroy(tbec_batch *b);
int tbec_batch_upload_psi(tbec_batch *b, int32_t i, const void *psi_host);
int tbec_batch_download_psi(tbec_batch *b, int32_t i, void *psi_host);
/* Imaginary-time ground states of all instances (TBEC_IMAG_RK3 semantics, each with its
 * own dtau = -dt_im); out receives n results. */
int tbec_batch_imag_solve(tbec_batch *b, int64_t max_steps, double tol, double norm,
                          int check_every, tbec_solve_result *out);
/* Tile and segment for every instance (seg = 0: automatic over the whole ensemble). */
int tbec_batch_set_tuning(tbec_batch *b, int tile, int seg);
/* tbec_set_residual_gate for every instance of the batch (next solve). */
int tbec_batch_set_residual_gate(tbec_batch *b, double rmax);
/* on = 1: during a solve, once the unfinished instances' work items number under one per
 * CU (256), they run on the shorter segment the automatic rule picks for them alone (the
 * batch's own segment is restored when the solve returns).  Speeds up a long tail of a few
 * slow instances; the FP64 reduction order of those instances then changes, so their
 * results match a single-handle solve to rounding rather than bit for bit.  Default 0. */
int tbec_batch_set_tail_segment(tbec_batch *b, int on);
void *tbec_batch_stream(tbec_batch *b);

/* ---- learned warm start ---------------------------------------------------------
 * No reference counterpart.  An MLP n_in -> depth x width (GELU) -> n_out whose GEMMs run
 * on BF16 MFMA with FP32 accumulation; FP32 master weights, Adam.  Parameters are laid
 * out per layer as W (fan_in x fan_out, row-major) then b (fan_out). */
typedef struct tbec_mlp tbec_mlp;
int tbec_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed,
                    tbec_mlp **out);
int tbec_mlp_destroy(tbec_mlp *m);
int64_t tbec_mlp_num_params(tbec_mlp *m);
int tbec_mlp_get_params(tbec_mlp *m, float *out);
int tbec_mlp_set_params(tbec_mlp *m, const float *in);
/* y (batch x n_out) = MLP(x (batch x n_in)); host buffers. */
int tbec_mlp_forward(tbec_mlp *m, const float *x, int32_t batch, float *y);
/* Mean-squared-error loss against t and its gradient w.r.t. every parameter. */
int tbec_mlp_loss_grad(tbec_mlp *m, const float *x, const float *t, int32_t batch,
                       double *loss, float *grads);
int tbec_mlp_set_data(tbec_mlp *m, const float *x, const float *t, int32_t n);
int tbec_mlp_train(tbec_mlp *m, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                   double *loss_hist);
/* Replace psi by the trilinear upsampling of a coarse (3, cn, cn, cn) field
 * (sqrt-density, cos phase, sin phase) on the same box; resets the imaginary-time state. */
int tbec_upload_coarse(tbec_handle *h, const float *coarse, int32_t cn);
/* Replace dst's psi by the trilinear interpolation of src's current psi (its
 * renormalisation applied) at dst's cell centres, coordinates clamped to src's outermost
 * cell centres; the two handles share one box, one precision and one device.  Resets
 * dst's imaginary-time state; synchronises.  No reference counterpart (the reference
 * solves on one grid): the step between the levels of a coarse-to-fine ground-state
 * solve, for callers that relax a state on a coarser grid first. */
int tbec_prolong(tbec_handle *dst, tbec_handle *src);
/* tbec_prolong for every instance of a batch: instance i of dst (its next solve's start)
 * from instance i of src; the batches hold as many instances, each pair on one box, in
 * one precision.  Synchronises.  No reference counterpart; the coarse-to-fine solve of
 * an ensemble. */
int tbec_batch_prolong(tbec_batch *dst, tbec_batch *src);

const char *tbec_last_error(void);
/* Library version string and the gfx target it was built for. */
const char *tbec_version(void);

#ifdef __cplusplus
}
#endif
#endif
