/* CPU oracle for the TrapBEC GPE stepper -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * liboracle (built by oracle/Makefile), and only as the checker.  The product library
 * libtbec never links or calls it.
 *
 * Part 1 (tbec_oracle_impl.h, instantiated for double and float) restates the
 * reference's stepper: get_rhs (src/finite_diff.f90:8-136), second_deriv and
 * cmplx_first_deriv (src/deriv.f90:12-25), the low-storage RK3 and imaginary Euler
 * updates (src/run.f90:20-55), coordinates (src/initial.f90:34-42), the potentials
 * of get_rhs (src/finite_diff.f90:30-124) with the mobius Trap table
 * (src/initial.f90:160-196), the vortex imprint of force_random_points
 * (src/force.f90:32-53) and Psi_info / Psi_norm (src/diagnostic.f90:7-25).  Pinned against the reference build bit for
 * bit (FP64, -ffp-contract=off) and against closed forms in tests/test_oracle.py.
 *
 * Part 2 (below, FP64 only) restates quantities that have no reference code:
 * dV-weighted norm, energy, chemical potential and GP residual, the renormalised
 * imaginary-time RK3 step and plaquette phase-winding vortex detection.  Pinned by
 * closed forms (g=0 harmonic ground state E = (wx+wy+wz)/2, Lz eigenstates, virial
 * theorem, brute-force winding of imprinted vortices) in tests/test_oracle.py.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int nx, ny, nz;        /* global interior grid */
  int dims_x, dims_y;    /* MPI process grid (1,1 for one domain) */
  int coord_x, coord_y;  /* this block's Cartesian coordinates */
  int two_dim;
  double Lx, Ly, Lz;
  double harm_w[3];      /* harm_wx, harm_wy, harm_wz */
  double omega_z;        /* reference-units frame coefficient */
  double c0, gnl;        /* (c0 - gnl*|psi|^2 - pot); both 1 in the reference */
  double dt_re, dt_im;   /* complex dt after the (1 - i*gamma) factor */
  double potential_off, t_spindown;
  int boundaries;        /* 0 zero, 1 solid (mirror about the edge point), 2 periodic */
  int periodicz;         /* z periodic whatever the boundaries */
  int potential;         /* ORC_POT_*: the `potential` select of src/finite_diff.f90:30-117 */
  int lbeam;             /* laser beam term (src/finite_diff.f90:118-123) */
  double wbeam;          /* its width */
  double advective_vx;   /* 2 i vx D1x psi in place of the frame term (:126-127) */
  const void *trap;      /* Trap(k,j,i) for ORC_POT_MOBIUS / ORC_POT_TABLE: interior
                            layout, working precision */
} orc_params;

/* Potential kinds, in the order the C ABI numbers them. */
enum {
  ORC_POT_ZERO = 0, ORC_POT_HARMONIC = 1, ORC_POT_HARMONIC_RING = 2,
  ORC_POT_HARMONIC_RING2 = 3, ORC_POT_BARRIER = 4, ORC_POT_BOX_TRAP = 5,
  ORC_POT_CHANNEL = 6, ORC_POT_QUARTIC = 7, ORC_POT_SPHERE = 8, ORC_POT_MOBIUS = 9,
  ORC_POT_TABLE = 10
};

#define T double
#define SUF _f64
#define HYPOT hypot
#define ATAN2 atan2
#define SQRT sqrt
#define SIN sin
#define COS cos
#define ACOS acos
#define EXP exp
#define POW pow
#define FLT_EPSILON_T DBL_EPSILON
#include "tbec_oracle_impl.h"
#undef T
#undef SUF
#undef HYPOT
#undef ATAN2
#undef SQRT
#undef SIN
#undef COS
#undef ACOS
#undef EXP
#undef POW
#undef FLT_EPSILON_T

#define T float
#define SUF _f32
#define HYPOT hypotf
#define ATAN2 atan2f
#define SQRT sqrtf
#define SIN sinf
#define COS cosf
#define ACOS acosf
#define EXP expf
#define POW powf
#define FLT_EPSILON_T FLT_EPSILON
#include "tbec_oracle_impl.h"
#undef T
#undef SUF
#undef HYPOT
#undef ATAN2
#undef SQRT
#undef SIN
#undef COS
#undef ACOS
#undef EXP
#undef POW
#undef FLT_EPSILON_T

/* ---------------- Part 2: FP64 quantities with no reference counterpart --------- */

/* Reductions over the interior of psi and rhs = get_rhs(psi):
 *   out[0] = sum |psi|^2 dV           (norm N)
 *   out[1] = Re sum conj(psi)*rhs dV
 *   out[2] = sum |psi|^4 dV
 *   out[3] = sum |rhs|^2 dV
 * In the stepper's units dpsi/dt = (i/2) rhs with H psi = -rhs/2 + (c0/2) psi, so
 *   <H> = -out[1]/2 + c0*out[0]/2,  E = <H> - gnl*out[2]/4,  mu = <H>/N,
 *   |H psi - mu psi|^2 = out[3]/4 - a*out[1] + a^2*out[0],  a = c0/2 - mu. */
void orc_sums_f64(const orc_params *p, const double *psi_pad, const double *rhs,
                  double out[4]) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const long sy = nz + 6, sx = (long)(ny + 6) * (nz + 6);
  const double dV = (p->Lx / nx) * (p->Ly / ny) * (p->Lz / nz);
  long double s[4] = {0, 0, 0, 0};
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const double *c = psi_pad + 2 * ((long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3));
        const double *r = rhs + 2 * (((long)i * ny + j) * nz + k);
        const double n2 = c[0] * c[0] + c[1] * c[1];
        s[0] += n2;
        s[1] += (long double)c[0] * r[0] + (long double)c[1] * r[1];
        s[2] += (long double)n2 * n2;
        s[3] += (long double)r[0] * r[0] + (long double)r[1] * r[1];
      }
  for (int q = 0; q < 4; ++q) out[q] = (double)(s[q] * dV);
}

/* Energy diagnostics from one get_rhs evaluation: out = {N, E, mu, residual}. */
void orc_energy_f64(const orc_params *p, double *psi_pad, double *rhs_scratch,
                    double t, double out[4]) {
  double s[4];
  orc_get_rhs_f64(p, psi_pad, rhs_scratch, t);
  orc_sums_f64(p, psi_pad, rhs_scratch, s);
  const double H = -0.5 * s[1] + 0.5 * p->c0 * s[0];
  const double mu = H / s[0];
  const double a = 0.5 * p->c0 - mu;
  double res2 = 0.25 * s[3] - a * s[1] + a * a * s[0];
  if (res2 < 0) res2 = 0;
  out[0] = s[0];
  out[1] = H - 0.25 * p->gnl * s[2];
  out[2] = mu;
  out[3] = sqrt(res2 / s[0]);
}

/* Renormalised imaginary-time RK3 step: the reference RK3 with dt = (0, -dtau)
 * (src/run.f90:46-55 with a complex dt, src/cparam.local:4), evaluated with the
 * constant coefficient c0_eff = c0 + 2*mu_prev (the chemical potential of the previous
 * step, so the ground state is an exact fixed point of the step), then psi scaled so
 * that sum |psi|^2 dV = norm_target.  diag receives {N, E, mu, residual} of psi_n from
 * the stage-1 right-hand side; *c0_eff is advanced to c0 + 2*mu_n. */
void orc_imag_rk3_step_f64(const orc_params *p, double *psi_pad, double *w, double *rhs,
                           double norm_target, double *c0_eff, double diag[4]) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const long n = (long)nx * ny * nz;
  const long sy = nz + 6, sx = (long)(ny + 6) * (nz + 6);
  const double dV = (p->Lx / nx) * (p->Ly / ny) * (p->Lz / nz);
  orc_params q = *p;
  q.c0 = *c0_eff;
  const double dr = q.dt_re, di = q.dt_im;
  const double cr = 0.0 * dr - 0.5 * di, ci = 0.0 * di + 0.5 * dr;
  double sums[4];
  for (int stage = 0; stage < 3; ++stage) {
    orc_get_rhs_f64(&q, psi_pad, rhs, 0.0);
    if (stage == 0) {
      orc_sums_f64(&q, psi_pad, rhs, sums);
      const double H = -0.5 * sums[1] + 0.5 * q.c0 * sums[0];
      const double mu = H / sums[0];
      const double a = 0.5 * q.c0 - mu;
      double res2 = 0.25 * sums[3] - a * sums[1] + a * a * sums[0];
      /* energy and mu of the state scaled exactly to norm_target: the quadratic part
       * of H scales with r = norm_target/N, the interaction part with r^2 */
      const double r = norm_target / sums[0], g4 = 0.25 * q.gnl * sums[2];
      const double hlin = H - 2.0 * g4;
      diag[0] = sums[0];
      diag[1] = r * hlin + r * r * g4;
      diag[2] = (r * hlin + r * r * 2.0 * g4) / norm_target;
      diag[3] = sqrt((res2 < 0 ? 0 : res2) / sums[0]);
    }
    for (long t = 0; t < n; ++t) {
      const double rr = rhs[2 * t], ri = rhs[2 * t + 1];
      const double crr = cr * rr - ci * ri, cri = cr * ri + ci * rr;
      if (stage == 0) {
        w[2 * t] = crr;
        w[2 * t + 1] = cri;
      } else if (stage == 1) {
        w[2 * t] = (-2.0 / 3.0) * w[2 * t] + crr;
        w[2 * t + 1] = (-2.0 / 3.0) * w[2 * t + 1] + cri;
      } else {
        w[2 * t] = -w[2 * t] + crr;
        w[2 * t + 1] = -w[2 * t + 1] + cri;
      }
    }
    const double b = stage == 0 ? 1.0 / 3.0 : (stage == 1 ? 1.0 : 0.5);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          double *c = psi_pad + 2 * ((long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3));
          const double *wv = w + 2 * (((long)i * ny + j) * nz + k);
          c[0] += b * wv[0];
          c[1] += b * wv[1];
        }
  }
  long double s = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const double *c = psi_pad + 2 * ((long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3));
        s += (long double)c[0] * c[0] + (long double)c[1] * c[1];
      }
  const double scale = sqrt(norm_target / (double)(s * dV));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        double *c = psi_pad + 2 * ((long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3));
        c[0] *= scale;
        c[1] *= scale;
      }
  *c0_eff = p->c0 + 2.0 * diag[2];
}

/* Vortex cores by phase winding around grid plaquettes.  orientation 0: z-normal
 * plaquettes (x,y)-(x+1,y)-(x+1,y+1)-(x,y+1) in every z-plane; 1: x-normal, corners
 * (y,z)-(y+1,z)-(y+1,z+1)-(y,z+1) in every x-plane; 2: y-normal, corners
 * (z,x)-(z+1,x)-(z+1,x+1)-(z,x+1) in every y-plane.  Only plaquettes whose four corners
 * have |psi|^2 >= dens_min count.  The winding is the sum of the four wrapped phase
 * differences / 2pi (an integer).  Writes up to cap entries (i, j, k, charge) of the
 * plaquette's lowest corner, in (i, j, k) order, and returns the total count. */
long orc_vortex_plaquettes_f64(const orc_params *p, const double *psi_pad, double dens_min,
                               int orientation, int *out, long cap) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const long sy = nz + 6, sx = (long)(ny + 6) * (nz + 6);
  long da, db;   /* offsets of the two in-plane axes, in (first, second) order */
  int ei = 0, ej = 0, ek = 0;
  if (orientation == 0) { da = sx; db = sy; ei = 1; ej = 1; }
  else if (orientation == 1) { da = sy; db = 1; ej = 1; ek = 1; }
  else { da = 1; db = sx; ek = 1; ei = 1; }
  long cnt = 0;
  for (int i = 0; i + ei < nx; ++i)
    for (int j = 0; j + ej < ny; ++j)
      for (int k = 0; k + ek < nz; ++k) {
        const long o0 = (long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3);
        const long o[4] = {o0, o0 + da, o0 + da + db, o0 + db};
        double ph[4];
        int ok = 1;
        for (int q = 0; q < 4; ++q) {
          const double a = psi_pad[2 * o[q]], b = psi_pad[2 * o[q] + 1];
          if (a * a + b * b < dens_min) ok = 0;
          ph[q] = atan2(b, a);
        }
        if (!ok) continue;
        double wsum = 0;
        for (int q = 0; q < 4; ++q) {
          double d = ph[(q + 1) & 3] - ph[q];
          while (d > M_PI) d -= 2 * M_PI;
          while (d <= -M_PI) d += 2 * M_PI;
          wsum += d;
        }
        const int n = (int)lround(wsum / (2 * M_PI));
        if (n != 0) {
          if (cnt < cap) {
            out[4 * cnt] = i;
            out[4 * cnt + 1] = j;
            out[4 * cnt + 2] = k;
            out[4 * cnt + 3] = n;
          }
          ++cnt;
        }
      }
  return cnt;
}
