/* CPU restatement of the TrapBEC GPE stepper -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * code, and only as the checker.  The product path (libtbec) never calls it.
 *
 * This header is included twice by tbec_oracle.c with T = double (suffix _f64) and
 * T = float (suffix _f32), mirroring the reference's FP64 (-fdefault-real-8) and FP32
 * (default-kind) builds.  Every arithmetic expression keeps the reference's operand
 * order so the FP64 build can be compared bit for bit with the reference compiled
 * with -ffp-contract=off.
 *
 * Layout: psi is padded by G=3 ghost cells on every side, (nx+6)*(ny+6)*(nz+6)
 * elements of interleaved (re, im), z fastest, matching Psi(-2:nz+3,-2:ny+3,-2:nx+3)
 * (src/initial.f90:16).  rhs and w are interior-only nx*ny*nz, z fastest
 * (src/initial.f90:24).
 *
 * Pinning (tests/test_oracle.py): every function here is checked against the
 * reference build (oracle/refbuild.py) and against closed forms; see the test file.
 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

typedef struct { T re, im; } FN(cplx);

/* Cell-centred coordinates, src/initial.f90:34-42 (per-rank offset form). */
static void FN(coords)(const orc_params *p, T *xx, T *yy, T *zz) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const T Lx = (T)p->Lx, Ly = (T)p->Ly, Lz = (T)p->Lz;
  for (int k = 1; k <= mz; ++k)
    zz[k - 1] = ((T)(2 * k - 1) / ((T)2 * (T)mz)) * Lz - Lz / (T)2;
  for (int j = 1; j <= my; ++j)
    yy[j - 1] = ((T)(2 * j - 1) / ((T)2 * (T)my)) * Ly / (T)p->dims_y +
                (T)p->coord_y * Ly / (T)p->dims_y - Ly / (T)2;
  for (int i = 1; i <= mx; ++i)
    xx[i - 1] = ((T)(2 * i - 1) / ((T)2 * (T)mx)) * Lx / (T)p->dims_x +
                (T)p->coord_x * Lx / (T)p->dims_x - Lx / (T)2;
}

/* Complex / real as flang evaluates it: the real operand is promoted to (d, 0) and the
 * textbook quotient ((a*d + b*0)/(d*d + 0*0), (b*d - a*0)/(d*d + 0*0)) is taken, i.e.
 * round(round(a*d)/round(d*d)).  Established by comparing against amdflang output
 * (tests/test_oracle.py::test_matches_reference_*). */
static inline FN(cplx) FN(cdiv_real)(FN(cplx) z, T d) {
  const T dd = d * d;
  FN(cplx) r;
  r.re = (z.re * d) / dd;
  r.im = (z.im * d) / dd;
  return r;
}

/* second_deriv, src/deriv.f90:19-25: left-to-right 7-tap sum, then / (180*h**2). */
static inline FN(cplx) FN(d2)(const FN(cplx) *v, long s, T h) {
  const FN(cplx) m3 = v[-3 * s], m2 = v[-2 * s], m1 = v[-s], c = v[0], p1 = v[s],
                 p2 = v[2 * s], p3 = v[3 * s];
  const T den = (T)180 * (h * h);
  FN(cplx) r;
  r.re = ((((((T)2 * m3.re - (T)27 * m2.re) + (T)270 * m1.re) - (T)490 * c.re) +
           (T)270 * p1.re) - (T)27 * p2.re) + (T)2 * p3.re;
  r.im = ((((((T)2 * m3.im - (T)27 * m2.im) + (T)270 * m1.im) - (T)490 * c.im) +
           (T)270 * p1.im) - (T)27 * p2.im) + (T)2 * p3.im;
  return FN(cdiv_real)(r, den);
}

/* cmplx_first_deriv, src/deriv.f90:12-18. */
static inline FN(cplx) FN(d1)(const FN(cplx) *v, long s, T h) {
  const FN(cplx) m3 = v[-3 * s], m2 = v[-2 * s], m1 = v[-s], p1 = v[s], p2 = v[2 * s],
                 p3 = v[3 * s];
  const T den = (T)60 * h;
  FN(cplx) r;
  r.re = ((((-m3.re + (T)9 * m2.re) - (T)45 * m1.re) + (T)45 * p1.re) - (T)9 * p2.re) +
         p3.re;
  r.im = ((((-m3.im + (T)9 * m2.im) - (T)45 * m1.im) + (T)45 * p1.im) - (T)9 * p2.im) +
         p3.im;
  return FN(cdiv_real)(r, den);
}

/* ghost_comm, src/ghost.f90:27-192, on one domain (or the global boundary faces of a
 * rank whose inner faces the caller filled): faces only, over the interior of the other
 * two axes; corners are left as they are.  zero: ghosts 0; solid: mirror about the
 * edge point (ghost 0 = psi 2, -1 = 3, -2 = 4); periodic: the opposite side (ghost 0 =
 * psi n, n+1 = 1, ...).  z follows `boundaries` unless periodicz.  With more than one
 * rank only the global faces are written, which is complete for zero and solid. */
void FN(orc_fill_ghosts)(const orc_params *p, T *psi_pad) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  FN(cplx) *P = (FN(cplx) *)psi_pad;
  const int bc = p->boundaries;
  const int bz = p->periodicz ? 2 : bc;
  const int n[3] = {mx, my, mz};
  const long st[3] = {sx, sy, 1};
  const int lo_edge[3] = {p->coord_x == 0, p->coord_y == 0, 1};
  const int hi_edge[3] = {p->coord_x == p->dims_x - 1, p->coord_y == p->dims_y - 1, 1};
  for (int a = 0; a < 3; ++a) {
    const int b = a == 2 ? bz : bc;
    const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;   /* the two other axes */
    for (int iu = 1; iu <= n[u]; ++iu)
      for (int iv = 1; iv <= n[v]; ++iv) {
        FN(cplx) *row = P + (long)(iu + 2) * st[u] + (long)(iv + 2) * st[v] + 2 * st[a];
        /* row[q * st[a]] is the point with index q (1..n) along axis a */
        for (int g = 1; g <= 3; ++g) {
          FN(cplx) zero = {(T)0, (T)0};
          FN(cplx) lo = zero, hi = zero;
          if (b == 1) {
            lo = row[(1 + g) * st[a]];
            hi = row[(n[a] - g) * st[a]];
          } else if (b == 2) {
            lo = row[(n[a] + 1 - g) * st[a]];
            hi = row[g * st[a]];
          }
          if (lo_edge[a] || a == 2) row[(1 - g) * st[a]] = lo;
          if (hi_edge[a] || a == 2) row[(n[a] + g) * st[a]] = hi;
        }
      }
  }
}

/* get_rhs, src/finite_diff.f90:8-136, for the harmonic potential (:33-34), the
 * rotating frame (:128-130) and the nonlinearity (:132), with zero ghosts
 * (src/ghost.f90:55,88,121,154,179-181: the padded halo is never written).
 * Generalised by two coefficients that equal 1 in the reference:
 *   rhs += (c0 - gnl*abs(psi**2) - pot)*psi
 * (gnl*x == x exactly when gnl == 1, so the reference arithmetic is unchanged). */
void FN(orc_get_rhs)(const orc_params *p, T *psi_pad, T *rhs_out, double t) {
  FN(orc_fill_ghosts)(p, psi_pad);
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sz = 1, sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  const T hx = (T)p->Lx / (T)p->nx, hy = (T)p->Ly / (T)p->ny, hz = (T)p->Lz / (T)p->nz;
  T *xx = malloc(sizeof(T) * mx), *yy = malloc(sizeof(T) * my), *zz = malloc(sizeof(T) * mz);
  FN(coords)(p, xx, yy, zz);
  const FN(cplx) *P = (const FN(cplx) *)psi_pad;
  FN(cplx) *R = (FN(cplx) *)rhs_out;
  const T wx = (T)p->harm_w[0], wy = (T)p->harm_w[1], wz = (T)p->harm_w[2];
  const T Lx = (T)p->Lx, Ly = (T)p->Ly, Lz = (T)p->Lz;
  const T om = (T)p->omega_z, c0 = (T)p->c0, gnl = (T)p->gnl;
  const int rot = (fabs((double)om) > (double)FLT_EPSILON_T) && (t < p->t_spindown);
  const int pot_on = !(t > p->potential_off);
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= mx; ++i) {
    for (int j = 1; j <= my; ++j) {
      for (int k = 1; k <= mz; ++k) {
        const long o = (long)(i + 2) * sx + (long)(j + 2) * sy + (k + 2);
        const FN(cplx) *c = P + o;
        FN(cplx) r = FN(d2)(c, sx, hx);
        FN(cplx) q = FN(d2)(c, sy, hy);
        r.re = r.re + q.re;
        r.im = r.im + q.im;
        if (!p->two_dim) {
          q = FN(d2)(c, sz, hz);
          r.re = r.re + q.re;
          r.im = r.im + q.im;
        }
        T pot = (T)0;
        if (pot_on) {
          const T ax = xx[i - 1] / Lx, ay = yy[j - 1] / Ly, az = zz[k - 1] / Lz;
          pot = wx * (ax * ax) + wy * (ay * ay) + wz * (az * az);
        }
        if (rot) {
          const FN(cplx) dy1 = FN(d1)(c, sy, hy), dx1 = FN(d1)(c, sx, hx);
          const T pr = xx[i - 1] * dy1.re - yy[j - 1] * dx1.re;
          const T pi_ = xx[i - 1] * dy1.im - yy[j - 1] * dx1.im;
          /* (0, om) * (pr, pi_) */
          r.re = r.re + ((T)0 * pr - om * pi_);
          r.im = r.im + ((T)0 * pi_ + om * pr);
        }
        const T a = c->re, b = c->im;
        const T s_re = a * a - b * b, s_im = a * b + b * a; /* psi**2 */
        const T mod = HYPOT(s_re, s_im);                       /* abs(psi**2) */
        const T f = (c0 - gnl * mod) - pot;
        r.re = r.re + f * a;
        r.im = r.im + f * b;
        R[((long)(i - 1) * my + (j - 1)) * mz + (k - 1)] = r;
      }
    }
  }
  free(xx);
  free(yy);
  free(zz);
}

/* Interior update psi += a*w for stage coefficient a (division form for 1/3). */
static void FN(psi_axpy)(const orc_params *p, T *psi_pad, const T *w_, int stage) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  FN(cplx) *P = (FN(cplx) *)psi_pad;
  const FN(cplx) *W = (const FN(cplx) *)w_;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < mx; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 0; k < mz; ++k) {
        FN(cplx) *c = P + (long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3);
        const FN(cplx) wv = W[((long)i * my + j) * mz + k];
        if (stage == 0) {        /* Psi + (w/3.)  src/run.f90:48 */
          const FN(cplx) q = FN(cdiv_real)(wv, (T)3);
          c->re = c->re + q.re;
          c->im = c->im + q.im;
        } else if (stage == 1) { /* Psi + w       src/run.f90:51 */
          c->re = c->re + wv.re;
          c->im = c->im + wv.im;
        } else {                 /* Psi + 0.5*w   src/run.f90:54 */
          c->re = c->re + (T)0.5 * wv.re;
          c->im = c->im + (T)0.5 * wv.im;
        }
      }
}

/* One low-storage RK3 step, src/run.f90:46-55.  dt is the complex step after the
 * dissipation factor dt*(1-gamma*i) (src/run.f90:42).  rhs is scratch.  Returns t+Re(dt). */
double FN(orc_rk3_step)(const orc_params *p, T *psi_pad, T *w_, T *rhs_, double t) {
  const long n = (long)(p->nx / p->dims_x) * (p->ny / p->dims_y) * p->nz;
  const T dr = (T)p->dt_re, di = (T)p->dt_im;
  /* eye*0.5*dt = ((0,1)*0.5)*dt = (0,0.5)*(dr,di) */
  const T cr = (T)0 * dr - (T)0.5 * di, ci = (T)0 * di + (T)0.5 * dr;
  const T m23 = (T)(-2.0 / 3.0);
  FN(cplx) *W = (FN(cplx) *)w_;
  const FN(cplx) *R = (const FN(cplx) *)rhs_;
  for (int stage = 0; stage < 3; ++stage) {
    FN(orc_get_rhs)(p, psi_pad, rhs_, t);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < n; ++q) {
      const T crr = cr * R[q].re - ci * R[q].im, cri = cr * R[q].im + ci * R[q].re;
      if (stage == 0) {          /* w = eye*0.5*dt*rhs */
        W[q].re = crr;
        W[q].im = cri;
      } else if (stage == 1) {   /* w = (-2./3.)*w + eye*0.5*dt*rhs */
        W[q].re = m23 * W[q].re + crr;
        W[q].im = m23 * W[q].im + cri;
      } else {                   /* w = -w + eye*0.5*dt*rhs */
        W[q].re = -W[q].re + crr;
        W[q].im = -W[q].im + cri;
      }
    }
    FN(psi_axpy)(p, psi_pad, w_, stage);
  }
  return (double)((T)t + dr);
}

/* Reference imaginary-time step, src/run.f90:21-22: Psi = Psi + dt*0.5*rhs. */
void FN(orc_euler_step)(const orc_params *p, T *psi_pad, T *rhs_, double t) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  const T dr = (T)p->dt_re * (T)0.5, di = (T)p->dt_im * (T)0.5;
  FN(orc_get_rhs)(p, psi_pad, rhs_, t);
  FN(cplx) *P = (FN(cplx) *)psi_pad;
  const FN(cplx) *R = (const FN(cplx) *)rhs_;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < mx; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 0; k < mz; ++k) {
        FN(cplx) *c = P + (long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3);
        const FN(cplx) r = R[((long)i * my + j) * mz + k];
        c->re = c->re + (dr * r.re - di * r.im);
        c->im = c->im + (dr * r.im + di * r.re);
      }
}

/* Psi_info / Psi_norm, src/diagnostic.f90:7-25 (single domain): max|psi|, min|psi|,
 * sum|psi|/(Nx*Ny*Nz), sum|psi|**2, accumulated here in long double, memory order. */
void FN(orc_psi_info)(const orc_params *p, const T *psi_pad, double out[4]) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  const FN(cplx) *P = (const FN(cplx) *)psi_pad;
  double mxv = -1.0, mnv = HUGE_VAL;
  long double s1 = 0, s2 = 0;
  for (int i = 0; i < mx; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 0; k < mz; ++k) {
        const FN(cplx) c = P[(long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3)];
        const double a = (double)HYPOT(c.re, c.im);
        if (a > mxv) mxv = a;
        if (a < mnv) mnv = a;
        s1 += a;
        s2 += (long double)a * a;
      }
  out[0] = mxv;
  out[1] = mnv;
  out[2] = (double)(s1 / ((long double)p->nx * p->ny * p->nz));
  out[3] = (double)s2;
}

/* get_phase, src/diagnostic.f90:27-46: phase = atan2(Im psi, Re psi) over the whole
 * padded grid (zero ghosts give phase 0), then vel(k,j,i,a) = first_deriv of the phase
 * along axis a (src/deriv.f90:5-11: 6th order, real arithmetic, no unwrapping, so a
 * 2*pi jump in the phase shows up as a large velocity).  vel_out: 3*mx*my*mz values,
 * component slowest, then x, y, z fastest (the reference's vel(k,j,i,a) order). */
void FN(orc_velocity)(const orc_params *p, T *psi_pad, T *vel_out) {
  FN(orc_fill_ghosts)(p, psi_pad);
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6), np = (long)(mx + 6) * sx;
  const T hx = (T)p->Lx / (T)p->nx, hy = (T)p->Ly / (T)p->ny, hz = (T)p->Lz / (T)p->nz;
  const FN(cplx) *P = (const FN(cplx) *)psi_pad;
  T *ph = (T *)malloc(sizeof(T) * np);
  for (long q = 0; q < np; ++q) ph[q] = ATAN2(P[q].im, P[q].re);
  const long n = (long)mx * my * mz;
  const long st[3] = {sx, sy, 1};
  const T h[3] = {hx, hy, hz};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < mx; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 0; k < mz; ++k) {
        const T *c = ph + (long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3);
        const long o = ((long)i * my + j) * mz + k;
        for (int a = 0; a < 3; ++a) {
          const long s = st[a];
          const T v = ((((-c[-3 * s] + (T)9 * c[-2 * s]) - (T)45 * c[-s]) + (T)45 * c[s]) -
                       (T)9 * c[2 * s]) + c[3 * s];
          vel_out[a * n + o] = v / ((T)60 * h[a]);
        }
      }
  free(ph);
}

/* get_qpressure, src/diagnostic.f90:50-83: density = abs(psi) over the padded grid,
 * laplace_rho on the box extended by one cell, 4th order (real_second_deriv,
 * src/deriv.f90:26-32, summed x + y + z), then qpressure(k,j,i,a) =
 * low_first_deriv of laplace_rho along a (src/deriv.f90:33-39, 2nd order).  Output
 * layout as orc_velocity. */
void FN(orc_qpressure)(const orc_params *p, T *psi_pad, T *qp_out) {
  FN(orc_fill_ghosts)(p, psi_pad);
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6), np = (long)(mx + 6) * sx;
  const T hx = (T)p->Lx / (T)p->nx, hy = (T)p->Ly / (T)p->ny, hz = (T)p->Lz / (T)p->nz;
  const FN(cplx) *P = (const FN(cplx) *)psi_pad;
  T *dn = (T *)malloc(sizeof(T) * np);
  T *lp = (T *)calloc((size_t)np, sizeof(T));
  for (long q = 0; q < np; ++q) dn[q] = HYPOT(P[q].re, P[q].im);
  const long st[3] = {sx, sy, 1};
  const T h[3] = {hx, hy, hz};
#pragma omp parallel for schedule(static)
  for (int i = -1; i <= mx; ++i)
    for (int j = -1; j <= my; ++j)
      for (int k = -1; k <= mz; ++k) {
        const long o = (long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3);
        const T *c = dn + o;
        T d[3];
        for (int a = 0; a < 3; ++a) {
          const long s = st[a];
          const T v = ((((-c[-2 * s] + (T)16 * c[-s]) - (T)30 * c[0]) + (T)16 * c[s]) -
                       c[2 * s]);
          d[a] = v / ((T)12 * (h[a] * h[a]));
        }
        lp[o] = (d[0] + d[1]) + d[2];
      }
  const long n = (long)mx * my * mz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < mx; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 0; k < mz; ++k) {
        const T *c = lp + (long)(i + 3) * sx + (long)(j + 3) * sy + (k + 3);
        const long o = ((long)i * my + j) * mz + k;
        for (int a = 0; a < 3; ++a) {
          const long s = st[a];
          qp_out[a * n + o] = (-c[-s] + c[s]) / ((T)2 * h[a]);
        }
      }
  free(dn);
  free(lp);
}

#undef FN
#undef CAT
#undef CAT_
