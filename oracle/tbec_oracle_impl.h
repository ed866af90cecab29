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

/* Time-dependent scalars of the potential select (src/finite_diff.f90:30-124), in the
 * working precision: s = max(0, 5 - 0.1 t) (ring, ring2, barrier openings), b =
 * 5 sin(0.1 t) (box_trap tilt) and the laser centre (las_freq = 0). */
typedef struct { T s, b, las_x, las_y, t; } FN(pot_time);

static FN(pot_time) FN(pot_time_of)(T t) {
  FN(pot_time) q;
  const T m = (T)5 - (T)0.1 * t;
  q.s = (T)0 > m ? (T)0 : m;                              /* max(0., 5.-0.1*t) */
  q.b = (T)5 * SIN((T)0.1 * t);                           /* 5*sin(0.1*t) */
  const T lf = (T)0;                                      /* las_freq=0.0 */
  q.las_x = ((T)70 * COS(lf * t)) * ((T)1 - SIN(lf * t)); /* :120 */
  q.las_y = ((T)30 * SIN(lf * t)) * COS(lf * t);          /* :121 */
  q.t = t;
  return q;
}

/* pot at interior point (i, j, k) (1-based, this block), src/finite_diff.f90:30-124,
 * operand for operand.  Integer powers x**n are n-1 multiplications left to right
 * (x*x, r*r*r*r*r*r*r*r), as amdflang emits them; real powers (harmonic_ring2's
 * pow = 6.) call pow. */
static T FN(pot_at)(const orc_params *p, int i, int j, int k, const T *xx, const T *yy,
                    const T *zz, const FN(pot_time) *q) {
  const T x = xx[i - 1], y = yy[j - 1], z = zz[k - 1];
  const T Lx = (T)p->Lx, Ly = (T)p->Ly, Lz = (T)p->Lz;
  const T wx = (T)p->harm_w[0], wy = (T)p->harm_w[1], wz = (T)p->harm_w[2];
  T pot = (T)0;
  switch (p->potential) {
    case ORC_POT_ZERO:
      break;
    case ORC_POT_HARMONIC: {                                  /* :33-34 */
      const T ax = x / Lx, ay = y / Ly, az = z / Lz;
      pot = wx * (ax * ax) + wy * (ay * ay) + wz * (az * az);
      break;
    }
    case ORC_POT_HARMONIC_RING: {                             /* :35-40 */
      const T ax = x / Lx, ay = y / Lx;
      const T r = SQRT(ax * ax + ay * ay);
      const T d = r - (T)0.3;
      pot = wx * (d * d);
      if (fabs((double)(r - (T)0.3)) < (double)(T)0.01) pot = q->s;
      break;
    }
    case ORC_POT_HARMONIC_RING2: {                            /* :41-65 */
      const T pw = (T)6;
      const T ax = x / Lx;
      const T a1 = POW((T)0.25, pw);
      const T u1 = POW(ax, pw), v1 = POW(y / Ly - (T)0.2, pw), v2 = POW(y / Ly + (T)0.2, pw);
      T r1 = POW((T)fabs((double)((a1 - (T)0.15 * u1) - (T)4 * v1)), (T)1 / pw);
      T r2 = POW((T)fabs((double)((a1 - (T)0.15 * u1) - (T)4 * v2)), (T)1 / pw);
      r1 = r1 < (T)0.2495 ? (T)0 : (T)5;
      r2 = r2 < (T)0.2495 ? (T)0 : (T)5;
      pot = r1 * r2;
      if (pot < (T)0) pot = (T)0;
      if (fabs((double)(y / Ly)) < (double)(T)0.01 && pot < (T)0.01) pot = q->s;
      break;
    }
    case ORC_POT_BARRIER:                                     /* :66-71 */
      pot = (T)0;
      if (fabs((double)(y / Ly)) > (double)(T)0.3) pot = (T)5;
      if (fabs((double)(y / Ly)) < (double)(T)0.01) pot = q->s;
      break;
    case ORC_POT_BOX_TRAP: {                                  /* :72-78 */
      pot = (T)10;
      const T r = SQRT(x * x + y * y);
      if (r < Lx / (T)3 && fabs((double)z) < (double)(Lz / (T)2.5)) pot = (T)0;
      pot = pot + (q->b * z) / Lz;
      break;
    }
    case ORC_POT_CHANNEL: {                                   /* :79-96 */
      pot = (T)0;
      if (fabs((double)y) > (double)(((T)0.75 * Ly) / (T)2) ||
          fabs((double)z) > (double)(((T)0.75 * Lz) / (T)2))
        pot = (T)5;
      const T hxg = Lx / (T)p->nx, hyg = Ly / (T)p->ny;
      for (int l1 = -1; l1 <= 1; ++l1)
        if (fabs((double)x) < (double)hxg &&
            fabs((double)(y - ((T)l1 * Ly) / (T)5)) < (double)hyg)
          pot = (T)7.5;
      for (int l1 = -1; l1 <= 1; ++l1)
        if (fabs((double)x) < (double)hxg &&
            fabs((double)(z - ((T)l1 * Lz) / (T)5)) < (double)hyg)
          pot = (T)7.5;
      break;
    }
    case ORC_POT_QUARTIC: {                                   /* :97-99 */
      const T ax = x / Lx, ay = y / Ly;
      const T r = SQRT(ax * ax + ay * ay);
      /* r**8: amdflang multiplies out left to right (pinned by the reference build) */
      const T r8 = r * r * r * r * r * r * r * r;
      pot = wx * r8;
      break;
    }
    case ORC_POT_SPHERE: {                                    /* :100-111 */
      const T theta = ATAN2(y, x);
      const T ax = x / Lx, ay = y / Ly, az = z / Lz;
      const T r1 = SQRT((ax * ax + ay * ay) + az * az);
      const T phi = ACOS(az / r1);
      const T lim = ((T)0.2 + (T)0.01 * SIN((T)20 * phi)) +
                    (T)0.01 * COS((T)30 * theta + (T)0.03);
      pot = r1 < lim ? (T)10 : (T)0;
      break;
    }
    case ORC_POT_MOBIUS:                                      /* :112-116 */
    case ORC_POT_TABLE: {
      const int my = p->ny / p->dims_y, mz = p->nz;
      pot = ((const T *)p->trap)[((long)(i - 1) * my + (j - 1)) * mz + (k - 1)];
      if (p->potential == ORC_POT_MOBIUS && fabs((double)y) < 1.0 &&
          fabs((double)(x - (T)25)) < 1.0)
        pot = pot + (T)10;
      break;
    }
    default:
      break;
  }
  (void)wy;
  (void)wz;
  if (p->lbeam) {                                             /* :118-123 */
    const T wb = (T)p->wbeam;
    const T u = (x - q->las_x) / wb, v = (y - q->las_y) / wb;
    pot = pot + (T)20 * EXP(-(u * u) - v * v);
  }
  if (q->t > (T)p->potential_off) pot = (T)0;                 /* :124 */
  return pot;
}

/* The potential over this block's interior at time t (interior layout, z fastest). */
void FN(orc_potential)(const orc_params *p, double t, T *out) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  T *xx = malloc(sizeof(T) * mx), *yy = malloc(sizeof(T) * my), *zz = malloc(sizeof(T) * mz);
  FN(coords)(p, xx, yy, zz);
  const FN(pot_time) q = FN(pot_time_of)((T)t);
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= mx; ++i)
    for (int j = 1; j <= my; ++j)
      for (int k = 1; k <= mz; ++k)
        out[((long)(i - 1) * my + (j - 1)) * mz + (k - 1)] =
            FN(pot_at)(p, i, j, k, xx, yy, zz, &q);
  free(xx);
  free(yy);
  free(zz);
}

/* The mobius Trap table, src/initial.f90:160-196: 200 x 200 points of the strip
 * ((25 + 5 r cos(th/2)) cos th, (25 + 5 r cos(th/2)) sin th, 5 r sin(th/2)),
 * Trap = sum of exp(-1.5 |x - m|^2) over the strip in (ii1, ii2) order, capped at 3
 * after every term, then Trap = max(Trap) - Trap.  This block's max only (one block). */
void FN(orc_mobius_trap)(const orc_params *p, T *trap) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const T pi = (T)3.14159265359;   /* src/cdata.f90:13 */
  T *m = malloc(sizeof(T) * 200 * 200 * 3);
  for (int i = 1; i <= 200; ++i)
    for (int j = 1; j <= 200; ++j) {
      const T r = ((T)(i - 1) * (T)2) / (T)200 - (T)1;
      const T th = (((T)(j - 1) * (T)2) * pi) / (T)200;
      T *e = m + ((long)(i - 1) * 200 + (j - 1)) * 3;
      e[0] = ((T)25 + ((T)5 * r) * COS(th / (T)2)) * COS(th);
      e[1] = ((T)25 + ((T)5 * r) * COS(th / (T)2)) * SIN(th);
      e[2] = ((T)5 * r) * SIN(th / (T)2);
    }
  T *xx = malloc(sizeof(T) * mx), *yy = malloc(sizeof(T) * my), *zz = malloc(sizeof(T) * mz);
  FN(coords)(p, xx, yy, zz);
#pragma omp parallel for schedule(dynamic)
  for (int i = 1; i <= mx; ++i)
    for (int j = 1; j <= my; ++j)
      for (int k = 1; k <= mz; ++k) {
        T acc = (T)0;
        for (int a = 0; a < 200 * 200; ++a) {
          const T *e = m + (long)a * 3;
          const T dx = xx[i - 1] - e[0], dy = yy[j - 1] - e[1], dz = zz[k - 1] - e[2];
          acc = acc + EXP(-((T)1.5 * (dx * dx)) - (T)1.5 * (dy * dy) - (T)1.5 * (dz * dz));
          if (acc > (T)3) acc = (T)3;
        }
        trap[((long)(i - 1) * my + (j - 1)) * mz + (k - 1)] = acc;
      }
  const long n = (long)mx * my * mz;
  T mxv = trap[0];
  for (long q = 1; q < n; ++q) mxv = trap[q] > mxv ? trap[q] : mxv;
  for (long q = 0; q < n; ++q) trap[q] = mxv - trap[q];
  free(m);
  free(xx);
  free(yy);
  free(zz);
}

/* The vortex imprint of force_random_points at a given centre and sense,
 * src/force.f90:32-53: psi *= pade(r) exp(i dir theta) about (x0, y0), with
 * pade(r) = sqrt(r^2 (c1 + c2 r^2) / (1 + c3 r^2 + c2 r^4)), c1 = 11/32. */
void FN(orc_imprint_vortex)(const orc_params *p, T *psi_pad, double x0, double y0,
                            double dir) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  T *xx = malloc(sizeof(T) * mx), *yy = malloc(sizeof(T) * my), *zz = malloc(sizeof(T) * mz);
  FN(coords)(p, xx, yy, zz);
  const T vx = (T)x0, vy = (T)y0, d = (T)dir;
  const T c1 = (T)11 / (T)32;
  const T c3 = ((T)5 - (T)32 * c1) / ((T)48 - (T)192 * c1);
  const T c2 = c1 * (c3 - (T)0.25);
  FN(cplx) *P = (FN(cplx) *)psi_pad;
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= mx; ++i)
    for (int j = 1; j <= my; ++j) {
      const T ddx = xx[i - 1] - vx, ddy = yy[j - 1] - vy;
      const T r = SQRT(ddx * ddx + ddy * ddy);
      const T theta = ATAN2(ddy, ddx);
      const T r2 = r * r, r4 = r * r * r * r;
      const T func = SQRT((r2 * (c1 + c2 * r2)) / (((T)1 + c3 * r2) + c2 * r4));
      const T ang = d * theta;
      const T fr = func * COS(ang), fi = func * SIN(ang);
      for (int k = 1; k <= mz; ++k) {
        FN(cplx) *c = P + (long)(i + 2) * sx + (long)(j + 2) * sy + (k + 2);
        const T a = c->re, b = c->im;
        c->re = a * fr - b * fi;
        c->im = a * fi + b * fr;
      }
    }
  free(xx);
  free(yy);
  free(zz);
}

/* get_rhs, src/finite_diff.f90:8-136: the Laplacian, the potential select (:30-124,
 * pot_at), the advective (:126-127) or rotating-frame (:128-130) term and the
 * nonlinearity (:132), after ghost_comm (src/ghost.f90:27-192, orc_fill_ghosts).
 * Generalised by two coefficients that equal 1 in the reference:
 *   rhs += (c0 - gnl*abs(psi**2) - pot)*psi
 * (gnl*x == x exactly when gnl == 1, so the reference arithmetic is unchanged).  t, the
 * switch-off times and epsilon(0.) are compared in the working precision. */
void FN(orc_get_rhs)(const orc_params *p, T *psi_pad, T *rhs_out, double t) {
  FN(orc_fill_ghosts)(p, psi_pad);
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sz = 1, sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  const T hx = (T)p->Lx / (T)p->nx, hy = (T)p->Ly / (T)p->ny, hz = (T)p->Lz / (T)p->nz;
  T *xx = malloc(sizeof(T) * mx), *yy = malloc(sizeof(T) * my), *zz = malloc(sizeof(T) * mz);
  FN(coords)(p, xx, yy, zz);
  const FN(cplx) *P = (const FN(cplx) *)psi_pad;
  FN(cplx) *R = (FN(cplx) *)rhs_out;
  const T om = (T)p->omega_z, c0 = (T)p->c0, gnl = (T)p->gnl;
  const T tt = (T)t;
  const T vx = (T)p->advective_vx;
  const int adv = fabs((double)vx) > (double)FLT_EPSILON_T;
  const int rot = !adv && (fabs((double)om) > (double)FLT_EPSILON_T) &&
                  (tt < (T)p->t_spindown);
  const FN(pot_time) ptime = FN(pot_time_of)(tt);
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
        const T pot = FN(pot_at)(p, i, j, k, xx, yy, zz, &ptime);
        if (adv) {
          /* 2*eye*advective_vx*D1x: (0, 2 vx) * (d.re, d.im) */
          const FN(cplx) dx1 = FN(d1)(c, sx, hx);
          const T v2 = (T)2 * vx;
          r.re = r.re + ((T)0 * dx1.re - v2 * dx1.im);
          r.im = r.im + ((T)0 * dx1.im + v2 * dx1.re);
        } else if (rot) {
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


/* ---- initial states and insertions (src/initial.f90:58-156,202-626,637-870) ------
 * kind / what numbers follow the C ABI (TBEC_IC_*, TBEC_INSERT_*).  prm carries the
 * values the reference draws at run time or forms once (T values held in doubles):
 *   random_phase : 20 x (wave1, wave2, wave3, amp1, amp2)   (setup_quench's draws)
 *   random_points: 20 x (x0, y0, dir)
 *   vortex_rings : count x (c11 c12 c13 c21 c22 c23 c31 c32 c33, tx, ty, tz): the
 *                  rotation coefficients as the reference groups them, and translations
 *   bundle       : 21 x (p1, p2): the three bundles' centres
 *   soliton      : (Re, Im) of exp(eye*pi)
 * Integer powers are multiplied out left to right; x**1.15 calls pow. */

static FN(cplx) FN(cmul)(FN(cplx) a, FN(cplx) b) {
  FN(cplx) r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

/* func*exp(s*eye*theta) of a straight line along the third axis at (d1, d2) from it:
 * func = 1 - exp(-c*r**1.15), theta = atan2(d2, d1). */
static FN(cplx) FN(line_factor)(T d1, T d2, T c, T s) {
  const T r = SQRT(d1 * d1 + d2 * d2);
  const T theta = ATAN2(d2, d1);
  const T func = (T)1 - EXP(-(c * POW(r, (T)1.15)));
  const T a = s * theta;
  FN(cplx) f;
  f.re = func * COS(a);
  f.im = func * SIN(a);
  return f;
}

/* ring_density (src/initial.f90:628-635) */
static T FN(ring_density)(T r) {
  const T c1 = (T)11 / (T)32;
  const T c3 = ((T)5 - (T)32 * c1) / ((T)48 - (T)192 * c1);
  const T c2 = c1 * (c3 - (T)0.25);
  return SQRT(((r * r) * (c1 + c2 * (r * r))) / (((T)1 + c3 * (r * r)) + c2 * (r * r * r * r)));
}

/* ring_density(r)*exp(eye*theta) of a ring pair term: axial offset a, radial distance q
 * from the ring axis, sign sg = +1 (q + R) or -1 (q - R). */
static FN(cplx) FN(ring_term)(T a, T q, T R, int sg) {
  const T qq = sg > 0 ? q + R : q - R;
  const T theta = ATAN2(a, qq);
  const T r = SQRT(a * a + qq * qq);
  const T d = FN(ring_density)(r);
  FN(cplx) f;
  f.re = d * COS(theta);
  f.im = d * SIN(theta);
  return f;
}

static FN(cplx) FN(conj)(FN(cplx) a) { a.im = -a.im; return a; }

/* setup_thomas_fermi's TF_dense before the 1/sqrt(2) (src/initial.f90:637-652) */
static T FN(tf_dense)(const orc_params *p, T x, T y, T z) {
  const T Lx = (T)p->Lx, Ly = (T)p->Ly, Lz = (T)p->Lz;
  const T ax = x / Lx, ay = y / Ly, az = z / Lz;
  return (((T)1 - (T)p->harm_w[0] * (ax * ax)) - (T)p->harm_w[1] * (ay * ay)) -
         (T)p->harm_w[2] * (az * az);
}

static FN(cplx) FN(state_at)(const orc_params *p, int kind, int winding, T R,
                             const double *prm, int nprm, T x, T y, T z) {
  const T Lx = (T)p->Lx, Ly = (T)p->Ly, Lz = (T)p->Lz;
  const T pi = (T)3.14159265359;
  FN(cplx) psi = {(T)0, (T)0};
  switch (kind) {
    case 0: {                                          /* constant :59-66 */
      const T c = (T)1 / SQRT((T)2);
      psi.re = c; psi.im = c;
      break;
    }
    case 1: {                                          /* soliton :67-80 */
      T tf = FN(tf_dense)(p, x, y, z) / SQRT((T)2);
      if (tf < (T)0) tf = (T)0;
      const T ld = HYPOT(tf, tf);
      if (x < (T)0) { psi.re = ld; psi.im = (T)0 * ld; }
      else { psi.re = ld * (T)prm[0]; psi.im = ld * (T)prm[1]; }
      break;
    }
    case 2: {                                          /* random_phase: setup_quench :346-375 */
      for (int q = 0; q < 20; ++q) {
        const double *w = prm + 5 * q;
        T s = x * (T)w[0] + y * (T)w[1];
        if (!p->two_dim) s = s + z * (T)w[2];
        const T arg = (((s * (T)2) * pi) * Lx) / (Lx * Lx);   /* eye*s*2*pi/Lx */
        FN(cplx) e = {COS(arg), SIN(arg)};
        FN(cplx) amp = {(T)w[3], (T)w[4]};
        const FN(cplx) t = FN(cmul)(amp, e);
        psi.re = psi.re + t.re;
        psi.im = psi.im + t.im;
      }
      T tf = FN(tf_dense)(p, x, y, z);                   /* multiply_thomas_fermi :654-667 */
      if (tf < (T)0) tf = (T)0;
      psi.re = psi.re * tf;
      psi.im = psi.im * tf;
      break;
    }
    case 3: {                                          /* crow :81-108 */
      const T v2 = -Ly / (T)4;
      const T v1 = (T)(-4) + COS((pi * z) / Lz);
      psi = FN(line_factor)(x - v1, y - v2, (T)0.7, (T)1);
      const T w1 = (T)4 - COS((pi * z) / Lz);
      psi = FN(cmul)(psi, FN(line_factor)(x - w1, y - v2, (T)0.7, (T)(-1)));
      break;
    }
    case 4:                                            /* single_vortex :109-121 */
      psi = FN(line_factor)(x - (T)0, y - (T)0, (T)0.7, (T)winding);
      break;
    case 5: {                                          /* sausage :122-137 */
      const T c = fabs((double)z) < 4.0 ? (T)0.1 : (T)0.7;
      psi = FN(line_factor)(x - (T)0, y - (T)0, c, (T)1);
      break;
    }
    case 6: {                                          /* random_points :268-302 */
      psi.re = (T)0.7071; psi.im = (T)0.7071;
      for (int q = 0; q < 20; ++q) {
        const double *v = prm + 3 * q;
        psi = FN(cmul)(psi, FN(line_factor)(x - (T)v[0], y - (T)v[1], (T)0.7, (T)v[2]));
      }
      break;
    }
    case 7: {                                          /* leap_frog :304-344 */
      psi.re = (T)0.7071; psi.im = (T)0.7071;
      const T sh = (T)3, s1 = (T)3, s2 = (T)12;
      const T px[4] = {-sh, sh, -sh, sh}, py[4] = {s1, s2, -s1, -s2};
      const T dir[4] = {(T)1, (T)1, (T)(-1), (T)(-1)};
      for (int q = 0; q < 4; ++q)
        psi = FN(cmul)(psi, FN(line_factor)(x - px[q], y - py[q], (T)0.7, dir[q]));
      break;
    }
    case 8: {                                          /* vortex_rings :202-266 */
      psi.re = (T)1; psi.im = (T)0;
      for (int q = 0; q < nprm / 12; ++q) {
        const double *c = prm + 12 * q;
        const T tx = ((T)c[0] * x + (T)c[1] * y) + (T)c[2] * z;
        const T ty = ((T)c[3] * x + (T)c[4] * y) + (T)c[5] * z;
        const T tz = ((T)c[6] * x - (T)c[7] * y) + (T)c[8] * z;
        const T dx = tx - (T)c[9], dy = ty - (T)c[10], dz = tz - (T)c[11];
        const T qd = SQRT(dy * dy + dz * dz);
        const T theta = ATAN2(dz, dy);
        FN(cplx) p12[2];
        for (int sgi = 0; sgi < 2; ++sgi) {
          const T qq = sgi == 0 ? qd + R : qd - R;
          const T r = SQRT(dx * dx + qq * qq);
          const T d = SQRT(((r * r) * ((T)0.347 + (T)0.0286 * (r * r))) /
                           (((T)1 + (T)0.3333 * (r * r)) + (T)0.0286 * (r * r * r * r)));
          FN(cplx) e = {COS(theta), SIN(theta)};
          FN(cplx) dd = {d, (T)0};
          p12[sgi] = FN(cmul)(dd, e);
        }
        psi = FN(cmul)(psi, FN(cmul)(p12[0], FN(conj)(p12[1])));
      }
      break;
    }
    case 9: {                                          /* line_of_lines :541-585 */
      const T c = (T)1 / SQRT((T)2);
      psi.re = c; psi.im = c;
      const T a = x - (T)0, qd = SQRT(y * y + z * z);
      psi = FN(cmul)(psi, FN(ring_term)(a, qd, R, 1));
      psi = FN(cmul)(psi, FN(conj)(FN(ring_term)(a, qd, R, -1)));
      break;
    }
    case 10: {                                         /* linked_rings :587-626 */
      const T t1 = R * (T)0;
      const T a = x - (T)0, q1 = SQRT(y * y + (z - t1) * (z - t1));
      psi = FN(cmul)(FN(ring_term)(a, q1, R, 1), FN(conj)(FN(ring_term)(a, q1, R, -1)));
      const T t2 = R * (T)1.2;
      const T b = y - (T)0, q2 = SQRT(x * x + (z - t2) * (z - t2));
      psi = FN(cmul)(psi, FN(ring_term)(b, q2, R, 1));
      psi = FN(cmul)(psi, FN(conj)(FN(ring_term)(b, q2, R, -1)));
      break;
    }
    case 11:                                           /* lattice :377-420 (ramp = 0) */
      psi.re = (T)1; psi.im = (T)1;
      for (int v1 = 1; v1 <= 10; ++v1)
        for (int v2 = 1; v2 <= 10; ++v2) {
          const T p1 = (T)v1 * (Lx / (T)16) - Lx / (T)3;
          const T p2 = (T)v2 * (Ly / (T)16) - Ly / (T)3;
          psi = FN(cmul)(psi, FN(line_factor)(x - p1, y - p2, (T)0.7, (T)1));
        }
      break;
    case 12: {                                         /* lattice2 :422-464 */
      T tf = FN(tf_dense)(p, x, y, z) / SQRT((T)2);
      if (tf < (T)0) tf = (T)0;
      psi.re = tf; psi.im = tf;
      for (int v1 = 1; v1 <= 3; ++v1)
        for (int v2 = 1; v2 <= 3; ++v2) {
          const T p1 = (T)v1 * (Lx / (T)4) - Lx / (T)2;
          const T p2 = (T)v2 * (Ly / (T)4) - Ly / (T)2;
          psi = FN(cmul)(psi, FN(line_factor)(x - p1, y - p2, (T)0.7, (T)1));
        }
      for (int v1 = 1; v1 <= 3; ++v1)
        for (int v2 = 1; v2 <= 3; ++v2) {
          const T p1 = ((T)v1 * (Lx / (T)4) + (T)10) - Lx / (T)2;
          const T p2 = ((T)v2 * (Lz / (T)4) + (T)10) - Lz / (T)2;
          psi = FN(cmul)(psi, FN(line_factor)(x - p1, z - p2, (T)0.7, (T)1));
        }
      break;
    }
    case 13: {                                         /* bundle :466-539 */
      T tf = FN(tf_dense)(p, x, y, z) / SQRT((T)2);
      if (tf < (T)0) tf = (T)0;
      psi.re = tf; psi.im = tf;
      for (int q = 0; q < 7; ++q)
        psi = FN(cmul)(psi, FN(line_factor)(x - (T)prm[2 * q], y - (T)prm[2 * q + 1],
                                            (T)0.7, (T)1));
      for (int q = 7; q < 14; ++q)
        psi = FN(cmul)(psi, FN(line_factor)(x - (T)prm[2 * q], z - (T)prm[2 * q + 1],
                                            (T)0.7, (T)1));
      for (int q = 14; q < 21; ++q)
        psi = FN(cmul)(psi, FN(line_factor)(y - (T)prm[2 * q], z - (T)prm[2 * q + 1],
                                            (T)0.7, (T)1));
      break;
    }
    default:
      break;
  }
  return psi;
}

/* One of the reference's init_cond states over this block's interior. */
void FN(orc_reference_state)(const orc_params *p, int kind, int winding, double ring_rad,
                             const double *prm, int nprm, T *psi_pad) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  T *xx = malloc(sizeof(T) * mx), *yy = malloc(sizeof(T) * my), *zz = malloc(sizeof(T) * mz);
  FN(coords)(p, xx, yy, zz);
  FN(cplx) *P = (FN(cplx) *)psi_pad;
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= mx; ++i)
    for (int j = 1; j <= my; ++j)
      for (int k = 1; k <= mz; ++k)
        P[(long)(i + 2) * sx + (long)(j + 2) * sy + (k + 2)] =
            FN(state_at)(p, kind, winding, (T)ring_rad, prm, nprm, xx[i - 1], yy[j - 1],
                         zz[k - 1]);
  free(xx);
  free(yy);
  free(zz);
}

/* The insertions applied to an existing state (src/run.f90:26,37-40): what 0
 * setup_pcurrent (:669-739, by potential kind), 1 setup_add_ring (:759-794, option 1),
 * 2 setup_add_dipole (:839-870), 3 phase_imprint (:820-836), 4 add_noise (:741-757, prm
 * holds 2 uniform draws per point in [0, 1), point-major). */
void FN(orc_reference_insert)(const orc_params *p, int what, int winding, double ring_rad,
                              const double *prm, T *psi_pad) {
  const int mx = p->nx / p->dims_x, my = p->ny / p->dims_y, mz = p->nz;
  const long sy = mz + 6, sx = (long)(my + 6) * (mz + 6);
  T *xx = malloc(sizeof(T) * mx), *yy = malloc(sizeof(T) * my), *zz = malloc(sizeof(T) * mz);
  FN(coords)(p, xx, yy, zz);
  FN(cplx) *P = (FN(cplx) *)psi_pad;
  const T Lx = (T)p->Lx, Ly = (T)p->Ly, R = (T)ring_rad;
  const T pi = (T)3.14159265359;
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= mx; ++i)
    for (int j = 1; j <= my; ++j)
      for (int k = 1; k <= mz; ++k) {
        FN(cplx) *c = P + (long)(i + 2) * sx + (long)(j + 2) * sy + (k + 2);
        const T x = xx[i - 1], y = yy[j - 1], z = zz[k - 1];
        if (what == 0 || what == 3) {
          T theta = (T)0, n = (T)1;
          int set = 1;
          if (what == 3) {
            theta = ATAN2(y - (T)0, x - (T)0);
            n = (T)winding;
          } else {
            switch (p->potential) {
              case ORC_POT_BARRIER:
                theta = y < (T)0 ? ((x * (T)50) * pi) / Lx
                                 : (T)20 * pi - ((x * (T)50) * pi) / Lx;
                break;
              case ORC_POT_MOBIUS:
                theta = (T)20 * ATAN2(y, x);
                break;
              case ORC_POT_HARMONIC_RING: {
                const T ax = x / Lx, ay = y / Lx;
                const T r = SQRT(ax * ax + ay * ay);
                theta = r < (T)0.3 ? ATAN2(y, x) : ATAN2(-y, x);
                n = (T)10;
                break;
              }
              case ORC_POT_HARMONIC_RING2:
                theta = y < (T)0 ? ATAN2((T)0.5 * (y + Ly / (T)5), x)
                                 : ATAN2((T)0.5 * (y - Ly / (T)5), x) - pi;
                n = (T)20;
                break;
              case ORC_POT_CHANNEL:
                theta = (((T)20 * pi) * (x + (T)0.5 * Lx)) / Lx;
                n = (T)2;
                break;
              default:
                set = 0;
            }
          }
          if (set) {
            const T ld = HYPOT(c->re, c->im);
            const T a = n * theta;
            c->re = ld * COS(a);
            c->im = ld * SIN(a);
          }
        } else if (what == 1) {
          const T tx = (T)(-0.25) * Lx;
          const T a = x - tx, qd = SQRT(y * y + (z - (T)0) * (z - (T)0));
          FN(cplx) v = FN(cmul)(*c, FN(ring_term)(a, qd, R, 1));
          *c = FN(cmul)(v, FN(conj)(FN(ring_term)(a, qd, R, -1)));
        } else if (what == 2) {
          *c = FN(cmul)(*c, FN(line_factor)(x - (T)(-20), y - (T)5, (T)0.7, (T)1));
          *c = FN(cmul)(*c, FN(line_factor)(x - (T)(-20), y - (T)(-5), (T)0.7, (T)(-1)));
        } else if (what == 4) {
          const long q = ((long)(i - 1) * my + (j - 1)) * mz + (k - 1);
          const T e1 = (T)prm[2 * q] * (T)2 - (T)1, e2 = (T)prm[2 * q + 1] * (T)2 - (T)1;
          c->re = c->re + (T)0.01 * e1;
          c->im = c->im + (T)0.01 * e2;
        }
      }
  free(xx);
  free(yy);
  free(zz);
}

#undef FN
#undef CAT
#undef CAT_
