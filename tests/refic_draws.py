"""The values libtbec draws or forms once for the reference's initial states, replicated
in numpy for the oracle (test infrastructure): the counter-based uniform of
trapbec_amd.initial mapped by runif(a, b) = a + u*(b - a) in the working precision
(src/statistics.f90:29-35), and cos / sin from the C library (as the reference's runtime
and libtbec's host code call them)."""
import ctypes
import math

import numpy as np

from trapbec_amd import initial

_LIBM = ctypes.CDLL("libm.so.6")
for _f in ("cosf", "sinf", "sqrtf"):
    getattr(_LIBM, _f).restype = ctypes.c_float
    getattr(_LIBM, _f).argtypes = [ctypes.c_float]


def tdt(prec):
    return np.float32 if prec == 32 else np.float64


def tcos(v, prec):
    return tdt(prec)(_LIBM.cosf(float(v)) if prec == 32 else math.cos(float(v)))


def tsin(v, prec):
    return tdt(prec)(_LIBM.sinf(float(v)) if prec == 32 else math.sin(float(v)))


def runif(seed, stream, idx, lo, hi, prec):
    T = tdt(prec)
    u = T(initial.uniform(seed, stream, np.array([idx]))[0])
    return T(lo) + u * (T(hi) - T(lo))


PI = 3.14159265359   # src/cdata.f90:13


def random_phase(seed, prec):
    """setup_quench's 20 x (wave1, wave2, wave3, amp1, amp2), stream 201."""
    T = tdt(prec)
    out = []
    for q in range(20):
        for w in range(3):
            out.append(float(np.floor(runif(seed, 201, 5 * q + w, T(-20), T(20), prec))))
        for w in range(3, 5):
            out.append(float(runif(seed, 201, 5 * q + w, T(-0.1), T(0.1), prec)))
    return np.array(out)


def random_points(seed, Lx, Ly, prec):
    """setup_random_points' 20 x (x0, y0, dir), stream 202."""
    T = tdt(prec)
    Lx, Ly = T(Lx), T(Ly)
    out = []
    for q in range(20):
        out.append(float(runif(seed, 202, 3 * q, -Lx / T(5), Lx / T(5), prec)))
        out.append(float(runif(seed, 202, 3 * q + 1, -Ly / T(5), Ly / T(5), prec)))
        d = runif(seed, 202, 3 * q + 2, T(0.9), T(1), prec)
        out.append(-1.0 if d < T(0.5) else 1.0)
    return np.array(out)


def vortex_rings(seed, count, Lx, prec):
    """setup_vortex_rings' rotation coefficients and translations per ring, stream 203."""
    T = tdt(prec)
    pi = T(PI)
    out = []
    for v in range(count):
        b = 6 * v
        ax = runif(seed, 203, b, T(0), T(0), prec)
        ay = runif(seed, 203, b + 1, T(0), T(2) * pi, prec)
        az = runif(seed, 203, b + 2, T(0), T(2) * pi, prec)
        tu = runif(seed, 203, b + 3, T(-1), T(1), prec)
        tt = runif(seed, 203, b + 4, T(0), T(2) * pi, prec)
        tr = runif(seed, 203, b + 5, T(0), T(Lx) / T(6), prec)
        cax, sax, cay, say = tcos(ax, prec), tsin(ax, prec), tcos(ay, prec), tsin(ay, prec)
        caz, saz = tcos(az, prec), tsin(az, prec)
        root = T(_LIBM.sqrtf(float(T(1) - tu * tu))) if prec == 32 else \
            T(math.sqrt(float(T(1) - tu * tu)))
        c = [cay * caz, cax * saz + (sax * say) * caz, sax * saz - (cax * say) * caz,
             -(cay * saz), cax * caz - (sax * say) * saz, sax * cay + (cax * say) * saz,
             say, sax * cay, cax * cay,
             (tr * root) * tcos(tt, prec), (tr * root) * tsin(tt, prec), tr * tu]
        out += [float(T(x)) for x in c]
    return np.array(out)


def bundle_centres(prec):
    """setup_bundle's 21 line centres (src/initial.f90:476-535)."""
    T = tdt(prec)
    pi = T(PI)
    out = []
    for b in range(3):
        for v1 in range(1, 8):
            rpos, tp = (T(0), T(0)) if v1 == 1 else (T(10), ((T(v1 - 2) * T(2)) * pi) / T(6))
            p1, p2 = rpos * tcos(tp, prec), rpos * tsin(tp, prec)
            if b == 1:
                p2 = p2 - T(30)
            else:
                p1 = p1 - T(30)
            out += [float(p1), float(p2)]
    return np.array(out)


def exp_i_pi(prec):
    """exp(eye*pi) as libtbec forms it (the reference folds it at compile time)."""
    T = tdt(prec)
    return np.array([float(tcos(T(PI), prec)), float(tsin(T(PI), prec))])


def noise(seed, n):
    """add_noise's two uniforms per point (point-major), stream 204."""
    return initial.uniform(seed, 204, np.arange(2 * n))
