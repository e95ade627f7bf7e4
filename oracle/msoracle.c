/*
 * msoracle.c -- CPU restatement of the MSModel Foldy-Lax path (test infrastructure).
 *
 * This file is the parity oracle for the HIP library in msmodel_amd/.  It is
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, always as the checker.  It is written from the published algorithms the
 * reference cites (Numerical Recipes 3rd ed. "Ran" and normal deviates;
 * Abramowitz & Stegun 9.6 / 9.7 for K_nu; Steed/Temme continued fraction CF2
 * for K_nu at intermediate |z|; LAPACK's partial-pivoting LU; Aberth's method)
 * and from the behaviour of the reference functions cited per function.
 * Paths in citations are relative to the MSModel source tree.
 */
#include "msoracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define O_PI      3.1415926535897932385
#define O_TWOPI   6.2831853071795864769
#define O_FOURPI  12.566370614359172954
#define O_SQRTPI  1.7724538509055160273
#define O_EULER   0.5772156649015328606
#define O_SQRT2PI 2.506628274631000502

/* ------------------------------------------------------------------------- */
/* NR3 "Ran": LCG + 64-bit xorshift + multiply-with-carry, combined output.   */
/* nru_random.c:24-73                                                         */
/* ------------------------------------------------------------------------- */
uint64_t orc_rng_next(uint64_t st[4]) {
    uint64_t u = st[0], v = st[1], w = st[2];
    u = u * 2862933555777941757ULL + 7046029254386353087ULL;
    v ^= v >> 17; v ^= v << 31; v ^= v >> 8;
    w = 4294957665ULL * (w & 0xffffffffULL) + (w >> 32);
    uint64_t x = u ^ (u << 21);
    x ^= x >> 35; x ^= x << 4;
    st[0] = u; st[1] = v; st[2] = w;
    return (x + v) ^ w;
}

void orc_rng_init(uint64_t st[4], uint64_t seed) {
    st[3] = seed;
    st[1] = 4101842887655102017ULL;
    st[2] = 1;
    st[0] = seed ^ st[1];
    orc_rng_next(st);
    st[1] = st[0];
    orc_rng_next(st);
    st[2] = st[1];
    orc_rng_next(st);
}

/* 64-bit output scaled by 2^-64 (the reference multiplies by the double 2^-64). */
double orc_rng_double(uint64_t st[4]) {
    return 0x1p-64 * (double)orc_rng_next(st);
}

/* Leva's ratio-of-uniforms normal deviate as given in NR3 (Normaldev). */
double orc_rng_normal(uint64_t st[4]) {
    double u, v, x, y, q;
    for (;;) {
        u = orc_rng_double(st);
        v = 1.7156 * (orc_rng_double(st) - 0.5);
        x = u - 0.449871;
        y = fabs(v) + 0.386595;
        q = x * x + y * (0.19600 * y - 0.25472 * x);
        if (q <= 0.27597) break;
        if (q <= 0.27846 && v * v <= -4. * u * u * log(u)) break;
    }
    return v / u;
}

/* ------------------------------------------------------------------------- */
/* Unit-ball volume/surface by the two-step recursion (green_bessel.c:41-80). */
/* The multiplication order is kept because radius_of_ball feeds positions.   */
/* ------------------------------------------------------------------------- */
double orc_uball_volume(int d) {
    if (d == 0) return 1.;
    double v = (d % 2 == 0) ? O_PI : 2.;
    for (int m = d; m > 2; m -= 2) v *= O_TWOPI / m;
    return v;
}

double orc_uball_surface(int d) {
    if (d == 0) return 0.;
    double s = (d % 2 == 0) ? O_TWOPI : 2.;
    for (int m = d - 2; m > 0; m -= 2) s *= O_TWOPI / m;
    return s;
}

double orc_radius_of_ball(int d, int n) {
    return pow(n / orc_uball_volume(d), 1. / d);
}

/* fill_medium with its shape dispatch and aspect-ratio stretch (medium_util.c:95-193). */
int orc_fill_medium(int d, int n, int shape, double ratio, uint64_t seed, double* pos) {
    uint64_t st[4];
    if (d == 1 && shape == ORC_BALL) shape = ORC_CUBE;
    orc_rng_init(st, seed);
    int ntot = n * d;
    if (shape == ORC_CUBE) {
        double side = pow(n, 1. / d);
        for (int l = 0; l < ntot; l++) pos[l] = side * orc_rng_double(st);
    } else if (shape == ORC_BALL) {
        double rad = orc_radius_of_ball(d, n), rad2 = rad * rad;
        for (int i = 0; i < n; i++) {
            double* x = pos + (size_t)i * d;
            double s2;
            do {  /* rejection sampling in the bounding cube */
                s2 = 0.;
                for (int c = 0; c < d; c++) {
                    x[c] = rad * (2 * orc_rng_double(st) - 1);
                    s2 += x[c] * x[c];
                }
            } while (s2 > rad2);
        }
    } else if (shape == ORC_LATTICE) {
        int side = (int)ceil(pow(n, 1. / d));
        for (int i = 0; i < n; i++) {
            int rem = i;
            for (int c = 0; c < d; c++) { pos[i * d + c] = rem % side; rem /= side; }
        }
    } else { /* gaussian, density 1 at the centre */
        double sig = pow(n, 1. / d) / O_SQRT2PI;
        for (int l = 0; l < ntot; l++) pos[l] = sig * orc_rng_normal(st);
    }
    if (ratio != 1. && d != 1) {
        double f2 = pow(ratio, -1. / d), f1 = ratio * f2;
        for (int l = 0; l < ntot; l++) pos[l] *= (l % d == 0) ? f1 : f2;
    }
    return shape;
}

/* ------------------------------------------------------------------------- */
/* Modified Bessel K_nu(z), complex z, integer or half-integer nu.           */
/* Regimes as in green_bessel.c:208-218: Hankel asymptotic series for        */
/* |z| > 16; ascending series inside the parabola Re z < 3 - (Im z)^2/12;    */
/* otherwise the Steed/Temme continued fraction CF2 (NR3 6.6.2).             */
/* ------------------------------------------------------------------------- */

/* Hankel asymptotic expansion B_nu(z) = K_nu(z) / (sqrt(pi/2z) e^-z), A&S 9.7.2,
 * truncated at the smallest term (the series diverges beyond |2z|). */
static orc_cplx asym_b(int twonu, orc_cplx z) {
    double mu = (double)twonu * twonu;         /* 4 nu^2 */
    orc_cplx term = 1., sum = 1.;
    int kmax = (int)floor(cabs(2. * z));
    for (int k = 1; k < kmax + 1 && k < 200; k++) {
        double odd = 2. * k - 1.;
        orc_cplx nt = term * (mu - odd * odd) / (8. * k * z);
        if (cabs(nt) > cabs(term)) break;
        term = nt;
        sum += term;
        if (cabs(term) < 1e-17 * cabs(sum)) break;
    }
    return sum;
}

/* Ascending series for K_0 and K_1 (A&S 9.6.13 and 9.6.11 with n=1). */
static void series_k01(orc_cplx z, orc_cplx* k0, orc_cplx* k1) {
    orc_cplx q = 0.25 * z * z;                 /* z^2/4 */
    orc_cplx lg = clog(0.5 * z) + O_EULER;     /* ln(z/2) + gamma */
    /* K0 = -(ln(z/2)+gamma) I0 + sum_{k>=1} H_k q^k/(k!)^2 */
    orc_cplx t = 1., i0 = 1., s0 = 0.;
    double hk = 0.;
    /* K1 = 1/z + ln(z/2) I1 - (z/4) sum_{k>=0} (psi(k+1)+psi(k+2)) q^k/(k!(k+1)!) */
    orc_cplx t1 = 1., i1s = 1., s1;
    double psk = -O_EULER, psk1 = 1. - O_EULER;  /* psi(1), psi(2) */
    s1 = psk + psk1;
    for (int k = 1; k < 200; k++) {
        t *= q / ((double)k * k);
        hk += 1. / k;
        i0 += t;
        orc_cplx c0 = hk * t;
        s0 += c0;
        t1 *= q / ((double)k * (k + 1));
        i1s += t1;
        psk += 1. / k;
        psk1 += 1. / (k + 1);
        orc_cplx c1 = (psk + psk1) * t1;
        s1 += c1;
        if (cabs(t) < 1e-18 * (cabs(i0) + cabs(s0)) && cabs(t1) < 1e-18 * (cabs(i1s) + cabs(s1))) break;
    }
    *k0 = -lg * i0 + s0;
    orc_cplx i1 = 0.5 * z * i1s;               /* I1 = (z/2) sum q^k/(k!(k+1)!) */
    *k1 = 1. / z + (lg - O_EULER) * i1 - 0.25 * z * s1;
}

/* Steed's continued fraction CF2 with Temme's normalisation: returns K_nu and
 * K_{nu+1} for 0 <= nu <= 1/2 (used with nu = 0 or 1/2). NR3 section 6.6. */
static void cf2_k(double nu, orc_cplx z, orc_cplx* knu, orc_cplx* knu1) {
    orc_cplx b = 2. * (1. + z), dd = 1. / b, h = dd, delh = dd;
    orc_cplx q1 = 0., q2 = 1.;
    double a1 = 0.25 - nu * nu;
    orc_cplx q = a1, c = a1, a = -a1;
    orc_cplx s = 1. + q * delh;
    for (int i = 1; i < 100000; i++) {
        a -= 2 * i;
        c = -a * c / (i + 1.);
        orc_cplx qn = (q1 - b * q2) / a;
        q1 = q2; q2 = qn;
        q += c * qn;
        b += 2.;
        dd = 1. / (b + a * dd);
        delh = (b * dd - 1.) * delh;
        h += delh;
        orc_cplx dels = q * delh;
        s += dels;
        if (cabs(dels) < 1e-17 * cabs(s)) break;
    }
    h = a1 * h;
    *knu = csqrt(O_PI / (2. * z)) * cexp(-z) / s;
    *knu1 = *knu * (nu + z + 0.5 - h) / z;
}

static int in_series_region(orc_cplx z) {
    return creal(z) < 3.0 - cimag(z) * cimag(z) / 12.0;
}

/* K_0 and K_1 for any z off the negative real axis. */
static void bessel_k01(orc_cplx z, orc_cplx* k0, orc_cplx* k1) {
    double az2 = creal(z) * creal(z) + cimag(z) * cimag(z);
    if (az2 > 256.) {
        orc_cplx pre = csqrt(0.5 * O_PI / z) * cexp(-z);
        *k0 = pre * asym_b(0, z);
        *k1 = pre * asym_b(2, z);
    } else if (in_series_region(z)) {
        series_k01(z, k0, k1);
    } else {
        cf2_k(0., z, k0, k1);
    }
}

/* Finite sum for half-integer order nu = m + 1/2:
 * K_nu(z) = sqrt(pi/2z) e^-z sum_{j=0}^{m} (m+j)! / (j! (m-j)!) (2z)^-j  (A&S 10.2.15). */
static orc_cplx half_k(int m, orc_cplx z) {
    orc_cplx sum = 1., term = 1.;
    for (int j = 1; j <= m; j++) {
        term *= (double)(m + j) * (m - j + 1) / (j * 2. * z);
        sum += term;
    }
    return csqrt(O_PI / (2. * z)) * cexp(-z) * sum;
}

orc_cplx orc_bessel_k(int twonu, orc_cplx z) {
    if (twonu < 0) twonu = -twonu;
    if (twonu % 2) return half_k((twonu - 1) / 2, z);
    orc_cplx k0, k1;
    bessel_k01(z, &k0, &k1);
    int nu = twonu / 2;
    if (nu == 0) return k0;
    for (int m = 1; m < nu; m++) {   /* K_{m+1} = K_{m-1} + (2m/z) K_m, stable upward */
        orc_cplx k2 = k0 + (2. * m / z) * k1;
        k0 = k1; k1 = k2;
    }
    return k1;
}

/* Regularised 0F1(;a;z) = sum z^n / (n! Gamma(a+n)) with the reflection for
 * integer a <= 0 (green_bessel.c:254-272). */
orc_cplx orc_hyp0f1_reg(double a, orc_cplx z) {
    orc_cplx t = 1.;
    if (a < 1. && fmod(a, 1.) == 0.) {
        t = cpow(z, 1. - a);
        a = 2. - a;
    }
    t /= tgamma(a);
    orc_cplx s = t;
    for (int n = 1; n <= 200; n++) {
        t *= z / (n * a);
        s += t;
        a += 1.;
        if (cabs(t) < 1e-17 * cabs(s)) break;
    }
    return s;
}

/* ------------------------------------------------------------------------- */
/* Free Green function and its relatives (green_bessel.c:280-313).           */
/* ------------------------------------------------------------------------- */
orc_cplx orc_green(int d, orc_cplx k, double r) {
    orc_cplx ik = I * k;
    if (d == 1) return cexp(ik * r) / (2. * ik);
    if (d == 3) return -cexp(ik * r) / (O_FOURPI * r);
    orc_cplx z = -ik * r;
    if (d == 2) return -orc_bessel_k(0, z) / O_TWOPI;
    return -cpow(-ik / (O_TWOPI * r), (d - 2.) / 2.) * orc_bessel_k(d - 2, z) / O_TWOPI;
}

orc_cplx orc_green_imag(int d, orc_cplx k, double r) {
    if (cabs(k) * r < 3.0) {
        orc_cplx x = k * r;
        return 0.25 * cpow(k / (2. * O_SQRTPI), d - 2.) * orc_hyp0f1_reg(d / 2., -x * x / 4.);
    }
    double sgn = (creal(k) < 0 && d % 2 == 0) ? -1. : 1.;
    return sgn * (orc_green(d, -k, r) - orc_green(d, k, r)) / (2. * I);
}

orc_cplx orc_dk_green(int d, orc_cplx k, double r) {
    return (k / O_TWOPI) * orc_green(d - 2, k, r);
}

/* Inverse single-atom amplitude F(k)^-1 per model (scattering_model_util.c:101-169). */
static orc_cplx i0_factor(int d, orc_cplx k) {
    return orc_uball_surface(d) * cpow(k / O_TWOPI, d - 2) / (8 * O_PI);
}

orc_cplx orc_invf(int model, const double* p, int d, orc_cplx k) {
    switch (model) {
    case ORC_HARDSPHERE:
        return -i0_factor(d, k) * orc_green(d, k, p[0]) / orc_green_imag(d, k, p[0]);
    case ORC_SOFTSPHERE: {
        double al = p[0];
        if (d == 2) return clog(I / (al * k)) / O_TWOPI;
        double pre = (d % 2 == 0) ? 0.25 : al / 2.;
        for (int m = d - 2; m > 0; m -= 2) pre /= O_TWOPI * al * al * m;
        return pre * (1. + I * cpow(al * k, d - 2));
    }
    case ORC_RESONANT:
        return i0_factor(d, k) * (I - (k - p[0]) / p[1]);
    case ORC_MAXIMUM:
        return I * i0_factor(d, k);
    default:
        return 0.;
    }
}

/* ------------------------------------------------------------------------- */
/* Matrix assembly (ms_matrix_util.c:32-106).                                */
/* ------------------------------------------------------------------------- */
static double pair_distance(int d, const double* x, const double* y) {
    double s = 0.;
    for (int c = 0; c < d; c++) { double t = x[c] - y[c]; s += t * t; }
    return sqrt(s);
}

void orc_build_matrix(int d, int n, const double* pos, int model, const double* p, orc_cplx k, orc_cplx* m) {
    orc_cplx dg = orc_invf(model, p, d, k);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++) {
            orc_cplx e = -orc_green(d, k, pair_distance(d, pos + i * d, pos + j * d));
            m[i + (size_t)j * n] = e;
            m[j + (size_t)i * n] = e;
        }
        m[j + (size_t)j * n] = dg;
    }
}

void orc_build_dk_matrix(int d, int n, const double* pos, int model, const double* p, orc_cplx k, orc_cplx* m) {
    const orc_cplx h = 1.e-4;   /* central difference step of the diagonal */
    orc_cplx dg = (orc_invf(model, p, d, k + h) - orc_invf(model, p, d, k - h)) / (2. * h);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++) {
            orc_cplx e = -orc_dk_green(d, k, pair_distance(d, pos + i * d, pos + j * d));
            m[i + (size_t)j * n] = e;
            m[j + (size_t)i * n] = e;
        }
        m[j + (size_t)j * n] = dg;
    }
}

void orc_build_normalized_matrix(int d, int n, const double* pos, orc_cplx k, orc_cplx* m) {
    orc_cplx iz = orc_green_imag(d, k, 0.);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < j; i++) {
            orc_cplx e = -orc_green(d, k, pair_distance(d, pos + i * d, pos + j * d)) / iz;
            m[i + (size_t)j * n] = e;
            m[j + (size_t)i * n] = e;
        }
        m[j + (size_t)j * n] = I;
    }
}

void orc_build_plane_wave(int d, int n, const double* pos, orc_cplx k, double theta_deg, orc_cplx* v) {
    double th = theta_deg * (O_PI / 180);
    double ct = cos(th), s = sin(th);
    if (d == 1) {
        ct = (ct > 0.) ? 1. : -1.;
        for (int i = 0; i < n; i++) v[i] = cexp(I * k * ct * pos[i]);
    } else {
        for (int i = 0; i < n; i++) v[i] = cexp(I * k * (ct * pos[i * d] + s * pos[i * d + 1]));
    }
}

/* ------------------------------------------------------------------------- */
/* Partial-pivoting LU in LAPACK's conventions (zgetf2 semantics):           */
/* pivot = first row of max |Re|+|Im| (izamax), scale by the reciprocal,     */
/* zero pivot recorded in info and skipped.                                  */
/* ------------------------------------------------------------------------- */
static double cabs1(orc_cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

int orc_getrf(int n, orc_cplx* a, int lda, int* ipiv) {
    int info = 0;
    for (int j = 0; j < n; j++) {
        orc_cplx* cj = a + (size_t)j * lda;
        int p = j;
        double best = cabs1(cj[j]);
        for (int i = j + 1; i < n; i++) {
            double v = cabs1(cj[i]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[j] = p + 1;
        if (cj[p] != 0.) {
            if (p != j)
                for (int c = 0; c < n; c++) {
                    orc_cplx* col = a + (size_t)c * lda;
                    orc_cplx t = col[j]; col[j] = col[p]; col[p] = t;
                }
            orc_cplx r = 1. / cj[j];
            for (int i = j + 1; i < n; i++) cj[i] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; c++) {
            orc_cplx* col = a + (size_t)c * lda;
            orc_cplx u = col[j];
            if (u != 0.)
                for (int i = j + 1; i < n; i++) col[i] -= cj[i] * u;
        }
    }
    return info;
}

void orc_getrs(int n, int nrhs, const orc_cplx* lu, int lda, const int* ipiv, orc_cplx* b, int ldb) {
    for (int r = 0; r < nrhs; r++) {
        orc_cplx* x = b + (size_t)r * ldb;
        for (int i = 0; i < n; i++) {
            int p = ipiv[i] - 1;
            if (p != i) { orc_cplx t = x[i]; x[i] = x[p]; x[p] = t; }
        }
        for (int j = 0; j < n; j++) {           /* L y = P b, unit diagonal */
            orc_cplx xj = x[j];
            if (xj != 0.) {
                const orc_cplx* col = lu + (size_t)j * lda;
                for (int i = j + 1; i < n; i++) x[i] -= xj * col[i];
            }
        }
        for (int j = n - 1; j >= 0; j--) {      /* U x = y */
            const orc_cplx* col = lu + (size_t)j * lda;
            if (x[j] != 0.) {
                x[j] /= col[j];
                orc_cplx xj = x[j];
                for (int i = 0; i < j; i++) x[i] -= xj * col[i];
            }
        }
    }
}

int orc_gesv(int n, int nrhs, orc_cplx* a, int lda, int* ipiv, orc_cplx* b, int ldb) {
    int info = orc_getrf(n, a, lda, ipiv);
    if (info == 0) orc_getrs(n, nrhs, a, lda, ipiv, b, ldb);
    return info;
}

/* Trace-log (1/N) [sum log U_ii + i*pi*(odd number of interchanges)], ms_matrix_util.c:185-202. */
orc_cplx orc_tracelog(int n, orc_cplx* a) {
    int* ipiv = (int*)malloc(sizeof(int) * n);
    orc_getrf(n, a, n, ipiv);
    orc_cplx s = 0.;
    int odd = 0;
    for (int i = 0; i < n; i++) {
        s += clog(a[i + (size_t)i * n]);
        if (ipiv[i] != i + 1) odd ^= 1;
    }
    if (odd) s += I * O_PI;
    free(ipiv);
    return s / n;
}

/* tr(M^-1 dM/dk) with the N-column solve of the reference (ms_matrix_util.c:242-263). */
orc_cplx orc_dk_log_det(int d, int n, const double* pos, int model, const double* p, orc_cplx k) {
    orc_cplx* m = (orc_cplx*)malloc(sizeof(orc_cplx) * (size_t)n * n);
    orc_cplx* dm = (orc_cplx*)malloc(sizeof(orc_cplx) * (size_t)n * n);
    int* ipiv = (int*)malloc(sizeof(int) * n);
    orc_build_matrix(d, n, pos, model, p, k, m);
    orc_build_dk_matrix(d, n, pos, model, p, k, dm);
    orc_getrf(n, m, n, ipiv);
    orc_getrs(n, n, m, n, ipiv, dm, n);
    orc_cplx tr = 0.;
    for (int i = 0; i < n; i++) tr += dm[i + (size_t)i * n];
    free(m); free(dm); free(ipiv);
    return tr;
}

/* Tr(A^-1) via LU and n unit right-hand sides (ms_matrix_util.c:319-335 uses zgetri). */
orc_cplx orc_traceinv(int n, orc_cplx* a) {
    int* ipiv = (int*)malloc(sizeof(int) * n);
    orc_cplx* e = (orc_cplx*)calloc((size_t)n, sizeof(orc_cplx));
    orc_getrf(n, a, n, ipiv);
    orc_cplx tr = 0.;
    for (int j = 0; j < n; j++) {
        memset(e, 0, sizeof(orc_cplx) * n);
        e[j] = 1.;
        orc_getrs(n, 1, a, n, ipiv, e, n);
        tr += e[j];
    }
    free(ipiv); free(e);
    return tr;
}

/* Normalisation with the phase convention of cnormalize (complex_vector_util.c:47-60). */
static orc_cplx cnormalize(int n, orc_cplx* x) {
    int i = 0;
    while (i < n && x[i] == 0.) i++;
    if (i == n) return 1.;
    double nrm = 0.;
    for (int j = 0; j < n; j++) nrm += creal(x[j]) * creal(x[j]) + cimag(x[j]) * cimag(x[j]);
    orc_cplx fac = sqrt(nrm) * x[i] / cabs(x[i]);
    for (int j = 0; j < n; j++) x[j] /= fac;
    return fac;
}

int orc_inverse_iteration(int n, orc_cplx* a, orc_cplx* vec, orc_cplx* lambda, int maxit, double toler) {
    int* ipiv = (int*)malloc(sizeof(int) * n);
    orc_cplx* old = (orc_cplx*)malloc(sizeof(orc_cplx) * n);
    orc_getrf(n, a, n, ipiv);
    orc_cplx fac = cnormalize(n, vec);
    int conv = 0;
    for (int s = 0; s < maxit; s++) {
        memcpy(old, vec, sizeof(orc_cplx) * n);
        orc_getrs(n, 1, a, n, ipiv, vec, n);
        fac = cnormalize(n, vec);
        double d2 = 0.;
        for (int i = 0; i < n; i++) {
            orc_cplx t = vec[i] - old[i];
            d2 += creal(t) * creal(t) + cimag(t) * cimag(t);
        }
        if (sqrt(d2) < toler) { conv = 1; break; }
    }
    *lambda = 1. / fac;
    free(ipiv); free(old);
    return conv;
}

double orc_total_cross_section(int d, int n, const double* pos, int model, const double* p, double k) {
    orc_cplx* m = (orc_cplx*)malloc(sizeof(orc_cplx) * (size_t)n * n);
    orc_cplx* b = (orc_cplx*)malloc(sizeof(orc_cplx) * n);
    orc_cplx* phi = (orc_cplx*)malloc(sizeof(orc_cplx) * n);
    int* ipiv = (int*)malloc(sizeof(int) * n);
    orc_build_matrix(d, n, pos, model, p, k, m);
    for (int i = 0; i < n; i++) { phi[i] = cexp(I * k * pos[i * d]); b[i] = phi[i]; }
    orc_gesv(n, 1, m, n, ipiv, b, n);
    orc_cplx dot = 0.;
    for (int i = 0; i < n; i++) dot += conj(phi[i]) * b[i];
    free(m); free(b); free(phi); free(ipiv);
    return -cimag(dot) / k;
}

/* ------------------------------------------------------------------------- */
/* Incident wave, total wave and the psi grid (incident_wave_util.c:73-96,   */
/* wave_function_util.c:76-81).                                              */
/* ------------------------------------------------------------------------- */
orc_cplx orc_incident(int source, orc_cplx k, int d, const double* x) {
    if (source == ORC_PLANE) return cexp(I * k * x[0]);
    if (source == ORC_SPHERICAL) {
        double s = 0.;
        for (int c = 0; c < d; c++) s += x[c] * x[c];
        return orc_green(d, k, sqrt(s));
    }
    return 0.;
}

orc_cplx orc_total_wave(int source, orc_cplx k, int d, int n, const double* pos, const orc_cplx* a, const double* x) {
    orc_cplx acc = 0.;
    for (int i = 0; i < n; i++) acc += a[i] * orc_green(d, k, pair_distance(d, pos + i * d, x));
    return acc + orc_incident(source, k, d, x);
}

void orc_wavefunction_grid(int source, orc_cplx k, int d, int n, const double* pos, const orc_cplx* a,
                           int nx, int ny, double xmin, double xmax, double ymin, double ymax, orc_cplx* psi) {
    double dx = (xmax - xmin) / (nx - 1), dy = (ymax - ymin) / (ny - 1);
    double x[32] = {0};
    for (int l = 0; l < nx * ny; l++) {
        int ix = l % nx, iy = l / nx;
        x[0] = xmin + ix * dx;
        if (d > 1) x[1] = ymax - iy * dy;
        psi[l] = orc_total_wave(source, k, d, n, pos, a, x);
    }
}

/* ------------------------------------------------------------------------- */
/* k-grids (complex_map_util.c:27-31, complex_cut_util.c:32-36,              */
/* log_cut_util.c:72-101).                                                   */
/* ------------------------------------------------------------------------- */
void orc_kgrid_map(int nx, int ny, double xmin, double xmax, double ymin, double ymax, orc_cplx* k) {
    double dx = (xmax - xmin) / (nx - 1), dy = (ymax - ymin) / (ny - 1);
    for (int l = 0; l < nx * ny; l++) {
        int ix = l % nx, iy = l / nx;
        k[l] = (xmin + ix * dx) + I * (ymax - iy * dy);
    }
}

void orc_kgrid_cut(int ns, orc_cplx zmin, orc_cplx zmax, orc_cplx* k) {
    /* the transverse offsets are +-i dz with the complex step dz = (zmax - zmin)/(ns - 1)
     * (complex_cut_util.c:68): perpendicular to the cut */
    orc_cplx dz = (zmax - zmin) / (ns - 1);
    for (int l = 0; l < 3 * (ns + 2); l++) {
        int il = l / 3, it = l % 3;
        k[l] = ((ns - il) * zmin + (il - 1) * zmax) / (ns - 1) + I * dz * (it - 1);
    }
}

void orc_kgrid_logcut(int ns, double rek, double imkmin, double imkmax, orc_cplx* k) {
    double lf = log(imkmax / imkmin);
    for (int i = 0; i < ns; i++) k[i] = rek + I * imkmin * exp((lf * i) / (ns - 1));
}

/* ------------------------------------------------------------------------- */
/* Halton points and the Aberth root finder (domain_util.c:156-185,          */
/* root_finder.c:161-172, 522-577).                                          */
/* ------------------------------------------------------------------------- */
static double vdc(int i, int b) {
    double h = 0., f = 1.;
    while (i > 0) { f /= b; h += f * (i % b); i /= b; }
    return h;
}

void orc_halton(int n, double xmin, double xmax, double ymin, double ymax, orc_cplx* z) {
    for (int i = 0; i < n; i++)
        z[i] = xmin + (xmax - xmin) * vdc(i + 1, 2) + I * (ymin + (ymax - ymin) * vdc(i + 1, 3));
}

/* Least squares min |A x - b| for a small complex m x nc system by Householder QR. */
static void small_lstsq(int m, int nc, orc_cplx* a, orc_cplx* b) {
    for (int j = 0; j < nc; j++) {
        double nrm = 0.;
        for (int i = j; i < m; i++) nrm += creal(a[i + j * m] * conj(a[i + j * m]));
        nrm = sqrt(nrm);
        if (nrm == 0.) continue;
        orc_cplx ajj = a[j + j * m];
        orc_cplx alpha = (cabs(ajj) > 0 ? -ajj / cabs(ajj) : -1.) * nrm;
        orc_cplx* v = (orc_cplx*)malloc(sizeof(orc_cplx) * m);
        for (int i = 0; i < m; i++) v[i] = (i < j) ? 0. : a[i + j * m];
        v[j] -= alpha;
        double vn = 0.;
        for (int i = j; i < m; i++) vn += creal(v[i] * conj(v[i]));
        if (vn > 0.) {
            for (int c = j; c < nc; c++) {
                orc_cplx s = 0.;
                for (int i = j; i < m; i++) s += conj(v[i]) * a[i + c * m];
                s = 2. * s / vn;
                for (int i = j; i < m; i++) a[i + c * m] -= s * v[i];
            }
            orc_cplx s = 0.;
            for (int i = j; i < m; i++) s += conj(v[i]) * b[i];
            s = 2. * s / vn;
            for (int i = j; i < m; i++) b[i] -= s * v[i];
        }
        free(v);
    }
    for (int j = nc - 1; j >= 0; j--) {
        orc_cplx s = b[j];
        for (int c = j + 1; c < nc; c++) s -= a[j + c * m] * b[c];
        b[j] = s / a[j + j * m];
    }
}

static int in_open_domain(orc_cplx z, double xmin, double xmax, double ymin, double ymax) {
    return xmin < creal(z) && creal(z) < xmax && ymin < cimag(z) && cimag(z) < ymax;
}

static int cmp_real(const void* x, const void* y) {
    double a = creal(*(const orc_cplx*)x), b = creal(*(const orc_cplx*)y);
    return (a > b) - (a < b);
}

void orc_aberth(int d, int n, const double* pos, int model, const double* p,
                double xmin, double xmax, double ymin, double ymax,
                int* nz, orc_cplx* z, int nexpsup, int maxit, double toler) {
    /* duplicate removal and domain filter of the starting points */
    int cnt = 1;
    for (int i = 1; i < *nz; i++) {
        int uniq = 1;
        for (int j = 0; j < cnt; j++) {
            double dd = cabs(z[i] - z[j]);
            if (dd < toler * cabs(z[i]) || dd < toler) { uniq = 0; break; }
        }
        if (uniq) { orc_cplx t = z[i]; z[i] = z[cnt]; z[cnt] = t; cnt++; }
    }
    int nin = 0;
    for (int i = 0; i < cnt; i++)
        if (in_open_domain(z[i], xmin, xmax, ymin, ymax)) { orc_cplx t = z[i]; z[i] = z[nin]; z[nin] = t; nin++; }
    *nz = nin;
    if (nin <= 1) return;
    int deg = nexpsup - 1;
    orc_cplx poly[16] = {0};
    if (nexpsup > 0) {
        int ns = 6 * nexpsup;
        orc_cplx* zs = (orc_cplx*)malloc(sizeof(orc_cplx) * ns);
        orc_cplx* a = (orc_cplx*)malloc(sizeof(orc_cplx) * ns * nexpsup);
        orc_cplx* b = (orc_cplx*)malloc(sizeof(orc_cplx) * ns);
        orc_halton(ns, xmin, xmax, ymin, ymax, zs);
        for (int i = 0; i < ns; i++) {
            orc_cplx zp = 1.;
            for (int j = 0; j < nexpsup; j++) { a[i + j * ns] = zp; zp *= zs[i]; }
            b[i] = orc_dk_log_det(d, n, pos, model, p, zs[i]);
        }
        small_lstsq(ns, nexpsup, a, b);
        for (int j = 0; j < nexpsup; j++) poly[j] = b[j];
        free(zs); free(a); free(b);
    }
    int* flag = (int*)calloc(nin, sizeof(int));
    for (int s = 0; s < maxit; s++) {
        int busy = 0;
        for (int i = 0; i < nin; i++) {
            if (flag[i] != 0) continue;
            orc_cplx pv = 0.;
            for (int j = deg; j >= 0; j--) pv = pv * z[i] + poly[j];
            orc_cplx inv = orc_dk_log_det(d, n, pos, model, p, z[i]) - pv;
            for (int j = 0; j < nin; j++) if (j != i) inv -= 1. / (z[i] - z[j]);
            orc_cplx dz = 1. / inv, zn = z[i] - dz;
            if (!in_open_domain(zn, xmin, xmax, ymin, ymax)) flag[i] = -1;
            else if (cabs(dz) < toler * cabs(z[i])) flag[i] = 1;
            else z[i] = zn;
        }
        for (int i = 0; i < nin; i++) if (flag[i] == 0) busy = 1;
        if (!busy) break;
    }
    int kept = 0;
    for (int i = 0; i < nin; i++)
        if (flag[i] == 1) { orc_cplx t = z[i]; z[i] = z[kept]; z[kept] = t; kept++; }
    *nz = kept;
    qsort(z, kept, sizeof(orc_cplx), cmp_real);
    free(flag);
}

/* ------------------------------------------------------------------------- */
/* Damped Muller root finder on another function of M(k) (root_finder.c:     */
/* 342-407) with the drivers find_eigenstates_serial_2/_3 (ms_core.c:335-401) */
/* and the vector filters filter_cvector / uniq_cvector                      */
/* (complex_vector_util.c:130-177).                                          */
/* ------------------------------------------------------------------------- */
typedef struct {
    int d, n, method;
    const double* pos;
    int model;
    const double* p;
    orc_cplx* m;      /* n x n work matrix */
    orc_cplx* vec;    /* the inverse-iteration vector (mineigval) */
    orc_cplx* vsave;  /* method 3: the start's vector before its first three points */
} MullerCtx;

/* f(k) = the eigenvalue reached by inverse iteration (50 steps, 1e-10) from the shared
 * vector, which keeps the last result (method 1); or 1 / tr M(k)^-1 (method 2). */
static orc_cplx muller_f(MullerCtx* c, orc_cplx k) {
    orc_build_matrix(c->d, c->n, c->pos, c->model, c->p, k, c->m);
    if (c->method == 1 || c->method == 3) {
        orc_cplx mu;
        orc_inverse_iteration(c->n, c->m, c->vec, &mu, 50, 1e-10);
        return mu;
    }
    return 1. / orc_traceinv(c->n, c->m);
}

static int muller_in_domain(orc_cplx z, double xmin, double xmax, double ymin, double ymax) {
    return xmin < creal(z) && creal(z) < xmax && ymin < cimag(z) && cimag(z) < ymax;
}

/* One start: returns 1 on convergence (root in *z), 0 otherwise. */
static int muller_one(MullerCtx* c, double xmin, double xmax, double ymin, double ymax, orc_cplx* z, int maxit,
                      double toler) {
    if (maxit <= 1 || toler <= 0.) return 0;
    const double h = 1e-4;                        /* relative offsets of the first two points */
    orc_cplx za = (1. - h * I) * (*z), zb = *z, zc = (1. + h * I) * (*z);
    orc_cplx fa, fb, fc;
    if (c->method == 3) {     /* the three first points all start from the start's vector */
        memcpy(c->vsave, c->vec, sizeof(orc_cplx) * c->n);
        fa = muller_f(c, za);
        memcpy(c->vec, c->vsave, sizeof(orc_cplx) * c->n);
        fb = muller_f(c, zb);
        memcpy(c->vec, c->vsave, sizeof(orc_cplx) * c->n);
        fc = muller_f(c, zc);
    } else {
        fa = muller_f(c, za);
        fb = muller_f(c, zb);
        fc = muller_f(c, zc);
    }
    double absfc = cabs(fc);
    for (int s = 0; s < maxit; s++) {
        /* divided differences; the quadratic through the three points, taken at zc */
        orc_cplx fab = (fa - fb) / (za - zb), fac = (fa - fc) / (za - zc), fbc = (fb - fc) / (zb - zc);
        orc_cplx d1 = fbc + fac - fab, d2 = (fac - fbc) / (za - zb);
        orc_cplx sq = csqrt(d1 * d1 - 4. * fc * d2);
        orc_cplx den = d1 + sq;
        if (cabs(d1 - sq) > cabs(den)) den = d1 - sq;
        if (den == 0. || den != den) return 0;
        orc_cplx dz = 2. * fc / den;
        double adz = cabs(dz), tol = toler * cabs(zc);
        if (!muller_in_domain(zc, xmin, xmax, ymin, ymax)) return 0;
        if (adz < tol || absfc == 0.) { *z = zc; return 1; }
        /* damping: mu = 1, 1/2, 1/8, 1/64, ... while |f| grows */
        double mu = 1., div = 2., absfn;
        orc_cplx zn, fn;
        do {
            zn = zc - mu * dz;
            fn = muller_f(c, zn);
            absfn = cabs(fn);
            mu /= div;
            div *= 2.;
        } while (absfn > absfc && mu * adz > tol);
        za = zb; fa = fb;
        zb = zc; fb = fc;
        zc = zn; fc = fn; absfc = absfn;
    }
    return 0;
}

int orc_roots_muller(int d, int n, const double* pos, int model, const double* p, double xmin, double xmax,
                     double ymin, double ymax, int* nz, orc_cplx* z, int maxit, double toler, int method) {
    MullerCtx c = {d, n, method, pos, model, p, NULL, NULL, NULL};
    c.m = (orc_cplx*)malloc(sizeof(orc_cplx) * (size_t)n * n);
    c.vec = (orc_cplx*)malloc(sizeof(orc_cplx) * n);
    c.vsave = (orc_cplx*)malloc(sizeof(orc_cplx) * n);
    for (int i = 0; i < n; i++) c.vec[i] = 1. / sqrt((double)n);      /* constant_unit_cvector */
    int* conv = (int*)calloc(*nz, sizeof(int));
    for (int i = 0; i < *nz; i++) {
        if (method == 3)                     /* each start from the constant unit vector */
            for (int j = 0; j < n; j++) c.vec[j] = 1. / sqrt((double)n);
        conv[i] = muller_one(&c, xmin, xmax, ymin, ymax, &z[i], maxit, toler);
        if (method != 2) conv[i] *= (cabs(muller_f(&c, z[i])) < sqrt(toler));
    }
    /* keep the converged roots in order, then the first of each group closer than 1e-14 */
    int size = 0;
    for (int i = 0; i < *nz; i++)
        if (conv[i] == 1) { orc_cplx t = z[i]; z[i] = z[size]; z[size] = t; size++; }
    int u = size > 0 ? 1 : 0;
    for (int i = 1; i < size; i++) {
        int uniq = 1;
        for (int j = 0; j < u; j++)
            if (cabs(z[i] - z[j]) < 1e-14 * cabs(z[i]) || cabs(z[i] - z[j]) < 1e-14) { uniq = 0; break; }
        if (uniq) { orc_cplx t = z[i]; z[i] = z[u]; z[u] = t; u++; }
    }
    *nz = u;
    free(c.m); free(c.vec); free(c.vsave); free(conv);
    return u;
}
