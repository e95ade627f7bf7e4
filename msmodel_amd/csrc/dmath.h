// dmath.h -- complex arithmetic, Green functions and scattering amplitudes for
// the device (and host), templated on the real type (double or float).
//
// Green function conventions follow green_bessel.c:280-313 of the reference:
//   G_1 =  e^{ikr}/(2ik),  G_3 = -e^{ikr}/(4 pi r),  G_2 = -K_0(-ikr)/(2 pi),
//   G_4 = -(-ik/(2 pi r)) K_1(-ikr)/(2 pi),  dG_d/dk = (k/2pi) G_{d-2}.
// The MS matrix stores -G off the diagonal (ms_matrix_util.c:32-44).
// K_0/K_1 use three regimes (as green_bessel.c:208-218): the Hankel asymptotic
// series for |z|>16, the ascending series (Abramowitz & Stegun 9.6.11/9.6.13)
// inside the parabola Re z < 3 - (Im z)^2/12, and Steed's continued fraction
// CF2 with Temme's normalisation elsewhere.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define MSG_HD __host__ __device__ __forceinline__

namespace msg {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr double kFourPi = 12.5663706143591729539;
constexpr double kEuler = 0.57721566490153286061;

template <class T> struct cx {
    T x, y;
};

template <class T> MSG_HD cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <class T> MSG_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> MSG_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T> MSG_HD cx<T> operator-(cx<T> a) { return {-a.x, -a.y}; }
template <class T> MSG_HD cx<T> operator*(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <class T> MSG_HD cx<T> operator*(T s, cx<T> a) { return {s * a.x, s * a.y}; }
template <class T> MSG_HD cx<T> operator*(cx<T> a, T s) { return {s * a.x, s * a.y}; }
template <class T> MSG_HD cx<T> operator+(cx<T> a, T s) { return {a.x + s, a.y}; }
template <class T> MSG_HD cx<T> operator-(cx<T> a, T s) { return {a.x - s, a.y}; }
template <class T> MSG_HD T abs1(cx<T> a) { return fabs(a.x) + fabs(a.y); }
template <class T> MSG_HD T cabs(cx<T> a) { return hypot(a.x, a.y); }
template <class T> MSG_HD T norm2(cx<T> a) { return a.x * a.x + a.y * a.y; }
template <class T> MSG_HD cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// Smith's division, robust against overflow of |b|^2.
template <class T> MSG_HD cx<T> operator/(cx<T> a, cx<T> b) {
    if (fabs(b.x) >= fabs(b.y)) {
        T r = b.y / b.x, den = b.x + b.y * r;
        return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
    }
    T r = b.x / b.y, den = b.x * r + b.y;
    return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}
template <class T> MSG_HD cx<T> operator/(cx<T> a, T s) { return {a.x / s, a.y / s}; }
template <class T> MSG_HD cx<T> inv(cx<T> b) { return cx<T>{T(1), T(0)} / b; }
MSG_HD void sincos_t(double x, double* s, double* c) { sincos(x, s, c); }
MSG_HD void sincos_t(float x, float* s, float* c) { sincosf(x, s, c); }
template <class T> MSG_HD cx<T> cexp(cx<T> z) {
    T e = exp(z.x), s, c;
    sincos_t(z.y, &s, &c);
    return {e * c, e * s};
}
template <class T> MSG_HD cx<T> clog(cx<T> z) { return {log(hypot(z.x, z.y)), atan2(z.y, z.x)}; }
// Principal square root.
template <class T> MSG_HD cx<T> csqrt(cx<T> z) {
    T m = hypot(z.x, z.y);
    if (m == T(0)) return {T(0), T(0)};
    T a = sqrt((m + fabs(z.x)) * T(0.5));
    if (z.x >= T(0)) return {a, z.y / (T(2) * a)};
    return {fabs(z.y) / (T(2) * a), z.y >= T(0) ? a : -a};
}
// Principal power z^p for real p (p may be half-integer).
template <class T> MSG_HD cx<T> cpowr(cx<T> z, T p) {
    cx<T> l = clog(z);
    return cexp(cx<T>{p * l.x, p * l.y});
}
template <class T> MSG_HD cx<T> ipow(cx<T> z, int e) {   // integer power by products
    cx<T> r{T(1), T(0)};
    if (e < 0) { z = inv(z); e = -e; }
    for (int i = 0; i < e; i++) r = r * z;
    return r;
}

template <class T> struct eps_of;
template <> struct eps_of<double> { static constexpr double v = 1e-17; };
template <> struct eps_of<float> { static constexpr float v = 3e-8f; };

// Hankel asymptotic series B_nu(z) = K_nu(z)/(sqrt(pi/2z) e^{-z}) for nu=0,1 at once,
// truncated at its smallest term (A&S 9.7.2).
template <class T> MSG_HD void asym_b01(cx<T> z, cx<T>& b0, cx<T>& b1) {
    cx<T> t0{T(1), T(0)}, t1{T(1), T(0)};
    b0 = t0; b1 = t1;
    cx<T> iz8 = inv(T(8) * z);
    int kmax = (int)floor(T(2) * cabs(z));
    if (kmax > 80) kmax = 80;
    bool live0 = true, live1 = true;
    const T e2 = eps_of<T>::v * eps_of<T>::v;      // term tests on squared moduli
    for (int k = 1; k <= kmax && (live0 || live1); k++) {
        T odd = T(2 * k - 1), q = odd * odd, rk = T(1) / T(k);
        if (live0) {
            cx<T> n0 = t0 * iz8 * (-q * rk);
            if (norm2(n0) > norm2(t0)) live0 = false;
            else { t0 = n0; b0 = b0 + t0; if (norm2(t0) < e2 * norm2(b0)) live0 = false; }
        }
        if (live1) {
            cx<T> n1 = t1 * iz8 * ((T(4) - q) * rk);
            if (norm2(n1) > norm2(t1)) live1 = false;
            else { t1 = n1; b1 = b1 + t1; if (norm2(t1) < e2 * norm2(b1)) live1 = false; }
        }
    }
}

// Ascending series of K_0 and K_1 (A&S 9.6.13 and 9.6.11 with n=1).
template <class T> MSG_HD void series_k01(cx<T> z, cx<T>& k0, cx<T>& k1) {
    cx<T> q = T(0.25) * (z * z);
    cx<T> lz = clog(T(0.5) * z);
    cx<T> t{T(1), T(0)}, i0 = t, s0{T(0), T(0)};
    cx<T> t1 = t, i1s = t;
    T hk = T(0), pk = T(-kEuler), pk1 = T(1.0 - kEuler);
    cx<T> s1{pk + pk1, T(0)};
    // one division per term: rk = 1/k carried from the previous term's 1/(k+1); the stop
    // test |t| < eps (|i0| + |s0|) is taken on squared moduli (never earlier)
    const T e2 = eps_of<T>::v * eps_of<T>::v;
    T rk = T(1);
    for (int k = 1; k < 120; k++) {
        const T rk1 = T(1) / T(k + 1);
        t = t * q * (rk * rk);
        hk += rk;
        i0 = i0 + t;
        s0 = s0 + hk * t;
        t1 = t1 * q * (rk * rk1);
        i1s = i1s + t1;
        pk += rk;
        pk1 += rk1;
        s1 = s1 + (pk + pk1) * t1;
        if (norm2(t) < e2 * (norm2(i0) + norm2(s0)) && norm2(t1) < e2 * (norm2(i1s) + norm2(s1))) break;
        rk = rk1;
    }
    k0 = s0 - (lz + T(kEuler)) * i0;
    cx<T> i1 = T(0.5) * (z * i1s);
    k1 = inv(z) + lz * i1 - T(0.25) * (z * s1);
}

// Steed's CF2 with Temme's sum for nu=0: returns K_0(z), K_1(z).
template <class T> MSG_HD void cf2_k01(cx<T> z, cx<T>& k0, cx<T>& k1) {
    cx<T> b = T(2) * (z + T(1));
    cx<T> dd = inv(b), h = dd, delh = dd;
    cx<T> q1{T(0), T(0)}, q2{T(1), T(0)};
    const T a1 = T(0.25);
    cx<T> q{a1, T(0)}, c{a1, T(0)};
    T a = -a1;
    cx<T> s = q * delh + T(1);
    for (int i = 1; i < 600; i++) {
        a -= T(2 * i);
        c = c * (-a / T(i + 1));
        cx<T> qn = (q1 - b * q2) * (T(1) / a);
        q1 = q2; q2 = qn;
        q = q + c * qn;
        b = b + T(2);
        dd = inv(b + a * dd);
        delh = (b * dd - T(1)) * delh;
        h = h + delh;
        cx<T> dels = q * delh;
        s = s + dels;
        if (norm2(dels) < eps_of<T>::v * eps_of<T>::v * norm2(s)) break;
    }
    h = a1 * h;
    k0 = csqrt(cx<T>{T(kPi), T(0)} / (T(2) * z)) * cexp(-z) / s;
    k1 = k0 * (z + T(0.5) - h) / z;
}

template <class T> MSG_HD void bessel_k01(cx<T> z, cx<T>& k0, cx<T>& k1) {
    T a2 = z.x * z.x + z.y * z.y;
    if (a2 > T(256)) {
        cx<T> b0, b1;
        asym_b01(z, b0, b1);
        cx<T> pre = csqrt(cx<T>{T(0.5 * kPi), T(0)} / z) * cexp(-z);
        k0 = pre * b0;
        k1 = pre * b1;
    } else if (z.x < T(3) - z.y * z.y / T(12)) {
        series_k01(z, k0, k1);
    } else {
        cf2_k01(z, k0, k1);
    }
}

// -G_d(k,r) (the off-diagonal MS matrix element) for d = 1..4.
template <class T> MSG_HD cx<T> neg_green(int d, cx<T> k, T r) {
    cx<T> ikr{-k.y * r, k.x * r};               // i k r
    if (d == 3) return cexp(ikr) * (T(1) / (T(kFourPi) * r));
    if (d == 1) return -(cexp(ikr) / cx<T>{-T(2) * k.y, T(2) * k.x});
    cx<T> z = -ikr, k0, k1;                       // z = -ikr
    bessel_k01(z, k0, k1);
    if (d == 2) return k0 * T(1.0 / kTwoPi);
    // d == 4: -G_4 = (-ik/(2 pi r)) K_1(-ikr) / (2 pi)
    cx<T> w{k.y / (T(kTwoPi) * r), -k.x / (T(kTwoPi) * r)};
    return (w * k1) * T(1.0 / kTwoPi);
}

// -dG_d/dk (the off-diagonal element of dM/dk) for d = 1..4, in closed form:
//   d=1: -e^{ikr}(kr+i)/(2k^2); d=3: i e^{ikr}/(4pi);
//   d=2: i r K_1(-ikr)/(2pi);   d=4: k K_0(-ikr)/(4 pi^2).
template <class T> MSG_HD cx<T> neg_dk_green(int d, cx<T> k, T r) {
    cx<T> ikr{-k.y * r, k.x * r};
    if (d == 3) { cx<T> e = cexp(ikr); return cx<T>{-e.y, e.x} * T(1.0 / kFourPi); }
    if (d == 1) {
        cx<T> e = cexp(ikr);
        cx<T> num = e * cx<T>{k.x * r, k.y * r + T(1)};
        return -(num / (T(2) * (k * k)));
    }
    cx<T> k0, k1;
    bessel_k01(-ikr, k0, k1);
    if (d == 2) return cx<T>{-k1.y, k1.x} * (r / T(kTwoPi));
    return (k * k0) * T(1.0 / (kFourPi * kPi));
}

// Full free Green function G_d(k,r) for d=1..4 (used by the wave sum and invf).
template <class T> MSG_HD cx<T> green(int d, cx<T> k, T r) { return -neg_green(d, k, r); }

// Unit-ball surface S_d for d = 1..4 (green_bessel.c:63-80).
MSG_HD double uball_surface(int d) {
    double s = (d % 2 == 0) ? kTwoPi : 2.0;
    for (int m = d - 2; m > 0; m -= 2) s *= kTwoPi / m;
    return d == 0 ? 0.0 : s;
}

// Regularised 0F1(;a;z) for a > 0 (green_bessel.c:254-272 without the reflection,
// which only integer a <= 0 needs; here a = d/2 >= 1/2).
MSG_HD cx<double> hyp0f1_reg(double a, cx<double> z) {
    cx<double> t{1.0 / tgamma(a), 0.0}, s = t;
    for (int n = 1; n <= 200; n++) {
        t = t * z * (1.0 / (n * a));
        s = s + t;
        a += 1.0;
        if (cabs(t) < 1e-17 * cabs(s)) break;
    }
    return s;
}

// I_d(k,r), the analytic continuation of -Im G (green_bessel.c:298-305).
MSG_HD cx<double> green_imag(int d, cx<double> k, double r) {
    if (cabs(k) * r < 3.0) {
        cx<double> x = k * r;
        cx<double> pre = ipow(k * (1.0 / (2.0 * 1.7724538509055160273)), d - 2);
        return 0.25 * (pre * hyp0f1_reg(d / 2.0, -0.25 * (x * x)));
    }
    double sgn = (k.x < 0 && d % 2 == 0) ? -1.0 : 1.0;
    cx<double> dif = green(d, -k, r) - green(d, k, r);   // (G(-k) - G(k)) / (2i)
    return cx<double>{sgn * dif.y * 0.5, -sgn * dif.x * 0.5};
}

// Scattering model parameters as carried to the device.
struct ScModel {
    int type;      // 0 hardsphere, 1 softsphere, 2 resonant, 3 maximum
    double p0, p1;
};

// F(k)^-1 (scattering_model_util.c:101-169), FP64.
MSG_HD cx<double> invf(const ScModel& m, int d, cx<double> k) {
    cx<double> i0 = (uball_surface(d) / (8.0 * kPi)) * ipow(k * (1.0 / kTwoPi), d - 2);
    switch (m.type) {
    case 0: return -(i0 * green(d, k, m.p0) / green_imag(d, k, m.p0));
    case 1: {
        double al = m.p0;
        if (d == 2) return clog(cx<double>{0.0, 1.0} / (al * k)) * (1.0 / kTwoPi);
        double pre = (d % 2 == 0) ? 0.25 : al / 2.0;
        for (int q = d - 2; q > 0; q -= 2) pre /= kTwoPi * al * al * q;
        cx<double> ak = ipow(al * k, d - 2);
        return pre * cx<double>{1.0 - ak.y, ak.x};
    }
    case 2: return i0 * (cx<double>{0.0, 1.0} - (k - m.p0) * (1.0 / m.p1));
    case 3: return i0 * cx<double>{0.0, 1.0};
    default: return cx<double>{0.0, 0.0};
    }
}

}  // namespace msg
