// medium.cpp -- host side of the medium boundary: the NR3 "Ran" generator,
// media filling, scattering-model construction, distances.  Positions are pure
// FP64 host arithmetic in the reference's operation order, so they match the
// reference bit for bit; the device only ever receives finished positions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "capi_internal.h"

namespace msg {

namespace {
constexpr double kPiD = 3.141592653589793238;
constexpr double kTwoPiD = 6.2831853071795864769;
constexpr double kSqrtTwoPi = 2.506628274631000502;

struct Ran {   // Numerical Recipes 3rd ed. section 7.1.3, "Ran"
    uint64_t u, v, w;
    uint64_t step() {
        u = u * 2862933555777941757ULL + 7046029254386353087ULL;
        v ^= v >> 17;
        v ^= v << 31;
        v ^= v >> 8;
        w = 4294957665ULL * (w & 0xffffffffULL) + (w >> 32);
        uint64_t x = u ^ (u << 21);
        x ^= x >> 35;
        x ^= x << 4;
        return (x + v) ^ w;
    }
    double uniform() { return 0x1p-64 * (double)step(); }
    double normal() {   // Leva's ratio of uniforms (NR3 Normaldev)
        for (;;) {
            double a = uniform(), b = 1.7156 * (uniform() - 0.5);
            double x = a - 0.449871, y = std::fabs(b) + 0.386595;
            double q = x * x + y * (0.19600 * y - 0.25472 * x);
            if (q <= 0.27597) return b / a;
            if (q <= 0.27846 && b * b <= -4. * a * a * std::log(a)) return b / a;
        }
    }
};

void ran_seed(msgpu_urandom* st, uint64_t seed) {
    Ran r{seed ^ 4101842887655102017ULL, 4101842887655102017ULL, 1};
    r.step();
    r.v = r.u;
    r.step();
    r.w = r.v;
    r.step();
    st->u = r.u; st->v = r.v; st->w = r.w; st->seed = seed;
}
}  // namespace

double ball_volume(int d) {
    if (d == 0) return 1.;
    double v = (d % 2 == 0) ? kPiD : 2.;
    for (int m = d; m > 2; m -= 2) v *= kTwoPiD / m;
    return v;
}

double ball_radius(int d, int n) { return std::pow(n / ball_volume(d), 1. / d); }

void fill_positions(msgpu_medium* med, uint64_t seed) {
    ran_seed(med->rnd, seed);
    Ran r{med->rnd->u, med->rnd->v, med->rnd->w};
    const int d = med->d, n = med->n, ntot = n * d;
    double* pos = med->pos;
    switch (med->shape) {
    case MSGPU_CUBE: {
        const double side = std::pow(n, 1. / d);
        for (int l = 0; l < ntot; l++) pos[l] = side * r.uniform();
        break;
    }
    case MSGPU_BALL: {
        const double rad = ball_radius(d, n), rad2 = rad * rad;
        for (int i = 0; i < n; i++) {
            double* x = pos + (size_t)i * d;
            double s2;
            do {
                s2 = 0.;
                for (int c = 0; c < d; c++) {
                    x[c] = rad * (2 * r.uniform() - 1);
                    s2 += x[c] * x[c];
                }
            } while (s2 > rad2);
        }
        break;
    }
    case MSGPU_LATTICE: {
        const int side = (int)std::ceil(std::pow(n, 1. / d));
        for (int i = 0; i < n; i++) {
            int rem = i;
            for (int c = 0; c < d; c++) { pos[i * d + c] = rem % side; rem /= side; }
        }
        break;
    }
    default: {
        const double sig = std::pow(n, 1. / d) / kSqrtTwoPi;
        for (int l = 0; l < ntot; l++) pos[l] = sig * r.normal();
    }
    }
    if (med->ratio != 1. && d != 1) {
        const double f2 = std::pow(med->ratio, -1. / d), f1 = med->ratio * f2;
        for (int l = 0; l < ntot; l++) pos[l] *= (l % d == 0) ? f1 : f2;
    }
    med->rnd->u = r.u; med->rnd->v = r.v; med->rnd->w = r.w;
}

ScModel scmodel_of(const msgpu_medium* med) {
    ScModel m{};
    const msgpu_scmodel* s = med->scmodel;
    m.type = s->type;
    m.p0 = s->np > 0 ? s->p[0] : 0.0;
    m.p1 = s->np > 1 ? s->p[1] : 0.0;
    return m;
}

}  // namespace msg

using namespace msg;

extern "C" {

MSGPU_API int init_scmodel(msgpu_scmodel* sm, int type, int np, double* p) {
    static const int expected[4] = {1, 1, 2, 0};
    if (type < 0 || type > 3 || np != expected[type]) {
        printf("[ERROR] Unexpected number of scattering parameters. Given %d but expected %d, aborting...\n", np,
               (type >= 0 && type <= 3) ? expected[type] : 0);
        exit(EXIT_FAILURE);
    }
    static const char* names[4] = {"hardsphere", "softsphere", "resonant", "maximum"};
    memset(sm->dir, 0, sizeof sm->dir);
    memset(sm->param_ascii, 0, sizeof sm->param_ascii);
    memset(sm->param_latex, 0, sizeof sm->param_latex);
    snprintf(sm->name, sizeof sm->name, "%s", names[type]);
    switch (type) {
    case 0:
        snprintf(sm->dir, sizeof sm->dir, "a%g", p[0]);
        snprintf(sm->param_ascii, sizeof sm->param_ascii, "alpha=%gsp", p[0]);
        snprintf(sm->param_latex, sizeof sm->param_latex, "$\\alpha=%g\\,\\varsigma$", p[0]);
        break;
    case 1:
        snprintf(sm->dir, sizeof sm->dir, "b%g", p[0]);
        snprintf(sm->param_ascii, sizeof sm->param_ascii, "alpha=%gsp", p[0]);
        snprintf(sm->param_latex, sizeof sm->param_latex, "$\\alpha=%g\\,\\varsigma$", p[0]);
        break;
    case 2:
        snprintf(sm->dir, sizeof sm->dir, "r%g", p[0]);
        snprintf(sm->param_ascii, sizeof sm->param_ascii, "p=%g%+gi sp^-1", p[0], p[1]);
        snprintf(sm->param_latex, sizeof sm->param_latex, "$p=(%g%+g\\,{\\rm i})\\varsigma^{-1}$", p[0], p[1]);
        break;
    default:
        snprintf(sm->dir, sizeof sm->dir, "max");
    }
    sm->type = type;
    sm->np = np;
    sm->p = (double*)calloc(np > 0 ? np : 1, sizeof(double));
    for (int i = 0; i < np; i++) sm->p[i] = p[i];
    return 0;
}

MSGPU_API void init_medium(msgpu_medium* med, int d, int n, msgpu_scmodel* sm, int shape, double ratio) {
    if (n < 2) { printf("[ERROR] The number of atoms should be strictly greater than 1, aborting...\n"); exit(EXIT_FAILURE); }
    if (n > 5e4) { printf("[ERROR] The number of atoms n=%d is too large, aborting...\n", n); exit(EXIT_FAILURE); }
    if (d < 1) { printf("[ERROR] The dimension should be strictly positive, aborting...\n"); exit(EXIT_FAILURE); }
    if (d > 20) { printf("[ERROR] The dimension d=%d is too large, aborting...\n", d); exit(EXIT_FAILURE); }
    if (ratio < 1.) { printf("[ERROR] The aspect ratio should be larger or equal to 1, aborting...\n"); exit(EXIT_FAILURE); }
    if (d == 1 && shape == MSGPU_BALL) {
        shape = MSGPU_CUBE;
        printf("[WARN] Changed medium shape 'ball' to 'cube' since 1D.\n");
    }
    med->d = d;
    med->n = n;
    med->ntot = n * d;
    med->shape = shape;
    med->ratio = ratio;
    med->alpha = 0.0;
    med->pos = (double*)calloc((size_t)n * d, sizeof(double));
    med->rnd = (msgpu_urandom*)calloc(1, sizeof(msgpu_urandom));
    med->scmodel = (msgpu_scmodel*)calloc(1, sizeof(msgpu_scmodel));
    init_scmodel(med->scmodel, sm->type, sm->np, sm->p);
}

MSGPU_API void init_copy_medium(msgpu_medium* src, msgpu_medium* dst) {
    init_medium(dst, src->d, src->n, src->scmodel, src->shape, src->ratio);
}

MSGPU_API void del_medium(msgpu_medium* med) {
    if (med->scmodel) free(med->scmodel->p);
    free(med->scmodel);
    free(med->pos);
    free(med->rnd);
}

MSGPU_API double radius_of_ball(msgpu_medium* med) { return ball_radius(med->d, med->n); }

MSGPU_API void fill_medium(msgpu_medium* med, uint64_t seed) { fill_positions(med, seed); }

MSGPU_API double atom_distance(msgpu_medium* med, int i, int j) {
    const double* x = med->pos + (size_t)i * med->d;
    const double* y = med->pos + (size_t)j * med->d;
    double s = 0.;
    for (int c = 0; c < med->d; c++) { double t = x[c] - y[c]; s += t * t; }
    return std::sqrt(s);
}

}  // extern "C"
