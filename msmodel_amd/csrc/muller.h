// muller.h -- the damped Muller iteration (root_finder.c:342-407, find_root_muller) as a
// step-wise state machine, so that one driver can advance the searches of many starts
// with one batched evaluation of f per round.
//
// Per start: f at za = (1 - 1e-4 i) z, zb = z, zc = (1 + 1e-4 i) z; then each iteration
// takes the quadratic through the three points (divided differences), the root of
// larger denominator, dz = 2 f(zc) / denom; it stops with failure on a zero or NaN
// denominator or when zc has left the open domain (domain_util.c:119-121), with success
// when |dz| < toler |zc| or f(zc) = 0, and otherwise damps the step: zn = zc - mu dz with
// mu = 1, 1/2, 1/8, 1/64, ... (mu /= 2, 4, 8, ...) while |f(zn)| > |f(zc)| and
// mu |dz| > toler |zc|.  At most maxit iterations; a failed start keeps its initial z.
#pragma once
#include <complex>
#include <vector>

namespace msg {

class MullerRun {
public:
    using cd = std::complex<double>;
    MullerRun(cd z0, double xmin, double xmax, double ymin, double ymax, int maxit, double toler);
    // The points at which f is wanted next (empty once finished), then their values.
    std::vector<cd> points() const;
    void take(const std::vector<cd>& f);
    bool finished() const { return state_ == kDone; }
    bool converged() const { return conv_; }
    cd root() const { return z_; }

private:
    enum { kInit, kDamp, kDone };
    void begin_iteration();
    double x0_, x1_, y0_, y1_, toler_;
    int maxit_, s_ = 0, state_ = kInit;
    bool conv_ = false;
    cd z_, za_, zb_, zc_, fa_, fb_, fc_, dz_, zn_;
    double absfc_ = 0.0, absdz_ = 0.0, tol_ = 0.0, mu_ = 1.0, pow2_ = 2.0;
    bool in_open(const cd& z) const { return x0_ < z.real() && z.real() < x1_ && y0_ < z.imag() && z.imag() < y1_; }
};

}  // namespace msg
