// aberth.h -- the Maehly-Aberth-Ehrlich iteration on det M(k) as a step-wise state
// machine, so that one driver can advance many searches (media, requests) with one
// batched evaluation of f'/f = tr(M^-1 M') per sweep.
//
// Follows root_finder.c:522-577 (solve_aberth): starts closer than toler are merged
// and starts outside the open domain (domain_util.c:119-121) dropped; the exponential
// background f'/f ~ poly(z) is fitted by least squares to 6*nexpsup Halton samples of
// the domain (root_finder.c:161-172, Halton index i+1); every sweep moves each live
// root by dz = 1/(f'/f - poly(z) - sum_{j != i} 1/(z_i - z_j)); a root converges when
// |dz| < toler |z| (keeping its position) and is abandoned when its step leaves the
// domain; at most maxit sweeps, roots not converged by then are dropped.
// With mixed precision the sweeps use complex64 factors until every live step is
// below max(toler, 1e-5)|z| (at most maxit sweeps), then up to 10 complex128 sweeps
// polish to toler.
#pragma once
#include <complex>
#include <vector>

namespace msg {

class AberthRun {
public:
    using cd = std::complex<double>;
    AberthRun(const cd* starts, int nstart, double xmin, double xmax, double ymin, double ymax, int nexpsup,
              int maxit, double toler, bool mixed);
    // Phase 1: background fit.  fit_points() is empty when no fit is needed.
    std::vector<cd> fit_points() const;
    void take_fit(const std::vector<cd>& fvals);
    // Phase 2: sweeps.  sweep_points() lists the live roots (empty once finished) and
    // the precision they are to be evaluated in (1 = complex64, 0 = complex128).
    std::vector<cd> sweep_points(int* precision) const;
    void take_sweep(const std::vector<cd>& fvals);
    bool finished() const;
    int sweeps() const { return sweeps_; }
    // Converged roots sorted by real part.
    std::vector<cd> roots() const;

private:
    double x0_, x1_, y0_, y1_, toler_;
    int nexpsup_, maxit_;
    std::vector<cd> live_, poly_;
    std::vector<int> flag_;     // 0 live, 1 converged, -1 left the domain
    std::vector<int> act_;      // indices evaluated by the current sweep
    int prec_, sweeps_ = 0;
    bool mixed_, fitted_;
    bool in_open(const cd& z) const { return x0_ < z.real() && z.real() < x1_ && y0_ < z.imag() && z.imag() < y1_; }
};

}  // namespace msg
