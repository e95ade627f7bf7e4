/*
 * msoracle.h -- CPU restatement of the MSModel Foldy-Lax path, in plain C99.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in msmodel_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function restates the algorithm of the reference function cited next
 * to it (paths relative to the MSModel source tree).  All arithmetic is FP64.
 *
 * Pinning (how each function is checked; see tests/test_oracle_*.py):
 *   orc_rng_*          NR3 "Ran" bit streams vs the compiled reference   (fixtures)
 *   orc_fill_medium    positions bit-exact vs the compiled reference      (fixtures)
 *   orc_bessel_k       Mathematica K_nu table of the reference tests, scipy kv
 *   orc_green*         Mathematica G / I / dG tables of the reference tests,
 *                      closed forms (d=1,3), scipy hankel1 (d=2)
 *   orc_invf           closed forms (d=3 cot, d=2 Hankel/J0), reference fixtures
 *   orc_build_*        symmetry M=M^T, element formula, reference fixtures
 *   orc_getrf/getrs    numpy/LAPACK slogdet and solve, residuals, izamax ties
 *   orc_tracelog       Mathematica values of test_ms_matrix.c:52-57 (Re),
 *                      arg det mod 2pi vs numpy slogdet, N=2 closed form
 *   orc_dk_log_det     test_ms_matrix.c:99-103 (loose: FD diagonal), finite
 *                      difference of orc_tracelog
 *   orc_gesv           test_ms_matrix.c:229-231 scattering amplitudes, residual
 *   orc_total_wave     direct sum identity at scatterer-free points, fixtures
 *   orc_aberth         N=2 closed-form resonances (mpmath), det M(k*)~0,
 *                      root sets of the compiled reference (fixtures, d=2 and d=3)
 *   orc_roots_muller   root sets of the compiled reference's [eigenstate_list]
 *                      driver, mineigval and invtraceinv (fixtures)
 *   orc_inverse_iteration  Mathematica values of test_ms_matrix.c:137-138 (3x3),
 *                      the compiled reference's eigenvalues and eigenvectors at roots
 *   orc_kgrid_*        bit-exact vs the compiled reference's get_arg_cmap,
 *                      get_arg_ccut and init_logcut/get_arg_lcut (fixtures)
 */
#ifndef MSORACLE_H
#define MSORACLE_H
#include <complex.h>
#include <stdint.h>

typedef double complex orc_cplx;

/* Scattering models, same numbering as ScatteringType (scattering_model_type.h:17). */
enum { ORC_HARDSPHERE = 0, ORC_SOFTSPHERE = 1, ORC_RESONANT = 2, ORC_MAXIMUM = 3 };
/* Medium shapes, same numbering as Shape (medium_type.h:19). */
enum { ORC_CUBE = 0, ORC_BALL = 1, ORC_LATTICE = 2, ORC_GAUSSIAN = 3 };
/* Incident waves, same numbering as IncidentWaveSource (incident_wave_type.h:17). */
enum { ORC_PLANE = 0, ORC_SPHERICAL = 1 };

/* NR3 Ran generator; state = {u, v, w, seed} as Urandom (nru_random_type.h:16-21). */
void     orc_rng_init(uint64_t st[4], uint64_t seed);     /* nru_random.c:63-73 */
uint64_t orc_rng_next(uint64_t st[4]);                    /* nru_random.c:24-34 */
double   orc_rng_double(uint64_t st[4]);                  /* nru_random.c:39-42 */
double   orc_rng_normal(uint64_t st[4]);                  /* nru_random.c:48-58 */

double orc_uball_volume(int d);                           /* green_bessel.c:41-58 */
double orc_uball_surface(int d);                          /* green_bessel.c:63-80 */
/* Fills pos[n*d] (row-major [atom][coord]) as fill_medium (medium_util.c:164-193).
 * Returns the shape actually used (ball becomes cube when d==1, medium_util.c:57-60). */
int    orc_fill_medium(int d, int n, int shape, double ratio, uint64_t seed, double* pos);
double orc_radius_of_ball(int d, int n);                  /* medium_util.c:107-109 */

orc_cplx orc_bessel_k(int twonu, orc_cplx z);             /* green_bessel.c:224-246 */
orc_cplx orc_hyp0f1_reg(double a, orc_cplx z);            /* green_bessel.c:254-272 */
orc_cplx orc_green(int d, orc_cplx k, double r);          /* green_bessel.c:280-292 */
orc_cplx orc_green_imag(int d, orc_cplx k, double r);     /* green_bessel.c:298-305 */
orc_cplx orc_dk_green(int d, orc_cplx k, double r);       /* green_bessel.c:311-313 */
orc_cplx orc_invf(int model, const double* p, int d, orc_cplx k); /* scattering_model_util.c:101-169 */

/* Matrix assembly, column-major n x n (ms_matrix_util.c:32-81). */
void orc_build_matrix(int d, int n, const double* pos, int model, const double* p, orc_cplx k, orc_cplx* m);
void orc_build_dk_matrix(int d, int n, const double* pos, int model, const double* p, orc_cplx k, orc_cplx* m);
void orc_build_normalized_matrix(int d, int n, const double* pos, orc_cplx k, orc_cplx* m);
void orc_build_plane_wave(int d, int n, const double* pos, orc_cplx k, double theta_deg, orc_cplx* v); /* ms_matrix_util.c:88-106 */

/* Dense complex LU with partial pivoting, LAPACK zgetrf conventions (1-based ipiv,
 * pivot = first index of max |Re|+|Im|, info = first exactly-zero pivot). */
int  orc_getrf(int n, orc_cplx* a, int lda, int* ipiv);
void orc_getrs(int n, int nrhs, const orc_cplx* lu, int lda, const int* ipiv, orc_cplx* b, int ldb);
int  orc_gesv(int n, int nrhs, orc_cplx* a, int lda, int* ipiv, orc_cplx* b, int ldb);

orc_cplx orc_tracelog(int n, orc_cplx* a);                /* ms_matrix_util.c:185-202 */
orc_cplx orc_dk_log_det(int d, int n, const double* pos, int model, const double* p, orc_cplx k); /* ms_matrix_util.c:242-263 */
orc_cplx orc_traceinv(int n, orc_cplx* a);                /* ms_matrix_util.c:319-335 */
int      orc_inverse_iteration(int n, orc_cplx* a, orc_cplx* vec, orc_cplx* lambda, int maxit, double toler); /* ms_matrix_util.c:279-314 */
double   orc_total_cross_section(int d, int n, const double* pos, int model, const double* p, double k); /* ms_matrix_util.c:344-366 */

/* Incident and total wave (incident_wave_util.c:73-96). */
orc_cplx orc_incident(int source, orc_cplx k, int d, const double* x);
orc_cplx orc_total_wave(int source, orc_cplx k, int d, int n, const double* pos, const orc_cplx* a, const double* x);
/* psi on an nx x ny grid in the (x0,x1) plane, row-major from the top-left corner:
 * x0 = xmin + ix*dx, x1 = ymax - iy*dy (wave_function_util.c:76-81). */
void orc_wavefunction_grid(int source, orc_cplx k, int d, int n, const double* pos, const orc_cplx* a,
                           int nx, int ny, double xmin, double xmax, double ymin, double ymax, orc_cplx* psi);

/* k-grids (complex_map_util.c:27-31, complex_cut_util.c:32-36, log_cut_util.c:72-101). */
void orc_kgrid_map(int nx, int ny, double xmin, double xmax, double ymin, double ymax, orc_cplx* k);
void orc_kgrid_cut(int ns, orc_cplx zmin, orc_cplx zmax, orc_cplx* k);   /* 3 (ns + 2) points */
void orc_kgrid_logcut(int ns, double rek, double imkmin, double imkmax, orc_cplx* k);

/* Halton (2,3) start points, i = 1..n (domain_util.c:156-185). */
void orc_halton(int n, double xmin, double xmax, double ymin, double ymax, orc_cplx* z);

/* Maehly-Aberth-Ehrlich iteration with exponential-background suppression
 * (root_finder.c:522-577) on f(k)=det M(k), dlogf = orc_dk_log_det.
 * On exit *nz = number of converged roots, z[0..*nz) sorted by real part. */
void orc_aberth(int d, int n, const double* pos, int model, const double* p,
                double xmin, double xmax, double ymin, double ymax,
                int* nz, orc_cplx* z, int nexpsup, int maxit, double toler);

/* Roots of det M(k) by the damped Muller method (root_finder.c:342-407) on the smallest
 * eigenvalue of M(k) by inverse iteration from one vector shared by the starts in turn
 * (method 1, find_eigenstates_serial_2, ms_core.c:335-367) or on 1/tr M(k)^-1 (method 2,
 * find_eigenstates_serial_3, ms_core.c:374-401).  The starts z[0..*nz) are run in order;
 * on exit the converged, distinct (1e-14) roots in that order.  Returns their count.
 * Method 3 is method 1 with one vector per start instead of the shared one: every start
 * begins from the constant unit vector, its first three points (za, zb, zc) all start
 * from it, and each later evaluation from the vector of that start's previous point
 * (the order in which the batched msgpu_roots_muller evaluates, all starts at once). */
int orc_roots_muller(int d, int n, const double* pos, int model, const double* p, double xmin, double xmax,
                     double ymin, double ymax, int* nz, orc_cplx* z, int maxit, double toler, int method);

#endif
