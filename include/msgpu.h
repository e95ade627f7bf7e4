/*
 * msgpu.h -- C ABI of the MI355X Foldy-Lax library (libmsgpu.so).
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements that keep MSModel's C signatures and semantics
 *     (host buffers in and out, input matrices destroyed as LAPACK does, LAPACK
 *     errors printed with the "[LAPACK]" tag and never fatal).  A program that
 *     links libmsgpu.so ahead of the reference objects
 *     runs its matrix work on the GPU unchanged.  Each declaration cites the
 *     reference declaration it replaces (paths relative to the MSModel tree).
 *
 *  2. Batched, device-resident entry points (msgpu_*) that the task drivers
 *     of ms_core.c call with all k points of a map, all Newton starts, or a
 *     whole psi grid at once.  These are the fast path.
 *
 * Complex numbers are C99 "double _Complex" (interleaved re, im FP64), the
 * reference's dcomplex (dcomplex_type.h:16).  Structures are laid out exactly
 * as the reference's (same member order, types and enum numbering), under
 * prefixed names so this header can be included beside the reference headers.
 */
#ifndef MSGPU_H
#define MSGPU_H
#include <stdint.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> msgpu_cplx;   /* same layout and return ABI as double _Complex */
extern "C" {
#else
#include <complex.h>
typedef double _Complex msgpu_cplx;
#endif

/* ------------------------------------------------------------------ types */
/* Urandom, nru_random_type.h:16-21 */
typedef struct { uint64_t u, v, w, seed; } msgpu_urandom;
/* ScatteringType, scattering_model_type.h:17 */
enum { MSGPU_HARDSPHERE = 0, MSGPU_SOFTSPHERE = 1, MSGPU_RESONANT = 2, MSGPU_MAXIMUM = 3 };
/* ScatteringModel, scattering_model_type.h:22-34 */
typedef struct {
    int type;
    int np;
    double* p;
    char dir[50];
    char name[150];
    char param_ascii[150];
    char param_latex[150];
} msgpu_scmodel;
/* Shape, medium_type.h:19 */
enum { MSGPU_CUBE = 0, MSGPU_BALL = 1, MSGPU_LATTICE = 2, MSGPU_GAUSSIAN = 3 };
/* Medium, medium_type.h:24-36 */
typedef struct {
    int d;
    int n;
    int ntot;
    int shape;
    double ratio;
    msgpu_scmodel* scmodel;
    double alpha;
    msgpu_urandom* rnd;
    double* pos;
} msgpu_medium;
/* IncidentWaveSource / IncidentWave, incident_wave_type.h:17-31 */
enum { MSGPU_PLANE = 0, MSGPU_SPHERICAL = 1 };
typedef struct {
    int source;
    msgpu_cplx k;
    char desc_ascii[80];
    char desc_latex[80];
} msgpu_iwave;

/* Arithmetic of the batched entry points. */
enum {
    MSGPU_FP64 = 0,    /* complex128 everywhere                                      */
    MSGPU_MIXED = 1,   /* complex64 LU on MFMA; FP64 re-evaluation of flagged cells  */
    MSGPU_FP32 = 2,    /* complex64 LU only (error reported, not corrected)          */
    MSGPU_MIXED_GATED = 3, /* complex64 LU; FP64 where its pivot growth could exceed tol */
    MSGPU_FP64_LDLT = 4    /* complex128; pivot-free LDL^T where its error estimate stays
                            * small, pivoted LU elsewhere (trace-log batches; complex128
                            * in the other entry points)                               */
};

/* ------------------------------------------------- 1. drop-in (reference ABI) */
/* medium_util.h:55-64 and medium_util.c:36 (init_medium is undeclared there). */
void   init_medium(msgpu_medium* med, int d, int n, msgpu_scmodel* scmodel, int shape, double ratio);
void   init_copy_medium(msgpu_medium* src, msgpu_medium* dst);
void   del_medium(msgpu_medium* med);
double radius_of_ball(msgpu_medium* med);
void   fill_medium(msgpu_medium* med, uint64_t seed);
double atom_distance(msgpu_medium* med, int i, int j);
msgpu_cplx invf(msgpu_medium* med, msgpu_cplx k);                       /* medium_util.h:64 */
/* scattering_model_util.h:17 */
int    init_scmodel(msgpu_scmodel* scmodel, int type, int np, double* p);

/* ms_matrix_util.h:17-30 */
void       build_ms_matrix(msgpu_medium* med, msgpu_cplx k, msgpu_cplx* matrix);
void       build_normalized_matrix(msgpu_medium* med, msgpu_cplx k, msgpu_cplx* matrix);
void       build_dk_ms_matrix(msgpu_medium* med, msgpu_cplx k, msgpu_cplx* matrix);
void       build_plane_wave(msgpu_medium* med, msgpu_cplx k, double theta, msgpu_cplx* vec);
/* tracelog_lu: callable from any number of threads at once; calls of the same n that
 * wait together share one batched factorization and each returns what it returns alone.
 * The matrix is overwritten with zgetrf_'s LAPACK-format factor, as the reference's is;
 * with the environment variable MSGPU_TRACELOG_KEEP_INPUT=1 it is left as given, which
 * saves a third of the host-link traffic (the reference's callers never read it). */
msgpu_cplx tracelog_lu(int n, msgpu_cplx* matrix);
msgpu_cplx dk_log_det_ms(msgpu_medium* med, msgpu_cplx k);
msgpu_cplx traceinv(int n, msgpu_cplx* matrix);
int        inverse_iteration(int n, msgpu_cplx* a, msgpu_cplx* vec, msgpu_cplx* lambda, int maxit, double toler,
                             int verb);
double     total_cross_section(msgpu_medium* med, double k);

/* incident_wave_util.h:18-19.  eval_total_wave reads pos and a at each call, as the
 * reference does; their device copy is kept per content (compared byte for byte).  The
 * calls (msmain's pixel loop calls it once per pixel from every thread,
 * ms_core.c:629-640) are served by a resident kernel that polls a ring of request slots
 * in pinned host memory: about 10 us per call, 1e6 calls/s from 16 threads at N = 1000.
 * The kernel ends itself 20 ms after its last request (or after 2 s) and is relaunched by
 * the next call.  MSGPU_WAVE_RESIDENT=0 selects instead one launch per batch of the
 * concurrent calls. */
msgpu_cplx eval_incident_wave(msgpu_iwave* iwave, int d, double* pos);
msgpu_cplx eval_total_wave(msgpu_iwave* iwave, msgpu_medium* med, msgpu_cplx* a, double* pos);

/* LAPACK Fortran ABI (LP64) as declared at ms_matrix_util.c:137-142 and ms_core.c:47. */
void zgetrf_(int* m, int* n, msgpu_cplx* a, int* lda, int* ipiv, int* info);
void zgetrs_(char* trans, int* n, int* nrhs, msgpu_cplx* a, int* lda, int* ipiv, msgpu_cplx* b, int* ldb,
             int* info);
void zgesv_(int* n, int* nrhs, msgpu_cplx* a, int* lda, int* ipiv, msgpu_cplx* b, int* ldb, int* info);

/* ---------------------------------------------- 2. batched / device-resident */
/* Library set-up; called implicitly by every entry point.  Returns 0 on success. */
int msgpu_init(void);
/* Last error message of the calling thread ("" if none). */
const char* msgpu_last_error(void);
/* Kernel timing (no counterpart in the reference; measurement hook for bench.py):
 * between begin and end every batched launch is bracketed by HIP events on its own
 * stream.  end() fills, per category c < ncat (0 panel, 1 swap+trsm, 2 gemm,
 * 3 assembly, 4 tracelog), the summed milliseconds, algorithmic flops and bytes and
 * the launch count. */
int msgpu_profile_begin(void);
int msgpu_profile_end(double* ms, double* flops, double* bytes, int* count, int ncat);

/* Autotuner (no counterpart in the reference): times the batched trace-log of nitems
 * matrices of size N (synthetic d=3 unit-density medium) for each blocking (64 or 128
 * columns) and storage order of the factorisation, swap + TRSM tile shape, panel shape,
 * chunk size, stream count and HIP-graph replay of the chunk pipeline, and keeps the fastest for N's padded size and
 * the mode's precision (complex128 for MSGPU_FP64, complex64 otherwise); later calls
 * use it.  report (may be NULL) = {swap variant, chunk, streams, best ms, row-major,
 * panel variant, graph replay, 128-column blocking}.
 * msgpu_tune_save/_load write/read the table as text; MSGPU_TUNE_FILE is loaded at
 * initialisation.  _load returns the number of entries read. */
int msgpu_autotune(int n, int mode, int nitems, double* report);
int msgpu_tune_save(const char* path);
int msgpu_tune_load(const char* path);

/* Trace-log of M(k) for nk wave numbers of one medium: out[l] = tracelog_lu(M(ks[l]))
 * as in run_tracelog_omp (ms_core.c:122-127).  mode = MSGPU_FP64/MIXED/FP32.
 * For MSGPU_MIXED, cells with N*Re(tracelog) < max_l N*Re(tracelog) + ln(reeval_rel)
 * (|det| below reeval_rel of the map maximum) and non-finite cells are recomputed in
 * FP64 (reeval_rel<=0 disables the first test); the other cells keep the complex64
 * value (typical per-N error 1e-7..1e-6 on dense media).  flags[l] (may be NULL) = 1
 * for re-evaluated cells.  Returns the number of re-evaluated cells (>= 0) or <0 on
 * error (message in msgpu_last_error).
 * For MSGPU_MIXED_GATED, reeval_rel is instead the tolerance tol on |d tracelog| (<= 0:
 * 1e-8): every cell is factored in complex64 and recomputed in FP64 where the error bound
 * 16 eps32 exp(max_i log|U_ii| - min_i log|U_ii|) of its complex64 factor (pivot growth
 * across the diagonal; eps32 = 2^-24) exceeds tol, or where the result is not finite.
 * The bound held for every one of 11520 cells measured on five media (d=2,3, N=200-1000;
 * the largest error was 13.8 eps32 exp(...)); since exp(...) >= 1, tol below 16 eps32 ~
 * 1e-6 recomputes every cell. */
int msgpu_tracelog_batch(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out, int mode,
                         double reeval_rel, int* flags);
/* Allocation-free variant of msgpu_tracelog_batch, mode MSGPU_FP64 or MSGPU_FP32.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory
 * (code under HIP-graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc/hipFree or allocating library routine: every device buffer lives
 * in one static 256 MiB device array per GPU, held until the process ends.  Limit: one
 * matrix and its workspace must fit in it, N <= 4032 (complex128) or N <= 5760
 * (complex64); larger N returns -5 with a message.  Calls serialise on one lock per
 * process (they share each GPU's array).  Results equal msgpu_tracelog_batch's bit for
 * bit.  Returns 0, or < 0 on error (message in msgpu_last_error). */
int msgpu_tracelog_batch_noalloc(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out, int mode);
/* As msgpu_tracelog_batch in FP64 or FP32 (no re-evaluation), with the pivot statistics
 * stats[2l] = max_i log|U_ii|, stats[2l+1] = min_i log|U_ii| of cell l.  With
 * MSGPU_FP64_LDLT every cell is the pivot-free LDL^T (scatterers in order of increasing
 * distance from their centroid; no LU fallback) and stats[2l] = max |L|,
 * stats[2l+1] = rho = sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r|, whose product with the
 * double epsilon estimates the rounding error of log det (msgpu_tracelog_batch keeps the
 * LDL^T of a cell when rho <= 1e7).  No counterpart in the reference; a conditioning
 * indicator for the caller. */
int msgpu_tracelog_stats(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out, double* stats,
                         int mode);
/* d/dk log det M(k) = tr(M^-1 M') for nk wave numbers (ms_matrix_util.c:242-263). */
int msgpu_dk_log_det_batch(const msgpu_medium* med, int nk, const msgpu_cplx* ks, msgpu_cplx* out, int mode);
/* Foldy-Lax amplitudes a = M(k)^-1 phi with phi the incident wave at the scatterers
 * (ms_core.c:616-620).  mode MSGPU_FP64: complex128 LU; MSGPU_MIXED: complex64 LU
 * with FP64 iterative refinement until ||M a - phi|| <= tol ||phi|| (tol<=0: 1e-12).
 * When the refinement stalls (the residual fails to halve in two consecutive steps,
 * as near a resonance where cond(M) eps32 >~ 1) or has not met tol after 30 steps,
 * the amplitudes are recomputed with the complex128 LU, which counts as one more step.
 * Returns the number of steps (>=0) or <0 on error. */
int msgpu_solve(const msgpu_medium* med, const msgpu_iwave* iwave, msgpu_cplx* a, int mode, double tol);
/* Roots of det M(k) in the open domain (xmin, xmax) x (ymin, ymax) by the damped Muller
 * method (find_root_muller, root_finder.c:342-407) on f(k) = smallest eigenvalue of M(k)
 * (MSGPU_ROOT_MINEIGVAL, find_eigenstates_serial_2, ms_core.c:335-367) or f(k) =
 * 1/tr(M(k)^-1) (MSGPU_ROOT_INVTRACEINV, find_eigenstates_serial_3, ms_core.c:374-401),
 * all starts advanced together (one batched evaluation of f per round).  On entry
 * z[0..*nz) holds the starts; on exit the converged, distinct roots (filter_cvector,
 * uniq_cvector with relative tolerance 1e-14) and *nz their count.  For MINEIGVAL each
 * start carries its own inverse-iteration vector from one evaluation to the next (the
 * reference shares one vector across its sequential loop over starts).  Returns the
 * number of rounds, or < 0 on error. */
enum { MSGPU_ROOT_MINEIGVAL = 1, MSGPU_ROOT_INVTRACEINV = 2 };
int msgpu_roots_muller(const msgpu_medium* med, double xmin, double xmax, double ymin, double ymax, int* nz,
                       msgpu_cplx* z, int maxit, double toler, int method);
/* Random functions and cross sections of many (medium, real k) items in one batch, for
 * the task drivers run_random_function_omp (ms_core.c:476-519) and
 * run_differential_cross_section_omp (ms_core.c:526-587); the incident wave is the unit
 * plane wave along 1x.  item i is (meds[med_of[i]], ks[i]) (med_of may be NULL: medium 0).
 *   MSGPU_CRSEC: out[i] = total cross section -Im(phi^H M^-1 phi)/k (total_cross_section,
 *                ms_matrix_util.c:344-366)
 *   MSGPU_DOS:   out[i] = density of states -Im tr(M^-1 M')/(2 pi k) (ms_core.c:503-505)
 *   MSGPU_DCS:   out[i*nang + t] = (k/2pi)^(d-3)/(16 pi^2) |b_t^H M^-1 phi|^2 with b_t the
 *                plane wave at theta_deg[t] degrees (build_plane_wave, ms_matrix_util.c:88-106)
 * Returns 0, or < 0 on error. */
enum { MSGPU_CRSEC = 0, MSGPU_DOS = 1, MSGPU_DCS = 2 };
int msgpu_scattering_multi(int nmed, const msgpu_medium* const* meds, int nk, const int* med_of, const double* ks,
                           int type, int nang, const double* theta_deg, double* out);
/* Eigenstates at resonances (replaces the per-root loop of compute_eigenstates,
 * ms_core.c:296-313, over inverse_iteration, ms_matrix_util.c:279-314): for each root
 * k_r, the eigenvector of M(k_r) of smallest eigenvalue by inverse power iteration from
 * the constant unit vector, at most maxit steps, stopping when |v - v_old| < toler; all
 * roots are factored and iterated in one batch.  vecs: nroot x N (root-major); mu[r] =
 * the eigenvalue (1/fac of the last step); conv[r] = 1 if converged.  suspect (may be
 * NULL) holds the flags of check_eigenstate_list (eigenstate_list_util.c:130-163):
 * |mu| > 1e-6, or an earlier root within 1e-6 |k_r|.  Returns the number of suspect
 * states, or < 0 on error. */
int msgpu_eigenstates(const msgpu_medium* med, int nroot, const msgpu_cplx* kroot, int maxit, double toler,
                      msgpu_cplx* vecs, msgpu_cplx* mu, int* conv, int* suspect);
/* psi(r) = phi(r) + sum_i a_i G(r|x_i) on an nx x ny grid of the (x0,x1) plane,
 * row-major from the top-left corner (wave_function_util.c:76-81). */
int msgpu_wavefunction_grid(const msgpu_medium* med, const msgpu_iwave* iwave, const msgpu_cplx* a, int nx,
                            int ny, double xmin, double xmax, double ymin, double ymax, msgpu_cplx* psi);

/* Batches over several media.  No single counterpart in the reference: its task
 * drivers loop over seeds and media one matrix at a time (run_tracelog_omp,
 * ms_core.c:96-148; find_eigenstates_serial_1, ms_core.c:320-327).  Item i is the
 * wave number ks[i] on medium meds[med_of[i]] (med_of may be NULL: all on meds[0]).
 * Media may differ in N, d and scattering model; items are grouped by padded size
 * and each result is the one the single-medium call gives for that item.
 * MSGPU_MIXED re-evaluates, per medium, the cells below reeval_rel of that medium's
 * largest |det| in the batch. */
int msgpu_tracelog_multi(int nmed, const msgpu_medium* const* meds, int nk, const int* med_of,
                         const msgpu_cplx* ks, msgpu_cplx* out, int mode, double reeval_rel, int* flags);
int msgpu_dk_log_det_multi(int nmed, const msgpu_medium* const* meds, int nk, const int* med_of,
                           const msgpu_cplx* ks, msgpu_cplx* out, int mode);
/* Resonances of several media at once (the uneven-batch search): medium m searches
 * the open rectangle doms[4m..4m+3] = {xmin, xmax, ymin, ymax} from nz[m] starts
 * stored in z at offset sum_{j<m} nz[j]; on return nz[m] roots of medium m are at
 * that same offset, sorted by real part.  Every sweep evaluates f'/f for the live
 * starts of all media in one batched call per precision.  Returns the sweep count. */
int msgpu_resonances_multi(int nmed, const msgpu_medium* const* meds, const double* doms, int* nz, msgpu_cplx* z,
                           int nexpsup, int maxit, double toler, int mode);

/* Resonances: Maehly-Aberth-Ehrlich on det M(k) (root_finder.c:522-577) with every
 * live start's f'/f evaluated in one batched launch per sweep.  z[0..*nz) holds the
 * starts on entry and the converged roots (sorted by real part) on exit.
 * mode MSGPU_MIXED iterates in complex64 then polishes every root in FP64. */
int msgpu_resonances(const msgpu_medium* med, double xmin, double xmax, double ymin, double ymax, int* nz,
                     msgpu_cplx* z, int nexpsup, int maxit, double toler, int mode);

/* k-point generators of the reference's k-plane tasks, bit for bit: the nx*ny points of a
 * complex map (get_arg_cmap, complex_map_util.c:27-31), the 3(ns+2) points of a complex
 * cut (get_arg_ccut, complex_cut_util.c:32-36) and the ns points of a log cut
 * (init_logcut, log_cut_util.c:72-101), in their storage order.  Each returns the number
 * of points written, or < 0 on bad arguments. */
int msgpu_kgrid_map(int nx, int ny, double xmin, double xmax, double ymin, double ymax, msgpu_cplx* k);
int msgpu_kgrid_cut(int ns, msgpu_cplx zmin, msgpu_cplx zmax, msgpu_cplx* k);
int msgpu_kgrid_logcut(int ns, double rek, double imkmin, double imkmax, msgpu_cplx* k);

/* ------------------------------------------------------- 3. serving loop */
/* No counterpart in the reference (msmain runs one task at a time).  A server admits
 * log-det tiles, Foldy-Lax solves (+ psi grid) and resonance searches of mixed N and d
 * as they arrive, by device footprint (16 N^2 bytes per complex128 matrix, 8 N^2 per
 * complex64 one, times the matrices a request keeps live), keeps each admitted
 * request's positions on the device, and batches the current stage of every admitted
 * log-det and resonance request into shared multi-medium launches (continuous
 * batching); solves run concurrently on their own thread.  Inputs are copied at
 * submit; output buffers must stay valid until msgpu_server_wait returns.  A result
 * does not depend on the other requests in flight. */
enum { MSGPU_REQ_LOGDET = 0, MSGPU_REQ_SOLVE = 1, MSGPU_REQ_RESONANCE = 2 };
typedef struct {
    int kind;                   /* MSGPU_REQ_*                                               */
    const msgpu_medium* med;    /* positions and model are copied at submit                  */
    int mode;                   /* MSGPU_FP64 / MSGPU_MIXED / MSGPU_FP32                      */
    /* LOGDET: out[l] = tracelog(M(ks[l])), l < nk, as msgpu_tracelog_batch            */
    int nk;
    const msgpu_cplx* ks;
    double reeval_rel;
    msgpu_cplx* out;
    /* SOLVE: a[N] as msgpu_solve; psi[nx*ny] as msgpu_wavefunction_grid when nx*ny > 0 */
    msgpu_cplx k;
    int source;
    double tol;
    msgpu_cplx* a;
    int nx, ny;
    double xmin, xmax, ymin, ymax;
    msgpu_cplx* psi;
    /* RESONANCE: starts z[0..nz) in the open domain dom = {xmin, xmax, ymin, ymax}, as
     * msgpu_resonances; the roots go to roots[0..*nroots)                             */
    int nz;
    const msgpu_cplx* z;
    double dom[4];
    int nexpsup, maxit;
    double toler;
    msgpu_cplx* roots;
    int* nroots;
} msgpu_request;
typedef struct {
    int status;                 /* 0 ok, < 0 error (message in msgpu_last_error after wait)  */
    int count;                  /* LOGDET cells re-evaluated, SOLVE refinement steps,
                                   RESONANCE sweeps                                          */
    double queue_ms;            /* submit -> admitted                                        */
    double total_ms;            /* submit -> done                                            */
} msgpu_result;
typedef struct msgpu_server msgpu_server;
/* hbm_gb: admission budget in GiB; max_active: most requests admitted at once. */
msgpu_server* msgpu_server_create(double hbm_gb, int max_active);
/* Returns the request id (>= 0) or < 0 on a malformed request. */
long long msgpu_server_submit(msgpu_server* s, const msgpu_request* r);
/* Blocks until the request is done; returns its status.  Each id is waited for once. */
int msgpu_server_wait(msgpu_server* s, long long id, msgpu_result* res);
/* stats[0..nstats): steps, batched matrices, most requests in one step, peak admitted
 * bytes (kept factors included), budget bytes, complex128 solves that used a kept factor,
 * complex128 solves that factored M(k), bytes of kept factors now, resident media now. */
int msgpu_server_stats(msgpu_server* s, double* stats, int nstats);
/* Serves every submitted request, then stops the loops and frees the server. */
void msgpu_server_destroy(msgpu_server* s);

#ifdef __cplusplus
}
#endif
#endif
