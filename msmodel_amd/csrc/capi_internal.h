// capi_internal.h -- declarations shared by the host translation units of
// libmsgpu.so (medium.cpp, engine.cpp, capi.cpp).
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/msgpu.h"
#include "dmath.h"

#define MSGPU_API __attribute__((visibility("default")))

namespace msg {

// Cells of a complex64 trace-log map of one medium (N = n) to recompute in complex128:
// |det| below rel of the map's largest |det| (N Re t_l < max_l N Re t_l + ln rel; rel <= 0
// disables this test) and non-finite cells.  Returns their indices in increasing order.
inline std::vector<int> reeval_cells(int n, int nk, const std::complex<double>* t, double rel) {
    double best = -INFINITY;
    for (int l = 0; l < nk; l++)
        if (std::isfinite(t[l].real())) best = std::max(best, n * t[l].real());
    const double cut = rel > 0 ? best + std::log(rel) : -INFINITY;
    std::vector<int> idx;
    for (int l = 0; l < nk; l++)
        if (n * t[l].real() < cut || !std::isfinite(t[l].real())) idx.push_back(l);
    return idx;
}

// Cells of a complex64 trace-log map to recompute in complex128 under the pivot-growth
// gate: the error bound 16 eps32 exp(max log|U_ii| - min log|U_ii|) of the cell's complex64
// factor (stats[2l], stats[2l+1]) exceeds tol (<= 0: 1e-8), or the value is not finite.
constexpr double kGateFactor = 16.0;
inline double gate_bound(const double* st) { return kGateFactor * 5.9604644775390625e-08 * std::exp(st[0] - st[1]); }
inline std::vector<int> gated_cells(int nk, const double* stats, const std::complex<double>* t, double tol) {
    if (tol <= 0) tol = 1e-8;
    std::vector<int> idx;
    for (int l = 0; l < nk; l++) {
        const double b = gate_bound(stats + 2 * l);
        if (!(b <= tol) || !std::isfinite(t[l].real()) || !std::isfinite(t[l].imag())) idx.push_back(l);
    }
    return idx;
}

// medium.cpp
double ball_volume(int d);
double ball_radius(int d, int n);
void fill_positions(msgpu_medium* med, uint64_t seed);
ScModel scmodel_of(const msgpu_medium* med);

// What the engine needs to know about a medium.
struct MediumView {
    int d = 0, n = 0, np = 0;
    ScModel m{};
    const double* pos = nullptr;   // host, row-major [n][d]
    const double* dpos = nullptr;  // device copy of pos when the caller keeps one resident
};
MediumView view_of(const msgpu_medium* med);
inline int padded(int n) { return (n + 63) / 64 * 64; }

// engine.cpp -- every function returns 0 on success, <0 on error (message via set_error)
void set_error(const std::string& msg);
const char* last_error();
int engine_init();
// noalloc.hip: the device address of the allocation-free scratch on the current device
int noalloc_setup();
int profile_begin();
int profile_end(double* ms, double* flops, double* bytes, int* count, int ncat);

// Batched trace-log for nk wave numbers; precision 0 = FP64, 1 = FP32.
// stats (may be null) receives {max log|U_ii|, min log|U_ii|} per cell.
int tracelog_points(const MediumView& mv, int nk, const cx<double>* ks, cx<double>* out, double* stats,
                    int precision);
// The same over several media: item i is (meds[midx[i]], ks[i]) (midx null: medium 0).
// Items of different padded sizes are grouped; results come back in input order.
int tracelog_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                   cx<double>* out, double* stats, int precision);
int dk_log_det_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                     cx<double>* out, int precision);
// tr(M^-1 dM/dk) for nk wave numbers.
int dk_log_det_points(const MediumView& mv, int nk, const cx<double>* ks, cx<double>* out, int precision);
// Amplitudes a = M(k)^-1 phi; refine: complex64 LU + FP64 refinement to tol.  Returns #steps or <0.
int solve_amplitudes(const MediumView& mv, cx<double> k, int source, cx<double>* a, int mixed, double tol);
// A complex128 factor of one M(k) kept on the device: the row-major LU of the symmetric M
// (left interchanges applied) and the row map P of M^-1 v = P V.
struct ResidentFactor;
size_t resident_factor_bytes(const ResidentFactor& f);
// The complex128 solve of solve_amplitudes (same kernels and launch order, so the same
// amplitudes bit for bit).  With reuse, that factor is used: M(k) is neither assembled nor
// factored.  With keep, the factor used is handed back for later calls.  mv.dpos, when set,
// is a device copy of the positions.
int solve_resident(const MediumView& mv, cx<double> k, int source, cx<double>* a,
                   const std::shared_ptr<ResidentFactor>& reuse, std::shared_ptr<ResidentFactor>* keep);
int eigenstates(const MediumView& mv, int nr, const cx<double>* kr, int maxit, double toler, cx<double>* vecs,
                cx<double>* mu, int* conv, const cx<double>* v_init = nullptr);
int trace_inv_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                    cx<double>* out, int precision);
int amplitudes_multi(const std::vector<MediumView>& meds, int nk, const int* midx, const cx<double>* ks,
                     cx<double>* out);
// psi grid
int wave_grid(const MediumView& mv, cx<double> k, int source, const cx<double>* a, int nx, int ny, double xmin,
              double xmax, double ymin, double ymax, cx<double>* psi);
// Matrix assembly into a host buffer (n x n, ld n): kind 0 M, 1 dM/dk, 2 normalised.
int build_matrix_host(const MediumView& mv, cx<double> k, int kind, cx<double>* out);
int plane_wave_host(const MediumView& mv, cx<double> k, double theta_deg, cx<double>* out);
int incident_host(const MediumView& mv, cx<double> k, int source, cx<double>* out);
// Host matrix (n x n, ld lda) -> LU on the device -> written back in LAPACK format.
// Optionally solves nrhs right-hand sides b (ld ldb) in place and returns the tracelog.
int lu_host(int n, cx<double>* a, int lda, int* ipiv, int nrhs, cx<double>* b, int ldb, cx<double>* tracelog,
            int* info);
// tracelog_lu of one host matrix (n x n, ld lda; overwritten by its LAPACK-format
// factor), coalesced with the calls of other threads into one batched factorization.
int tracelog_host_coalesced(int n, cx<double>* a, int lda, int* ipiv, cx<double>* tracelog, int* info);
// Solve with an existing LAPACK-format LU from the host (zgetrs 'N').
int lu_solve_host(int n, const cx<double>* lu, int lda, const int* ipiv, int nrhs, cx<double>* b, int ldb);
// Autotuner (see engine.cpp): report = {swap variant, chunk, streams, best ms, row-major,
// panel variant, graph, 128-column blocking}.
int autotune(int np, int prec, int nitems, double* report);
int save_tuning(const char* path);
int load_tuning(const char* path);
// Device-resident copy of a medium's positions (released with release_positions).
int upload_positions(const MediumView& mv, const double** dpos, size_t* token);
void release_positions(const double* dpos, size_t token);
// psi at one point, coalesced with the concurrent calls of other threads (same content of
// positions, amplitudes, k and source) into one launch; the content's device copy is
// reused from call to call.
int wave_point_coalesced(const MediumView& mv, cx<double> k, int source, const cx<double>* a, const double* x,
                         cx<double>* psi);
// psi at arbitrary points (npt points, row-major [npt][d]).
int wave_points(const MediumView& mv, cx<double> k, int source, const cx<double>* a, int npt, const double* pts,
                cx<double>* psi);

}  // namespace msg
