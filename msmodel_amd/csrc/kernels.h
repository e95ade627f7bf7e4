// kernels.h -- launch wrappers of the HIP kernels (kern_*.hip), called by the
// host engine.  All matrices are column-major interleaved complex, ld = np.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include "dmath.h"

namespace msg {

// One matrix of a batch whose matrices may come from different media (same padded
// size np): where its scatterers are, its wave number and its diagonal.
struct MatItem {
    const double* pos;     // device, row-major [n][d]
    int n, d;
    ScModel m;
    cx<double> k;
    cx<double> diag0;      // F(k)^-1                       (filled by launch_item_diag)
    cx<double> diag1;      // (F^-1(k+h) - F^-1(k-h)) / 2h  (filled by launch_item_diag)
};

// kern_assemble.hip
void launch_item_diag(MatItem* items, int cnt, hipStream_t st);
// phi[b np + i] = e^{i k_b x_i0} for i < n_b, 0 below np (the plane wave along 1x)
void launch_plane_items(const MatItem* items, int np, cx<double>* phi, int nbatch, hipStream_t st);
// kind 0: M, kind 1: dM/dk, from items[0..nbatch); dall: the dimension d shared by every
// item, or 0 when they differ (dall = 3 with kind 0 launches the kernel specialised for d = 3);
// lower: the lower triangle only (the diagonal tiles in full; the rest is left as it was)
template <class T>
void launch_assemble_items(const MatItem* items, int np, int kind, cx<T>* out, long long mstride, int nbatch,
                           hipStream_t st, int dall = 0, bool lower = false);
// out[b] = tr(M_b^-1 M_b') from V_b, the row-major factor's solve against the identity
// (M^-1 = P V, perm from launch_build_perm), with M' generated on the fly; inv (np ints
// per matrix) receives the inverse permutation, partial dk_contract_partials(np) sums per
// matrix
size_t dk_contract_partials(int np);
template <class T>
void launch_dk_contract(const cx<T>* V, long long sv, int np, const MatItem* items, const int* perm, int* inv,
                        cx<double>* partial, cx<double>* out, int nbatch, hipStream_t st);
void launch_diag(const cx<double>* ks, int nk, ScModel m, int d, int kind, cx<double>* diag, cx<double>* izero,
                 hipStream_t st);
template <class T>
void launch_assemble(const double* pos, int n, int d, int np, const cx<double>* ks, const cx<double>* diag,
                     const cx<double>* izero, int kind, cx<T>* out, long long mstride, int nbatch, hipStream_t st);
void launch_incident(const double* pos, int n, int d, int np, cx<double> k, int source, double ct, double st_,
                     cx<double>* phi, hipStream_t st);
// psi at npt points (row-major [npt][d]), one workgroup per point (small batches)
void launch_wave_points(const double* pos, int n, int d, const cx<double>* a, cx<double> k, int source, int npt,
                        const double* pts, cx<double>* psi, hipStream_t st);
// Persistent trace-log by the pivot-free LDL^T (kern_persist.hip): nwg workgroups take
// items[0..nitems) from *counter (0 at launch) and each runs an item's assembly,
// factorisation and reduction in its own workspace matrix (ws + w * wstride, at least
// Np^2 elements for the largest item; accws + w * astride: Np doubles); out[i] =
// sum log D_ii / N, growth[i] = the bits of its largest |L|, rho[i] = its error estimate
// sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r|.  d3: every item is d = 3.
int persist_ldlt_blocks_per_cu(bool d3);
void launch_ldlt_persist(const MatItem* items, int nitems, bool d3, cx<double>* ws, long long wstride, double* accws,
                         long long astride, int nwg, int* counter, cx<double>* out, unsigned long long* growth,
                         double* rho, hipStream_t st);
// One request of the resident psi kernel (a slot of its ring in coherent pinned host
// memory): the host fills the request fields and then publishes seq = index + 1; the
// device writes psi and then done = index + 1.
struct WaveSlot {
    const double* pos;          // device copies of the positions [n][d] and amplitudes [n]
    const cx<double>* a;
    double x[4];                // the point
    cx<double> k;
    int n, d, source, pad_;
    unsigned long long seq;
    cx<double> psi;
    unsigned long long done;
    unsigned long long pad2_[2];
};
static_assert(sizeof(WaveSlot) == 128, "one slot per 128 bytes");
// Workgroup w serves requests w, w + nwg, ... from resume[w] until *stop, an idle
// launch (idle_ticks of the wall clock without a request served) or life_ticks; then
// resume[w] holds its next request.  *closing and *last must be 0 at launch.
void launch_wave_resident(WaveSlot* ring, int nring, const unsigned* stop, unsigned long long* resume,
                          unsigned* closing, unsigned long long* last, long long idle_ticks, long long life_ticks,
                          int nwg, hipStream_t st);
// pts == nullptr: rows row0 .. row0+ny-1 of a grid nx wide (psi holds those rows);
// otherwise nx points given row-major [nx][d] (ny and row0 ignored)
void launch_wave(const double* pos, int n, int d, const cx<double>* a, cx<double> k, int source, int nx, int ny,
                 int row0, double xmin, double dx, double ymax, double dy, const double* pts, cx<double>* psi,
                 hipStream_t st);

// kern_lu.hip
// rm: the matrices are stored row-major (see at<> in kern_lu.hip); variant: panel
// shape index < kPanelVariants (0 = default for the size range; 5-7 the left-looking forms of
// shapes 3, 0 and 4)
constexpr int kPanelVariants = 8;
template <class T>
void launch_panel(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb, int nbatch,
                  hipStream_t st, bool rm = false, int variant = 0, int xa = 0, int xb = 0);
// (xa, xb): columns outside the panel (at most 64) to which the panel's interchanges are
// applied at the end of the same launch (xb <= xa: none)
// variant: tile shape index < swap_variants() (0 = 256-column tiles, 8-column rounds)
template <class T>
void launch_swap_trsm(cx<T>* A, int np, long long mstride, const int* pmap, int k0, int nb, int apply_left,
                      int nbatch, hipStream_t st, int variant = 0, bool rm = false);
int swap_variants();
// Row map of panel k0 on explicit column ranges: left tiles cover [lb, k0) (swap only),
// right tiles [rb, re) (swap if do_swap, then U12 = L11^-1 A12 if do_trsm).
// k0b >= 0: apply the interchanges of panel k0 and then those of panel k0b (> k0) as one
// composed row map (panel k0b's rows are all >= k0b, so its interchanges commute with
// the TRSM of rows [k0, k0 + nb))
struct SwapRange {
    int lb, nleft_t, rb, re, do_swap, do_trsm, k0b;
};
template <class T>
void launch_swap_range(cx<T>* A, int np, long long mstride, const int* pmap, int k0, int nb, int lb, int rb, int re,
                       int do_swap, int do_trsm, int nbatch, hipStream_t st, int variant = 0, bool rm = false,
                       int k0b = -1);
template <class T>
void launch_gemm_sub(cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int n, int k, int lda, int ldb, int ldc,
                     long long sC, long long sA, long long sB, int nbatch, hipStream_t st);
// C -= A B on the tiles on and below the diagonal of the square m x m C only (the
// trailing update of the symmetric factorisation)
template <class T>
void launch_gemm_sub_lower(cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int k, int lda, int ldb, int ldc,
                           long long sC, long long sA, long long sB, int nbatch, hipStream_t st);
// Pivot-free LDL^T of symmetric matrices (kern_ldlt.hip), column-major, ld np: the 64 x 64
// diagonal block at k0 becomes L11 (strictly lower part) and D11 (diagonal), zeros above;
// then the rows below get L21 = A21 L11^-T D11^-1 and the rows k0 .. k0+63 right of the
// block U12 = D11 L21^T.
// growth[b] (bits of a non-negative double) = the largest |L| seen; acc (np doubles per
// matrix, zero before the first block) and rho[b] (zero before the first block) carry the
// error estimate rho[b] = sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r|.
template <class T>
void launch_ldlt_diag(cx<T>* A, int np, long long ms, int k0, unsigned long long* growth, const double* acc,
                      double* rho, int nbatch, hipStream_t st);
template <class T>
void launch_ldlt_rows(cx<T>* A, int np, long long ms, int k0, unsigned long long* growth, double* acc, int nbatch,
                      hipStream_t st);
// n of matrix b is items[b].n when items is given, else n; ipiv null: no interchanges
template <class T>
void launch_tracelog(const cx<T>* A, int n, const MatItem* items, int np, long long mstride, const int* ipiv,
                     cx<double>* out, double* stats, int nbatch, hipStream_t st);
void launch_build_perm(const int* pmap, int np, int nb, int* perm, int nbatch, hipStream_t st);
template <class T>
void launch_gather_rows(cx<T>* dst, const cx<T>* src, const int* perm, int np, int ncol, long long sdst,
                        long long ssrc, int nbatch, hipStream_t st);
// rm = false: unit lower / non-unit upper (column-major LU); rm = true: non-unit lower /
// unit upper (row-major LU of a symmetric matrix read column-major)
template <class T>
void launch_trsm_block(const cx<T>* LU, cx<T>* B, int np, int ncol, int j0, int upper, long long slu, long long sb,
                       int nbatch, hipStream_t st, bool rm = false);
// x[i] = v, i < n
void launch_fill_int(int* x, int v, int n, hipStream_t st);
// every matrix of the batch set to the identity (not for captured graphs: a memset)
template <class T>
hipError_t launch_set_identity(cx<T>* A, int np, long long ms, int nbatch, hipStream_t st);
// inverse power iteration: v <- y / fac, fac = |y| y_i0/|y_i0|, delta = |v_new - v_old|,
// done once delta < toler (vectors of ld entries, n used; done vectors untouched)
void launch_power_step(cx<double>* V, cx<double>* Y, int n, int ld, double toler, int* done, cx<double>* fac,
                       double* delta, int nbatch, hipStream_t st);
// dst[perm[i], c] = src[i, c]
template <class T>
void launch_scatter_rows(cx<T>* dst, const cx<T>* src, const int* perm, int np, int ncol, long long sdst,
                         long long ssrc, int nbatch, hipStream_t st);
// out[b] = sum_{i<n} X[i, perm[i]] (perm null: X[i, i])
template <class T>
void launch_diag_sum(const cx<T>* X, int n, const MatItem* items, int ld, long long sx, cx<double>* out, int nbatch,
                     hipStream_t st, const int* perm = nullptr);
template <class T1, class T2>
void launch_convert(const cx<T1>* x, cx<T2>* y, long long count, hipStream_t st);
void launch_residual(const cx<double>* M, const cx<double>* a, const cx<double>* phi, int np, cx<double>* r,
                     hipStream_t st);
template <class T>
void launch_axpy(cx<double>* a, const cx<T>* d, int n, hipStream_t st);
template <class T>
void launch_apply_map(cx<T>* X, int np, int ncol, const int* pmap, int nb, int q, long long sx, int nbatch,
                      hipStream_t st);
template <class T>
void launch_swap_left(cx<T>* A, int np, long long mstride, const int* pmap, int k0, int nb, int nbatch,
                      hipStream_t st);
template <class T>
void launch_pad_identity(cx<T>* A, int n, int np, int nbatch, hipStream_t st);

// pmap ints per matrix for a given (np, nb)
inline size_t pmap_ints(int np, int nb) { return (size_t)(np / nb) * (1 + 4 * (size_t)nb); }
// The factorisation workspace of a batch: the row maps of every matrix, then (16-byte
// aligned) one 64 x 128 complex128 block per matrix for the 128-column blocking.
constexpr size_t kBlockWInts = 64 * 128 * 4;
inline size_t wblock_offset(int np, int nb, int nbatch) { return ((size_t)nbatch * pmap_ints(np, nb) + 63) / 64 * 64; }
inline size_t factor_ws_ints(int np, int nb, int nbatch) { return wblock_offset(np, nb, nbatch) + (size_t)nbatch * kBlockWInts; }
// Wb of the second panel of the 128-column block at K0 (see kern_lu.hip)
template <class T>
void launch_block_w(const cx<T>* A, int np, long long mstride, int K0, cx<T>* W, int nbatch, hipStream_t st, bool rm);

}  // namespace msg
