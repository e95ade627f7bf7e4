// kern_lu.hip -- the trailing-update GEMM of the batched blocked LU (C -= A B on
// MFMA) and the kernels that consume the factorisation: trace-log reduction,
// permutations, blocked triangular solves, the inverse power step and small
// elementwise helpers.  The panel kernel is in kern_panel.h, the interchanges
// and U12 solves in kern_swap.hip.
//
// Per block step k0 (panel width nb):
//   lu_panel    one workgroup per matrix factors the Np x nb panel, S columns
//               at a time held in registers (row g in thread (g-c0) % PT).  Pivot
//               search: a DPP arg-max in each wave, then across the waves' keys
//               after one barrier per column.  Pivot rule = LAPACK izamax: largest
//               |Re|+|Im|, first index on ties.  The panel's interchanges are also
//               folded into a composite row map (<= 2nb rows) that the next kernel
//               applies as one gather.
//   swap_trsm   one workgroup per 64-column tile: applies the row map, and for
//               tiles right of the panel solves U12 = L11^-1 A12 in registers
//               (4 lanes per column, wave shuffles, no barriers).
//   gemm_sub    C -= A*B on MFMA (v_mfma_{f64,f32}_16x16x4), A and B read in their
//               natural column-major complex layout: three real products (3M) for
//               complex128 at K >= 128, else the real product of the interleaved C
//               (2m x n) with [[Ar,-Ai],[Ai,Ar]] (2m x 2k) and the interleaved B.
#include "lu_device.h"
#include "gemm_dev.h"
#include "kern_panel.h"

namespace msg {

// (three waves per SIMD: the register budget of the complex128 3M tile)
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void gemm_sub_kernel(
    cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int n, int kk, int lda, int ldb, int ldc, long long sC,
    long long sA, long long sB, int tiles_m, int tiles_n, int nbatch, int lower) {
    __shared__ __align__(16) T lds[KC * (2 * GT + PADA) + GT * (2 * KC + PADB)];
    // lower != 0: only the tiles on and below the diagonal of a square C (tiles_m == tiles_n)
    const int tiles = lower ? tiles_m * (tiles_m + 1) / 2 : tiles_m * tiles_n;
    int g = blockIdx.x, b, t;
    if ((nbatch & 7) == 0) {   // keep every tile of a matrix on one XCD (dispatch is round-robin over 8)
        const int xcd = g & 7, j = g >> 3;
        b = xcd + 8 * (j / tiles);
        t = j % tiles;
    } else {
        b = g / tiles;
        t = g % tiles;
    }
    int ti = t % tiles_m, tj = t / tiles_m;
    if (lower) {            // column tj holds tiles_m - tj tiles, rows tj .. tiles_m - 1
        tj = 0;
        while (t >= tiles_m - tj) { t -= tiles_m - tj; tj++; }
        ti = tj + t;
    }
    const int row0 = ti * GT, col0 = tj * GT;
    C += (size_t)b * sC; A += (size_t)b * sA; B += (size_t)b * sB;
    if constexpr (sizeof(T) == 8) {
        if (kk >= K3M) {
            gemm_tile<T, true, KC3>(lds, C, A, B, m, n, kk, lda, ldb, ldc, row0, col0);
            return;
        }
    }
    gemm_tile<T, false, KC>(lds, C, A, B, m, n, kk, lda, ldb, ldc, row0, col0);
}

// ------------------------------------------------------------------ tracelog
// out[b] = (sum_i log U_ii + i pi [odd interchanges]) / n  (ms_matrix_util.c:185-202),
// accumulated in FP64; stats[b] = {max_i log|U_ii|, min_i log|U_ii|}.
template <class T>
__global__ __launch_bounds__(256) void tracelog_kernel(const cx<T>* __restrict__ A, int n,
                                                       const MatItem* __restrict__ items, int np,
                                                       long long mstride, const int* __restrict__ ipiv_all,
                                                       cx<double>* __restrict__ out, double* __restrict__ stats) {
    __shared__ double sre[256], sim[256], smax[256], smin[256];
    __shared__ int sodd[256];
    const int b = blockIdx.x;
    if (items) n = items[b].n;
    const int* ipiv = ipiv_all ? ipiv_all + (size_t)b * np : nullptr;      // null: no interchanges
    tracelog_block<T>(A + (size_t)b * mstride, n, np, ipiv, sre, sim, smax, smin, sodd, out + b,
                      stats ? stats + 2 * b : nullptr);
}

// ------------------------------------------------------------------ permutations
// perm[i] = source row of row i of P*B, from the panels' composite maps.
__global__ __launch_bounds__(256) void build_perm_kernel(const int* __restrict__ pmap_all, int np, int nb,
                                                         int* __restrict__ perm_all) {
    extern __shared__ int sp[];
    const int b = blockIdx.x, tid = threadIdx.x, nq = np / nb;
    for (int i = tid; i < np; i += 256) sp[i] = i;
    __syncthreads();
    for (int q = 0; q < nq; q++) {
        const int* pm = pmap_all + ((size_t)b * nq + q) * (1 + 4 * nb);
        const int cnt = pm[0];
        int v = 0;
        if (tid < cnt) v = sp[pm[1 + 2 * nb + tid]];
        __syncthreads();
        if (tid < cnt) sp[pm[1 + tid]] = v;
        __syncthreads();
    }
    for (int i = tid; i < np; i += 256) perm_all[(size_t)b * np + i] = sp[i];
}

// dst[b][i, c] = src[b][perm[b][i], c] for i < np, c < ncol (column-major, ld np).
template <class T>
__global__ void gather_rows_kernel(cx<T>* __restrict__ dst, const cx<T>* __restrict__ src,
                                   const int* __restrict__ perm_all, int np, int ncol, long long sdst,
                                   long long ssrc) {
    const int b = blockIdx.z;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)np * ncol) return;
    const int i = (int)(e % np), c = (int)(e / np);
    const int* perm = perm_all + (size_t)b * np;
    dst[(size_t)b * sdst + (size_t)i + (size_t)c * np] = src[(size_t)b * ssrc + (size_t)perm[i] + (size_t)c * np];
}

// ------------------------------------------------------------------ triangular block solve
// Rows [j0, j0+64) of B (np x ncol, ld np) against the 64x64 diagonal block of the
// factor read column-major: UPPER=0 lower forward, UPPER=1 upper backward, with a
// unit or stored diagonal (UNIT).  Column-major LU: unit lower, non-unit upper; the
// row-major LU of a symmetric matrix read column-major: non-unit lower, unit upper.
template <class T, int UPPER, bool UNIT>
__global__ __launch_bounds__(256) void trsm_block_kernel(const cx<T>* __restrict__ LU, cx<T>* __restrict__ B,
                                                         int np, int ncol, int j0, long long slu, long long sb) {
    __shared__ cx<T> sL[64 * 65 / 2];          // the triangle with its diagonal, packed
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
    const cx<T>* F = LU + (size_t)b * slu;
    cx<T>* X = B + (size_t)b * sb;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int i = e % 64, m = e / 64;
        if (UPPER ? i <= m : i >= m)
            sL[UPPER ? (m * (m + 1) >> 1) + i : (i * (i + 1) >> 1) + m] = F[(size_t)(j0 + i) + (size_t)(j0 + m) * np];
    }
    __syncthreads();
    const int cl = tid >> 2, qq = tid & 3;
    const int col = blockIdx.x * 64 + cl;
    const bool live = col < ncol;
    cx<T> xv[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int i = 4 * m + qq;
        xv[m] = live ? X[(size_t)(j0 + i) + (size_t)col * np] : cx<T>{T(0), T(0)};
    }
    if (!UPPER) lower_unit_solve<T, UNIT>(xv, PackedLD<T>{sL}, 64, qq, lane);
    else upper_solve<T, UNIT>(xv, PackedUD<T>{sL}, qq, lane);
    if (live) {
#pragma unroll
        for (int m = 0; m < 16; m++) X[(size_t)(j0 + 4 * m + qq) + (size_t)col * np] = xv[m];
    }
}

// out[b] = sum_{i<n} X[b][i, p(i)] in FP64, p(i) = perm[b][i] (perm given, ld = np) or i.
template <class T>
__global__ __launch_bounds__(256) void diag_sum_kernel(const cx<T>* __restrict__ X, int n,
                                                       const MatItem* __restrict__ items, int ld, long long sx,
                                                       cx<double>* __restrict__ out, const int* __restrict__ perm) {
    __shared__ double sre[256], sim[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    if (items) n = items[b].n;
    double re = 0.0, im = 0.0;
    for (int i = tid; i < n; i += 256) {
        const int c = perm ? perm[(size_t)b * ld + i] : i;
        const cx<T> v = X[(size_t)b * sx + (size_t)i + (size_t)c * ld];
        re += (double)v.x;
        im += (double)v.y;
    }
    sre[tid] = re; sim[tid] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) { sre[tid] += sre[tid + s]; sim[tid] += sim[tid + s]; }
        __syncthreads();
    }
    if (tid == 0) out[b] = cx<double>{sre[0], sim[0]};
}

// y (T2) = x (T1) elementwise, count complex entries.
template <class T1, class T2>
__global__ void convert_kernel(const cx<T1>* __restrict__ x, cx<T2>* __restrict__ y, long long count) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) y[e] = cx<T2>{(T2)x[e].x, (T2)x[e].y};
}

// r = phi - M a for one matrix (np x np, FP64), one thread per row.
__global__ void residual_kernel(const cx<double>* __restrict__ M, const cx<double>* __restrict__ a,
                                const cx<double>* __restrict__ phi, int np, cx<double>* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    cx<double> s = phi[i];
    for (int j = 0; j < np; j++) s = s - M[(size_t)i + (size_t)j * np] * a[j];
    r[i] = s;
}

// a += d (FP64 += FP32-or-FP64 correction)
template <class T>
__global__ void axpy_kernel(cx<double>* __restrict__ a, const cx<T>* __restrict__ d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = a[i] + cx<double>{(double)d[i].x, (double)d[i].y};
}


// Apply panel q's composite row map to X (np x ncol, ld np): 8 columns per workgroup.
template <class T>
__global__ __launch_bounds__(256) void apply_map_kernel(cx<T>* __restrict__ X, int np, int ncol,
                                                        const int* __restrict__ pmap_all, int nb, int q,
                                                        long long sx) {
    __shared__ int s_dst[128], s_src[128];
    const int b = blockIdx.y, tid = threadIdx.x, nq = np / nb;
    const int* pm = pmap_all + ((size_t)b * nq + q) * (1 + 4 * nb);
    cx<T>* M = X + (size_t)b * sx;
    const int cnt = pm[0];
    for (int e = tid; e < cnt; e += 256) { s_dst[e] = pm[1 + e]; s_src[e] = pm[1 + 2 * nb + e]; }
    __syncthreads();
    const int c0 = blockIdx.x * 8, nc = min(8, ncol - c0), pairs = cnt * nc;
    cx<T> tmp[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int idx = tid + 256 * u;
        if (idx < pairs) tmp[u] = M[(size_t)s_src[idx % cnt] + (size_t)(c0 + idx / cnt) * np];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int idx = tid + 256 * u;
        if (idx < pairs) M[(size_t)s_dst[idx % cnt] + (size_t)(c0 + idx / cnt) * np] = tmp[u];
    }
}

// Identity padding: M[i][i] = 1 for n <= i < np (the rest of the padding is already zero).
template <class T>
__global__ void pad_identity_kernel(cx<T>* __restrict__ A, int n, int np, long long ms) {
    const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) A[(size_t)blockIdx.y * ms + (size_t)i + (size_t)i * np] = cx<T>{T(1), T(0)};
}

template <class T>
void launch_gemm_sub(cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int n, int k, int lda, int ldb, int ldc,
                     long long sC, long long sA, long long sB, int nbatch, hipStream_t st) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    const int tm = (m + GT - 1) / GT, tn = (n + GT - 1) / GT;
    hipLaunchKernelGGL(gemm_sub_kernel<T>, dim3(tm * tn * nbatch), dim3(256), 0, st, C, A, B, m, n, k, lda, ldb, ldc,
                       sC, sA, sB, tm, tn, nbatch, 0);
}

template <class T>
void launch_gemm_sub_lower(cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int k, int lda, int ldb, int ldc,
                           long long sC, long long sA, long long sB, int nbatch, hipStream_t st) {
    if (m <= 0 || k <= 0) return;
    const int tm = (m + GT - 1) / GT;
    hipLaunchKernelGGL(gemm_sub_kernel<T>, dim3(tm * (tm + 1) / 2 * nbatch), dim3(256), 0, st, C, A, B, m, m, k, lda,
                       ldb, ldc, sC, sA, sB, tm, tm, nbatch, 1);
}

template <class T>
void launch_tracelog(const cx<T>* A, int n, const MatItem* items, int np, long long mstride, const int* ipiv,
                     cx<double>* out, double* stats, int nbatch, hipStream_t st) {
    hipLaunchKernelGGL(tracelog_kernel<T>, dim3(nbatch), dim3(256), 0, st, A, n, items, np, mstride, ipiv, out,
                       stats);
}

void launch_build_perm(const int* pmap, int np, int nb, int* perm, int nbatch, hipStream_t st) {
    hipLaunchKernelGGL(build_perm_kernel, dim3(nbatch), dim3(256), sizeof(int) * (size_t)np, st, pmap, np, nb, perm);
}

template <class T>
void launch_gather_rows(cx<T>* dst, const cx<T>* src, const int* perm, int np, int ncol, long long sdst,
                        long long ssrc, int nbatch, hipStream_t st) {
    const long long tot = (long long)np * ncol;
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3((unsigned)((tot + 255) / 256), 1, nbatch), dim3(256), 0, st, dst,
                       src, perm, np, ncol, sdst, ssrc);
}

template <class T>
void launch_trsm_block(const cx<T>* LU, cx<T>* B, int np, int ncol, int j0, int upper, long long slu, long long sb,
                       int nbatch, hipStream_t st, bool rm) {
    dim3 grid((ncol + 63) / 64, nbatch);
    if (upper && !rm)
        hipLaunchKernelGGL((trsm_block_kernel<T, 1, false>), grid, dim3(256), 0, st, LU, B, np, ncol, j0, slu, sb);
    else if (upper)
        hipLaunchKernelGGL((trsm_block_kernel<T, 1, true>), grid, dim3(256), 0, st, LU, B, np, ncol, j0, slu, sb);
    else if (!rm)
        hipLaunchKernelGGL((trsm_block_kernel<T, 0, true>), grid, dim3(256), 0, st, LU, B, np, ncol, j0, slu, sb);
    else
        hipLaunchKernelGGL((trsm_block_kernel<T, 0, false>), grid, dim3(256), 0, st, LU, B, np, ncol, j0, slu, sb);
}

// One inverse power step's normalisation for every live vector of a batch
// (complex_vector_util.c:47-74): fac = |y| y_i0 / |y_i0| with i0 the first nonzero entry
// (fac = 1 for y = 0), y /= fac, delta = |y - v|; then v = y, fac and delta stored, and
// done set once delta < toler.  Vectors already done are left as they are.  One
// 256-thread workgroup per vector; sums in FP64.
__global__ __launch_bounds__(256) void power_step_kernel(cx<double>* __restrict__ V, cx<double>* __restrict__ Y,
                                                         int n, int ld, double toler, int* __restrict__ done,
                                                         cx<double>* __restrict__ fac_out,
                                                         double* __restrict__ delta_out) {
    __shared__ double s_r[256];
    __shared__ int s_i[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    if (done[b]) return;
    cx<double>* v = V + (size_t)b * ld;
    cx<double>* y = Y + (size_t)b * ld;
    double r2 = 0.0;
    int i0 = n;
    for (int i = tid; i < n; i += 256) {
        const cx<double> t = y[i];
        r2 += t.x * t.x + t.y * t.y;
        if (i < i0 && (t.x != 0.0 || t.y != 0.0)) i0 = i;
    }
    s_r[tid] = r2;
    s_i[tid] = i0;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) { s_r[tid] += s_r[tid + h]; s_i[tid] = min(s_i[tid], s_i[tid + h]); }
        __syncthreads();
    }
    const int first = s_i[0];
    cx<double> fac{1.0, 0.0};
    if (first < n) {
        const cx<double> p = y[first];
        fac = (p * (sqrt(s_r[0]) / hypot(p.x, p.y)));
    }
    __syncthreads();
    double d2 = 0.0;
    for (int i = tid; i < n; i += 256) {
        const cx<double> t = first < n ? y[i] / fac : y[i];
        const cx<double> e = t - v[i];
        d2 += e.x * e.x + e.y * e.y;
        v[i] = t;
    }
    s_r[tid] = d2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) s_r[tid] += s_r[tid + h];
        __syncthreads();
    }
    if (tid == 0) {
        const double delta = sqrt(s_r[0]);
        fac_out[b] = fac;
        delta_out[b] = delta;
        if (delta < toler) done[b] = 1;
    }
}

void launch_power_step(cx<double>* V, cx<double>* Y, int n, int ld, double toler, int* done, cx<double>* fac,
                       double* delta, int nbatch, hipStream_t st) {
    if (nbatch > 0)
        hipLaunchKernelGGL(power_step_kernel, dim3(nbatch), dim3(256), 0, st, V, Y, n, ld, toler, done, fac, delta);
}

// A[b] = I (np x np) for every matrix of a batch: zero fill, then the diagonal.
template <class T>
__global__ void unit_diag_kernel(cx<T>* __restrict__ A, int np, long long ms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) A[(size_t)blockIdx.y * ms + (size_t)i * (np + 1)] = cx<T>{T(1), T(0)};
}

template <class T>
hipError_t launch_set_identity(cx<T>* A, int np, long long ms, int nbatch, hipStream_t st) {
    if (nbatch <= 0) return hipSuccess;
    const hipError_t e = hipMemsetAsync(A, 0, sizeof(cx<T>) * (size_t)ms * nbatch, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(unit_diag_kernel<T>, dim3((np + 255) / 256, nbatch), dim3(256), 0, st, A, np, ms);
    return hipGetLastError();
}
template hipError_t launch_set_identity<double>(cx<double>*, int, long long, int, hipStream_t);
template hipError_t launch_set_identity<float>(cx<float>*, int, long long, int, hipStream_t);

// x[i] = v for i < n (a kernel, so that it may be captured in a HIP graph).
__global__ void fill_int_kernel(int* __restrict__ x, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

void launch_fill_int(int* x, int v, int n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, st, x, v, n);
}

// dst[b][perm[i], c] = src[b][i, c] for i < np, c < ncol (column-major, ld np).
template <class T>
__global__ void scatter_rows_kernel(cx<T>* __restrict__ dst, const cx<T>* __restrict__ src,
                                    const int* __restrict__ perm_all, int np, int ncol, long long sdst,
                                    long long ssrc) {
    const int b = blockIdx.z;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)np * ncol) return;
    const int i = (int)(e % np), c = (int)(e / np);
    const int* perm = perm_all + (size_t)b * np;
    dst[(size_t)b * sdst + (size_t)perm[i] + (size_t)c * np] = src[(size_t)b * ssrc + (size_t)i + (size_t)c * np];
}

template <class T>
void launch_scatter_rows(cx<T>* dst, const cx<T>* src, const int* perm, int np, int ncol, long long sdst,
                         long long ssrc, int nbatch, hipStream_t st) {
    const long long tot = (long long)np * ncol;
    hipLaunchKernelGGL(scatter_rows_kernel<T>, dim3((unsigned)((tot + 255) / 256), 1, nbatch), dim3(256), 0, st, dst,
                       src, perm, np, ncol, sdst, ssrc);
}

template <class T>
void launch_diag_sum(const cx<T>* X, int n, const MatItem* items, int ld, long long sx, cx<double>* out, int nbatch,
                     hipStream_t st, const int* perm) {
    hipLaunchKernelGGL(diag_sum_kernel<T>, dim3(nbatch), dim3(256), 0, st, X, n, items, ld, sx, out, perm);
}

template <class T1, class T2>
void launch_convert(const cx<T1>* x, cx<T2>* y, long long count, hipStream_t st) {
    hipLaunchKernelGGL((convert_kernel<T1, T2>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, x, y, count);
}

void launch_residual(const cx<double>* M, const cx<double>* a, const cx<double>* phi, int np, cx<double>* r,
                     hipStream_t st) {
    hipLaunchKernelGGL(residual_kernel, dim3((np + 63) / 64), dim3(64), 0, st, M, a, phi, np, r);
}

template <class T>
void launch_axpy(cx<double>* a, const cx<T>* d, int n, hipStream_t st) {
    hipLaunchKernelGGL(axpy_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, a, d, n);
}


template <class T>
void launch_apply_map(cx<T>* X, int np, int ncol, const int* pmap, int nb, int q, long long sx, int nbatch,
                      hipStream_t st) {
    hipLaunchKernelGGL(apply_map_kernel<T>, dim3((ncol + 7) / 8, nbatch), dim3(256), 0, st, X, np, ncol, pmap, nb, q,
                       sx);
}

template <class T>
void launch_pad_identity(cx<T>* A, int n, int np, int nbatch, hipStream_t st) {
    if (np <= n) return;
    hipLaunchKernelGGL(pad_identity_kernel<T>, dim3((np - n + 63) / 64, nbatch), dim3(64), 0, st, A, n, np,
                       (long long)np * np);
}

template <class T>
void launch_panel(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb, int nbatch,
                  hipStream_t st, bool rm, int variant, int xa, int xb) {
    if (rm) launch_panel_rm<T, true>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, variant, xa, xb);
    else launch_panel_rm<T, false>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, variant, xa, xb);
}

#define MSG_INST(T)                                                                                                   \
    template void launch_panel<T>(cx<T>*, int, long long, int*, int*, int*, int, int, int, hipStream_t, bool, int, \
                                  int, int);                                                                       \
    template void launch_gemm_sub<T>(cx<T>*, const cx<T>*, const cx<T>*, int, int, int, int, int, int, long long,      \
                                     long long, long long, int, hipStream_t);                                          \
    template void launch_gemm_sub_lower<T>(cx<T>*, const cx<T>*, const cx<T>*, int, int, int, int, int, long long,    \
                                           long long, long long, int, hipStream_t);                                    \
    template void launch_tracelog<T>(const cx<T>*, int, const MatItem*, int, long long, const int*, cx<double>*,     \
                                     double*, int, hipStream_t);                                                       \
    template void launch_gather_rows<T>(cx<T>*, const cx<T>*, const int*, int, int, long long, long long, int,       \
                                        hipStream_t);                                                                  \
    template void launch_trsm_block<T>(const cx<T>*, cx<T>*, int, int, int, int, long long, long long, int,           \
                                       hipStream_t, bool);                                                             \
    template void launch_scatter_rows<T>(cx<T>*, const cx<T>*, const int*, int, int, long long, long long, int,      \
                                         hipStream_t);                                                                 \
    template void launch_diag_sum<T>(const cx<T>*, int, const MatItem*, int, long long, cx<double>*, int,             \
                                     hipStream_t, const int*);                                                         \
    template void launch_axpy<T>(cx<double>*, const cx<T>*, int, hipStream_t);                                       \
    template void launch_apply_map<T>(cx<T>*, int, int, const int*, int, int, long long, int, hipStream_t);           \
    template void launch_pad_identity<T>(cx<T>*, int, int, int, hipStream_t);
MSG_INST(double)
MSG_INST(float)
template void launch_convert<double, float>(const cx<double>*, cx<float>*, long long, hipStream_t);
template void launch_convert<float, double>(const cx<float>*, cx<double>*, long long, hipStream_t);

}  // namespace msg
