// gemm_dev.h -- one 64 x 64 complex tile of C -= A B by a 256-thread workgroup on MFMA,
// shared by the batched GEMM launches (kern_lu.hip) and the persistent trace-log kernel
// (kern_persist.hip).
#pragma once
#include "lu_device.h"

namespace msg {

// ------------------------------------------------------------------ GEMM: C -= A B
constexpr int GT = 64;     // complex tile edge (rows and columns)
#ifndef MSG_GEMM_KC
#define MSG_GEMM_KC 16
#endif
constexpr int KC = MSG_GEMM_KC;     // complex K chunk staged in LDS
#ifndef MSG_GEMM_PADA
#define MSG_GEMM_PADA 2
#endif
#ifndef MSG_GEMM_PADB
#define MSG_GEMM_PADB 2
#endif
constexpr int PADA = MSG_GEMM_PADA, PADB = MSG_GEMM_PADB;   // LDS row padding (elements of T)


// One 64 x 64 complex tile of C -= A B (256 threads, 2 x 2 waves of 32 x 32), K staged
// through LDS in chunks of KCT complex columns (A) / rows (B); the next chunk's global
// loads are in flight while the current one computes.
//   M3 (complex128, K >= 128): the 3M (Gauss) product.  With A = Ar + i Ai, B = Br + i Bi,
//     the real products X = Ar Br, Y = Ai Bi, Z = (Ar + Ai)(Br + Bi) give
//     AB = (X - Y) + i (Z - X - Y): three real MFMA products per complex product instead
//     of four.  The accumulators start at zero and C - AB = (Re C + (Y - X)) +
//     i (Im C + (X + Y - Z)) is formed at the end (C read there), so that exact-zero
//     products (padding) leave C bit for bit.  MFMA row i of a 16 x 16 tile is
//     complex row arow(i): a lane's four registers are four consecutive rows of one
//     column (one 64-byte access of C).
//   otherwise: the real product of the interleaved C (2m x n) with [[Ar,-Ai],[Ai,Ar]]
//     (2m x 2k) and the interleaved B (2k x n); a lane's four registers are two
//     consecutive complex rows of one column.
template <class T, bool M3, int KCT>
__device__ __forceinline__ void gemm_tile(T* lds, cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int n, int kk,
                                          int lda, int ldb, int ldc, int row0, int col0) {
    T (*lA)[2 * GT + PADA] = reinterpret_cast<T (*)[2 * GT + PADA]>(lds);
    T (*lB)[2 * KCT + PADB] = reinterpret_cast<T (*)[2 * KCT + PADB]>(lds + KCT * (2 * GT + PADA));
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w & 1, wc = w >> 1, kq = lane >> 4;
    const int ar = mfma<T>::arow(lane & 15);
    using acc_t = typename mfma<T>::acc;
    constexpr int LU = KCT * GT / 256;     // complex elements per thread per operand and chunk
    cx<T> pa[LU], pb[LU];
    // interior tiles (the common case) load without per-element bounds tests
    const bool full_mn = row0 + GT <= m && col0 + GT <= n;
    auto fetch = [&](int kc0) {
        if (full_mn && kc0 + KCT <= kk) {
#pragma unroll
            for (int u = 0; u < LU; u++) {
                const int e = tid + 256 * u;
                pa[u] = A[(size_t)(row0 + e % GT) + (size_t)(kc0 + e / GT) * lda];
                pb[u] = B[(size_t)(kc0 + e % KCT) + (size_t)(col0 + e / KCT) * ldb];
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < LU; u++) {
            const int e = tid + 256 * u;
            const int kcol = e / GT, row = e % GT;
            pa[u] = (row0 + row < m && kc0 + kcol < kk) ? A[(size_t)(row0 + row) + (size_t)(kc0 + kcol) * lda]
                                                       : cx<T>{T(0), T(0)};
            const int col = e / KCT, kr = e % KCT;
            pb[u] = (col0 + col < n && kc0 + kr < kk) ? B[(size_t)(kc0 + kr) + (size_t)(col0 + col) * ldb]
                                                     : cx<T>{T(0), T(0)};
        }
    };
    auto stage = [&] {
#pragma unroll
        for (int u = 0; u < LU; u++) {
            const int e = tid + 256 * u;
            const int kcol = e / GT, row = e % GT;
            lA[kcol][2 * row] = pa[u].x;
            lA[kcol][2 * row + 1] = pa[u].y;
            const int col = e / KCT, kr = e % KCT;
            lB[col][2 * kr] = pb[u].x;
            lB[col][2 * kr + 1] = pb[u].y;
        }
    };
    fetch(0);
    if constexpr (M3) {
        acc_t X[2][2], Y[2][2], Z[2][2];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    X[rb][cb][r] = T(0);
                    Y[rb][cb][r] = T(0);
                    Z[rb][cb][r] = T(0);
                }
        for (int kc0 = 0; kc0 < kk; kc0 += KCT) {
            __syncthreads();
            stage();
            __syncthreads();
            if (kc0 + KCT < kk) fetch(kc0 + KCT);     // in flight while this chunk computes
#pragma unroll
            for (int ks = 0; ks < KCT; ks += 4) {
                T a_r[2], a_i[2], a_s[2], b_r[2], b_i[2], b_s[2];
#pragma unroll
                for (int rb = 0; rb < 2; rb++) {
                    const cx<T> v = *reinterpret_cast<const cx<T>*>(&lA[ks + kq][2 * (wr * 32 + rb * 16 + ar)]);
                    a_r[rb] = v.x; a_i[rb] = v.y; a_s[rb] = v.x + v.y;
                }
#pragma unroll
                for (int cb = 0; cb < 2; cb++) {
                    const cx<T> v = *reinterpret_cast<const cx<T>*>(&lB[wc * 32 + cb * 16 + (lane & 15)][2 * (ks + kq)]);
                    b_r[cb] = v.x; b_i[cb] = v.y; b_s[cb] = v.x + v.y;
                }
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int cb = 0; cb < 2; cb++) {
                        X[rb][cb] = mfma<T>::op(a_r[rb], b_r[cb], X[rb][cb]);
                        Y[rb][cb] = mfma<T>::op(a_i[rb], b_i[cb], Y[rb][cb]);
                        Z[rb][cb] = mfma<T>::op(a_s[rb], b_s[cb], Z[rb][cb]);
                    }
            }
        }
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const int crow = row0 + wr * 32 + rb * 16 + 4 * kq;
                const int ccol = col0 + wc * 32 + cb * 16 + (lane & 15);
                if (ccol < n) {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (crow + r < m) {
                            const T x = X[rb][cb][r], y = Y[rb][cb][r], z = Z[rb][cb][r];
                            cx<T>* cp = C + (size_t)(crow + r) + (size_t)ccol * ldc;
                            const cx<T> c = *cp;
                            *cp = cx<T>{c.x + (y - x), c.y + ((x + y) - z)};
                        }
                }
            }
    } else {
        acc_t acc[4][2];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const int crow = row0 + wr * 32 + rb * 8 + 2 * kq;
                const int ccol = col0 + wc * 32 + cb * 16 + (lane & 15);
                cx<T> v0{T(0), T(0)}, v1{T(0), T(0)};
                if (ccol < n) {
                    const cx<T>* cp = C + (size_t)crow + (size_t)ccol * ldc;
                    if (crow + 1 < m) { v0 = cp[0]; v1 = cp[1]; }
                    else if (crow < m) v0 = cp[0];
                }
                acc[rb][cb][0] = v0.x; acc[rb][cb][1] = v0.y; acc[rb][cb][2] = v1.x; acc[rb][cb][3] = v1.y;
            }
        const int rot = kq & 1;
        // A fragment of [[Ar, -Ai], [Ai, Ar]]: element (rr, p) of the rotation is
        // -lA[p][rr] (rot = 0), -lA[p][rr-1] (rot = 1, rr odd) or +lA[p][rr+1] (rot = 1, rr
        // even).  rr's parity is ar's, so offset and sign are fixed per lane: one LDS read
        // and no branch per fragment.
        const int aoff = rot ? ((ar & 1) ? -1 : 1) : 0;
        const T asgn = (rot && !(ar & 1)) ? T(1) : T(-1);
        for (int kc0 = 0; kc0 < kk; kc0 += KCT) {
            __syncthreads();
            stage();
            __syncthreads();
            if (kc0 + KCT < kk) fetch(kc0 + KCT);     // in flight while this chunk computes
#pragma unroll
            for (int ks = 0; ks < 2 * KCT; ks += 4) {
                const int p = (ks >> 1) + (kq >> 1);
                T av[4], bv[2];
#pragma unroll
                for (int rb = 0; rb < 4; rb++) av[rb] = asgn * lA[p][wr * 64 + rb * 16 + ar + aoff];
#pragma unroll
                for (int cb = 0; cb < 2; cb++) bv[cb] = lB[wc * 32 + cb * 16 + (lane & 15)][ks + kq];
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int cb = 0; cb < 2; cb++) acc[rb][cb] = mfma<T>::op(av[rb], bv[cb], acc[rb][cb]);
            }
        }
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++) {
                const int crow = row0 + wr * 32 + rb * 8 + 2 * kq;
                const int ccol = col0 + wc * 32 + cb * 16 + (lane & 15);
                if (ccol < n) {
                    cx<T>* cp = C + (size_t)crow + (size_t)ccol * ldc;
                    if (crow < m) cp[0] = cx<T>{acc[rb][cb][0], acc[rb][cb][1]};
                    if (crow + 1 < m) cp[1] = cx<T>{acc[rb][cb][2], acc[rb][cb][3]};
                }
            }
    }
}

// 3M chunk (complex columns of K staged at once) and the K from which complex128 uses 3M
constexpr int KC3 = 8, K3M = 128;

}  // namespace msg
