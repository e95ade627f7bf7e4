// assemble_dev.h -- the multiple-scattering matrix, one pair of 32 x 32 tiles at a time
// by a 256-thread workgroup: shared by the assembly launches (kern_assemble.hip) and the
// persistent trace-log kernel (kern_persist.hip).
//
// Layout: Np x Np, column-major, interleaved complex (re, im), leading dimension Np.  Np
// is N rounded up to a multiple of 64; the padding is the identity.
#pragma once
#include "kernels.h"

namespace msg {

// ------------------------------------------------------------------ d = 3 Green function
__device__ __forceinline__ double rsqrt_t(double x) { return rsqrt(x); }
__device__ __forceinline__ float rsqrt_t(float x) { return rsqrtf(x); }

// -G_3(k, r) = e^{ikr} / (4 pi r) from s = r^2, with 1/r from one reciprocal square root
// and r = s * (1/r), in place of a square root and a division; REALK (Im k = 0) leaves out
// the factor exp(-Im k r) = 1.  At s = 0 it returns what the general form does there.
template <class T, bool REALK>
__device__ __forceinline__ cx<T> neg_green3_s(cx<T> k, T s) {
    const T ri = rsqrt_t(s);
    const T r = s > T(0) ? s * ri : T(0);
    T sn, cs;
    sincos_t(k.x * r, &sn, &cs);
    T w = ri * T(1.0 / kFourPi);
    if (!REALK) w *= exp(-k.y * r);
    return {cs * w, sn * w};
}

// ------------------------------------------------------------------ assembly
// The tile pair (I, J), I <= J, of one matrix: the Green function is evaluated once per
// unordered pair of scatterers and written to both (I,J) and, through an LDS transpose,
// (J,I), so M = M^T exactly (ms_matrix_util.c:36-41 mirrors the upper triangle the same
// way).  kind 0: M (diagonal dg = F^-1(k)); kind 1: dM/dk (dg the finite-difference
// derivative of F^-1); kind 2: the normalised matrix (divided by iz).
//
// D3: the matrix is d = 3 and kind 0.  The general code carries the Bessel-function code
// of d = 2 and 4 and needs 208 VGPRs (two waves per SIMD), which leaves the d = 3 stores
// little to overlap with; the specialised one has only the d = 3 Green function.
//
// sx (2 x AT x 4 doubles) and tile (AT x (AT + 1)) are shared memory; the caller
// separates two calls that reuse them by a barrier.
constexpr int AT = 32;                 // tile edge (256 threads: AT / (256 / AT) elements each)
constexpr int ARW = 256 / AT;           // thread rows of a tile

// lower: the tile below the diagonal only (J, I) for I < J (the pivot-free LDL^T reads no
// other); the diagonal tiles in full either way.
template <class T, bool D3>
__device__ __forceinline__ void assemble_pair(const double* __restrict__ pos, int n, int d, int np, cx<double> kd,
                                              cx<double> dg, cx<double> iz, int kind, cx<T>* __restrict__ M,
                                              int I, int J, double (*sx)[AT][4], cx<T> (*tile)[AT + 1],
                                              bool lower = false) {
    const int tid = threadIdx.x;
    if (tid < 2 * AT) {
        int side = tid / AT, a = tid % AT;
        int atom = (side == 0 ? I : J) * AT + a;
        for (int c = 0; c < 4; c++) sx[side][a][c] = (atom < n && c < d) ? pos[(size_t)atom * d + c] : 0.0;
    }
    __syncthreads();
    const cx<T> k{(T)kd.x, (T)kd.y};
    const int ii = tid % AT, j0 = tid / AT;
#pragma unroll
    for (int q = 0; q < AT / ARW; q++) {
        const int jj = j0 + ARW * q;
        const int gi = I * AT + ii, gj = J * AT + jj;
        cx<T> v;
        if (gi >= n || gj >= n) {
            v = cx<T>{gi == gj ? T(1) : T(0), T(0)};
        } else if (gi == gj) {
            v = cx<T>{(T)dg.x, (T)dg.y};
        } else if (D3 || (d == 3 && kind == 0)) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double e = sx[0][ii][c] - sx[1][jj][c];
                s += e * e;
            }
            v = neg_green3_s<T, false>(k, (T)s);
        } else if constexpr (!D3) {
            double s = 0.0;
            for (int c = 0; c < d; c++) {
                double e = sx[0][ii][c] - sx[1][jj][c];
                s += e * e;
            }
            T r = (T)sqrt(s);
            if (kind == 1) {
                v = neg_dk_green<T>(d, k, r);
            } else if (kind == 2) {
                const cx<T> ng = neg_green<T>(d, k, r);
                const cx<double> g = cx<double>{(double)ng.x, (double)ng.y} / iz;
                v = cx<T>{(T)g.x, (T)g.y};
            } else {
                v = neg_green<T>(d, k, r);
            }
        }
        if (!lower || I == J) M[(size_t)gi + (size_t)gj * np] = v;
        tile[jj][ii] = v;
    }
    if (I == J) return;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < AT / ARW; q++) {
        const int bb = j0 + ARW * q;      // column inside tile (J, I) is I*AT + bb
        M[(size_t)(J * AT + ii) + (size_t)(I * AT + bb) * np] = tile[ii][bb];
    }
}

// The upper-triangular tile pair of index t (0 <= t < nt (nt + 1) / 2, nt = np / AT).
__device__ __forceinline__ void tile_pair(int t, int nt, int& I, int& J) {
    I = 0;
    while (t >= nt - I) { t -= nt - I; I++; }
    J = I + t;
}

}  // namespace msg
