// ldlt_dev.h -- the phases of the pivot-free LDL^T of a complex symmetric matrix
// (column-major, ld np) for a 256-thread workgroup: the 64 x 64 diagonal block and the
// rows below it.  Shared by the per-phase launches (kern_ldlt.hip) and the persistent
// trace-log kernel (kern_persist.hip).
#pragma once
#include "lu_device.h"

namespace msg {

template <class T>
__device__ __forceinline__ T growth_of(cx<T> l) {
    const T a = sqrt(l.x * l.x + l.y * l.y);
    return a == a ? a : T(INFINITY);      // NaN counts as unbounded growth
}

// The error estimate of the factorisation: rho_r = sum_{k<r} |L_rk|^2 |D_k| / |D_r| is the
// diagonal of |L| |D| |L|^T relative to D_r, so eps rho_r bounds the relative rounding
// error of D_r, and eps sum_r rho_r that of log det M (up to a modest factor).  The sums
// run along with the factorisation: acc[r] (one double per row, zero at the start)
// collects |L_rk|^2 |D_k| = |L_rk| |L_rk D_k| over the columns of the finished panels, and
// the diagonal block of r adds its own columns and divides by |D_r|.
template <class T>
__device__ __forceinline__ T mag(cx<T> z) {
    return sqrt(z.x * z.x + z.y * z.y);
}

// v as held by lane src of the wave (src the same in every lane).
__device__ __forceinline__ double lane_value(double v, int src) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((int)b, src), hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_value(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Largest g over the workgroup (NW waves of 64), in every thread; sred: NW doubles of
// shared memory.
template <int NW = 4>
__device__ __forceinline__ double group_max(double g, double* sred) {
    for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_xor(g, off));
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) sred[tid >> 6] = g;
    __syncthreads();
    double m = sred[0];
#pragma unroll
    for (int w = 1; w < NW; w++) m = fmax(m, sred[w]);
    return m;
}

// The 64 x 64 diagonal block at M (its origin; ld np), held in registers by NW waves of
// CW = 64 / NW columns: lane r of wave w holds a(r, CW w + j) in x[j]; only the lower
// triangle is factored.  Step c: the wave holding column c forms l_r = y_r / y_c from
// y = a(:, c), keeps l_r for r > c and passes y and l through LDS (double-buffered: one
// barrier per step); every wave updates a(r, j) -= l_r y_j for r >= j > c.  The update takes
// y_j = a(j, c) from the column, not a(c, j) from the row: on a block whose two
// triangles agree only to rounding, an LU-style update (row from the upper triangle)
// makes L11 D11 L11^T drift from the block, and the rows below, formed against that
// L11, lose accuracy in proportion to the growth of L (a numpy model of the blocked
// scheme on the N = 1000 map: 1.3e-8 in log det at k = 11.35 - 1.08i against 5e-11 with
// the symmetric update).  Every element sees the same operations in the same order for
// any NW, and rho's partial sums are kept per 8 columns, so NW changes no bit.  Returns
// this thread's largest |l|; rho receives sum_r rho_r over the block's 64 rows in the
// first wave (acc: the block's rows, acc[r] for row r of the block); sbuf: 512 doubles of
// shared memory.
template <class T, int NW = 4>
__device__ __forceinline__ T ldlt_diag_block(cx<T>* __restrict__ M, int np, double* sbuf,
                                             const double* __restrict__ acc, double& rho) {
    constexpr int CW = 64 / NW, NP = CW / 8;
    static_assert(NW == 4 || NW == 8, "four or eight waves");
    static_assert(4 * 64 * sizeof(cx<T>) <= 512 * sizeof(double), "two buffers of y and l");
    // step c: the wave owning column c divides it by the pivot and publishes the column
    // before (sy) and after (sl) the division; the other waves only read them
    auto sy = reinterpret_cast<cx<T> (*)[64]>(sbuf);
    auto sl = sy + 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // the block as its lower triangle defines it: the updates of the lower tiles leave the
    // diagonal tile's two triangles equal only to rounding (L21 D U12 products); the upper
    // triangle is neither read nor kept
    cx<T> x[CW];
#pragma unroll
    for (int j = 0; j < CW; j++) {
        const int col = CW * wv + j;
        x[j] = lane >= col ? M[(size_t)lane + (size_t)col * np] : cx<T>{T(0), T(0)};
    }
    T g = T(0);
    double part[NP];                         // sums of |L_rc|^2 |D_c| over 8 columns c each
#pragma unroll
    for (int q = 0; q < NP; q++) part[q] = 0.0;
#pragma unroll 1
    for (int cw = 0; cw < NW; cw++) {
#pragma unroll
        for (int jc = 0; jc < CW; jc++) {
            const int c = CW * cw + jc;
            if (c == 63) break;
            const int buf = jc & 1;
            if (wv == cw) {
                const cx<T> y = x[jc];                    // column c before the division
                const cx<T> pv = {lane_value(y.x, c), lane_value(y.y, c)};
                const cx<T> l = y * inv(pv);
                sy[buf][lane] = y;
                sl[buf][lane] = l;
                if (lane > c) {                           // column c of L11
                    const T a = growth_of(l);
                    part[jc / 8] += (double)(a * mag(y));
                    x[jc] = l;
                    g = fmax(g, a);
                }
            }
            __syncthreads();
            if (CW * wv + CW - 1 > c) {      // this wave has columns right of c
                const cx<T> l = sl[buf][lane];
#pragma unroll
                for (int j = 0; j < CW; j++) {
                    const int col = CW * wv + j;
                    if (col > c && lane >= col) x[j] = x[j] - l * sy[buf][col];
                }
            }
        }
    }
    __syncthreads();                         // the last step's reads of sl are done
    double* sp = sbuf;
#pragma unroll
    for (int q = 0; q < NP; q++) sp[64 * (NP * wv + q) + lane] = part[q];
#pragma unroll
    for (int j = 0; j < CW; j++) M[(size_t)lane + (size_t)(CW * wv + j) * np] = x[j];
    __syncthreads();                         // the parts and the block's stores are visible
    if (wv == 0) {
        double s = acc[lane];
#pragma unroll
        for (int q = 0; q < 8; q++) s += sp[64 * q + lane];
        double r = s / (double)mag(M[(size_t)lane + (size_t)lane * np]);
        r = r == r ? r : INFINITY;
        for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off);
        rho = r;
    }
    return g;
}

// Rows below the diagonal block at k0, with A21 = L21 D11 L11^T: the row y of
// Y = A21 L11^-T solves y L11^T = a by unit forward substitution against L11, then
// U12 = Y^T (= D11 L21^T) and L21 = Y D11^-1.
// Stage: L11's strictly lower part, packed by rows (sL, 64 * 63 / 2), and the inverse of
// D11 (sD, 64), so that L21 = Y D11^-1 multiplies as the diagonal block forms L11.
template <class T>
__device__ __forceinline__ void ldlt_rows_stage(const cx<T>* __restrict__ M, int np, int k0, cx<T>* sL, cx<T>* sD) {
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        if (r > c) sL[(r * (r - 1) >> 1) + c] = M[(size_t)(k0 + r) + (size_t)(k0 + c) * np];
        else if (r == c) sD[r] = inv(M[(size_t)(k0 + r) + (size_t)(k0 + r) * np]);
    }
}

// Solve: the 64 rows k0 + 64 + 64 jb + [0, 64) (those below np are skipped), four lanes
// per row.  Returns this thread's largest |l|; adds sum_k |L_rk|^2 |D_k| over the panel's
// columns to acc[r] (acc: the matrix's rows).
template <class T>
__device__ __forceinline__ T ldlt_rows_solve(cx<T>* __restrict__ M, int np, int k0, int jb, const cx<T>* sL,
                                             const cx<T>* sD, double* __restrict__ acc) {
    const int tid = threadIdx.x, qq = tid & 3;
    const int W = np - k0 - 64;
    const int j = jb * 64 + (tid >> 2);
    const bool live = j < W;
    const size_t row = (size_t)(k0 + 64 + (live ? j : 0));
    cx<T> xv[16];
#pragma unroll
    for (int m = 0; m < 16; m++)
        xv[m] = live ? M[row + (size_t)(k0 + 4 * m + qq) * np] : cx<T>{T(0), T(0)};
    lower_unit_solve64<T, 1>(xv, PackedL<T>{sL}, qq);
    T g = T(0);
    double part = 0.0;
    if (live) {
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int i = 4 * m + qq;
            const cx<T> l = xv[m] * sD[i];
            M[(size_t)(k0 + i) + row * np] = xv[m];                                   // U12 = D11 L21^T
            M[row + (size_t)(k0 + i) * np] = l;                                       // L21
            const T a = growth_of(l);
            part += (double)(a * mag(xv[m]));
            g = fmax(g, a);
        }
    }
    part += __shfl_xor(part, 1);             // the row's four lanes
    part += __shfl_xor(part, 2);
    if (live && qq == 0) acc[row] += part;
    return g;
}

}  // namespace msg
