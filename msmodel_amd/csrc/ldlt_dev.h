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

// Largest g over the workgroup (256 threads, four waves), in thread 0; sred: 4 doubles
// of shared memory.
__device__ __forceinline__ double group_max(double g, double* sred) {
    for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_xor(g, off));
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) sred[tid >> 6] = g;
    __syncthreads();
    return fmax(fmax(sred[0], sred[1]), fmax(sred[2], sred[3]));
}

// The 64 x 64 diagonal block at M (its origin; ld np), held in registers: lane r of
// wave w holds a(r, 16 w + j) in x[j].  Step c: the wave holding column c forms
// l_r = a(r, c) / a(c, c) for r > c (the pivot from lane c by readlane) and passes the
// column through LDS (sl: 2 x 64, double-buffered: one barrier per step); every wave then
// updates a(r, j) -= l_r a(c, j) for r, j > c, with row c of its own columns read from
// its lane c.  The arithmetic is the unblocked LDL^T's, element by element.  Returns this
// thread's largest |l|.
template <class T>
__device__ __forceinline__ T ldlt_diag_block(cx<T>* __restrict__ M, int np, cx<T> (*sl)[64]) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // the block as its lower triangle defines it: the updates of the lower tiles leave the
    // diagonal tile's two triangles equal only to rounding (L21 D U12 products), and
    // factoring that slightly unsymmetric block while U12 is formed from L21 and D (which
    // assumes a symmetric block) loses accuracy in proportion to the growth of L
    cx<T> x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int col = 16 * wv + j;
        x[j] = lane >= col ? M[(size_t)lane + (size_t)col * np] : M[(size_t)col + (size_t)lane * np];
    }
    T g = T(0);
#pragma unroll 1
    for (int cw = 0; cw < 4; cw++) {
#pragma unroll
        for (int jc = 0; jc < 16; jc++) {
            const int c = 16 * cw + jc;
            if (c == 63) break;
            const int buf = jc & 1;
            if (wv == cw) {                  // column c of L11
                const cx<T> piv{rdlane(x[jc].x, c), rdlane(x[jc].y, c)};
                if (lane > c) {
                    const cx<T> l = x[jc] * inv(piv);
                    x[jc] = l;
                    g = fmax(g, growth_of(l));
                }
                sl[buf][lane] = x[jc];
            }
            __syncthreads();
            if (16 * wv + 15 > c) {          // this wave has columns right of c
                const cx<T> l = sl[buf][lane];
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int col = 16 * wv + j;
                    const cx<T> u{rdlane(x[j].x, c), rdlane(x[j].y, c)};
                    if (col > c && lane > c) x[j] = x[j] - l * u;
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 16; j++) M[(size_t)lane + (size_t)(16 * wv + j) * np] = x[j];
    return g;
}

// Rows below the diagonal block at k0, with A21 = L21 D11 L11^T: the row y of
// Y = A21 L11^-T solves y L11^T = a by unit forward substitution against L11, then
// U12 = Y^T (= D11 L21^T) and L21 = Y D11^-1.
// Stage: L11's strictly lower part, packed by rows (sL, 64 * 63 / 2), and D11 (sD, 64).
template <class T>
__device__ __forceinline__ void ldlt_rows_stage(const cx<T>* __restrict__ M, int np, int k0, cx<T>* sL, cx<T>* sD) {
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        if (r > c) sL[(r * (r - 1) >> 1) + c] = M[(size_t)(k0 + r) + (size_t)(k0 + c) * np];
        else if (r == c) sD[r] = M[(size_t)(k0 + r) + (size_t)(k0 + r) * np];
    }
}

// Solve: the 64 rows k0 + 64 + 64 jb + [0, 64) (those below np are skipped), four lanes
// per row.  Returns this thread's largest |l|.
template <class T>
__device__ __forceinline__ T ldlt_rows_solve(cx<T>* __restrict__ M, int np, int k0, int jb, const cx<T>* sL,
                                             const cx<T>* sD) {
    const int tid = threadIdx.x, qq = tid & 3;
    const int W = np - k0 - 64;
    const int j = jb * 64 + (tid >> 2);
    const bool live = j < W;
    const size_t row = (size_t)(k0 + 64 + (live ? j : 0));
    cx<T> xv[16];
#pragma unroll
    for (int m = 0; m < 16; m++)
        xv[m] = live ? M[row + (size_t)(k0 + 4 * m + qq) * np] : cx<T>{T(0), T(0)};
    // (the block edge as a value the compiler cannot see through: known to be the constant
    // 64, it schedules the unrolled solve into more registers than the file holds)
    int nb = min(64, np - k0);
    asm volatile("" : "+v"(nb));
    lower_unit_solve_unrolled<T>(xv, PackedL<T>{sL}, nb, qq);
    T g = T(0);
    if (live) {
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int i = 4 * m + qq;
            const cx<T> l = xv[m] / sD[i];
            M[(size_t)(k0 + i) + row * np] = xv[m];                                   // U12 = D11 L21^T
            M[row + (size_t)(k0 + i) * np] = l;                                       // L21
            g = fmax(g, growth_of(l));
        }
    }
    return g;
}

}  // namespace msg
