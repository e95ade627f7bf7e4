// kern_ldlt.hip -- pivot-free LDL^T of the complex symmetric multiple-scattering matrix,
// the blocked right-looking form whose panels need no pivot search:
//
//   for k0 = 0, 64, ...:
//     diagonal block A11 = L11 D11 L11^T      (ldlt_diag_kernel: one workgroup per matrix, in
//                                              registers; the block is taken from its lower triangle)
//     Y = A21 L11^-T; U12 = Y^T = D11 L21^T;  (ldlt_rows_kernel: four lanes per row, unit forward
//     L21 = Y D11^-1                           substitution against L11 staged in LDS; A12 = A21^T,
//                                              so U12 needs no solve of its own)
//     A22 -= L21 U12 on and below the diagonal (launch_gemm_sub_lower: half the tiles of the LU's update)
//
// The strictly upper tiles of A22 are never read again: every later step reads the lower
// part, its own diagonal block in full (a diagonal tile of the update), and U12 from the
// rows kernel.  log det M = sum_i log D_ii (the trace-log kernel without interchanges).
//
// Without pivoting the factors can grow; each kernel folds the largest |L| of its matrix
// into growth[b] (the bits of a non-negative double: atomicMax on the integer order), and
// a NaN or infinite |L| counts as infinite growth.  The caller accepts a matrix only when
// that growth is small and recomputes the others with the pivoted LU.
#include "lu_device.h"

namespace msg {

namespace {

template <class T>
__device__ __forceinline__ T growth_of(cx<T> l) {
    const T a = sqrt(l.x * l.x + l.y * l.y);
    return a == a ? a : T(INFINITY);      // NaN counts as unbounded growth
}

// Largest value over the workgroup (256 threads, four waves), folded into growth[b].
__device__ __forceinline__ void fold_growth(double g, unsigned long long* growth, int b, double* sred) {
    for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_xor(g, off));
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) sred[tid >> 6] = g;
    __syncthreads();
    if (tid == 0) {
        const double m = fmax(fmax(sred[0], sred[1]), fmax(sred[2], sred[3]));
        atomicMax(growth + b, (unsigned long long)__double_as_longlong(m));
    }
}

// The 64 x 64 diagonal block, one workgroup per matrix, held in registers: lane r of
// wave w holds a(r, 16 w + j) in x[j].  Step c: the wave holding column c forms
// l_r = a(r, c) / a(c, c) for r > c (the pivot from lane c by readlane) and passes the
// column through LDS (double-buffered: one barrier per step); every wave then updates
// a(r, j) -= l_r a(c, j) for r, j > c, with row c of its own columns read from its
// lane c.  The arithmetic is the unblocked LDL^T's, element by element.
template <class T>
__global__ __launch_bounds__(256) void ldlt_diag_kernel(cx<T>* __restrict__ A, int np, long long ms, int k0,
                                                        unsigned long long* __restrict__ growth) {
    __shared__ cx<T> sl[2][64];
    __shared__ double sred[4];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    cx<T>* M = A + (size_t)b * ms + (size_t)k0 + (size_t)k0 * np;
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
    fold_growth((double)g, growth, b, sred);
}

// Rows below the diagonal block, 64 per workgroup, four lanes per row: with
// A21 = L21 D11 L11^T, the row y of Y = A21 L11^-T solves y L11^T = a by unit forward
// substitution against L11 (staged packed in LDS), then U12 = Y^T (= D11 L21^T) and
// L21 = Y D11^-1.
template <class T>
__global__ __launch_bounds__(256) void ldlt_rows_kernel(cx<T>* __restrict__ A, int np, long long ms, int k0,
                                                        unsigned long long* __restrict__ growth) {
    __shared__ cx<T> sL[64 * 63 / 2];        // L11 strictly lower, packed by rows
    __shared__ cx<T> sD[64];
    __shared__ double sred[4];
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63, qq = tid & 3;
    (void)lane;
    cx<T>* M = A + (size_t)b * ms;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        if (r > c) sL[(r * (r - 1) >> 1) + c] = M[(size_t)(k0 + r) + (size_t)(k0 + c) * np];
        else if (r == c) sD[r] = M[(size_t)(k0 + r) + (size_t)(k0 + r) * np];
    }
    __syncthreads();
    const int W = np - k0 - 64;
    const int j = blockIdx.x * 64 + (tid >> 2);
    const bool live = j < W;
    const size_t row = (size_t)(k0 + 64 + (live ? j : 0));
    cx<T> xv[16];
#pragma unroll
    for (int m = 0; m < 16; m++)
        xv[m] = live ? M[row + (size_t)(k0 + 4 * m + qq) * np] : cx<T>{T(0), T(0)};
    // (the block edge as a run-time value: with the constant 64 the compiler schedules
    // the unrolled solve into more registers than the file holds)
    lower_unit_solve_unrolled<T>(xv, PackedL<T>{sL}, min(64, np - k0), qq);
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
    fold_growth((double)g, growth, b, sred);
}

}  // namespace

template <class T>
void launch_ldlt_diag(cx<T>* A, int np, long long ms, int k0, unsigned long long* growth, int nbatch, hipStream_t st) {
    hipLaunchKernelGGL(ldlt_diag_kernel<T>, dim3(nbatch), dim3(256), 0, st, A, np, ms, k0, growth);
}

template <class T>
void launch_ldlt_rows(cx<T>* A, int np, long long ms, int k0, unsigned long long* growth, int nbatch, hipStream_t st) {
    const int W = np - k0 - 64;
    if (W <= 0) return;
    hipLaunchKernelGGL(ldlt_rows_kernel<T>, dim3((W + 63) / 64, nbatch), dim3(256), 0, st, A, np, ms, k0, growth);
}

template void launch_ldlt_diag<double>(cx<double>*, int, long long, int, unsigned long long*, int, hipStream_t);
template void launch_ldlt_rows<double>(cx<double>*, int, long long, int, unsigned long long*, int, hipStream_t);

}  // namespace msg
