// kern_ldlt.hip -- pivot-free LDL^T of the complex symmetric multiple-scattering matrix,
// the blocked right-looking form whose panels need no pivot search:
//
//   for k0 = 0, 64, ...:
//     diagonal block A11 = L11 D11 L11^T      (ldlt_diag_kernel: one workgroup per matrix, in LDS;
//                                              the block is taken from its lower triangle)
//     L21 = A21 U11^-1, U11 = D11 L11^T       (ldlt_rows_kernel: four lanes per row, U11^T staged in LDS)
//     U12 = D11 L21^T                         (same kernel: A12 = A21^T, so U12 needs no solve)
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

template <class T>
__global__ __launch_bounds__(256) void ldlt_diag_kernel(cx<T>* __restrict__ A, int np, long long ms, int k0,
                                                        unsigned long long* __restrict__ growth) {
    __shared__ cx<T> s[64][65];              // s[c][r] = block (r, c)
    __shared__ double sred[4];
    const int b = blockIdx.x, tid = threadIdx.x;
    cx<T>* M = A + (size_t)b * ms + (size_t)k0 + (size_t)k0 * np;
    // the block as its lower triangle defines it: the updates of the lower tiles leave the
    // diagonal tile's two triangles equal only to rounding (L21 D U12 products), and
    // factoring that slightly unsymmetric block while U12 is formed as D11 L21^T (which
    // assumes U11 = D11 L11^T) loses accuracy in proportion to the growth of L
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        if (r >= c) {
            const cx<T> v = M[(size_t)r + (size_t)c * np];
            s[c][r] = v;
            s[r][c] = v;
        }
    }
    __syncthreads();
    T g = T(0);
    for (int c = 0; c < 63; c++) {
        if (tid > c && tid < 64) {           // column c of L11
            const cx<T> l = s[c][tid] * inv(s[c][c]);
            s[c][tid] = l;
            g = fmax(g, growth_of(l));
        }
        __syncthreads();
        const int w = 63 - c;                // rows and columns c+1 .. 63 (U part of row c unchanged)
        for (int e = tid; e < w * w; e += 256) {
            const int r = c + 1 + e % w, cc = c + 1 + e / w;
            s[cc][r] = s[cc][r] - s[c][r] * s[cc][c];
        }
        __syncthreads();
    }
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        M[(size_t)r + (size_t)c * np] = s[c][r];
    }
    fold_growth((double)g, growth, b, sred);
}

template <class T>
__global__ __launch_bounds__(256) void ldlt_rows_kernel(cx<T>* __restrict__ A, int np, long long ms, int k0,
                                                        unsigned long long* __restrict__ growth) {
    __shared__ cx<T> sL[64 * 65 / 2];        // U11^T, packed lower with its diagonal: (i, m) = U11(m, i)
    __shared__ double sred[4];
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63, qq = tid & 3;
    cx<T>* M = A + (size_t)b * ms;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int m = e & 63, i = e >> 6;
        if (m <= i) sL[(i * (i + 1) >> 1) + m] = M[(size_t)(k0 + m) + (size_t)(k0 + i) * np];
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
    // x U11 = a, i.e. U11^T x^T = a^T: forward substitution with the stored diagonal
    lower_unit_solve<T, false>(xv, PackedLD<T>{sL}, 64, qq, lane);
    T g = T(0);
    if (live) {
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int i = 4 * m + qq;
            M[row + (size_t)(k0 + i) * np] = xv[m];                                  // L21
            M[(size_t)(k0 + i) + row * np] = sL[(i * (i + 1) >> 1) + i] * xv[m];    // U12 = D11 L21^T
            g = fmax(g, growth_of(xv[m]));
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
