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
// part (of its diagonal block too) and U12 from the rows kernel; the factored diagonal
// block keeps L11 below its diagonal, D11 on it and zeros above.  log det M =
// sum_i log D_ii (the trace-log kernel without interchanges).
//
// Without pivoting the factors can grow; each kernel folds the largest |L| of its matrix
// into growth[b] (the bits of a non-negative double: atomicMax on the integer order), and
// a NaN or infinite |L| counts as infinite growth.  Along with it runs the error estimate
// rho[b] = sum_r sum_{k<r} |L_rk|^2 |D_k| / |D_r| (ldlt_dev.h; acc: one double per row of
// every matrix, zero at the start).  The caller accepts a matrix only when that estimate
// is small and recomputes the others with the pivoted LU.
#include "ldlt_dev.h"

namespace msg {

namespace {

// Largest value over the workgroup (NW waves), folded into growth[b] (the bits of a
// non-negative double: atomicMax on the integer order).
template <int NW = 4>
__device__ __forceinline__ void fold_growth(double g, unsigned long long* growth, int b, double* sred) {
    const double m = group_max<NW>(g, sred);
    if (threadIdx.x == 0) atomicMax(growth + b, (unsigned long long)__double_as_longlong(m));
}

// One workgroup per matrix: the diagonal block at k0 (ldlt_diag_block), eight waves of
// eight columns: a launch has one workgroup per matrix, and with four waves of sixteen
// columns each SIMD held a single wave of dependent steps.  The blocks of a matrix are
// factored by successive launches, so rho[b] takes a plain sum.
constexpr int kDiagWaves = 8;
template <class T>
__global__ __launch_bounds__(64 * kDiagWaves) void ldlt_diag_kernel(cx<T>* __restrict__ A, int np, long long ms,
                                                                    int k0, unsigned long long* __restrict__ growth,
                                                                    const double* __restrict__ acc,
                                                                    double* __restrict__ rho) {
    __shared__ __align__(16) double sbuf[512];
    __shared__ double sred[kDiagWaves];
    const int b = blockIdx.x;
    double r = 0.0;
    const T g = ldlt_diag_block<T, kDiagWaves>(A + (size_t)b * ms + (size_t)k0 + (size_t)k0 * np, np, sbuf,
                                               acc + (size_t)b * np + k0, r);
    fold_growth<kDiagWaves>((double)g, growth, b, sred);
    if (threadIdx.x == 0) rho[b] += r;
}

// Rows below the diagonal block, 64 per workgroup (ldlt_rows_stage, ldlt_rows_solve);
// three waves per SIMD (168 registers).
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void ldlt_rows_kernel(cx<T>* __restrict__ A, int np, long long ms, int k0,
                                                        unsigned long long* __restrict__ growth,
                                                        double* __restrict__ acc) {
    __shared__ cx<T> sL[64 * 63 / 2];        // L11 strictly lower, packed by rows
    __shared__ cx<T> sD[64];
    __shared__ double sred[4];
    const int b = blockIdx.y;
    cx<T>* M = A + (size_t)b * ms;
    ldlt_rows_stage<T>(M, np, k0, sL, sD);
    __syncthreads();
    const T g = ldlt_rows_solve<T>(M, np, k0, blockIdx.x, sL, sD, acc + (size_t)b * np);
    fold_growth((double)g, growth, b, sred);
}

}  // namespace

template <class T>
void launch_ldlt_diag(cx<T>* A, int np, long long ms, int k0, unsigned long long* growth, const double* acc,
                      double* rho, int nbatch, hipStream_t st) {
    hipLaunchKernelGGL(ldlt_diag_kernel<T>, dim3(nbatch), dim3(64 * kDiagWaves), 0, st, A, np, ms, k0, growth, acc,
                       rho);
}

template <class T>
void launch_ldlt_rows(cx<T>* A, int np, long long ms, int k0, unsigned long long* growth, double* acc, int nbatch,
                      hipStream_t st) {
    const int W = np - k0 - 64;
    if (W <= 0) return;
    hipLaunchKernelGGL(ldlt_rows_kernel<T>, dim3((W + 63) / 64, nbatch), dim3(256), 0, st, A, np, ms, k0, growth,
                       acc);
}

template void launch_ldlt_diag<double>(cx<double>*, int, long long, int, unsigned long long*, const double*, double*,
                                       int, hipStream_t);
template void launch_ldlt_rows<double>(cx<double>*, int, long long, int, unsigned long long*, double*, int,
                                       hipStream_t);

}  // namespace msg
