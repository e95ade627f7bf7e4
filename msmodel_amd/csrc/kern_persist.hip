// kern_persist.hip -- the persistent trace-log kernel of the pivot-free LDL^T path.
//
// One launch covers a whole batch of (medium, k) items of any sizes: every workgroup
// takes the next item from an atomic counter and runs its whole pipeline alone, in its
// own workspace matrix:
//   assembly of M(k) (Np = N rounded up to 64, identity padding),
//   the blocked LDL^T in 128-column blocks (per 64-column panel the diagonal block and
//   the rows below it; the K = 64 update of the block's second panel; the K = 128 update
//   of the lower tiles right of the block, on MFMA),
//   sum_i log D_ii / N, the largest |L| of the item and its error estimate (ldlt_dev.h;
//   the row sums in the workgroup's own Np doubles of accws).
// These are the device functions of the per-phase launches (kern_assemble.hip,
// kern_ldlt.hip, kern_lu.hip) in the same order, so each item's results are those of
// the batched LDL^T bit for bit.  Items depend on nothing but themselves: no workgroup
// waits for another, and each leaves when the counter passes the last item.  A
// workgroup's phases are separated by barriers (its global writes are visible to its
// own later reads after one).
#include "assemble_dev.h"
#include "gemm_dev.h"
#include "ldlt_dev.h"

namespace msg {

namespace {

// shared memory of the largest phase, in doubles: the K = 64 GEMM tile's staging
constexpr int kGemmLds = KC * (2 * GT + PADA) + GT * (2 * KC + PADB);
constexpr int kPersistLds = kGemmLds;
static_assert(2 * AT * 4 + AT * (AT + 1) * 2 <= kPersistLds, "assembly staging");
static_assert((64 * 63 / 2 + 64) * 2 <= kPersistLds, "rows staging");
static_assert(KC3 * (2 * GT + PADA) + GT * (2 * KC3 + PADB) <= kPersistLds, "3M staging");
static_assert(4 * 256 + 128 <= kPersistLds, "trace-log reduction");
static_assert(512 <= kPersistLds, "diagonal block");

// Three workgroups per CU (the register budget of the 3M tile, as for the batched GEMM),
// so that one workgroup's latency-bound phases overlap another's GEMM tiles.
template <bool D3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void ldlt_persist_kernel(const MatItem* __restrict__ items, int nitems,
                                                           cx<double>* __restrict__ ws, long long wstride,
                                                           double* __restrict__ accws, long long astride,
                                                           int* __restrict__ counter, cx<double>* __restrict__ out,
                                                           unsigned long long* __restrict__ growth,
                                                           double* __restrict__ rho) {
    __shared__ __align__(16) double smem[kPersistLds];
    __shared__ double sred[4];
    __shared__ int s_item;
    const int tid = threadIdx.x;
    cx<double>* M = ws + (size_t)blockIdx.x * wstride;
    double* acc = accws + (size_t)blockIdx.x * astride;
    auto at = [&](int r, int c, int np) { return M + r + (size_t)c * np; };
    for (;;) {
        if (tid == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) break;                           // uniform: every workgroup leaves here
        const MatItem item = items[it];
        const int n = item.n, np = (n + 63) / 64 * 64;
        for (int i = tid; i < np; i += 256) acc[i] = 0.0;  // ordered before use by the barriers below
        {   // assembly, tile pair by tile pair
            auto sx = reinterpret_cast<double (*)[AT][4]>(smem);
            auto tile = reinterpret_cast<cx<double> (*)[AT + 1]>(smem + 2 * AT * 4);
            const int nt = np / AT;
            for (int t = 0; t < nt * (nt + 1) / 2; t++) {
                int I, J;
                tile_pair(t, nt, I, J);
                assemble_pair<double, D3>(item.pos, n, item.d, np, item.k, item.diag0, cx<double>{1.0, 0.0}, 0, M,
                                          I, J, sx, tile, true);
                __syncthreads();
            }
        }
        // 64-column panels kp = 0, 64, ...: the diagonal block, then the rows below it; after
        // the first panel of a 128-column block, the K = 64 update of its second panel; after
        // the second, the K = 128 update of the lower tiles right of the block (3M)
        double g = 0.0, rs = 0.0;
        for (int kp = 0; kp < np; kp += 64) {
            double r = 0.0;
            g = fmax(g, (double)ldlt_diag_block<double, 4>(at(kp, kp, np), np, smem, acc + kp, r));
            rs += r;
            __syncthreads();
            const int W = np - kp - 64;
            if (W <= 0) break;
            cx<double>* sL = reinterpret_cast<cx<double>*>(smem);
            cx<double>* sD = sL + 64 * 63 / 2;
            ldlt_rows_stage<double>(M, np, kp, sL, sD);
            __syncthreads();
            // (a barrier per 64 rows: the triangle's reads are the same for every block of rows,
            // and hoisted out of the loop they would not fit in registers)
            for (int jb = 0; jb < (W + 63) / 64; jb++) {
                g = fmax(g, (double)ldlt_rows_solve<double>(M, np, kp, jb, sL, sD, acc));
                __syncthreads();
            }
            if ((kp & 127) == 0) {
                for (int ti = 0; ti < (W + GT - 1) / GT; ti++)
                    gemm_tile<double, false, KC>(smem, at(kp + 64, kp + 64, np), at(kp + 64, kp, np),
                                                 at(kp, kp + 64, np), W, 64, 64, np, np, np, ti * GT, 0);
            } else {
                const int k0 = kp - 64, tm = (W + GT - 1) / GT;
                for (int tj = 0; tj < tm; tj++)
                    for (int ti = tj; ti < tm; ti++)
                        gemm_tile<double, true, KC3>(smem, at(k0 + 128, k0 + 128, np), at(k0 + 128, k0, np),
                                                     at(k0, k0 + 128, np), W, W, 128, np, np, np, ti * GT, tj * GT);
            }
            __syncthreads();
        }
        double* r = smem;
        tracelog_block<double>(M, n, np, nullptr, r, r + 256, r + 512, r + 768, reinterpret_cast<int*>(r + 1024),
                               out + it, nullptr);
        const double gm = group_max(g, sred);
        if (tid == 0) {
            growth[it] = (unsigned long long)__double_as_longlong(gm);
            rho[it] = rs;
        }
        __syncthreads();                                   // s_item, smem and sred are reused
    }
}

}  // namespace

int persist_ldlt_blocks_per_cu(bool d3) {
    int nb = 0;
    const hipError_t e = d3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ldlt_persist_kernel<true>, 256, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ldlt_persist_kernel<false>, 256, 0);
    return e == hipSuccess ? nb : 0;
}

void launch_ldlt_persist(const MatItem* items, int nitems, bool d3, cx<double>* ws, long long wstride, double* accws,
                         long long astride, int nwg, int* counter, cx<double>* out, unsigned long long* growth,
                         double* rho, hipStream_t st) {
    if (d3)
        hipLaunchKernelGGL(ldlt_persist_kernel<true>, dim3(nwg), dim3(256), 0, st, items, nitems, ws, wstride, accws,
                           astride, counter, out, growth, rho);
    else
        hipLaunchKernelGGL(ldlt_persist_kernel<false>, dim3(nwg), dim3(256), 0, st, items, nitems, ws, wstride, accws,
                           astride, counter, out, growth, rho);
}

}  // namespace msg
