// kern_swap.hip -- row interchanges outside the panel with the U12 triangular
// solve (swap_trsm_kernel), and the 64 x 128 update block of the 128-column
// blocking (block_w_kernel).
#include "lu_device.h"

namespace msg {

// ------------------------------------------------------------------ swap + TRSM
// grid.x = column tiles of TW columns: the first nleft_t tiles cover [0, k0)
// (LAPACK-format L), the rest cover [kend, np).  The row map is applied RW columns
// per round (every read of a round precedes its writes; up to 128 x RW scattered
// elements in flight per workgroup).  Right tiles then get U12 = L11^-1 A12, 64
// columns at a time against L11 staged once in LDS.
template <class T, int TW, int RW, bool RM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void swap_trsm_kernel(cx<T>* __restrict__ A, int np, long long mstride,
                                                        const int* __restrict__ pmap_all, int k0, int nb,
                                                        SwapRange sr) {
    const int nleft_t = sr.nleft_t;
    constexpr int U = (128 * RW + 255) / 256;
    __shared__ int s_dst[256], s_src[256];
    __shared__ int s_d1[128], s_s1[128], s_d2[128], s_s2[128], s_wc[4];
    __shared__ cx<T> sL[64 * 63 / 2];          // strictly lower part of the unit-lower L11, packed
    const int b = blockIdx.y;
    cx<T>* M = A + (size_t)b * mstride;
    const int nq = np / nb, q = k0 / nb;
    const int* pmap = pmap_all + ((size_t)b * nq + q) * (1 + 4 * nb);
    const int kend = min(k0 + nb, np);
    const int tid = threadIdx.x;
    const bool left = (int)blockIdx.x < nleft_t;
    const int col0 = left ? sr.lb + blockIdx.x * TW : sr.rb + (blockIdx.x - nleft_t) * TW;
    const int colend = left ? min(col0 + TW, k0) : min(col0 + TW, sr.re);
    const int do_trsm = sr.do_trsm;
    const int ncol = colend - col0;
    int cnt;
    if (sr.k0b < 0) {
        cnt = pmap[0];
        for (int e = tid; e < cnt; e += 256) { s_dst[e] = pmap[1 + e]; s_src[e] = pmap[1 + 2 * nb + e]; }
    } else {
        // compose map 1 (panel k0) then map 2 (panel k0b): the row that ends at y is
        // w = src1'(src2'(y)) (src' = identity off the map's rows), for y in dst1 U dst2;
        // one thread per candidate (dst2 entries, then dst1 entries not in dst2)
        const int* pm2 = pmap_all + ((size_t)b * nq + sr.k0b / nb) * (1 + 4 * nb);
        const int c1 = pmap[0], c2 = pm2[0];
        for (int e = tid; e < c1; e += 256) { s_d1[e] = pmap[1 + e]; s_s1[e] = pmap[1 + 2 * nb + e]; }
        for (int e = tid; e < c2; e += 256) { s_d2[e] = pm2[1 + e]; s_s2[e] = pm2[1 + 2 * nb + e]; }
        __syncthreads();
        int y = -1, w = -1;
        bool act = false;
        if (tid < c1 + c2) {
            const bool from2 = tid < c2;
            y = from2 ? s_d2[tid] : s_d1[tid - c2];
            int j2 = from2 ? tid : -1;
            if (!from2)
                for (int e = 0; e < c2; e++) j2 = (s_d2[e] == y) ? e : j2;
            if (from2 || j2 < 0) {
                const int z = j2 >= 0 ? s_s2[j2] : y;
                int j1 = -1;
                for (int e = 0; e < c1; e++) j1 = (s_d1[e] == z) ? e : j1;
                w = j1 >= 0 ? s_s1[j1] : z;
                act = w != y;
            }
        }
        const unsigned long long m = __ballot(act);
        const int lane = tid & 63, wv = tid >> 6;
        if (lane == 0) s_wc[wv] = __popcll(m);
        __syncthreads();
        int off = 0;
        for (int v = 0; v < wv; v++) off += s_wc[v];
        if (act) {
            const int slot = off + __popcll(m & ((1ull << lane) - 1ull));
            s_dst[slot] = y;
            s_src[slot] = w;
        }
        cnt = s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
    }
    if (!left && do_trsm) {
        // unit-lower L11 (nb x nb) of the panel
        for (int e = tid; e < nb * nb; e += 256) {
            const int i = RM ? e / nb : e % nb, m = RM ? e % nb : e / nb;
            if (i > m) sL[(i * (i - 1) >> 1) + m] = M[at<RM>(k0 + i, k0 + m, np)];
        }
    }
    __syncthreads();
    if (cnt > 0 && sr.do_swap) {
        // columns per round: all rows of a column are read before any is written, and a
        // round holds at most 256 U elements (a composed map has up to 256 rows)
        const int rw = max(1, min(RW, (256 * U) / cnt));
        for (int cb = 0; cb < ncol; cb += rw) {
            const int w = min(rw, ncol - cb), pairs = cnt * w;
            cx<T> tmp[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int idx = tid + 256 * u;
                tmp[u] = cx<T>{T(0), T(0)};     // written on every path: the array stays in registers
                if (idx < pairs) {
                    const int e = RM ? idx / w : idx % cnt, c = col0 + cb + (RM ? idx % w : idx / cnt);
                    tmp[u] = M[at<RM>(s_src[e], c, np)];
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int idx = tid + 256 * u;
                if (idx < pairs) {
                    const int e = RM ? idx / w : idx % cnt, c = col0 + cb + (RM ? idx % w : idx / cnt);
                    M[at<RM>(s_dst[e], c, np)] = tmp[u];
                }
            }
            __syncthreads();
        }
    }
    if (left || !do_trsm) return;
    // TRSM: 4 lanes per column; lane q of a group holds rows i = 4m + q, m < nb/4
    const int cl = tid >> 2, qq = tid & 3, lane = tid & 63;
    for (int sub = 0; sub < ncol; sub += 64) {
        const int col = col0 + sub + cl;
        const bool live = sub + cl < ncol;
        cx<T> xv[16];
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const int i = 4 * m + qq;
            xv[m] = (live && i < nb) ? M[at<RM>(k0 + i, col, np)] : cx<T>{T(0), T(0)};
        }
        lower_unit_solve64<T, 1>(xv, PackedL<T>{sL}, qq);          // nb = 64: every panel is full
        if (live) {
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int i = 4 * m + qq;
                if (i < nb) M[at<RM>(k0 + i, col, np)] = xv[m];
            }
        }
    }
}

// ------------------------------------------------------------------ second-panel update block
// For a 128-column block [K0, K0+128) with second panel k1 = K0 + 64: the rows k1..ke
// of the trailing columns become U12(k1) = L11^-1 (A - L21 U12(K0)) with L11 the unit
// lower block of k1 and L21 = M(k1:ke, K0:k1).  That is one update of K = 128,
//   C <- C - Wb [U12(K0); C],   Wb = [L11^-1 L21, I - L11^-1]   (64 x 128),
// on the rows K0..ke of the trailing columns.  One 256-thread workgroup per matrix forms
// Wb by forward substitution (4 lanes per column) against L11 packed in LDS (32 KB, so
// that the workgroup shares a CU with others): L11^-1 L21 from the columns of L21, then
// L11^-1 from those of the identity.  Wb is stored column-major (ld 64) for column-major
// factors and row-major (row stride 128) for row-major ones, where the update is applied
// to the transposed blocks and Wb^T is the right operand.
template <class T, bool RM>
__global__ __launch_bounds__(256) void block_w_kernel(const cx<T>* __restrict__ A, int np, long long mstride, int K0,
                                                      cx<T>* __restrict__ W) {
    __shared__ cx<T> sL[64 * 63 / 2];          // strictly lower part of the unit-lower L11, packed
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const cx<T>* M = A + (size_t)b * mstride;
    cx<T>* Wb = W + (size_t)b * (64 * 128);
    const int k1 = K0 + 64;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int i = RM ? e / 64 : e % 64, m = RM ? e % 64 : e / 64;
        if (i > m) sL[(i * (i - 1) >> 1) + m] = M[at<RM>(k1 + i, k1 + m, np)];
    }
    __syncthreads();
    // thread: column cl of the right-hand sides, rows 4 m + qq
    const int cl = tid >> 2, qq = tid & 3;
    cx<T> xv[16];
    // Wb(:, cl) = L11^-1 L21(:, cl), with L21 = M(k1:ke, K0:k1)
#pragma unroll
    for (int m = 0; m < 16; m++) xv[m] = M[at<RM>(k1 + 4 * m + qq, K0 + cl, np)];
    lower_unit_solve<T>(xv, PackedL<T>{sL}, 64, qq, lane);
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int i = 4 * m + qq;
        if (RM) Wb[i * 128 + cl] = xv[m];
        else Wb[i + cl * 64] = xv[m];
    }
    // Wb(:, 64 + cl) = (I - L11^-1)(:, cl)
#pragma unroll
    for (int m = 0; m < 16; m++) xv[m] = (4 * m + qq == cl) ? cx<T>{T(1), T(0)} : cx<T>{T(0), T(0)};
    lower_unit_solve<T>(xv, PackedL<T>{sL}, 64, qq, lane);
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const int i = 4 * m + qq;
        const cx<T> v = (i == cl) ? cx<T>{T(0), T(0)} : cx<T>{-xv[m].x, -xv[m].y};
        if (RM) Wb[i * 128 + 64 + cl] = v;
        else Wb[i + (64 + cl) * 64] = v;
    }
}

template <class T>
void launch_block_w(const cx<T>* A, int np, long long mstride, int K0, cx<T>* W, int nbatch, hipStream_t st, bool rm) {
    if (nbatch <= 0) return;
    if (rm) hipLaunchKernelGGL((block_w_kernel<T, true>), dim3(nbatch), dim3(256), 0, st, A, np, mstride, K0, W);
    else hipLaunchKernelGGL((block_w_kernel<T, false>), dim3(nbatch), dim3(256), 0, st, A, np, mstride, K0, W);
}
template void launch_block_w<double>(const cx<double>*, int, long long, int, cx<double>*, int, hipStream_t, bool);
template void launch_block_w<float>(const cx<float>*, int, long long, int, cx<float>*, int, hipStream_t, bool);

// Tile width and round width of the swap + TRSM variants (the autotuner picks one).
struct SwapShape { int tw, rw; };
constexpr SwapShape kSwapShapes[] = {{256, 8}, {128, 16}, {64, 16}, {64, 32}, {128, 32}, {256, 32}};
constexpr int kNumSwapShapes = sizeof(kSwapShapes) / sizeof(kSwapShapes[0]);

template <class T, bool RM>
static void swap_go_rm(int variant, dim3 grid, hipStream_t st, cx<T>* A, int np, long long mstride, const int* pmap,
                       int k0, int nb, SwapRange sr) {
    switch (variant) {
        case 1: hipLaunchKernelGGL((swap_trsm_kernel<T, 128, 16, RM>), grid, dim3(256), 0, st, A, np, mstride, pmap, k0, nb, sr); break;
        case 2: hipLaunchKernelGGL((swap_trsm_kernel<T, 64, 16, RM>), grid, dim3(256), 0, st, A, np, mstride, pmap, k0, nb, sr); break;
        case 3: hipLaunchKernelGGL((swap_trsm_kernel<T, 64, 32, RM>), grid, dim3(256), 0, st, A, np, mstride, pmap, k0, nb, sr); break;
        case 4: hipLaunchKernelGGL((swap_trsm_kernel<T, 128, 32, RM>), grid, dim3(256), 0, st, A, np, mstride, pmap, k0, nb, sr); break;
        case 5: hipLaunchKernelGGL((swap_trsm_kernel<T, 256, 32, RM>), grid, dim3(256), 0, st, A, np, mstride, pmap, k0, nb, sr); break;
        default: hipLaunchKernelGGL((swap_trsm_kernel<T, 256, 8, RM>), grid, dim3(256), 0, st, A, np, mstride, pmap, k0, nb, sr); break;
    }
}

template <class T>
static void swap_go(int variant, dim3 grid, hipStream_t st, cx<T>* A, int np, long long mstride, const int* pmap,
                    int k0, int nb, SwapRange sr, bool rm) {
    if (rm) swap_go_rm<T, true>(variant, grid, st, A, np, mstride, pmap, k0, nb, sr);
    else swap_go_rm<T, false>(variant, grid, st, A, np, mstride, pmap, k0, nb, sr);
}

int swap_variants() { return kNumSwapShapes; }

template <class T>
void launch_swap_trsm(cx<T>* A, int np, long long mstride, const int* pmap, int k0, int nb, int apply_left,
                      int nbatch, hipStream_t st, int variant, bool rm) {
    const int kend = std::min(k0 + nb, np);
    launch_swap_range<T>(A, np, mstride, pmap, k0, nb, apply_left ? 0 : k0, kend, np, 1, 1, nbatch, st, variant, rm);
}

template <class T>
void launch_swap_range(cx<T>* A, int np, long long mstride, const int* pmap, int k0, int nb, int lb, int rb, int re,
                       int do_swap, int do_trsm, int nbatch, hipStream_t st, int variant, bool rm, int k0b) {
    if (variant < 0 || variant >= kNumSwapShapes) variant = 0;
    const int tw = kSwapShapes[variant].tw;
    SwapRange sr;
    sr.lb = lb;
    sr.nleft_t = (do_swap && lb < k0) ? (k0 - lb + tw - 1) / tw : 0;
    sr.rb = rb;
    sr.re = re;
    sr.do_swap = do_swap;
    sr.do_trsm = do_trsm;
    sr.k0b = (do_swap && k0b > k0) ? k0b : -1;
    const int nright_t = re > rb ? (re - rb + tw - 1) / tw : 0;
    if (sr.nleft_t + nright_t == 0) return;
    swap_go<T>(variant, dim3(sr.nleft_t + nright_t, nbatch), st, A, np, mstride, pmap, k0, nb, sr, rm);
}

template <class T>
void launch_swap_left(cx<T>* A, int np, long long mstride, const int* pmap, int k0, int nb, int nbatch,
                      hipStream_t st) {
    launch_swap_range<T>(A, np, mstride, pmap, k0, nb, 0, 0, 0, 1, 0, nbatch, st, 0, false);
}


#define MSG_INST_SWAP(T)                                                                                              \
    template void launch_swap_trsm<T>(cx<T>*, int, long long, const int*, int, int, int, int, hipStream_t, int, bool);\
    template void launch_swap_range<T>(cx<T>*, int, long long, const int*, int, int, int, int, int, int, int, int,    \
                                       hipStream_t, int, bool, int);                                                   \
    template void launch_swap_left<T>(cx<T>*, int, long long, const int*, int, int, int, hipStream_t);
MSG_INST_SWAP(double)
MSG_INST_SWAP(float)

}  // namespace msg
