// kern_lu.hip -- batched right-looking blocked LU with partial pivoting for
// complex matrices (complex128 or complex64), and the kernels that consume it.
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
#include "kernels.h"

namespace msg {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
template <class T> struct mfma;
// Operands (both types): lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Result: lane l, register r holds D[i][l%16] with i = 4*(l/16)+r for f32 and
// i = l/16 + 4*r for f64 (checked on gfx950 with an in-bounds probe).  For f64 the
// GEMM feeds MFMA row i with real row arow(i) = 4*(i%4) + i/4 of the tile, so that
// for both types the four registers of a lane are real rows 4*(l/16)..+3, i.e. two
// consecutive complex elements of one column (one 32- or 16-byte access).
template <> struct mfma<double> {
    using acc = d4v;
    static __device__ __forceinline__ acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int arow(int i) { return 4 * (i & 3) + (i >> 2); }
};
template <> struct mfma<float> {
    using acc = f4v;
    static __device__ __forceinline__ acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int arow(int i) { return i; }
};

// Element (r, c) of an np x np matrix: column-major (RM = false) or row-major (RM =
// true).  The row-major factorisation of a symmetric M is the column-pivoted one of
// the same buffer read column-major: its row interchanges move contiguous rows.
template <bool RM>
__device__ __forceinline__ size_t at(int r, int c, int np) {
    return RM ? (size_t)r * np + (size_t)c : (size_t)r + (size_t)c * np;
}

// ------------------------------------------------------------------ panel
// pmap layout per matrix and panel q: [cnt, dst[2*nb], src[2*nb]] (1 + 4*nb ints): the
// panel's interchanges as one move of row src[e] to row dst[e] (rows that end where
// they started are left out).

// DPP lane exchange (all rows and banks enabled; the controls used have a source lane
// in the same row for every lane).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
__device__ __forceinline__ float dpp_v(float v) { return __int_as_float(dpp_i<CTRL>(__float_as_int(v))); }
template <int CTRL>
__device__ __forceinline__ double dpp_v(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)b), hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rdlane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double rdlane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Pivot order of LAPACK izamax: the larger |Re|+|Im| first, the smaller row on ties.
template <class T>
__device__ __forceinline__ void key_max(T& v, int& i, T ov, int oi) {
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}
template <int CTRL, class T>
__device__ __forceinline__ void key_dpp(T& v, int& i) { key_max(v, i, dpp_v<CTRL>(v), dpp_i<CTRL>(i)); }
// Largest key of each row of 16 lanes, in every lane of the row: quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror.  Called with all lanes active.
template <class T>
__device__ __forceinline__ void key_row16(T& v, int& i) {
    key_dpp<0xB1>(v, i);
    key_dpp<0x4E>(v, i);
    key_dpp<0x141>(v, i);
    key_dpp<0x140>(v, i);
}
// Largest key of the wave (wave-uniform result).
template <class T>
__device__ __forceinline__ void key_wave(T& v, int& i) {
    key_row16(v, i);
    T bv = rdlane(v, 0);
    int bi = rdlane(i, 0);
    key_max(bv, bi, rdlane(v, 16), rdlane(i, 16));
    key_max(bv, bi, rdlane(v, 32), rdlane(i, 32));
    key_max(bv, bi, rdlane(v, 48), rdlane(i, 48));
    v = bv;
    i = bi;
}

// One workgroup of PT threads per matrix factors the panel's columns [k0, k0+nb),
// S columns (a sub-panel) at a time: row g of the sub-panel [c0, c0+S) lives in thread
// (g-c0) % PT, register row (g-c0) / PT.  Per column: the pivot search is a DPP
// arg-max in each wave and across the waves' results after one barrier; the pivot row
// travels through LDS.  Per sub-panel: its interchanges are applied to the panel's
// other columns, U12 of the columns right of it is solved, and those columns get the
// rank-S update.  The pivots are recorded and the row maps built from them afterwards,
// one lane per touched row.
template <class T, int S, int R, int PT, bool RM>
__global__ __launch_bounds__(PT) void lu_panel_kernel(cx<T>* __restrict__ A, int np, long long mstride,
                                                      int* __restrict__ ipiv_all, int* __restrict__ info_all,
                                                      int* __restrict__ pmap_all, int k0, int nb, int xa, int xb) {
    constexpr int PW = PT / 64;
    static_assert(PW <= 16 && 2 * S <= 64, "panel shape");
    __shared__ T s_abs[2][PW];
    __shared__ int s_idx[2][PW];
    __shared__ cx<T> s_row[2][PW][S];
    __shared__ cx<T> s_rowj[2][S];
    __shared__ cx<T> s_L[S][S];
    __shared__ cx<T> s_U[S][64];
    __shared__ int s_piv[64];
    __shared__ int s_sd[2 * S], s_ss[2 * S], s_scnt;
    __shared__ int s_pd[128], s_ps[128], s_pcnt;        // the panel's composite map, for columns [xa, xb)
    // L rows of the current sub-panel for the LDS-broadcast rank-S update, PT rows at a
    // time, whenever they fit beside the arrays above within the 160 KB of LDS.  The
    // 256-thread shapes then need 45 KB or less and half the register file, which leaves
    // room on the CU for the GEMM workgroups of other streams.
    constexpr bool LDSU = (size_t)PT * S * sizeof(cx<T>) <= 139264;
    __shared__ cx<T> sL[LDSU ? S : 1][LDSU ? PT : 1];

    const int b = blockIdx.x;
    cx<T>* M = A + (size_t)b * mstride;
    int* ipiv = ipiv_all + (size_t)b * np;
    const int nq = np / nb, q = k0 / nb;
    int* pmap = pmap_all + ((size_t)b * nq + q) * (1 + 4 * nb);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int kend = min(k0 + nb, np);

    for (int c0 = k0; c0 < kend; c0 += S) {
        cx<T> x[R][S];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int g = c0 + tid + PT * r;
#pragma unroll
            for (int c = 0; c < S; c++)
                x[r][c] = (g < np) ? M[at<RM>(g, c0 + c, np)] : cx<T>{T(0), T(0)};
        }
        // Column loop with a rotating register file: at step jj, x[r][c] holds column
        // c0 + (jj + c) % S, so the active column is always x[r][0] and every register
        // index is a compile-time constant.  The loop is unrolled, so the rotation is
        // register renaming (rolled, its moves cost ~10% of the column phase).
#pragma unroll
        for (int jj = 0; jj < S; jj++) {
            const int j = c0 + jj, buf = jj & 1;
            T best = T(-1);
            int bidx = 0x7fffffff, br = -1;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int g = c0 + tid + PT * r;
                if (g >= j && g < np) {
                    T v = abs1(x[r][0]);
                    if (v != v) v = T(INFINITY);    // NaN competes as infinite: the pivot is always a row in [j, np)
                    if (v > best) { best = v; bidx = g; br = r; }
                }
            }
            const int mine = bidx;
            key_wave(best, bidx);
            if (lane == 0) { s_abs[buf][wv] = best; s_idx[buf][wv] = bidx; }
            if (br >= 0 && mine == bidx) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (r == br)
#pragma unroll
                        for (int c = 0; c < S; c++) s_row[buf][wv][(jj + c) % S] = x[r][c];
            }
            if (tid == jj) {        // row j = c0 + jj lives in thread jj, register row 0
#pragma unroll
                for (int c = 0; c < S; c++) s_rowj[buf][(jj + c) % S] = x[0][c];
            }
            __syncthreads();
            T wvv = lane < PW ? s_abs[buf][lane] : T(-1);
            int wvi = lane < PW ? s_idx[buf][lane] : 0x7fffffff;
            key_row16(wvv, wvi);                        // the PW <= 16 wave keys sit in lanes 0..15
            const int p = rdlane(wvi, 0);
            const cx<T>* piv = s_row[buf][((p - c0) % PT) >> 6];
            const cx<T> pv = piv[jj];
            const bool zero = (pv.x == T(0) && pv.y == T(0));
            if (tid == 0) {
                ipiv[j] = p + 1;
                s_piv[j - k0] = p;
                if (zero && info_all[b] == 0) info_all[b] = j + 1;
            }
            if (p != j) {
                const int op = (p - c0) % PT, rp = (p - c0) / PT;
                if (tid == op) {
#pragma unroll
                    for (int r = 0; r < R; r++)
                        if (r == rp)
#pragma unroll
                            for (int c = 0; c < S; c++) x[r][c] = s_rowj[buf][(jj + c) % S];
                }
                if (tid == jj) {
#pragma unroll
                    for (int c = 0; c < S; c++) x[0][c] = piv[(jj + c) % S];
                }
            }
            const cx<T> rinv = zero ? cx<T>{T(1), T(0)} : inv(pv);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int g = c0 + tid + PT * r;
                if (g > j && g < np) {
                    const cx<T> l = x[r][0] * rinv;
                    x[r][0] = l;
#pragma unroll
                    for (int c = 1; c < S; c++)
                        if (jj + c < S) x[r][c] = x[r][c] - l * piv[jj + c];
                }
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                const cx<T> t = x[r][0];
#pragma unroll
                for (int c = 0; c < S - 1; c++) x[r][c] = x[r][c + 1];
                x[r][S - 1] = t;
            }
        }
        // write the factored sub-panel back
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int g = c0 + tid + PT * r;
            if (g < np)
#pragma unroll
                for (int c = 0; c < S; c++) M[at<RM>(g, c0 + c, np)] = x[r][c];
        }
        if (tid < S) {
#pragma unroll
            for (int c = 0; c < S; c++) s_L[tid][c] = x[0][c];
        }
        __syncthreads();
        const int nleft = c0 - k0, nright = kend - c0 - S;
        if (nleft + nright > 0) {
            // composite map of this sub-panel's S interchanges (rows c0 + k <-> pk): lane
            // k < S takes row c0 + k, lane S + k row pk if below the sub-panel and not
            // named by an earlier pivot; each traces its row back through the
            // interchanges to the row whose content ends there
            if (wv == 0) {
                int pk[S];
#pragma unroll
                for (int k = 0; k < S; k++) pk[k] = s_piv[c0 - k0 + k];
                int pos = -1;
                if (lane < S) {
                    pos = c0 + lane;
                } else if (lane < 2 * S) {
                    const int l2 = lane - S, qv = s_piv[c0 - k0 + l2];
                    bool dup = qv < c0 + S;
#pragma unroll
                    for (int k = 0; k < S; k++) dup = dup || (k < l2 && pk[k] == qv);
                    if (!dup) pos = qv;
                }
                int src = pos;
#pragma unroll
                for (int k = S - 1; k >= 0; k--) {
                    const int a = c0 + k;
                    if (src == a) src = pk[k];
                    else if (src == pk[k]) src = a;
                }
                const bool act = pos >= 0 && src != pos;
                const unsigned long long m = __ballot(act);
                if (act) {
                    const int slot = __popcll(m & ((1ull << lane) - 1ull));
                    s_sd[slot] = pos;
                    s_ss[slot] = src;
                }
                if (lane == 0) s_scnt = __popcll(m);
            }
            __syncthreads();
            // apply it to the panel columns outside the sub-panel: gather all, then scatter
            const int cnt = s_scnt, ncol = nleft + nright, pairs = cnt * ncol;
            cx<T> tmp[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int idx = tid + PT * u;
                tmp[u] = cx<T>{T(0), T(0)};     // written on every path: the array stays in registers
                if (idx < pairs) {
                    const int e = RM ? idx / ncol : idx % cnt, ci = RM ? idx % ncol : idx / cnt;
                    const int col = ci < nleft ? k0 + ci : c0 + S + (ci - nleft);
                    tmp[u] = M[at<RM>(s_ss[e], col, np)];
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int idx = tid + PT * u;
                if (idx < pairs) {
                    const int e = RM ? idx / ncol : idx % cnt, ci = RM ? idx % ncol : idx / cnt;
                    const int col = ci < nleft ? k0 + ci : c0 + S + (ci - nleft);
                    M[at<RM>(s_sd[e], col, np)] = tmp[u];
                }
            }
            __syncthreads();
        }
        if (nright > 0) {
            // U12' = L11'^-1 A12' for the panel columns right of the sub-panel
            if (tid < nright) {          // solved in LDS: the registers hold the sub-panel
                const int col = c0 + S + tid;
                for (int i = 0; i < S; i++) s_U[i][tid] = M[at<RM>(c0 + i, col, np)];
                for (int i = 1; i < S; i++) {
                    cx<T> acc = s_U[i][tid];
                    for (int m = 0; m < i; m++) acc = acc - s_L[i][m] * s_U[m][tid];
                    s_U[i][tid] = acc;
                }
                for (int i = 0; i < S; i++) M[at<RM>(c0 + i, col, np)] = s_U[i][tid];
            }
            __syncthreads();
            if constexpr (LDSU) {
                // rank-S update with the sub-panel's L rows broadcast through LDS, one
                // register row (PT rows of the panel) at a time
#pragma unroll
                for (int r = 0; r < R; r++) {
                    // rows c0 + [lo, hi) are those of register row r below the sub-panel
                    const int lo = max(PT * r, S), hi = min(PT * (r + 1), np - c0);
                    if (lo >= hi) continue;                 // uniform
#pragma unroll
                    for (int m = 0; m < S; m++) sL[m][tid] = x[r][m];
                    __syncthreads();
                    const int base = PT * r;
                    if constexpr (RM) {
                        // one column per thread (U column in registers), rows strided by the
                        // number of row groups, UR rows in flight; consecutive threads take
                        // consecutive columns of a row: contiguous in row-major storage
                        constexpr int UR = 8;
                        const int rg = PT / nright, ci = tid % nright, rq = tid / nright;
                        if (rq < rg) {
                            cx<T> u[S];
#pragma unroll
                            for (int m = 0; m < S; m++) u[m] = s_U[m][ci];
                            for (int rb = lo + rq; rb < hi; rb += rg * UR) {
                                cx<T> a[UR];
#pragma unroll
                                for (int t = 0; t < UR; t++) {
                                    const int ri = rb + t * rg;
                                    a[t] = cx<T>{T(0), T(0)};
                                    if (ri < hi) a[t] = M[at<true>(c0 + ri, c0 + S + ci, np)];
                                }
#pragma unroll
                                for (int t = 0; t < UR; t++) {
                                    const int ri = rb + t * rg;
                                    if (ri < hi) {
                                        cx<T> v = a[t];
#pragma unroll
                                        for (int m = 0; m < S; m++) v = v - sL[m][ri - base] * u[m];
                                        M[at<true>(c0 + ri, c0 + S + ci, np)] = v;
                                    }
                                }
                            }
                        }
                    } else {
                        // every thread updates whole (row, column) elements in storage order
                        const int nr = hi - lo, tot = nr * nright;
#pragma unroll 2
                        for (int e = tid; e < tot; e += PT) {
                            const int ri = lo + e % nr, ci = e / nr;
                            const size_t at_e = at<false>(c0 + ri, c0 + S + ci, np);
                            cx<T> a = M[at_e];
#pragma unroll
                            for (int m = 0; m < S; m++) a = a - sL[m][ri - base] * s_U[m][ci];
                            M[at_e] = a;
                        }
                    }
                    __syncthreads();
                }
                continue;
            }
            // rank-S update of rows below the sub-panel, WR columns per round (the loads of a
            // round are in flight together; shapes of <= 512 threads with few register rows
            // have the registers for 16)
            constexpr int WR = (PT <= 512 && R * S * (int)sizeof(T) <= 128) ? 16 : (sizeof(T) == 8 ? 4 : 8);
            for (int w0 = 0; w0 < nright; w0 += WR) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int g = c0 + tid + PT * r;
                    if (g < c0 + S || g >= np) continue;
                    cx<T> av[WR];
#pragma unroll
                    for (int w = 0; w < WR; w++)
                        if (w0 + w < nright) av[w] = M[at<RM>(g, c0 + S + w0 + w, np)];
#pragma unroll
                    for (int w = 0; w < WR; w++) {
                        if (w0 + w < nright) {
                            cx<T> a = av[w];
#pragma unroll
                            for (int m = 0; m < S; m++) a = a - x[r][m] * s_U[m][w0 + w];
                            M[at<RM>(g, c0 + S + w0 + w, np)] = a;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    // publish the panel's composite row map (wave 0; the last barrier of the loop makes
    // every pivot visible): lane k takes row k0 + k and row pk when pk >= kend and no
    // earlier pivot named it, and traces both back through the npv interchanges
    if (wv == 0) {
        const int npv = kend - k0;
        int pa = lane < npv ? k0 + lane : -1, pb = -1;
        if (lane < npv) {
            const int qv = s_piv[lane];
            if (qv >= kend) pb = qv;
        }
        int sa = pa, sb = pb;
        bool dup = false;
#pragma unroll 8
        for (int k = npv - 1; k >= 0; k--) {
            const int a = k0 + k, pk = s_piv[k];
            if (sa == a) sa = pk;
            else if (sa == pk) sa = a;
            if (sb == a) sb = pk;
            else if (sb == pk) sb = a;
            dup = dup || (k < lane && pk == pb);
        }
        const bool acta = pa >= 0 && sa != pa, actb = pb >= 0 && !dup && sb != pb;
        const unsigned long long ma = __ballot(acta), mb = __ballot(actb);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (acta) {
            const int slot = __popcll(ma & below);
            pmap[1 + slot] = pa;
            pmap[1 + 2 * nb + slot] = sa;
            s_pd[slot] = pa;
            s_ps[slot] = sa;
        }
        if (actb) {
            const int slot = __popcll(ma) + __popcll(mb & below);
            pmap[1 + slot] = pb;
            pmap[1 + 2 * nb + slot] = sb;
            s_pd[slot] = pb;
            s_ps[slot] = sb;
        }
        if (lane == 0) {
            pmap[0] = __popcll(ma) + __popcll(mb);
            s_pcnt = __popcll(ma) + __popcll(mb);
        }
    }
    if (xb > xa) {
        // the panel's interchanges on the extra columns [xa, xb) (at most 64, outside the
        // panel): every row of a round's columns is read before any is written
        __syncthreads();
        const int cnt = s_pcnt, ncol = xb - xa;
        if (cnt == 0) return;
        const int rw = max(1, min(ncol, (4 * PT) / cnt));
        for (int cb = 0; cb < ncol; cb += rw) {
            const int w = min(rw, ncol - cb), pairs = cnt * w;
            cx<T> tmp[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int idx = tid + PT * u;
                tmp[u] = cx<T>{T(0), T(0)};     // written on every path: the array stays in registers
                if (idx < pairs) {
                    const int e = RM ? idx / w : idx % cnt, c = xa + cb + (RM ? idx % w : idx / cnt);
                    tmp[u] = M[at<RM>(s_ps[e], c, np)];
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int idx = tid + PT * u;
                if (idx < pairs) {
                    const int e = RM ? idx / w : idx % cnt, c = xa + cb + (RM ? idx % w : idx / cnt);
                    M[at<RM>(s_pd[e], c, np)] = tmp[u];
                }
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------ in-register triangular solves
// Value of lane q of each quad (DPP quad_perm [q,q,q,q]); q is a constant after unrolling.
template <class T>
__device__ __forceinline__ T quad_bcast(T v, int q) {
    switch (q) {
        case 0: return dpp_v<0x00>(v);
        case 1: return dpp_v<0x55>(v);
        case 2: return dpp_v<0xAA>(v);
        default: return dpp_v<0xFF>(v);
    }
}

// L(row, i) from packed strictly-lower storage (row r at offset r (r - 1) / 2; unit
// diagonal only).
template <class T>
struct PackedL {
    const cx<T>* s;
    __device__ __forceinline__ cx<T> operator()(int r, int c) const { return s[(r * (r - 1) >> 1) + c]; }
};
// Packed triangles with their diagonal: lower (c <= r) by rows, upper (r <= c) by columns.
template <class T>
struct PackedLD {
    const cx<T>* s;
    __device__ __forceinline__ cx<T> operator()(int r, int c) const { return s[(r * (r + 1) >> 1) + c]; }
};
template <class T>
struct PackedUD {
    const cx<T>* s;
    __device__ __forceinline__ cx<T> operator()(int r, int c) const { return s[(c * (c + 1) >> 1) + r]; }
};

// Four lanes per right-hand side (a quad): lane qq of a quad holds rows 4m+qq in xv[m],
// and each solved row reaches the other lanes of its quad by a DPP broadcast.  The
// register file is rotated by one after every 4 rows, so the row being eliminated is
// always xv[0] (forward) or xv[15] (backward) and every register index is a
// compile-time constant while the 16-step loop stays rolled.  Called by whole waves.
template <class T, bool UNIT = true, class LA>
__device__ __forceinline__ void lower_unit_solve(cx<T> (&xv)[16], LA sL, int nb, int qq, int /*lane*/) {
#pragma unroll 1
    for (int ib = 0; ib < 16; ib++) {
#pragma unroll
        for (int iq = 0; iq < 4; iq++) {
            const int i = 4 * ib + iq;
            if (!UNIT && qq == iq && i < nb) xv[0] = xv[0] / sL(i, i);
            const cx<T> xi{quad_bcast(xv[0].x, iq), quad_bcast(xv[0].y, iq)};
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int row = 4 * (m + ib) + qq;
                if (m + ib < 16 && row > i && row < nb) xv[m] = xv[m] - sL(row, i) * xi;
            }
        }
        const cx<T> t = xv[0];
#pragma unroll
        for (int m = 0; m < 15; m++) xv[m] = xv[m + 1];
        xv[15] = t;
    }
}

template <class T, bool UNIT = false, class LA>
__device__ __forceinline__ void upper_solve(cx<T> (&xv)[16], LA sL, int qq, int /*lane*/) {
#pragma unroll 1
    for (int ib = 15; ib >= 0; ib--) {
#pragma unroll
        for (int iq = 3; iq >= 0; iq--) {
            const int i = 4 * ib + iq;
            if (!UNIT && qq == iq) xv[15] = xv[15] / sL(i, i);
            const cx<T> xi{quad_bcast(xv[15].x, iq), quad_bcast(xv[15].y, iq)};
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const int row = 4 * (m - (15 - ib)) + qq;
                if (m >= 15 - ib && row < i) xv[m] = xv[m] - sL(row, i) * xi;
            }
        }
        const cx<T> t = xv[15];
#pragma unroll
        for (int m = 15; m > 0; m--) xv[m] = xv[m - 1];
        xv[0] = t;
    }
}

// ------------------------------------------------------------------ swap + TRSM
// grid.x = column tiles of TW columns: the first nleft_t tiles cover [0, k0)
// (LAPACK-format L), the rest cover [kend, np).  The row map is applied RW columns
// per round (every read of a round precedes its writes; up to 128 x RW scattered
// elements in flight per workgroup).  Right tiles then get U12 = L11^-1 A12, 64
// columns at a time against L11 staged once in LDS.
template <class T, int TW, int RW, bool RM>
__global__ __launch_bounds__(256) void swap_trsm_kernel(cx<T>* __restrict__ A, int np, long long mstride,
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
        lower_unit_solve<T>(xv, PackedL<T>{sL}, nb, qq, lane);
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

// (three waves per SIMD: the register budget of the complex128 3M tile)
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void gemm_sub_kernel(
    cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int n, int kk, int lda, int ldb, int ldc, long long sC,
    long long sA, long long sB, int tiles_m, int tiles_n, int nbatch) {
    __shared__ __align__(16) T lds[KC * (2 * GT + PADA) + GT * (2 * KC + PADB)];
    const int tiles = tiles_m * tiles_n;
    int g = blockIdx.x, b, t;
    if ((nbatch & 7) == 0) {   // keep every tile of a matrix on one XCD (dispatch is round-robin over 8)
        const int xcd = g & 7, j = g >> 3;
        b = xcd + 8 * (j / tiles);
        t = j % tiles;
    } else {
        b = g / tiles;
        t = g % tiles;
    }
    const int row0 = (t % tiles_m) * GT, col0 = (t / tiles_m) * GT;
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
    const int b = blockIdx.x, tid = threadIdx.x;
    if (items) n = items[b].n;
    const cx<T>* M = A + (size_t)b * mstride;
    const int* ipiv = ipiv_all + (size_t)b * np;
    double re = 0.0, im = 0.0, mx = -INFINITY, mn = INFINITY;
    int odd = 0;
    for (int i = tid; i < n; i += 256) {
        const cx<T> u = M[(size_t)i + (size_t)i * np];
        const double la = log(hypot((double)u.x, (double)u.y));
        re += la;
        im += atan2((double)u.y, (double)u.x);
        mx = fmax(mx, la);
        mn = fmin(mn, la);
        odd ^= (ipiv[i] != i + 1);
    }
    sre[tid] = re; sim[tid] = im; smax[tid] = mx; smin[tid] = mn; sodd[tid] = odd;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            sre[tid] += sre[tid + s];
            sim[tid] += sim[tid + s];
            smax[tid] = fmax(smax[tid], smax[tid + s]);
            smin[tid] = fmin(smin[tid], smin[tid + s]);
            sodd[tid] ^= sodd[tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        double imt = sim[0] + (sodd[0] ? kPi : 0.0);
        out[b] = cx<double>{sre[0] / n, imt / n};
        if (stats) { stats[2 * b] = smax[0]; stats[2 * b + 1] = smin[0]; }
    }
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

// ------------------------------------------------------------------ launchers
// Panel shapes (threads PT, sub-panel width S, register rows R = rows / PT), chosen
// per precision and padded size from measurements of the panel microbenchmark.
template <class T, int S, int R, int PT>
static void panel_go(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb,
                     int nbatch, hipStream_t st, bool rm, int xa, int xb) {
    if (rm)
        hipLaunchKernelGGL((lu_panel_kernel<T, S, R, PT, true>), dim3(nbatch), dim3(PT), 0, st,
                           A, np, mstride, ipiv, info, pmap, k0, nb, xa, xb);
    else
        hipLaunchKernelGGL((lu_panel_kernel<T, S, R, PT, false>), dim3(nbatch), dim3(PT), 0, st, A, np, mstride,
                           ipiv, info, pmap, k0, nb, xa, xb);
}

// Panel shapes <S, R, PT> per size range, tried by the autotuner; variant 0 is the
// default.  PT x R must cover np rows.  The 1024-thread shapes are limited to 128
// VGPRs per lane; the 512-thread ones may use 256.
template <class T>
void launch_panel(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb, int nbatch,
                  hipStream_t st, bool rm, int variant, int xa, int xb) {
    constexpr bool D = sizeof(T) == 8;
    if (variant < 0 || variant >= kPanelVariants) variant = 0;
    if (variant == 4) {          // 16-column sub-panels: half the passes over the panel's right part
        if (np <= 1024) {
            if (D) panel_go<T, 16, 2, 512>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
            else panel_go<T, 16, 1, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
            return;
        }
        variant = 0;
    }
    if (variant == 3) {          // 256 threads: leaves room on the CU for GEMM waves of another stream
        if (np <= 1024) panel_go<T, 8, 4, 256>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else if (np <= 2048) panel_go<T, 4, 8, 256>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else if (np <= 4096) panel_go<T, 2, 16, 256>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else panel_go<T, 1, 32, 256>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        return;
    }
    // up to N = 2048 the defaults are the 1024-thread shapes whose rank-S update reads the
    // L rows from LDS (at N = 1024-2048 they take 0.4-0.7 of the 512-thread shapes' time)
    if (np <= 1024) {
        if (variant == 1) panel_go<T, 8, 2, 512>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else if (variant == 2) panel_go<T, 4, 1, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else panel_go<T, 8, 1, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
    } else if (np <= 2048) {
        if (variant == 2) panel_go<T, 8, 4, 512>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else if ((variant == 0) == D) panel_go<T, 4, 2, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else panel_go<T, 8, 2, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
    } else if (np <= 4096) {     // complex128 default: the 512-thread shape, which does not spill
        if (variant == 1) panel_go<T, 2, 4, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else if ((variant == 0) == D) panel_go<T, 4, 8, 512>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else panel_go<T, 4, 4, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
    } else {
        if (variant == 1) panel_go<T, 1, 8, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else if ((variant == 0) == D) panel_go<T, 2, 16, 512>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
        else panel_go<T, 2, 8, 1024>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, rm, xa, xb);
    }
}

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
void launch_gemm_sub(cx<T>* C, const cx<T>* A, const cx<T>* B, int m, int n, int k, int lda, int ldb, int ldc,
                     long long sC, long long sA, long long sB, int nbatch, hipStream_t st) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    const int tm = (m + GT - 1) / GT, tn = (n + GT - 1) / GT;
    hipLaunchKernelGGL(gemm_sub_kernel<T>, dim3(tm * tn * nbatch), dim3(256), 0, st, C, A, B, m, n, k, lda, ldb, ldc,
                       sC, sA, sB, tm, tn, nbatch);
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
void launch_swap_left(cx<T>* A, int np, long long mstride, const int* pmap, int k0, int nb, int nbatch,
                      hipStream_t st) {
    launch_swap_range<T>(A, np, mstride, pmap, k0, nb, 0, 0, 0, 1, 0, nbatch, st, 0, false);
}

template <class T>
void launch_pad_identity(cx<T>* A, int n, int np, int nbatch, hipStream_t st) {
    if (np <= n) return;
    hipLaunchKernelGGL(pad_identity_kernel<T>, dim3((np - n + 63) / 64, nbatch), dim3(64), 0, st, A, n, np,
                       (long long)np * np);
}

#define MSG_INST(T)                                                                                                   \
    template void launch_panel<T>(cx<T>*, int, long long, int*, int*, int*, int, int, int, hipStream_t, bool, int, \
                                  int, int);                                                                       \
    template void launch_swap_trsm<T>(cx<T>*, int, long long, const int*, int, int, int, int, hipStream_t, int, bool);\
    template void launch_swap_range<T>(cx<T>*, int, long long, const int*, int, int, int, int, int, int, int, int,    \
                                       hipStream_t, int, bool, int);                                                   \
    template void launch_gemm_sub<T>(cx<T>*, const cx<T>*, const cx<T>*, int, int, int, int, int, int, long long,      \
                                     long long, long long, int, hipStream_t);                                          \
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
    template void launch_swap_left<T>(cx<T>*, int, long long, const int*, int, int, int, hipStream_t);                \
    template void launch_pad_identity<T>(cx<T>*, int, int, int, hipStream_t);
MSG_INST(double)
MSG_INST(float)
template void launch_convert<double, float>(const cx<double>*, cx<float>*, long long, hipStream_t);
template void launch_convert<float, double>(const cx<float>*, cx<double>*, long long, hipStream_t);

}  // namespace msg
