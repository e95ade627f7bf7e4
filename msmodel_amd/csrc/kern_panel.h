// kern_panel.h -- the panel factorisation kernel of the blocked LU and its
// launcher.  With MSG_PANEL_IMPL defined it holds the kernel, instantiated per
// element type and storage order by kern_panel_{c128,c64}_{rm,cm}.hip (translation
// units that compile in parallel); without, the launcher's declaration only.
#pragma once
#include "lu_device.h"

namespace msg {

// The panel launcher of one element type and storage order (RM: row-major); each
// (type, order) pair is instantiated in its own translation unit.
template <class T, bool RM>
void launch_panel_rm(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb,
                     int nbatch, hipStream_t st, int variant, int xa, int xb);

#ifdef MSG_PANEL_IMPL


// One workgroup of PT threads per matrix factors the panel's columns [k0, k0+nb),
// S columns (a sub-panel) at a time: row g of the sub-panel [c0, c0+S) lives in thread
// (g-c0) % PT, register row (g-c0) / PT.  Per column: the pivot search is a DPP
// arg-max in each wave and across the waves' results after one barrier; the pivot row
// travels through LDS.  Per sub-panel: its interchanges are applied to the panel's
// other columns, U12 of the columns right of it is solved, and those columns get the
// rank-S update.  The pivots are recorded and the row maps built from them afterwards,
// one lane per touched row.
//
// LL (left-looking, Crout order within the panel): a sub-panel receives all updates of
// the panel's earlier sub-panels when it is loaded instead of after each of them, so the
// columns right of a sub-panel are only read once and written once (their row
// interchanges aside).  For the sub-panel [c0, c0+S) with K = c0 - k0 factored columns:
//   U(k0:c0, c0:c0+S) = L(k0:c0, k0:c0)^-1 A(k0:c0, c0:c0+S)    (unit lower, staged in LDS;
//        one lane per row, the solved rows broadcast by readlane)
//   A(c0:np, c0:c0+S) -= L(c0:np, k0:c0) U(k0:c0, c0:c0+S)       (L read from the stored,
//        already interchanged columns; U broadcast from LDS)
// then the column loop as in the right-looking form.  The panel moves about 2.8 times its
// algorithmic bytes with S = 8 instead of about 4.5.
template <class T, int S, int R, int PT, bool RM, bool LL>
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
    __shared__ cx<T> s_U[LL ? 1 : S][64];
    __shared__ int s_piv[64];
    __shared__ int s_sd[2 * S], s_ss[2 * S], s_scnt;
    __shared__ int s_pd[128], s_ps[128], s_pcnt;        // the panel's composite map, for columns [xa, xb)
    // L rows of the current sub-panel for the LDS-broadcast rank-S update, PT rows at a
    // time, whenever they fit beside the arrays above within the 160 KB of LDS.  The
    // 256-thread shapes then need 45 KB or less and half the register file, which leaves
    // room on the CU for the GEMM workgroups of other streams.
    constexpr bool LDSU = !LL && (size_t)PT * S * sizeof(cx<T>) <= 139264;
    __shared__ cx<T> sL[LDSU ? S : 1][LDSU ? PT : 1];
    // LL: the strictly lower part of the panel's factored top rows (packed, row r at
    // r (r - 1) / 2) and the current sub-panel's top rows
    __shared__ cx<T> sLt[LL ? 64 * 63 / 2 : 1];
    __shared__ cx<T> sUt[LL ? 64 : 1][S];

    const int b = blockIdx.x;
    cx<T>* M = A + (size_t)b * mstride;
    int* ipiv = ipiv_all + (size_t)b * np;
    const int nq = np / nb, q = k0 / nb;
    int* pmap = pmap_all + ((size_t)b * nq + q) * (1 + 4 * nb);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int kend = min(k0 + nb, np);

    for (int c0 = k0; c0 < kend; c0 += S) {
        const int K = c0 - k0;
        if constexpr (LL) {
            if (K > 0) {
                // stage the unit-lower L of rows [k0, c0) and the sub-panel's top rows
                for (int e = tid; e < K * K; e += PT) {
                    const int i = RM ? e / K : e % K, m = RM ? e % K : e / K;
                    if (i > m) sLt[(i * (i - 1) >> 1) + m] = M[at<RM>(k0 + i, k0 + m, np)];
                }
                for (int e = tid; e < K * S; e += PT) {
                    const int i = RM ? e / S : e % K, c = RM ? e % S : e / K;
                    sUt[i][c] = M[at<RM>(k0 + i, c0 + c, np)];
                }
                __syncthreads();
                // U(k0:c0, c) by forward substitution: lane i holds row i; row m, once
                // final, reaches the rows below it by readlane
                for (int c = wv; c < S; c += PW) {
                    cx<T> u = lane < K ? sUt[lane][c] : cx<T>{T(0), T(0)};
                    for (int m = 0; m + 1 < K; m++) {
                        const cx<T> um{rdlane(u.x, m), rdlane(u.y, m)};
                        if (lane > m && lane < K) u = u - sLt[(lane * (lane - 1) >> 1) + m] * um;
                    }
                    if (lane < K) {
                        sUt[lane][c] = u;
                        M[at<RM>(k0 + lane, c0 + c, np)] = u;
                    }
                }
                __syncthreads();
            }
        }
        cx<T> x[R][S];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int g = c0 + tid + PT * r;
#pragma unroll
            for (int c = 0; c < S; c++)
                x[r][c] = (g < np) ? M[at<RM>(g, c0 + c, np)] : cx<T>{T(0), T(0)};
        }
        if constexpr (LL && RM) {
            // rows [c0, np): A(g, c0:c0+S) -= L(g, k0:c0) U(k0:c0, c0:c0+S).  Row-major: a
            // thread reads eight consecutive entries of its row (one 128-byte line) at a
            // time, so every line fetched is used whole
            for (int m0 = 0; m0 < K; m0 += 8) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int g = c0 + tid + PT * r;
                    if (g >= np) continue;
                    cx<T> l[8];
#pragma unroll
                    for (int h = 0; h < 8; h++)
                        l[h] = (m0 + h < K) ? M[at<true>(g, k0 + m0 + h, np)] : cx<T>{T(0), T(0)};
#pragma unroll
                    for (int h = 0; h < 8; h++) {
                        if (m0 + h >= K) break;
#pragma unroll
                        for (int c = 0; c < S; c++) x[r][c] = x[r][c] - l[h] * sUt[m0 + h][c];
                    }
                }
            }
        }
        if constexpr (LL && !RM) {
            // rows [c0, np): A(g, c0:c0+S) -= L(g, k0:c0) U(k0:c0, c0:c0+S).  Column-major:
            // the threads' rows are consecutive in memory, two L columns per round in flight
            for (int m = 0; m < K; m += 2) {
                cx<T> l[R][2];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int g = c0 + tid + PT * r;
#pragma unroll
                    for (int h = 0; h < 2; h++)
                        l[r][h] = (g < np && m + h < K) ? M[at<RM>(g, k0 + m + h, np)] : cx<T>{T(0), T(0)};
                }
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    if (m + h >= K) break;
#pragma unroll
                    for (int c = 0; c < S; c++) {
                        const cx<T> u = sUt[m + h][c];
#pragma unroll
                        for (int r = 0; r < R; r++) x[r][c] = x[r][c] - l[r][h] * u;
                    }
                }
            }
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
        if (!LL && nright > 0) {
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

// ------------------------------------------------------------------ launchers
// Panel shapes (threads PT, sub-panel width S, register rows R = rows / PT), chosen
// per precision and padded size from measurements of the panel microbenchmark.
template <class T, bool RM, int S, int R, int PT, bool LL = false>
static void panel_go(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb,
                     int nbatch, hipStream_t st, int xa, int xb) {
    hipLaunchKernelGGL((lu_panel_kernel<T, S, R, PT, RM, LL>), dim3(nbatch), dim3(PT), 0, st, A, np, mstride,
                       ipiv, info, pmap, k0, nb, xa, xb);
}

// Panel shapes <S, R, PT> per size range, tried by the autotuner; variant 0 is the
// default.  PT x R must cover np rows.  The 1024-thread shapes are limited to 128
// VGPRs per lane; the 512-thread ones may use 256.
template <class T, bool RM, bool LL>
static void panel_dispatch(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb,
                     int nbatch, hipStream_t st, int variant, int xa, int xb) {
    constexpr bool D = sizeof(T) == 8;
    if (variant == 4) {          // 16-column sub-panels: half the passes over the panel's right part
        if (np <= 1024) {
            if (D) panel_go<T, RM, 16, 2, 512, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
            else panel_go<T, RM, 16, 1, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
            return;
        }
        variant = 0;
    }
    if (variant == 3) {          // 256 threads: leaves room on the CU for GEMM waves of another stream
        // (as many register rows as the padded size needs: rows beyond np would still cost
        // issue cycles in the pivot search and the rank-1 updates)
        if (np <= 256) panel_go<T, RM, 8, 1, 256, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if (np <= 512) panel_go<T, RM, 8, 2, 256, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if (np <= 1024) panel_go<T, RM, 8, 4, 256, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if (np <= 2048) panel_go<T, RM, 4, 8, 256, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if (np <= 4096) panel_go<T, RM, 2, 16, 256, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else panel_go<T, RM, 1, 32, 256, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        return;
    }
    // up to N = 2048 the defaults are the 1024-thread shapes whose rank-S update reads the
    // L rows from LDS (at N = 1024-2048 they take 0.4-0.7 of the 512-thread shapes' time)
    if (np <= 1024) {
        if (variant == 1) panel_go<T, RM, 8, 2, 512, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if (variant == 2) panel_go<T, RM, 4, 1, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else panel_go<T, RM, 8, 1, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
    } else if (np <= 2048) {
        if (variant == 2) panel_go<T, RM, 8, 4, 512, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if ((variant == 0) == D) panel_go<T, RM, 4, 2, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else panel_go<T, RM, 8, 2, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
    } else if (np <= 4096) {     // complex128 default: the 512-thread shape, which does not spill
        if (variant == 1) panel_go<T, RM, 2, 4, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if ((variant == 0) == D) panel_go<T, RM, 4, 8, 512, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else panel_go<T, RM, 4, 4, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
    } else {
        if (variant == 1) panel_go<T, RM, 1, 8, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else if ((variant == 0) == D) panel_go<T, RM, 2, 16, 512, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
        else panel_go<T, RM, 2, 8, 1024, LL>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, xa, xb);
    }
}

// Variants 0-4: right-looking shapes; 5, 6, 7: the left-looking form (LL) of variants 3,
// 0 and 4.
template <class T, bool RM>
void launch_panel_rm(cx<T>* A, int np, long long mstride, int* ipiv, int* info, int* pmap, int k0, int nb,
                     int nbatch, hipStream_t st, int variant, int xa, int xb) {
    if (variant < 0 || variant >= kPanelVariants) variant = 0;
    if (variant >= 5)
        panel_dispatch<T, RM, true>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, variant == 5 ? 3 : variant == 6 ? 0 : 4,
                                    xa, xb);
    else
        panel_dispatch<T, RM, false>(A, np, mstride, ipiv, info, pmap, k0, nb, nbatch, st, variant, xa, xb);
}

#endif  // MSG_PANEL_IMPL

}  // namespace msg
