// Dense LU factorisation with partial pivoting (FP64) for the AIC solve.
//
// Replaces DGETRF (inv2, libMath.f90:48-83; the native Doolittle routine of
// libMath_noBLAS.f90 in no-BLAS builds).  Column-major, 0-based pivots,
// info = first zero pivot (1-based) or 0.  Pivot choice follows IDAMAX:
// the first row of largest magnitude.
//
// n <= LU_SMALL (the AIC of one rotor): one workgroup, unblocked.
// Larger n: right-looking blocked LU with panels of NB columns.
//   1. panel: one workgroup factorises A[k:n, k:k+NB] column by column,
//      holding the panel in registers (LDS for the row exchange); the row
//      swaps then reach the other columns through one composed permutation;
//   2. U12 = L11^-1 A12 (unit lower triangular solve, one lane per column);
//   3. A22 -= L21 U12 on the matrix cores: 64x64 output tiles, four waves
//      of 2x2 v_mfma_f64_16x16x4f64 blocks, operands staged in LDS.
//      Lane layout of the 16x16x4 FP64 MFMA (checked on gfx950): lane l
//      supplies A[l%16][l/16] and B[l/16][l%16]; result register i of lane
//      l is D[4i + l/16][l%16].
#include "vlm_internal.h"

namespace vlm {
namespace {

constexpr int LU_THREADS = 1024;
constexpr int LU_WAVES = LU_THREADS / 64;
constexpr int LU_SMALL = 256;  // one-workgroup path up to this order
constexpr int NB = 16;         // panel width of the blocked path
constexpr int REG_THREADS = 512;  // workgroup of the register panel (256 VGPRs per lane)
constexpr int REG_WAVES = REG_THREADS / 64;
constexpr int RPT = 3;            // panel rows per lane in the register panel (m <= 1536)
// Row-move list of one panel: [count, dst0, src0, dst1, src1, ...], at most
// 2 NB moved rows (absolute row indices; the panel's swaps composed).
constexpr int LSTRIDE = 1 + 4 * NB;

// IDAMAX order: larger magnitude wins, ties go to the lower row.
__device__ __forceinline__ void better(double &bv, int &bi, double v, int i)
{
    const bool take = v > bv || (v == bv && i < bi);
    bv = take ? v : bv;
    bi = take ? i : bi;
}

// Unblocked right-looking LU of columns [k0, k1) over rows [k0, n).  Row
// swaps are applied to columns [s0, s1) (the whole matrix for the
// one-workgroup path, the panel only in the blocked path), the rank-1
// updates to columns (k, k1).
__global__ __launch_bounds__(LU_THREADS) void lu_panel_kernel(int n, double *A, int *ipiv,
                                                              int *info, int k0, int k1, int s0,
                                                              int s1)
{
    __shared__ double sv[LU_WAVES];
    __shared__ int si[LU_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const size_t ld = (size_t)n;
    if (*info != 0) {  // an earlier panel hit a zero pivot: keep pivots in range
        for (int k = k0 + tid; k < k1; k += blockDim.x) ipiv[k] = k;
        return;
    }
    // 2-D split of the trailing update: 64 consecutive rows per wave
    // (coalesced), LU_WAVES column groups.
    for (int k = k0; k < k1; ++k) {
        double bv = -1.0;
        int bi = k;
        for (int i = k + tid; i < n; i += LU_THREADS) better(bv, bi, fabs(A[i + ld * k]), i);
        for (int off = 32; off > 0; off >>= 1)
            better(bv, bi, __shfl_xor(bv, off), __shfl_xor(bi, off));
        if (lane == 0) {
            sv[w] = bv;
            si[w] = bi;
        }
        __syncthreads();
        bv = sv[0];
        bi = si[0];
        for (int q = 1; q < LU_WAVES; ++q) better(bv, bi, sv[q], si[q]);
        const int p = bi;
        if (tid == 0) ipiv[k] = p;
        if (bv == 0.0) {
            if (tid == 0 && *info == 0) *info = k + 1;
            for (int j = k + 1 + tid; j < k1; j += LU_THREADS) ipiv[j] = j;
            return;
        }
        if (p != k)
            for (int j = s0 + tid; j < s1; j += LU_THREADS) {
                const double t = A[k + ld * j];
                A[k + ld * j] = A[p + ld * j];
                A[p + ld * j] = t;
            }
        __syncthreads();
        const double inv = 1.0 / A[k + ld * k];
        for (int j = k + 1 + w; j < k1; j += LU_WAVES) {
            const double u = A[k + ld * j];
            for (int i = k + 1 + lane; i < n; i += 64) A[i + ld * j] -= (A[i + ld * k] * inv) * u;
        }
        __syncthreads();
        for (int i = k + 1 + tid; i < n; i += LU_THREADS) A[i + ld * k] *= inv;
    }
}


// Panel of the blocked path held on chip: lane t owns the physical panel
// rows t + REG_THREADS*q, RPT of them in registers and RL more in LDS
// (columns [k0, k0+jb), jb <= NB).  Rows never move during the panel:
// pos is the logical row (after the swaps so far) of each owned row and
// phys[] in LDS maps logical rows back, so a row swap is two index
// updates.  Pivot search and rank-1 update touch no global memory; each row
// is read once and written once, to its final logical position.  The
// column loop is unrolled by template recursion so that every register
// index is a constant.
struct PanelCtx {
    double *sv, *prow, *L;   // L: LDS rows, L[(q*NB + j)*REG_THREADS + tid]
    int *si, *phys, *ipiv, *info;
    int tid, lane, w, k0, m, jb;
    bool dead;
};

template <int RL, int C>
__device__ __forceinline__ void panel_column(double (&a)[RPT][NB], int (&pos)[RPT + RL],
                                             PanelCtx &P)
{
    double *L = P.L + P.tid;
    if (C < P.jb && !P.dead) {
        // pivot: largest |a| among logical rows >= C, first logical row on ties
        double bv = -1.0;
        int bi = P.m;
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (pos[q] >= C && pos[q] < P.m) better(bv, bi, fabs(a[q][C]), pos[q]);
#pragma unroll
        for (int q = 0; q < RL; ++q)
            if (pos[RPT + q] >= C && pos[RPT + q] < P.m)
                better(bv, bi, fabs(L[(q * NB + C) * REG_THREADS]), pos[RPT + q]);
        for (int off = 32; off > 0; off >>= 1)
            better(bv, bi, __shfl_xor(bv, off), __shfl_xor(bi, off));
        if (P.lane == 0) {
            P.sv[P.w] = bv;
            P.si[P.w] = bi;
        }
        __syncthreads();
        bv = P.sv[0];
        bi = P.si[0];
        for (int q = 1; q < REG_WAVES; ++q) better(bv, bi, P.sv[q], P.si[q]);
        const int lp = bi;                 // logical row of the pivot
        const int pp = P.phys[lp], pc = P.phys[C];
        if (P.tid == 0) P.ipiv[P.k0 + C] = P.k0 + lp;
        if (bv == 0.0) {
            P.dead = true;
            if (P.tid == 0 && *P.info == 0) *P.info = P.k0 + C + 1;
            for (int j = C + 1 + P.tid; j < P.jb; j += REG_THREADS) P.ipiv[P.k0 + j] = P.k0 + j;
        } else {
#pragma unroll
            for (int q = 0; q < RPT + RL; ++q) {
                const int r = P.tid + REG_THREADS * q;
                if (r == pp) {
#pragma unroll
                    for (int j = 0; j < NB; ++j)
                        P.prow[j] = q < RPT ? a[q < RPT ? q : 0][j]
                                            : L[((q - RPT) * NB + j) * REG_THREADS];
                    pos[q] = C;
                } else if (r == pc) {
                    pos[q] = lp;
                }
            }
            __syncthreads();   // prow published; every lane has read phys[]
            if (P.tid == 0) {
                P.phys[C] = pp;
                P.phys[lp] = pc;
            }
            const double inv = 1.0 / P.prow[C];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (pos[q] > C && pos[q] < P.m) {
                    const double l = a[q][C] * inv;
                    a[q][C] = l;
#pragma unroll
                    for (int j = C + 1; j < NB; ++j) a[q][j] -= l * P.prow[j];
                }
            }
#pragma unroll
            for (int q = 0; q < RL; ++q) {
                if (pos[RPT + q] > C && pos[RPT + q] < P.m) {
                    double *row = L + (q * NB) * REG_THREADS;
                    const double l = row[C * REG_THREADS] * inv;
                    row[C * REG_THREADS] = l;
#pragma unroll
                    for (int j = C + 1; j < NB; ++j) row[j * REG_THREADS] -= l * P.prow[j];
                }
            }
        }
    }
    if constexpr (C + 1 < NB) panel_column<RL, C + 1>(a, pos, P);
}

template <int RL>
__global__ __launch_bounds__(REG_THREADS) void lu_panel_reg_kernel(int n, double *A, int *ipiv,
                                                                   int *info, int k0, int jb,
                                                                   int *list)
{
    constexpr int ROWS = (RPT + RL) * REG_THREADS;
    __shared__ double sv[REG_WAVES];
    __shared__ int si[REG_WAVES];
    __shared__ double prow[NB];
    __shared__ int phys[ROWS];
    __shared__ double Lrows[RL > 0 ? RL * NB * REG_THREADS : 1];
    __shared__ int moved;
    const int tid = threadIdx.x;
    const size_t ld = (size_t)n;
    const int m = n - k0;
    if (*info != 0) {
        for (int k = tid; k < jb; k += REG_THREADS) ipiv[k0 + k] = k0 + k;
        if (tid == 0) list[0] = 0;
        return;
    }
    if (tid == 0) moved = 0;
    double a[RPT][NB];
    int pos[RPT + RL];
#pragma unroll
    for (int q = 0; q < RPT + RL; ++q) {
        const int r = tid + REG_THREADS * q;
        pos[q] = r < m ? r : m;   // m marks an empty slot
        phys[r] = r;
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        const double *col = A + ld * (size_t)(k0 + (c < jb ? c : 0));  // uniform base
#pragma unroll
        for (int q = 0; q < RPT + RL; ++q) {
            const int r = tid + REG_THREADS * q;
            const double v = (r < m && c < jb) ? col[k0 + r] : 0.0;
            if (q < RPT) a[q < RPT ? q : 0][c] = v;
            else Lrows[((q - RPT) * NB + c) * REG_THREADS + tid] = v;
        }
    }
    __syncthreads();
    PanelCtx P{sv, prow, Lrows, si, phys, ipiv, info, tid, tid & 63, tid >> 6, k0, m, jb, false};
    panel_column<RL, 0>(a, pos, P);
    if (P.dead) {
        if (tid == 0) list[0] = 0;
        return;
    }
    __syncthreads();   // final phys[] visible
#pragma unroll
    for (int q = 0; q < RPT + RL; ++q) {
        const int r = tid + REG_THREADS * q;
        if (r < m && phys[r] != r) {
            const int e = atomicAdd(&moved, 1);
            if (e < 2 * NB) {
                list[1 + 2 * e] = k0 + r;
                list[2 + 2 * e] = k0 + phys[r];
            }
        }
    }
    __syncthreads();
    if (tid == 0) list[0] = moved < 2 * NB ? moved : 2 * NB;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < jb) {
            double *col = A + ld * (size_t)(k0 + c);
#pragma unroll
            for (int q = 0; q < RPT + RL; ++q)
                if (pos[q] < m)
                    col[k0 + pos[q]] = q < RPT ? a[q < RPT ? q : 0][c]
                                               : Lrows[((q - RPT) * NB + c) * REG_THREADS + tid];
        }
    }
}

// Panel of the blocked path too tall for the register panel: one lane per
// row (strided), the row's NB panel entries loaded together, scaled and
// updated in one pass; the pivot row is read by broadcast loads.
__global__ __launch_bounds__(LU_THREADS) void lu_panel_tall_kernel(int n, double *A, int *ipiv,
                                                                   int *info, int k0, int jb,
                                                                   int *list)
{
    __shared__ double sv[LU_WAVES];
    __shared__ int si[LU_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const size_t ld = (size_t)n;
    if (*info != 0) {
        for (int k = tid; k < jb; k += LU_THREADS) ipiv[k0 + k] = k0 + k;
        if (tid == 0) list[0] = 0;
        return;
    }
    double *P0 = A + ld * (size_t)k0;  // panel column 0
    for (int c = 0; c < jb; ++c) {
        const int k = k0 + c;
        double *Pc = P0 + ld * (size_t)c;
        double bv = -1.0;
        int bi = k;
        for (int i = k + tid; i < n; i += LU_THREADS) better(bv, bi, fabs(Pc[i]), i);
        for (int off = 32; off > 0; off >>= 1)
            better(bv, bi, __shfl_xor(bv, off), __shfl_xor(bi, off));
        if (lane == 0) {
            sv[w] = bv;
            si[w] = bi;
        }
        __syncthreads();
        bv = sv[0];
        bi = si[0];
        for (int q = 1; q < LU_WAVES; ++q) better(bv, bi, sv[q], si[q]);
        const int p = bi;
        if (tid == 0) ipiv[k] = p;
        if (bv == 0.0) {
            if (tid == 0 && *info == 0) *info = k + 1;
            for (int j = c + 1 + tid; j < jb; j += LU_THREADS) ipiv[k0 + j] = k0 + j;
            if (tid == 0) list[0] = 0;
            return;
        }
        if (p != k && tid < jb) {
            double *col = P0 + ld * (size_t)tid;
            const double t = col[k];
            col[k] = col[p];
            col[p] = t;
        }
        __syncthreads();
        const double inv = 1.0 / Pc[k];
        double u[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) u[j] = (j > c && j < jb) ? P0[(size_t)k + ld * j] : 0.0;
        for (int i = k + 1 + tid; i < n; i += LU_THREADS) {
            double v[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) v[j] = (j > c && j < jb) ? P0[(size_t)i + ld * j] : 0.0;
            const double l = Pc[i] * inv;
            Pc[i] = l;
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (j > c && j < jb) P0[(size_t)i + ld * j] = v[j] - l * u[j];
        }
        __syncthreads();
    }
    if (tid == 0) {   // compose the panel's swaps into its row-move list
        int rows[2 * NB], src[2 * NB];
        int cnt = 0;
        for (int k = k0; k < k0 + jb; ++k) {
            const int p = ipiv[k];
            if (p == k) continue;
            int ik = -1, ip = -1;
            for (int e = 0; e < cnt; ++e) {
                if (rows[e] == k) ik = e;
                if (rows[e] == p) ip = e;
            }
            if (ik < 0) {
                ik = cnt++;
                rows[ik] = k;
                src[ik] = k;
            }
            if (ip < 0) {
                ip = cnt++;
                rows[ip] = p;
                src[ip] = p;
            }
            const int t = src[ik];
            src[ik] = src[ip];
            src[ip] = t;
        }
        for (int e = 0; e < cnt; ++e) {
            list[1 + 2 * e] = rows[e];
            list[2 + 2 * e] = src[e];
        }
        list[0] = cnt;
    }
}

// Applies row-move lists, in panel order, to one column per 64-lane block:
// the moves of every panel right of the column's own (its swaps reach the
// columns left of it only at the end, as in a deferred DLASWP).
__device__ __forceinline__ void apply_moves(int n, double *col, const int *L)
{
    const int na = min(L[0], 2 * NB);
    const int lane = threadIdx.x;
    int dst = -1;
    double v = 0.0;
    if (lane < na) {
        dst = L[1 + 2 * lane];
        const int src = L[2 + 2 * lane];
        if (dst >= 0 && dst < n && src >= 0 && src < n) v = col[src];
        else dst = -1;
    }
    __syncthreads();
    if (dst >= 0) col[dst] = v;
    __syncthreads();
}

__global__ __launch_bounds__(64) void lu_swap_left_kernel(int n, double *A, const int *lists,
                                                          int npanels)
{
    const int c = blockIdx.x;
    double *col = A + (size_t)n * c;
    for (int p = c / NB + 1; p < npanels; ++p) apply_moves(n, col, lists + (size_t)p * LSTRIDE);
}

__global__ void zero_info_kernel(int *info) { *info = 0; }

constexpr int GT = 64;       // output tile of the trailing update
constexpr int AS_LD = NB + 1;
constexpr int BS_LD = GT + 1;

// U12 <- L11^-1 U12 for the columns [k+jb, n): rows k..k+jb of those columns
// hold A12 on entry; L11 is the unit lower triangle of the panel.  Each
// group of NB lanes (one per row) solves one column by forward
// substitution, the solved entry broadcast by a lane shuffle.
constexpr int TRSM_COLS = 256 / NB;  // columns per 256-thread block
__global__ __launch_bounds__(256) void lu_trsm_kernel(int n, double *A, int k, int jb,
                                                     const int *list)
{
    __shared__ double L[NB][NB + 1];
    __shared__ int mv[LSTRIDE];
    const size_t ld = (size_t)n;
    const int tid = threadIdx.x;
    // this panel's row moves, applied to the block's columns
    if (tid == 0) mv[0] = min(list[0], 2 * NB);
    __syncthreads();
    const int na = mv[0];
    if (tid < 2 * na) mv[1 + tid] = list[1 + tid];
    __syncthreads();
    {
        const int g = tid / NB, e0 = tid % NB;   // NB lanes per column
        const int c = k + jb + blockIdx.x * TRSM_COLS + g;
        double v0 = 0.0, v1 = 0.0;
        int d0 = -1, d1 = -1;
        if (c < n) {
            double *col = A + ld * c;
            if (e0 < na) {
                d0 = mv[1 + 2 * e0];
                const int s0 = mv[2 + 2 * e0];
                if (d0 >= 0 && d0 < n && s0 >= 0 && s0 < n) v0 = col[s0];
                else d0 = -1;
            }
            if (e0 + NB < na) {
                d1 = mv[1 + 2 * (e0 + NB)];
                const int s1 = mv[2 + 2 * (e0 + NB)];
                if (d1 >= 0 && d1 < n && s1 >= 0 && s1 < n) v1 = col[s1];
                else d1 = -1;
            }
        }
        __syncthreads();
        if (c < n) {
            double *col = A + ld * c;
            if (d0 >= 0) col[d0] = v0;
            if (d1 >= 0) col[d1] = v1;
        }
        __syncthreads();
    }
    for (int e = tid; e < NB * NB; e += 256) {
        const int i = e % NB, m = e / NB;
        L[i][m] = (i < jb && m < jb) ? A[(size_t)(k + i) + ld * (k + m)] : 0.0;
    }
    __syncthreads();
    const int row = tid % NB;
    const int c = k + jb + blockIdx.x * TRSM_COLS + tid / NB;
    const bool live = c < n && row < jb;
    double x = live ? A[(size_t)(k + row) + ld * c] : 0.0;
    const int base = (tid & 63) & ~(NB - 1);  // first lane of this group
    for (int i = 0; i < jb - 1; ++i) {
        const double xi = __shfl(x, base + i);
        if (row > i) x -= L[row][i] * xi;
    }
    if (live) A[(size_t)(k + row) + ld * c] = x;
}

typedef double dx4 __attribute__((ext_vector_type(4)));

// A22 -= L21 U12 for one 64x64 tile of the trailing matrix.
__global__ __launch_bounds__(256) void lu_gemm_mfma_kernel(int n, double *A, int k, int jb)
{
    __shared__ double As[GT * AS_LD];   // L21 tile: As[r*AS_LD + kk]
    __shared__ double Bs[NB * BS_LD];   // U12 tile: Bs[kk*BS_LD + c]
    const size_t ld = (size_t)n;
    const int tid = threadIdx.x;
    const int t0 = k + jb;
    const int R0 = t0 + blockIdx.y * GT, C0 = t0 + blockIdx.x * GT;
    for (int e = tid; e < GT * NB; e += 256) {
        const int r = e % GT, kk = e / GT;
        const int row = R0 + r;
        As[r * AS_LD + kk] = (row < n && kk < jb) ? A[(size_t)row + ld * (k + kk)] : 0.0;
    }
    for (int e = tid; e < GT * NB; e += 256) {
        const int kk = e % NB, c = e / NB;
        const int col = C0 + c;
        Bs[kk * BS_LD + c] = (col < n && kk < jb) ? A[(size_t)(k + kk) + ld * col] : 0.0;
    }
    __syncthreads();
    const int lane = tid & 63, w = tid >> 6;
    const int rw = (w & 1) * 32, cw = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    dx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = dx4{0.0, 0.0, 0.0, 0.0};
    const int kend = (jb + 3) & ~3;
    for (int kk = 0; kk < kend; kk += 4) {
        const double a0 = As[(rw + li) * AS_LD + kk + lk];
        const double a1 = As[(rw + 16 + li) * AS_LD + kk + lk];
        const double b0 = Bs[(kk + lk) * BS_LD + cw + li];
        const double b1 = Bs[(kk + lk) * BS_LD + cw + 16 + li];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = C0 + cw + 16 * b + li;
            if (col >= n) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = R0 + rw + 16 * a + 4 * i + lk;
                if (row < n) A[(size_t)row + ld * col] -= acc[a][b][i];
            }
        }
}

}  // namespace

namespace {

int lu_launch(vlm_ctx *ctx, int n, double *A, int *ipiv, int *info_dev)
{
    const int npanels = (n + NB - 1) / NB;
    if (n > LU_SMALL) VLM_HIP(ctx->lu_work.reserve((size_t)npanels * LSTRIDE));
    int *lists = ctx->lu_work.p;
    hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, ctx->stream, info_dev);
    if (n <= LU_SMALL) {
        hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(LU_THREADS), 0, ctx->stream, n, A,
                           ipiv, info_dev, 0, n, 0, n);
        VLM_HIP(hipGetLastError());
        return VLM_OK;
    }
    for (int k = 0; k < n; k += NB) {
        const int jb = n - k < NB ? n - k : NB;
        int *list = lists + (size_t)(k / NB) * LSTRIDE;
        if (n - k <= RPT * REG_THREADS)
            hipLaunchKernelGGL(lu_panel_reg_kernel<0>, dim3(1), dim3(REG_THREADS), 0, ctx->stream,
                               n, A, ipiv, info_dev, k, jb, list);
        else if (n - k <= (RPT + 2) * REG_THREADS)
            hipLaunchKernelGGL(lu_panel_reg_kernel<2>, dim3(1), dim3(REG_THREADS), 0, ctx->stream,
                               n, A, ipiv, info_dev, k, jb, list);
        else
            hipLaunchKernelGGL(lu_panel_tall_kernel, dim3(1), dim3(LU_THREADS), 0, ctx->stream, n,
                               A, ipiv, info_dev, k, jb, list);
        const int rest = n - k - jb;
        if (rest > 0) {
            hipLaunchKernelGGL(lu_trsm_kernel, dim3((rest + TRSM_COLS - 1) / TRSM_COLS),
                               dim3(256), 0, ctx->stream, n, A, k, jb, list);
            const int tiles = (rest + GT - 1) / GT;
            hipLaunchKernelGGL(lu_gemm_mfma_kernel, dim3(tiles, tiles), dim3(256), 0,
                               ctx->stream, n, A, k, jb);
        }
        VLM_HIP(hipGetLastError());
    }
    if (npanels > 1) {
        hipLaunchKernelGGL(lu_swap_left_kernel, dim3((npanels - 1) * NB), dim3(64),
                           0, ctx->stream, n, A, lists, npanels);
        VLM_HIP(hipGetLastError());
    }
    return VLM_OK;
}

}  // namespace

int lu_factor_device(vlm_ctx *ctx, int n, double *A, int *ipiv, int *info_dev)
{
    if (!ctx->profile) return lu_launch(ctx, n, A, ipiv, info_dev);
    Timer tm;
    VLM_HIP(hipEventCreate(&tm.a));
    VLM_HIP(hipEventCreate(&tm.b));
    tm.pairs = 2.0 * (double)n * n * n / 3.0;
    VLM_HIP(hipEventRecord(tm.a, ctx->stream));
    const int rc = lu_launch(ctx, n, A, ipiv, info_dev);
    VLM_HIP(hipEventRecord(tm.b, ctx->stream));
    ctx->lu_timers.push_back(tm);
    return rc;
}

}  // namespace vlm
