// Dense LU factorisation with partial pivoting (FP64) for the AIC solve.
//
// Replaces DGETRF (inv2, libMath.f90:48-83; the native Doolittle routine of
// libMath_noBLAS.f90 in no-BLAS builds).  Column-major, 0-based pivots,
// info = first zero pivot (1-based) or 0.  Pivot choice follows IDAMAX:
// the first row of largest magnitude.
//
// n <= LU_SMALL (the AIC of one rotor): one workgroup, unblocked.
// Larger n: right-looking blocked LU.  Panels are 32 columns wide where the
// panel fits on chip (m <= 2048 rows), 16 columns above that.
//   1. panel: one workgroup factorises A[k:n, k:k+jb] column by column,
//      the panel in registers; rows never move during the panel, each lane
//      tracks the logical position of its rows, and one barrier per column
//      publishes every wave's pivot candidate with its row.  16-column
//      panels retire each pivot row from the registers, 32-column panels
//      keep it and mask it (the faster of the two at each width, measured);
//      panels too tall for registers run from global memory.  The panel's
//      row swaps reach the other columns through one composed row-move
//      list;
//   2. U12 = L11^-1 A12 (unit lower triangular solve, jb lanes per column,
//      after the panel's row moves);
//   3. A22 -= L21 U12 on the matrix cores, K = jb: 64x64 output tiles, four
//      waves of 2x2 v_mfma_f64_16x16x4f64 blocks, operands staged in LDS.
//      Lane layout of the 16x16x4 FP64 MFMA (checked on gfx950): lane l
//      supplies A[l%16][l/16] and B[l/16][l%16]; result register i of lane
//      l is D[4i + l/16][l%16].  The tile is computed transposed, so that
//      each result register holds 16 consecutive rows of one column: the
//      accumulators load A22 and store it back in 128-byte runs.
#include <climits>
#include <vector>

#include "vlm_internal.h"

namespace vlm {
namespace {

constexpr int LU_THREADS = 1024;
constexpr int LU_WAVES = LU_THREADS / 64;
constexpr int LU_SMALL = 256;  // one-workgroup path up to this order
constexpr int NB = 16;         // panel width of tall panels
constexpr int NBMAX = 32;      // panel width where the panel fits on chip
constexpr int PT = 512;        // workgroup of the on-chip panel
constexpr int PW = PT / 64;
constexpr int WIDE_ROWS = 2 * PT;  // 32-column panels up to this many rows (2 per lane)
constexpr int MID_ROWS = 5 * PT;   // 16-column panels held in registers up to this many
// Row-move list of one panel: [count, dst0, src0, dst1, src1, ..., k0, jb],
// at most 2 jb moved rows (absolute row indices; the panel's swaps
// composed), then the panel's first column and width.
constexpr int LPAIRS = 2 * NBMAX;
constexpr int LSTRIDE = 3 + 2 * LPAIRS;
constexpr int LK0 = LSTRIDE - 2;
constexpr int LJB = LSTRIDE - 1;

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


__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// max over the lanes that DPP control CTRL pairs (unsigned 32-bit)
template <int CTRL>
__device__ __forceinline__ unsigned dpp_umax(unsigned v)
{
    return max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ int dpp_min(int i)
{
    return min(i, __builtin_amdgcn_update_dpp(INT_MAX, i, CTRL, 0xF, 0xF, false));
}

// max(v, v of the lane that DPP control CTRL selects)
template <int CTRL>
__device__ __forceinline__ double dpp_max(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return fmax(v, __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false),
                                    __builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false)));
}

// Maximum over the wave (every lane active), returned to every lane.
// Within each row of 16 lanes: xor 1, xor 2, half-row mirror, row mirror;
// then the four rows through readlane.
__device__ __forceinline__ double wave_max(double v)
{
    v = dpp_max<0xB1>(v);    // quad_perm [1,0,3,2]
    v = dpp_max<0x4E>(v);    // quad_perm [2,3,0,1]
    v = dpp_max<0x141>(v);   // row_half_mirror
    v = dpp_max<0x140>(v);   // row_mirror
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)),
                fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Maximum / minimum over the wave (every lane active), uniform result.
// Within each row of 16 lanes: xor 1, xor 2, half-row mirror, row mirror;
// then the four rows through readlane.
__device__ __forceinline__ unsigned wave_umax(unsigned v)
{
    v = dpp_umax<0xB1>(v);    // quad_perm [1,0,3,2]
    v = dpp_umax<0x4E>(v);    // quad_perm [2,3,0,1]
    v = dpp_umax<0x141>(v);   // row_half_mirror
    v = dpp_umax<0x140>(v);   // row_mirror
    return max(max((unsigned)__builtin_amdgcn_readlane((int)v, 0),
                   (unsigned)__builtin_amdgcn_readlane((int)v, 16)),
               max((unsigned)__builtin_amdgcn_readlane((int)v, 32),
                   (unsigned)__builtin_amdgcn_readlane((int)v, 48)));
}

__device__ __forceinline__ int wave_min(int i)
{
    i = dpp_min<0xB1>(i);
    i = dpp_min<0x4E>(i);
    i = dpp_min<0x141>(i);
    i = dpp_min<0x140>(i);
    return min(min(__builtin_amdgcn_readlane(i, 0), __builtin_amdgcn_readlane(i, 16)),
               min(__builtin_amdgcn_readlane(i, 32), __builtin_amdgcn_readlane(i, 48)));
}

// A wave's pivot candidate as the other waves read it: the bit pattern of
// its |value| (lo, hi: for non-negative doubles the unsigned 64-bit order
// is the numeric order), its logical row times 8 plus the wave, and the
// owning thread times 8 plus the register slot of the row.
struct PivRec {
    unsigned lo, hi;
    int key, owner;
};

// Keep the better of this lane's record and the one DPP control CTRL
// pairs it with: larger |value|, then the lower logical row (IDAMAX).
template <int CTRL>
__device__ __forceinline__ void rec_best(PivRec &r)
{
    const unsigned lo = __builtin_amdgcn_update_dpp((int)r.lo, (int)r.lo, CTRL, 0xF, 0xF, false);
    const unsigned hi = __builtin_amdgcn_update_dpp((int)r.hi, (int)r.hi, CTRL, 0xF, 0xF, false);
    const int key = __builtin_amdgcn_update_dpp(r.key, r.key, CTRL, 0xF, 0xF, false);
    const int own = __builtin_amdgcn_update_dpp(r.owner, r.owner, CTRL, 0xF, 0xF, false);
    const unsigned long long a = ((unsigned long long)r.hi << 32) | r.lo;
    const unsigned long long b = ((unsigned long long)hi << 32) | lo;
    const bool t = (b > a) | ((b == a) & (key < r.key));
    r.lo = t ? lo : r.lo;
    r.hi = t ? hi : r.hi;
    r.key = t ? key : r.key;
    r.owner = t ? own : r.owner;
}

__device__ __forceinline__ bool slot_is(int q, int s)
{
    asm("" : "+s"(q));
    return q == s;
}

constexpr int NOROW = 0x0FFFFFFF;   // logical position of a retired or empty slot

// Panel held in registers, retiring design (the 16-column panels, up to
// 2560 rows): lane t owns the physical panel rows t + PT*q,
// RPT of them (columns [k0, k0+JB)).  Rows never move during the panel:
// pos[q] is the logical row (after the swaps so far) of each owned row.
// Per column, each wave finds its candidate (a DPP maximum of the high
// words of |value|, the low words and the row index only on ties), its
// owner lane publishes the candidate row and record, and one barrier later
// every wave takes the best of the PW records by a DPP butterfly.  The
// owner of the winning row then retires it: the row goes to LDS (Ur, at
// its final position C) and its registers are zeroed, so every later
// column treats all slots alike (a zero row scales to l = 0 and is never
// updated) and only the row at logical C, which takes the pivot's old
// position, changes its pos.  The slots of the candidate rows alternate
// by column parity, so one barrier per column suffices.  The column loop
// is unrolled by template recursion so that every register index is a
// constant; the arithmetic per element (the reciprocal of the pivot, the
// scaling, the fused rank-1 updates in column order) is DGETF2's, so both
// designs give the same factors bit for bit.  Against the masking design
// below this needs fewer registers (no per-row bookkeeping at 5 rows per
// lane: no spills), but the owner's copy and zeroing of the row grows
// with the panel width, and at 32 columns it costs more than it saves.
template <int JB>
struct PanelCtx {
    double *cand;   // [2][PW][JB] candidate rows
    PivRec *rec;    // [2][PW]
    double *Ur;     // [JB][JB] the row retired at column C
    int *Uphys;     // [JB] its physical row in the panel
    int *ipiv, *info;
    int tid, lane, w, k0;
};

template <int JB, int RPT, int C>
__device__ __forceinline__ void panel_column(double (&a)[RPT][JB], int (&pos)[RPT],
                                             PanelCtx<JB> &P)
{
    constexpr int par = C & 1;
    // this lane's candidate: largest |a| over its rows, the lowest logical
    // row on ties (retired and empty slots hold 0 at logical NOROW)
    double bv = -1.0;
    int bi = NOROW, bq = 0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const double x = fabs(a[q][C]);
        const int p = pos[q];
        const bool t = (x > bv) | ((x == bv) & (p < bi));
        bv = t ? x : bv;
        bi = t ? p : bi;
        bq = t ? q : bq;
    }
    // the wave's candidate lane: the high words decide unless tied
    const unsigned hi = (unsigned)__double2hiint(bv);
    const unsigned M = wave_umax(hi);
    unsigned long long hit = __ballot(hi == M);
    if (__popcll(hit) != 1) {
        const unsigned lo = hi == M ? (unsigned)__double2loint(bv) : 0u;
        const unsigned Ml = wave_umax(lo);
        const bool top = (hi == M) & (lo == Ml);
        hit = __ballot(top);
        if (__popcll(hit) != 1) {
            const int r = wave_min(top ? bi : INT_MAX);
            hit = __ballot(top & (bi == r));
        }
    }
    const int wl = __ffsll((long long)hit) - 1;
    const int uq = __builtin_amdgcn_readlane(bq, wl);
    if (P.lane == wl) {
        double *crow = P.cand + (par * PW + P.w) * JB;
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (slot_is(q, uq)) {
#pragma unroll
                for (int j = C; j < JB; ++j) crow[j] = a[q][j];
            }
        P.rec[par * PW + P.w] = PivRec{(unsigned)__double2loint(bv), hi, bi * 8 + P.w,
                                       P.tid * 8 + uq};
    }
    __syncthreads();
    // the block's pivot: butterfly over the PW records in each group of
    // 8 lanes (half-row mirror, then xor 2 and xor 1)
    static_assert(PW == 8, "butterfly over 8 waves");
    PivRec r = P.rec[par * PW + (P.lane & 7)];
    rec_best<0x141>(r);
    rec_best<0x4E>(r);
    rec_best<0xB1>(r);
    const int gkey = __builtin_amdgcn_readfirstlane(r.key);
    const int gown = __builtin_amdgcn_readfirstlane(r.owner);
    const bool zero = (__builtin_amdgcn_readfirstlane((int)r.hi) |
                       __builtin_amdgcn_readfirstlane((int)r.lo)) == 0;
    const int gi = gkey >> 3, gw = gkey & 7;
    const double *prow = P.cand + (par * PW + gw) * JB;
    if (P.tid == 0) {
        P.ipiv[P.k0 + C] = P.k0 + gi;
        // a zero column (every candidate 0: gi = C, no swap) is recorded
        // and left as it is, as DGETF2 does; the factorisation reports it
        if (zero && *P.info == 0) *P.info = P.k0 + C + 1;
    }
    const double inv = zero ? 0.0 : 1.0 / prow[C];
    // the row at logical C takes the pivot's logical position
#pragma unroll
    for (int q = 0; q < RPT; ++q) pos[q] = pos[q] == C ? gi : pos[q];
    // the owner retires the pivot row
    const int otid = gown >> 3, oq = gown & 7;
    if (P.tid == otid) {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (slot_is(q, oq)) {
#pragma unroll
                for (int j = 0; j < JB; ++j) P.Ur[C * JB + j] = a[q][j];
#pragma unroll
                for (int j = C; j < JB; ++j) a[q][j] = 0.0;
                pos[q] = NOROW;
            }
        P.Uphys[C] = otid + PT * oq;
    }
    // scale and rank-1 update, CH pivot-row entries at a time (the
    // scheduling barrier keeps later entries' loads from piling up in
    // registers)
    double l[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        l[q] = a[q][C] * inv;
        a[q][C] = l[q];
    }
    constexpr int CH = RPT >= 3 ? 4 : 16;
#pragma unroll
    for (int j0 = C + 1; j0 < JB; j0 += CH) {
        double u[CH];
#pragma unroll
        for (int t = 0; t < CH; ++t) u[t] = j0 + t < JB ? prow[j0 + t < JB ? j0 + t : 0] : 0.0;
#pragma unroll
        for (int t = 0; t < CH; ++t) {
            if (j0 + t < JB) {
                const int j = j0 + t < JB ? j0 + t : 0;
#pragma unroll
                for (int q = 0; q < RPT; ++q) a[q][j] -= l[q] * u[t];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (C + 1 < JB) panel_column<JB, RPT, C + 1>(a, pos, P);
}

// Exactly JB columns of at most PT*RPT rows (the host gives ragged last
// panels to the tall kernel).
template <int JB, int RPT>
__global__ __launch_bounds__(PT) void lu_panel_fast_kernel(int n, double *A, int *ipiv, int *info,
                                                           int k0, int *list)
{
    __shared__ double cand[2 * PW * JB];
    __shared__ PivRec rec[2 * PW];
    __shared__ double Ur[JB * JB];
    __shared__ int Uphys[JB];
    __shared__ int moved;
    const int tid = threadIdx.x;
    const size_t ld = (size_t)n;
    const int m = n - k0;
    if (tid == 0) {
        list[LK0] = k0;
        list[LJB] = JB;
    }
    if (*info != 0) {
        for (int k = tid; k < JB; k += PT) ipiv[k0 + k] = k0 + k;
        if (tid == 0) list[0] = 0;
        return;
    }
    if (tid == 0) moved = 0;
    double a[RPT][JB];
    int pos[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = tid + PT * q;
        pos[q] = r < m ? r : NOROW;
    }
    const double *col = A + ld * (size_t)k0 + k0;
#pragma unroll
    for (int c = 0; c < JB; ++c) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            // unconditional load from a clamped row, then the select: no
            // branch, so every load of the panel is in flight at once
            const int r = tid + PT * q;
            const double x = col[ld * c + (r < m ? r : m - 1)];
            a[q][c] = r < m ? x : 0.0;
        }
        __builtin_amdgcn_sched_barrier(0);   // one column's addresses live at a time
    }
    __syncthreads();   // moved = 0 visible
    PanelCtx<JB> P{cand, rec, Ur, Uphys, ipiv, info, tid, tid & 63, tid >> 6, k0};
    panel_column<JB, RPT, 0>(a, pos, P);
    // rows still in registers go to their logical rows, the retired ones
    // (from LDS) to rows 0..JB-1 of the panel; every row that changed
    // place enters the row-move list
    double *out = A + ld * (size_t)k0 + k0;
#pragma unroll
    for (int c = 0; c < JB; ++c) {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (pos[q] != NOROW) out[ld * c + pos[q]] = a[q][c];
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = tid + PT * q;
        if (pos[q] != NOROW && pos[q] != r) {
            const int e = atomicAdd(&moved, 1);
            if (e < LPAIRS) {
                list[1 + 2 * e] = k0 + pos[q];   // destination (logical) row
                list[2 + 2 * e] = k0 + r;        // source (physical) row
            }
        }
    }
    __syncthreads();   // Ur, Uphys complete
    if (tid < JB && Uphys[tid] != tid) {
        const int e = atomicAdd(&moved, 1);
        if (e < LPAIRS) {
            list[1 + 2 * e] = k0 + tid;
            list[2 + 2 * e] = k0 + Uphys[tid];
        }
    }
    for (int e = tid; e < JB * JB; e += PT) {
        const int rr = e % JB, c = e / JB;
        out[ld * c + rr] = Ur[rr * JB + c];
    }
    __syncthreads();
    if (tid == 0) list[0] = moved < LPAIRS ? moved : LPAIRS;
}


// Panel held in registers, masking design (the 32-column panels, up to
// 1024 rows): lane t owns the physical panel rows t + PT*q, RPT of
// them in registers and RL more in LDS (columns [k0, k0+jb), jb <= JB).
// A pivot row stays in its lane and is masked out of later columns by its
// logical position.
// Rows never move during the panel: pos[q] is the logical row (after the
// swaps so far) of each owned row, so a row swap is two index updates.
// Per column, each wave finds its candidate (wave argmax), the owning lane
// copies the candidate row into the wave's LDS slot, and one barrier later
// every wave picks the winning record (a DPP butterfly over the waves'
// records) and reads the pivot row from its slot.
// The slots alternate between two buffers by column parity, so a fast wave
// writing the next column's slot never overwrites one still being read.
// The column loop is unrolled by template recursion so that every register
// index is a constant.
template <int JB>
struct WideCtx {
    double *cand;   // [2][PW][JB]
    PivRec *rec;    // [2][PW] candidate records
    double *cinv;   // [2][PW] 1 / candidate value
    double *L;      // LDS rows: L[(q*JB + j)*PT + tid]
    int *ipiv, *info, *list;
    int tid, lane, w, k0, m;
};

template <int JB, int RPT, int RL, int C>
__device__ __forceinline__ void wide_column(double (&a)[RPT][JB], int (&pos)[RPT + RL],
                                            WideCtx<JB> &P)
{
    constexpr int par = C & 1;
    double *L = P.L + P.tid;
    // this lane's candidate: largest |a| among its rows at logical >= C,
    // the lowest logical row on ties (branch-free)
    double bv = -1.0;
    int bi = INT_MAX;
#pragma unroll
    for (int q = 0; q < RPT + RL; ++q) {
        const int p = pos[q];
        const double x = fabs(q < RPT ? a[q < RPT ? q : 0][C] : L[((q - RPT) * JB + C) * PT]);
        const bool t = (p >= C) & (p < P.m) & ((x > bv) | ((x == bv) & (p < bi)));
        bv = t ? x : bv;
        bi = t ? p : bi;
    }
    // the wave's candidate: its maximum, and the lowest row holding it
    const double M = wave_max(bv);
    const unsigned long long hit = __ballot(bv == M);
    const int wi = __popcll(hit) == 1 ? __builtin_amdgcn_readlane(bi, __ffsll((long long)hit) - 1)
                                      : wave_min(bv == M ? bi : INT_MAX);
    // its owner publishes the row (columns C..JB-1) and the reciprocal
    if (bi == wi) {
        double *crow = P.cand + (par * PW + P.w) * JB;
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (pos[q] == wi) {
                P.cinv[par * PW + P.w] = 1.0 / a[q][C];
#pragma unroll
                for (int j = C; j < JB; ++j) crow[j] = a[q][j];
            }
#pragma unroll
        for (int q = 0; q < RL; ++q)
            if (pos[RPT + q] == wi) {
                P.cinv[par * PW + P.w] = 1.0 / L[(q * JB + C) * PT];
#pragma unroll
                for (int j = C; j < JB; ++j) crow[j] = L[(q * JB + j) * PT];
            }
    }
    if (P.lane == 0) {
        // a wave without candidate rows (M = -1) loses to every other
        const bool none = M < 0.0;
        P.rec[par * PW + P.w] = PivRec{none ? 0u : (unsigned)__double2loint(M),
                                       none ? 0u : (unsigned)__double2hiint(M),
                                       none ? INT_MAX : wi * 8 + P.w, 0};
    }
    __syncthreads();
    // the block's pivot: largest wave maximum, lowest row among the waves
    // holding it, by a butterfly over the PW records in each group of 8
    // lanes (half-row mirror, then xor 2 and xor 1)
    static_assert(PW == 8, "butterfly over 8 waves");
    PivRec r = P.rec[par * PW + (P.lane & 7)];
    rec_best<0x141>(r);
    rec_best<0x4E>(r);
    rec_best<0xB1>(r);
    const double gv = __hiloint2double(__builtin_amdgcn_readfirstlane((int)r.hi),
                                       __builtin_amdgcn_readfirstlane((int)r.lo));
    const int gkey = __builtin_amdgcn_readfirstlane(r.key);
    const int gw = gkey & 7;
    const int lp = gkey >> 3;   // logical row of the pivot
    if (P.tid == 0) {
        P.ipiv[P.k0 + C] = P.k0 + lp;
        // a zero column (every candidate 0: lp = C, no swap) is recorded
        // and left as it is, as DGETF2 does; the factorisation reports it
        if (gv == 0.0 && *P.info == 0) *P.info = P.k0 + C + 1;
    }
    const double *prow = P.cand + (par * PW + gw) * JB;
    const double inv = gv == 0.0 ? 0.0 : P.cinv[par * PW + gw];
#pragma unroll
    for (int q = 0; q < RPT + RL; ++q) {
        if (pos[q] == lp) pos[q] = C;
        else if (pos[q] == C) pos[q] = lp;
    }
    double l[RPT + RL];
#pragma unroll
    for (int q = 0; q < RPT + RL; ++q) {
        const bool upd = pos[q] > C && pos[q] < P.m;
        if (q < RPT) {
            const double x = a[q < RPT ? q : 0][C];
            l[q] = upd ? x * inv : 0.0;
            a[q < RPT ? q : 0][C] = upd ? l[q] : x;
        } else {
            double *e = L + ((q - RPT) * JB + C) * PT;
            const double x = *e;
            l[q] = upd ? x * inv : 0.0;
            if (upd) *e = l[q];
        }
    }
    // rank-1 update, CH pivot-row entries at a time (the scheduling barrier
    // keeps later entries' loads from piling up in registers when the
    // panel fills them)
    constexpr int CH = RPT >= 3 ? 4 : 16;
#pragma unroll
    for (int j0 = C + 1; j0 < JB; j0 += CH) {
        double u[CH];
#pragma unroll
        for (int t = 0; t < CH; ++t) u[t] = j0 + t < JB ? prow[j0 + t < JB ? j0 + t : 0] : 0.0;
#pragma unroll
        for (int t = 0; t < CH; ++t) {
            if (j0 + t < JB) {
                const int j = j0 + t < JB ? j0 + t : 0;
#pragma unroll
                for (int q = 0; q < RPT; ++q) a[q][j] -= l[q] * u[t];
#pragma unroll
                for (int q = 0; q < RL; ++q)
                    if (l[RPT + q] != 0.0) L[(q * JB + j) * PT] -= l[RPT + q] * u[t];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (C + 1 < JB) wide_column<JB, RPT, RL, C + 1>(a, pos, P);
}

// Exactly JB columns (the host gives ragged last panels to the tall kernel).
template <int JB, int RPT, int RL>
__global__ __launch_bounds__(PT) void lu_panel_wide_kernel(int n, double *A, int *ipiv, int *info,
                                                           int k0, int *list)
{
    constexpr int S = RPT + RL;
    constexpr int EMPTY = 1 << 30;   // logical position of a slot past the panel
    __shared__ double cand[2 * PW * JB];
    __shared__ PivRec rec[2 * PW];
    __shared__ double cinv[2 * PW];
    __shared__ double Lrows[RL > 0 ? RL * JB * PT : 1];
    __shared__ int moved;
    const int tid = threadIdx.x;
    const size_t ld = (size_t)n;
    const int m = n - k0;
    if (tid == 0) {
        list[LK0] = k0;
        list[LJB] = JB;
    }
    if (*info != 0) {
        for (int k = tid; k < JB; k += PT) ipiv[k0 + k] = k0 + k;
        if (tid == 0) list[0] = 0;
        return;
    }
    if (tid == 0) moved = 0;
    double a[RPT][JB];
    int pos[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int r = tid + PT * q;
        pos[q] = r < m ? r : EMPTY;
    }
    const double *col = A + ld * (size_t)k0 + k0;
#pragma unroll
    for (int c = 0; c < JB; ++c) {
#pragma unroll
        for (int q = 0; q < S; ++q) {
            // unconditional load from a clamped row, then the select: no
            // branch, so every load of the panel is in flight at once
            const int r = tid + PT * q;
            const double x = col[ld * c + (r < m ? r : m - 1)];
            const double v = r < m ? x : 0.0;
            if (q < RPT) a[q < RPT ? q : 0][c] = v;
            else Lrows[((q - RPT) * JB + c) * PT + tid] = v;
        }
        __builtin_amdgcn_sched_barrier(0);   // one column's addresses live at a time
    }
    __syncthreads();   // moved = 0 visible
    WideCtx<JB> P{cand, rec, cinv, Lrows, ipiv, info, list, tid, tid & 63, tid >> 6, k0, m};
    wide_column<JB, RPT, RL, 0>(a, pos, P);
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const int r = tid + PT * q;
        if (r < m && pos[q] != r) {
            const int e = atomicAdd(&moved, 1);
            if (e < LPAIRS) {
                list[1 + 2 * e] = k0 + pos[q];   // destination (logical) row
                list[2 + 2 * e] = k0 + r;        // source (physical) row
            }
        }
    }
    __syncthreads();
    if (tid == 0) list[0] = moved < LPAIRS ? moved : LPAIRS;
    double *out = A + ld * (size_t)k0 + k0;
#pragma unroll
    for (int c = 0; c < JB; ++c) {
#pragma unroll
        for (int q = 0; q < S; ++q)
            if (pos[q] < m)
                out[ld * c + pos[q]] = q < RPT ? a[q < RPT ? q : 0][c]
                                               : Lrows[((q - RPT) * JB + c) * PT + tid];
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Panel of the blocked path too tall to be held on chip (jb <= NB): one
// lane per row (strided), the row's NB panel entries loaded together,
// scaled and updated in one pass; the pivot row is read by broadcast loads.
__global__ __launch_bounds__(LU_THREADS) void lu_panel_tall_kernel(int n, double *A, int *ipiv,
                                                                   int *info, int k0, int jb,
                                                                   int *list)
{
    __shared__ double sv[LU_WAVES];
    __shared__ int si[LU_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const size_t ld = (size_t)n;
    if (tid == 0) {
        list[LK0] = k0;
        list[LJB] = jb;
    }
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

// Applies one row-move list to one column (64-lane block): every moved
// row is read before any is written.
__device__ __forceinline__ void apply_moves(int n, double *col, const int *L)
{
    const int na = min(L[0], LPAIRS);
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

// The deferred DLASWP: column c receives, in panel order, the row moves of
// every panel that starts right of it.
__global__ __launch_bounds__(64) void lu_swap_left_kernel(int n, double *A, const int *lists,
                                                          int npanels)
{
    const int c = blockIdx.x;
    double *col = A + (size_t)n * c;
    for (int p = 0; p < npanels; ++p) {
        const int *L = lists + (size_t)p * LSTRIDE;
        if (L[LK0] > c) apply_moves(n, col, L);
    }
}

__global__ void zero_info_kernel(int *info) { *info = 0; }

// U12 <- L11^-1 U12 for the columns [k+jb, n): rows k..k+jb of those columns
// hold A12 on entry, after this panel's row moves (applied here to the
// block's columns); L11 is the unit lower triangle of the panel.  Each
// group of JB lanes (one per row) solves one column by forward
// substitution, the solved entry broadcast from its lane by readlane.
template <int JB>
__global__ __launch_bounds__(256) void lu_trsm_kernel(int n, double *A, int k, int jb,
                                                     const int *list)
{
    constexpr int CPB = 256 / JB;   // columns per block
    __shared__ double L[JB][JB + 1];
    __shared__ int mv[1 + 2 * LPAIRS];
    const size_t ld = (size_t)n;
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid == 0) mv[0] = min(list[0], 2 * JB);
    __syncthreads();
    const int na = mv[0];
    if (tid < 2 * na) mv[1 + tid] = list[1 + tid];
    for (int e = tid; e < JB * JB; e += 256) {
        const int i = e % JB, c = e / JB;
        L[i][c] = (i < jb && c < jb) ? A[(size_t)(k + i) + ld * (k + c)] : 0.0;
    }
    __syncthreads();
    const int row = tid % JB;
    const int c = k + jb + blockIdx.x * CPB + tid / JB;
    {
        // the moves of this panel, at most 2 per lane of the column's group
        double v0 = 0.0, v1 = 0.0;
        int d0 = -1, d1 = -1;
        if (c < n) {
            const double *col = A + ld * c;
            if (row < na) {
                d0 = mv[1 + 2 * row];
                const int s0 = mv[2 + 2 * row];
                if (d0 >= 0 && d0 < n && s0 >= 0 && s0 < n) v0 = col[s0];
                else d0 = -1;
            }
            if (row + JB < na) {
                d1 = mv[1 + 2 * (row + JB)];
                const int s1 = mv[2 + 2 * (row + JB)];
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
    const bool live = c < n && row < jb;
    double x = live ? A[(size_t)(k + row) + ld * c] : 0.0;
    const int grp = lane / JB;
    for (int i = 0; i < jb - 1; ++i) {
        double xi = 0.0;
#pragma unroll
        for (int g = 0; g < 64 / JB; ++g) {
            const double t = readlane_f64(x, g * JB + i);
            xi = grp == g ? t : xi;
        }
        if (row > i) x -= L[row][i] * xi;
    }
    if (live) A[(size_t)(k + row) + ld * c] = x;
}

typedef double dx4 __attribute__((ext_vector_type(4)));
constexpr int GT = 64;            // output tile of the trailing update
constexpr int GS_LD = GT + 1;

// A22 -= L21 U12 for one 64x64 tile of the trailing matrix, K = jb <= 32.
// Computed transposed, D' = U12^T (-L21)^T + A22^T, so that result register
// i of lane l is A22[row l%16][column 4i + l/16]: the accumulators start
// from A22 and go back to it in runs of 16 consecutive rows.
__global__ __launch_bounds__(256) void lu_gemm_mfma_kernel(int n, double *A, int k, int jb)
{
    __shared__ double Ls[NBMAX * GS_LD];   // -L21^T: Ls[kk*GS_LD + r]
    __shared__ double Us[NBMAX * GS_LD];   // U12:    Us[kk*GS_LD + c]
    const size_t ld = (size_t)n;
    const int tid = threadIdx.x;
    const int t0 = k + jb;
    const int R0 = t0 + blockIdx.y * GT, C0 = t0 + blockIdx.x * GT;
    const int kend = (jb + 3) & ~3;
    for (int e = tid; e < GT * kend; e += 256) {
        const int r = e % GT, kk = e / GT;
        const int row = R0 + r;
        Ls[kk * GS_LD + r] = (row < n && kk < jb) ? -A[(size_t)row + ld * (k + kk)] : 0.0;
    }
    for (int e = tid; e < GT * kend; e += 256) {
        const int kk = e % kend, c = e / kend;
        const int col = C0 + c;
        Us[kk * GS_LD + c] = (col < n && kk < jb) ? A[(size_t)(k + kk) + ld * col] : 0.0;
    }
    const int lane = tid & 63, w = tid >> 6;
    const int rw = (w & 1) * 32, cw = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    dx4 acc[2][2];   // [column block][row block]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = R0 + rw + 16 * b + li;
                const int col = C0 + cw + 16 * a + 4 * i + lk;
                acc[a][b][i] = (row < n && col < n) ? A[(size_t)row + ld * col] : 0.0;
            }
    __syncthreads();
    for (int kk = 0; kk < kend; kk += 4) {
        const double u0 = Us[(kk + lk) * GS_LD + cw + li];
        const double u1 = Us[(kk + lk) * GS_LD + cw + 16 + li];
        const double l0 = Ls[(kk + lk) * GS_LD + rw + li];
        const double l1 = Ls[(kk + lk) * GS_LD + rw + 16 + li];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(u0, l0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(u0, l1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(u1, l0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(u1, l1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = R0 + rw + 16 * b + li;
                const int col = C0 + cw + 16 * a + 4 * i + lk;
                if (row < n && col < n) A[(size_t)row + ld * col] = acc[a][b][i];
            }
}

}  // namespace

namespace {

int lu_launch(vlm_ctx *ctx, int n, double *A, int *ipiv, int *info_dev)
{
    hipLaunchKernelGGL(zero_info_kernel, dim3(1), dim3(1), 0, ctx->stream, info_dev);
    if (n <= LU_SMALL) {
        hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(LU_THREADS), 0, ctx->stream, n, A,
                           ipiv, info_dev, 0, n, 0, n);
        VLM_HIP(hipGetLastError());
        return VLM_OK;
    }
    // panel schedule: 16 columns while the panel is taller than the
    // 32-column on-chip panel, 32 after; a last panel of fewer than 32
    // rows goes 16 columns at a time through the tall kernel
    std::vector<int> k0s;
    for (int k = 0; k < n;) {
        k0s.push_back(k);
        const int m = n - k;
        k += m > WIDE_ROWS ? NB : (m >= NBMAX ? NBMAX : (m < NB ? m : NB));
    }
    const int npanels = (int)k0s.size();
    VLM_HIP(ctx->lu_work.reserve((size_t)npanels * LSTRIDE));
    int *lists = ctx->lu_work.p;
    for (int p = 0; p < npanels; ++p) {
        const int k = k0s[p], m = n - k;
        const int jb = (p + 1 < npanels ? k0s[p + 1] : n) - k;
        int *list = lists + (size_t)p * LSTRIDE;
        const dim3 g1(1), bp(PT);
        if (m > MID_ROWS || jb < NB || (m <= WIDE_ROWS && jb < NBMAX))
            hipLaunchKernelGGL(lu_panel_tall_kernel, g1, dim3(LU_THREADS), 0, ctx->stream, n, A,
                               ipiv, info_dev, k, jb, list);
        else if (m > 4 * PT)
            hipLaunchKernelGGL((lu_panel_fast_kernel<NB, 5>), g1, bp, 0, ctx->stream, n, A,
                               ipiv, info_dev, k, list);
        else if (m > 3 * PT)
            hipLaunchKernelGGL((lu_panel_fast_kernel<NB, 4>), g1, bp, 0, ctx->stream, n, A,
                               ipiv, info_dev, k, list);
        else if (m > WIDE_ROWS)
            hipLaunchKernelGGL((lu_panel_fast_kernel<NB, 3>), g1, bp, 0, ctx->stream, n, A,
                               ipiv, info_dev, k, list);
        else if (m > PT)
            hipLaunchKernelGGL((lu_panel_wide_kernel<NBMAX, 2, 0>), g1, bp, 0, ctx->stream, n,
                               A, ipiv, info_dev, k, list);
        else
            hipLaunchKernelGGL((lu_panel_wide_kernel<NBMAX, 1, 0>), g1, bp, 0, ctx->stream, n,
                               A, ipiv, info_dev, k, list);
        const int rest = n - k - jb;
        if (rest > 0) {
            if (jb <= NB)
                hipLaunchKernelGGL(lu_trsm_kernel<NB>, dim3((rest + 256 / NB - 1) / (256 / NB)),
                                   dim3(256), 0, ctx->stream, n, A, k, jb, list);
            else
                hipLaunchKernelGGL(lu_trsm_kernel<NBMAX>,
                                   dim3((rest + 256 / NBMAX - 1) / (256 / NBMAX)), dim3(256), 0,
                                   ctx->stream, n, A, k, jb, list);
            const int tiles = (rest + GT - 1) / GT;
            hipLaunchKernelGGL(lu_gemm_mfma_kernel, dim3(tiles, tiles), dim3(256), 0,
                               ctx->stream, n, A, k, jb);
        }
        VLM_HIP(hipGetLastError());
    }
    if (npanels > 1) {
        hipLaunchKernelGGL(lu_swap_left_kernel, dim3(k0s.back()), dim3(64), 0, ctx->stream, n,
                           A, lists, npanels);
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
