// qt_inv.hip -- batched complex128 matrix inversion, blocked Gauss-Jordan
// with partial pivoting, for the Schur complements of the RGF sweeps and the
// lead decimation (the role of ZGETRF+ZGETRI in linalg.f90:42-62).
//
// Rows are never moved.  Pivoting is kept as the list pivp[j] = physical
// row chosen at step j ("lazy" pivoting); all row operations of
// Gauss-Jordan act on physical rows, so the working array ends up holding
// inv(Q A) with row j stored at physical row pivp[j], where Q moves row
// pivp[j] of A to position j.  The output gather undoes both permutations:
//     inv(A)[r][c] = W[pivp[r]][pinv[c]],   pinv = inverse of pivp.
//
// For each panel of NB columns [kb, kb+NB):
//   1. panel kernel (one workgroup per matrix, one row per thread, the
//      panel in registers): unblocked Gauss-Jordan on the n x NB panel; the
//      pivot of column j is the first-indexed row of maximal |re|+|im|
//      (the IZAMAX choice of ZGETRF) among the rows not yet pivoted; the
//      pivot row is broadcast through LDS, two barriers per column.  Every
//      row's multiplier g_rc = a_rc / pivot_c of step c is also recorded
//      (G, n x NB per member), with 1/pivot_c at the pivot row's entry.
//   2. update kernel (one workgroup per 32-column tile outside the panel
//      and member) replays the panel's NB elementary row operations on its
//      columns in the order unblocked Gauss-Jordan applies them:
//        y   = forward substitution of the pivot rows' values with the
//              multipliers pivot rows received before their own step,
//        row r (not a pivot of the panel): a_r - sum_c g_rc y_c,
//        pivot row of step c: (y_c - sum_{c' > c} g_rc' y_c') / pivot_c,
//      the sums on the FP64 matrix cores.  The pivot rows keep the y_c the
//      other rows were eliminated with, which is what gives Gauss-Jordan its
//      small residual |A X - I|; multiplying by the panel's explicit
//      inv(A_KK) instead loses up to cond(A_KK) of it.
// Blocks of n <= 64 are inverted by one kernel holding the whole matrix in
// registers (gj_small_kernel: the same pivot rule and row operations, no
// panel records).
// A zero pivot sets info (first column, 1-based) and the output block is
// zeroed (linalg.f90:50-58).
#include <algorithm>
#include <climits>
#include <cstdlib>

#include "qt_internal.h"

namespace qt {

namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double cabs1(double2 v) { return fabs(v.x) + fabs(v.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 crecip(double2 a) {
    // Smith's algorithm for 1 / (a.x + i a.y)
    if (fabs(a.x) >= fabs(a.y)) {
        const double r = a.y / a.x, d = a.x + a.y * r;
        return make_double2(1.0 / d, -r / d);
    } else {
        const double r = a.x / a.y, d = a.y + a.x * r;
        return make_double2(r / d, -1.0 / d);
    }
}

// ------------------------------------------------------------ panel (rolled)
// One DPP step of a wave-wide maximum: lane i takes max(x_i, x_src(i)); lanes
// without a source (or outside ROWMASK) keep x_i.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_max_step(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    const int tlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xF, false);
    const int thi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xF, false);
    return fmax(x, __hiloint2double(thi, tlo));
}
// maximum over the 64 lanes (row shifts 1, 2, 4, 8 leave each 16-lane
// row's maximum in its last lane; row broadcasts 15 and 31 carry it to lane 63)
__device__ __forceinline__ double wave_max(double x) {
    x = dpp_max_step<0x111, 0xF>(x);
    x = dpp_max_step<0x112, 0xF>(x);
    x = dpp_max_step<0x114, 0xF>(x);
    x = dpp_max_step<0x118, 0xF>(x);
    x = dpp_max_step<0x142, 0xA>(x);
    x = dpp_max_step<0x143, 0xC>(x);
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
    return __hiloint2double(hi, lo);
}

// Same elimination and pivot rule as gj_panel_kernel, with the column loop
// kept rolled (the fully unrolled loop is ~170 KB of code, streamed through
// the instruction cache once per panel): the panel column is picked from the
// registers by selects, the pivot search is a DPP maximum plus a ballot for
// the first row holding it, and row updates are four FMAs per element.
// The pivot row is not divided by its pivot when chosen: its thread keeps the
// row w unscaled with the factor s = 1/pivot (true row = s w, w[c] = 1), and
// multiplies once at the write-back.  Later eliminations keep their form on
// w (the multiplier g = w[c'] d' and w[c'] <- -g carry the factor), so the
// thread owning the pivot does no extra work per column; the panel is FP64
// vector-bound and every wave waits at the barrier for the slowest one.
template <int NB, int NT>
__global__ __launch_bounds__(NT) void gj_panel_loop_kernel(double2* A, int64_t sa, double2* Gm, int64_t sg,
                                                           int n, int kb, int* pivp, int* info,
                                                           const int* active) {
    constexpr int NW = NT / 64;
    const int b = blockIdx.x;
    if (active && !active[b]) return;
    __shared__ unsigned used[32];          // rows already pivoted (n <= 1024)
    __shared__ double red_v[2][NW];
    __shared__ int red_i[2][NW];
    __shared__ double2 prow[2][NB];
    __shared__ int pcol[NB];               // pivot rows chosen in this panel
    double2* Ab = A + b * sa;
    double2* Gb = Gm + b * sg;             // multiplier of row r at step c: Gb[c*n + r]
    int* pv = pivp + (int64_t)b * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncols = min(NB, n - kb);
    const int r = tid;                     // this thread's physical row (n <= NT)

    if (tid < 32) used[tid] = 0u;
    __syncthreads();
    for (int j = tid; j < kb; j += NT) atomicOr(&used[pv[j] >> 5], 1u << (pv[j] & 31));
    __syncthreads();
    bool mine = r < n && !((used[r >> 5] >> (r & 31)) & 1u);
    int zero_col = 0;                      // first zero pivot (1-based), tid 0
    double2 srow = make_double2(1.0, 0.0);  // this row's pending factor (true row = srow * v)

    double2 v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        v[c] = (r < n && c < ncols) ? Ab[r + (int64_t)(kb + c) * n] : make_double2(0.0, 0.0);

    // columns in groups of 8: the group loop stays rolled, the 8 columns of
    // a group are unrolled, so column c = 8 g + i is picked out of the
    // registers by NB / 8 selects on the uniform group index g
    constexpr int NG = NB / 8;
#pragma unroll 1
    for (int g = 0; g < NG; ++g) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int c = g * 8 + i;
        if (c >= ncols) break;  // uniform over the workgroup
        const int buf = c & 1;
        double2 vc = v[i];
#pragma unroll
        for (int gg = 1; gg < NG; ++gg)
            if (gg == g) vc = v[gg * 8 + i];
        // pivot search among rows not yet pivoted: wave maximum, first lane
        // holding it, then the first-indexed maximum over the waves
        const double cand = mine ? cabs1(vc) : -1.0;
        const double wmax = wave_max(cand);
        const unsigned long long hit = __ballot(cand == wmax);
        if (lane == 0) {
            red_v[buf][wave] = wmax;
            red_i[buf][wave] = wave * 64 + (__ffsll((long long)hit) - 1);
        }
        __syncthreads();
        // lane w < NW takes wave w's candidate; the first lane holding the
        // maximum has the smallest row (wave w's rows are 64w..64w+63)
        const double rvv = lane < NW ? red_v[buf][lane] : -2.0;
        const int riv = lane < NW ? red_i[buf][lane] : 0;
        const double bv = wave_max(rvv);
        const unsigned long long hitw = __ballot(rvv == bv);
        int p = __builtin_amdgcn_readlane(riv, __ffsll((long long)hitw) - 1);
        const int j = kb + c;
        if (p < 0 || p >= n) p = 0;  // unreachable: n - j rows are always unpivoted
        if (r == p) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) prow[buf][cc] = v[cc];
            mine = false;
        }
        // the pivot list and info go to global memory after the loop: a
        // store here would be waited for at every barrier
        if (tid == 0) {
            pcol[c] = p;
            if (!(bv > 0.0) && zero_col == 0) zero_col = j + 1;
        }
        __syncthreads();
        const double2 pivv = prow[buf][c];
        const double2 d = (pivv.x == 0.0 && pivv.y == 0.0) ? make_double2(0.0, 0.0) : crecip(pivv);
        double2 vnew;
        const double2 gf = cmul(vc, d);  // f / pivot
        // the record the update kernel replays (stores are not waited for
        // at the barriers, which wait on LDS traffic only); 32-bit offsets
        // (c n + r < 2^15), components selected one by one
        if (r < n) {
            const bool piv_row = r == p;
            double2 rec;
            rec.x = piv_row ? d.x : gf.x;
            rec.y = piv_row ? d.y : gf.y;
            Gb[(unsigned)(c * n + r)] = rec;
        }
        if (r == p) {
            srow = d;  // the row stays unscaled; true row = d * v, true v[c] = d
            vnew = make_double2(1.0, 0.0);
        } else {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) {
                const double2 u = prow[buf][cc];
                double2 t = v[cc];
                t.x = fma(-gf.x, u.x, t.x);
                t.x = fma(gf.y, u.y, t.x);
                t.y = fma(-gf.x, u.y, t.y);
                t.y = fma(-gf.y, u.x, t.y);
                v[cc] = t;
            }
            vnew = make_double2(-gf.x, -gf.y);
        }
#pragma unroll
        for (int gg = 0; gg < NG; ++gg)
            if (gg == g) v[gg * 8 + i] = vnew;
    }
    }
    if (r < n) {
        const bool scaled = srow.x != 1.0 || srow.y != 0.0;
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < ncols) Ab[r + (int64_t)(kb + c) * n] = scaled ? cmul(v[c], srow) : v[c];
    }
    if (tid < ncols) pv[kb + tid] = pcol[tid];
    if (tid == 0 && zero_col != 0 && info[b] == 0) info[b] = zero_col;
}

// ------------------------------------------------------------ update (narrow tiles)
// Workgroup = (TC-column tile, member), all n rows, NT/64 waves (NT = 512,
// TC = 32).  Phase 1: wave w owns columns 4w..4w+3 of the tile, 16 lanes per
// column, lane l holding the values of the panel's pivot rows l and l + 16 in
// that column; the forward substitution y_c' -= g_{p_c', c} y_c runs in
// registers with y_c broadcast by a lane shuffle (no barrier), then y goes to
// LDS.  Phase 2: wave w takes the 16-row blocks w, w + NT/64, ...; the
// multipliers G are read from global memory straight into MFMA operand
// registers (written by the panel kernel of the same member just before and
// shared by all its tiles, so they come from L2), with a pivot row's entries
// of its own and earlier steps zeroed; y is the other operand.  Columns are
// owned by exactly one workgroup, so the old pivot-row values a workgroup
// needs are never written by another one.  Workgroup ids are spread so that
// all tiles of member m run on the XCD the panel kernel of m ran on (ids go
// round-robin over 8 XCDs).
template <int NB, int TC, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void gj_update_cols_kernel(
    double2* A, int64_t sa, const double2* Gm, int64_t sg, int n, int batch, int ntiles, int kb, int w,
    const int* pivp, const int* active) {
    static_assert(TC % 16 == 0 && NB % 4 == 0 && NB <= 32, "tile");
    static_assert(NT / 64 * 4 == TC, "phase 1 maps 4 columns to each wave");
    constexpr int NJ = TC / 16, NW = NT / 64, KS = NB / 4;
    const int lin = blockIdx.x;
    const int b = ((lin >> 3) / ntiles) * 8 + (lin & 7);
    const int c0 = ((lin >> 3) % ntiles) * TC;
    if (b >= batch || (active && !active[b])) return;
    if (c0 >= kb && c0 + TC <= kb + w) return;  // tile entirely inside the panel
    __shared__ double2 Ys[NB][TC];     // y of the pivot rows, tile columns
    __shared__ double2 Lk[NB][NB];     // Lk[t][c] = g_{p_t, c} for c < t
    __shared__ double2 dpv[NB];        // 1 / pivot of step t
    __shared__ signed char pos[1024];  // step of row r in this panel, -1 if none
    double2* Ab = A + b * sa;
    const double2* Gb = Gm + b * sg;
    const int* pv = pivp + (int64_t)b * n;
    const int tid = threadIdx.x;
    // offsets below are 32-bit: n <= 1024, so row + col n < 2^20
    for (int r = tid; r < n; r += NT) pos[r] = -1;
    for (int e = tid; e < NB * NB; e += NT) {
        const int t = e / NB, c = e % NB;
        Lk[t][c] = (t < w && c < t) ? Gb[(unsigned)(c * n + pv[kb + t])] : make_double2(0.0, 0.0);
    }
    if (tid < w) dpv[tid] = Gb[(unsigned)(tid * n + pv[kb + tid])];
    __syncthreads();
    if (tid < w) pos[pv[kb + tid]] = (signed char)tid;

    const int wave = tid >> 6, lane = tid & 63;
    {
        // phase 1: y = forward substitution, column c0 + 4 wave + lane / 16
        const int cc = wave * 4 + (lane >> 4), col = c0 + cc;
        const bool cok = col < n && !(col >= kb && col < kb + w);
        const int ta = lane & 15, tb = ta + 16;
        double2 ya = (cok && ta < w) ? Ab[(unsigned)(pv[kb + ta] + col * n)] : make_double2(0.0, 0.0);
        double2 yb = (cok && tb < w) ? Ab[(unsigned)(pv[kb + tb] + col * n)] : make_double2(0.0, 0.0);
        const int src0 = lane & ~15;
        for (int c = 0; c + 1 < w; ++c) {
            const double2 own = c < 16 ? ya : yb;  // uniform choice
            const int src = src0 + (c & 15);
            double2 yc;
            yc.x = __shfl(own.x, src, 64);
            yc.y = __shfl(own.y, src, 64);
            if (ta > c) {
                const double2 l = Lk[ta][c];
                ya.x -= l.x * yc.x - l.y * yc.y;
                ya.y -= l.x * yc.y + l.y * yc.x;
            }
            if (tb > c && tb < NB) {
                const double2 l = Lk[tb][c];
                yb.x -= l.x * yc.x - l.y * yc.y;
                yb.y -= l.x * yc.y + l.y * yc.x;
            }
        }
        Ys[ta][cc] = ya;
        if (tb < NB) Ys[tb][cc] = yb;
    }
    __syncthreads();

    // phase 2: out_r = base_r - sum_k g'_rk y_k; every global load of a row
    // block is issued before the LDS lookups that decide how it is used
    const int li = lane & 15, lk = lane >> 4;
    for (int r0 = wave * 16; r0 < n; r0 += NW * 16) {
        const int row = r0 + li;
        const bool rok = row < n;
        double2 av[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = ks * 4 + lk;
            av[ks] = (rok && k < w) ? Gb[(unsigned)(k * n + row)] : make_double2(0.0, 0.0);
        }
        double2 base[NJ][4];
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = c0 + j * 16 + 4 * q + lk;
                base[j][q] = (rok && col < n) ? Ab[(unsigned)(row + col * n)] : make_double2(0.0, 0.0);
            }
        const int prow = rok ? pos[row] : -1;  // step of this pivot row, -1 otherwise
        const int py = prow < 0 ? 0 : prow;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const bool keep = ks * 4 + lk > prow;  // a pivot row keeps the later steps only
            av[ks].x = keep ? av[ks].x : 0.0;
            av[ks].y = keep ? av[ks].y : 0.0;
        }
        const double2 dp = dpv[py];
        // lane index through an empty asm: the y operands are re-read from
        // LDS for every row block instead of being hoisted into registers
        int ln = lane;
        asm volatile("" : "+v"(ln));
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 yv = Ys[py][j * 16 + 4 * q + lk];  // pivot rows start from y
                base[j][q].x = prow >= 0 ? yv.x : base[j][q].x;
                base[j][q].y = prow >= 0 ? yv.y : base[j][q].y;
            }
            double4_t accR = {0.0, 0.0, 0.0, 0.0}, accI = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const double2 bv = Ys[ks * 4 + (ln >> 4)][j * 16 + (ln & 15)];
                accR = __builtin_amdgcn_mfma_f64_16x16x4f64(bv.x, av[ks].x, accR, 0, 0, 0);
                accR = __builtin_amdgcn_mfma_f64_16x16x4f64(-bv.y, av[ks].y, accR, 0, 0, 0);
                accI = __builtin_amdgcn_mfma_f64_16x16x4f64(bv.y, av[ks].x, accI, 0, 0, 0);
                accI = __builtin_amdgcn_mfma_f64_16x16x4f64(bv.x, av[ks].y, accI, 0, 0, 0);
            }
            if (!rok) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = c0 + j * 16 + 4 * q + lk;
                if (col >= n || (col >= kb && col < kb + w)) continue;
                const double2 o = make_double2(base[j][q].x - accR[q], base[j][q].y - accI[q]);
                const double2 os = cmul(o, dp);
                Ab[(unsigned)(row + col * n)] = make_double2(prow >= 0 ? os.x : o.x, prow >= 0 ? os.y : o.y);
            }
        }
    }
}

// ------------------------------------------------------------ output gather
// X[r][c] = W[pivp[r]][pinv[c]]; zero block when the factorisation failed.
__global__ __launch_bounds__(256) void gj_finish_kernel(const double2* W, int64_t sw, double2* X,
                                                        int64_t sx, int n, const int* pivp,
                                                        const int* info, const int* active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    __shared__ int rows[1024];
    __shared__ int pinv[1024];
    const int* pv = pivp + (int64_t)b * n;
    for (int i = threadIdx.x; i < n; i += 256) {
        const int p = pv[i];
        rows[i] = p;
        pinv[p] = i;
    }
    __syncthreads();
    const bool bad = info[b] != 0;
    const double2* Wb = W + b * sw;
    double2* Xb = X + b * sx;
    for (int cc = 0; cc < 8; ++cc) {
        const int c = blockIdx.x * 8 + cc;
        if (c >= n) break;
        const double2* src = Wb + (int64_t)pinv[c] * n;
        double2* dst = Xb + (int64_t)c * n;
        for (int r = threadIdx.x; r < n; r += 256) dst[r] = bad ? make_double2(0.0, 0.0) : src[rows[r]];
    }
}

// ------------------------------------------------------------ blocks of n <= 64
// The whole inversion in one launch: one 256-thread workgroup per member,
// wave w holding columns 16 w .. 16 w + 15, lane r holding row r, in
// registers.  Unblocked Gauss-Jordan with the pivot rule and the row
// operations of the panel kernel (first-indexed maximal |re| + |im| among the
// rows not yet pivoted, the pivot row kept unscaled with its factor 1/pivot
// applied at the write-back).  Column j lives in one wave, which finds the
// pivot with a wave maximum and publishes the column (every row's
// multiplier) through LDS: one barrier per column (double-buffered, so the
// next owner never overwrites a buffer still being read).  Each wave takes
// the pivot row's values in its columns from lane p by readlane.  The output
// gather X[pinv[R]][pivp[C]] = W[R][C] (W: the working array, row R scaled
// by its factor) and the zero block of a failed factorisation are written by
// the same workgroup.  A and X may be the same array: every element of A is
// read before the first barrier and X is written after the last one.
__device__ __forceinline__ double readlane_f64(double x, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(256) void gj_small_kernel(const double2* A, int64_t sa, double2* X, int64_t sx,
                                                       int n, int* info, const int* active) {
    const int b = blockIdx.x;
    if (active && !active[b]) {
        if (threadIdx.x == 0) info[b] = 0;
        return;
    }
    __shared__ double2 fcol[2][64];  // column j of the working array
    __shared__ int piv_s[2];         // pivot row of step j
    __shared__ double pmax_s[2];     // its |re| + |im|
    __shared__ int pvs[64];          // physical row pivoted at step j
    __shared__ int pinv[64];         // step at which physical row R was pivoted
    __shared__ int zero_s;
    const int t = threadIdx.x, r = t & 63, w = t >> 6;
    const double2* Ab = A + b * sa;
    double2 v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int c = 16 * w + k;
        v[k] = (r < n && c < n) ? Ab[r + c * n] : make_double2(0.0, 0.0);
    }
    bool mine = r < n;  // row not yet pivoted
    double2 srow = make_double2(1.0, 0.0);
    int zero_col = 0;   // first zero pivot (1-based), thread 0
#pragma unroll 1
    for (int jw = 0; jw < 4; ++jw) {
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int j = 16 * jw + jj;
            if (j >= n) continue;  // uniform over the workgroup
            const int buf = j & 1;
            if (w == jw) {  // the wave holding column j
                const double cand = mine ? cabs1(v[jj]) : -1.0;
                const double wmax = wave_max(cand);
                const unsigned long long hit = __ballot(cand == wmax);
                fcol[buf][r] = v[jj];
                if (r == 0) {
                    piv_s[buf] = __ffsll((long long)hit) - 1;
                    pmax_s[buf] = wmax;
                }
            }
            __syncthreads();
            int p = __builtin_amdgcn_readfirstlane(piv_s[buf]);
            if (p < 0 || p >= n) p = 0;  // unreachable: n - j rows are always unpivoted
            const double2 f = fcol[buf][r];
            const double2 pivv = fcol[buf][p];
            const double2 d = (pivv.x == 0.0 && pivv.y == 0.0) ? make_double2(0.0, 0.0) : crecip(pivv);
            if (t == 0) {
                pvs[j] = p;
                pinv[p] = j;
                if (!(pmax_s[buf] > 0.0) && zero_col == 0) zero_col = j + 1;
            }
            const double2 gf = cmul(f, d);  // f / pivot
            const bool piv = r == p;
            // every row but the pivot row: v -= gf * (pivot row); the pivot
            // row's value of column k is read from lane p before it changes
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const double ux = readlane_f64(v[k].x, p), uy = readlane_f64(v[k].y, p);
                double2 x = v[k];
                x.x = fma(-gf.x, ux, x.x);
                x.x = fma(gf.y, uy, x.x);
                x.y = fma(-gf.x, uy, x.y);
                x.y = fma(-gf.y, ux, x.y);
                v[k].x = piv ? v[k].x : x.x;
                v[k].y = piv ? v[k].y : x.y;
            }
            if (w == jw) {
                v[jj].x = piv ? 1.0 : -gf.x;
                v[jj].y = piv ? 0.0 : -gf.y;
            }
            if (piv) {
                srow = d;  // the row stays unscaled; true row = d * v, true v[j] = d
                mine = false;
            }
        }
    }
    if (t == 0) {
        zero_s = zero_col;
        info[b] = zero_col;
    }
    __syncthreads();
    const bool bad = zero_s != 0;
    if (r >= n) return;
    double2* Xb = X + b * sx;
    const int orow = pinv[r];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int c = 16 * w + k;
        if (c >= n) continue;
        const double2 o = bad ? make_double2(0.0, 0.0) : cmul(v[k], srow);
        Xb[orow + (int64_t)pvs[c] * n] = o;
    }
}

__global__ void zero_int_kernel(int* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int NB, int NT>
void run_panels(hipStream_t s, const InvArgs& a) {
    const int n = a.n;
    for (int kb = 0; kb < n; kb += NB) {
        const int w = std::min(NB, n - kb);
        hipLaunchKernelGGL((gj_panel_loop_kernel<NB, NT>), dim3(a.batch), dim3(NT), 0, s, a.A, a.sa, a.G, a.sg,
                           n, kb, a.piv, a.info, a.active);
        QT_HIP(hipGetLastError());
        if (w == n) continue;
        constexpr int TC = 32;
        const int ntiles = (n + TC - 1) / TC;
        const int grid = ntiles * ((a.batch + 7) / 8) * 8;
        hipLaunchKernelGGL((gj_update_cols_kernel<NB, TC, 512>), dim3(grid), dim3(512), 0, s, a.A, a.sa, a.G,
                           a.sg, n, a.batch, ntiles, kb, w, a.piv, a.active);
        QT_HIP(hipGetLastError());
    }
}

}  // namespace

namespace {

void zinv_part(hipStream_t s, const InvArgs& a) {
    const int n = a.n;
    if (n <= 64) {
        hipLaunchKernelGGL(gj_small_kernel, dim3(a.batch), dim3(256), 0, s, a.A, a.sa, a.X, a.sx, n, a.info,
                           a.active);
        QT_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(zero_int_kernel, dim3((a.batch + 255) / 256), dim3(256), 0, s, a.info, a.batch);
    if (n <= 256)
        run_panels<32, 256>(s, a);
    else if (n <= 512)
        run_panels<32, 512>(s, a);
    else if (n <= 768)
        run_panels<32, 768>(s, a);  // 3 waves per SIMD leave room for the 32-column panel
    else
        run_panels<16, 1024>(s, a);
    hipLaunchKernelGGL(gj_finish_kernel, dim3((n + 7) / 8, a.batch), dim3(256), 0, s, a.A, a.sa, a.X,
                       a.sx, n, a.piv, a.info, a.active);
    QT_HIP(hipGetLastError());
}

InvArgs members(const InvArgs& a, int b0, int cnt) {
    InvArgs h = a;
    h.batch = cnt;
    h.A = a.A + b0 * a.sa;
    h.X = a.X + b0 * a.sx;
    h.piv = a.piv + (int64_t)b0 * a.n;
    h.G = a.G + b0 * a.sg;
    h.info = a.info + b0;
    h.active = a.active ? a.active + b0 : nullptr;
    return h;
}

// a second stream (default priority) and two events for splitting a batch
// up to three more streams (default priority) and their events for splitting
// a batch; QT_INV_PARTS (2..4, default 2) sets the number of parts
constexpr int kMaxParts = 4;
struct Side {
    hipStream_t s[kMaxParts - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxParts - 1] = {};
    int device = -1;
};
Side& side() {
    static Side sd;
    int dev = 0;
    QT_HIP(hipGetDevice(&dev));
    if (sd.device != dev) {
        QT_HIP(hipEventCreateWithFlags(&sd.fork, hipEventDisableTiming));
        for (int i = 0; i < kMaxParts - 1; ++i) {
            QT_HIP(hipStreamCreate(&sd.s[i]));
            QT_HIP(hipEventCreateWithFlags(&sd.join[i], hipEventDisableTiming));
        }
        sd.device = dev;
    }
    return sd;
}
int inv_parts() {
    static const int v = [] {
        const char* e = std::getenv("QT_INV_PARTS");
        const int p = e ? std::atoi(e) : 2;
        return p < 2 ? 2 : (p > kMaxParts ? kMaxParts : p);
    }();
    return v;
}

}  // namespace

// Blocks of n <= 64 take the one-launch kernel above on stream s.  For larger
// ones, the panel kernel keeps one workgroup per member busy for a whole latency-
// bound column sweep, leaving most CUs idle; the batch is therefore split in
// parts on as many streams so that the panels of one part run while the
// MFMA updates of the others fill the remaining CUs.
void zinv_single(hipStream_t s, const InvArgs& a) {
    if (a.n <= 0 || a.batch <= 0) return;
    if (a.n > 1024) throw hip_error("zinv: block size above 1024", 6);
    if (!a.G || a.sg < (int64_t)a.n * kInvPanelMax) throw hip_error("zinv: multiplier workspace missing", 1);
    zinv_part(s, a);
}

void zinv(hipStream_t s, const InvArgs& a, Timer* t) {
    const int n = a.n;
    if (n <= 0 || a.batch <= 0) return;
    if (n > 1024) throw hip_error("zinv: block size above 1024", 6);
    if (!a.G || a.sg < (int64_t)n * kInvPanelMax) throw hip_error("zinv: multiplier workspace missing", 1);
    if (t && t->enabled) t->begin(s, 1);
    if (a.batch >= 8 && n > 64) {
        Side& sd = side();
        const int parts = std::min(inv_parts(), a.batch / 4);
        QT_HIP(hipEventRecord(sd.fork, s));
        // part p holds members [p*batch/parts, (p+1)*batch/parts); part 0 stays on s
        for (int p = 1; p < parts; ++p) {
            const int b0 = p * a.batch / parts, b1 = (p + 1) * a.batch / parts;
            QT_HIP(hipStreamWaitEvent(sd.s[p - 1], sd.fork, 0));
            zinv_part(sd.s[p - 1], members(a, b0, b1 - b0));
        }
        zinv_part(s, members(a, 0, a.batch / parts));
        for (int p = 1; p < parts; ++p) {
            QT_HIP(hipEventRecord(sd.join[p - 1], sd.s[p - 1]));
            QT_HIP(hipStreamWaitEvent(s, sd.join[p - 1], 0));
        }
    } else {
        zinv_part(s, a);
    }
    if (t && t->enabled)
        t->end(s, 1, 8.0 * (double)n * n * n * (a.flop_batch >= 0 ? a.flop_batch : a.batch));
}

}  // namespace qt
