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
//      pivot row is broadcast through LDS, two barriers per column.
//   2. update kernel (one workgroup per 32-column tile outside the panel
//      and member): copies the NB pivot rows of its columns to LDS (R), then
//      for every 16-row block adds P R on the FP64 matrix cores, with P the
//      new panel columns and the pivot rows' old values dropped:
//        A_KO <- inv(A_KK) A_KO,  A_OO <- A_OO - A_OK inv(A_KK) A_KO.
// The first versions of both kernels (fully unrolled panel, 64-column
// update tiles) are kept behind QT_INV_PANEL=unrolled / QT_INV_UPDATE=tiles64.
// A zero pivot sets info (first column, 1-based) and the output block is
// zeroed (linalg.f90:50-58).
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <string>

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

// ------------------------------------------------------------ panel
template <int NB, int NT>
__global__ __launch_bounds__(NT) void gj_panel_kernel(double2* A, int64_t sa, int n, int kb,
                                                      int* pivp, int* info, const int* active) {
    constexpr int NW = NT / 64;
    const int b = blockIdx.x;
    if (active && !active[b]) return;
    __shared__ unsigned used[32];          // rows already pivoted (n <= 1024)
    __shared__ double red_v[2][NW];
    __shared__ int red_i[2][NW];
    __shared__ double2 prow[2][NB];
    double2* Ab = A + b * sa;
    int* pv = pivp + (int64_t)b * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncols = min(NB, n - kb);
    const int r = tid;                     // this thread's physical row (n <= NT)

    if (tid < 32) used[tid] = 0u;
    __syncthreads();
    for (int j = tid; j < kb; j += NT) atomicOr(&used[pv[j] >> 5], 1u << (pv[j] & 31));
    __syncthreads();
    bool mine = r < n && !((used[r >> 5] >> (r & 31)) & 1u);

    double2 v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
        v[c] = (r < n && c < ncols) ? Ab[r + (int64_t)(kb + c) * n] : make_double2(0.0, 0.0);

#pragma clang loop unroll(full)
    for (int c = 0; c < NB; ++c) {
        if (c >= ncols) continue;  // uniform over the workgroup
        const int buf = c & 1;
        // pivot search among rows not yet pivoted
        double best = -1.0;
        int bi = INT_MAX;
        if (mine) {
            best = cabs1(v[c]);
            bi = r;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if (lane == 0) { red_v[buf][wave] = best; red_i[buf][wave] = bi; }
        __syncthreads();
        double bv = red_v[buf][0];
        int p = red_i[buf][0];
#pragma unroll
        for (int w = 1; w < NW; ++w)
            if (red_v[buf][w] > bv || (red_v[buf][w] == bv && red_i[buf][w] < p)) {
                bv = red_v[buf][w];
                p = red_i[buf][w];
            }
        const int j = kb + c;
        if (p < 0 || p >= n) p = 0;  // unreachable: n - j rows are always unpivoted
        if (r == p) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) prow[buf][cc] = v[cc];
            mine = false;
        }
        if (tid == 0) {
            pv[j] = p;
            if (!(bv > 0.0) && info[b] == 0) info[b] = j + 1;
        }
        __syncthreads();
        const double2 pivv = prow[buf][c];
        const double2 d = (pivv.x == 0.0 && pivv.y == 0.0) ? make_double2(0.0, 0.0) : crecip(pivv);
        if (r == p) {
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) v[cc] = (cc == c) ? d : cmul(v[cc], d);
        } else if (r < n) {
            const double2 g = cmul(v[c], d);  // f / pivot
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) {
                if (cc == c) {
                    v[cc] = make_double2(-g.x, -g.y);
                } else {
                    const double2 u = prow[buf][cc];
                    v[cc].x -= g.x * u.x - g.y * u.y;
                    v[cc].y -= g.x * u.y + g.y * u.x;
                }
            }
        }
    }
    if (r < n) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < ncols) Ab[r + (int64_t)(kb + c) * n] = v[c];
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
__global__ __launch_bounds__(NT) void gj_panel_loop_kernel(double2* A, int64_t sa, int n, int kb,
                                                           int* pivp, int* info, const int* active) {
    constexpr int NW = NT / 64;
    const int b = blockIdx.x;
    if (active && !active[b]) return;
    __shared__ unsigned used[32];          // rows already pivoted (n <= 1024)
    __shared__ double red_v[2][NW];
    __shared__ int red_i[2][NW];
    __shared__ double2 prow[2][NB];
    __shared__ int pcol[NB];               // pivot rows chosen in this panel
    double2* Ab = A + b * sa;
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
        if (r == p) {
            srow = d;  // the row stays unscaled; true row = d * v, true v[c] = d
            vnew = make_double2(1.0, 0.0);
        } else {
            const double2 gf = cmul(vc, d);  // f / pivot
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

// ------------------------------------------------------------ update
// Workgroup = (64-column tile, member); 4 waves, each a 32x32 sub-tile of
// every 64-row block; K = panel width, complex product as four
// v_mfma_f64_16x16x4 (operand order as in the GEMM kernel).
template <int NB>
__global__ __launch_bounds__(256) void gj_update_kernel(double2* A, int64_t sa, int n, int kb, int w,
                                                        const int* pivp, const int* active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int c0 = blockIdx.x * 64;
    if (c0 >= kb && c0 + 64 <= kb + w) return;  // tile entirely inside the panel
    __shared__ double2 Rs[NB][64];
    __shared__ double2 Ps[NB][64];
    __shared__ unsigned pmask[32];
    double2* Ab = A + b * sa;
    const int* pv = pivp + (int64_t)b * n;
    const int tid = threadIdx.x;
    if (tid < 32) pmask[tid] = 0u;
    __syncthreads();
    if (tid < w) atomicOr(&pmask[pv[kb + tid] >> 5], 1u << (pv[kb + tid] & 31));
    // R = pivot rows of this tile's columns (old values)
    for (int e = tid; e < NB * 64; e += 256) {
        const int t = e >> 6, c = e & 63, col = c0 + c;
        Rs[t][c] = (t < w && col < n && !(col >= kb && col < kb + w))
                       ? Ab[pv[kb + t] + (int64_t)col * n]
                       : make_double2(0.0, 0.0);
    }
    __syncthreads();

    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;
    for (int rb = 0; rb < n; rb += 64) {
        // P block: rows rb..rb+63 of the panel columns, stored [k][i]
        for (int e = tid; e < NB * 64; e += 256) {
            const int k = e >> 6, i = e & 63, row = rb + i;
            Ps[k][i] = (k < w && row < n) ? Ab[row + (int64_t)(kb + k) * n] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        double4_t accR[2][2], accI[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                accR[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
                accI[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
            }
#pragma unroll
        for (int kq = 0; kq < NB / 4; ++kq) {
            const int k = kq * 4 + lk;
            double2 av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = Ps[k][wm * 32 + i * 16 + li];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Rs[k][wn * 32 + j * 16 + li];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].x, accR[i][j], 0, 0, 0);
                    accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bv[j].y, av[i].y, accR[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].y, av[i].x, accI[i][j], 0, 0, 0);
                    accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].y, accI[i][j], 0, 0, 0);
                }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int row = rb + wm * 32 + i * 16 + li;
            if (row >= n) continue;
            const bool prow_ = (pmask[row >> 5] >> (row & 31)) & 1u;
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int col = c0 + wn * 32 + j * 16 + 4 * q + lk;
                    if (col >= n || (col >= kb && col < kb + w)) continue;
                    double2* dst = Ab + row + (int64_t)col * n;
                    double2 base = prow_ ? make_double2(0.0, 0.0) : *dst;
                    *dst = make_double2(base.x + accR[i][j][q], base.y + accI[i][j][q]);
                }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ update (narrow tiles)
// Workgroup = (TC-column tile, member), all n rows, NT/64 waves.  Wave w
// takes the 16-row blocks w, w + NT/64, ...; the panel operand P is read
// from global memory straight into MFMA operand registers (it is written by
// the panel kernel of the same member just before and is shared by all the
// member's tiles, so it comes from L2), the pivot rows R of the tile's
// columns sit in LDS.  Columns are owned by exactly one workgroup, so the
// old pivot-row values each workgroup needs are never written by another
// one.  Workgroup ids are spread so that all tiles of member m run on the
// XCD the panel kernel of m ran on (ids go round-robin over 8 XCDs).
template <int NB, int TC, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void gj_update_cols_kernel(double2* A, int64_t sa, int n, int batch,
                                                            int ntiles, int kb, int w,
                                                            const int* pivp, const int* active) {
    static_assert(TC % 16 == 0 && NB % 4 == 0, "tile");
    constexpr int NJ = TC / 16, NW = NT / 64, KS = NB / 4;
    const int lin = blockIdx.x;
    const int b = ((lin >> 3) / ntiles) * 8 + (lin & 7);
    const int c0 = ((lin >> 3) % ntiles) * TC;
    if (b >= batch || (active && !active[b])) return;
    if (c0 >= kb && c0 + TC <= kb + w) return;  // tile entirely inside the panel
    __shared__ double2 Rs[NB][TC];
    __shared__ unsigned pmask[32];
    double2* Ab = A + b * sa;
    const int* pv = pivp + (int64_t)b * n;
    const int tid = threadIdx.x;
    if (tid < 32) pmask[tid] = 0u;
    __syncthreads();
    if (tid < w) atomicOr(&pmask[pv[kb + tid] >> 5], 1u << (pv[kb + tid] & 31));
    for (int e = tid; e < NB * TC; e += NT) {
        const int t = e / TC, c = e % TC, col = c0 + c;
        Rs[t][c] = (t < w && col < n && !(col >= kb && col < kb + w))
                       ? Ab[pv[kb + t] + (int64_t)col * n]
                       : make_double2(0.0, 0.0);
    }
    __syncthreads();

    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    for (int r0 = wave * 16; r0 < n; r0 += NW * 16) {
        const int row = r0 + li;
        const bool rok = row < n;
        double2 av[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = ks * 4 + lk;
            av[ks] = (rok && k < w) ? Ab[row + (int64_t)(kb + k) * n] : make_double2(0.0, 0.0);
        }
        const bool prow_ = rok && ((pmask[row >> 5] >> (row & 31)) & 1u);
        // lane index through an empty asm: the R operands are re-read from
        // LDS for every row block instead of being hoisted into registers
        int ln = lane;
        asm volatile("" : "+v"(ln));
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            double2 base[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = c0 + j * 16 + 4 * q + lk;
                base[q] = (rok && !prow_ && col < n) ? Ab[row + (int64_t)col * n] : make_double2(0.0, 0.0);
            }
            double4_t accR = {0.0, 0.0, 0.0, 0.0}, accI = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const double2 bv = Rs[ks * 4 + (ln >> 4)][j * 16 + (ln & 15)];
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
                Ab[row + (int64_t)col * n] = make_double2(base[q].x + accR[q], base[q].y + accI[q]);
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

__global__ void zero_int_kernel(int* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Kernel choices, read once from the environment: QT_INV_UPDATE=tiles64
// selects the 64-column update kernel (default: 32-column tiles),
// QT_INV_PANEL=unrolled the fully unrolled panel kernel (default: rolled).
bool env_is(const char* name, const char* value) {
    const char* e = std::getenv(name);
    return e && std::string(e) == value;
}
int update_variant() {
    static const int v = env_is("QT_INV_UPDATE", "tiles64") ? 0 : 1;
    return v;
}
int panel_variant() {
    static const int v = env_is("QT_INV_PANEL", "unrolled") ? 0 : 1;
    return v;
}

template <int NB, int NT>
void run_panels(hipStream_t s, const InvArgs& a) {
    const int n = a.n;
    for (int kb = 0; kb < n; kb += NB) {
        const int w = std::min(NB, n - kb);
        if (panel_variant() == 1)
            hipLaunchKernelGGL((gj_panel_loop_kernel<NB, NT>), dim3(a.batch), dim3(NT), 0, s, a.A, a.sa, n,
                               kb, a.piv, a.info, a.active);
        else
            hipLaunchKernelGGL((gj_panel_kernel<NB, NT>), dim3(a.batch), dim3(NT), 0, s, a.A, a.sa, n, kb,
                               a.piv, a.info, a.active);
        QT_HIP(hipGetLastError());
        if (w == n) continue;
        if (update_variant() == 1) {
            constexpr int TC = 32;
            const int ntiles = (n + TC - 1) / TC;
            const int grid = ntiles * ((a.batch + 7) / 8) * 8;
            hipLaunchKernelGGL((gj_update_cols_kernel<NB, TC, 512>), dim3(grid), dim3(512), 0, s, a.A, a.sa,
                               n, a.batch, ntiles, kb, w, a.piv, a.active);
        } else {
            hipLaunchKernelGGL((gj_update_kernel<NB>), dim3((n + 63) / 64, a.batch), dim3(256), 0, s, a.A,
                               a.sa, n, kb, w, a.piv, a.active);
        }
        QT_HIP(hipGetLastError());
    }
}

}  // namespace

namespace {

void zinv_part(hipStream_t s, const InvArgs& a) {
    const int n = a.n;
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

// The panel kernel keeps one workgroup per member busy for a whole latency-
// bound column sweep, leaving most CUs idle; the batch is therefore split in
// parts on as many streams so that the panels of one part run while the
// MFMA updates of the others fill the remaining CUs.
void zinv_single(hipStream_t s, const InvArgs& a) {
    if (a.n <= 0 || a.batch <= 0) return;
    if (a.n > 1024) throw hip_error("zinv: block size above 1024", 6);
    zinv_part(s, a);
}

void zinv(hipStream_t s, const InvArgs& a, Timer* t) {
    const int n = a.n;
    if (n <= 0 || a.batch <= 0) return;
    if (n > 1024) throw hip_error("zinv: block size above 1024", 6);
    if (t && t->enabled) t->begin(s, 1);
    if (a.batch >= 8 && n >= 64) {
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
