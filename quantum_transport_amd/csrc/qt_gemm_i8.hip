// qt_gemm_i8.hip -- batched complex128 GEMM emulated on the gfx950 int8
// matrix cores (the product kernel of QT_PREC_MIXED).
//
// Each row of op(A) (each column of op(B)) gets a power-of-two scale that
// brings its largest |re| or |im| into [2^(Q-1), 2^Q); the scaled real and
// imaginary parts are rounded to integers and written as ND balanced
// base-128 digits in [-64, 63] (most significant first), so a value is kept
// to Q = 7 ND - 2 bits relative to its row maximum (ND = 6 for QT_PREC_MIXED,
// Q = 40; ND = 7 for QT_PREC_MIXED47, Q = 47 at 28/21 of the MFMA work, on
// the 64x64 kernel only).  A third digit set holds re + im digit
// by digit (range [-128, 126]) for the Karatsuba product
//     Re C = P1 - P2,  Im C = P3 - P1 - P2,
//     P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi).
// Each real product is a sum over digit pairs (i, j) of int8 GEMMs with
// weight 128^(2(ND-1) - i - j); pairs with i + j <= ND - 1 are kept (21 of 36
// for ND = 6: the dropped ones weigh at most 2^-42 of the leading term), and
// pairs of equal i + j share one int32 accumulator (|sum| < ND * 2^14 * K <
// 2^31 for K <= 1024).  The accumulators are converted to FP64 and combined per
// product, so the only errors are the 40-bit rounding of the operands, the
// dropped pairs and the FP64 additions of the epilogue.
//
// Planes live in a per-stream workspace: for member m, plane p (re digits,
// im digits, digit sums: ND each), row r, k:
//     P[((m * 3 ND + p) * Rp + r) * Kp + k],  Rp, Kp = R, K rounded up to the tile,
// with zero digits in the padding, so the GEMM kernel needs no bounds checks
// on its operands.  Operands shared by every member (no member or k-point
// stride) are split once.
#include <cstdlib>
#include <map>
#include <mutex>

#include "qt_internal.h"

namespace qt {

namespace {

// ND digits per value: 3 ND planes per operand; a row maximum is scaled
// into [2^(QB-1), 2^QB), QB = 7 ND - 2 (top digit <= 32); the leading digit
// pair weighs 2^EW, EW = 7 (ND - 1).
template <int ND>
constexpr int npl() { return 3 * ND; }
template <int ND>
constexpr int qbits() { return 7 * ND - 2; }
template <int ND>
constexpr int ew() { return 7 * (ND - 1); }
constexpr int TM = 64, TN = 64, TK = 64;
constexpr int LROW = 80;       // LDS row stride in bytes (64 + 16: conflict-free 16-byte reads)
constexpr int NXCD = 8;
constexpr int KMAX = 1024;     // split kernel holds a row of up to 1024 values in registers

typedef int v4i __attribute__((ext_vector_type(4)));

// Writes the 3 ND digit planes of four consecutive values of one row (k0 a
// multiple of 4): one 32-bit store per plane.
template <int ND>
__device__ __forceinline__ void write_digits(int8_t* __restrict__ base, int64_t pstride, int k0,
                                             const double2 (&v)[4], int s) {
    constexpr int NPL = npl<ND>();
    unsigned w[NPL];
#pragma unroll
    for (int p = 0; p < NPL; ++p) w[p] = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double xr = rint(ldexp(v[q].x, s)), xi = rint(ldexp(v[q].y, s));
#pragma unroll
        for (int d = ND - 1; d >= 0; --d) {
            const double hr = floor((xr + 64.0) * (1.0 / 128.0));
            const double hi = floor((xi + 64.0) * (1.0 / 128.0));
            const int dr = (int)(xr - 128.0 * hr), di = (int)(xi - 128.0 * hi);
            xr = hr;
            xi = hi;
            w[d] |= (unsigned)(dr & 0xff) << (8 * q);
            w[ND + d] |= (unsigned)(di & 0xff) << (8 * q);
            w[2 * ND + d] |= (unsigned)((dr + di) & 0xff) << (8 * q);
        }
    }
#pragma unroll
    for (int p = 0; p < NPL; ++p) *reinterpret_cast<unsigned*>(base + p * pstride + k0) = w[p];
}

template <int ND>
__device__ __forceinline__ int row_scale(double mx) { return mx > 0.0 ? (qbits<ND>() - 1) - ilogb(mx) : 0; }

// Rows stored contiguously, Y[r][k] = X[k + r ld] (conj: imaginary part
// negated): one wave per row, each lane four consecutive k per 256.
template <int ND>
__global__ __launch_bounds__(256) void split_rows_kernel(ZOp X, int conj, int R, int K, int Rp, int Kp,
                                                         const int* kidx, int kshift, const int* active,
                                                         int8_t* __restrict__ out, int* __restrict__ sc) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    const int m = blockIdx.y;
    if (r >= Rp) return;  // wave-uniform
    if (active && !active[m]) return;
    const int kz = kidx ? kidx[m >> kshift] : 0;
    const double2* __restrict__ Xp = X.p + m * X.sb + kz * X.sk;

    double2 v[KMAX / 256][4];
    double mx = 0.0;
#pragma unroll
    for (int j = 0; j < KMAX / 256; ++j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = lane * 4 + 256 * j + q;
            double2 x = make_double2(0.0, 0.0);
            if (r < R && k < K) x = Xp[k + (int64_t)r * X.ld];
            if (conj) x.y = -x.y;
            v[j][q] = x;
            mx = fmax(mx, fmax(fabs(x.x), fabs(x.y)));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    const int s = row_scale<ND>(mx);
    if (lane == 0) sc[(int64_t)m * Rp + r] = s;

    int8_t* __restrict__ base = out + (int64_t)m * npl<ND>() * Rp * Kp + (int64_t)r * Kp;
    const int64_t pstride = (int64_t)Rp * Kp;
#pragma unroll
    for (int j = 0; j < KMAX / 256; ++j) {
        const int k0 = lane * 4 + 256 * j;
        if (k0 >= Kp) break;
        write_digits<ND>(base, pstride, k0, v[j], s);
    }
}

// Rows strided, Y[r][k] = X[r + k ld]: eight rows per workgroup, thread t
// takes row t % 8 and four consecutive k per 128 (k = 4 (t / 8) + 128 j), so
// the eight lanes of a row group read 128 contiguous bytes.
template <int ND>
__global__ __launch_bounds__(256) void split_cols_kernel(ZOp X, int conj, int R, int K, int Rp, int Kp,
                                                         const int* kidx, int kshift, const int* active,
                                                         int8_t* __restrict__ out, int* __restrict__ sc) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int rr = t & 7, kq = t >> 3;
    const int r = blockIdx.x * 8 + rr;
    const int m = blockIdx.y;
    if (active && !active[m]) return;
    const int kz = kidx ? kidx[m >> kshift] : 0;
    const double2* __restrict__ Xp = X.p + m * X.sb + kz * X.sk;
    __shared__ double red[4][8];

    double2 v[KMAX / 128][4];
    double mx = 0.0;
#pragma unroll
    for (int j = 0; j < KMAX / 128; ++j) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = kq * 4 + 128 * j + q;
            double2 x = make_double2(0.0, 0.0);
            if (r < R && k < K) x = Xp[r + (int64_t)k * X.ld];
            if (conj) x.y = -x.y;
            v[j][q] = x;
            mx = fmax(mx, fmax(fabs(x.x), fabs(x.y)));
        }
    }
    // maximum over the lanes of this row (lane % 8 == rr), then over waves
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if (lane < 8) red[wave][lane] = mx;
    __syncthreads();
    mx = fmax(fmax(red[0][rr], red[1][rr]), fmax(red[2][rr], red[3][rr]));
    const int s = row_scale<ND>(mx);
    if (r >= Rp) return;
    if (kq == 0) sc[(int64_t)m * Rp + r] = s;

    int8_t* __restrict__ base = out + (int64_t)m * npl<ND>() * Rp * Kp + (int64_t)r * Kp;
    const int64_t pstride = (int64_t)Rp * Kp;
#pragma unroll
    for (int j = 0; j < KMAX / 128; ++j) {
        const int k0 = kq * 4 + 128 * j;
        if (k0 >= Kp) break;
        write_digits<ND>(base, pstride, k0, v[j], s);
    }
}

// One workgroup (4 waves) per 64x64 tile of C and member; each wave a 32x32
// sub-tile = 2x2 blocks of 16x16.  For each of the three real products the
// K loop stages the 6 digit planes of A and B (64 rows x 64 k each) in LDS and
// issues 21 digit-pair MFMAs per block and chunk into 6 int32 accumulators
// (one per digit weight); the product is then folded into FP64 Re/Im sums.
// v_mfma_i32_16x16x64_i8 is fed (B fragment, A fragment), so lane l,
// register q holds C[row = l % 16][col = 4 * (l / 16) + q] of a block.
template <int ND>
__global__ __launch_bounds__(256, 2) void zgemm_i8_kernel(ZGemmArgs a, const int8_t* __restrict__ PA,
                                                           const int* __restrict__ SA, int64_t pa_stride,
                                                           int64_t sa_stride, const int8_t* __restrict__ PB,
                                                           const int* __restrict__ SB, int64_t pb_stride,
                                                           int64_t sb_stride, int Mp, int Np, int Kp, int tiles_m,
                                                           int tiles) {
    const int L = blockIdx.x;
    const int rest = L / NXCD;
    const int tile = rest % tiles;
    const int b = L % NXCD + NXCD * (rest / tiles);
    if (b >= a.batch) return;
    if (a.active && !a.active[b]) return;
    const int m0 = (tile % tiles_m) * TM;
    const int n0 = (tile / tiles_m) * TN;

    const int8_t* __restrict__ Ab = PA + b * pa_stride + (int64_t)m0 * Kp;
    const int8_t* __restrict__ Bb = PB + b * pb_stride + (int64_t)n0 * Kp;
    const int64_t pla = (int64_t)Mp * Kp, plb = (int64_t)Np * Kp;

    __shared__ __attribute__((aligned(16))) int8_t As[ND][TM][LROW];
    __shared__ __attribute__((aligned(16))) int8_t Bs[ND][TN][LROW];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;

    double cr[2][2][4], ci[2][2][4];
    const int nk = Kp / TK;

#pragma unroll 1
    for (int prod = 0; prod < 3; ++prod) {
        v4i acc[ND][2][2];
#pragma unroll
        for (int g = 0; g < ND; ++g)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[g][i][j] = v4i{0, 0, 0, 0};
        const int8_t* __restrict__ Ap = Ab + prod * ND * pla;
        const int8_t* __restrict__ Bp = Bb + prod * ND * plb;

#pragma unroll 1
        for (int kt = 0; kt < nk; ++kt) {
            const int k0 = kt * TK;
            v4i ra[ND], rb[ND];
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                // segment q*256 + tid: digit q, row tid/4, 16-byte part tid%4
                const int row = tid >> 2, part = tid & 3;
                ra[q] = *reinterpret_cast<const v4i*>(Ap + q * pla + (int64_t)row * Kp + k0 + 16 * part);
                rb[q] = *reinterpret_cast<const v4i*>(Bp + q * plb + (int64_t)row * Kp + k0 + 16 * part);
            }
            __syncthreads();  // previous chunk's reads done
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                const int row = tid >> 2, part = tid & 3;
                *reinterpret_cast<v4i*>(&As[q][row][16 * part]) = ra[q];
                *reinterpret_cast<v4i*>(&Bs[q][row][16 * part]) = rb[q];
            }
            __syncthreads();
            v4i bf[ND][2];
#pragma unroll
            for (int d = 0; d < ND; ++d)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    bf[d][j] = *reinterpret_cast<const v4i*>(&Bs[d][wn * 32 + j * 16 + li][16 * lk]);
#pragma unroll
            for (int di = 0; di < ND; ++di) {
                v4i af[2];
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    af[i] = *reinterpret_cast<const v4i*>(&As[di][wm * 32 + i * 16 + li][16 * lk]);
#pragma unroll
                for (int dj = 0; dj + di < ND; ++dj)
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
                            acc[di + dj][i][j] =
                                __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[dj][j], af[i], acc[di + dj][i][j], 0, 0, 0);
            }
        }
        // fold: value = 2^EW * sum_g acc_g * 2^(7 (ND - 1 - g)) / (2^s_row 2^t_col)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    double R = 0.0;
#pragma unroll
                    for (int g = ND - 1; g >= 0; --g) R += (double)acc[g][i][j][q] * (double)(1LL << (7 * (ND - 1 - g)));
                    if (prod == 0) {
                        cr[i][j][q] = R;
                        ci[i][j][q] = -R;
                    } else if (prod == 1) {
                        cr[i][j][q] -= R;
                        ci[i][j][q] -= R;
                    } else {
                        ci[i][j][q] += R;
                    }
                }
    }

    const int kz = a.kidx ? a.kidx[b >> a.kshift] : 0;
    double2* __restrict__ Cp = a.C + b * a.sbc + kz * a.skc;
    const double2* __restrict__ Dp = a.D.p ? a.D.p + b * a.D.sb + kz * a.D.sk : nullptr;
    const int* __restrict__ sa = SA + b * sa_stride;
    const int* __restrict__ sb = SB + b * sb_stride;
    const double alpha = a.alpha, beta = a.beta;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int gi = m0 + wm * 32 + i * 16 + li;
        if (gi >= a.M) continue;
        const int es = ew<ND>() - sa[gi];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gj = n0 + wn * 32 + j * 16 + 4 * lk + q;
                if (gj >= a.N || (gj >= a.skip_lo && gj < a.skip_hi)) continue;
                const int e = es - sb[gj];
                double re = alpha * ldexp(cr[i][j][q], e), im = alpha * ldexp(ci[i][j][q], e);
                if (Dp) {
                    const double2 d = Dp[gi + (int64_t)gj * a.D.ld];
                    re += beta * d.x;
                    im += beta * d.y;
                }
                Cp[gi + (int64_t)gj * a.ldc] = make_double2(re, im);
            }
        }
    }
}

// 128x64 tiles, 8 waves (4 x 2 sub-tiles of 32x32), digit planes copied from
// global memory straight into LDS (global_load_lds_dwordx4: lane L of a wave
// writes 16 bytes at base + 16 L), double buffered, one barrier per chunk.
// Row r of a plane tile keeps its four 16-byte parts in the order
// part ^ ((r >> 1) & 3), so the 16-byte fragment reads of eight consecutive
// rows fall in distinct banks.
constexpr int VM = 128, VN = 64;
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void gbl_void;

__device__ __forceinline__ int swz(int row) { return (row >> 1) & 3; }

template <int ND>
__global__ __launch_bounds__(512, 1) void zgemm_i8_lds_kernel(ZGemmArgs a, const int8_t* __restrict__ PA,
                                                               const int* __restrict__ SA, int64_t pa_stride,
                                                               int64_t sa_stride, const int8_t* __restrict__ PB,
                                                               const int* __restrict__ SB, int64_t pb_stride,
                                                               int64_t sb_stride, int Mp, int Np, int Kp,
                                                               int tiles_m, int tiles) {
    constexpr int VA = VM * ND * 64, VB = VN * ND * 64;  // bytes of the A and B tiles of one chunk
    // two buffers of 192 rows x 64 k x ND digits must fit the 160 KB LDS
    // (ND <= 6), and the 4 ND row groups of B spread evenly over 8 waves
    static_assert(2 * (VA + VB) <= 160 * 1024 && (4 * ND) % 8 == 0, "digit count outside the LDS kernel");
    const int L = blockIdx.x;
    const int rest = L / NXCD;
    const int tile = rest % tiles;
    const int b = L % NXCD + NXCD * (rest / tiles);
    if (b >= a.batch) return;
    if (a.active && !a.active[b]) return;
    const int m0 = (tile % tiles_m) * VM;
    const int n0 = (tile / tiles_m) * VN;

    const int8_t* __restrict__ Ab = PA + b * pa_stride + (int64_t)m0 * Kp;
    const int8_t* __restrict__ Bb = PB + b * pb_stride + (int64_t)n0 * Kp;
    const int64_t pla = (int64_t)Mp * Kp, plb = (int64_t)Np * Kp;

    __shared__ __attribute__((aligned(16))) int8_t lds[2][VA + VB];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 3, wn = wave >> 2;
    const int li = lane & 15, lk = lane >> 4;
    const int nk = Kp / TK, total = 3 * nk;

    // chunk c = (product c / nk, k-chunk c % nk) into buffer buf
    auto issue = [&](int c, int buf) {
        const int prod = c / nk, k0 = (c % nk) * TK;
        const int pp = lane & 3;
        const int ra = 16 * wave + (lane >> 2);
        const int8_t* srcA = Ab + (int64_t)prod * ND * pla + (int64_t)ra * Kp + k0 + 16 * (pp ^ swz(ra));
#pragma unroll
        for (int d = 0; d < ND; ++d)
            __builtin_amdgcn_global_load_lds((gbl_void*)(srcA + d * pla), (lds_void*)&lds[buf][d * VM * 64 + wave * 1024],
                                             16, 0, 0);
#pragma unroll
        for (int i = 0; i < ND / 2; ++i) {
            const int J = wave + 8 * i, d = J >> 2, grp = J & 3;
            const int rb = 16 * grp + (lane >> 2);
            const int8_t* srcB = Bb + (int64_t)(prod * ND + d) * plb + (int64_t)rb * Kp + k0 + 16 * (pp ^ swz(rb));
            __builtin_amdgcn_global_load_lds((gbl_void*)srcB, (lds_void*)&lds[buf][VA + d * VN * 64 + grp * 1024], 16,
                                             0, 0);
        }
    };

    double cr[2][2][4], ci[2][2][4];
    v4i acc[ND][2][2];
    issue(0, 0);
#pragma unroll 1
    for (int c = 0; c < total; ++c) {
        const int buf = c & 1, kt = c % nk, prod = c / nk;
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's copies of chunk c are in LDS
        __syncthreads();                     // everyone's copies landed; buffer buf^1 is free
        if (c + 1 < total) issue(c + 1, buf ^ 1);
        if (kt == 0) {
#pragma unroll
            for (int g = 0; g < ND; ++g)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) acc[g][i][j] = v4i{0, 0, 0, 0};
        }
        const int8_t* As = &lds[buf][0];
        const int8_t* Bs = &lds[buf][VA];
        v4i bf[ND][2];
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int r = wn * 32 + j * 16 + li;
                bf[d][j] = *reinterpret_cast<const v4i*>(Bs + d * VN * 64 + r * 64 + 16 * (lk ^ swz(r)));
            }
#pragma unroll
        for (int di = 0; di < ND; ++di) {
            v4i af[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int r = wm * 32 + i * 16 + li;
                af[i] = *reinterpret_cast<const v4i*>(As + di * VM * 64 + r * 64 + 16 * (lk ^ swz(r)));
            }
#pragma unroll
            for (int dj = 0; dj + di < ND; ++dj)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[di + dj][i][j] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[dj][j], af[i], acc[di + dj][i][j], 0, 0, 0);
        }
        if (kt == nk - 1) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        double R = 0.0;
#pragma unroll
                        for (int g = ND - 1; g >= 0; --g)
                            R += (double)acc[g][i][j][q] * (double)(1LL << (7 * (ND - 1 - g)));
                        if (prod == 0) {
                            cr[i][j][q] = R;
                            ci[i][j][q] = -R;
                        } else if (prod == 1) {
                            cr[i][j][q] -= R;
                            ci[i][j][q] -= R;
                        } else {
                            ci[i][j][q] += R;
                        }
                    }
        }
    }

    const int kz = a.kidx ? a.kidx[b >> a.kshift] : 0;
    double2* __restrict__ Cp = a.C + b * a.sbc + kz * a.skc;
    const double2* __restrict__ Dp = a.D.p ? a.D.p + b * a.D.sb + kz * a.D.sk : nullptr;
    const int* __restrict__ sa = SA + b * sa_stride;
    const int* __restrict__ sb = SB + b * sb_stride;
    const double alpha = a.alpha, beta = a.beta;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int gi = m0 + wm * 32 + i * 16 + li;
        if (gi >= a.M) continue;
        const int es = ew<ND>() - sa[gi];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gj = n0 + wn * 32 + j * 16 + 4 * lk + q;
                if (gj >= a.N || (gj >= a.skip_lo && gj < a.skip_hi)) continue;
                const int e = es - sb[gj];
                double re = alpha * ldexp(cr[i][j][q], e), im = alpha * ldexp(ci[i][j][q], e);
                if (Dp) {
                    const double2 d = Dp[gi + (int64_t)gj * a.D.ld];
                    re += beta * d.x;
                    im += beta * d.y;
                }
                Cp[gi + (int64_t)gj * a.ldc] = make_double2(re, im);
            }
        }
    }
}

// 2: the 128x64 LDS kernel (6 digits), 1: the 64x64 kernel (QT_I8_KERNEL=1
// forces it for 6 digits; 7 digits always run it)
int kernel_variant(int nd) {
    static const int v = [] {
        const char* e = std::getenv("QT_I8_KERNEL");
        return (e && e[0] == '1') ? 1 : 2;
    }();
    return nd == 6 ? v : 1;
}

// per-stream workspace for the digit planes, grown on demand: growth waits
// for the stream's earlier products before the old planes are freed
struct Ws {
    int8_t* planes = nullptr;
    size_t pcap = 0;
    int* scales = nullptr;
    size_t scap = 0;
};
std::mutex ws_mu;
std::map<std::pair<int, hipStream_t>, Ws>& ws_map() {
    static std::map<std::pair<int, hipStream_t>, Ws> m;
    return m;
}
Ws& workspace(hipStream_t s, size_t pbytes, size_t sints) {
    int dev = 0;
    QT_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(ws_mu);
    Ws& w = ws_map()[{dev, s}];
    if ((w.pcap < pbytes && w.planes) || (w.scap < sints && w.scales)) QT_HIP(hipStreamSynchronize(s));
    if (w.pcap < pbytes) {
        if (w.planes) QT_HIP(hipFree(w.planes));
        w.planes = nullptr;
        w.pcap = 0;
        QT_HIP(hipMalloc(&w.planes, pbytes));
        w.pcap = pbytes;
    }
    if (w.scap < sints) {
        if (w.scales) QT_HIP(hipFree(w.scales));
        w.scales = nullptr;
        w.scap = 0;
        QT_HIP(hipMalloc(&w.scales, sints * sizeof(int)));
        w.scap = sints;
    }
    return w;
}

int64_t up64(int64_t x) { return (x + 63) / 64 * 64; }

template <int ND>
void split(hipStream_t s, const ZOp& X, bool strided, bool conj, int R, int K, int Rp, int Kp, int members,
           const int* kidx, int kshift, const int* active, int8_t* out, int* sc) {
    if (strided)
        hipLaunchKernelGGL(split_cols_kernel<ND>, dim3((unsigned)(Rp / 8), members), dim3(256), 0, s, X,
                           conj ? 1 : 0, R, K, Rp, Kp, kidx, kshift, active, out, sc);
    else
        hipLaunchKernelGGL(split_rows_kernel<ND>, dim3((unsigned)(Rp / 4), members), dim3(256), 0, s, X,
                           conj ? 1 : 0, R, K, Rp, Kp, kidx, kshift, active, out, sc);
    QT_HIP(hipGetLastError());
}

template <int ND>
void zgemm_i8_nd(hipStream_t s, int opa, int opb, const ZGemmArgs& a) {
    constexpr int NPL = npl<ND>();
    const int var = kernel_variant(ND);
    const int tm = var == 2 ? VM : TM, tn = var == 2 ? VN : TN;
    const int64_t Mp = (a.M + tm - 1) / tm * tm, Np = (a.N + tn - 1) / tn * tn, Kp = up64(a.K);
    const bool sharedA = a.A.sb == 0 && (a.A.sk == 0 || !a.kidx);
    const bool sharedB = a.B.sb == 0 && (a.B.sk == 0 || !a.kidx);
    const int ma = sharedA ? 1 : a.batch, mb = sharedB ? 1 : a.batch;
    const size_t pa = (size_t)ma * NPL * Mp * Kp, pb = (size_t)mb * NPL * Np * Kp;
    const size_t sa = (size_t)ma * Mp, sb = (size_t)mb * Np;
    Ws& w = workspace(s, pa + pb, sa + sb);
    int8_t* PA = w.planes;
    int8_t* PB = w.planes + pa;
    int* SA = w.scales;
    int* SB = w.scales + sa;
    // op(A) rows: N -> X[r + k ld] (strided); T, C -> X[k + r ld]
    split<ND>(s, a.A, opa == OP_N, opa == OP_C, a.M, a.K, (int)Mp, (int)Kp, sharedA ? 1 : a.batch,
              sharedA ? nullptr : a.kidx, a.kshift, sharedA ? nullptr : a.active, PA, SA);
    // op(B) columns: N -> X[k + c ld]; T, C -> X[c + k ld] (strided)
    split<ND>(s, a.B, opb != OP_N, opb == OP_C, a.N, a.K, (int)Np, (int)Kp, sharedB ? 1 : a.batch,
              sharedB ? nullptr : a.kidx, a.kshift, sharedB ? nullptr : a.active, PB, SB);
    const int tiles_m = (int)(Mp / tm), tiles_n = (int)(Np / tn);
    const int tiles = tiles_m * tiles_n;
    const int batch_pad = (a.batch + NXCD - 1) / NXCD * NXCD;
    const int64_t blocks = (int64_t)tiles * batch_pad;
    if (blocks > 0x7fffffff) throw hip_error("zgemm_i8: grid too large", 1);
    const int64_t psa = sharedA ? 0 : (int64_t)NPL * Mp * Kp, ssa = sharedA ? 0 : Mp;
    const int64_t psb = sharedB ? 0 : (int64_t)NPL * Np * Kp, ssb = sharedB ? 0 : Np;
    if constexpr (ND == 6) {
        if (var == 2) {
            hipLaunchKernelGGL(zgemm_i8_lds_kernel<ND>, dim3((unsigned)blocks), dim3(512), 0, s, a, PA, SA, psa, ssa,
                               PB, SB, psb, ssb, (int)Mp, (int)Np, (int)Kp, tiles_m, tiles);
            QT_HIP(hipGetLastError());
            return;
        }
    }
    hipLaunchKernelGGL(zgemm_i8_kernel<ND>, dim3((unsigned)blocks), dim3(256), 0, s, a, PA, SA, psa, ssa, PB, SB,
                       psb, ssb, (int)Mp, (int)Np, (int)Kp, tiles_m, tiles);
    QT_HIP(hipGetLastError());
}

}  // namespace

bool zgemm_i8_fits(const ZGemmArgs& a) { return a.M >= 64 && a.N >= 64 && a.K >= 64 && a.K <= KMAX; }

void zgemm_i8(hipStream_t s, int opa, int opb, const ZGemmArgs& a, int digits) {
    if (!zgemm_i8_fits(a)) throw hip_error("zgemm_i8: shape outside 64 <= M, N, K and K <= 1024", 6);
    if (digits == 6)
        zgemm_i8_nd<6>(s, opa, opb, a);
    else if (digits == 7)
        zgemm_i8_nd<7>(s, opa, opb, a);
    else
        throw hip_error("zgemm_i8: digit count must be 6 or 7", 6);
}

}  // namespace qt
