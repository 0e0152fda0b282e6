// qt_gemm.hip -- batched complex128 GEMM on the gfx950 FP64 matrix cores.
//
// One workgroup (4 waves, 256 threads) computes a 64x64 complex tile of
// C = alpha*op(A)op(B) + beta*D for one batch member.  K is staged through
// LDS in chunks of 16 complex values, double buffered (2 x 32 KB).  Each
// wave owns a 32x32 sub-tile = 2x2 MFMA tiles of 16x16; a complex product is
// four v_mfma_f64_16x16x4 (re*re, -im*im into Re; re*im, im*re into Im), or,
// in the three-multiplication form (G3, QT_PREC_FP64_3M), three:
//     P1 = Ar Br,  P2 = Ai Bi,  P3 = (Ar + Ai)(Br + Bi)  (sums rounded to FP64),
//     Re = P1 - P2,  Im = (P3 - P1) - P2  (in the epilogue),
// 3/4 of the matrix-core work, with the imaginary part's rounding error
// bounded by |Ar + Ai||Br + Bi| + |Ar||Br| + |Ai||Bi| instead of
// |Ar||Bi| + |Ai||Br| (Higham, Accuracy and Stability of Numerical
// Algorithms, 2nd ed., sec. 23.2.4).
//
// Operand order: the MFMA is fed (B-fragment, A-fragment), so that the
// 16x16 result lands transposed in the lanes.  v_mfma_f64_16x16x4f64 leaves
// element (r, c) of its product in lane c + 16*(r%4), register r/4 (measured
// on gfx950); with the swapped operands lane l, register q holds
// C[row = l%16][col = 4*q + l/16], so the 16 lanes of a quarter wave write
// 16 consecutive rows of one column (256 B, coalesced for column-major C).
//
// Work mapping: the 1-D grid is ordered so that all tiles of one batch
// member are dispatched to the same XCD (workgroups are dealt round-robin
// over the 8 XCDs), keeping that member's operands in one L2.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "qt_internal.h"

namespace qt {

namespace {

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int NXCD = 8;

typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cconj(double2 v) { return make_double2(v.x, -v.y); }

template <int OPA, int OPB, bool G3>
__global__ __launch_bounds__(256, 2) void zgemm_mfma_kernel(ZGemmArgs a, int tiles_m, int tiles,
                                                            int batch_pad) {
    // XCD-aware decode: linear id L -> (member, tile), member%8 == L%8
    const int L = blockIdx.x;
    const int lane8 = L % NXCD;
    const int rest = L / NXCD;
    const int tile = rest % tiles;
    const int b = lane8 + NXCD * (rest / tiles);
    if (b >= a.batch) return;
    if (a.active && !a.active[b]) return;
    (void)batch_pad;

    const int kz = a.kidx ? a.kidx[b >> a.kshift] : 0;
    const double2* __restrict__ Ap = a.A.p + b * a.A.sb + kz * a.A.sk;
    const double2* __restrict__ Bp = a.B.p + b * a.B.sb + kz * a.B.sk;
    const int64_t lda = a.A.ld, ldb = a.B.ld;
    const int M = a.M, N = a.N, K = a.K;
    const int m0 = (tile % tiles_m) * BM;
    const int n0 = (tile / tiles_m) * BN;

    __shared__ double2 As[2][BK][BM];
    __shared__ double2 Bs[2][BK][BN];

    const int tid = threadIdx.x;
    double2 ra[4], rb[4];

    auto loadA = [&](int k0) {
        if (OPA == OP_N) {
            const int kk = tid >> 4, ii = (tid & 15) * 4;
            const int gk = k0 + kk;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gi = m0 + ii + q;
                ra[q] = (gi < M && gk < K) ? Ap[gi + gk * lda] : make_double2(0.0, 0.0);
            }
        } else {
            const int ii = tid >> 2, kk = (tid & 3) * 4;
            const int gi = m0 + ii;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gk = k0 + kk + q;
                ra[q] = (gi < M && gk < K) ? Ap[gk + gi * lda] : make_double2(0.0, 0.0);
            }
        }
    };
    auto storeA = [&](int buf) {
        if (OPA == OP_N) {
            const int kk = tid >> 4, ii = (tid & 15) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) As[buf][kk][ii + q] = ra[q];
        } else {
            const int ii = tid >> 2, kk = (tid & 3) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) As[buf][kk + q][ii] = OPA == OP_C ? cconj(ra[q]) : ra[q];
        }
    };
    auto loadB = [&](int k0) {
        if (OPB == OP_N) {
            const int jj = tid >> 2, kk = (tid & 3) * 4;
            const int gj = n0 + jj;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gk = k0 + kk + q;
                rb[q] = (gj < N && gk < K) ? Bp[gk + gj * ldb] : make_double2(0.0, 0.0);
            }
        } else {
            const int kk = tid >> 4, jj = (tid & 15) * 4;
            const int gk = k0 + kk;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gj = n0 + jj + q;
                rb[q] = (gj < N && gk < K) ? Bp[gj + gk * ldb] : make_double2(0.0, 0.0);
            }
        }
    };
    auto storeB = [&](int buf) {
        if (OPB == OP_N) {
            const int jj = tid >> 2, kk = (tid & 3) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[buf][kk + q][jj] = rb[q];
        } else {
            const int kk = tid >> 4, jj = (tid & 15) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[buf][kk][jj + q] = OPB == OP_C ? cconj(rb[q]) : rb[q];
        }
    };

    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;

    // G3: accR = P1, accI = P2, acc3 = P3
    double4_t accR[2][2], accI[2][2], acc3[G3 ? 2 : 1][G3 ? 2 : 1];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accR[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
            accI[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
            if constexpr (G3) acc3[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
        }

    const int nk = (K + BK - 1) / BK;
    loadA(0);
    loadB(0);
    storeA(0);
    storeB(0);
    __syncthreads();

    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) {
            loadA((kt + 1) * BK);
            loadB((kt + 1) * BK);
        }
#pragma unroll
        for (int kq = 0; kq < BK / 4; ++kq) {
            const int k = kq * 4 + lk;
            double2 av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = As[buf][k][wm * 32 + i * 16 + li];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Bs[buf][k][wn * 32 + j * 16 + li];
            if constexpr (G3) {
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].x, accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].y, av[i].y, accI[i][j], 0, 0, 0);
                        acc3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x + bv[j].y, av[i].x + av[i].y,
                                                                         acc3[i][j], 0, 0, 0);
                    }
            } else {
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
        }
        if (kt + 1 < nk) {
            storeA(buf ^ 1);
            storeB(buf ^ 1);
        }
        __syncthreads();
    }

    // epilogue
    double2* __restrict__ Cp = a.C + b * a.sbc + kz * a.skc;
    const double2* __restrict__ Dp = a.D.p ? a.D.p + b * a.D.sb + kz * a.D.sk : nullptr;
    const double alpha = a.alpha, beta = a.beta;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int gi = m0 + wm * 32 + i * 16 + li;
        if (gi >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gj = n0 + wn * 32 + j * 16 + 4 * q + lk;
                if (gj >= N || (gj >= a.skip_lo && gj < a.skip_hi)) continue;
                double cr = accR[i][j][q], ci = accI[i][j][q];
                if constexpr (G3) {
                    const double p1 = accR[i][j][q], p2 = accI[i][j][q];
                    cr = p1 - p2;
                    ci = (acc3[i][j][q] - p1) - p2;
                }
                double re = alpha * cr, im = alpha * ci;
                if (Dp) {
                    const double2 d = Dp[gi + gj * a.D.ld];
                    re += beta * d.x;
                    im += beta * d.y;
                }
                Cp[gi + gj * a.ldc] = make_double2(re, im);
            }
        }
    }
}

// Wider variant: 128x64 block tile, K staged 8 at a time (2 x 24 KB LDS),
// each wave a 64x32 sub-tile (4x2 MFMA tiles): twice the MFMAs per LDS read
// of the 64x64 kernel and fewer global loads per flop.
constexpr int WBM = 128, WBN = 64, WBK = 8;

template <int OPA, int OPB, bool G3>
__global__ __launch_bounds__(256, 2) void zgemm_mfma_wide_kernel(ZGemmArgs a, int tiles_m, int tiles) {
    const int L = blockIdx.x;
    const int lane8 = L % NXCD;
    const int rest = L / NXCD;
    const int tile = rest % tiles;
    const int b = lane8 + NXCD * (rest / tiles);
    if (b >= a.batch) return;
    if (a.active && !a.active[b]) return;

    const int kz = a.kidx ? a.kidx[b >> a.kshift] : 0;
    const double2* __restrict__ Ap = a.A.p + b * a.A.sb + kz * a.A.sk;
    const double2* __restrict__ Bp = a.B.p + b * a.B.sb + kz * a.B.sk;
    const int64_t lda = a.A.ld, ldb = a.B.ld;
    const int M = a.M, N = a.N, K = a.K;
    const int m0 = (tile % tiles_m) * WBM;
    const int n0 = (tile / tiles_m) * WBN;

    // Every load pattern below has consecutive lanes either along the LDS row
    // (contiguous stores) or along k (8 lanes = the 8 rows of a stage).  The
    // rows are padded to 137 and 73 complex (row stride = 144 mod 256 bytes),
    // so 8 lanes writing 8 different rows hit 8 distinct 16-byte bank groups.
    __shared__ double2 As[2][WBK][WBM + 9];
    __shared__ double2 Bs[2][WBK][WBN + 9];

    const int tid = threadIdx.x;
    double2 ra[4], rb[2];

    auto loadA = [&](int k0) {
        if (OPA == OP_N) {
            // element e = tid + 256 q: rows contiguous across lanes
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = tid + 256 * q, kk = e >> 7, ii = e & 127;
                const int gi = m0 + ii, gk = k0 + kk;
                ra[q] = (gi < M && gk < K) ? Ap[gi + gk * lda] : make_double2(0.0, 0.0);
            }
        } else {
            // 8 lanes cover the 8 k of one row of op(A) (128 B contiguous)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int kk = tid & 7, ii = (tid >> 3) + 32 * q;
                const int gi = m0 + ii, gk = k0 + kk;
                ra[q] = (gi < M && gk < K) ? Ap[gk + gi * lda] : make_double2(0.0, 0.0);
            }
        }
    };
    auto storeA = [&](int buf) {
        if (OPA == OP_N) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = tid + 256 * q;
                As[buf][e >> 7][e & 127] = ra[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                As[buf][tid & 7][(tid >> 3) + 32 * q] = OPA == OP_C ? cconj(ra[q]) : ra[q];
        }
    };
    auto loadB = [&](int k0) {
        if (OPB == OP_N) {
            // 8 lanes cover the 8 k of one column (128 B contiguous)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int kk = tid & 7, jj = (tid >> 3) + 32 * q;
                const int gj = n0 + jj, gk = k0 + kk;
                rb[q] = (gj < N && gk < K) ? Bp[gk + gj * ldb] : make_double2(0.0, 0.0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int jj = tid & 63, kk = (tid >> 6) + 4 * q;
                const int gj = n0 + jj, gk = k0 + kk;
                rb[q] = (gj < N && gk < K) ? Bp[gj + gk * ldb] : make_double2(0.0, 0.0);
            }
        }
    };
    auto storeB = [&](int buf) {
        if (OPB == OP_N) {
#pragma unroll
            for (int q = 0; q < 2; ++q) Bs[buf][tid & 7][(tid >> 3) + 32 * q] = rb[q];
        } else {
#pragma unroll
            for (int q = 0; q < 2; ++q)
                Bs[buf][(tid >> 6) + 4 * q][tid & 63] = OPB == OP_C ? cconj(rb[q]) : rb[q];
        }
    };

    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;

    // G3: accR = P1, accI = P2, acc3 = P3
    double4_t accR[4][2], accI[4][2], acc3[G3 ? 4 : 1][G3 ? 2 : 1];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accR[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
            accI[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
            if constexpr (G3) acc3[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
        }

    const int nk = (K + WBK - 1) / WBK;
    loadA(0);
    loadB(0);
    storeA(0);
    storeB(0);
    __syncthreads();

    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) {
            loadA((kt + 1) * WBK);
            loadB((kt + 1) * WBK);
        }
#pragma unroll
        for (int kq = 0; kq < WBK / 4; ++kq) {
            const int k = kq * 4 + lk;
            if constexpr (G3) {
                // one A fragment at a time keeps the registers within two
                // workgroups per CU
                double2 bv[2];
                double bs[2];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    bv[j] = Bs[buf][k][wn * 32 + j * 16 + li];
                    bs[j] = bv[j].x + bv[j].y;
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const double2 av = As[buf][k][wm * 64 + i * 16 + li];
                    const double as = av.x + av.y;
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av.x, accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].y, av.y, accI[i][j], 0, 0, 0);
                        acc3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bs[j], as, acc3[i][j], 0, 0, 0);
                    }
                }
            } else {
                double2 av[4], bv[2];
#pragma unroll
                for (int i = 0; i < 4; ++i) av[i] = As[buf][k][wm * 64 + i * 16 + li];
#pragma unroll
                for (int j = 0; j < 2; ++j) bv[j] = Bs[buf][k][wn * 32 + j * 16 + li];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].x, accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].y, av[i].x, accI[i][j], 0, 0, 0);
                    }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bv[j].y, av[i].y, accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].y, accI[i][j], 0, 0, 0);
                    }
            }
        }
        if (kt + 1 < nk) {
            storeA(buf ^ 1);
            storeB(buf ^ 1);
        }
        __syncthreads();
    }

    double2* __restrict__ Cp = a.C + b * a.sbc + kz * a.skc;
    const double2* __restrict__ Dp = a.D.p ? a.D.p + b * a.D.sb + kz * a.D.sk : nullptr;
    const double alpha = a.alpha, beta = a.beta;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = m0 + wm * 64 + i * 16 + li;
        if (gi >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gj = n0 + wn * 32 + j * 16 + 4 * q + lk;
                if (gj >= N || (gj >= a.skip_lo && gj < a.skip_hi)) continue;
                double cr = accR[i][j][q], ci = accI[i][j][q];
                if constexpr (G3) {
                    const double p1 = accR[i][j][q], p2 = accI[i][j][q];
                    cr = p1 - p2;
                    ci = (acc3[i][j][q] - p1) - p2;
                }
                double re = alpha * cr, im = alpha * ci;
                if (Dp) {
                    const double2 d = Dp[gi + gj * a.D.ld];
                    re += beta * d.x;
                    im += beta * d.y;
                }
                Cp[gi + gj * a.ldc] = make_double2(re, im);
            }
        }
    }
}

// 128x64 tiles with 8 waves of 32x32 (2x2 MFMA tiles each): the 128x64
// kernel's LDS staging and global loads, the 64x64 kernel's per-wave register
// tile, so the accumulators of the three-multiplication form fit beside
// more waves per CU.  Every output element sees the MFMAs of the other two
// kernels in the same order.
template <int OPA, int OPB, bool G3>
__global__ __launch_bounds__(512, G3 ? 1 : 2) void zgemm_mfma_w8_kernel(ZGemmArgs a, int tiles_m, int tiles) {
    const int L = blockIdx.x;
    const int lane8 = L % NXCD;
    const int rest = L / NXCD;
    const int tile = rest % tiles;
    const int b = lane8 + NXCD * (rest / tiles);
    if (b >= a.batch) return;
    if (a.active && !a.active[b]) return;

    const int kz = a.kidx ? a.kidx[b >> a.kshift] : 0;
    const double2* __restrict__ Ap = a.A.p + b * a.A.sb + kz * a.A.sk;
    const double2* __restrict__ Bp = a.B.p + b * a.B.sb + kz * a.B.sk;
    const int64_t lda = a.A.ld, ldb = a.B.ld;
    const int M = a.M, N = a.N, K = a.K;
    const int m0 = (tile % tiles_m) * WBM;
    const int n0 = (tile / tiles_m) * WBN;

    __shared__ double2 As[2][WBK][WBM + 9];
    __shared__ double2 Bs[2][WBK][WBN + 9];

    const int tid = threadIdx.x;
    double2 ra[2], rb;

    auto loadA = [&](int k0) {
        if (OPA == OP_N) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int e = tid + 512 * q, kk = e >> 7, ii = e & 127;
                const int gi = m0 + ii, gk = k0 + kk;
                ra[q] = (gi < M && gk < K) ? Ap[gi + gk * lda] : make_double2(0.0, 0.0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int kk = tid & 7, ii = (tid >> 3) + 64 * q;
                const int gi = m0 + ii, gk = k0 + kk;
                ra[q] = (gi < M && gk < K) ? Ap[gk + gi * lda] : make_double2(0.0, 0.0);
            }
        }
    };
    auto storeA = [&](int buf) {
        if (OPA == OP_N) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const int e = tid + 512 * q;
                As[buf][e >> 7][e & 127] = ra[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 2; ++q)
                As[buf][tid & 7][(tid >> 3) + 64 * q] = OPA == OP_C ? cconj(ra[q]) : ra[q];
        }
    };
    auto loadB = [&](int k0) {
        if (OPB == OP_N) {
            const int kk = tid & 7, jj = tid >> 3;
            const int gj = n0 + jj, gk = k0 + kk;
            rb = (gj < N && gk < K) ? Bp[gk + gj * ldb] : make_double2(0.0, 0.0);
        } else {
            const int jj = tid & 63, kk = tid >> 6;
            const int gj = n0 + jj, gk = k0 + kk;
            rb = (gj < N && gk < K) ? Bp[gj + gk * ldb] : make_double2(0.0, 0.0);
        }
    };
    auto storeB = [&](int buf) {
        if (OPB == OP_N)
            Bs[buf][tid & 7][tid >> 3] = rb;
        else
            Bs[buf][tid >> 6][tid & 63] = OPB == OP_C ? cconj(rb) : rb;
    };

    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 3, wn = wave >> 2;
    const int li = lane & 15, lk = lane >> 4;

    double4_t accR[2][2], accI[2][2], acc3[G3 ? 2 : 1][G3 ? 2 : 1];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            accR[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
            accI[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
            if constexpr (G3) acc3[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
        }

    const int nk = (K + WBK - 1) / WBK;
    loadA(0);
    loadB(0);
    storeA(0);
    storeB(0);
    __syncthreads();

    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) {
            loadA((kt + 1) * WBK);
            loadB((kt + 1) * WBK);
        }
#pragma unroll
        for (int kq = 0; kq < WBK / 4; ++kq) {
            const int k = kq * 4 + lk;
            double2 av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = As[buf][k][wm * 32 + i * 16 + li];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Bs[buf][k][wn * 32 + j * 16 + li];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if constexpr (G3) {
                        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].x, accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].y, av[i].y, accI[i][j], 0, 0, 0);
                        acc3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x + bv[j].y, av[i].x + av[i].y,
                                                                         acc3[i][j], 0, 0, 0);
                    } else {
                        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].x, accR[i][j], 0, 0, 0);
                        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bv[j].y, av[i].y, accR[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].y, av[i].x, accI[i][j], 0, 0, 0);
                        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j].x, av[i].y, accI[i][j], 0, 0, 0);
                    }
                }
        }
        if (kt + 1 < nk) {
            storeA(buf ^ 1);
            storeB(buf ^ 1);
        }
        __syncthreads();
    }

    double2* __restrict__ Cp = a.C + b * a.sbc + kz * a.skc;
    const double2* __restrict__ Dp = a.D.p ? a.D.p + b * a.D.sb + kz * a.D.sk : nullptr;
    const double alpha = a.alpha, beta = a.beta;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int gi = m0 + wm * 32 + i * 16 + li;
        if (gi >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gj = n0 + wn * 32 + j * 16 + 4 * q + lk;
                if (gj >= N || (gj >= a.skip_lo && gj < a.skip_hi)) continue;
                double cr = accR[i][j][q], ci = accI[i][j][q];
                if constexpr (G3) {
                    const double p1 = accR[i][j][q], p2 = accI[i][j][q];
                    cr = p1 - p2;
                    ci = (acc3[i][j][q] - p1) - p2;
                }
                double re = alpha * cr, im = alpha * ci;
                if (Dp) {
                    const double2 d = Dp[gi + gj * a.D.ld];
                    re += beta * d.x;
                    im += beta * d.y;
                }
                Cp[gi + gj * a.ldc] = make_double2(re, im);
            }
        }
    }
}

template <int OPA, int OPB, bool G3>
void launch_w8(hipStream_t s, const ZGemmArgs& a) {
    const int tiles_m = (a.M + WBM - 1) / WBM;
    const int tiles_n = (a.N + WBN - 1) / WBN;
    const int tiles = tiles_m * tiles_n;
    const int batch_pad = (a.batch + NXCD - 1) / NXCD * NXCD;
    const int64_t blocks = (int64_t)tiles * batch_pad;
    if (blocks > 0x7fffffff) throw hip_error("zgemm: grid too large", 1);
    hipLaunchKernelGGL((zgemm_mfma_w8_kernel<OPA, OPB, G3>), dim3((unsigned)blocks), dim3(512), 0, s, a, tiles_m,
                       tiles);
    QT_HIP(hipGetLastError());
}

template <int OPA, int OPB, bool G3>
void launch_wide(hipStream_t s, const ZGemmArgs& a) {
    const int tiles_m = (a.M + WBM - 1) / WBM;
    const int tiles_n = (a.N + WBN - 1) / WBN;
    const int tiles = tiles_m * tiles_n;
    const int batch_pad = (a.batch + NXCD - 1) / NXCD * NXCD;
    const int64_t blocks = (int64_t)tiles * batch_pad;
    if (blocks > 0x7fffffff) throw hip_error("zgemm: grid too large", 1);
    hipLaunchKernelGGL((zgemm_mfma_wide_kernel<OPA, OPB, G3>), dim3((unsigned)blocks), dim3(256), 0, s, a,
                       tiles_m, tiles);
    QT_HIP(hipGetLastError());
}

// Tile choice.  QT_GEMM_TILE=64x64, =128x64 or =128x64w8 fixes the kernel;
// otherwise the autotuner below picks per shape (QT_GEMM_TUNE=0: the
// row-waste rule of use_wide alone).  All three kernels accumulate every
// output element in the same order (k ascending in groups of 4, the same
// MFMAs per group and component, the same epilogue), so the choice never
// changes a result.
enum { TILE_AUTO = 0, TILE_NARROW = 1, TILE_WIDE = 2, TILE_W8 = 3 };
int tile_forced() {
    static const int v = [] {
        const char* e = getenv("QT_GEMM_TILE");
        if (e && std::string(e) == "64x64") return (int)TILE_NARROW;
        if (e && std::string(e) == "128x64") return (int)TILE_WIDE;
        if (e && std::string(e) == "128x64w8") return (int)TILE_W8;
        return (int)TILE_AUTO;
    }();
    return v;
}
bool tune_enabled() {
    static const bool v = [] {
        const char* e = getenv("QT_GEMM_TUNE");
        return !(e && std::string(e) == "0");
    }();
    return v;
}
bool tune_log() {
    static const bool v = getenv("QT_GEMM_TUNE_LOG") != nullptr;
    return v;
}

// 128-row tiles unless they leave more rows idle than 64-row tiles do
// (M = 192, 320, 448, ...: the half-empty last tile costs more than the
// 64x64 kernel's lower peak)
bool use_wide(int M) {
    const int waste_wide = (M + WBM - 1) / WBM * WBM - M;
    const int waste_narrow = (M + BM - 1) / BM * BM - M;
    return M >= WBM && waste_wide <= waste_narrow;
}

template <int OPA, int OPB, bool G3>
void launch_narrow(hipStream_t s, const ZGemmArgs& a) {
    const int tiles_m = (a.M + BM - 1) / BM;
    const int tiles_n = (a.N + BN - 1) / BN;
    const int tiles = tiles_m * tiles_n;
    const int batch_pad = (a.batch + NXCD - 1) / NXCD * NXCD;
    const int64_t blocks = (int64_t)tiles * batch_pad;
    if (blocks > 0x7fffffff) throw hip_error("zgemm: grid too large", 1);
    hipLaunchKernelGGL((zgemm_mfma_kernel<OPA, OPB, G3>), dim3((unsigned)blocks), dim3(256), 0, s, a,
                       tiles_m, tiles, batch_pad);
    QT_HIP(hipGetLastError());
}

// Autotuner: the first product of each shape (op pair, M, N, K, batch) runs
// the three kernels twice into a scratch output (the operands and the addend are
// only read; at most 512 MiB of scratch, fewer members where the full batch
// needs more) after the work already queued on the stream, and the faster
// kernel is kept for the shape for the life of the process.  Where the
// scratch cannot be allocated the row-waste rule decides.
struct TuneState {
    std::mutex mu;
    std::map<std::tuple<int, int, int, int, int>, int> tile;
};
TuneState& tune_state() {
    static TuneState t;
    return t;
}

template <int OPA, int OPB, bool G3>
void launch_tile(hipStream_t s, const ZGemmArgs& a, int tile) {
    if (tile == TILE_W8)
        launch_w8<OPA, OPB, G3>(s, a);
    else if (tile == TILE_WIDE)
        launch_wide<OPA, OPB, G3>(s, a);
    else
        launch_narrow<OPA, OPB, G3>(s, a);
}

template <int OPA, int OPB, bool G3>
int tuned_tile(hipStream_t s, const ZGemmArgs& a) {
    const auto key = std::make_tuple((OPA * 3 + OPB) * 2 + (G3 ? 1 : 0), a.M, a.N, a.K, a.batch);
    TuneState& ts = tune_state();
    {
        std::lock_guard<std::mutex> lk(ts.mu);
        auto it = ts.tile.find(key);
        if (it != ts.tile.end()) return it->second;
    }
    const size_t per = sizeof(double2) * (size_t)a.M * (size_t)a.N;
    const size_t cap = (size_t)512 << 20;
    const int members = (int)std::max<size_t>(1, std::min<size_t>((size_t)a.batch, cap / std::max<size_t>(per, 1)));
    ZGemmArgs t = a;
    t.batch = members;
    t.ldc = a.M;
    t.sbc = (int64_t)a.M * a.N;
    t.skc = 0;
    t.skip_lo = t.skip_hi = 0;
    void* scratch = nullptr;
    hipEvent_t ev[3];
    float ms[3] = {1e30f, 1e30f, 1e30f};  // 64x64, 128x64, 128x64 with 8 waves
    if (hipMalloc(&scratch, per * (size_t)members) != hipSuccess) {
        (void)hipGetLastError();  // no room for the scratch output: the row-waste rule, kept
        const int tile = use_wide(a.M) ? TILE_WIDE : TILE_NARROW;
        std::lock_guard<std::mutex> lk(ts.mu);
        ts.tile[key] = tile;
        return tile;
    }
    for (auto& e : ev) QT_HIP(hipEventCreate(&e));
    t.C = (double2*)scratch;
    for (int rep = 0; rep < 2; ++rep)
        for (int v = 0; v < 3; ++v) {
            QT_HIP(hipEventRecord(ev[0], s));
            launch_tile<OPA, OPB, G3>(s, t, v + 1);
            QT_HIP(hipEventRecord(ev[1], s));
            QT_HIP(hipEventSynchronize(ev[1]));
            float m = 0.0f;
            QT_HIP(hipEventElapsedTime(&m, ev[0], ev[1]));
            ms[v] = std::min(ms[v], m);
        }
    for (auto& e : ev) QT_HIP(hipEventDestroy(e));
    QT_HIP(hipFree(scratch));
    int best = 0;
    for (int v = 1; v < 3; ++v)
        if (ms[v] < ms[best]) best = v;
    const int tile = best + 1;
    if (tune_log())
        fprintf(stderr,
                "qt gemm tune: ops %d%s M %d N %d K %d batch %d (%d timed): 64x64 %.4f ms, 128x64 %.4f ms, "
                "128x64w8 %.4f ms -> %s\n",
                OPA * 3 + OPB, G3 ? " (3m)" : "", a.M, a.N, a.K, a.batch, members, ms[0], ms[1], ms[2],
                tile == TILE_W8 ? "128x64w8" : (tile == TILE_WIDE ? "128x64" : "64x64"));
    std::lock_guard<std::mutex> lk(ts.mu);
    ts.tile[key] = tile;
    return tile;
}

template <int OPA, int OPB, bool G3>
void launch_g(hipStream_t s, const ZGemmArgs& a) {
    const int forced = tile_forced();
    int tile;
    if (forced != TILE_AUTO)
        tile = forced;
    else if (tune_enabled())
        tile = tuned_tile<OPA, OPB, G3>(s, a);
    else
        tile = use_wide(a.M) ? TILE_WIDE : TILE_NARROW;
    launch_tile<OPA, OPB, G3>(s, a, tile);
}

template <int OPA, int OPB>
void launch(hipStream_t s, const ZGemmArgs& a) {
    if (a.prec == 3)
        launch_g<OPA, OPB, true>(s, a);
    else
        launch_g<OPA, OPB, false>(s, a);
}

}  // namespace

void zgemm(hipStream_t s, int opa, int opb, const ZGemmArgs& a, Timer* t) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return;
    if (t && t->enabled) t->begin(s, 0);
    const int key = opa * 3 + opb;
    if ((a.prec == 6 || a.prec == 7) && zgemm_i8_fits(a)) {
        zgemm_i8(s, opa, opb, a, a.prec);
        if (t && t->enabled)
            t->end(s, 0, 8.0 * a.M * a.N * (double)a.K * (a.flop_batch >= 0 ? a.flop_batch : a.batch));
        return;
    }
    switch (key) {
        case 0: launch<OP_N, OP_N>(s, a); break;
        case 2: launch<OP_N, OP_C>(s, a); break;
        case 6: launch<OP_C, OP_N>(s, a); break;
        case 8: launch<OP_C, OP_C>(s, a); break;
        case 1: launch<OP_N, OP_T>(s, a); break;
        case 3: launch<OP_T, OP_N>(s, a); break;
        default: throw hip_error("zgemm: unsupported op pair", 6);
    }
    if (t && t->enabled)
        t->end(s, 0, 8.0 * a.M * a.N * (double)a.K * (a.flop_batch >= 0 ? a.flop_batch : a.batch));
}

}  // namespace qt
