// qt_gw.hip -- GW polarisation and self-energy assembly on the device
// (gw_mod.f90:444-519 P^<,>,R;  :641-743 Sigma^<,>,R;  :747-749 mixing).
//
// Arrays are element-major, the by_blocks view of gw_mod.f90:246-255:
// series (ij, ix, ik) of an array starts at ((ij*nx + ix)*nk + ik)*nen,
// ij = row + col*nm.  Only entries with |row - col| <= ndiag are assembled;
// the others are zero, as after the reference's band zeroing.
//
// The reference evaluates every correlation / convolution with its own
// three DFTs.  Here the k-sums are taken in the frequency domain: one
// workgroup owns one element pair ij/ji (Sigma; P at one k point) or one
// element (P over several k points) at one slab, transforms each input series once per term with the LDS FFT of
// qt_fft.h (length L = power of two >= 2 nen - 1, which gives the same
// linear products), accumulates the products over k in registers and
// inverse-transforms once per output series:
//   corr(X, Y) = F^-1[ F(X) F(Yf) ],  Yf[(-q) mod L] = Y[q]
//   conv(X, Y) = F^-1[ F(X) F(Yr) ],  Yr[l mod L] = Y[l + n/2]
// G enters as i Im G (gw_mod.f90:444-446).
#include <vector>

#include "qt_fft.h"
#include "qt_internal.h"

namespace qt {

namespace {

using fft::cadd;
using fft::cmul;
using fft::pad;

constexpr double kPi = 3.14159265359;  // the reference's constant (static.f90:15)

struct GwDims {
    int nen, nky, nkz, nk, nx, nm;
};

// k' paired with (k, q) (gw_mod.f90:478-489, 656-667; 1-based there)
__device__ __forceinline__ int kd_index(int ik, int iq, int nky, int nkz) {
    const int ikz = ik % nkz + 1, iky = ik / nkz + 1;
    const int iqz = iq % nkz + 1, iqy = iq / nkz + 1;
    int ikzd = ikz - iqz + nkz / 2;
    int ikyd = iky - iqy + nky / 2;
    if (ikzd < 1) ikzd += nkz;
    if (ikzd > nkz) ikzd -= nkz;
    if (ikyd < 1) ikyd += nky;
    if (ikyd > nky) ikyd -= nky;
    if (nky == 1) ikyd = 1;
    if (nkz == 1) ikzd = 1;
    return (ikzd - 1) + (ikyd - 1) * nkz;
}

__device__ __forceinline__ const double2* series(const double2* A, const GwDims& d, int ij, int ix, int ik) {
    return A + (((int64_t)ij * d.nx + ix) * d.nk + ik) * (int64_t)d.nen;
}
__device__ __forceinline__ double2* series(double2* A, const GwDims& d, int ij, int ix, int ik) {
    return A + (((int64_t)ij * d.nx + ix) * d.nk + ik) * (int64_t)d.nen;
}

// the spectrum samples i = tid + q*NT (i < L) of one thread
template <int L, int NT>
struct Spec {
    static constexpr int PER = L >= NT ? L / NT : 1;
    double2 v[PER];
    __device__ __forceinline__ static bool ok(int i) { return L >= NT || i < L; }
    __device__ __forceinline__ void zero() {
#pragma unroll
        for (int q = 0; q < PER; ++q) v[q] = make_double2(0.0, 0.0);
    }
};

// The twiddle table is copied into LDS behind the transform buffer (L
// entries; 64 KB at L = 4096, which the one-workgroup-per-CU occupancy of
// these kernels leaves free), so the passes read it with LDS latency.
#ifndef QT_GW_TW_LDS
#define QT_GW_TW_LDS 1
#endif
template <int L>
constexpr size_t gw_lds_elems() {
    return (size_t)fft::padded_len(L) + (QT_GW_TW_LDS ? (size_t)L : 0);
}
template <int L, int NT>
__device__ __forceinline__ const double2* stage_twiddles(double2* s, const double2* __restrict__ tw) {
#if QT_GW_TW_LDS
    double2* t = s + fft::padded_len(L);
    for (int i = threadIdx.x; i < L; i += NT) t[i] = tw[i];
    __syncthreads();
    return t;
#else
    return tw;
#endif
}

// radix of the FFT passes (8: register budget of the assembly kernels)
#ifndef QT_GW_RADIX
#define QT_GW_RADIX 8
#endif
constexpr int RADIX = QT_GW_RADIX;

// inverse of acc, scaled by c / L, lags written to z (corr) or samples (conv)
// (the inverse reads acc from registers and writes z directly where the
// transform allows it)
template <int L, int NT>
__device__ __forceinline__ void finish(double2* s, const double2* tw, const Spec<L, NT>& acc,
                                       double2 c, int n, bool corr, double2* z) {
    const double inv = 1.0 / (double)L;
    const int h = n / 2;
    fft::transform_io<L, NT, true, RADIX>(
        s, tw, [&](int, int q) { return acc.v[q]; },
        [&](int m, int, double2 v) {
            // sample m is lag m (corr: output i = m + n/2 mod L) or output m (conv)
            const int i = corr ? (m + h) & (L - 1) : m;
            if (i < n) z[i] = make_double2((c.x * v.x - c.y * v.y) * inv, (c.x * v.y + c.y * v.x) * inv);
        });
}

// ---------------------------------------------------------------- packed transforms
// Every G operand enters as i Im G, i.e. i times a real series y.  Two such
// series share one complex transform: with Z = F(y1 + i y2),
//   F(i y1)[k] = i (Z[k] + conj Z[-k]) / 2,   F(i y2)[k] = (Z[k] - conj Z[-k]) / 2,
// and the flipped operand of a correlation needs no transform of its own,
// since F(Yf)[k] = F(Y)[-k].
// s <- F(Im y1 + i Im y2)
template <int L, int NT>
__device__ __forceinline__ void fwd_pack(double2* s, const double2* tw, const double2* y1, const double2* y2,
                                         int n) {
    fft::transform_load<L, NT, false, RADIX>(s, tw, [&](int i, int) {
        return i < n ? make_double2(y1[i].y, y2 ? y2[i].y : 0.0) : make_double2(0.0, 0.0);
    });
}
// Z[k] and Z[-k] of the thread's sample q
template <int L, int NT>
__device__ __forceinline__ void zpair(const double2* s, int q, double2& zk, double2& zm) {
    const int i = threadIdx.x + q * NT;
    zk = s[pad(i)];
    zm = s[pad((L - i) & (L - 1))];
}
// F(i y1) and F(i y2) at k from zk = Z[k], zm = Z[-k] (swap them for -k)
__device__ __forceinline__ double2 spec1(double2 zk, double2 zm) {
    return make_double2(0.5 * (zm.y - zk.y), 0.5 * (zk.x + zm.x));
}
__device__ __forceinline__ double2 spec2(double2 zk, double2 zm) {
    return make_double2(0.5 * (zk.x - zm.x), 0.5 * (zk.y + zm.y));
}

// ---------------------------------------------------------------- P
// Two passes over k per q, each with two packed transforms per k (the
// reference takes eight forward transforms per k):
//   1. (G^<_ji, G^>_ji) at k' and (G^<_ij, G^>_ij) at k give P^< and P^>,
//      correlations of real series that share one inverse transform (real
//      and imaginary parts);
//   2. (G^<_ij at k, G^<_ji at k') and (G^R_ij at k, k') give P^R.
template <int L, int NT>
__global__ __launch_bounds__(NT) void gw_pol_kernel(GwDims d, const int* __restrict__ elems, int nelem,
                                                    const double2* __restrict__ Gr,
                                                    const double2* __restrict__ Gl,
                                                    const double2* __restrict__ Gg, double2* Pr,
                                                    double2* Pl, double2* Pg, double2 dE,
                                                    const double2* __restrict__ tw) {
    using S = Spec<L, NT>;
    constexpr int PER = S::PER;
    extern __shared__ double2 s[];
    const int w = blockIdx.x;
    if (w >= nelem * d.nx) return;
    tw = stage_twiddles<L, NT>(s, tw);
    const int e = w / d.nx, ix = w % d.nx;
    const int ij = elems[e];
    const int row = ij % d.nm, col = ij / d.nm;
    const int ji = col + row * d.nm;
    const int n = d.nen;
    for (int iq = 0; iq < d.nk; ++iq) {
        S accl, accg;
        accl.zero();
        accg.zero();
        for (int ik = 0; ik < d.nk; ++ik) {
            const int kd = kd_index(ik, iq, d.nky, d.nkz);
            // F(i Im G^<_ji(k'))(-k), F(i Im G^>_ji(k'))(-k)
            fwd_pack<L, NT>(s, tw, series(Gl, d, ji, ix, kd), series(Gg, d, ji, ix, kd), n);
            double2 bl[PER], bg[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (!S::ok(threadIdx.x + q * NT)) continue;
                double2 zk, zm;
                zpair<L, NT>(s, q, zk, zm);
                bl[q] = spec1(zm, zk);
                bg[q] = spec2(zm, zk);
            }
            // P^< += F(i Im G^<_ij(k)) F(i Im G^>_ji(k'))(-k), P^> likewise
            fwd_pack<L, NT>(s, tw, series(Gl, d, ij, ix, ik), series(Gg, d, ij, ix, ik), n);
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (!S::ok(threadIdx.x + q * NT)) continue;
                double2 zk, zm;
                zpair<L, NT>(s, q, zk, zm);
                accl.v[q] = cadd(accl.v[q], cmul(spec1(zk, zm), bg[q]));
                accg.v[q] = cadd(accg.v[q], cmul(spec2(zk, zm), bl[q]));
            }
        }
        // P^< (real part) and P^> (imaginary part) from one inverse transform;
        // sample m is lag m, i.e. output i = m + n/2 (mod L)
        {
            const double inv = 1.0 / (double)L;
            const int h = n / 2;
            double2* zl = series(Pl, d, ij, ix, iq);
            double2* zg = series(Pg, d, ij, ix, iq);
            fft::transform_io<L, NT, true, RADIX>(
                s, tw,
                [&](int, int q) {
                    return make_double2(accl.v[q].x - accg.v[q].y, accl.v[q].y + accg.v[q].x);
                },
                [&](int m, int, double2 v) {
                    const int i = (m + h) & (L - 1);
                    if (i < n) {
                        zl[i] = make_double2(dE.x * v.x * inv, dE.y * v.x * inv);
                        zg[i] = make_double2(dE.x * v.y * inv, dE.y * v.y * inv);
                    }
                });
        }
        // P^R += F(i Im G^<_ij(k)) F(-i Im G^R_ij(k'))(-k) + F(i Im G^R_ij(k)) F(i Im G^<_ji(k'))(-k)
        S accr;
        accr.zero();
        for (int ik = 0; ik < d.nk; ++ik) {
            const int kd = kd_index(ik, iq, d.nky, d.nkz);
            fwd_pack<L, NT>(s, tw, series(Gl, d, ij, ix, ik), series(Gl, d, ji, ix, kd), n);
            double2 al[PER], bl[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (!S::ok(threadIdx.x + q * NT)) continue;
                double2 zk, zm;
                zpair<L, NT>(s, q, zk, zm);
                al[q] = spec1(zk, zm);
                bl[q] = spec2(zm, zk);
            }
            const bool same = kd == ik;
            fwd_pack<L, NT>(s, tw, series(Gr, d, ij, ix, ik), same ? nullptr : series(Gr, d, ij, ix, kd), n);
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (!S::ok(threadIdx.x + q * NT)) continue;
                double2 zk, zm;
                zpair<L, NT>(s, q, zk, zm);
                const double2 rk = spec1(zk, zm);
                const double2 rd = same ? spec1(zm, zk) : spec2(zm, zk);
                const double2 t1 = cmul(al[q], rd), t2 = cmul(rk, bl[q]);
                accr.v[q] = make_double2(accr.v[q].x - t1.x + t2.x, accr.v[q].y - t1.y + t2.y);
            }
        }
        finish<L, NT>(s, tw, accr, dE, n, true, series(Pr, d, ij, ix, iq));
    }
}

// -x* : the spectrum at -k of i times a real series, from its value x at k
__device__ __forceinline__ double2 at_minus(double2 x) { return make_double2(-x.x, x.y); }

// P of one element pair (ij, ji), row <= col, at one k point (nk == 1, so
// k' = k): the six series of the pair in three packed transforms,
//   (G^<_ij, G^>_ij), (G^<_ji, G^>_ji), (G^R_ij, G^R_ji),
// and four inverse transforms (P^< + i P^> and P^R of each element), where
// the element kernel above takes six transforms per element.  The spectrum
// of i y at -k is -conj of its value at k (y real), bit for bit the value
// the unpacking gives at -k, so a thread needs only its own samples k.
// A diagonal pair (ij == ji) takes (G^<, G^>) and (G^R, 0) and two inverses.
template <int L, int NT>
__global__ __launch_bounds__(NT) void gw_pol_pair_kernel(GwDims d, const int* __restrict__ pairs, int npair,
                                                         const double2* __restrict__ Gr,
                                                         const double2* __restrict__ Gl,
                                                         const double2* __restrict__ Gg, double2* Pr,
                                                         double2* Pl, double2* Pg, double2 dE,
                                                         const double2* __restrict__ tw) {
    using S = Spec<L, NT>;
    constexpr int PER = S::PER;
    extern __shared__ double2 s[];
    const int w = blockIdx.x;
    if (w >= npair * d.nx) return;
    tw = stage_twiddles<L, NT>(s, tw);
    const int p = w / d.nx, ix = w % d.nx;
    const int ij = pairs[p];
    const int row = ij % d.nm, col = ij / d.nm;
    const int ji = col + row * d.nm;
    const bool diag = ij == ji;
    const int n = d.nen;
    // P^< (real part) and P^> (imaginary part) of c from one inverse transform;
    // sample m is lag m, i.e. output i = m + n/2 (mod L)
    auto lesser_greater = [&](const S& c, int e) {
        const double inv = 1.0 / (double)L;
        const int h = n / 2;
        double2* zl = series(Pl, d, e, ix, 0);
        double2* zg = series(Pg, d, e, ix, 0);
        fft::transform_io<L, NT, true, RADIX>(
            s, tw, [&](int, int q) { return c.v[q]; },
            [&](int m, int, double2 v) {
                const int i = (m + h) & (L - 1);
                if (i < n) {
                    zl[i] = make_double2(dE.x * v.x * inv, dE.y * v.x * inv);
                    zg[i] = make_double2(dE.x * v.y * inv, dE.y * v.y * inv);
                }
            });
    };
    // spectra at k of i Im G^<_ij, i Im G^<_ji
    double2 lij[PER], lji[PER];
    {
        S cij, cji;  // P^< + i P^> of ij and ji
        double2 gij[PER];
        fwd_pack<L, NT>(s, tw, series(Gl, d, ij, ix, 0), series(Gg, d, ij, ix, 0), n);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (!S::ok(threadIdx.x + q * NT)) continue;
            double2 zk, zm;
            zpair<L, NT>(s, q, zk, zm);
            lij[q] = spec1(zk, zm);
            gij[q] = spec2(zk, zm);
        }
        if (!diag) fwd_pack<L, NT>(s, tw, series(Gl, d, ji, ix, 0), series(Gg, d, ji, ix, 0), n);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (!S::ok(threadIdx.x + q * NT)) continue;
            double2 l = lij[q], g = gij[q];
            if (!diag) {
                double2 zk, zm;
                zpair<L, NT>(s, q, zk, zm);
                l = spec1(zk, zm);
                g = spec2(zk, zm);
            }
            lji[q] = l;
            // P^<_ij ~ G^<_ij(k) G^>_ji(-k), P^>_ij ~ G^>_ij(k) G^<_ji(-k)
            double2 a = cmul(lij[q], at_minus(g)), b = cmul(gij[q], at_minus(l));
            cij.v[q] = make_double2(a.x - b.y, a.y + b.x);
            // P^<_ji ~ G^<_ji(k) G^>_ij(-k), P^>_ji ~ G^>_ji(k) G^<_ij(-k)
            a = cmul(l, at_minus(gij[q]));
            b = cmul(g, at_minus(lij[q]));
            cji.v[q] = make_double2(a.x - b.y, a.y + b.x);
        }
        lesser_greater(cij, ij);
        if (!diag) lesser_greater(cji, ji);
    }
    // P^R_ij ~ -G^<_ij(k) G^R_ij(-k) + G^R_ij(k) G^<_ji(-k), and ji likewise
    S rij, rji;
    fwd_pack<L, NT>(s, tw, series(Gr, d, ij, ix, 0), diag ? nullptr : series(Gr, d, ji, ix, 0), n);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if (!S::ok(threadIdx.x + q * NT)) continue;
        double2 zk, zm;
        zpair<L, NT>(s, q, zk, zm);
        const double2 r1 = spec1(zk, zm);
        double2 t1 = cmul(lij[q], at_minus(r1)), t2 = cmul(r1, at_minus(lji[q]));
        rij.v[q] = make_double2(t2.x - t1.x, t2.y - t1.y);
        const double2 r2 = spec2(zk, zm);
        t1 = cmul(lji[q], at_minus(r2));
        t2 = cmul(r2, at_minus(lij[q]));
        rji.v[q] = make_double2(t2.x - t1.x, t2.y - t1.y);
    }
    finish<L, NT>(s, tw, rij, dE, n, true, series(Pr, d, ij, ix, 0));
    if (!diag) finish<L, NT>(s, tw, rji, dE, n, true, series(Pr, d, ji, ix, 0));
}

// ---------------------------------------------------------------- Sigma
// Sigma of the element pair (ij, ji), row <= col, at slab ix and point k.
// The transpose fold (S + S^T)/2 (gw_mod.f90:732-743) is linear and so is
// the Sigma^R rule before it (:694-698), so the spectra of both elements are
// summed and each output takes one inverse transform, written to ij and ji.
// The q-sums of each element in two passes:
//   1. i Im G^< and i Im G^R at k' share one packed transform; with the
//      transforms of W^< and W^R they give Sigma^< and the three terms of
//      Sigma^R (W^< + W^R enters by linearity);
//   2. i Im G^> at k' and W^> give Sigma^>.
// The reference takes eight forward transforms per q and element and three
// inverses per element; these take four and a half per q and element (the
// i Im G^> of both elements share one) and three inverses per pair.
template <int L, int NT>
__device__ __forceinline__ void sigma_pair(double2* s, const double2* tw, const GwDims& d, int ij, int ji,
                                           int ix, int ik, const double2* Gr, const double2* Gl,
                                           const double2* Gg, const double2* Wr, const double2* Wl,
                                           const double2* Wg, double2 dE, double2* Sr, double2* Sl,
                                           double2* Sg) {
    using Sp = Spec<L, NT>;
    constexpr int PER = Sp::PER;
    constexpr int EXT = (L / 2 + NT - 1) / NT;
    const int n = d.nen;
    const bool diag = ij == ji;
    const int nel = diag ? 1 : 2;
    // dE/L (one element) or dE/(2L) (the mean of two)
    const double inv = (diag ? 1.0 : 0.5) / (double)L;
    // the scaled inverse of acc at the samples this thread owns
    // (the inverse takes acc from registers and hands back the thread's own
    // samples, where the transform allows it)
    auto extract = [&](const Sp& acc, double2* t) {
        fft::transform_io<L, NT, true, RADIX>(
            s, tw, [&](int, int q) { return acc.v[q]; },
            [&](int i, int q, double2 v) {
                if (q >= EXT) return;
                if (i >= n) v = make_double2(0.0, 0.0);
                t[q] = make_double2((dE.x * v.x - dE.y * v.y) * inv, (dE.x * v.y + dE.y * v.x) * inv);
            });
    };
    auto store = [&](double2* S, const double2* t) {
        double2* a = series(S, d, ij, ix, ik);
        double2* b = series(S, d, ji, ix, ik);
#pragma unroll
        for (int q = 0; q < EXT; ++q) {
            const int i = threadIdx.x + q * NT;
            if (i >= n) continue;
            a[i] = t[q];
            if (!diag) b[i] = t[q];
        }
    };
    // F(Yr), Yr[l mod L] = Y[l + n/2]  (conv)
    auto fwd_w = [&](const double2* w) {
        const int h = n / 2;
        fft::transform_load<L, NT, false, RADIX>(s, tw, [&](int i, int) {
            int src = -1;
            if (i < n - h) src = i + h;
            else if (i >= L - h) src = i - (L - h);
            return src >= 0 ? w[src] : make_double2(0.0, 0.0);
        });
    };
    double2 t[EXT], rre[EXT];
    double dim[EXT];
    {
        Sp accl, accr;
        accl.zero();
        accr.zero();
        for (int m = 0; m < nel; ++m) {
            const int e = m == 0 ? ij : ji;
            for (int iq = 0; iq < d.nk; ++iq) {
                const int kd = kd_index(ik, iq, d.nky, d.nkz);
                fwd_pack<L, NT>(s, tw, series(Gl, d, e, ix, kd), series(Gr, d, e, ix, kd), n);
                double2 gl[PER], gr[PER];
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    if (!Sp::ok(threadIdx.x + q * NT)) continue;
                    double2 zk, zm;
                    zpair<L, NT>(s, q, zk, zm);
                    gl[q] = spec1(zk, zm);
                    gr[q] = spec2(zk, zm);
                }
                // Sigma^< += F(G^<) F(W^<r), Sigma^R += F(G^R) F(W^<r)
                fwd_w(series(Wl, d, e, ix, iq));
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int i = threadIdx.x + q * NT;
                    if (!Sp::ok(i)) continue;
                    const double2 wv = s[pad(i)];
                    accl.v[q] = cadd(accl.v[q], cmul(gl[q], wv));
                    accr.v[q] = cadd(accr.v[q], cmul(gr[q], wv));
                }
                // Sigma^R += (F(G^<) + F(G^R)) F(W^Rr)
                fwd_w(series(Wr, d, e, ix, iq));
#pragma unroll
                for (int q = 0; q < PER; ++q) {
                    const int i = threadIdx.x + q * NT;
                    if (!Sp::ok(i)) continue;
                    accr.v[q] = cadd(accr.v[q], cmul(cadd(gl[q], gr[q]), s[pad(i)]));
                }
            }
        }
        extract(accl, t);
#pragma unroll
        for (int q = 0; q < EXT; ++q) dim[q] = -t[q].y;
        store(Sl, t);
        extract(accr, t);
#pragma unroll
        for (int q = 0; q < EXT; ++q) rre[q] = t[q];
    }
    {
        // i Im G^> of both elements in one packed transform
        Sp accg;
        accg.zero();
        for (int iq = 0; iq < d.nk; ++iq) {
            const int kd = kd_index(ik, iq, d.nky, d.nkz);
            fwd_pack<L, NT>(s, tw, series(Gg, d, ij, ix, kd), diag ? nullptr : series(Gg, d, ji, ix, kd), n);
            double2 g1[PER], g2[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                if (!Sp::ok(threadIdx.x + q * NT)) continue;
                double2 zk, zm;
                zpair<L, NT>(s, q, zk, zm);
                g1[q] = spec1(zk, zm);
                g2[q] = spec2(zk, zm);
            }
            fwd_w(series(Wg, d, ij, ix, iq));
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int i = threadIdx.x + q * NT;
                if (!Sp::ok(i)) continue;
                accg.v[q] = cadd(accg.v[q], cmul(g1[q], s[pad(i)]));
            }
            if (diag) continue;
            fwd_w(series(Wg, d, ji, ix, iq));
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int i = threadIdx.x + q * NT;
                if (!Sp::ok(i)) continue;
                accg.v[q] = cadd(accg.v[q], cmul(g2[q], s[pad(i)]));
            }
        }
        extract(accg, t);
    }
    store(Sg, t);
    // Sigma^R = Re Sigma^R + i Im(Sigma^> - Sigma^<) / 2  (gw_mod.f90:694-698)
#pragma unroll
    for (int q = 0; q < EXT; ++q) t[q] = make_double2(rre[q].x, 0.5 * (dim[q] + t[q].y));
    store(Sr, t);
}

// One workgroup per element pair (ij, ji) with row <= col and slab ix.
template <int L, int NT>
__global__ __launch_bounds__(NT) void gw_sigma_kernel(GwDims d, const int* __restrict__ pairs, int npair,
                                                      const double2* __restrict__ Gr,
                                                      const double2* __restrict__ Gl,
                                                      const double2* __restrict__ Gg,
                                                      const double2* __restrict__ Wr,
                                                      const double2* __restrict__ Wl,
                                                      const double2* __restrict__ Wg, double2* Sr,
                                                      double2* Sl, double2* Sg, double2 dE,
                                                      const double2* __restrict__ tw) {
    extern __shared__ double2 s[];
    const int w = blockIdx.x;
    if (w >= npair * d.nx) return;
    tw = stage_twiddles<L, NT>(s, tw);
    const int p = w / d.nx, ix = w % d.nx;
    const int ij = pairs[p];
    const int row = ij % d.nm, col = ij / d.nm;
    const int ji = col + row * d.nm;
    for (int ik = 0; ik < d.nk; ++ik)
        sigma_pair<L, NT>(s, tw, d, ij, ji, ix, ik, Gr, Gl, Gg, Wr, Wl, Wg, dE, Sr, Sl, Sg);
}

// e-photon self-energy, monochromatic field: the two energy-shifted blocks
// summed (S = G(E - hw) + G(E + hw), one contiguous stream shifted by
// 2 nop energies), then Sigma^R = i Im(Sigma^> - Sigma^<) / 2.
__global__ void eph_sum_kernel(int64_t n, int64_t shift, const double2* __restrict__ G, double2* S) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = G[t], b = G[t + shift];
        S[t] = make_double2(a.x + b.x, a.y + b.y);
    }
}

__global__ void eph_retarded_kernel(int64_t n, const double2* __restrict__ Sl, const double2* __restrict__ Sg,
                                    double2* Sr) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        Sr[t] = make_double2(0.0, 0.5 * (Sg[t].y - Sl[t].y));
}

// ---------------------------------------------------------------- currents, collision integrals
// U, L: the cell-superdiagonal blocks (j, j+1) and the cell-subdiagonal
// blocks (j+1, j) of each slab's diagonal block H (nm x nm, nm = nb ns),
// every other entry zero
__global__ void cell_bands_kernel(int nm, int nb, int nx, const double2* __restrict__ H, double2* U,
                                  double2* L) {
    const int64_t nn = (int64_t)nm * nm, total = nn * nx;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = t % nn;
        const int r = (int)(e % nm), c = (int)(e / nm);
        const int cr = r / nb, cc = c / nb;
        const double2 h = H[t];
        U[t] = cc == cr + 1 ? h : make_double2(0.0, 0.0);
        L[t] = cr == cc + 1 ? h : make_double2(0.0, 0.0);
    }
}

// Orbital-resolved current of every cell bond c = ix ns + j (c -> c + 1) of
// member m = ix + nx e, nb x nb, into jd (nb, nb, nx ns, cnt):
//   j < ns - 1:   Re (U G^<)_{jj} - Re (L G^<)_{j+1,j+1}
//                 (= Re H_{j,j+1} G^<_{j+1,j} - Re H_{j+1,j} G^<_{j,j+1});
//   j = ns - 1:   sum over the cells k of the diagonal blocks (k, k) of the
//                 slab bond's current matrix cur (nm, nm, nx, cnt), zero in
//                 the last slab.
__global__ void cell_current_kernel(int nb, int ns, int nx, int cnt, const double2* __restrict__ C1,
                                    const double2* __restrict__ C2, const double* __restrict__ cur, double* jd) {
    const int nm = nb * ns;
    const int64_t nn = (int64_t)nm * nm, bb = (int64_t)nb * nb;
    const int64_t total = bb * ns * nx * (int64_t)cnt;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(t % nb), b = (int)((t / nb) % nb);
        const int64_t cm = t / bb;  // cell bond of member: c + nx ns e
        const int j = (int)(cm % ns);
        const int64_t m = cm / ns;  // member ix + nx e
        const int ix = (int)(m % nx);
        double v = 0.0;
        if (j < ns - 1) {
            v = C1[m * nn + (j * nb + a) + (int64_t)(j * nb + b) * nm].x -
                C2[m * nn + ((j + 1) * nb + a) + (int64_t)((j + 1) * nb + b) * nm].x;
        } else if (ix < nx - 1) {
            for (int k = 0; k < ns; ++k) v += cur[m * nn + (k * nb + a) + (int64_t)(k * nb + b) * nm];
        }
        jd[t] = v;
    }
}

// acc[i] += sum_e X[i + n e] w[e], e = 0 .. cnt-1 in order
__global__ void energy_sum_kernel(int64_t n, int cnt, const double* __restrict__ X, const double* __restrict__ w,
                                  double* acc) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = acc[i];
        for (int e = 0; e < cnt; ++e) s += X[i + n * e] * w[e];
        acc[i] = s;
    }
}

__global__ void gw_mix_kernel(int64_t n, double alpha, double2* S, const double2* __restrict__ Snew) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = S[t], b = Snew[t];
        S[t] = make_double2(a.x + alpha * (b.x - a.x), a.y + alpha * (b.y - a.y));
    }
}

// one butterfly per thread where L allows (512 radix-8 threads at L = 4096),
// at most QT_GW_NTMAX threads
#ifndef QT_GW_NTMAX
#define QT_GW_NTMAX 512
#endif
template <int L>
struct GwCfg {
    static constexpr int NT = L / RADIX >= QT_GW_NTMAX ? QT_GW_NTMAX : (L / RADIX >= 64 ? L / RADIX : 64);
};

// elems: every in-band element (any nk), or, with pairs == true, the
// in-band elements with row <= col (nk == 1)
template <int L>
void launch_pol(hipStream_t st, const GwDims& d, const int* elems, int nelem, bool pairs, const double2* Gr,
                const double2* Gl, const double2* Gg, double2* Pr, double2* Pl, double2* Pg, double2 dE) {
    constexpr int NT = GwCfg<L>::NT;
    const size_t lds = sizeof(double2) * gw_lds_elems<L>();
    static bool attr = false;
    if (!attr) {
        QT_HIP(hipFuncSetAttribute((const void*)gw_pol_kernel<L, NT>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        QT_HIP(hipFuncSetAttribute((const void*)gw_pol_pair_kernel<L, NT>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    const int64_t blocks = (int64_t)nelem * d.nx;
    if (blocks > 0x7fffffff) throw hip_error("gw polarization: grid too large", 1);
    if (pairs) {
        if (d.nk != 1) throw hip_error("gw polarization: element pairs need nk == 1", 1);
        hipLaunchKernelGGL((gw_pol_pair_kernel<L, NT>), dim3((unsigned)blocks), dim3(NT), lds, st, d, elems,
                           nelem, Gr, Gl, Gg, Pr, Pl, Pg, dE, fft_twiddles(L));
    } else {
        hipLaunchKernelGGL((gw_pol_kernel<L, NT>), dim3((unsigned)blocks), dim3(NT), lds, st, d, elems, nelem,
                           Gr, Gl, Gg, Pr, Pl, Pg, dE, fft_twiddles(L));
    }
    QT_HIP(hipGetLastError());
}

template <int L>
void launch_sigma(hipStream_t st, const GwDims& d, const int* pairs, int npair, const double2* Gr,
                  const double2* Gl, const double2* Gg, const double2* Wr, const double2* Wl,
                  const double2* Wg, double2* Sr, double2* Sl, double2* Sg, double2 dE) {
    constexpr int NT = GwCfg<L>::NT;
    const size_t lds = sizeof(double2) * gw_lds_elems<L>();
    static bool attr = false;
    if (!attr) {
        QT_HIP(hipFuncSetAttribute((const void*)gw_sigma_kernel<L, NT>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    const int64_t blocks = (int64_t)npair * d.nx;
    if (blocks > 0x7fffffff) throw hip_error("gw self-energy: grid too large", 1);
    hipLaunchKernelGGL((gw_sigma_kernel<L, NT>), dim3((unsigned)blocks), dim3(NT), lds, st, d, pairs, npair,
                       Gr, Gl, Gg, Wr, Wl, Wg, Sr, Sl, Sg, dE, fft_twiddles(L));
    QT_HIP(hipGetLastError());
}

int fft_len(int nen) {
    int L = 2;
    while (L < 2 * nen - 1) L <<= 1;
    return L;
}

}  // namespace

void gw_polarization(hipStream_t st, const GwArgs& a, const int* elems, int nelem, const int* pairs, int npair) {
    GwDims d{a.nen, a.nky, a.nkz, a.nky * a.nkz, a.nx, a.nm};
    // dE = -i dEn/(2 pi) spindeg/nk  (gw_mod.f90:516)
    const double2 dE = make_double2(0.0, -a.dEn / 2.0 / kPi * a.spindeg / (double)d.nk);
    const int L = fft_len(a.nen);
    // one k point: element pairs share their transforms
    const bool use_pairs = d.nk == 1 && pairs != nullptr;
    const int* list = use_pairs ? pairs : elems;
    const int cnt = use_pairs ? npair : nelem;
#define QT_POL(LL) launch_pol<LL>(st, d, list, cnt, use_pairs, a.Gr, a.Gl, a.Gg, a.Pr, a.Pl, a.Pg, dE)
    switch (L) {
        case 2: QT_POL(2); break;
        case 4: QT_POL(4); break;
        case 8: QT_POL(8); break;
        case 16: QT_POL(16); break;
        case 32: QT_POL(32); break;
        case 64: QT_POL(64); break;
        case 128: QT_POL(128); break;
        case 256: QT_POL(256); break;
        case 512: QT_POL(512); break;
        case 1024: QT_POL(1024); break;
        case 2048: QT_POL(2048); break;
        case 4096: QT_POL(4096); break;
        default: throw hip_error("gw: more than 2048 energies", 6);
    }
#undef QT_POL
}

void gw_selfenergy(hipStream_t st, const GwArgs& a, const int* pairs, int npair) {
    GwDims d{a.nen, a.nky, a.nkz, a.nky * a.nkz, a.nx, a.nm};
    // dE = i dEn/(2 pi)/nk  (gw_mod.f90:694)
    const double2 dE = make_double2(0.0, a.dEn / 2.0 / kPi / (double)d.nk);
    const int L = fft_len(a.nen);
#define QT_SIG(LL) \
    launch_sigma<LL>(st, d, pairs, npair, a.Gr, a.Gl, a.Gg, a.Wr, a.Wl, a.Wg, a.Sr, a.Sl, a.Sg, dE)
    switch (L) {
        case 2: QT_SIG(2); break;
        case 4: QT_SIG(4); break;
        case 8: QT_SIG(8); break;
        case 16: QT_SIG(16); break;
        case 32: QT_SIG(32); break;
        case 64: QT_SIG(64); break;
        case 128: QT_SIG(128); break;
        case 256: QT_SIG(256); break;
        case 512: QT_SIG(512); break;
        case 1024: QT_SIG(1024); break;
        case 2048: QT_SIG(2048); break;
        case 4096: QT_SIG(4096); break;
        default: throw hip_error("gw: more than 2048 energies", 6);
    }
#undef QT_SIG
}

void gw_mix(hipStream_t st, int64_t n, double alpha, double2* S, const double2* Snew) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(gw_mix_kernel, dim3((unsigned)blocks), dim3(256), 0, st, n, alpha, S, Snew);
    QT_HIP(hipGetLastError());
}

void eph_sum(hipStream_t st, int64_t n, int64_t shift, const double2* G, double2* S) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(eph_sum_kernel, dim3((unsigned)blocks), dim3(256), 0, st, n, shift, G, S);
    QT_HIP(hipGetLastError());
}

void cell_bands(hipStream_t st, int nm, int nb, int nx, const double2* H, double2* U, double2* L) {
    const int64_t n = (int64_t)nm * nm * nx;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(cell_bands_kernel, dim3((unsigned)blocks), dim3(256), 0, st, nm, nb, nx, H, U, L);
    QT_HIP(hipGetLastError());
}

void cell_current(hipStream_t st, int nb, int ns, int nx, int cnt, const double2* C1, const double2* C2,
                  const double* cur, double* jd) {
    const int64_t n = (int64_t)nb * nb * ns * nx * cnt;
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(cell_current_kernel, dim3((unsigned)blocks), dim3(256), 0, st, nb, ns, nx, cnt, C1, C2,
                       cur, jd);
    QT_HIP(hipGetLastError());
}

void energy_sum(hipStream_t st, int64_t n, int cnt, const double* X, const double* w, double* acc) {
    if (n <= 0 || cnt <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(energy_sum_kernel, dim3((unsigned)blocks), dim3(256), 0, st, n, cnt, X, w, acc);
    QT_HIP(hipGetLastError());
}

void eph_retarded(hipStream_t st, int64_t n, const double2* Sl, const double2* Sg, double2* Sr) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(eph_retarded_kernel, dim3((unsigned)blocks), dim3(256), 0, st, n, Sl, Sg, Sr);
    QT_HIP(hipGetLastError());
}

}  // namespace qt
