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
// workgroup owns one element (P) or one element pair ij/ji (Sigma) at one
// slab, transforms each input series once per term with the LDS FFT of
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

// value transforms applied while loading
enum Load { L_PLAIN = 0, L_IMAG = 1, L_IMAG_CONJ = 2 };
__device__ __forceinline__ double2 xform(double2 v, int how) {
    if (how == L_IMAG) return make_double2(0.0, v.y);
    if (how == L_IMAG_CONJ) return make_double2(0.0, -v.y);
    return v;
}

// s <- X zero-padded (the X operand of corr / conv)
template <int L, int NT>
__device__ __forceinline__ void load_x(double2* s, const double2* __restrict__ x, int n, int how) {
    for (int i = threadIdx.x; i < L; i += NT) s[pad(i)] = i < n ? xform(x[i], how) : make_double2(0.0, 0.0);
}
// s <- Yf, Yf[(-q) mod L] = Y[q]  (corr)
template <int L, int NT>
__device__ __forceinline__ void load_flip(double2* s, const double2* __restrict__ y, int n, int how) {
    for (int i = threadIdx.x; i < L; i += NT) {
        double2 v = make_double2(0.0, 0.0);
        if (i == 0) v = xform(y[0], how);
        else if (i > L - n) v = xform(y[L - i], how);
        s[pad(i)] = v;
    }
}
// s <- Yr (+ Yr of y2 when given), Yr[l mod L] = Y[l + n/2]  (conv)
template <int L, int NT>
__device__ __forceinline__ void load_rot(double2* s, const double2* __restrict__ y,
                                         const double2* __restrict__ y2, int n) {
    const int h = n / 2;
    for (int i = threadIdx.x; i < L; i += NT) {
        int src = -1;
        if (i < n - h) src = i + h;
        else if (i >= L - h) src = i - (L - h);
        double2 v = make_double2(0.0, 0.0);
        if (src >= 0) {
            v = y[src];
            if (y2) v = cadd(v, y2[src]);
        }
        s[pad(i)] = v;
    }
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
    __device__ __forceinline__ void take(const double2* s) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            if (ok(i)) v[q] = s[pad(i)];
        }
    }
    // this += s * other
    __device__ __forceinline__ void fma_lds(const double2* s, const Spec& o) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            if (ok(i)) v[q] = cadd(v[q], cmul(s[pad(i)], o.v[q]));
        }
    }
    __device__ __forceinline__ void put(double2* s) const {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            if (ok(i)) s[pad(i)] = v[q];
        }
    }
};

constexpr int RADIX = 8;

// F(X) and F(Yf) of one corr term, accumulated into acc
template <int L, int NT>
__device__ __forceinline__ void corr_term(double2* s, const double2* tw, Spec<L, NT>& acc,
                                          Spec<L, NT>& hold, const double2* x, int hx,
                                          const double2* y, int hy, int n) {
    __syncthreads();
    load_x<L, NT>(s, x, n, hx);
    __syncthreads();
    fft::transform<L, NT, false, RADIX>(s, tw);
    hold.take(s);
    __syncthreads();
    load_flip<L, NT>(s, y, n, hy);
    __syncthreads();
    fft::transform<L, NT, false, RADIX>(s, tw);
    acc.fma_lds(s, hold);
}

// inverse of acc, scaled by c / L, lags written to z (corr) or samples (conv)
template <int L, int NT>
__device__ __forceinline__ void finish(double2* s, const double2* tw, const Spec<L, NT>& acc,
                                       double2 c, int n, bool corr, double2* z) {
    __syncthreads();
    acc.put(s);
    __syncthreads();
    fft::transform<L, NT, true, RADIX>(s, tw);
    const double inv = 1.0 / (double)L;
    const int h = n / 2;
    for (int i = threadIdx.x; i < n; i += NT) {
        int src = i;
        if (corr) {
            const int l = i - h;
            src = l >= 0 ? l : l + L;
        }
        const double2 v = s[pad(src)];
        z[i] = make_double2((c.x * v.x - c.y * v.y) * inv, (c.x * v.y + c.y * v.x) * inv);
    }
}

// ---------------------------------------------------------------- P
template <int L, int NT>
__global__ __launch_bounds__(NT) void gw_pol_kernel(GwDims d, const int* __restrict__ elems, int nelem,
                                                    const double2* __restrict__ Gr,
                                                    const double2* __restrict__ Gl,
                                                    const double2* __restrict__ Gg, double2* Pr,
                                                    double2* Pl, double2* Pg, double2 dE,
                                                    const double2* __restrict__ tw) {
    extern __shared__ double2 s[];
    const int w = blockIdx.x;
    if (w >= nelem * d.nx) return;
    const int e = w / d.nx, ix = w % d.nx;
    const int ij = elems[e];
    const int row = ij % d.nm, col = ij / d.nm;
    const int ji = col + row * d.nm;
    const int n = d.nen;
    Spec<L, NT> acc, hold;
    for (int iq = 0; iq < d.nk; ++iq) {
        // P^< = sum_k corr(G^<_ij(k), G^>_ji(k'))
        acc.zero();
        for (int ik = 0; ik < d.nk; ++ik) {
            const int kd = kd_index(ik, iq, d.nky, d.nkz);
            corr_term<L, NT>(s, tw, acc, hold, series(Gl, d, ij, ix, ik), L_IMAG, series(Gg, d, ji, ix, kd),
                             L_IMAG, n);
        }
        finish<L, NT>(s, tw, acc, dE, n, true, series(Pl, d, ij, ix, iq));
        // P^> = sum_k corr(G^>_ij(k), G^<_ji(k'))
        acc.zero();
        for (int ik = 0; ik < d.nk; ++ik) {
            const int kd = kd_index(ik, iq, d.nky, d.nkz);
            corr_term<L, NT>(s, tw, acc, hold, series(Gg, d, ij, ix, ik), L_IMAG, series(Gl, d, ji, ix, kd),
                             L_IMAG, n);
        }
        finish<L, NT>(s, tw, acc, dE, n, true, series(Pg, d, ij, ix, iq));
        // P^R = sum_k corr(G^<_ij(k), conj G^R_ij(k')) + corr(G^R_ij(k), G^<_ji(k'))
        acc.zero();
        for (int ik = 0; ik < d.nk; ++ik) {
            const int kd = kd_index(ik, iq, d.nky, d.nkz);
            corr_term<L, NT>(s, tw, acc, hold, series(Gl, d, ij, ix, ik), L_IMAG, series(Gr, d, ij, ix, kd),
                             L_IMAG_CONJ, n);
            corr_term<L, NT>(s, tw, acc, hold, series(Gr, d, ij, ix, ik), L_IMAG, series(Gl, d, ji, ix, kd),
                             L_IMAG, n);
        }
        finish<L, NT>(s, tw, acc, dE, n, true, series(Pr, d, ij, ix, iq));
    }
}

// ---------------------------------------------------------------- Sigma
// One workgroup per element pair (ij, ji) with row <= col and slab ix.  The
// Sigma of ij is kept in registers (EXT samples per thread, the same sample
// mapping for every output), the Sigma of ji is folded into it as
// (S + S^T)/2 as soon as each output is ready, so no thread reads back
// global memory.
template <int L, int NT>
__device__ __forceinline__ void sigma_elem(double2* s, const double2* tw, const GwDims& d, int ij, int ix,
                                           int ik, const double2* Gr, const double2* Gl, const double2* Gg,
                                           const double2* Wr, const double2* Wl, const double2* Wg,
                                           double2 dE, Spec<L, NT>& acc, Spec<L, NT>& hold, bool second,
                                           double2* aR, double2* aL, double2* aG) {
    constexpr int EXT = (L / 2 + NT - 1) / NT;
    const int n = d.nen;
    double2 t[EXT];
    double dim[EXT];
    // t = dE/L * F^-1[acc] at the samples this thread owns
    auto extract = [&]() {
        __syncthreads();
        acc.put(s);
        __syncthreads();
        fft::transform<L, NT, true, RADIX>(s, tw);
        const double inv = 1.0 / (double)L;
#pragma unroll
        for (int q = 0; q < EXT; ++q) {
            const int i = threadIdx.x + q * NT;
            const double2 v = i < n ? s[pad(i)] : make_double2(0.0, 0.0);
            t[q] = make_double2((dE.x * v.x - dE.y * v.y) * inv, (dE.x * v.y + dE.y * v.x) * inv);
        }
    };
    auto fold = [&](double2* a) {
#pragma unroll
        for (int q = 0; q < EXT; ++q)
            a[q] = second ? make_double2(0.5 * (a[q].x + t[q].x), 0.5 * (a[q].y + t[q].y)) : t[q];
    };
    // conv term: acc += F(X) F(Yr); X = i Im G, Yr the rotated W (+ W2)
    auto conv_term = [&](const double2* g, const double2* w, const double2* w2) {
        __syncthreads();
        load_rot<L, NT>(s, w, w2, n);
        __syncthreads();
        fft::transform<L, NT, false, RADIX>(s, tw);
        hold.take(s);
        __syncthreads();
        load_x<L, NT>(s, g, n, L_IMAG);
        __syncthreads();
        fft::transform<L, NT, false, RADIX>(s, tw);
        acc.fma_lds(s, hold);
    };
    // Sigma^< = sum_q conv(G^<(k'), W^<(q))
    acc.zero();
    for (int iq = 0; iq < d.nk; ++iq) {
        const int kd = kd_index(ik, iq, d.nky, d.nkz);
        conv_term(series(Gl, d, ij, ix, kd), series(Wl, d, ij, ix, iq), nullptr);
    }
    extract();
#pragma unroll
    for (int q = 0; q < EXT; ++q) dim[q] = -t[q].y;
    fold(aL);
    // Sigma^> = sum_q conv(G^>(k'), W^>(q))
    acc.zero();
    for (int iq = 0; iq < d.nk; ++iq) {
        const int kd = kd_index(ik, iq, d.nky, d.nkz);
        conv_term(series(Gg, d, ij, ix, kd), series(Wg, d, ij, ix, iq), nullptr);
    }
    extract();
#pragma unroll
    for (int q = 0; q < EXT; ++q) dim[q] += t[q].y;
    fold(aG);
    // Sigma^R = sum_q conv(G^<, W^R) + conv(G^R, W^< + W^R), then
    // Re Sigma^R + i Im(Sigma^> - Sigma^<) / 2  (gw_mod.f90:694-698)
    acc.zero();
    for (int iq = 0; iq < d.nk; ++iq) {
        const int kd = kd_index(ik, iq, d.nky, d.nkz);
        conv_term(series(Gl, d, ij, ix, kd), series(Wr, d, ij, ix, iq), nullptr);
        conv_term(series(Gr, d, ij, ix, kd), series(Wl, d, ij, ix, iq), series(Wr, d, ij, ix, iq));
    }
    extract();
#pragma unroll
    for (int q = 0; q < EXT; ++q) t[q].y = 0.5 * dim[q];
    fold(aR);
}

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
    constexpr int EXT = (L / 2 + NT - 1) / NT;
    extern __shared__ double2 s[];
    const int w = blockIdx.x;
    if (w >= npair * d.nx) return;
    const int p = w / d.nx, ix = w % d.nx;
    const int ij = pairs[p];
    const int row = ij % d.nm, col = ij / d.nm;
    const int ji = col + row * d.nm;
    const int n = d.nen;
    Spec<L, NT> acc, hold;
    for (int ik = 0; ik < d.nk; ++ik) {
        double2 aR[EXT], aL[EXT], aG[EXT];
        sigma_elem<L, NT>(s, tw, d, ij, ix, ik, Gr, Gl, Gg, Wr, Wl, Wg, dE, acc, hold, false, aR, aL, aG);
        if (ji != ij)
            sigma_elem<L, NT>(s, tw, d, ji, ix, ik, Gr, Gl, Gg, Wr, Wl, Wg, dE, acc, hold, true, aR, aL, aG);
        double2 *r1 = series(Sr, d, ij, ix, ik), *l1 = series(Sl, d, ij, ix, ik), *g1 = series(Sg, d, ij, ix, ik);
        double2 *r2 = series(Sr, d, ji, ix, ik), *l2 = series(Sl, d, ji, ix, ik), *g2 = series(Sg, d, ji, ix, ik);
#pragma unroll
        for (int q = 0; q < EXT; ++q) {
            const int i = threadIdx.x + q * NT;
            if (i < n) {
                r1[i] = aR[q];
                l1[i] = aL[q];
                g1[i] = aG[q];
                r2[i] = aR[q];
                l2[i] = aL[q];
                g2[i] = aG[q];
            }
        }
    }
}

__global__ void gw_mix_kernel(int64_t n, double alpha, double2* S, const double2* __restrict__ Snew) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = S[t], b = Snew[t];
        S[t] = make_double2(a.x + alpha * (b.x - a.x), a.y + alpha * (b.y - a.y));
    }
}

// one radix-8 butterfly per thread where L allows (512 threads at L = 4096)
template <int L>
struct GwCfg {
    static constexpr int NT = L / 8 >= 512 ? 512 : (L / 8 >= 64 ? L / 8 : 64);
};

template <int L>
void launch_pol(hipStream_t st, const GwDims& d, const int* elems, int nelem, const double2* Gr,
                const double2* Gl, const double2* Gg, double2* Pr, double2* Pl, double2* Pg, double2 dE) {
    constexpr int NT = GwCfg<L>::NT;
    const size_t lds = sizeof(double2) * fft::padded_len(L);
    static bool attr = false;
    if (!attr) {
        QT_HIP(hipFuncSetAttribute((const void*)gw_pol_kernel<L, NT>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    const int64_t blocks = (int64_t)nelem * d.nx;
    if (blocks > 0x7fffffff) throw hip_error("gw polarization: grid too large", 1);
    hipLaunchKernelGGL((gw_pol_kernel<L, NT>), dim3((unsigned)blocks), dim3(NT), lds, st, d, elems, nelem,
                       Gr, Gl, Gg, Pr, Pl, Pg, dE, fft_twiddles(L));
    QT_HIP(hipGetLastError());
}

template <int L>
void launch_sigma(hipStream_t st, const GwDims& d, const int* pairs, int npair, const double2* Gr,
                  const double2* Gl, const double2* Gg, const double2* Wr, const double2* Wl,
                  const double2* Wg, double2* Sr, double2* Sl, double2* Sg, double2 dE) {
    constexpr int NT = GwCfg<L>::NT;
    const size_t lds = sizeof(double2) * fft::padded_len(L);
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

void gw_polarization(hipStream_t st, const GwArgs& a, const int* elems, int nelem) {
    GwDims d{a.nen, a.nky, a.nkz, a.nky * a.nkz, a.nx, a.nm};
    // dE = -i dEn/(2 pi) spindeg/nk  (gw_mod.f90:516)
    const double2 dE = make_double2(0.0, -a.dEn / 2.0 / kPi * a.spindeg / (double)d.nk);
    const int L = fft_len(a.nen);
#define QT_POL(LL) launch_pol<LL>(st, d, elems, nelem, a.Gr, a.Gl, a.Gg, a.Pr, a.Pl, a.Pg, dE)
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

}  // namespace qt
