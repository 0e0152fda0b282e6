// qt_fft.hip -- energy-axis correlation and convolution of many series
// (fft_mod.f90:58-91 corr1d2, :129-168 conv1d2, used by the GW polarisation
// and self-energy loops, gw_mod.f90:496-503, 672-692).
//
// The reference zero-pads two length-n series to 2n-1 and multiplies their
// DFTs (one MKL descriptor per call).  Here one workgroup handles one series
// pair: both are placed in a length-L buffer (L = the power of two >= 2n-1,
// which yields the same linear correlation / convolution), transformed by
// the LDS FFT of qt_fft.h, multiplied pointwise, inverse-transformed and the
// n wanted lags written.
//   corr: Z[i] = sum_e X[e] Y[e - l],            l = i - n/2
//         = (X (*) Yf)[l mod L],  Yf[(-q) mod L] = Y[q]
//   conv: Z[e] = sum_{l=-n/2}^{n-1-n/2} X[e - l] Y[l + n/2]
//         = (Yr (*) X)[e],        Yr[l mod L] = Y[l + n/2]
// Twiddles W_L^t = exp(-2 pi i t / L), t < L, come from a table built once per L.
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "qt_fft.h"
#include "qt_internal.h"

namespace qt {

namespace {

using fft::pad;

// mode 0: corr1d2, mode 1: conv1d2
template <int L, int NT>
__global__ __launch_bounds__(NT) void corrconv_kernel(int n, int mode, const double2* __restrict__ X,
                                                      const double2* __restrict__ Y, double2* __restrict__ Z,
                                                      const double2* __restrict__ tw, int64_t nseries) {
    constexpr int PER = L >= NT ? L / NT : 1;
    extern __shared__ double2 s[];
    const int64_t sidx = blockIdx.x;
    if (sidx >= nseries) return;
    const double2* x = X + sidx * n;
    const double2* y = Y + sidx * n;
    const int h = n / 2;
    // ---- transform of the first operand, X (corr) or the rotated Y (conv),
    // kept in registers: Yr[l mod L] = Y[l + n/2], l in [-n/2, n-1-n/2]
    double2 fx[PER];
    fft::transform_io<L, NT, false, 16>(
        s, tw,
        [&](int i, int) {
            double2 v = make_double2(0.0, 0.0);
            if (mode == 0) {
                if (i < n) v = x[i];
            } else {
                if (i < n - h) v = y[i + h];
                else if (i >= L - h) v = y[i - (L - h)];
            }
            return v;
        },
        [&](int, int q, double2 v) { fx[q] = v; });
    // ---- times the transform of the flipped Y (corr, Yf[(-q) mod L] = Y[q]) or X (conv)
    fft::transform_io<L, NT, false, 16>(
        s, tw,
        [&](int i, int) {
            double2 v = make_double2(0.0, 0.0);
            if (mode == 0) {
                if (i == 0) v = y[0];
                else if (i > L - n) v = y[L - i];
            } else {
                if (i < n) v = x[i];
            }
            return v;
        },
        [&](int, int q, double2 v) { fx[q] = fft::cmul(v, fx[q]); });
    // ---- inverse; sample m is lag m (corr: output i = m + n/2 mod L) or output m (conv)
    const double inv = 1.0 / (double)L;
    double2* z = Z + sidx * n;
    fft::transform_io<L, NT, true, 16>(
        s, tw, [&](int, int q) { return fx[q]; },
        [&](int m, int, double2 v) {
            const int i = mode == 0 ? (m + h) & (L - 1) : m;
            if (i < n) z[i] = make_double2(v.x * inv, v.y * inv);
        });
}

std::mutex tw_mu;
std::map<int, double2*> tw_tables;

}  // namespace

const double2* fft_twiddles(int L) {
    std::lock_guard<std::mutex> lk(tw_mu);
    auto it = tw_tables.find(L);
    if (it != tw_tables.end()) return it->second;
    std::vector<double2> h(L);
    for (int k = 0; k < L; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)L;
        h[k] = make_double2(std::cos(ang), std::sin(ang));
    }
    double2* d = nullptr;
    QT_HIP(hipMalloc((void**)&d, sizeof(double2) * L));
    QT_HIP(hipMemcpy(d, h.data(), sizeof(double2) * L, hipMemcpyHostToDevice));
    tw_tables[L] = d;
    return d;
}


namespace {

template <int L>
void launch_cc(hipStream_t s, int n, int mode, const double2* X, const double2* Y, double2* Z,
               int64_t nseries) {
    constexpr int NT = fft::threads_for(L);
    const size_t lds = sizeof(double2) * fft::padded_len(L);
    static bool attr = false;
    if (!attr) {
        QT_HIP(hipFuncSetAttribute((const void*)corrconv_kernel<L, NT>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        attr = true;
    }
    const double2* tw = fft_twiddles(L);
    for (int64_t s0 = 0; s0 < nseries; s0 += 0x7fffffff) {
        const int64_t cnt = std::min<int64_t>(0x7fffffff, nseries - s0);
        hipLaunchKernelGGL((corrconv_kernel<L, NT>), dim3((unsigned)cnt), dim3(NT), lds, s, n, mode,
                           X + s0 * n, Y + s0 * n, Z + s0 * n, tw, cnt);
        QT_HIP(hipGetLastError());
    }
}

}  // namespace

void fft_release() {
    std::lock_guard<std::mutex> lk(tw_mu);
    for (auto& kv : tw_tables) (void)hipFree(kv.second);
    tw_tables.clear();
}

void corr_conv(hipStream_t s, int n, int mode, const double2* X, const double2* Y, double2* Z,
               int64_t nseries) {
    if (n <= 0 || nseries <= 0) return;
    int L = 2;
    while (L < 2 * n - 1) L <<= 1;
    switch (L) {
        case 2: launch_cc<2>(s, n, mode, X, Y, Z, nseries); break;
        case 4: launch_cc<4>(s, n, mode, X, Y, Z, nseries); break;
        case 8: launch_cc<8>(s, n, mode, X, Y, Z, nseries); break;
        case 16: launch_cc<16>(s, n, mode, X, Y, Z, nseries); break;
        case 32: launch_cc<32>(s, n, mode, X, Y, Z, nseries); break;
        case 64: launch_cc<64>(s, n, mode, X, Y, Z, nseries); break;
        case 128: launch_cc<128>(s, n, mode, X, Y, Z, nseries); break;
        case 256: launch_cc<256>(s, n, mode, X, Y, Z, nseries); break;
        case 512: launch_cc<512>(s, n, mode, X, Y, Z, nseries); break;
        case 1024: launch_cc<1024>(s, n, mode, X, Y, Z, nseries); break;
        case 2048: launch_cc<2048>(s, n, mode, X, Y, Z, nseries); break;
        case 4096: launch_cc<4096>(s, n, mode, X, Y, Z, nseries); break;
        default: throw hip_error("corr/conv: series length above 2048", 6);
    }
}

}  // namespace qt
