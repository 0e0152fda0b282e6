// qt_fft.h -- workgroup-level complex128 FFT held in LDS (gfx950).
//
// A length-L (power of two, 2..4096) transform is done by NT threads in
// place on one LDS buffer with Stockham autosort passes of radix 16 (last
// pass radix 2, 4 or 8):
//   pass (R, Ns): butterfly j in [0, L/R), k = j mod Ns
//     v[r] = s[j + r L/R] * W_{Ns R}^{r k}        r = 0..R-1
//     v    = DFT_R(v)                              (in registers)
//     s[(j - k) R + k + r Ns] = v[r]
// (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC 2008).  Every pass reads its operands into
// registers, synchronises, and writes them back, so one buffer suffices.
// LDS indices are padded by one element per 16 (pad()), which makes the
// radix-16 scatter of the first pass (stride 16 elements) conflict free.
//
// Twiddles come from a table tw[t] = exp(-2 pi i t / L), t < L, in global
// memory (W^r of a butterfly is built from W^1, W^2, W^4, W^8 with at most
// three products); the small-radix DFTs use exact constants of the 16th
// roots of 1.
#pragma once

#include <hip/hip_runtime.h>

namespace qt {
namespace fft {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

__host__ __device__ constexpr int pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int padded_len(int L) { return pad(L - 1) + 1; }

// cos and sin of 2 pi j / 16
struct Root16 {
    double c[16], s[16];
};
__device__ constexpr Root16 root16() {
    constexpr double c1 = 0.92387953251128675613;  // cos(pi/8)
    constexpr double s1 = 0.38268343236508977173;  // sin(pi/8)
    constexpr double h = 0.70710678118654752440;   // sqrt(1/2)
    return Root16{{1.0, c1, h, s1, 0.0, -s1, -h, -c1, -1.0, -c1, -h, -s1, 0.0, s1, h, c1},
                  {0.0, s1, h, c1, 1.0, c1, h, s1, 0.0, -s1, -h, -c1, -1.0, -c1, -h, -s1}};
}

// multiply by exp(sign 2 pi i j / 16), sign = -1 forward, +1 inverse
template <int J, bool INV>
__device__ __forceinline__ double2 rot16(double2 v) {
    constexpr Root16 w = root16();
    constexpr double c = w.c[J & 15];
    constexpr double s = INV ? w.s[J & 15] : -w.s[J & 15];
    if constexpr ((J & 15) == 0) return v;
    else if constexpr ((J & 15) == 4 || (J & 15) == 12)  // +-i
        return make_double2(-s * v.y, s * v.x);
    else if constexpr ((J & 15) == 8)
        return make_double2(-v.x, -v.y);
    else
        return make_double2(c * v.x - s * v.y, c * v.y + s * v.x);
}

// In-register DFT of R = 2, 4, 8, 16 points, natural order in and out
// (radix-2 decimation in frequency, then the bit-reversal renaming).
template <int R, bool INV>
__device__ __forceinline__ void dft(double2* v) {
#pragma unroll
    for (int half = R / 2; half >= 1; half >>= 1) {
#pragma unroll
        for (int st = 0; st < R; st += 2 * half) {
#pragma unroll
            for (int i = 0; i < half; ++i) {
                const double2 u = v[st + i], t = v[st + i + half];
                v[st + i] = cadd(u, t);
                const double2 d = csub(u, t);
                // W_{2 half}^i = root16^(i * 16 / (2 half))
                const int j = i * (16 / (2 * half));
                double2 r;
                switch (j) {
                    case 0: r = d; break;
                    case 1: r = rot16<1, INV>(d); break;
                    case 2: r = rot16<2, INV>(d); break;
                    case 3: r = rot16<3, INV>(d); break;
                    case 4: r = rot16<4, INV>(d); break;
                    case 5: r = rot16<5, INV>(d); break;
                    case 6: r = rot16<6, INV>(d); break;
                    default: r = rot16<7, INV>(d); break;
                }
                v[st + i + half] = r;
            }
        }
    }
    double2 t[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        int b = 0;
#pragma unroll
        for (int q = 1, kk = k; q < R; q <<= 1, kk >>= 1) b = (b << 1) | (kk & 1);
        t[k] = v[b];
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = t[k];
}

// threadIdx.x through an empty asm: values derived from it are recomputed
// in every pass instead of being hoisted out of the callers' loops (dozens
// of loop-invariant LDS and twiddle addresses would otherwise stay live)
__device__ __forceinline__ int opaque_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

template <int L, int NT, int R, int NS, bool INV>
__device__ __forceinline__ void pass(double2* s, const double2* __restrict__ tw) {
    constexpr int NBF = L / R;                              // butterflies
    constexpr int PER = NBF >= NT ? NBF / NT : 1;           // per thread
    double2 v[PER][R];
    const int tid = opaque_tid();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = tid + b * NT;
        if (NBF >= NT || j < NBF) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = s[pad(j + r * NBF)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = tid + b * NT;
        if (NBF >= NT || j < NBF) {
            const int k = j & (NS - 1);
            if constexpr (NS > 1) {
                // W^r from the table entries of W^1, W^2, W^4, W^8 (at most
                // three products per factor)
                constexpr int step = L / (NS * R);
                double2 wp[4];
#pragma unroll
                for (int bit = 0; (1 << bit) < R; ++bit) wp[bit] = tw[(k << bit) * step];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    double2 w = make_double2(1.0, 0.0);
                    bool first = true;
#pragma unroll
                    for (int bit = 0; (1 << bit) < R; ++bit)
                        if (r & (1 << bit)) {
                            w = first ? wp[bit] : cmul(w, wp[bit]);
                            first = false;
                        }
                    v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
                }
            }
            dft<R, INV>(v[b]);
            const int d = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) s[pad(d + r * NS)] = v[b][r];
        }
    }
    __syncthreads();
}

template <int L, int NT, int NS, bool INV, int RMAX>
__device__ __forceinline__ void passes(double2* s, const double2* __restrict__ tw) {
    if constexpr (NS < L) {
        constexpr int rem = L / NS;
        constexpr int R = rem >= RMAX ? RMAX : rem;
        pass<L, NT, R, NS, INV>(s, tw);
        passes<L, NT, NS * R, INV, RMAX>(s, tw);
    }
}

// In-place unnormalised transform of s[pad(0..L-1)]; sign -1 (forward) or
// +1 (INV), passes of radix RMAX (16 or 8) and a smaller last one.  The
// caller synchronises before (data written) and the last pass synchronises
// after.
template <int L, int NT, bool INV, int RMAX = 16>
__device__ __forceinline__ void transform(double2* s, const double2* __restrict__ tw) {
    static_assert(RMAX == 2 || RMAX == 4 || RMAX == 8 || RMAX == 16, "radix");
    passes<L, NT, 1, INV, RMAX>(s, tw);
}

// ---- transforms with fused input / output
// A thread's own samples are i = tid + q NT, q < L / NT.  When every pass
// has at least NT butterflies (L / R >= NT for its radix R), the first pass
// reads exactly the thread's samples (j = tid + b NT, inputs j + r L/R,
// q = b + r PER) and the last pass (NS = L / R) produces them, so the
// inputs can come straight from registers or global memory and the outputs
// go straight to registers or global memory, saving an LDS round trip and a
// barrier at each end.  Loaders are called as ld(i, q), stores as
// st(i, q, value); i is the natural sample index.

// first pass (NS = 1: no twiddles), inputs from ld
template <int L, int NT, int R, bool INV, class Ld>
__device__ __forceinline__ void pass_first(double2* s, Ld&& ld) {
    constexpr int NBF = L / R;
    constexpr int PER = NBF >= NT ? NBF / NT : 1;
    double2 v[PER][R];
    const int tid = opaque_tid();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = tid + b * NT;
        if (NBF >= NT || j < NBF) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = ld(j + r * NBF, b + r * PER);
        }
    }
    __syncthreads();  // earlier readers of s are done
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = tid + b * NT;
        if (NBF >= NT || j < NBF) {
            dft<R, INV>(v[b]);
#pragma unroll
            for (int r = 0; r < R; ++r) s[pad(j * R + r)] = v[b][r];
        }
    }
    __syncthreads();
}

// last pass (NS R = L, so k = j and the outputs are j + r NS), outputs to st;
// s is only read
template <int L, int NT, int R, int NS, bool INV, class St>
__device__ __forceinline__ void pass_last(const double2* s, const double2* __restrict__ tw, St&& st) {
    static_assert(NS * R == L, "last pass");
    constexpr int NBF = L / R;
    constexpr int PER = NBF >= NT ? NBF / NT : 1;
    double2 v[PER][R];
    const int tid = opaque_tid();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = tid + b * NT;
        if (NBF >= NT || j < NBF) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = s[pad(j + r * NBF)];
        }
    }
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int j = tid + b * NT;
        if (NBF >= NT || j < NBF) {
            if constexpr (NS > 1) {
                constexpr int step = L / (NS * R);
                double2 wp[4];
#pragma unroll
                for (int bit = 0; (1 << bit) < R; ++bit) wp[bit] = tw[(j << bit) * step];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    double2 w = make_double2(1.0, 0.0);
                    bool first = true;
#pragma unroll
                    for (int bit = 0; (1 << bit) < R; ++bit)
                        if (r & (1 << bit)) {
                            w = first ? wp[bit] : cmul(w, wp[bit]);
                            first = false;
                        }
                    v[b][r] = INV ? cmulc(v[b][r], w) : cmul(v[b][r], w);
                }
            }
            dft<R, INV>(v[b]);
#pragma unroll
            for (int r = 0; r < R; ++r) st(j + r * NS, b + r * PER, v[b][r]);
        }
    }
}

template <int L, int NT, int NS, bool INV, int RMAX, class St>
__device__ __forceinline__ void passes_to(double2* s, const double2* __restrict__ tw, St&& st) {
    constexpr int rem = L / NS;
    constexpr int R = rem >= RMAX ? RMAX : rem;
    if constexpr (NS * R == L) {
        pass_last<L, NT, R, NS, INV>(s, tw, st);
    } else {
        pass<L, NT, R, NS, INV>(s, tw);
        passes_to<L, NT, NS * R, INV, RMAX>(s, tw, st);
    }
}

// radix of the last pass of transform<L, NT, INV, RMAX>
template <int L, int RMAX, int NS = 1>
constexpr int last_radix() {
    if constexpr (L / NS <= RMAX) return L / NS;
    else return last_radix<L, RMAX, NS * RMAX>();
}
// true when the first and last passes cover exactly the threads' own samples
template <int L, int NT, int RMAX>
constexpr bool fused_io() {
    return L > RMAX && L / RMAX >= NT && L / last_radix<L, RMAX>() >= NT;
}

// transform of the samples ld(i, q) left in s (natural order); ld is only
// asked for global-memory style values (its q may be ignored)
template <int L, int NT, bool INV, int RMAX, class Ld>
__device__ __forceinline__ void transform_load(double2* s, const double2* __restrict__ tw, Ld&& ld) {
    constexpr int R0 = L >= RMAX ? RMAX : L;
    if constexpr (R0 < L) {
        pass_first<L, NT, R0, INV>(s, ld);
        passes<L, NT, R0, INV, RMAX>(s, tw);
    } else {
        __syncthreads();
        for (int i = threadIdx.x; i < L; i += NT) s[pad(i)] = ld(i, -1);
        __syncthreads();
        transform<L, NT, INV, RMAX>(s, tw);
    }
}

// transform of the thread's samples ld(i, q) with the results handed to
// st(i, q, value) for the thread's own samples; s is scratch
template <int L, int NT, bool INV, int RMAX, class Ld, class St>
__device__ __forceinline__ void transform_io(double2* s, const double2* __restrict__ tw, Ld&& ld, St&& st) {
    if constexpr (fused_io<L, NT, RMAX>()) {
        pass_first<L, NT, RMAX, INV>(s, ld);
        passes_to<L, NT, RMAX, INV, RMAX>(s, tw, st);
    } else {
        constexpr int PER = L >= NT ? L / NT : 1;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            if (L >= NT || i < L) s[pad(i)] = ld(i, q);
        }
        __syncthreads();
        transform<L, NT, INV, RMAX>(s, tw);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            if (L >= NT || i < L) st(i, q, s[pad(i)]);
        }
    }
}

// threads per workgroup used for length L
__host__ __device__ constexpr int threads_for(int L) { return L >= 4096 ? 256 : (L >= 32 ? (L / 16 > 64 ? L / 16 : 64) : 64); }

}  // namespace fft
}  // namespace qt
