// qt_inv.hip -- batched complex128 matrix inversion, blocked Gauss-Jordan
// with partial (row) pivoting, for the Schur complements of the RGF sweeps
// and the lead decimation (the role of ZGETRF+ZGETRI in linalg.f90:42-62).
//
// For each panel of NB columns [kb, kb+NB):
//   1. panel kernel (one workgroup per matrix, panel resident in LDS):
//      unblocked Gauss-Jordan on the n x NB panel; pivot = first row of
//      maximal |re|+|im| among the rows not yet pivoted (the IZAMAX choice
//      of ZGETRF); rows are swapped inside the panel and recorded in piv.
//   2. swap kernel: applies the panel's row swaps to every other column,
//      copies the pivot block row A[K, O] into R (NB x n, zero in the panel
//      columns) and zeroes A[K, O].
//   3. update: A += P * R on MFMA (P = the new panel columns), panel columns
//      masked from the write.  This is the blocked form of the GJ step
//        A_KO <- inv(A_KK) A_KO,  A_OO <- A_OO - A_OK inv(A_KK) A_KO.
// After the last panel the working matrix holds inv(P A); the inverse is
// gathered with the column interchanges applied in reverse order.  A zero
// pivot sets info and the output block is zeroed (linalg.f90:50-58).
#include <algorithm>

#include "qt_internal.h"

namespace qt {

namespace {

__device__ __forceinline__ double cabs1(double2 v) { return fabs(v.x) + fabs(v.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 crecip(double2 a) {
    // Smith's algorithm for 1/(a.x + i a.y)
    if (fabs(a.x) >= fabs(a.y)) {
        const double r = a.y / a.x, d = a.x + a.y * r;
        return make_double2(1.0 / d, -r / d);
    } else {
        const double r = a.x / a.y, d = a.y + a.x * r;
        return make_double2(r / d, -1.0 / d);
    }
}

template <int NB>
__global__ __launch_bounds__(256) void gj_panel_kernel(double2* A, int64_t sa, int n, int kb,
                                                       int* piv, int* info, const int* active) {
    const int b = blockIdx.x;
    if (active && !active[b]) return;
    extern __shared__ double2 P[];  // [NB][n], column c at P + c*n
    __shared__ double red_v[4];
    __shared__ int red_i[4];
    __shared__ int s_p;
    __shared__ double2 s_d;

    double2* Ab = A + b * sa;
    int* pv = piv + (int64_t)b * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ncols = min(NB, n - kb);

    for (int idx = tid; idx < n * ncols; idx += 256) {
        const int c = idx / n, r = idx - c * n;
        P[idx] = Ab[r + (int64_t)(kb + c) * n];
    }
    __syncthreads();

    for (int c = 0; c < ncols; ++c) {
        const int j = kb + c;
        double2* Pc = P + c * n;
        // --- pivot search over rows r >= j (first maximum wins)
        double best = -1.0;
        int bi = n;
        for (int r = j + tid; r < n; r += 256) {
            const double v = cabs1(Pc[r]);
            if (v > best) { best = v; bi = r; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if (lane == 0) { red_v[wave] = best; red_i[wave] = bi; }
        __syncthreads();
        if (tid == 0) {
            double bv = red_v[0];
            int bidx = red_i[0];
            for (int w = 1; w < 4; ++w)
                if (red_v[w] > bv || (red_v[w] == bv && red_i[w] < bidx)) { bv = red_v[w]; bidx = red_i[w]; }
            if (bidx >= n) bidx = j;
            s_p = bidx;
            pv[j] = bidx;
            if (!(bv > 0.0) && info[b] == 0) info[b] = j + 1;
        }
        __syncthreads();
        const int p = s_p;
        if (p != j && tid < ncols) {
            double2 t = P[tid * n + j];
            P[tid * n + j] = P[tid * n + p];
            P[tid * n + p] = t;
        }
        __syncthreads();
        if (tid == 0) {
            const double2 pivv = Pc[j];
            s_d = (pivv.x == 0.0 && pivv.y == 0.0) ? make_double2(0.0, 0.0) : crecip(pivv);
        }
        __syncthreads();
        const double2 d = s_d;
        // scale the pivot row
        if (tid < ncols) {
            if (tid == c) P[c * n + j] = d;
            else P[tid * n + j] = cmul(P[tid * n + j], d);
        }
        __syncthreads();
        // eliminate column c from every other row
        for (int r = tid; r < n; r += 256) {
            if (r == j) continue;
            const double2 f = Pc[r];
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) {
                if (cc >= ncols) break;
                if (cc == c) {
                    const double2 m = cmul(f, d);
                    Pc[r] = make_double2(-m.x, -m.y);
                } else {
                    const double2 u = P[cc * n + j];
                    double2 v = P[cc * n + r];
                    v.x -= f.x * u.x - f.y * u.y;
                    v.y -= f.x * u.y + f.y * u.x;
                    P[cc * n + r] = v;
                }
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < n * ncols; idx += 256) {
        const int c = idx / n, r = idx - c * n;
        Ab[r + (int64_t)(kb + c) * n] = P[idx];
    }
}

// One thread per (member, column outside the panel): apply the panel's row
// swaps, copy the pivot block row to R and zero it in A.
__global__ __launch_bounds__(256) void gj_swap_kernel(double2* A, int64_t sa, int n, int kb, int nb,
                                                      const int* piv, double2* R, int64_t sr,
                                                      const int* active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    double2* Rb = R + b * sr;
    if (c >= kb && c < kb + nb) {
        for (int t = 0; t < nb; ++t) Rb[t + (int64_t)c * nb] = make_double2(0.0, 0.0);
        return;
    }
    double2* col = A + b * sa + (int64_t)c * n;
    const int* pv = piv + (int64_t)b * n;
    for (int t = 0; t < nb; ++t) {
        const int j = kb + t, p = pv[j];
        if (p != j) {
            const double2 x = col[j];
            col[j] = col[p];
            col[p] = x;
        }
    }
    for (int t = 0; t < nb; ++t) {
        Rb[t + (int64_t)c * nb] = col[kb + t];
        col[kb + t] = make_double2(0.0, 0.0);
    }
}

// Output gather: X[:, c] = W[:, perm[c]] with the column interchanges
// applied in reverse order; zero block when the factorisation failed.
__global__ __launch_bounds__(256) void gj_finish_kernel(const double2* W, int64_t sw, double2* X,
                                                        int64_t sx, int n, const int* piv,
                                                        const int* info, const int* active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    extern __shared__ int perm[];
    const int* pv = piv + (int64_t)b * n;
    if (threadIdx.x == 0) {
        for (int c = 0; c < n; ++c) perm[c] = c;
        for (int j = n - 1; j >= 0; --j) {
            const int p = pv[j];
            const int t = perm[j];
            perm[j] = perm[p];
            perm[p] = t;
        }
    }
    __syncthreads();
    const bool bad = info[b] != 0;
    const double2* Wb = W + b * sw;
    double2* Xb = X + b * sx;
    // columns of this block: blockIdx.x covers 8 columns
    for (int cc = 0; cc < 8; ++cc) {
        const int c = blockIdx.x * 8 + cc;
        if (c >= n) break;
        const double2* src = Wb + (int64_t)perm[c] * n;
        double2* dst = Xb + (int64_t)c * n;
        for (int r = threadIdx.x; r < n; r += 256)
            dst[r] = bad ? make_double2(0.0, 0.0) : src[r];
    }
}

__global__ void zero_int_kernel(int* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

void zinv(hipStream_t s, const InvArgs& a, Timer* t) {
    const int n = a.n;
    if (n <= 0 || a.batch <= 0) return;
    if (n > 1024) throw hip_error("zinv: block size above 1024", 6);
    if (t && t->enabled) t->begin(s, 1);
    const int nb = (n <= 512) ? 16 : 8;
    const int64_t sr = (int64_t)16 * n;  // R member stride (sized for nb = 16)
    hipLaunchKernelGGL(zero_int_kernel, dim3((a.batch + 255) / 256), dim3(256), 0, s, a.info, a.batch);
    const size_t lds = (size_t)n * nb * sizeof(double2);
    static bool attr_set = false;
    if (!attr_set) {
        QT_HIP(hipFuncSetAttribute((const void*)gj_panel_kernel<16>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 512 * 16 * 16));
        QT_HIP(hipFuncSetAttribute((const void*)gj_panel_kernel<8>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 1024 * 8 * 16));
        attr_set = true;
    }
    for (int kb = 0; kb < n; kb += nb) {
        const int w = std::min(nb, n - kb);
        if (nb == 16)
            hipLaunchKernelGGL(gj_panel_kernel<16>, dim3(a.batch), dim3(256), lds, s, a.A, a.sa, n,
                               kb, a.piv, a.info, a.active);
        else
            hipLaunchKernelGGL(gj_panel_kernel<8>, dim3(a.batch), dim3(256), lds, s, a.A, a.sa, n,
                               kb, a.piv, a.info, a.active);
        QT_HIP(hipGetLastError());
        if (w == n) continue;  // a single panel holds the whole matrix
        hipLaunchKernelGGL(gj_swap_kernel, dim3((n + 255) / 256, a.batch), dim3(256), 0, s, a.A,
                           a.sa, n, kb, w, a.piv, a.R, sr, a.active);
        QT_HIP(hipGetLastError());
        // A[:, :] += A[:, kb:kb+w] * R   (panel columns masked)
        ZGemmArgs g;
        g.M = n;
        g.N = n;
        g.K = w;
        g.batch = a.batch;
        g.A.p = a.A + (int64_t)kb * n;
        g.A.ld = n;
        g.A.sb = a.sa;
        g.B.p = a.R;
        g.B.ld = w;
        g.B.sb = sr;
        g.C = a.A;
        g.ldc = n;
        g.sbc = a.sa;
        g.D.p = a.A;
        g.D.ld = n;
        g.D.sb = a.sa;
        g.alpha = 1.0;
        g.beta = 1.0;
        g.active = a.active;
        g.skip_lo = kb;
        g.skip_hi = kb + w;
        zgemm(s, OP_N, OP_N, g, nullptr);
    }
    hipLaunchKernelGGL(gj_finish_kernel, dim3((n + 7) / 8, a.batch), dim3(256), n * sizeof(int), s,
                       a.A, a.sa, a.X, a.sx, n, a.piv, a.info, a.active);
    QT_HIP(hipGetLastError());
    if (t && t->enabled)
        t->end(s, 1, 8.0 * (double)n * n * n * (a.flop_batch >= 0 ? a.flop_batch : a.batch));
}

}  // namespace qt
