// qt_elem.hip -- memory-bound companions of the RGF sweep: shifted blocks
// (z S - H - Sigma), anti-Hermitian parts, contact spectral currents,
// diagonal extraction, the Sancho-Rubio step update and small utilities.
// All are HBM-bound; blocks needing a transposed read stage a 32x32 tile
// pair through LDS so both reads stay coalesced.
#include <algorithm>

#include "qt_internal.h"

namespace qt {

namespace {

__device__ __forceinline__ const double2* member(const ZOp& o, int b, const int* kidx, int kshift) {
    if (!o.p) return nullptr;
    const int kz = kidx ? kidx[b >> kshift] : 0;
    return o.p + b * o.sb + kz * o.sk;
}

// ---------------------------------------------------------------- shifted
__global__ __launch_bounds__(256) void shifted_kernel(ShiftArgs a) {
    const int b = blockIdx.y;
    if (a.active && !a.active[b]) return;
    const int n = a.n;
    const int64_t nn = (int64_t)n * n;
    const double2* S = member(a.S, b, a.kidx, a.kshift);
    const double2* M1 = member(a.M1, b, a.kidx, a.kshift);
    const double2* M2 = member(a.M2, b, a.kidx, a.kshift);
    const double2* M3 = member(a.M3, b, a.kidx, a.kshift);
    const double zr = a.zre ? a.zre[b] : 0.0, zi = a.zim;
    double2* out = a.out + b * a.sout;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
        const int j = (int)(e / n), i = (int)(e - (int64_t)j * n);
        double re = 0.0, im = 0.0;
        if (S) {
            const double2 s = S[i + j * a.S.ld];
            re = zr * s.x - zi * s.y;
            im = zr * s.y + zi * s.x;
        } else if (i == j) {
            re = zr;
            im = zi;
        }
        if (M1) { const double2 v = M1[i + j * a.M1.ld]; re += a.c1 * v.x; im += a.c1 * v.y; }
        if (M2) { const double2 v = M2[i + j * a.M2.ld]; re += a.c2 * v.x; im += a.c2 * v.y; }
        if (M3) { const double2 v = M3[i + j * a.M3.ld]; re += a.c3 * v.x; im += a.c3 * v.y; }
        out[i + (int64_t)j * n] = make_double2(re, im);
    }
}

// ---------------------------------------------------------------- anti-Hermitian
// out = a_b (X - X^H) + c1 Y + c2 Z.  Workgroup = tile pair (ti, tj), ti <= tj.
constexpr int TT = 32;

__global__ __launch_bounds__(256) void antiherm_kernel(AntiArgs a, int ntile) {
    const int b = blockIdx.y;
    if (a.active && !a.active[b]) return;
    // decode upper-triangular tile index
    int t = blockIdx.x, ti = 0;
    while (t >= ntile - ti) { t -= ntile - ti; ++ti; }
    const int tj = ti + t;
    __shared__ double2 T1[TT][TT + 1];
    __shared__ double2 T2[TT][TT + 1];
    const int n = a.n;
    const double2* X = member(a.X, b, a.kidx, a.kshift);
    const double2* Y = member(a.Y, b, a.kidx, a.kshift);
    const double2* Z = member(a.Z, b, a.kidx, a.kshift);
    const double af = a.afac ? a.afac[b] : a.aconst;
    double2* out = a.out + b * a.sout;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    // T1 = X[ti block rows, tj block cols], T2 = X[tj rows, ti cols]
    for (int r = ty; r < TT; r += 8) {
        const int i1 = ti * TT + tx, j1 = tj * TT + r;
        T1[r][tx] = (i1 < n && j1 < n) ? X[i1 + j1 * a.X.ld] : make_double2(0.0, 0.0);
        const int i2 = tj * TT + tx, j2 = ti * TT + r;
        T2[r][tx] = (i2 < n && j2 < n) ? X[i2 + j2 * a.X.ld] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int r = ty; r < TT; r += 8) {
        // element (i1, j1) of the (ti, tj) block: X(i1,j1) - conj(X(j1,i1))
        {
            const int i1 = ti * TT + tx, j1 = tj * TT + r;
            if (i1 < n && j1 < n) {
                const double2 x = T1[r][tx], xt = T2[tx][r];
                double re = af * (x.x - xt.x), im = af * (x.y + xt.y);
                if (Y) { const double2 v = Y[i1 + j1 * a.Y.ld]; re += a.c1 * v.x; im += a.c1 * v.y; }
                if (Z) { const double2 v = Z[i1 + j1 * a.Z.ld]; re += a.c2 * v.x; im += a.c2 * v.y; }
                out[i1 + (int64_t)j1 * n] = make_double2(re, im);
            }
        }
        if (ti != tj) {
            const int i2 = tj * TT + tx, j2 = ti * TT + r;
            if (i2 < n && j2 < n) {
                const double2 x = T2[r][tx], xt = T1[tx][r];
                double re = af * (x.x - xt.x), im = af * (x.y + xt.y);
                if (Y) { const double2 v = Y[i2 + j2 * a.Y.ld]; re += a.c1 * v.x; im += a.c1 * v.y; }
                if (Z) { const double2 v = Z[i2 + j2 * a.Z.ld]; re += a.c2 * v.x; im += a.c2 * v.y; }
                out[i2 + (int64_t)j2 * n] = make_double2(re, im);
            }
        }
    }
}

// ---------------------------------------------------------------- contact current
// cur = Re sum_{j,k} Om[j,k] W[k,j],  Om = -(Sig - Sig^H),
// W = G^< + f (G - G^H)  (= f G^> - (f-1) G^<).
// Workgroup = tile pair (ti, tj) of the sum; partial sums per workgroup,
// reduced in a fixed order by the second kernel (deterministic).
__global__ __launch_bounds__(256) void current_partial_kernel(CurrentArgs a, int ntile) {
    const int b = blockIdx.y;
    if (a.active && !a.active[b]) return;
    const int ti = blockIdx.x % ntile, tj = blockIdx.x / ntile;
    const int n = a.n;
    const double2* S = a.Sig.p + b * a.Sig.sb;
    const double2* G = a.G.p + b * a.G.sb;
    const double2* L = a.Gl.p + b * a.Gl.sb;
    const double f = a.f[b];
    __shared__ double2 sS[TT][TT + 1], sSt[TT][TT + 1], sG[TT][TT + 1], sGt[TT][TT + 1], sL[TT][TT + 1];
    __shared__ double red[256];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // (j,k) in tile (ti,tj): j = ti*TT + ., k = tj*TT + .
    for (int r = ty; r < TT; r += 8) {
        const int j = ti * TT + tx, k = tj * TT + r;     // Sig(j,k), G(j,k)
        const bool in = j < n && k < n;
        sS[r][tx] = in ? S[j + k * a.Sig.ld] : make_double2(0.0, 0.0);
        sG[r][tx] = in ? G[j + k * a.G.ld] : make_double2(0.0, 0.0);
        const int k2 = tj * TT + tx, j2 = ti * TT + r;   // Sig(k,j), G(k,j), Gl(k,j)
        const bool in2 = k2 < n && j2 < n;
        sSt[r][tx] = in2 ? S[k2 + j2 * a.Sig.ld] : make_double2(0.0, 0.0);
        sGt[r][tx] = in2 ? G[k2 + j2 * a.G.ld] : make_double2(0.0, 0.0);
        sL[r][tx] = in2 ? L[k2 + j2 * a.Gl.ld] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    double acc = 0.0;
    for (int r = ty; r < TT; r += 8) {
        // element j = ti*TT+tx, k = tj*TT+r
        const double2 sjk = sS[r][tx];          // Sig(j,k)
        const double2 skj = sSt[tx][r];         // Sig(k,j)
        const double omr = -(sjk.x - skj.x), omi = -(sjk.y + skj.y);  // Om(j,k)
        const double2 gkj = sGt[tx][r];         // G(k,j)
        const double2 gjk = sG[r][tx];          // G(j,k)
        const double2 lkj = sL[tx][r];          // G^<(k,j)
        const double wr = lkj.x + f * (gkj.x - gjk.x);
        const double wi = lkj.y + f * (gkj.y + gjk.y);
        acc += omr * wr - omi * wi;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[(int64_t)b * gridDim.x + blockIdx.x] = red[0];
}

__global__ void reduce_rows_kernel(const double* partial, int ncol, double* out, const int* active,
                                   double scale) {
    const int b = blockIdx.x;
    if (active && !active[b]) return;
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < ncol; i += 256) acc += partial[(int64_t)b * ncol + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = scale * red[0];
}

__global__ __launch_bounds__(256) void real_diff_kernel(int n, double* out, int64_t sout, ZOp X1, ZOp X2) {
    const int b = blockIdx.y;
    const int64_t nn = (int64_t)n * n;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
        const int j = (int)(e / n), i = (int)(e - (int64_t)j * n);
        out[b * sout + e] = X1.p[b * X1.sb + i + j * X1.ld].x - X2.p[b * X2.sb + i + j * X2.ld].x;
    }
}

// ---------------------------------------------------------------- diagonals
__global__ void diag_kernel(DiagArgs a) {
    const int b = blockIdx.y;
    if (a.active && !a.active[b]) return;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.n) return;
    const double2 g = a.G.p[b * a.G.sb + j + j * a.G.ld];
    const double2 l = a.Gl.p[b * a.Gl.sb + j + j * a.Gl.ld];
    const int64_t o = b * a.ld + a.off + j;
    if (a.gr) a.gr[o] = g;
    if (a.gl) a.gl[o] = l;
    if (a.gg) a.gg[o] = make_double2(l.x, l.y + 2.0 * g.y);  // G - G^H on the diagonal = 2i Im G
}

// ---------------------------------------------------------------- Sancho step
__device__ __forceinline__ double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void sancho_update_kernel(SanchoUpdArgs a) {
    const int b = blockIdx.y;
    if (!a.active[b]) return;
    const bool on1 = a.active1 ? a.active1[b] != 0 : true;
    const bool on2 = a.HSS2 && a.active2[b] != 0;
    const int64_t nn = (int64_t)a.n * a.n;
    const int64_t base = b * a.s;
    __shared__ double red[256];
    double acc1 = 0.0, acc2 = 0.0;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
        const double2 c = a.Cm[base + e];
        const double2 ap = a.Ap[base + e];
        double2 bb = a.HBB[base + e];
        bb.x += c.x + ap.x;
        bb.y += c.y + ap.y;
        a.HBB[base + e] = bb;
        if (on1) {
            double2 ss = a.HSS[base + e];
            ss.x += c.x;
            ss.y += c.y;
            a.HSS[base + e] = ss;
        }
        if (on2) {
            double2 s2 = a.HSS2[base + e];
            s2.x += ap.x;
            s2.y += ap.y;
            a.HSS2[base + e] = s2;
        }
        const double2 h = a.H01n[base + e];
        acc1 += sqrt(h.y * h.y + h.x * h.x);
        if (a.HSS2) {
            const double2 g = a.H10n[base + e];
            acc2 += sqrt(g.y * g.y + g.x * g.x);
        }
    }
    const double s1 = block_sum(acc1, red);
    const double s2 = a.HSS2 ? block_sum(acc2, red) : 0.0;
    if (threadIdx.x == 0) {
        a.partial[((int64_t)b * 2 + 0) * gridDim.x + blockIdx.x] = s1;
        a.partial[((int64_t)b * 2 + 1) * gridDim.x + blockIdx.x] = s2;
    }
}

// convergence decision after the partial sums: rgf_mod.f90:535-551
__global__ void sancho_decide_kernel(SanchoUpdArgs a, int ncol) {
    const int b = blockIdx.x;
    if (!a.active[b]) return;
    __shared__ double red[256];
    double acc1 = 0.0, acc2 = 0.0;
    for (int i = threadIdx.x; i < ncol; i += 256) {
        acc1 += a.partial[((int64_t)b * 2 + 0) * ncol + i];
        acc2 += a.partial[((int64_t)b * 2 + 1) * ncol + i];
    }
    const double err1 = block_sum(acc1, red);
    const double err2 = block_sum(acc2, red);
    const bool on1 = a.active1 ? a.active1[b] != 0 : true;
    const bool on2 = a.HSS2 && a.active2[b] != 0;
    const bool conv1 = fabs(err1) < a.tol;
    const bool conv2 = fabs(err2) < a.tol;
    // reset on the last iteration (gw_mod.f90:1440-1445)
    if (a.reset && ((on1 && !conv1) || (on2 && !conv2))) {
        const int64_t nn = (int64_t)a.n * a.n;
        const int kz = a.kidx ? a.kidx[b] : 0;
        const double2* H00 = a.H00 + b * a.sh00 + kz * a.skh00;
        for (int64_t e = threadIdx.x; e < nn; e += 256) {
            if (on1 && !conv1) a.HSS[b * a.s + e] = H00[e];
            if (on2 && !conv2) a.HSS2[b * a.s + e] = H00[e];
            a.HBB[b * a.s + e] = H00[e];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool still = false;
        if (on1) {
            a.iters[b] = a.iter;
            if (conv1) {
                if (a.active1) a.active1[b] = 0;
            } else {
                still = true;
            }
        }
        if (on2) {
            a.iters2[b] = a.iter;
            if (conv2) a.active2[b] = 0;
            else still = true;
        }
        a.active[b] = still ? 1 : 0;
    }
}

// ---------------------------------------------------------------- copies
__global__ __launch_bounds__(256) void copy_kernel(int n, double2* dst, int64_t sdst, ZOp src,
                                                   const int* kidx, int kshift) {
    const int b = blockIdx.y;
    const double2* S = member(src, b, kidx, kshift);
    const int64_t nn = (int64_t)n * n;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
        const int j = (int)(e / n), i = (int)(e - (int64_t)j * n);
        dst[b * sdst + e] = S[i + j * src.ld];
    }
}

__global__ __launch_bounds__(256) void dagger_kernel(int n, double2* dst, int64_t sdst, ZOp src,
                                                     const int* kidx, int kshift, int ntile) {
    const int b = blockIdx.y;
    const double2* S = member(src, b, kidx, kshift);
    const int ti = blockIdx.x % ntile, tj = blockIdx.x / ntile;
    __shared__ double2 T[TT][TT + 1];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < TT; r += 8) {
        const int i = ti * TT + tx, j = tj * TT + r;
        T[r][tx] = (i < n && j < n) ? S[i + j * src.ld] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int r = ty; r < TT; r += 8) {
        // dst(j', i') = conj(src(i', j')) with dst row = tj*TT+tx, col = ti*TT+r
        const int i = tj * TT + tx, j = ti * TT + r;
        if (i < n && j < n) {
            const double2 v = T[tx][r];
            dst[b * sdst + i + (int64_t)j * n] = make_double2(v.x, -v.y);
        }
    }
}

__global__ void count_kernel(const int* active, int batch, int* count) {
    __shared__ int red[256];
    int acc = 0;
    for (int i = threadIdx.x; i < batch; i += 256) acc += active[i] ? 1 : 0;
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = red[0];
}

__global__ void fill_int_kernel(int* p, int n, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int elem_blocks(int64_t nn) {
    return (int)std::min<int64_t>((nn + 255) / 256, 1024);
}

}  // namespace

void shifted(hipStream_t s, const ShiftArgs& a) {
    if (a.batch <= 0) return;
    const int64_t nn = (int64_t)a.n * a.n;
    hipLaunchKernelGGL(shifted_kernel, dim3(elem_blocks(nn), a.batch), dim3(256), 0, s, a);
    QT_HIP(hipGetLastError());
}

void antiherm(hipStream_t s, const AntiArgs& a) {
    if (a.batch <= 0) return;
    const int nt = (a.n + TT - 1) / TT;
    hipLaunchKernelGGL(antiherm_kernel, dim3(nt * (nt + 1) / 2, a.batch), dim3(256), 0, s, a, nt);
    QT_HIP(hipGetLastError());
}

void contact_current(hipStream_t s, const CurrentArgs& a) {
    if (a.batch <= 0) return;
    const int nt = (a.n + TT - 1) / TT;
    hipLaunchKernelGGL(current_partial_kernel, dim3(nt * nt, a.batch), dim3(256), 0, s, a, nt);
    QT_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(a.batch), dim3(256), 0, s, a.partial, nt * nt, a.out,
                       a.active, a.scale);
    QT_HIP(hipGetLastError());
}

void diag_extract(hipStream_t s, const DiagArgs& a) {
    if (a.batch <= 0) return;
    hipLaunchKernelGGL(diag_kernel, dim3((a.n + 255) / 256, a.batch), dim3(256), 0, s, a);
    QT_HIP(hipGetLastError());
}

void sancho_update(hipStream_t s, const SanchoUpdArgs& a) {
    if (a.batch <= 0) return;
    const int64_t nn = (int64_t)a.n * a.n;
    const int nblk = (int)std::min<int64_t>((nn + 255) / 256, 32);
    hipLaunchKernelGGL(sancho_update_kernel, dim3(nblk, a.batch), dim3(256), 0, s, a);
    QT_HIP(hipGetLastError());
    hipLaunchKernelGGL(sancho_decide_kernel, dim3(a.batch), dim3(256), 0, s, a, nblk);
    QT_HIP(hipGetLastError());
}

void real_diff(hipStream_t s, int n, int batch, double* out, int64_t sout, ZOp X1, ZOp X2) {
    if (batch <= 0) return;
    hipLaunchKernelGGL(real_diff_kernel, dim3(elem_blocks((int64_t)n * n), batch), dim3(256), 0, s, n,
                       out, sout, X1, X2);
    QT_HIP(hipGetLastError());
}

void copy_blocks(hipStream_t s, int n, int batch, double2* dst, int64_t sdst, ZOp src,
                 const int* kidx, int kshift, bool dagger) {
    if (batch <= 0) return;
    if (!dagger) {
        hipLaunchKernelGGL(copy_kernel, dim3(elem_blocks((int64_t)n * n), batch), dim3(256), 0, s, n,
                           dst, sdst, src, kidx, kshift);
    } else {
        const int nt = (n + TT - 1) / TT;
        hipLaunchKernelGGL(dagger_kernel, dim3(nt * nt, batch), dim3(256), 0, s, n, dst, sdst, src,
                           kidx, kshift, nt);
    }
    QT_HIP(hipGetLastError());
}

void count_active(hipStream_t s, const int* active, int batch, int* count) {
    hipLaunchKernelGGL(count_kernel, dim3(1), dim3(256), 0, s, active, batch, count);
    QT_HIP(hipGetLastError());
}

void fill_int(hipStream_t s, int* p, int n, int v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    QT_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- Timer
void Timer::begin(hipStream_t s, int kind) {
    if (!created) {
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < 4096; ++i) QT_HIP(hipEventCreate(&ev[k][i]));
        created = true;
    }
    if (n[kind] + 2 > 4096) flush();
    QT_HIP(hipEventRecord(ev[kind][n[kind]], s));
}

void Timer::end(hipStream_t s, int kind, double fl) {
    QT_HIP(hipEventRecord(ev[kind][n[kind] + 1], s));
    n[kind] += 2;
    flops[kind] += fl;
    launches[kind] += 1;
}

void Timer::flush() {
    for (int k = 0; k < 2; ++k) {
        for (int i = 0; i + 1 < n[k]; i += 2) {
            QT_HIP(hipEventSynchronize(ev[k][i + 1]));
            float t = 0.f;
            QT_HIP(hipEventElapsedTime(&t, ev[k][i], ev[k][i + 1]));
            ms[k] += t;
        }
        n[k] = 0;
    }
}

void Timer::reset() {
    n[0] = n[1] = 0;
    flops[0] = flops[1] = 0;
    launches[0] = launches[1] = 0;
    ms[0] = ms[1] = 0;
}

}  // namespace qt
