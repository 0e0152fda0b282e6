// qt_rgf.hip -- batched RGF sweeps and the C ABI (include/qt_rgf.h).
//
// A solve is cut into device batches of B energy points.  For one batch the
// work is a fixed sequence of batched launches (GEMM / inversion /
// elementwise) over all B points at once:
//   leads:   Sancho-Rubio for both contacts as one batch of 2B members,
//            then the contact self-energies;
//   pass 1:  left-connected (forward variant) or right-connected (backward
//            variant) Gl, Gln for every slab, kept resident in HBM;
//   pass 2:  G^R, G^<, G^> of every slab from the neighbour's result,
//            diagonals written out, contact spectral currents.
// The arithmetic of each step is that of rgf_mod.f90 (cited per step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/qt_rgf.h"
#include "qt_ctx.h"
#include "qt_internal.h"


namespace qt {

// Fixed scratch of qt_invert_batched_noalloc: status word, info flags, pivot
// lists and the inversion's panel multipliers, QT_NOALLOC_SCRATCH_BYTES per
// device, held until the process ends.  Host code uses the device address
// that qt_init takes with hipGetSymbolAddress (Ctx::scratch), never this name.
static __device__ __align__(256) int scratch[QT_NOALLOC_SCRATCH_BYTES / sizeof(int)];

Ctx* g_ctx = nullptr;
std::mutex g_init_mu;
static thread_local std::string g_last_error;

// product form of a precision mode (ZGemmArgs::prec): 0 for QT_PREC_FP64, 3
// for the three-multiplication FP64 form, the int8 digits per value for the
// mixed modes; -1 for a mode that does not exist
static int prec_digits(int precision) {
    switch (precision) {
        case QT_PREC_FP64: return 0;
        case QT_PREC_MIXED: return 6;
        case QT_PREC_MIXED47: return 7;
        case QT_PREC_FP64_3M: return 3;
        default: return -1;
    }
}

size_t work_bytes(const qt_ham& h, int B, bool sigma) {
    const int64_t nn = (int64_t)h.nmax * h.nmax;
    const int64_t tiles = (h.nmax + 31) / 32;
    size_t b = 0;
    auto add = [&](size_t bytes) { b += (bytes + 255) & ~size_t(255); };
    add(sizeof(double2) * (size_t)h.hset * B * (sigma ? 4 : 2));
    for (int i = 0; i < NTEMP; ++i) add(sizeof(double2) * (size_t)nn * 2 * B);
    add(sizeof(double2) * (size_t)nn * B * 2);
    add(sizeof(int) * (size_t)2 * B * h.nmax);
    add(sizeof(double2) * (size_t)2 * B * h.nmax * kInvPanelMax);
    int64_t ndiag = 0;
    for (int s = 0; s < h.nx; ++s) ndiag += h.nm[s];
    add(sizeof(double2) * (size_t)3 * B * ndiag);
    add(sizeof(double) * (size_t)2 * B * std::max<int64_t>(64, tiles * tiles));
    add(4096 + sizeof(double) * 16 * (size_t)B + sizeof(int) * 16 * (size_t)B);
    return b + (1 << 20);
}

void work_carve(Arena& a, Work& w, const qt_ham& h, int B, bool sigma) {
    a.reset();
    w.B = B;
    w.nmax = h.nmax;
    w.nn = (int64_t)h.nmax * h.nmax;
    w.ndiag = 0;
    for (int s = 0; s < h.nx; ++s) w.ndiag += h.nm[s];
    const int64_t tiles = (h.nmax + 31) / 32;
    w.Gl = a.take<double2>((size_t)h.hset * B);
    w.Gln = a.take<double2>((size_t)h.hset * B);
    w.H00 = sigma ? a.take<double2>((size_t)h.hset * B) : nullptr;
    w.SLS = sigma ? a.take<double2>((size_t)h.hset * B) : nullptr;
    for (int i = 0; i < NTEMP; ++i) w.T[i] = a.take<double2>((size_t)w.nn * 2 * B);
    w.sigF = a.take<double2>((size_t)w.nn * B);
    w.sigL = a.take<double2>((size_t)w.nn * B);
    w.piv = a.take<int>((size_t)2 * B * h.nmax);
    w.ginv = a.take<double2>((size_t)2 * B * h.nmax * kInvPanelMax);
    w.info = a.take<int>((size_t)2 * B);
    w.zre = a.take<double>((size_t)2 * B);
    w.kidx = a.take<int>((size_t)2 * B);
    w.fF = a.take<double>(B);
    w.fFc = a.take<double>(B);
    w.fL = a.take<double>(B);
    w.nfF = a.take<double>(B);
    w.nfL = a.take<double>(B);
    w.active = a.take<int>((size_t)2 * B);
    w.iters = a.take<int>((size_t)2 * B);
    w.lead_it = a.take<int>((size_t)2 * B);
    w.act1 = a.take<int>((size_t)2 * B);
    w.act2 = a.take<int>((size_t)2 * B);
    w.iters2 = a.take<int>((size_t)2 * B);
    w.count = a.take<int>(1);
    w.partial = a.take<double>((size_t)2 * B * std::max<int64_t>(64, tiles * tiles));
    w.infoacc = a.take<int>(B);
    w.tr = a.take<double>(B);
    w.tre = a.take<double>(B);
    w.diag = a.take<double2>((size_t)3 * B * w.ndiag);
}

void inv_batch(Ctx& c, Work& w, int n, int batch, double2* A, int64_t sa, double2* X,
               int64_t sx, const int* active, int flop_batch) {
    InvArgs a;
    a.n = n;
    a.batch = batch;
    a.A = A;
    a.sa = sa;
    a.X = X;
    a.sx = sx;
    a.piv = w.piv;
    a.info = w.info;
    a.G = w.ginv;
    a.sg = (int64_t)n * kInvPanelMax;
    a.active = active;
    a.flop_batch = flop_batch;
    zinv(c.stream, a, &c.timer);
}

// accumulate inversion failures of members into per-point info bits
__global__ void info_accum_kernel(const int* info, int batch, int* acc, int period) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b < batch && info[b] != 0) atomicOr(&acc[b % period], (int)QT_INFO_SINGULAR);
}

void note_inv_info(Ctx& c, Work& w, int batch, int period) {
    hipLaunchKernelGGL(info_accum_kernel, dim3((batch + 255) / 256), dim3(256), 0, c.stream, w.info,
                       batch, w.infoacc, period);
    QT_HIP(hipGetLastError());
}

// lead iteration counts -> info bit when nmax was reached without convergence
__global__ void lead_info_kernel(const int* active, int members, int* acc, int period) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m < members && active[m]) atomicOr(&acc[m % period], (int)QT_INFO_LEAD_NMAX);
}

void sancho_run(Ctx& c, Work& w, SanchoIO& io) {
    const int n = io.n, M = io.members;
    const int64_t nn = io.nn;
    hipStream_t s = c.stream;
    double2* const* T = w.T;
    double2 *HSS = T[0], *HBB = T[1], *H10 = io.H10, *H01 = T[2], *A = T[3], *Ai = T[4];
    double2 *B1 = T[5], *B2 = T[6], *C1 = T[7], *H01n = T[8], *Cp = T[9], *Ap = T[10];
    // H_BB = H00, H_10 = H10, H_01 = H10^H, H_SS = H00  (rgf_mod.f90:515-518)
    copy_blocks(s, n, M, HSS, nn, io.H00, io.kidx, 0, false);
    copy_blocks(s, n, M, HBB, nn, io.H00, io.kidx, 0, false);
    copy_blocks(s, n, M, H01, nn, op_member(H10, n, nn), nullptr, 0, true);
    const bool dual = io.G00b != nullptr;
    double2* HSS2 = T[14];
    if (dual) {
        copy_blocks(s, n, M, HSS2, nn, io.H00, io.kidx, 0, false);
        fill_int(s, w.act1, M, 1);
        fill_int(s, w.act2, M, 1);
        fill_int(s, w.iters2, M, 0);
    }
    fill_int(s, w.active, M, 1);
    fill_int(s, w.iters, M, 0);
    int nact = M;
    Gemm g{s, &c.timer, M, nullptr};
    g.active = w.active;
    io.total_iters = 0;
    for (int it = 1; it <= io.nmax && nact > 0; ++it) {
        io.total_iters += nact;
        g.flop_batch = nact;
        ShiftArgs sh;
        sh.n = n;
        sh.batch = M;
        sh.out = A;
        sh.sout = nn;
        sh.zre = io.zre;
        sh.zim = io.eta;
        sh.S = io.S;
        sh.M1 = op_member(HBB, n, nn);
        sh.c1 = -1.0;
        sh.kidx = io.kidx;
        sh.active = w.active;
        shifted(s, sh);                                   // A = z S00 - H_BB
        inv_batch(c, w, n, M, A, nn, Ai, nn, w.active, nact);  // invert(A)
        const ZOp a = op_member(Ai, n, nn), h10 = op_member(H10, n, nn), h01 = op_member(H01, n, nn);
        g(OP_N, OP_N, n, n, n, a, h10, B1, n, nn);                       // B = A H_10
        g(OP_N, OP_N, n, n, n, h01, op_member(B1, n, nn), C1, n, nn);    // C = H_01 B
        g(OP_N, OP_N, n, n, n, h10, op_member(B1, n, nn), Cp, n, nn);    // C = H_10 B
        g(OP_N, OP_N, n, n, n, a, h01, B2, n, nn);                       // B = A H_01
        g(OP_N, OP_N, n, n, n, h10, op_member(B2, n, nn), Ap, n, nn);    // A = H_10 B
        g(OP_N, OP_N, n, n, n, h01, op_member(B2, n, nn), H01n, n, nn);  // C = H_01 B
        SanchoUpdArgs u;
        u.n = n;
        u.batch = M;
        u.HSS = HSS;
        u.HBB = HBB;
        u.Cm = C1;
        u.Ap = Ap;
        u.H01n = H01n;
        if (dual) {
            u.HSS2 = HSS2;
            u.H10n = Cp;
            u.active1 = w.act1;
            u.active2 = w.act2;
            u.iters2 = w.iters2;
        }
        u.s = nn;
        u.partial = w.partial;
        u.active = w.active;
        u.iters = w.iters;
        u.iter = it;
        u.tol = io.tol;
        u.reset = (io.reset && it == io.nmax) ? 1 : 0;
        u.H00 = io.H00.p;
        u.sh00 = io.H00.sb;
        u.skh00 = io.H00.sk;
        u.kidx = io.kidx;
        sancho_update(s, u);  // H_SS += C, H_BB += C + A, error = sum |H_01|, stop test
        std::swap(H10, Cp);   // H_10 = C (new coupling)
        std::swap(H01, H01n); // H_01 = C
        count_active(s, w.active, M, w.count);
        QT_HIP(hipMemcpyAsync(c.host_count, w.count, sizeof(int), hipMemcpyDeviceToHost, s));
        QT_HIP(hipStreamSynchronize(s));
        nact = *c.host_count;
    }
    if (io.info_bits && nact > 0) {
        hipLaunchKernelGGL(lead_info_kernel, dim3((M + 255) / 256), dim3(256), 0, s, w.active, M,
                           io.info_bits, io.period);
        QT_HIP(hipGetLastError());
    }
    if (io.iters_out)
        QT_HIP(hipMemcpyAsync(io.iters_out, w.iters, sizeof(int) * M, hipMemcpyDeviceToDevice, s));
    if (dual && io.iters2_out)
        QT_HIP(hipMemcpyAsync(io.iters2_out, w.iters2, sizeof(int) * M, hipMemcpyDeviceToDevice, s));
    // G00 = inv(z S00 - H_SS), GBB = inv(z S00 - H_BB)  (rgf_mod.f90:553-557)
    if (dual) {
        ShiftArgs sh;
        sh.n = n;
        sh.batch = M;
        sh.out = A;
        sh.sout = nn;
        sh.zre = io.zre;
        sh.zim = io.eta;
        sh.S = io.S;
        sh.M1 = op_member(HSS2, n, nn);
        sh.c1 = -1.0;
        sh.kidx = io.kidx;
        shifted(s, sh);
        inv_batch(c, w, n, M, A, nn, io.G00b, nn, nullptr);
        if (io.info_bits) note_inv_info(c, w, M, io.period);
    }
    for (int pass = 0; pass < (io.GBB ? 2 : 1); ++pass) {
        ShiftArgs sh;
        sh.n = n;
        sh.batch = M;
        sh.out = A;
        sh.sout = nn;
        sh.zre = io.zre;
        sh.zim = io.eta;
        sh.S = io.S;
        sh.M1 = op_member(pass == 0 ? HSS : HBB, n, nn);
        sh.c1 = -1.0;
        sh.kidx = io.kidx;
        shifted(s, sh);
        inv_batch(c, w, n, M, A, nn, pass == 0 ? io.G00 : io.GBB, nn, nullptr);
        if (io.info_bits) note_inv_info(c, w, M, io.period);
    }
}

// ------------------------------------------------------------ full-block host outputs
struct FullOut {
    bool any = false;
    void* const* G_r = nullptr;
    void* const* G_lesser = nullptr;
    void* const* G_greater = nullptr;
    void* const* Jdens = nullptr;
    void* const* Gl = nullptr;
    void* const* Gln = nullptr;
    void* const* cur = nullptr;   // QT_RGF_GW extras
    void* const* GRi1 = nullptr;
    void* const* GLi1 = nullptr;
    void* const* GGi1 = nullptr;
    int nx = 0;
    int p0 = 0;  // first point of the batch
    // device element-major (nm*nm, nx, 1, nen) destinations of G^R, G^<, G^>
    // (point p0 + b at energy index p0 + b; members b < nb only)
    double2* em[3] = {nullptr, nullptr, nullptr};
    int nen = 0;
    int nb = 0;
};

// out[q*out_sq + x*out_sx + p] = in[p*in_sp + x*in_sx + q], p < P, q < Q, x < X:
// the transposes between element-major series and member-major blocks
__global__ __launch_bounds__(256) void tr3_kernel(const double2* __restrict__ in, int64_t in_sp, int64_t in_sx,
                                                  double2* __restrict__ out, int64_t out_sq, int64_t out_sx,
                                                  int P, int Q) {
    __shared__ double2 t[32][33];
    const int q0 = blockIdx.x * 32, p0 = blockIdx.y * 32;
    const double2* ib = in + (int64_t)blockIdx.z * in_sx;
    double2* ob = out + (int64_t)blockIdx.z * out_sx;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int pp = p0 + k, qq = q0 + threadIdx.x;
        if (pp < P && qq < Q) t[k][threadIdx.x] = ib[(int64_t)pp * in_sp + qq];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int qq = q0 + k, pp = p0 + threadIdx.x;
        if (pp < P && qq < Q) ob[(int64_t)qq * out_sq + pp] = t[threadIdx.x][k];
    }
}

static void transpose3(hipStream_t s, const double2* in, int64_t in_sp, int64_t in_sx, double2* out,
                       int64_t out_sq, int64_t out_sx, int P, int Q, int X) {
    if (P <= 0 || Q <= 0 || X <= 0) return;
    if ((P + 31) / 32 > 65535 || X > 65535) throw hip_error("transpose3: grid too large", 1);
    hipLaunchKernelGGL(tr3_kernel, dim3((Q + 31) / 32, (P + 31) / 32, X), dim3(32, 8), 0, s, in, in_sp, in_sx,
                       out, out_sq, out_sx, P, Q);
    QT_HIP(hipGetLastError());
}

static void d2h_real_blocks(Ctx& c, void* const* dst, int nx, int p0, int B, int s, int n,
                            const double* src, int64_t sb) {
    if (!dst) return;
    QT_HIP(hipStreamSynchronize(c.stream));
    for (int b = 0; b < B; ++b) {
        void* d = dst[(int64_t)(p0 + b) * nx + s];
        if (!d) continue;
        QT_HIP(hipMemcpy(d, src + b * sb, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost));
    }
}

static void d2h_blocks(Ctx& c, void* const* dst, int nx, int p0, int B, int s, int n,
                       const double2* src, int64_t sb) {
    if (!dst) return;
    QT_HIP(hipStreamSynchronize(c.stream));
    for (int b = 0; b < B; ++b) {
        void* d = dst[(int64_t)(p0 + b) * nx + s];
        if (!d) continue;
        QT_HIP(hipMemcpy(d, src + b * sb, sizeof(double2) * (size_t)n * n, hipMemcpyDeviceToHost));
    }
}

// ------------------------------------------------------------ one batch
struct Solve {
    Ctx& c;
    qt_ham& h;
    const qt_rgf_params& p;
    Work& w;
    int B;
    bool fwd;
    bool fix;
    bool sigma;
    bool gw;       // green_RGF_RS bookkeeping (QT_RGF_GW)
    FullOut full;
    double2* gr = nullptr;  // diag destinations (device), already offset to the batch
    double2* glo = nullptr;
    double2* gg = nullptr;
    int64_t lead_iterations = 0;

    std::vector<int64_t> doff;  // diag offset of slab s

    ZOp opH00(int s) const {
        if (sigma) return op_member(w.H00 + h.hoff[s] * B, h.nm[s], sq(s));
        return op_k(h.H + h.hoff[s], h.nm[s], h.hset);
    }
    ZOp opS(int s) const {
        if (!h.S) return ZOp();
        return op_k(h.S + h.hoff[s], h.nm[s], h.hset);
    }
    ZOp opC(int s) const {  // H1i[s]: nm[s] x nm[s-1]; s = 0 and s = nx are square
        const int rows = s < h.nx ? h.nm[s] : h.nm[h.nx - 1];
        return op_k(h.C + h.coff[s], rows, h.cset);
    }
    int64_t sq(int s) const { return (int64_t)h.nm[s] * h.nm[s]; }
    double2* gl(int s) const { return w.Gl + h.hoff[s] * B; }
    double2* gln(int s) const { return w.Gln + h.hoff[s] * B; }
    ZOp opGl(int s) const { return op_member(gl(s), h.nm[s], sq(s)); }
    ZOp opGln(int s) const { return op_member(gln(s), h.nm[s], sq(s)); }
    ZOp opSLS(int s) const {
        if (!w.SLS) return ZOp();
        return op_member(w.SLS + h.hoff[s] * B, h.nm[s], sq(s));
    }

    void diag_out(int s, const double2* G, const double2* L) {
        DiagArgs d;
        d.n = h.nm[s];
        d.batch = B;
        d.G = op_member(G, h.nm[s], w.nn);
        d.Gl = op_member(L, h.nm[s], w.nn);
        d.gr = gr;
        d.gl = glo;
        d.gg = gg;
        d.ld = w.ndiag;
        d.off = doff[s];
        diag_extract(c.stream, d);
    }

    // G_r, G_lesser, G_greater (and Jdens from `jd`) of slab s to the host
    void full_slab(int s, const double2* G, const double2* L, const double2* jd) {
        if (!full.any) return;
        const int n = h.nm[s];
        d2h_blocks(c, full.G_r, full.nx, full.p0, B, s, n, G, w.nn);
        d2h_blocks(c, full.G_lesser, full.nx, full.p0, B, s, n, L, w.nn);
        // element-major destinations: member b's block -> series (ij, s) at energy p0 + b
        const int64_t sser = (int64_t)full.nx * full.nen;
        auto scatter = [&](double2* em, const double2* src) {
            if (em)
                transpose3(c.stream, src, w.nn, 0, em + (int64_t)s * full.nen + full.p0, sser, 0, full.nb,
                           n * n, 1);
        };
        scatter(full.em[0], G);
        scatter(full.em[1], L);
        if (full.G_greater || full.em[2]) {
            AntiArgs an;
            an.n = n;
            an.batch = B;
            an.out = w.T[12];
            an.sout = w.nn;
            an.X = op_member(G, n, w.nn);
            an.aconst = 1.0;
            an.Y = op_member(L, n, w.nn);
            an.c1 = 1.0;
            antiherm(c.stream, an);  // G^> = G^< + (G - G^H)
            d2h_blocks(c, full.G_greater, full.nx, full.p0, B, s, n, w.T[12], w.nn);
            scatter(full.em[2], w.T[12]);
        }
        if (jd) d2h_blocks(c, full.Jdens, full.nx, full.p0, B, s, n, jd, w.nn);
    }

    // green_RGF_RS extras of slab sc (gw_mod.f90:1250-1291), with its solved
    // neighbour's G (Gc) and G^< (Lc), X = Gl W^H, Y = X G^<, Q = W Gln:
    //   GLi1 = G<_{l,l+1} = Y - (Gc Q)^H          (Gln anti-Hermitian)
    //   A    = G<_{l+1,l} = Lc X^H + Gc Q
    //   cur  = Re(W^H A) - Re(W GLi1),  GRi1 = GN0,  GGi1 = Gl^H W^H G^>_last
    void gw_extras(int sc, int n, const ZOp& W, double2* X, double2* Y, double2* Q, double2* GN0,
                   double2* Gc, double2* Lc, const Gemm& g) {
        hipStream_t s = c.stream;
        const int64_t nn = w.nn;
        double2* const* T = w.T;
        double2* R = T[10];
        double2* Bm = T[11];
        double2* Acur = T[14];
        double2* P1 = T[13] + (int64_t)B * nn;
        double2* P2 = T[12];
        g(OP_N, OP_N, n, n, n, op_member(Gc, n, nn), op_member(Q, n, nn), R, n, nn);   // Gc Q
        AntiArgs an;  // Bm = (R - R^H) + Y - R = Y - R^H
        an.n = n;
        an.batch = B;
        an.out = Bm;
        an.sout = nn;
        an.X = op_member(R, n, nn);
        an.aconst = 1.0;
        an.Y = op_member(Y, n, nn);
        an.c1 = 1.0;
        an.Z = op_member(R, n, nn);
        an.c2 = -1.0;
        antiherm(s, an);
        g(OP_N, OP_C, n, n, n, op_member(Lc, n, nn), op_member(X, n, nn), Acur, n, nn, 1.0,
          op_member(R, n, nn), 1.0);                                                    // Lc X^H + Gc Q
        if (full.cur) {
            g(OP_C, OP_N, n, n, n, W, op_member(Acur, n, nn), P1, n, nn);
            g(OP_N, OP_N, n, n, n, W, op_member(Bm, n, nn), P2, n, nn);
            double* curb = (double*)T[2];  // P is consumed by now; 2 nn doubles per member
            real_diff(s, n, B, curb, nn, op_member(P1, n, nn), op_member(P2, n, nn));
            d2h_real_blocks(c, full.cur, full.nx, full.p0, B, sc, n, curb, nn);
        }
        d2h_blocks(c, full.GLi1, full.nx, full.p0, B, sc, n, Bm, nn);
        d2h_blocks(c, full.GRi1, full.nx, full.p0, B, sc, n, GN0, nn);
        if (full.GGi1) {
            g(OP_C, OP_C, n, n, n, opGl(sc), W, R, n, nn);                                // Gl^H W^H
            g(OP_N, OP_N, n, n, n, op_member(R, n, nn), op_member(T[13], n, nn), P2, n, nn);
            d2h_blocks(c, full.GGi1, full.nx, full.p0, B, sc, n, P2, nn);
        }
    }

    void run(qt_timing& tm) {
        hipStream_t s = c.stream;
        const int nx = h.nx;
        const int64_t nn = w.nn;
        const int sf = fwd ? 0 : nx - 1;  // first slab of pass 1
        const int sl = fwd ? nx - 1 : 0;  // last slab of pass 1
        const int step = fwd ? 1 : -1;
        const int nF = h.nm[sf], nL = h.nm[sl];
        const bool want_j = full.any && full.Jdens && !gw;
        const bool want_gw = gw && full.any && (full.cur || full.GLi1 || full.GRi1 || full.GGi1);
        double2* const* T = w.T;
        Gemm g{s, &c.timer, B, w.kidx};
        QT_HIP(hipEventRecord(c.ev[0], s));

        // ---------------- leads
        // first contact: forward = left lead: sancho(H10 = H1i(1)^H),
        //   Sigma = H1i(1) G00 H1i(1)^H            (rgf_mod.f90:80-93)
        //   backward = right lead: sancho(H10 = H1i(nx+1)),
        //   Sigma = H1i(nx+1) G00 H1i(nx+1)^H      (rgf_mod.f90:302-314)
        // last contact: forward = right lead: sancho(H10 = H1i(nx+1)),
        //   Sigma = H1i(nx+1)^H G00 H1i(nx+1)      (rgf_mod.f90:157-161)
        //   backward = left lead: sancho(H10 = H1i(1)^H),
        //   Sigma = H1i(1)^H G00 H1i(1)            (rgf_mod.f90:370-374)
        const int cF = fwd ? 0 : nx, cL = fwd ? nx : 0;
        {
            double2* H10buf = T[11];
            double2* G00buf = T[12];
            double2* H00buf = T[13];
            copy_blocks(s, nF, B, H10buf, nn, opC(cF), w.kidx, 0, fwd);
            copy_blocks(s, nL, B, H10buf + B * nn, nn, opC(cL), w.kidx, 0, !fwd);
            copy_blocks(s, nF, B, H00buf, nn, opH00(sf), w.kidx, 0, false);
            copy_blocks(s, nL, B, H00buf + B * nn, nn, opH00(sl), w.kidx, 0, false);
            // Identical leads: the last contact's decimation is the first
            // one with H_01 and H_10 exchanged, so one run of B members
            // carries both surfaces.  Otherwise both contacts go in one batch
            // of 2B members when they share the block size and the overlap is
            // the identity.
            const bool dual = h.leads_same && !sigma;
            const int parts = dual ? 0 : ((nF == nL && !h.S) ? 1 : 2);
            if (dual) {
                SanchoIO io;
                io.n = nF;
                io.members = B;
                io.nn = nn;
                io.zre = w.zre;
                io.kidx = w.kidx;
                io.eta = p.lead_eta;
                io.nmax = p.lead_nmax;
                io.tol = p.lead_tol;
                io.reset = (p.flags & QT_F_LEAD_RESET) != 0;
                io.S = opS(sf);
                io.H00 = op_member(H00buf, nF, nn);
                io.H10 = H10buf;
                io.G00 = G00buf;
                io.G00b = G00buf + B * nn;
                io.iters_out = w.lead_it;
                io.iters2_out = w.lead_it + B;
                io.info_bits = w.infoacc;
                io.period = B;
                sancho_run(c, w, io);
                lead_iterations += io.total_iters;
            }
            for (int part = 0; part < parts; ++part) {
                SanchoIO io;
                io.n = parts == 1 ? nF : (part == 0 ? nF : nL);
                io.members = parts == 1 ? 2 * B : B;
                io.nn = nn;
                io.zre = w.zre + part * B;
                io.kidx = w.kidx + part * B;
                io.eta = p.lead_eta;
                io.nmax = p.lead_nmax;
                io.tol = p.lead_tol;
                io.reset = (p.flags & QT_F_LEAD_RESET) != 0;
                io.S = opS(part == 0 ? sf : sl);
                io.H00 = op_member(H00buf + part * B * nn, io.n, nn);
                io.H10 = H10buf + part * B * nn;
                io.G00 = G00buf + part * B * nn;
                io.iters_out = w.lead_it + part * B;
                io.info_bits = w.infoacc;
                io.period = B;
                sancho_run(c, w, io);
                lead_iterations += io.total_iters;
            }
            // Sigma_first = U G00 U^H, U = H1i(cF)
            g(OP_N, OP_N, nF, nF, nF, opC(cF), op_member(G00buf, nF, nn), T[0], nF, nn);
            g(OP_N, OP_C, nF, nF, nF, op_member(T[0], nF, nn), opC(cF), w.sigF, nF, nn);
            // Sigma_last = V^H G00 V, V = H1i(cL)
            g(OP_C, OP_N, nL, nL, nL, opC(cL), op_member(G00buf + B * nn, nL, nn), T[0], nL, nn);
            g(OP_N, OP_N, nL, nL, nL, op_member(T[0], nL, nn), opC(cL), w.sigL, nL, nn);
        }
        QT_HIP(hipEventRecord(c.ev[1], s));
        // mixed mode: the sweep's block products on the int8 matrix cores (the
        // decimation, its self-energies and every inversion stay in FP64)
        g.prec = prec_digits(p.precision);

        // ---------------- pass 1, first slab (rgf_mod.f90:91-104 / 312-325)
        {
            const int n = nF;
            ShiftArgs sh;
            sh.n = n;
            sh.batch = B;
            sh.out = T[0];
            sh.sout = nn;
            sh.zre = w.zre;
            sh.S = opS(sf);
            sh.M1 = opH00(sf);
            sh.c1 = -1.0;
            sh.M2 = op_member(w.sigF, n, nn);
            sh.c2 = -1.0;
            sh.kidx = w.kidx;
            shifted(s, sh);                                   // z S - H00 - Sigma
            inv_batch(c, w, n, B, T[0], nn, gl(sf), sq(sf), nullptr);
            note_inv_info(c, w, B, B);
            AntiArgs an;                                      // -(Sigma - Sigma^H) f + sl S
            an.n = n;
            an.batch = B;
            an.out = T[1];
            an.sout = nn;
            an.X = op_member(w.sigF, n, nn);
            an.afac = w.nfF;
            an.Y = opSLS(sf);
            an.c1 = 1.0;
            antiherm(s, an);
            g(OP_N, OP_N, n, n, n, opGl(sf), op_member(T[1], n, nn), T[2], n, nn);
            g(OP_N, OP_C, n, n, n, op_member(T[2], n, nn), opGl(sf), gln(sf), n, sq(sf));
        }
        // ---------------- pass 1, interior slabs (rgf_mod.f90:113-136 / 330-353)
        int sp = sf;
        for (int sc = sf + step; sc != sl; sc += step) {
            const int n = h.nm[sc], np = h.nm[sp];
            const ZOp U = opC(fwd ? sc : sc + 1);  // H1i(ii) forward, H1i(ii+1) backward
            g(OP_N, OP_N, n, np, np, U, opGl(sp), T[0], n, nn);          // U Gl
            ShiftArgs sh;
            sh.n = n;
            sh.batch = B;
            sh.out = T[1];
            sh.sout = nn;
            sh.zre = w.zre;
            sh.S = opS(sc);
            sh.M1 = opH00(sc);
            sh.c1 = -1.0;
            sh.kidx = w.kidx;
            shifted(s, sh);                                              // z S - H00
            g(OP_N, OP_C, n, n, np, op_member(T[0], n, nn), U, T[1], n, nn, -1.0,
              op_member(T[1], n, nn), 1.0);                              // - U Gl U^H
            inv_batch(c, w, n, B, T[1], nn, gl(sc), sq(sc), nullptr);
            note_inv_info(c, w, B, B);
            if (sigma) {
                g(OP_N, OP_N, n, np, np, U, opGln(sp), T[2], n, nn);     // U Gln
                g(OP_N, OP_C, n, n, np, op_member(T[2], n, nn), U, T[3], n, nn, 1.0, opSLS(sc),
                  1.0);                                                  // U Gln U^H + sl S
                g(OP_N, OP_N, n, n, n, opGl(sc), op_member(T[3], n, nn), T[4], n, nn);
                g(OP_N, OP_C, n, n, n, op_member(T[4], n, nn), opGl(sc), gln(sc), n, sq(sc));
            } else {
                // no interior injection: Gln = Gl U Gln' U^H Gl^H = V Gln' V^H with
                // V = Gl U, three products instead of four
                g(OP_N, OP_N, n, np, n, opGl(sc), U, T[2], n, nn);       // V = Gl U
                g(OP_N, OP_N, n, np, np, op_member(T[2], n, nn), opGln(sp), T[3], n, nn);
                g(OP_N, OP_C, n, n, np, op_member(T[3], n, nn), op_member(T[2], n, nn), gln(sc), n,
                  sq(sc));                                               // (V Gln') V^H
            }
            sp = sc;
        }
        // ---------------- last slab of pass 1 (rgf_mod.f90:172-202 / 385-415)
        double2* Gc = T[6];   // G of the slab just solved (n x n per point, stride nn)
        double2* Lc = T[7];   // its G^<
        double2* Gn = T[8];
        double2* Ln = T[9];
        {
            const int n = nL, np = h.nm[sp];
            const ZOp U = opC(fwd ? sl : sl + 1);
            g(OP_N, OP_N, n, np, np, U, opGl(sp), T[0], n, nn);
            ShiftArgs sh;
            sh.n = n;
            sh.batch = B;
            sh.out = T[1];
            sh.sout = nn;
            sh.zre = w.zre;
            sh.S = opS(sl);
            sh.M1 = opH00(sl);
            sh.c1 = -1.0;
            sh.M2 = op_member(w.sigL, n, nn);
            sh.c2 = -1.0;
            sh.kidx = w.kidx;
            shifted(s, sh);                                              // z S - H00 - Sigma
            g(OP_N, OP_C, n, n, np, op_member(T[0], n, nn), U, T[1], n, nn, -1.0,
              op_member(T[1], n, nn), 1.0);                              // - U Gl U^H
            inv_batch(c, w, n, B, T[1], nn, Gc, nn, nullptr);
            note_inv_info(c, w, B, B);
            copy_blocks(s, n, B, gl(sl), sq(sl), op_member(Gc, n, nn), nullptr, 0, false);
            g(OP_N, OP_N, n, np, np, U, opGln(sp), T[2], n, nn);
            g(OP_N, OP_C, n, n, np, op_member(T[2], n, nn), U, T[3], n, nn, 1.0, opSLS(sl), 1.0);
            AntiArgs an;  // sig = -(Sigma - Sigma^H) f + U Gln U^H + sl S
            an.n = n;
            an.batch = B;
            an.out = T[4];
            an.sout = nn;
            an.X = op_member(w.sigL, n, nn);
            an.afac = w.nfL;
            an.Y = op_member(T[3], n, nn);
            an.c1 = 1.0;
            antiherm(s, an);
            g(OP_N, OP_N, n, n, n, op_member(Gc, n, nn), op_member(T[4], n, nn), T[5], n, nn);
            g(OP_N, OP_C, n, n, n, op_member(T[5], n, nn), op_member(Gc, n, nn), Lc, n, nn);
            CurrentArgs ca;  // spectral current at the last contact
            ca.n = n;
            ca.batch = B;
            ca.Sig = op_member(w.sigL, n, nn);
            ca.G = op_member(Gc, n, nn);
            ca.Gl = op_member(Lc, n, nn);
            ca.f = w.fL;
            ca.partial = w.partial;
            ca.out = fwd ? w.tr : w.tre;
            ca.scale = gw ? -1.0 : 1.0;  // gw_mod.f90:1245
            contact_current(s, ca);
            diag_out(sl, Gc, Lc);
            if (want_gw && full.GGi1) {
                // G^> of this slab, used by every GGi1 block (gw_mod.f90:1237, 1276)
                AntiArgs gp;
                gp.n = n;
                gp.batch = B;
                gp.out = T[13];
                gp.sout = nn;
                gp.X = op_member(Gc, n, nn);
                gp.aconst = 1.0;
                gp.Y = op_member(Lc, n, nn);
                gp.c1 = 1.0;
                antiherm(s, gp);
            }
            const double2* jd = nullptr;
            if (want_j) {
                // Jdens = Om (G^< + f (G - G^H)),  Om = -(Sigma - Sigma^H)
                AntiArgs om;
                om.n = n;
                om.batch = B;
                om.out = T[0];
                om.sout = nn;
                om.X = op_member(w.sigL, n, nn);
                om.aconst = -1.0;
                antiherm(s, om);
                AntiArgs wv;
                wv.n = n;
                wv.batch = B;
                wv.out = T[1];
                wv.sout = nn;
                wv.X = op_member(Gc, n, nn);
                wv.afac = w.fL;
                wv.Y = op_member(Lc, n, nn);
                wv.c1 = 1.0;
                antiherm(s, wv);
                g(OP_N, OP_N, n, n, n, op_member(T[0], n, nn), op_member(T[1], n, nn), T[10], n, nn);
                jd = T[10];
            }
            full_slab(sl, Gc, Lc, jd);
        }
        QT_HIP(hipEventRecord(c.ev[2], s));

        // ---------------- pass 2 (rgf_mod.f90:212-250 / 420-458)
        int sn = sl;
        for (int sc = sl - step; sc != sf - step; sc -= step) {
            const int n = h.nm[sc], nq = h.nm[sn];
            // coupling between slab sc and its solved neighbour: the reference
            // uses H1i(ii) in both variants; QT_F_COUPLING_FIX selects H1i(ii+1)
            // in the forward variant
            const ZOp W = opC(fwd && fix ? sc + 1 : sc);  // nq x n
            double2 *X = T[0], *GN0 = T[1], *P = T[2], *Y = T[3], *Q = T[4], *U = T[5];
            // Z form: GN0 (W Gl) and GN0 (W Gln) as (GN0 W) Gl and (GN0 W) Gln,
            // three products instead of four, between slabs of equal size (the
            // RGF variants; the GW variant keeps GN0 and Q for its extra
            // blocks).  The choice depends on the device and the variant only,
            // so a point's results do not depend on the outputs requested.
            const bool zform = !gw && nq == n;
            double2* Z = P;  // n x n, ld n
            g(OP_N, OP_C, n, nq, n, opGl(sc), W, X, n, nn);                                   // X = Gl W^H
            g(OP_N, OP_N, n, nq, nq, op_member(X, n, nn), op_member(Gc, nq, nn), GN0, n, nn);  // GN0 = X G
            if (zform) {
                g(OP_N, OP_N, n, n, nq, op_member(GN0, n, nn), W, Z, n, nn);                  // Z = GN0 W
                g(OP_N, OP_N, n, n, n, op_member(Z, n, nn), opGl(sc), Gn, n, nn, 1.0, opGl(sc),
                  1.0);                                                                        // G = Gl + Z Gl
            } else {
                g(OP_N, OP_N, nq, n, n, W, opGl(sc), P, nq, nn);                               // P = W Gl
                g(OP_N, OP_N, n, n, nq, op_member(GN0, n, nn), op_member(P, nq, nn), Gn, n, nn, 1.0,
                  opGl(sc), 1.0);                                                              // G = Gl + GN0 P
            }
            g(OP_N, OP_N, n, nq, nq, op_member(X, n, nn), op_member(Lc, nq, nn), Y, n, nn);    // Y = X G<
            g(OP_N, OP_C, n, n, nq, op_member(Y, n, nn), op_member(X, n, nn), U, n, nn, 1.0,
              opGln(sc), 1.0);                                                                 // U = Gln + Y X^H
            if (!zform || want_j) g(OP_N, OP_N, nq, n, n, W, opGln(sc), Q, nq, nn);           // Q = W Gln
            if (want_gw) gw_extras(sc, n, W, X, Y, Q, GN0, Gc, Lc, g);
            if (!sigma) {
                if (zform)
                    g(OP_N, OP_N, n, n, n, op_member(Z, n, nn), opGln(sc), Y, n, nn);         // Tm = Z Gln
                else
                    g(OP_N, OP_N, n, n, nq, op_member(GN0, n, nn), op_member(Q, nq, nn), Y, n, nn);  // Tm = GN0 Q
                // G< = U + Tm - Tm^H: the reference's last term Gln W^H GN0^H equals
                // -(GN0 W Gln)^H because Gln is anti-Hermitian when every injection
                // is a contact term -(Sigma - Sigma^H) f
                AntiArgs an;
                an.n = n;
                an.batch = B;
                an.out = Ln;
                an.sout = nn;
                an.X = op_member(Y, n, nn);
                an.aconst = 1.0;
                an.Y = op_member(U, n, nn);
                an.c1 = 1.0;
                antiherm(s, an);
            } else {
                // scattering injections (sigma_lesser_ph S, or a Sigma^< that is not
                // anti-Hermitian) break that symmetry: the reference's four terms
                if (zform)
                    g(OP_N, OP_N, n, n, n, op_member(Z, n, nn), opGln(sc), Y, n, nn, 1.0, op_member(U, n, nn),
                      1.0);                                                                      // U + Z Gln
                else
                    g(OP_N, OP_N, n, n, nq, op_member(GN0, n, nn), op_member(Q, nq, nn), Y, n, nn, 1.0,
                      op_member(U, n, nn), 1.0);                                                 // U + GN0 Q
                g(OP_N, OP_C, n, nq, n, opGln(sc), W, P, n, nn);                                  // Gln W^H
                g(OP_N, OP_C, n, n, nq, op_member(P, n, nn), op_member(GN0, n, nn), Ln, n, nn, 1.0,
                  op_member(Y, n, nn), 1.0);                                                     // + Gln W^H GN0^H
            }
            const double2* jd = nullptr;
            if (want_j) {
                // A = G<(nbr) X^H + G(nbr) Q ; Jdens = -2 W^H A  (rgf_mod.f90:215-221)
                g(OP_N, OP_C, nq, n, nq, op_member(Lc, nq, nn), op_member(X, n, nn), U, nq, nn);
                g(OP_N, OP_N, nq, n, nq, op_member(Gc, nq, nn), op_member(Q, nq, nn), U, nq, nn, 1.0,
                  op_member(U, nq, nn), 1.0);
                g(OP_C, OP_N, n, n, nq, W, op_member(U, nq, nn), T[10], n, nn, -2.0);
                jd = T[10];
            }
            std::swap(Gc, Gn);
            std::swap(Lc, Ln);
            diag_out(sc, Gc, Lc);
            full_slab(sc, Gc, Lc, jd);
            sn = sc;
        }
        // ---------------- current at the first contact (rgf_mod.f90:256-266 / 459-469)
        {
            CurrentArgs ca;
            ca.n = nF;
            ca.batch = B;
            ca.Sig = op_member(w.sigF, nF, nn);
            ca.G = op_member(Gc, nF, nn);
            ca.Gl = op_member(Lc, nF, nn);
            ca.f = w.fFc;
            ca.partial = w.partial;
            ca.out = fwd ? w.tre : w.tr;
            contact_current(s, ca);
        }
        // Gl / Gln workspace lists of the reference interface
        if (full.any) {
            for (int sc = 0; sc < nx; ++sc) {
                d2h_blocks(c, full.Gl, full.nx, full.p0, B, sc, h.nm[sc], gl(sc), sq(sc));
                if (sc != sl) d2h_blocks(c, full.Gln, full.nx, full.p0, B, sc, h.nm[sc], gln(sc), sq(sc));
            }
        }
        QT_HIP(hipEventRecord(c.ev[3], s));
        QT_HIP(hipEventSynchronize(c.ev[3]));
        float t01 = 0, t12 = 0, t23 = 0;
        QT_HIP(hipEventElapsedTime(&t01, c.ev[0], c.ev[1]));
        QT_HIP(hipEventElapsedTime(&t12, c.ev[1], c.ev[2]));
        QT_HIP(hipEventElapsedTime(&t23, c.ev[2], c.ev[3]));
        tm.lead_ms += t01;
        tm.forward_ms += t12;
        tm.backward_ms += t23;
    }
};

int status_of(const hip_error& e) {
    g_last_error = e.what();
    return e.code;
}

}  // namespace qt

using namespace qt;

// ====================================================================== C ABI
extern "C" {

const char* qt_status_string(int st) {
    switch (st) {
        case QT_OK: return "ok";
        case QT_ERR_INVALID: return "invalid argument or shape";
        case QT_ERR_NOT_INIT: return "qt_init() has not been called";
        case QT_ERR_HIP: return g_last_error.empty() ? "HIP runtime error" : g_last_error.c_str();
        case QT_ERR_NOMEM: return "device memory exhausted";
        case QT_ERR_SINGULAR: return "singular block inversion (block zeroed)";
        case QT_ERR_UNSUPPORTED: return "unsupported shape or option";
        case QT_ERR_NO_DEVICE: return "no HIP device";
        case QT_ERR_TOO_LARGE:
            return "request exceeds the QT_NOALLOC_SCRATCH_BYTES scratch of the no-allocation entry point";
        default: return "unknown status";
    }
}

int qt_init(int device) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    try {
        if (g_ctx) {
            if (g_ctx->device == device) return QT_OK;
            return QT_ERR_INVALID;
        }
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return QT_ERR_NO_DEVICE;
        if (device < 0 || device >= n) return QT_ERR_INVALID;
        QT_HIP(hipSetDevice(device));
        Ctx* c = new Ctx();
        c->device = device;
        QT_HIP(hipStreamCreate(&c->stream));
        QT_HIP(hipGetSymbolAddress((void**)&c->scratch, HIP_SYMBOL(scratch)));
        if (!c->scratch) throw hip_error("hipGetSymbolAddress(scratch) returned null", QT_ERR_HIP);
        QT_HIP(hipHostMalloc((void**)&c->host_count, sizeof(int), hipHostMallocDefault));
        for (int i = 0; i < 4; ++i) QT_HIP(hipEventCreate(&c->ev[i]));
        g_ctx = c;
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_finish(void) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (!g_ctx) return QT_OK;
    Ctx* c = g_ctx;
    (void)hipStreamSynchronize(c->stream);
    c->arena.release();
    fft_release();
    if (c->timer.created)
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < 4096; ++i) (void)hipEventDestroy(c->timer.ev[k][i]);
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(c->ev[i]);
    (void)hipHostFree(c->host_count);
    if (c->staging) (void)hipHostFree(c->staging);
    (void)hipStreamDestroy(c->stream);
    delete c;
    g_ctx = nullptr;
    return QT_OK;
}

int qt_ham_create(int32_t nx, const int32_t* nm, int32_t nk, const void* const* Hii,
                  const void* const* H1i, const void* const* Sii, qt_ham** out) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (nx < 2 || nk < 1 || !nm || !Hii || !H1i || !out) return QT_ERR_INVALID;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    qt_ham* h = new qt_ham();
    try {
        h->nx = nx;
        h->nk = nk;
        h->nm.assign(nm, nm + nx);
        h->hoff.resize(nx);
        h->coff.resize(nx + 1);
        int64_t o = 0;
        for (int s = 0; s < nx; ++s) {
            if (nm[s] <= 0 || nm[s] > 1024) throw hip_error("block size out of range", QT_ERR_INVALID);
            h->hoff[s] = o;
            o += (int64_t)nm[s] * nm[s];
            h->nmax = std::max(h->nmax, nm[s]);
            if (nm[s] != nm[0]) h->uniform = false;
        }
        h->hset = o;
        o = 0;
        for (int s = 0; s <= nx; ++s) {
            h->coff[s] = o;
            const int r = s < nx ? nm[s] : nm[nx - 1];
            const int cc = s == 0 ? nm[0] : nm[s - 1];
            o += (int64_t)r * cc;
        }
        h->cset = o;
        QT_HIP(hipMalloc((void**)&h->H, sizeof(double2) * h->hset * nk));
        QT_HIP(hipMalloc((void**)&h->C, sizeof(double2) * h->cset * nk));
        if (Sii) QT_HIP(hipMalloc((void**)&h->S, sizeof(double2) * h->hset * nk));
        for (int k = 0; k < nk; ++k) {
            for (int s = 0; s < nx; ++s) {
                const size_t bytes = sizeof(double2) * (size_t)nm[s] * nm[s];
                if (!Hii[(int64_t)k * nx + s]) throw hip_error("null Hii block", QT_ERR_INVALID);
                QT_HIP(hipMemcpy(h->H + k * h->hset + h->hoff[s], Hii[(int64_t)k * nx + s], bytes,
                                 hipMemcpyHostToDevice));
                if (Sii) {
                    if (!Sii[(int64_t)k * nx + s]) throw hip_error("null Sii block", QT_ERR_INVALID);
                    QT_HIP(hipMemcpy(h->S + k * h->hset + h->hoff[s], Sii[(int64_t)k * nx + s],
                                     bytes, hipMemcpyHostToDevice));
                }
            }
            for (int s = 0; s <= nx; ++s) {
                const int r = s < nx ? nm[s] : nm[nx - 1];
                const int cc = s == 0 ? nm[0] : nm[s - 1];
                if (!H1i[(int64_t)k * (nx + 1) + s]) throw hip_error("null H1i block", QT_ERR_INVALID);
                QT_HIP(hipMemcpy(h->C + k * h->cset + h->coff[s], H1i[(int64_t)k * (nx + 1) + s],
                                 sizeof(double2) * (size_t)r * cc, hipMemcpyHostToDevice));
            }
        }
        // both contacts are the same lead when the outer slabs, the outer
        // couplings and the outer overlaps coincide for every k-point
        bool same = nm[0] == nm[nx - 1];
        for (int k = 0; same && k < nk; ++k) {
            const size_t bytes = sizeof(double2) * (size_t)nm[0] * nm[0];
            same = same && !memcmp(Hii[(int64_t)k * nx], Hii[(int64_t)k * nx + nx - 1], bytes);
            same = same && !memcmp(H1i[(int64_t)k * (nx + 1)], H1i[(int64_t)k * (nx + 1) + nx], bytes);
            if (Sii) same = same && !memcmp(Sii[(int64_t)k * nx], Sii[(int64_t)k * nx + nx - 1], bytes);
        }
        h->leads_same = same;
        *out = h;
        return QT_OK;
    } catch (const hip_error& e) {
        if (h->H) (void)hipFree(h->H);
        if (h->C) (void)hipFree(h->C);
        if (h->S) (void)hipFree(h->S);
        delete h;
        return status_of(e);
    }
}

int qt_ham_destroy(qt_ham* h) {
    if (!h) return QT_OK;
    if (g_ctx) {
        std::lock_guard<std::mutex> lk(g_ctx->mu);
        (void)hipStreamSynchronize(g_ctx->stream);
    }
    if (h->H) (void)hipFree(h->H);
    if (h->C) (void)hipFree(h->C);
    if (h->S) (void)hipFree(h->S);
    delete h;
    return QT_OK;
}

static int choose_batch(qt_ham& h, int npts, bool sigma) {
    const char* env = getenv("QT_BATCH");
    if (env && atoi(env) > 0) return std::min(npts, atoi(env));
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) freeb = (size_t)8 << 30;
    const size_t budget = (size_t)(0.80 * (double)(freeb + g_ctx->arena.cap));
    int B = std::min(npts, 128);
    while (B > 1 && work_bytes(h, B, sigma) > budget) B = B * 3 / 4;
    return std::max(B, 1);
}

int qt_rgf_solve(qt_ham* h, const qt_rgf_params* p, int32_t npts, const double* En,
                 const int32_t* kidx, const void* const* sigma_lesser_ph,
                 const void* const* sigma_r_ph, qt_rgf_outputs* out) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!h || !p || npts < 0 || (npts > 0 && !En) || !out) return QT_ERR_INVALID;
    if (npts == 0) return QT_OK;
    if (prec_digits(p->precision) < 0) return QT_ERR_UNSUPPORTED;
    if (p->variant != QT_RGF_FORWARD && p->variant != QT_RGF_BACKWARD && p->variant != QT_RGF_GW)
        return QT_ERR_INVALID;
    const bool gw = p->variant == QT_RGF_GW;
    const bool fwd = p->variant != QT_RGF_BACKWARD;
    const bool fix = (p->flags & QT_F_COUPLING_FIX) != 0;
    if (!h->uniform && (!fwd || !fix || gw)) return QT_ERR_UNSUPPORTED;
    if ((sigma_lesser_ph == nullptr) != (sigma_r_ph == nullptr)) return QT_ERR_INVALID;
    for (int i = 0; i < npts; ++i)
        if (kidx && (kidx[i] < 0 || kidx[i] >= h->nk)) return QT_ERR_INVALID;
    const bool dev_out = (p->flags & QT_F_DEVICE_OUTPUTS) != 0;
    const bool sigma = sigma_r_ph != nullptr;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const int B = choose_batch(*h, npts, sigma);
        c.arena.reserve(work_bytes(*h, B, sigma));
        Work w;
        work_carve(c.arena, w, *h, B, sigma);
        qt_timing tm{};
        tm.batch = B;
        c.timer.reset();
        hipEvent_t e0, e1;
        QT_HIP(hipEventCreate(&e0));
        QT_HIP(hipEventCreate(&e1));
        QT_HIP(hipEventRecord(e0, c.stream));
        std::vector<int64_t> doff(h->nx);
        int64_t od = 0;
        for (int s = 0; s < h->nx; ++s) { doff[s] = od; od += h->nm[s]; }
        const int64_t ndiag = od;
        std::vector<double> hz(2 * B), hf(5 * B);
        std::vector<int> hk(2 * B);
        std::vector<int> hinfo(B), hit(2 * B);
        std::vector<double> htr(B), htre(B);
        int status = QT_OK;
        for (int p0 = 0; p0 < npts; p0 += B) {
            const int nb = std::min(B, npts - p0);
            // pad the batch by repeating its last point (results discarded)
            for (int b = 0; b < B; ++b) {
                const int ip = p0 + std::min(b, nb - 1);
                hz[b] = hz[B + b] = En[ip];
                hk[b] = hk[B + b] = kidx ? kidx[ip] : 0;
                // occupations: rgf_mod.f90:100,182 (forward) / :321,395 (backward):
                // the first contact uses (mur, TEMPr), the last one (mul, TEMPl);
                // gw_mod.f90:1185,1232,1333: left (mul, TEMPl) injected, its
                // current with (mul, TEMPr), right (mur, TEMPl)
                double fF, fFc, fL;
                if (gw) {
                    fF = ferm((En[ip] - p->mul) / (BOLTZ * p->templ));
                    fFc = ferm((En[ip] - p->mul) / (BOLTZ * p->tempr));
                    fL = ferm((En[ip] - p->mur) / (BOLTZ * p->templ));
                } else {
                    fF = fFc = ferm((En[ip] - p->mur) / (BOLTZ * p->tempr));
                    fL = ferm((En[ip] - p->mul) / (BOLTZ * p->templ));
                }
                hf[b] = fF;
                hf[B + b] = fL;
                hf[2 * B + b] = -fF;
                hf[3 * B + b] = -fL;
                hf[4 * B + b] = fFc;
            }
            QT_HIP(hipMemcpyAsync(w.zre, hz.data(), sizeof(double) * 2 * B, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(w.kidx, hk.data(), sizeof(int) * 2 * B, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(w.fF, hf.data(), sizeof(double) * B, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(w.fL, hf.data() + B, sizeof(double) * B, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(w.nfF, hf.data() + 2 * B, sizeof(double) * B, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(w.nfL, hf.data() + 3 * B, sizeof(double) * B, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(w.fFc, hf.data() + 4 * B, sizeof(double) * B, hipMemcpyHostToDevice, c.stream));
            fill_int(c.stream, w.infoacc, B, 0);
            if (sigma) {
                // H00 = Hii + sigma_r_ph * S,  SLS = sigma_lesser_ph * S  (rgf_mod.f90:120-121, 132)
                for (int s = 0; s < h->nx; ++s) {
                    const int n = h->nm[s];
                    const size_t bytes = sizeof(double2) * (size_t)n * n;
                    double2* hr = w.H00 + h->hoff[s] * B;
                    double2* hl = w.SLS + h->hoff[s] * B;
                    for (int b = 0; b < B; ++b) {
                        const int ip = p0 + std::min(b, nb - 1);
                        const void* sr = sigma_r_ph[(int64_t)ip * h->nx + s];
                        const void* sl = sigma_lesser_ph[(int64_t)ip * h->nx + s];
                        QT_HIP(hipMemcpy(w.T[0] + b * w.nn, sr, bytes, hipMemcpyHostToDevice));
                        QT_HIP(hipMemcpy(w.T[1] + b * w.nn, sl, bytes, hipMemcpyHostToDevice));
                    }
                    Gemm g{c.stream, &c.timer, B, w.kidx};
                    ZOp S = h->S ? op_k(h->S + h->hoff[s], n, h->hset) : ZOp();
                    if (S.p) {
                        g(OP_N, OP_N, n, n, n, op_member(w.T[0], n, w.nn), S, w.T[2], n, w.nn);
                        g(OP_N, OP_N, n, n, n, op_member(w.T[1], n, w.nn), S, hl, n, (int64_t)n * n);
                    } else {
                        copy_blocks(c.stream, n, B, hl, (int64_t)n * n, op_member(w.T[1], n, w.nn),
                                    nullptr, 0, false);
                    }
                    ShiftArgs sh;
                    sh.n = n;
                    sh.batch = B;
                    sh.out = hr;
                    sh.sout = (int64_t)n * n;
                    sh.M1 = op_k(h->H + h->hoff[s], n, h->hset);
                    sh.c1 = 1.0;
                    sh.M2 = op_member(S.p ? w.T[2] : w.T[0], n, w.nn);
                    sh.c2 = 1.0;
                    sh.kidx = w.kidx;
                    shifted(c.stream, sh);
                }
            }
            Solve sv{c, *h, *p, w, B, fwd, fix, sigma, gw};
            sv.doff = doff;
            const bool want_diag = out->gr_diag || out->gl_diag || out->gg_diag;
            if (want_diag) {
                if (dev_out) {
                    sv.gr = out->gr_diag ? (double2*)out->gr_diag + (int64_t)p0 * ndiag : w.diag;
                    sv.glo = out->gl_diag ? (double2*)out->gl_diag + (int64_t)p0 * ndiag : w.diag + B * ndiag;
                    sv.gg = out->gg_diag ? (double2*)out->gg_diag + (int64_t)p0 * ndiag : w.diag + 2 * B * ndiag;
                    if (nb < B) {  // padded members must not write past the caller's arrays
                        sv.gr = w.diag;
                        sv.glo = w.diag + B * ndiag;
                        sv.gg = w.diag + 2 * B * ndiag;
                    }
                } else {
                    sv.gr = w.diag;
                    sv.glo = w.diag + B * ndiag;
                    sv.gg = w.diag + 2 * B * ndiag;
                }
            }
            sv.full.any = out->G_r || out->G_lesser || out->G_greater || out->Jdens || out->Gl || out->Gln ||
                          (gw && (out->cur || out->GRi1 || out->GLi1 || out->GGi1));
            if (sv.full.any) {
                if (nb < B) {
                    // full-block outputs are written per real point only
                }
                sv.full.G_r = out->G_r;
                sv.full.G_lesser = out->G_lesser;
                sv.full.G_greater = out->G_greater;
                sv.full.Jdens = out->Jdens;
                sv.full.Gl = out->Gl;
                sv.full.Gln = out->Gln;
                if (gw) {
                    sv.full.cur = out->cur;
                    sv.full.GRi1 = out->GRi1;
                    sv.full.GLi1 = out->GLi1;
                    sv.full.GGi1 = out->GGi1;
                }
                sv.full.nx = h->nx;
                sv.full.p0 = p0;
            }
            if (sv.full.any && nb < B) {
                // restrict host block copies to the real points of this batch
                static thread_local std::vector<void*> bufs[10];
                void* const* src[10] = {out->G_r, out->G_lesser, out->G_greater, out->Jdens, out->Gl, out->Gln,
                                        sv.full.cur, sv.full.GRi1, sv.full.GLi1, sv.full.GGi1};
                void* const** dst[10] = {&sv.full.G_r, &sv.full.G_lesser, &sv.full.G_greater,
                                         &sv.full.Jdens, &sv.full.Gl, &sv.full.Gln, &sv.full.cur,
                                         &sv.full.GRi1, &sv.full.GLi1, &sv.full.GGi1};
                for (int q = 0; q < 10; ++q) {
                    if (!src[q]) continue;
                    bufs[q].assign((size_t)(p0 + B) * h->nx, nullptr);
                    for (int ip = p0; ip < p0 + nb; ++ip)
                        for (int s = 0; s < h->nx; ++s)
                            bufs[q][(size_t)ip * h->nx + s] = src[q][(size_t)ip * h->nx + s];
                    *dst[q] = bufs[q].data();
                }
            }
            sv.run(tm);
            // per-point scalars
            if (dev_out) {
                if (out->tr) QT_HIP(hipMemcpyAsync(out->tr + p0, w.tr, sizeof(double) * nb, hipMemcpyDeviceToDevice, c.stream));
                if (out->tre) QT_HIP(hipMemcpyAsync(out->tre + p0, w.tre, sizeof(double) * nb, hipMemcpyDeviceToDevice, c.stream));
                if (out->info) QT_HIP(hipMemcpyAsync(out->info + p0, w.infoacc, sizeof(int) * nb, hipMemcpyDeviceToDevice, c.stream));
                if (want_diag && nb < B) {
                    if (out->gr_diag) QT_HIP(hipMemcpyAsync((double2*)out->gr_diag + (int64_t)p0 * ndiag, w.diag, sizeof(double2) * nb * ndiag, hipMemcpyDeviceToDevice, c.stream));
                    if (out->gl_diag) QT_HIP(hipMemcpyAsync((double2*)out->gl_diag + (int64_t)p0 * ndiag, w.diag + B * ndiag, sizeof(double2) * nb * ndiag, hipMemcpyDeviceToDevice, c.stream));
                    if (out->gg_diag) QT_HIP(hipMemcpyAsync((double2*)out->gg_diag + (int64_t)p0 * ndiag, w.diag + 2 * B * ndiag, sizeof(double2) * nb * ndiag, hipMemcpyDeviceToDevice, c.stream));
                }
                QT_HIP(hipMemcpyAsync(hinfo.data(), w.infoacc, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
            } else {
                QT_HIP(hipMemcpyAsync(htr.data(), w.tr, sizeof(double) * B, hipMemcpyDeviceToHost, c.stream));
                QT_HIP(hipMemcpyAsync(htre.data(), w.tre, sizeof(double) * B, hipMemcpyDeviceToHost, c.stream));
                QT_HIP(hipMemcpyAsync(hinfo.data(), w.infoacc, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
                if (out->gr_diag) QT_HIP(hipMemcpyAsync((double2*)out->gr_diag + (int64_t)p0 * ndiag, w.diag, sizeof(double2) * nb * ndiag, hipMemcpyDeviceToHost, c.stream));
                if (out->gl_diag) QT_HIP(hipMemcpyAsync((double2*)out->gl_diag + (int64_t)p0 * ndiag, w.diag + B * ndiag, sizeof(double2) * nb * ndiag, hipMemcpyDeviceToHost, c.stream));
                if (out->gg_diag) QT_HIP(hipMemcpyAsync((double2*)out->gg_diag + (int64_t)p0 * ndiag, w.diag + 2 * B * ndiag, sizeof(double2) * nb * ndiag, hipMemcpyDeviceToHost, c.stream));
            }
            QT_HIP(hipMemcpyAsync(hit.data(), w.lead_it, sizeof(int) * 2 * B, hipMemcpyDeviceToHost, c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
            for (int b = 0; b < nb; ++b) {
                if (!dev_out) {
                    if (out->tr) out->tr[p0 + b] = htr[b];
                    if (out->tre) out->tre[p0 + b] = htre[b];
                    if (out->info) out->info[p0 + b] = hinfo[b];
                }
                if (out->lead_iters && !dev_out) {
                    out->lead_iters[2 * (p0 + b)] = hit[b];
                    out->lead_iters[2 * (p0 + b) + 1] = hit[B + b];
                }
                if (hinfo[b] & QT_INFO_SINGULAR) status = QT_ERR_SINGULAR;
                tm.lead_iterations += hit[b] + hit[B + b];  // per lead, as the reference runs them
            }
        }
        QT_HIP(hipEventRecord(e1, c.stream));
        QT_HIP(hipEventSynchronize(e1));
        float tot = 0;
        QT_HIP(hipEventElapsedTime(&tot, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c.timer.flush();
        tm.total_ms = tot;
        tm.gemm_ms = c.timer.ms[0];
        tm.gemm_flops = c.timer.flops[0];
        tm.gemm_launches = c.timer.launches[0];
        tm.inv_ms = c.timer.ms[1];
        tm.inv_flops = c.timer.flops[1];
        tm.inv_launches = c.timer.launches[1];
        c.last = tm;
        return status;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_sancho(int32_t nm, int32_t npts, const double* En, double eta, int32_t nmax, double tol,
              int32_t reset_on_fail, const void* S00, const void* const* H00, const void* const* H10,
              void* const* G00, void* const* GBB, int32_t* iters, int32_t* converged) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (nm <= 0 || nm > 1024 || npts < 0 || !En || !H00 || !H10 || !G00) return QT_ERR_INVALID;
    if (npts == 0) return QT_OK;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        // a one-slab pseudo Hamiltonian sizes the workspace
        qt_ham h;
        h.nx = 1;
        h.nk = 1;
        h.nm = {nm};
        h.hoff = {0};
        h.hset = (int64_t)nm * nm;
        h.nmax = nm;
        int B = std::min(npts, 256);
        size_t freeb = 0, total = 0;
        (void)hipMemGetInfo(&freeb, &total);
        while (B > 1 && work_bytes(h, (B + 1) / 2, false) > 0.8 * (freeb + c.arena.cap)) B = B * 3 / 4;
        const int half = (B + 1) / 2;  // the workspace holds 2*half members
        // distinct H00 / H10 blocks of a batch (points usually share the
        // blocks of a few k-points) are uploaded once into UCAP staging slots
        // each and spread over the members on the device
        const int UCAP = std::min(B, 64);
        c.arena.reserve(work_bytes(h, half, false) + sizeof(double2) * (size_t)nm * nm * (2 * half + 2) +
                        sizeof(double2) * (size_t)nm * nm * 2 * UCAP + sizeof(int) * 2 * (size_t)B + 1024);
        Work w;
        work_carve(c.arena, w, h, half, false);
        double2* Sdev = S00 ? c.arena.take<double2>((size_t)nm * nm) : nullptr;
        double2* GBBdev = GBB ? c.arena.take<double2>((size_t)nm * nm * 2 * half) : nullptr;
        double2* U00 = c.arena.take<double2>((size_t)nm * nm * UCAP);
        double2* U10 = c.arena.take<double2>((size_t)nm * nm * UCAP);
        int* ukx = c.arena.take<int>(2 * (size_t)B);
        if (Sdev) QT_HIP(hipMemcpy(Sdev, S00, sizeof(double2) * nm * nm, hipMemcpyHostToDevice));
        const int64_t nn = (int64_t)nm * nm;
        std::vector<double> hz(2 * half);
        std::vector<int> hit(2 * half), hact(2 * half);
        c.timer.reset();
        qt_timing tm{};
        tm.batch = B;
        for (int p0 = 0; p0 < npts; p0 += B) {
            const int nb = std::min(B, npts - p0);
            std::unordered_map<const void*, int> u0, u1;
            std::vector<int> k01(2 * (size_t)nb);
            std::vector<const void*> l0, l1;
            for (int m = 0; m < nb; ++m) {
                hz[m] = En[p0 + m];
                auto i0 = u0.emplace(H00[p0 + m], (int)l0.size());
                if (i0.second) l0.push_back(H00[p0 + m]);
                auto i1 = u1.emplace(H10[p0 + m], (int)l1.size());
                if (i1.second) l1.push_back(H10[p0 + m]);
                k01[m] = i0.first->second;
                k01[nb + m] = i1.first->second;
            }
            if ((int)l0.size() <= UCAP && (int)l1.size() <= UCAP && ((int)l0.size() < nb || (int)l1.size() < nb)) {
                for (size_t i = 0; i < l0.size(); ++i)
                    QT_HIP(hipMemcpy(U00 + i * nn, l0[i], sizeof(double2) * nn, hipMemcpyHostToDevice));
                for (size_t i = 0; i < l1.size(); ++i)
                    QT_HIP(hipMemcpy(U10 + i * nn, l1[i], sizeof(double2) * nn, hipMemcpyHostToDevice));
                QT_HIP(hipMemcpy(ukx, k01.data(), sizeof(int) * 2 * nb, hipMemcpyHostToDevice));
                ZOp s0, s1;
                s0.p = U00;
                s0.ld = nm;
                s0.sk = nn;
                s1 = s0;
                s1.p = U10;
                copy_blocks(c.stream, nm, nb, w.T[13], nn, s0, ukx, 0, false);
                copy_blocks(c.stream, nm, nb, w.T[11], nn, s1, ukx + nb, 0, false);
            } else {
                for (int m = 0; m < nb; ++m) {
                    QT_HIP(hipMemcpy(w.T[13] + m * nn, H00[p0 + m], sizeof(double2) * nn, hipMemcpyHostToDevice));
                    QT_HIP(hipMemcpy(w.T[11] + m * nn, H10[p0 + m], sizeof(double2) * nn, hipMemcpyHostToDevice));
                }
            }
            QT_HIP(hipMemcpy(w.zre, hz.data(), sizeof(double) * nb, hipMemcpyHostToDevice));
            SanchoIO io;
            io.n = nm;
            io.members = nb;
            io.nn = nn;
            io.zre = w.zre;
            io.kidx = nullptr;
            io.eta = eta;
            io.nmax = nmax;
            io.tol = tol;
            io.reset = reset_on_fail != 0;
            io.S = Sdev ? op_k(Sdev, nm, 0) : ZOp();
            io.H00 = op_member(w.T[13], nm, nn);
            io.H10 = w.T[11];
            io.G00 = w.T[12];
            io.GBB = GBBdev;
            io.iters_out = w.lead_it;
            sancho_run(c, w, io);
            QT_HIP(hipStreamSynchronize(c.stream));
            QT_HIP(hipMemcpy(hit.data(), w.lead_it, sizeof(int) * nb, hipMemcpyDeviceToHost));
            QT_HIP(hipMemcpy(hact.data(), w.active, sizeof(int) * nb, hipMemcpyDeviceToHost));
            // one copy of the batch's blocks into pinned memory, then to the
            // caller's blocks (one pageable copy per block is several times slower)
            const size_t bbytes = sizeof(double2) * (size_t)nn;
            char* stg = c.host_staging(bbytes * nb * (GBB ? 2 : 1));
            QT_HIP(hipMemcpy(stg, w.T[12], bbytes * nb, hipMemcpyDeviceToHost));
            if (GBB) QT_HIP(hipMemcpy(stg + bbytes * nb, GBBdev, bbytes * nb, hipMemcpyDeviceToHost));
            for (int m = 0; m < nb; ++m) {
                std::memcpy(G00[p0 + m], stg + bbytes * m, bbytes);
                if (GBB && GBB[p0 + m]) std::memcpy(GBB[p0 + m], stg + bbytes * (nb + m), bbytes);
                if (iters) iters[p0 + m] = hit[m];
                if (converged) converged[p0 + m] = hact[m] ? 0 : 1;
                tm.lead_iterations += hit[m];
            }
        }
        // kernel timings of this call for qt_last_timing (GEMM and inversion launches)
        c.timer.flush();
        tm.gemm_ms = c.timer.ms[0];
        tm.gemm_flops = c.timer.flops[0];
        tm.gemm_launches = c.timer.launches[0];
        tm.inv_ms = c.timer.ms[1];
        tm.inv_flops = c.timer.flops[1];
        tm.inv_launches = c.timer.launches[1];
        c.last = tm;
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_zgemm_batched(int32_t opa, int32_t opb, int32_t m, int32_t n, int32_t k, int32_t batch,
                     double alpha, const void* A, int64_t lda, int64_t strideA, const void* B,
                     int64_t ldb, int64_t strideB, double beta, void* R, int64_t ldr,
                     int64_t strideR) {
    return qt_zgemm_batched_prec(opa, opb, m, n, k, batch, alpha, A, lda, strideA, B, ldb, strideB, beta, R,
                                 ldr, strideR, QT_PREC_FP64);
}

int qt_zgemm_batched_prec(int32_t opa, int32_t opb, int32_t m, int32_t n, int32_t k, int32_t batch,
                          double alpha, const void* A, int64_t lda, int64_t strideA, const void* B,
                          int64_t ldb, int64_t strideB, double beta, void* R, int64_t ldr,
                          int64_t strideR, int32_t precision) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (prec_digits(precision) < 0) return QT_ERR_INVALID;
    if (m <= 0 || n <= 0 || k < 0 || batch <= 0 || !A || !B || !R) return QT_ERR_INVALID;
    if (opa < 0 || opa > 2 || opb < 0 || opb > 2) return QT_ERR_INVALID;
    const int64_t ar = opa == OP_N ? m : k, ac = opa == OP_N ? k : m;
    const int64_t br = opb == OP_N ? k : n, bc = opb == OP_N ? n : k;
    if (lda < ar || ldb < br || ldr < m) return QT_ERR_INVALID;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        // extents of the three operands (elements)
        const int64_t ea = (batch - 1) * strideA + lda * (ac - 1) + ar;
        const int64_t eb = (batch - 1) * strideB + ldb * (bc - 1) + br;
        const int64_t er = (batch - 1) * strideR + ldr * (n - 1) + m;
        c.arena.reserve(sizeof(double2) * (size_t)(ea + eb + er) + 4096);
        c.arena.reset();
        double2* dA = c.arena.take<double2>(ea);
        double2* dB = c.arena.take<double2>(eb);
        double2* dR = c.arena.take<double2>(er);
        QT_HIP(hipMemcpy(dA, A, sizeof(double2) * ea, hipMemcpyHostToDevice));
        QT_HIP(hipMemcpy(dB, B, sizeof(double2) * eb, hipMemcpyHostToDevice));
        if (beta != 0.0) QT_HIP(hipMemcpy(dR, R, sizeof(double2) * er, hipMemcpyHostToDevice));
        ZGemmArgs g;
        g.M = m;
        g.N = n;
        g.K = k;
        g.batch = batch;
        g.A = op_member(dA, lda, strideA);
        g.B = op_member(dB, ldb, strideB);
        g.C = dR;
        g.ldc = ldr;
        g.sbc = strideR;
        if (beta != 0.0) g.D = op_member(dR, ldr, strideR);
        g.alpha = alpha;
        g.beta = beta;
        g.prec = prec_digits(precision);
        zgemm(c.stream, opa, opb, g, &c.timer);
        QT_HIP(hipStreamSynchronize(c.stream));
        QT_HIP(hipMemcpy(R, dR, sizeof(double2) * er, hipMemcpyDeviceToHost));
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_invert_batched(int32_t n, int32_t batch, void* A, int32_t* info) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (n <= 0 || n > 1024 || batch <= 0 || !A) return QT_ERR_INVALID;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const int64_t nn = (int64_t)n * n;
        c.arena.reserve(sizeof(double2) * (size_t)(2 * nn * batch) +
                        sizeof(double2) * (size_t)n * kInvPanelMax * batch +
                        sizeof(int) * (size_t)(n + 1) * batch + 8192);
        c.arena.reset();
        Work w;
        double2* dA = c.arena.take<double2>(nn * batch);
        double2* dX = c.arena.take<double2>(nn * batch);
        w.piv = c.arena.take<int>((int64_t)n * batch);
        w.ginv = c.arena.take<double2>((int64_t)n * kInvPanelMax * batch);
        w.info = c.arena.take<int>(batch);
        QT_HIP(hipMemcpy(dA, A, sizeof(double2) * nn * batch, hipMemcpyHostToDevice));
        inv_batch(c, w, n, batch, dA, nn, dX, nn, nullptr);
        QT_HIP(hipStreamSynchronize(c.stream));
        QT_HIP(hipMemcpy(A, dX, sizeof(double2) * nn * batch, hipMemcpyDeviceToHost));
        std::vector<int> hinfo(batch);
        QT_HIP(hipMemcpy(hinfo.data(), w.info, sizeof(int) * batch, hipMemcpyDeviceToHost));
        int st = QT_OK;
        for (int b = 0; b < batch; ++b) {
            if (info) info[b] = hinfo[b];
            if (hinfo[b]) st = QT_ERR_SINGULAR;
        }
        return st;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

// flag[0] |= (info[b] != 0) over the batch
__global__ void info_any_kernel(const int* info, int batch, int* flag) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b < batch && info[b] != 0) atomicOr(flag, 1);
}

// bytes of scratch one call needs: status word (1 KiB), info, pivots, and the
// panel multipliers (n x kInvPanelMax complex per block, 256-byte aligned)
static int64_t noalloc_bytes(int64_t n, int64_t batch) {
    return 1024 + 4 * batch + 4 * n * batch + 256 + (int64_t)sizeof(double2) * kInvPanelMax * n * batch;
}

int qt_invert_batched_noalloc(int32_t n, int32_t batch, void* A, void* X, int32_t* info) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (n <= 0 || n > 1024 || batch <= 0 || !A || !X) return QT_ERR_INVALID;
    if (noalloc_bytes(n, batch) > (int64_t)QT_NOALLOC_SCRATCH_BYTES) return QT_ERR_TOO_LARGE;
    const int64_t nn = (int64_t)n * n;
    const char *a0 = (const char*)A, *x0 = (const char*)X;
    const int64_t bytes = (int64_t)sizeof(double2) * nn * batch;
    if (a0 < x0 + bytes && x0 < a0 + bytes) return QT_ERR_INVALID;  // A and X overlap
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        char* base = (char*)c.scratch;
        int* flag = (int*)base;
        int* sinfo = info ? info : (int*)(base + 1024);
        int* piv = (int*)(base + 1024 + 4 * (int64_t)batch);
        const int64_t goff = (1024 + 4 * (int64_t)batch + 4 * (int64_t)n * batch + 255) & ~(int64_t)255;
        double2* G = (double2*)(base + goff);  // ends within noalloc_bytes(n, batch)
        fill_int(c.stream, flag, 1, 0);  // the array keeps the last call's contents
        InvArgs a;
        a.n = n;
        a.batch = batch;
        a.A = (double2*)A;
        a.sa = nn;
        a.X = (double2*)X;
        a.sx = nn;
        a.piv = piv;
        a.info = sinfo;  // cleared by the inversion itself
        a.G = G;
        a.sg = (int64_t)n * kInvPanelMax;
        zinv_single(c.stream, a);
        hipLaunchKernelGGL(info_any_kernel, dim3((batch + 255) / 256), dim3(256), 0, c.stream, sinfo, batch,
                           flag);
        QT_HIP(hipGetLastError());
        QT_HIP(hipMemcpyAsync(c.host_count, flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        QT_HIP(hipStreamSynchronize(c.stream));
        return *c.host_count ? QT_ERR_SINGULAR : QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

static int corr_conv_entry(int mode, int32_t n, int64_t nseries, const void* X, const void* Y,
                           void* Z, int32_t flags) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (n <= 0 || nseries < 0 || (nseries > 0 && (!X || !Y || !Z))) return QT_ERR_INVALID;
    if (n > 2048) return QT_ERR_UNSUPPORTED;
    if (nseries == 0) return QT_OK;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        if (flags & QT_F_DEVICE_OUTPUTS) {
            corr_conv(c.stream, n, mode, (const double2*)X, (const double2*)Y, (double2*)Z, nseries);
            QT_HIP(hipStreamSynchronize(c.stream));
            return QT_OK;
        }
        // host buffers: staged through the arena in chunks of at most 256 MiB
        const int64_t per = (int64_t)n * 3 * (int64_t)sizeof(double2);
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nseries, (256ll << 20) / per));
        c.arena.reserve((size_t)(chunk * per) + 4096);
        c.arena.reset();
        double2* dX = c.arena.take<double2>(chunk * n);
        double2* dY = c.arena.take<double2>(chunk * n);
        double2* dZ = c.arena.take<double2>(chunk * n);
        for (int64_t s0 = 0; s0 < nseries; s0 += chunk) {
            const int64_t cnt = std::min(chunk, nseries - s0);
            const size_t bytes = sizeof(double2) * (size_t)(cnt * n);
            QT_HIP(hipMemcpy(dX, (const double2*)X + s0 * n, bytes, hipMemcpyHostToDevice));
            QT_HIP(hipMemcpy(dY, (const double2*)Y + s0 * n, bytes, hipMemcpyHostToDevice));
            corr_conv(c.stream, n, mode, dX, dY, dZ, cnt);
            QT_HIP(hipStreamSynchronize(c.stream));
            QT_HIP(hipMemcpy((double2*)Z + s0 * n, dZ, bytes, hipMemcpyDeviceToHost));
        }
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_corr1d2_batched(int32_t n, int64_t nseries, const void* X, const void* Y, void* Z,
                       int32_t flags) {
    return corr_conv_entry(0, n, nseries, X, Y, Z, flags);
}

int qt_conv1d2_batched(int32_t n, int64_t nseries, const void* X, const void* Y, void* Z,
                       int32_t flags) {
    return corr_conv_entry(1, n, nseries, X, Y, Z, flags);
}

// in-band elements |row - col| <= ndiag (all when upper_only == false, else
// row <= col), as ij = row + col*nm
static std::vector<int> band_elements(int nm, int ndiag, bool upper_only) {
    std::vector<int> v;
    for (int col = 0; col < nm; ++col)
        for (int row = 0; row < nm; ++row) {
            if (std::abs(row - col) > ndiag) continue;
            if (upper_only && row > col) continue;
            v.push_back(row + col * nm);
        }
    return v;
}

// device working set allowed to a chunked entry point: 60% of the free
// memory plus what the arena already holds, or QT_WORKSET_BYTES when set
static size_t workset_budget(const Ctx& c) {
    const char* env = getenv("QT_WORKSET_BYTES");
    if (env && atoll(env) > 0) return (size_t)atoll(env);
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) freeb = (size_t)8 << 30;
    return (size_t)(0.6 * (double)(freeb + c.arena.cap));
}

// element-major (nm*nm, nx, nk, nen) arrays: slabs [x0, x0 + cx) of every
// element between the caller's host array and a compact device chunk
static void slabs_h2d(Ctx& c, double2* dst, const void* src, int64_t nij, int nx, int x0, int cx,
                      int64_t per_slab) {
    const size_t w = sizeof(double2) * (size_t)(per_slab * cx), sp = sizeof(double2) * (size_t)(per_slab * nx);
    QT_HIP(hipMemcpy2DAsync(dst, w, (const double2*)src + per_slab * x0, sp, w, (size_t)nij, hipMemcpyHostToDevice,
                            c.stream));
}

static void slabs_d2h(Ctx& c, void* dst, const double2* src, int64_t nij, int nx, int x0, int cx, int64_t per_slab) {
    const size_t w = sizeof(double2) * (size_t)(per_slab * cx), dp = sizeof(double2) * (size_t)(per_slab * nx);
    QT_HIP(hipMemcpy2DAsync((double2*)dst + per_slab * x0, dp, src, w, w, (size_t)nij, hipMemcpyDeviceToHost,
                            c.stream));
}

// slabs per chunk of a host-path GW call holding `arrays` chunk arrays
static int gw_slab_chunk(const Ctx& c, int64_t nij, int nx, int64_t per_slab, int arrays, size_t fixed) {
    const size_t per = sizeof(double2) * (size_t)(nij * per_slab) * (size_t)arrays + 256 * (size_t)arrays;
    const size_t budget = workset_budget(c);
    const int64_t cx = budget > fixed + per ? (int64_t)((budget - fixed) / per) : 1;
    return (int)std::min<int64_t>(nx, std::max<int64_t>(1, cx));
}

static bool gw_shape_ok(int32_t nen, int32_t nky, int32_t nkz, int32_t nx, int32_t nm, int32_t ndiag) {
    if (nen <= 0 || nky <= 0 || nkz <= 0 || nx <= 0 || nm <= 0 || ndiag < 0) return false;
    const int64_t nij = (int64_t)nm * nm;
    return nij <= 0x7fffffff && nij * nx <= 0x7fffffff;
}

int qt_gw_polarization(int32_t nen, int32_t nky, int32_t nkz, int32_t nx, int32_t nm, int32_t ndiag,
                       double dEn, double spindeg, const void* Gr, const void* Gl, const void* Gg,
                       void* Pr, void* Pl, void* Pg, int32_t flags) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!gw_shape_ok(nen, nky, nkz, nx, nm, ndiag) || !Gr || !Gl || !Gg || !Pr || !Pl || !Pg)
        return QT_ERR_INVALID;
    if (nen > 2048) return QT_ERR_UNSUPPORTED;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const std::vector<int> el = band_elements(nm, ndiag, false);
        const std::vector<int> pr = band_elements(nm, ndiag, true);
        const int64_t nij = (int64_t)nm * nm;
        const int64_t per_slab = (int64_t)nky * nkz * nen;
        const bool dev = (flags & QT_F_DEVICE_OUTPUTS) != 0;
        const size_t fixed = sizeof(int) * (el.size() + pr.size()) + 8192;
        // host path: slabs in chunks (every element of a slab range: 3 G and 3 P)
        const int cx = dev ? nx : gw_slab_chunk(c, nij, nx, per_slab, 6, fixed);
        const int64_t chunk = nij * cx * per_slab;
        c.arena.reserve(fixed + (dev ? 0 : 6 * (sizeof(double2) * (size_t)chunk + 256)));
        c.arena.reset();
        int* del = c.arena.take<int>((int64_t)el.size());
        int* dpr = c.arena.take<int>((int64_t)pr.size());
        QT_HIP(hipMemcpyAsync(del, el.data(), sizeof(int) * el.size(), hipMemcpyHostToDevice, c.stream));
        QT_HIP(hipMemcpyAsync(dpr, pr.data(), sizeof(int) * pr.size(), hipMemcpyHostToDevice, c.stream));
        GwArgs a;
        a.nen = nen;
        a.nky = nky;
        a.nkz = nkz;
        a.nm = nm;
        a.dEn = dEn;
        a.spindeg = spindeg;
        if (dev) {
            a.nx = nx;
            a.Gr = (const double2*)Gr;
            a.Gl = (const double2*)Gl;
            a.Gg = (const double2*)Gg;
            a.Pr = (double2*)Pr;
            a.Pl = (double2*)Pl;
            a.Pg = (double2*)Pg;
            const size_t bytes = sizeof(double2) * (size_t)(nij * nx * per_slab);
            if (ndiag < nm - 1)  // out-of-band series stay zero; in-band ones are written whole
                for (double2* p : {a.Pr, a.Pl, a.Pg}) QT_HIP(hipMemsetAsync(p, 0, bytes, c.stream));
            gw_polarization(c.stream, a, del, (int)el.size(), dpr, (int)pr.size());
            QT_HIP(hipStreamSynchronize(c.stream));
            return QT_OK;
        }
        double2* g[3];
        double2* pc[3];
        for (auto& p : g) p = c.arena.take<double2>(chunk);
        for (auto& p : pc) p = c.arena.take<double2>(chunk);
        const void* src[3] = {Gr, Gl, Gg};
        void* dst[3] = {Pr, Pl, Pg};
        for (int x0 = 0; x0 < nx; x0 += cx) {
            const int cnt = std::min(cx, nx - x0);
            const size_t bytes = sizeof(double2) * (size_t)(nij * cnt * per_slab);
            for (int i = 0; i < 3; ++i) {
                slabs_h2d(c, g[i], src[i], nij, nx, x0, cnt, per_slab);
                if (ndiag < nm - 1) QT_HIP(hipMemsetAsync(pc[i], 0, bytes, c.stream));
            }
            a.nx = cnt;
            a.Gr = g[0];
            a.Gl = g[1];
            a.Gg = g[2];
            a.Pr = pc[0];
            a.Pl = pc[1];
            a.Pg = pc[2];
            gw_polarization(c.stream, a, del, (int)el.size(), dpr, (int)pr.size());
            for (int i = 0; i < 3; ++i) slabs_d2h(c, dst[i], pc[i], nij, nx, x0, cnt, per_slab);
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_gw_selfenergy(int32_t nen, int32_t nky, int32_t nkz, int32_t nx, int32_t nm, int32_t ndiag,
                     double dEn, const void* Gr, const void* Gl, const void* Gg, const void* Wr,
                     const void* Wl, const void* Wg, void* Sr, void* Sl, void* Sg, int32_t flags) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!gw_shape_ok(nen, nky, nkz, nx, nm, ndiag) || !Gr || !Gl || !Gg || !Wr || !Wl || !Wg || !Sr ||
        !Sl || !Sg)
        return QT_ERR_INVALID;
    if (nen > 2048) return QT_ERR_UNSUPPORTED;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const std::vector<int> pr = band_elements(nm, ndiag, true);
        const int64_t nij = (int64_t)nm * nm;
        const int64_t per_slab = (int64_t)nky * nkz * nen;
        const bool dev = (flags & QT_F_DEVICE_OUTPUTS) != 0;
        const size_t fixed = sizeof(int) * pr.size() + 8192;
        // host path: slabs in chunks (6 inputs and 3 outputs of a slab range)
        const int cx = dev ? nx : gw_slab_chunk(c, nij, nx, per_slab, 9, fixed);
        const int64_t chunk = nij * cx * per_slab;
        c.arena.reserve(fixed + (dev ? 0 : 9 * (sizeof(double2) * (size_t)chunk + 256)));
        c.arena.reset();
        int* dpr = c.arena.take<int>((int64_t)pr.size());
        QT_HIP(hipMemcpyAsync(dpr, pr.data(), sizeof(int) * pr.size(), hipMemcpyHostToDevice, c.stream));
        GwArgs a;
        a.nen = nen;
        a.nky = nky;
        a.nkz = nkz;
        a.nm = nm;
        a.dEn = dEn;
        if (dev) {
            a.nx = nx;
            a.Gr = (const double2*)Gr;
            a.Gl = (const double2*)Gl;
            a.Gg = (const double2*)Gg;
            a.Wr = (const double2*)Wr;
            a.Wl = (const double2*)Wl;
            a.Wg = (const double2*)Wg;
            a.Sr = (double2*)Sr;
            a.Sl = (double2*)Sl;
            a.Sg = (double2*)Sg;
            const size_t bytes = sizeof(double2) * (size_t)(nij * nx * per_slab);
            if (ndiag < nm - 1)
                for (double2* p : {a.Sr, a.Sl, a.Sg}) QT_HIP(hipMemsetAsync(p, 0, bytes, c.stream));
            gw_selfenergy(c.stream, a, dpr, (int)pr.size());
            QT_HIP(hipStreamSynchronize(c.stream));
            return QT_OK;
        }
        double2* in[6];
        double2* sc[3];
        for (auto& p : in) p = c.arena.take<double2>(chunk);
        for (auto& p : sc) p = c.arena.take<double2>(chunk);
        const void* src[6] = {Gr, Gl, Gg, Wr, Wl, Wg};
        void* dst[3] = {Sr, Sl, Sg};
        for (int x0 = 0; x0 < nx; x0 += cx) {
            const int cnt = std::min(cx, nx - x0);
            const size_t bytes = sizeof(double2) * (size_t)(nij * cnt * per_slab);
            for (int i = 0; i < 6; ++i) slabs_h2d(c, in[i], src[i], nij, nx, x0, cnt, per_slab);
            if (ndiag < nm - 1)
                for (double2* p : sc) QT_HIP(hipMemsetAsync(p, 0, bytes, c.stream));
            a.nx = cnt;
            a.Gr = in[0];
            a.Gl = in[1];
            a.Gg = in[2];
            a.Wr = in[3];
            a.Wl = in[4];
            a.Wg = in[5];
            a.Sr = sc[0];
            a.Sl = sc[1];
            a.Sg = sc[2];
            gw_selfenergy(c.stream, a, dpr, (int)pr.size());
            for (int i = 0; i < 3; ++i) slabs_d2h(c, dst[i], sc[i], nij, nx, x0, cnt, per_slab);
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_gw_mix(int64_t n, double alpha, void* S, const void* Snew, int32_t flags) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (n < 0 || (n > 0 && (!S || !Snew))) return QT_ERR_INVALID;
    if (n == 0) return QT_OK;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        if (flags & QT_F_DEVICE_OUTPUTS) {
            gw_mix(c.stream, n, alpha, (double2*)S, (const double2*)Snew);
            QT_HIP(hipStreamSynchronize(c.stream));
            return QT_OK;
        }
        const size_t bytes = sizeof(double2) * (size_t)n;
        c.arena.reserve(2 * bytes + 4096);
        c.arena.reset();
        double2* dS = c.arena.take<double2>(n);
        double2* dN = c.arena.take<double2>(n);
        QT_HIP(hipMemcpyAsync(dS, S, bytes, hipMemcpyHostToDevice, c.stream));
        QT_HIP(hipMemcpyAsync(dN, Snew, bytes, hipMemcpyHostToDevice, c.stream));
        gw_mix(c.stream, n, alpha, dS, dN);
        QT_HIP(hipStreamSynchronize(c.stream));
        QT_HIP(hipMemcpy(S, dS, bytes, hipMemcpyDeviceToHost));
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

// ---------------------------------------------------------------- e-photon
// calc_sigma_ephoton_monochromatic_ext (gw_mod.f90:1075-1122): per slab ix
// and energy ie, Sigma^<> = f M_ix [G^<>(ie) + G^<>(ie + 2 nop)] M_ix with
// f = pre_fact intensity / hw^2, then Sigma^R = i Im(Sigma^> - Sigma^<) / 2.
// Members b = ix + length ie are consecutive nm x nm blocks; M_ix is picked
// per member through the GEMM's index table (kidx[b] = b mod length), so each
// product is one launch over all length x nen members.
int qt_gw_sigma_ephoton(int32_t nm, int32_t length, int32_t nen, int32_t nop, double pre_fact,
                        double intensity, double hw, const void* Mii, const void* Gl, const void* Gg,
                        void* Sr, void* Sl, void* Sg, int32_t flags) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (nm <= 0 || length <= 0 || nen < 0 || nop < 0 || !(hw != 0.0)) return QT_ERR_INVALID;
    if (nen > 0 && (!Mii || !Gl || !Gg || !Sr || !Sl || !Sg)) return QT_ERR_INVALID;
    if ((int64_t)length * nen > 0x7fffffff) return QT_ERR_UNSUPPORTED;
    if (nen == 0) return QT_OK;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const bool dev = (flags & QT_F_DEVICE_OUTPUTS) != 0;
        const int64_t nn = (int64_t)nm * nm;
        const int64_t per_e = nn * length;  // samples of one array at one energy
        const size_t z = sizeof(double2);
        // members (slab ix, energy e) are independent: energies go in chunks
        // sized to the working-set budget.  Per output energy the device holds
        // T, and on the host path also G^<, G^> and the three Sigma; fixed are
        // the 2 nop extra energies of each G and M.
        const size_t per = z * (size_t)per_e * (dev ? 1 : 6) + sizeof(int) * (size_t)length + 1024;
        const size_t fixed = (dev ? 0 : z * (size_t)(per_e * 4 * (int64_t)nop + nn * length)) + 8192;
        const size_t budget = workset_budget(c);
        const int ce = (int)std::min<int64_t>(nen, std::max<int64_t>(1, budget > fixed + per
                                                                             ? (int64_t)((budget - fixed) / per)
                                                                             : 1));
        c.arena.reserve(fixed + per * (size_t)ce);
        c.arena.reset();
        const int64_t ns_max = per_e * ce;
        double2* T = c.arena.take<double2>(ns_max);
        int* kidx = c.arena.take<int>((int64_t)length * ce);
        const double2* dM = (const double2*)Mii;
        double2 *gl = nullptr, *gg = nullptr, *hs[3] = {nullptr, nullptr, nullptr};
        if (!dev) {
            double2* m = c.arena.take<double2>(nn * length);
            QT_HIP(hipMemcpyAsync(m, Mii, z * (size_t)(nn * length), hipMemcpyHostToDevice, c.stream));
            dM = m;
            gl = c.arena.take<double2>(per_e * (ce + 2 * (int64_t)nop));
            gg = c.arena.take<double2>(per_e * (ce + 2 * (int64_t)nop));
            for (auto& q : hs) q = c.arena.take<double2>(ns_max);
        }
        {
            std::vector<int> k((size_t)length * ce);
            for (size_t b = 0; b < k.size(); ++b) k[b] = (int)(b % length);
            QT_HIP(hipMemcpyAsync(kidx, k.data(), sizeof(int) * k.size(), hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        const double f = pre_fact * intensity / (hw * hw);
        const int64_t shift = 2 * (int64_t)nop * per_e;
        for (int e0 = 0; e0 < nen; e0 += ce) {
            const int cnt = std::min(ce, nen - e0);
            const int64_t ns = per_e * cnt;
            const int64_t off = per_e * e0;  // of energy e0 in G (padded axis) and in Sigma
            const double2 *cGl, *cGg;
            double2 *cSr, *cSl, *cSg;
            if (dev) {
                cGl = (const double2*)Gl + off;
                cGg = (const double2*)Gg + off;
                cSr = (double2*)Sr + off;
                cSl = (double2*)Sl + off;
                cSg = (double2*)Sg + off;
            } else {
                const size_t gbytes = z * (size_t)(ns + shift);
                QT_HIP(hipMemcpyAsync(gl, (const double2*)Gl + off, gbytes, hipMemcpyHostToDevice, c.stream));
                QT_HIP(hipMemcpyAsync(gg, (const double2*)Gg + off, gbytes, hipMemcpyHostToDevice, c.stream));
                cGl = gl;
                cGg = gg;
                cSr = hs[0];
                cSl = hs[1];
                cSg = hs[2];
            }
            const double2* G[2] = {cGl, cGg};
            double2* S[2] = {cSl, cSg};
            for (int x = 0; x < 2; ++x) {
                eph_sum(c.stream, ns, shift, G[x], S[x]);  // S = G(E - hw) + G(E + hw)
                ZGemmArgs g;                                // T = M_ix S
                g.M = g.N = g.K = nm;
                g.batch = length * cnt;
                g.A = ZOp{dM, nm, 0, nn};
                g.B = ZOp{S[x], nm, nn, 0};
                g.C = T;
                g.ldc = nm;
                g.sbc = nn;
                g.kidx = kidx;
                zgemm(c.stream, 0, 0, g, &c.timer);
                g.A = ZOp{T, nm, nn, 0};                    // Sigma = f T M_ix
                g.B = ZOp{dM, nm, 0, nn};
                g.C = S[x];
                g.alpha = f;
                zgemm(c.stream, 0, 0, g, &c.timer);
            }
            eph_retarded(c.stream, ns, cSl, cSg, cSr);
            if (!dev) {
                const size_t sbytes = z * (size_t)ns;
                QT_HIP(hipMemcpyAsync((double2*)Sr + off, cSr, sbytes, hipMemcpyDeviceToHost, c.stream));
                QT_HIP(hipMemcpyAsync((double2*)Sl + off, cSl, sbytes, hipMemcpyDeviceToHost, c.stream));
                QT_HIP(hipMemcpyAsync((double2*)Sg + off, cSg, sbytes, hipMemcpyDeviceToHost, c.stream));
            }
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

// ---------------------------------------------------------------- block currents, collision integrals
// Weight of energy point e in an energy integral of the reference's
// terminal-current form (gw_mod.f90:424-428): dE e0 / (2 pi hbar) e0
// spindeg (A per unit of the integrand), dE = En[1] - En[0] (0 for one
// energy); times En[e] for the energy current.
static void current_weights(int32_t nen, const double* En, double spindeg, std::vector<double>& w,
                            std::vector<double>& we) {
    const double e0 = 1.6022e-19, hbar = 1.0546e-34, tpi = 3.14159265359 * 2.0;
    const double dE = nen > 1 ? En[1] - En[0] : 0.0;
    const double f = dE * e0 / tpi / hbar * e0 * spindeg;
    w.assign((size_t)nen, f);
    we.resize((size_t)nen);
    for (int32_t e = 0; e < nen; ++e) we[(size_t)e] = f * En[e];
}

int qt_gw_block_current(int32_t nen, int32_t nx, int32_t nb, int32_t ns, const double* En, double spindeg,
                        const void* Hii, const void* GL, const double* cur, double* jdens, double* tot_cur,
                        double* tot_ecur) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (nen <= 0 || nx <= 0 || nb <= 0 || ns <= 0 || !En || !Hii || !GL || !cur || !jdens || !tot_cur ||
        !tot_ecur)
        return QT_ERR_INVALID;
    const int64_t nm = (int64_t)nb * ns;
    if (nm > 4096 || nm * nm * nx > ((int64_t)1 << 40)) return QT_ERR_UNSUPPORTED;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const int64_t nn = nm * nm, per_e = nn * nx;  // samples of a (nm, nm, nx) array per energy
        const int64_t jper = (int64_t)nb * nb * ns * nx;  // jdens values per energy
        const size_t z = sizeof(double2);
        // energies in chunks: G^<, C1, C2 (complex), cur, jdens, weights per energy;
        // fixed: H, U, L, kidx and the two totals
        const size_t per = (3 * z + sizeof(double)) * (size_t)per_e + sizeof(double) * (size_t)jper +
                           sizeof(int) * (size_t)nx + 2 * sizeof(double) + 2048;
        const size_t fixed = 3 * z * (size_t)per_e + 2 * sizeof(double) * (size_t)jper + 16384;
        const size_t budget = workset_budget(c);
        const int ce = (int)std::min<int64_t>(
            nen, std::max<int64_t>(1, budget > fixed + per ? (int64_t)((budget - fixed) / per) : 1));
        c.arena.reserve(fixed + per * (size_t)ce);
        c.arena.reset();
        double2* H = c.arena.take<double2>(per_e);
        double2* U = c.arena.take<double2>(per_e);
        double2* L = c.arena.take<double2>(per_e);
        double* acc = c.arena.take<double>(2 * jper);  // tot_cur, tot_ecur
        double2* g = c.arena.take<double2>(per_e * ce);
        double2* C1 = c.arena.take<double2>(per_e * ce);
        double2* C2 = c.arena.take<double2>(per_e * ce);
        double* cu = c.arena.take<double>(per_e * ce);
        double* jd = c.arena.take<double>(jper * ce);
        double* dw = c.arena.take<double>(2 * (int64_t)ce);
        int* kidx = c.arena.take<int>((int64_t)nx * ce);
        std::vector<double> w, we;
        current_weights(nen, En, spindeg, w, we);
        QT_HIP(hipMemcpyAsync(H, Hii, z * (size_t)per_e, hipMemcpyHostToDevice, c.stream));
        cell_bands(c.stream, (int)nm, nb, nx, H, U, L);
        QT_HIP(hipMemsetAsync(acc, 0, sizeof(double) * (size_t)(2 * jper), c.stream));
        {
            std::vector<int> k((size_t)nx * ce);
            for (size_t b = 0; b < k.size(); ++b) k[b] = (int)(b % nx);
            QT_HIP(hipMemcpyAsync(kidx, k.data(), sizeof(int) * k.size(), hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        for (int e0 = 0; e0 < nen; e0 += ce) {
            const int cnt = std::min(ce, nen - e0);
            QT_HIP(hipMemcpyAsync(g, (const double2*)GL + per_e * e0, z * (size_t)(per_e * cnt),
                                  hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(cu, cur + per_e * e0, sizeof(double) * (size_t)(per_e * cnt),
                                  hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(dw, w.data() + e0, sizeof(double) * cnt, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(dw + ce, we.data() + e0, sizeof(double) * cnt, hipMemcpyHostToDevice, c.stream));
            if (ns > 1) {  // C1 = U_ix G^<, C2 = L_ix G^< for every member (ix, e)
                ZGemmArgs a;
                a.M = a.N = a.K = (int)nm;
                a.batch = nx * cnt;
                a.B = ZOp{g, nm, nn, 0};
                a.ldc = nm;
                a.sbc = nn;
                a.kidx = kidx;
                a.A = ZOp{U, nm, 0, nn};
                a.C = C1;
                zgemm(c.stream, 0, 0, a, &c.timer);
                a.A = ZOp{L, nm, 0, nn};
                a.C = C2;
                zgemm(c.stream, 0, 0, a, &c.timer);
            }
            cell_current(c.stream, nb, ns, nx, cnt, C1, C2, cu, jd);
            energy_sum(c.stream, jper, cnt, jd, dw, acc);
            energy_sum(c.stream, jper, cnt, jd, dw + ce, acc + jper);
            QT_HIP(hipMemcpyAsync(jdens + jper * e0, jd, sizeof(double) * (size_t)(jper * cnt),
                                  hipMemcpyDeviceToHost, c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        QT_HIP(hipMemcpyAsync(tot_cur, acc, sizeof(double) * (size_t)jper, hipMemcpyDeviceToHost, c.stream));
        QT_HIP(hipMemcpyAsync(tot_ecur, acc + jper, sizeof(double) * (size_t)jper, hipMemcpyDeviceToHost,
                              c.stream));
        QT_HIP(hipStreamSynchronize(c.stream));
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_gw_block_collision(int32_t nen, int32_t nx, int32_t nm, const double* En, double spindeg,
                          const void* Sl, const void* Sg, const void* Gl, const void* Gg, void* Itot,
                          void* Ispec) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (nen <= 0 || nx <= 0 || nm <= 0 || !En || !Sl || !Sg || !Gl || !Gg || !Itot || !Ispec)
        return QT_ERR_INVALID;
    if (nm > 4096 || (int64_t)nm * nm * nx > ((int64_t)1 << 40)) return QT_ERR_UNSUPPORTED;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const int64_t nn = (int64_t)nm * nm, per_e = nn * nx;
        const size_t z = sizeof(double2);
        // energies in chunks: Sigma^<, Sigma^>, G^<, G^>, I (complex) and a weight per energy
        const size_t per = 5 * z * (size_t)per_e + sizeof(double) + 1024;
        const size_t fixed = z * (size_t)per_e + 16384;
        const size_t budget = workset_budget(c);
        const int ce = (int)std::min<int64_t>(
            nen, std::max<int64_t>(1, budget > fixed + per ? (int64_t)((budget - fixed) / per) : 1));
        c.arena.reserve(fixed + per * (size_t)ce);
        c.arena.reset();
        double2* acc = c.arena.take<double2>(per_e);
        double2* a4[4];
        for (auto& p : a4) p = c.arena.take<double2>(per_e * ce);
        double2* I = c.arena.take<double2>(per_e * ce);
        double* dw = c.arena.take<double>(ce);
        // Itot = sum_e Ispec(e) dE spindeg / (2 pi)
        const double tpi = 3.14159265359 * 2.0;
        const double f = (nen > 1 ? En[1] - En[0] : 0.0) * spindeg / tpi;
        const std::vector<double> w((size_t)ce, f);
        QT_HIP(hipMemcpyAsync(dw, w.data(), sizeof(double) * ce, hipMemcpyHostToDevice, c.stream));
        QT_HIP(hipMemsetAsync(acc, 0, z * (size_t)per_e, c.stream));
        const void* src[4] = {Sl, Sg, Gl, Gg};
        for (int e0 = 0; e0 < nen; e0 += ce) {
            const int cnt = std::min(ce, nen - e0);
            for (int i = 0; i < 4; ++i)
                QT_HIP(hipMemcpyAsync(a4[i], (const double2*)src[i] + per_e * e0, z * (size_t)(per_e * cnt),
                                      hipMemcpyHostToDevice, c.stream));
            // I = Sigma^< G^> - Sigma^> G^< per (slab, energy) block
            ZGemmArgs a;
            a.M = a.N = a.K = nm;
            a.batch = nx * cnt;
            a.ldc = nm;
            a.sbc = nn;
            a.C = I;
            a.A = ZOp{a4[0], nm, nn, 0};
            a.B = ZOp{a4[3], nm, nn, 0};
            zgemm(c.stream, 0, 0, a, &c.timer);
            a.A = ZOp{a4[1], nm, nn, 0};
            a.B = ZOp{a4[2], nm, nn, 0};
            a.alpha = -1.0;
            a.D = ZOp{I, nm, nn, 0};
            a.beta = 1.0;
            zgemm(c.stream, 0, 0, a, &c.timer);
            energy_sum(c.stream, 2 * per_e, cnt, (const double*)I, dw, (double*)acc);
            QT_HIP(hipMemcpyAsync((double2*)Ispec + per_e * e0, I, z * (size_t)(per_e * cnt), hipMemcpyDeviceToHost,
                                  c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        QT_HIP(hipMemcpyAsync(Itot, acc, z * (size_t)per_e, hipMemcpyDeviceToHost, c.stream));
        QT_HIP(hipStreamSynchronize(c.stream));
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

// ---------------------------------------------------------------- W solve
// W^R = V (I - P^R V)^-1 and W^<,> = W^R P^<,> W^R^H (diagonal blocks) for B
// frequencies at once.  V (Vii, V1i: nm x nm x nx, V1i[i] = V_{i,i-1}) is
// shared; P and W are member-major (member w, slab i at (w*nx + i)*nn).
// With A = I - P^R V, G = A^-1, Z = G P^< G^H:
//   forward   g_i = (A_ii - A_{i,i-1} g_{i-1} A_{i-1,i})^-1,
//             z_i = g_i (P_i + A_{i,i-1} z_{i-1} A_{i,i-1}^H) g_i^H;
//   backward  G_{i,i+1} = -g_i A_{i,i+1} G_{i+1,i+1}, G_{i+1,i} = -G_{i+1,i+1} A_{i+1,i} g_i,
//             G_ii = g_i - G_{i,i+1} A_{i+1,i} g_i,
//             Z_{i,j} = -g_i A_{i,i+1} Z_{i+1,j} - z_i A_{i+1,i}^H G_{j,i+1}^H   (j > i),
//             Z_{j,i} = -Z_{j,i+1} (g_i A_{i,i+1})^H - G_{j,i+1} A_{i+1,i} z_i   (j > i),
//             Z_ii = -g_i A_{i,i+1} Z_{i+1,i} + z_i (I - G_{i,i+1} A_{i+1,i})^H;
//   W^R_ii = sum_j V_ij G_ji,  W^<_ii = sum_k (sum_j V_ij Z_jk) V_ik^H.
struct WSolve {
    Ctx& c;
    int nm, nx, B;
    int64_t nn;
    const double2 *V, *V1, *Id;
    Work w;
    Gemm mm;
    // slab-major internal blocks: slab i of member b at base + (i*B + b)*nn
    double2 *Aii, *Aup, *Alo, *g, *z[2], *Gd, *Gu, *Gl, *Z[2][5], *T[5];
    WSolve(Ctx& c_, int nm_, int nx_, int B_, const double2* V_, const double2* V1_, const double2* Id_)
        : c(c_), nm(nm_), nx(nx_), B(B_), nn((int64_t)nm_ * nm_), V(V_), V1(V1_), Id(Id_) {
        mm.s = c.stream;
        mm.t = &c.timer;
        mm.batch = B;
        mm.kidx = nullptr;
    }
    static size_t bytes(int nm, int nx, int B) {
        const size_t nn = (size_t)nm * nm;
        const size_t blk = sizeof(double2) * nn * B;
        const size_t per = (blk + 256) * (size_t)(19 * nx) + (blk + 256) * 5;
        return per + sizeof(int) * (size_t)B * (nm + 2) + sizeof(double2) * (size_t)B * nm * kInvPanelMax +
               8192;
    }
    void carve(Arena& a) {
        auto slabs = [&](int count) { return a.take<double2>((size_t)count * B * nn); };
        Aii = slabs(nx);
        Aup = slabs(nx);
        Alo = slabs(nx);
        g = slabs(nx);
        z[0] = slabs(nx);
        z[1] = slabs(nx);
        Gd = slabs(nx);
        Gu = slabs(nx);
        Gl = slabs(nx);
        for (int s = 0; s < 2; ++s)
            for (int d = 0; d < 5; ++d) Z[s][d] = slabs(nx);
        for (auto& t : T) t = slabs(1);
        w.piv = a.take<int>((size_t)B * nm);
        w.ginv = a.take<double2>((size_t)B * nm * kInvPanelMax);
        w.info = a.take<int>((size_t)B);
        w.infoacc = a.take<int>((size_t)B);
    }
    double2* at(double2* base, int i) const { return base + (int64_t)i * B * nn; }
    ZOp mem(const double2* base, int i) const { return op_member(at(const_cast<double2*>(base), i), nm, nn); }
    ZOp tmp(int t) const { return op_member(T[t], nm, nn); }
    ZOp vii(int i) const { return op_member(V + i * nn, nm, 0); }
    ZOp v1(int i) const { return op_member(V1 + i * nn, nm, 0); }
    // C = alpha op(A) op(B) + beta D on all members
    void g2(int opa, int opb, ZOp A, ZOp Bo, double2* C, double alpha = 1.0, ZOp D = ZOp(), double beta = 0.0) {
        mm(opa, opb, nm, nm, nm, A, Bo, C, nm, nn, alpha, D, beta);
    }
    // V_ij as an operand with its op (OP_N or OP_C of a stored block)
    void vop(int i, int j, ZOp& o, int& op) const {
        if (i == j) {
            o = vii(i);
            op = OP_N;
        } else if (i == j + 1) {
            o = v1(i);
            op = OP_N;
        } else {
            o = v1(j);
            op = OP_C;
        }
    }
    void run(const double2* PR, const double2* PX[2], double2* WR, double2* WX[2]) {
        const int64_t ms = (int64_t)nx * nn;  // member stride of P / W
        auto P = [&](const double2* base, int i) { return op_member(base + i * nn, nm, ms); };
        ZOp Iop = op_member(Id, nm, 0);
        fill_int(c.stream, w.infoacc, B, 0);
        // A blocks: A_ii = I - P_i V_ii, A_{i,i+1} = -P_i V_{i+1,i}^H, A_{i+1,i} = -P_{i+1} V_{i+1,i}
        for (int i = 0; i < nx; ++i) {
            g2(OP_N, OP_N, P(PR, i), vii(i), at(Aii, i), -1.0, Iop, 1.0);
            if (i + 1 < nx) {
                g2(OP_N, OP_C, P(PR, i), v1(i + 1), at(Aup, i), -1.0);
                g2(OP_N, OP_N, P(PR, i + 1), v1(i + 1), at(Alo, i), -1.0);
            }
        }
        // forward sweep (Alo[i-1] = A_{i,i-1}, Aup[i-1] = A_{i-1,i})
        for (int i = 0; i < nx; ++i) {
            if (i == 0) {
                copy_blocks(c.stream, nm, B, T[0], nn, mem(Aii, 0), nullptr, 0, false);
            } else {
                g2(OP_N, OP_N, mem(Alo, i - 1), mem(g, i - 1), T[1]);
                g2(OP_N, OP_N, tmp(1), mem(Aup, i - 1), T[0], -1.0, mem(Aii, i), 1.0);
            }
            inv_batch(c, w, nm, B, T[0], nn, at(g, i), nn, nullptr);
            note_inv_info(c, w, B, B);
            for (int s = 0; s < 2; ++s) {
                ZOp rhs = P(PX[s], i);
                if (i > 0) {
                    g2(OP_N, OP_N, mem(Alo, i - 1), mem(z[s], i - 1), T[1]);
                    g2(OP_N, OP_C, tmp(1), mem(Alo, i - 1), T[2], 1.0, P(PX[s], i), 1.0);
                    rhs = tmp(2);
                }
                g2(OP_N, OP_N, mem(g, i), rhs, T[3]);
                g2(OP_N, OP_C, tmp(3), mem(g, i), at(z[s], i));
            }
        }
        // backward sweep
        copy_blocks(c.stream, nm, B, at(Gd, nx - 1), nn, mem(g, nx - 1), nullptr, 0, false);
        for (int s = 0; s < 2; ++s)
            copy_blocks(c.stream, nm, B, at(Z[s][2], nx - 1), nn, mem(z[s], nx - 1), nullptr, 0, false);
        for (int i = nx - 2; i >= 0; --i) {
            g2(OP_N, OP_N, mem(g, i), mem(Aup, i), T[0]);                        // T0 = g A01
            g2(OP_N, OP_N, tmp(0), mem(Gd, i + 1), at(Gu, i), -1.0);             // G_{i,i+1}
            g2(OP_N, OP_N, mem(Alo, i), mem(g, i), T[1]);                        // T1 = A10 g
            g2(OP_N, OP_N, mem(Gd, i + 1), tmp(1), at(Gl, i), -1.0);             // G_{i+1,i}
            g2(OP_N, OP_N, mem(Gu, i), tmp(1), at(Gd, i), -1.0, mem(g, i), 1.0); // G_ii
            // M^H with M = I - G_{i,i+1} A10
            g2(OP_N, OP_N, mem(Gu, i), mem(Alo, i), T[4], -1.0, Iop, 1.0);
            for (int s = 0; s < 2; ++s) {
                double2** Zs = Z[s];
                const ZOp zi = mem(z[s], i);
                g2(OP_N, OP_C, zi, mem(Alo, i), T[2]);                             // T2 = z A10^H
                g2(OP_N, OP_N, mem(Alo, i), zi, T[3]);                             // T3 = A10 z
                // Z_{i,i+1}
                g2(OP_N, OP_C, tmp(2), mem(Gd, i + 1), at(Zs[3], i), -1.0);
                g2(OP_N, OP_N, tmp(0), mem(Zs[2], i + 1), at(Zs[3], i), -1.0, mem(Zs[3], i), 1.0);
                // Z_{i+1,i}
                g2(OP_N, OP_N, mem(Gd, i + 1), tmp(3), at(Zs[1], i + 1), -1.0);
                g2(OP_N, OP_C, mem(Zs[2], i + 1), tmp(0), at(Zs[1], i + 1), -1.0, mem(Zs[1], i + 1), 1.0);
                if (i + 2 < nx) {
                    // Z_{i,i+2}
                    g2(OP_N, OP_C, tmp(2), mem(Gl, i + 1), at(Zs[4], i), -1.0);
                    g2(OP_N, OP_N, tmp(0), mem(Zs[3], i + 1), at(Zs[4], i), -1.0, mem(Zs[4], i), 1.0);
                    // Z_{i+2,i}
                    g2(OP_N, OP_N, mem(Gl, i + 1), tmp(3), at(Zs[0], i + 2), -1.0);
                    g2(OP_N, OP_C, mem(Zs[1], i + 2), tmp(0), at(Zs[0], i + 2), -1.0, mem(Zs[0], i + 2), 1.0);
                }
                // Z_ii
                g2(OP_N, OP_C, zi, tmp(4), at(Zs[2], i));
                g2(OP_N, OP_N, tmp(0), mem(Zs[1], i + 1), at(Zs[2], i), -1.0, mem(Zs[2], i), 1.0);
            }
        }
        // W blocks
        for (int i = 0; i < nx; ++i) {
            double2* wr = WR + i * nn;
            bool first = true;
            for (int j = i - 1; j <= i + 1; ++j) {
                if (j < 0 || j >= nx) continue;
                ZOp vo;
                int op;
                vop(i, j, vo, op);
                const double2* Gji = j == i ? at(Gd, i) : (j == i - 1 ? at(Gu, i - 1) : at(Gl, i));
                mm(op, OP_N, nm, nm, nm, vo, op_member(Gji, nm, nn), wr, nm, ms, 1.0,
                   first ? ZOp() : op_member(wr, nm, ms), 1.0);
                first = false;
            }
            for (int s = 0; s < 2; ++s) {
                double2* wx = WX[s] + i * nn;
                bool firstk = true;
                for (int k = i - 1; k <= i + 1; ++k) {
                    if (k < 0 || k >= nx) continue;
                    bool firstj = true;
                    for (int j = i - 1; j <= i + 1; ++j) {
                        if (j < 0 || j >= nx) continue;
                        ZOp vo;
                        int op;
                        vop(i, j, vo, op);
                        mm(op, OP_N, nm, nm, nm, vo, mem(Z[s][2 + (k - j)], j), T[0], nm, nn, 1.0,
                           firstj ? ZOp() : tmp(0), 1.0);
                        firstj = false;
                    }
                    // wx += T0 V_ik^H
                    ZOp vo;
                    int op;
                    vop(i, k, vo, op);
                    const int opd = op == OP_N ? OP_C : OP_N;  // (V_ik)^H of the stored block
                    mm(OP_N, opd, nm, nm, nm, tmp(0), vo, wx, nm, ms, 1.0,
                       firstk ? ZOp() : op_member(wx, nm, ms), 1.0);
                    firstk = false;
                }
            }
        }
    }
};

int qt_gw_wsolve_batched(int32_t nm, int32_t nx, int32_t nw, const void* Vii, const void* V1i,
                         const void* PR, const void* PL, const void* PG, void* WR, void* WL, void* WG,
                         int32_t flags) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (nm <= 0 || nx <= 0 || nw < 0 || !Vii || (nx > 1 && !V1i)) return QT_ERR_INVALID;
    if (nw > 0 && (!PR || !PL || !PG || !WR || !WL || !WG)) return QT_ERR_INVALID;
    if (nm > 1024) return QT_ERR_UNSUPPORTED;
    if (nw == 0) return QT_OK;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        const bool dev = (flags & QT_F_DEVICE_OUTPUTS) != 0;
        const int64_t nn = (int64_t)nm * nm;
        const size_t vbytes = sizeof(double2) * (size_t)(nn * nx);
        // batch from free memory (each member also stages P and W when host buffers are given)
        size_t fr = 0, tot = 0;
        QT_HIP(hipMemGetInfo(&fr, &tot));
        const size_t per = WSolve::bytes(nm, nx, 1) + (dev ? 0 : 6 * vbytes);
        int B = (int)std::max<int64_t>(1, std::min<int64_t>(nw, (int64_t)(0.7 * (double)fr) / (int64_t)per));
        B = std::min(B, 256);
        const size_t need = WSolve::bytes(nm, nx, B) + (dev ? 0 : 6 * vbytes * (size_t)B) + 2 * vbytes +
                            sizeof(double2) * nn + 16384;
        c.arena.reserve(need);
        c.arena.reset();
        double2* dV = c.arena.take<double2>(nn * nx);
        double2* dV1 = c.arena.take<double2>(nn * nx);
        double2* dI = c.arena.take<double2>(nn);
        QT_HIP(hipMemcpyAsync(dV, Vii, vbytes, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c.stream));
        if (nx > 1)
            QT_HIP(hipMemcpyAsync(dV1, V1i, vbytes, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  c.stream));
        {
            std::vector<double2> id((size_t)nn, make_double2(0.0, 0.0));
            for (int d = 0; d < nm; ++d) id[(size_t)d * nm + d] = make_double2(1.0, 0.0);
            QT_HIP(hipMemcpyAsync(dI, id.data(), sizeof(double2) * nn, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        double2* stage[6] = {};
        if (!dev)
            for (auto& p : stage) p = c.arena.take<double2>(nn * nx * B);
        int st = QT_OK;
        for (int w0 = 0; w0 < nw; w0 += B) {
            const int cnt = std::min(B, nw - w0);
            WSolve ws(c, nm, nx, cnt, dV, dV1, dI);
            const size_t mark = c.arena.used;
            ws.carve(c.arena);
            const size_t off = (size_t)w0 * nn * nx;
            const size_t bytes = vbytes * cnt;
            const double2 *pr, *px[2];
            double2 *wr, *wx[2];
            if (dev) {
                pr = (const double2*)PR + off;
                px[0] = (const double2*)PL + off;
                px[1] = (const double2*)PG + off;
                wr = (double2*)WR + off;
                wx[0] = (double2*)WL + off;
                wx[1] = (double2*)WG + off;
            } else {
                QT_HIP(hipMemcpyAsync(stage[0], (const double2*)PR + off, bytes, hipMemcpyHostToDevice, c.stream));
                QT_HIP(hipMemcpyAsync(stage[1], (const double2*)PL + off, bytes, hipMemcpyHostToDevice, c.stream));
                QT_HIP(hipMemcpyAsync(stage[2], (const double2*)PG + off, bytes, hipMemcpyHostToDevice, c.stream));
                pr = stage[0];
                px[0] = stage[1];
                px[1] = stage[2];
                wr = stage[3];
                wx[0] = stage[4];
                wx[1] = stage[5];
            }
            ws.run(pr, px, wr, wx);
            QT_HIP(hipStreamSynchronize(c.stream));
            std::vector<int> info(cnt);
            QT_HIP(hipMemcpy(info.data(), ws.w.infoacc, sizeof(int) * cnt, hipMemcpyDeviceToHost));
            for (int v : info)
                if (v) st = QT_ERR_SINGULAR;
            if (!dev) {
                QT_HIP(hipMemcpy((double2*)WR + off, stage[3], bytes, hipMemcpyDeviceToHost));
                QT_HIP(hipMemcpy((double2*)WL + off, stage[4], bytes, hipMemcpyDeviceToHost));
                QT_HIP(hipMemcpy((double2*)WG + off, stage[5], bytes, hipMemcpyDeviceToHost));
            }
            c.arena.used = mark;
        }
        return st;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

// ---------------------------------------------------------------- GW-SCBA iteration
// One iteration of the loop of gw_ephoton_3d_ijs (gw_mod.f90:372-749) with
// every array on the device: G by the green_RGF_RS recursion with the
// current Sigma (gathered per batch from the element-major arrays, G
// scattered back into them), then P, W, Sigma_new and the mixing, each
// stage the kernels of its own entry point.
int qt_gw_scba_step(qt_ham* h, const qt_rgf_params* p, int32_t nen, const double* En, int32_t ndiag,
                    double spindeg, double alpha, const void* Vii, const void* V1i, void* const* S,
                    void* const* G, void* const* P, double* tr, double* tre, double* cur) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!h || !p || !En || !Vii || !S || !G || !P || nen < 2 || ndiag < 0) return QT_ERR_INVALID;
    if (p->variant != QT_RGF_GW || p->precision != QT_PREC_FP64) return QT_ERR_UNSUPPORTED;
    for (int i = 0; i < 3; ++i)
        if (!S[i] || !G[i] || !P[i]) return QT_ERR_INVALID;
    if (h->nk != 1 || h->S) return QT_ERR_UNSUPPORTED;
    const int nm = h->nm[0], nx = h->nx;
    for (int s = 0; s < nx; ++s)
        if (h->nm[s] != nm) return QT_ERR_UNSUPPORTED;
    if (nx > 1 && !V1i) return QT_ERR_INVALID;
    if (nen > 2048 || !gw_shape_ok(nen, 1, 1, nx, nm, ndiag)) return QT_ERR_UNSUPPORTED;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    void* tmp = nullptr;
    try {
        QT_HIP(hipSetDevice(c.device));
        const int64_t nn = (int64_t)nm * nm;
        const int64_t total = nn * nx * nen;  // samples of one array
        double2* const* Sd = (double2* const*)S;
        double2* const* Gd = (double2* const*)G;
        double2* const* Pd = (double2* const*)P;
        int status = QT_OK;

        // ---- G of every energy (qt_rgf_solve's batch loop, QT_RGF_GW with sigma)
        {
            const int B = choose_batch(*h, nen, true);
            c.arena.reserve(work_bytes(*h, B, true));
            Work w;
            work_carve(c.arena, w, *h, B, true);
            qt_timing tm{};
            std::vector<double> hz(2 * B), hf(5 * B), htr(B), htre(B);
            std::vector<int> hk(2 * B, 0), hinfo(B);
            std::vector<int64_t> doff(nx);
            for (int s = 0; s < nx; ++s) doff[s] = (int64_t)s * nm;
            // host blocks of the slab-bond current matrices, (nm, nm, nx, nen)
            std::vector<void*> curp(cur ? (size_t)nen * nx : 0);
            for (size_t i = 0; i < curp.size(); ++i) curp[i] = cur + (int64_t)i * nn;
            for (int p0 = 0; p0 < nen; p0 += B) {
                const int nb = std::min(B, nen - p0);
                for (int b = 0; b < B; ++b) {
                    const int ip = p0 + std::min(b, nb - 1);
                    hz[b] = hz[B + b] = En[ip];
                    // gw_mod.f90:1185,1232,1333, as in qt_rgf_solve
                    const double fF = ferm((En[ip] - p->mul) / (BOLTZ * p->templ));
                    const double fFc = ferm((En[ip] - p->mul) / (BOLTZ * p->tempr));
                    const double fL = ferm((En[ip] - p->mur) / (BOLTZ * p->templ));
                    hf[b] = fF;
                    hf[B + b] = fL;
                    hf[2 * B + b] = -fF;
                    hf[3 * B + b] = -fL;
                    hf[4 * B + b] = fFc;
                }
                QT_HIP(hipMemcpyAsync(w.zre, hz.data(), sizeof(double) * 2 * B, hipMemcpyHostToDevice, c.stream));
                QT_HIP(hipMemcpyAsync(w.kidx, hk.data(), sizeof(int) * 2 * B, hipMemcpyHostToDevice, c.stream));
                QT_HIP(hipMemcpyAsync(w.fF, hf.data(), sizeof(double) * B, hipMemcpyHostToDevice, c.stream));
                QT_HIP(hipMemcpyAsync(w.fL, hf.data() + B, sizeof(double) * B, hipMemcpyHostToDevice, c.stream));
                QT_HIP(hipMemcpyAsync(w.nfF, hf.data() + 2 * B, sizeof(double) * B, hipMemcpyHostToDevice,
                                      c.stream));
                QT_HIP(hipMemcpyAsync(w.nfL, hf.data() + 3 * B, sizeof(double) * B, hipMemcpyHostToDevice,
                                      c.stream));
                QT_HIP(hipMemcpyAsync(w.fFc, hf.data() + 4 * B, sizeof(double) * B, hipMemcpyHostToDevice,
                                      c.stream));
                fill_int(c.stream, w.infoacc, B, 0);
                // H00 = Hii + Sigma^R, SLS = Sigma^< of each member's energy (no overlap):
                // series (ij, s) of energies p0.. gathered into member blocks, the padding
                // members repeating the last one
                for (int s = 0; s < nx; ++s) {
                    const int64_t off = (int64_t)s * nen + p0;
                    transpose3(c.stream, Sd[0] + off, (int64_t)nx * nen, 0, w.T[0], w.nn, 0, (int)nn, nb, 1);
                    transpose3(c.stream, Sd[1] + off, (int64_t)nx * nen, 0, w.T[1], w.nn, 0, (int)nn, nb, 1);
                    for (int b = nb; b < B; ++b)
                        for (int q = 0; q < 2; ++q)
                            QT_HIP(hipMemcpyAsync(w.T[q] + b * w.nn, w.T[q] + (nb - 1) * w.nn,
                                                  sizeof(double2) * nn, hipMemcpyDeviceToDevice, c.stream));
                    double2* hr = w.H00 + h->hoff[s] * B;
                    double2* hl = w.SLS + h->hoff[s] * B;
                    copy_blocks(c.stream, nm, B, hl, nn, op_member(w.T[1], nm, w.nn), nullptr, 0, false);
                    ShiftArgs sh;
                    sh.n = nm;
                    sh.batch = B;
                    sh.out = hr;
                    sh.sout = nn;
                    sh.M1 = op_k(h->H + h->hoff[s], nm, h->hset);
                    sh.c1 = 1.0;
                    sh.M2 = op_member(w.T[0], nm, w.nn);
                    sh.c2 = 1.0;
                    sh.kidx = w.kidx;
                    shifted(c.stream, sh);
                }
                Solve sv{c, *h, *p, w, B, true, (p->flags & QT_F_COUPLING_FIX) != 0, true, true};
                sv.doff = doff;
                sv.full.any = true;
                sv.full.nx = nx;
                sv.full.p0 = p0;
                sv.full.nb = nb;
                sv.full.nen = nen;
                for (int i = 0; i < 3; ++i) sv.full.em[i] = Gd[i];
                if (cur) sv.full.cur = curp.data();
                sv.run(tm);
                QT_HIP(hipMemcpyAsync(htr.data(), w.tr, sizeof(double) * B, hipMemcpyDeviceToHost, c.stream));
                QT_HIP(hipMemcpyAsync(htre.data(), w.tre, sizeof(double) * B, hipMemcpyDeviceToHost, c.stream));
                QT_HIP(hipMemcpyAsync(hinfo.data(), w.infoacc, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
                QT_HIP(hipStreamSynchronize(c.stream));
                for (int b = 0; b < nb; ++b) {
                    if (tr) tr[p0 + b] = htr[b];
                    if (tre) tre[p0 + b] = htre[b];
                    if (hinfo[b] & QT_INFO_SINGULAR) status = QT_ERR_SINGULAR;
                }
            }
        }

        // ---- P, W, Sigma_new, mixing.  Scratch: A (3 arrays: P as blocks, then W
        // element-major), Bm (3: W as blocks, then Sigma_new), the element lists, I.
        const std::vector<int> el = band_elements(nm, ndiag, false);
        const std::vector<int> pr = band_elements(nm, ndiag, true);
        const size_t abytes = sizeof(double2) * (size_t)total;
        const size_t lists = sizeof(int) * (el.size() + pr.size()) + sizeof(double2) * (size_t)nn + 1024;
        QT_HIP(hipMalloc(&tmp, 6 * abytes + lists));
        double2* A[3];
        double2* Bm[3];
        for (int i = 0; i < 3; ++i) {
            A[i] = (double2*)tmp + (size_t)i * total;
            Bm[i] = (double2*)tmp + (size_t)(3 + i) * total;
        }
        double2* dI = (double2*)tmp + (size_t)6 * total;
        int* del = (int*)(dI + nn);
        int* dpr = del + el.size();
        QT_HIP(hipMemcpyAsync(del, el.data(), sizeof(int) * el.size(), hipMemcpyHostToDevice, c.stream));
        QT_HIP(hipMemcpyAsync(dpr, pr.data(), sizeof(int) * pr.size(), hipMemcpyHostToDevice, c.stream));
        {
            std::vector<double2> id((size_t)nn, make_double2(0.0, 0.0));
            for (int d = 0; d < nm; ++d) id[(size_t)d * nm + d] = make_double2(1.0, 0.0);
            QT_HIP(hipMemcpyAsync(dI, id.data(), sizeof(double2) * nn, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
        }
        const double dEn = En[1] - En[0];
        GwArgs a;
        a.nen = nen;
        a.nky = 1;
        a.nkz = 1;
        a.nx = nx;
        a.nm = nm;
        a.dEn = dEn;
        a.spindeg = spindeg;
        a.Gr = Gd[0];
        a.Gl = Gd[1];
        a.Gg = Gd[2];
        a.Pr = Pd[0];
        a.Pl = Pd[1];
        a.Pg = Pd[2];
        if (ndiag < nm - 1)
            for (int i = 0; i < 3; ++i) QT_HIP(hipMemsetAsync(Pd[i], 0, abytes, c.stream));
        gw_polarization(c.stream, a, del, (int)el.size(), dpr, (int)pr.size());
        // P element-major (ij, x, e) -> blocks (e, x, ij), the W solve's member-major layout
        for (int i = 0; i < 3; ++i)
            transpose3(c.stream, Pd[i], (int64_t)nx * nen, nen, A[i], (int64_t)nx * nn, nn, (int)nn, nen, nx);
        {
            size_t fr = 0, tot = 0;
            QT_HIP(hipMemGetInfo(&fr, &tot));
            const size_t per = WSolve::bytes(nm, nx, 1);
            int Bw = (int)std::max<int64_t>(1, std::min<int64_t>(nen, (int64_t)(0.7 * (double)fr) / (int64_t)per));
            Bw = std::min(Bw, 256);
            c.arena.reserve(WSolve::bytes(nm, nx, Bw) + 16384);
            c.arena.reset();
            for (int w0 = 0; w0 < nen; w0 += Bw) {
                const int cnt = std::min(Bw, nen - w0);
                WSolve ws(c, nm, nx, cnt, (const double2*)Vii, (const double2*)V1i, dI);
                const size_t mark = c.arena.used;
                ws.carve(c.arena);
                const size_t off = (size_t)w0 * nn * nx;
                const double2* px[2] = {A[1] + off, A[2] + off};
                double2* wx[2] = {Bm[1] + off, Bm[2] + off};
                ws.run(A[0] + off, px, Bm[0] + off, wx);
                QT_HIP(hipStreamSynchronize(c.stream));
                std::vector<int> info(cnt);
                QT_HIP(hipMemcpy(info.data(), ws.w.infoacc, sizeof(int) * cnt, hipMemcpyDeviceToHost));
                for (int v : info)
                    if (v) status = QT_ERR_SINGULAR;
                c.arena.used = mark;
            }
        }
        // W blocks (e, x, ij) -> element-major (ij, x, e) into A
        for (int i = 0; i < 3; ++i)
            transpose3(c.stream, Bm[i], (int64_t)nx * nn, nn, A[i], (int64_t)nx * nen, nen, nen, (int)nn, nx);
        a.Wr = A[0];
        a.Wl = A[1];
        a.Wg = A[2];
        a.Sr = Bm[0];
        a.Sl = Bm[1];
        a.Sg = Bm[2];
        if (ndiag < nm - 1)
            for (int i = 0; i < 3; ++i) QT_HIP(hipMemsetAsync(Bm[i], 0, abytes, c.stream));
        gw_selfenergy(c.stream, a, dpr, (int)pr.size());
        for (int i = 0; i < 3; ++i) gw_mix(c.stream, total, alpha, Sd[i], Bm[i]);
        QT_HIP(hipStreamSynchronize(c.stream));
        QT_HIP(hipFree(tmp));
        return status;
    } catch (const hip_error& e) {
        if (tmp) {
            (void)hipStreamSynchronize(c.stream);
            (void)hipFree(tmp);
        }
        return status_of(e);
    }
}

int qt_device_alloc(size_t bytes, void** ptr) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    try {
        QT_HIP(hipMalloc(ptr, bytes));
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}
int qt_device_free(void* ptr) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    return hipFree(ptr) == hipSuccess ? QT_OK : QT_ERR_HIP;
}
int qt_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? QT_OK : QT_ERR_HIP;
}
int qt_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? QT_OK : QT_ERR_HIP;
}
int qt_synchronize(void) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    return hipStreamSynchronize(g_ctx->stream) == hipSuccess ? QT_OK : QT_ERR_HIP;
}
int qt_set_kernel_timing(int enable) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    g_ctx->timer.enabled = enable != 0;
    return QT_OK;
}
int qt_last_timing(qt_timing* t) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!t) return QT_ERR_INVALID;
    *t = g_ctx->last;
    return QT_OK;
}

}  // extern "C"
