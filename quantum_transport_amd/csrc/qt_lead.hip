// qt_lead.hip -- learned warm start for the lead surface Green's function.
//
// The reference obtains every contact self-energy from plain Sancho-Rubio
// decimation (sancho, rgf_mod.f90:482-568): K ~ 25-30 steps of one inversion
// and six products each.  Here a small network (qt_mlp.hip) maps (E, kz) to
// g0 = diag(d) + U V^H (rank R), trained on the device on decimation results
// of the run's first energies; later points refine g0 by Newton's method on
//     F(g) = Phi(g) - g,   Phi(g) = (z - H00 - H01 g H10)^-1,  H01 = H10^H,
// whose correction delta solves the Stein equation
//     delta - P delta R = Phi(g) - g,   P = Phi(g) H01,  R = H10 Phi(g),
// by Smith doubling (delta = sum_k P^k r R^k, squaring P and R each step).
// A point is accepted once |Phi(g) - g|_F <= tol |Phi(g)|_F (g <- Phi(g));
// a point that has not got there within newton_max corrections, or whose
// residual grew, is handed to plain Sancho-Rubio, whose result stands.
// All block arithmetic is complex128 on the library's GEMM and inversion
// kernels; the network is BF16 with FP32 accumulation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "qt_ctx.h"

struct qt_lead {
    int n = 0, nk = 0, R = 32, hid = 256, nin = 32, W = 0;  // W: output row stride
    int64_t nn = 0;
    int bmax = 128;                 // rows of the activation buffers
    double eta = 1e-5, emin = -1.0, emax = 1.0;
    std::vector<double> kz;
    qt_lead_opts o{};
    double2* H00 = nullptr;         // [nk][nn]
    double2* H10 = nullptr;         // [nk][nn]
    int dims[5] = {0, 0, 0, 0, 0};  // layer widths: nin, hid, hid, hid, W
    float* master[4] = {};          // FP32 weights, row-major in x out
    uint16_t* wbf[4] = {};          // BF16 copies used by the products
    float* bias[4] = {};
    float *mw[4] = {}, *vw[4] = {}, *mb[4] = {}, *vb[4] = {}, *gw[4] = {}, *gb[4] = {};
    int step = 0;
    // activations (bmax rows)
    uint16_t* A[4] = {};            // layer inputs (BF16)
    float* Z[4] = {};               // pre-activations; Z[3] is the output row Y
    float* dZw = nullptr;           // output-row gradient (bmax x W)
    float* dZh = nullptr;           // hidden gradient (bmax x hid)
    float* dA = nullptr;            // gradient of a hidden layer's input (bmax x hid)
    uint16_t* dZb = nullptr;        // BF16 copy of a layer's dZ
    // complex factors and loss buffers (bmax members)
    double2 *d = nullptr, *U = nullptr, *V = nullptr, *dU = nullptr, *dV = nullptr, *E = nullptr,
            *G = nullptr;
    double *nE = nullptr, *nG = nullptr, *scale = nullptr, *lossacc = nullptr;
    std::vector<void*> mem;
};

namespace qt {

namespace {

// |X_b - Y_b|_F^2 (Y may be null) over `count` elements per member, one
// workgroup per member, deterministic tree sum
__global__ __launch_bounds__(256) void norm2_kernel(const double2* X, int64_t sx, const double2* Y, int64_t sy,
                                                    int64_t count, const int* active, double* out) {
    const int b = blockIdx.x;
    __shared__ double red[256];
    double s = 0.0;
    if (!active || active[b]) {
        const double2* x = X + b * sx;
        const double2* y = Y ? Y + b * sy : nullptr;
        for (int64_t e = threadIdx.x; e < count; e += 256) {
            double2 v = x[e];
            if (y) {
                v.x -= y[e].x;
                v.y -= y[e].y;
            }
            s += v.x * v.x + v.y * v.y;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0];
}

// out_b = a X_b + c Y_b for active members (n x n blocks, stride s)
__global__ void combine_kernel(int64_t nn, double2* out, double a, const double2* X, double c, const double2* Y,
                               int64_t s, const int* active) {
    const int b = blockIdx.y;
    if (active && !active[b]) return;
    for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
        const double2 x = X[b * s + e], y = Y[b * s + e];
        out[b * s + e] = make_double2(a * x.x + c * y.x, a * x.y + c * y.y);
    }
}

// out[b] = sum_i Im G_b(i, i) (one workgroup per member)
__global__ __launch_bounds__(256) void trace_im_kernel(int n, const double2* G, int64_t s, const int* active,
                                                       double* out) {
    const int b = blockIdx.x;
    __shared__ double red[256];
    double t = 0.0;
    if (!active || active[b])
        for (int i = threadIdx.x; i < n; i += 256) t += G[b * s + (int64_t)i * (n + 1)].y;
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0];
}

// G_b(i, i) += d_b(i)
__global__ void diag_add_kernel(int n, double2* G, int64_t s, const double2* d) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double2* g = G + b * s + (int64_t)i * (n + 1);
    const double2 v = d[(int64_t)b * n + i];
    g->x += v.x;
    g->y += v.y;
}

// scale_b = 2 / (rows |G_b|^2) for the gradient, and the batch's mean
// relative squared error added to *acc (one workgroup)
__global__ void loss_kernel(int rows, const double* nE, const double* nG, double* scale, double* acc) {
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < rows; b += 256) {
        scale[b] = 2.0 / ((double)rows * nG[b]);
        s += nE[b] / nG[b];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += red[0] / rows;
}

int elem_grid(int64_t nn) { return (int)std::min<int64_t>((nn + 255) / 256, 1024); }

void combine(hipStream_t s, int64_t nn, int batch, double2* out, double a, const double2* X, double c,
             const double2* Y, const int* active) {
    if (batch <= 0) return;
    hipLaunchKernelGGL(combine_kernel, dim3(elem_grid(nn), batch), dim3(256), 0, s, nn, out, a, X, c, Y, nn,
                       active);
    QT_HIP(hipGetLastError());
}

void norm2(hipStream_t s, int batch, const double2* X, const double2* Y, int64_t nn, const int* active,
           double* out) {
    if (batch <= 0) return;
    hipLaunchKernelGGL(norm2_kernel, dim3(batch), dim3(256), 0, s, X, nn, Y, nn, nn, active, out);
    QT_HIP(hipGetLastError());
}

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <class T>
T* dalloc(qt_lead* l, size_t count) {
    void* p = nullptr;
    QT_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    l->mem.push_back(p);
    QT_HIP(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(T)));
    return (T*)p;
}

// the network's inputs: scaled energy, cos kz, sin kz, and sin / cos of
// pi j e for j = 1..8 (19 values, zero-padded to 32)
void features(const qt_lead& L, double En, int k, uint16_t* f) {
    const double e = (2.0 * En - (L.emin + L.emax)) / (L.emax - L.emin);
    float v[32] = {};
    v[0] = (float)e;
    v[1] = (float)std::cos(L.kz[k]);
    v[2] = (float)std::sin(L.kz[k]);
    for (int j = 1; j <= 8; ++j) {
        v[1 + 2 * j] = (float)std::sin(M_PI * j * e);
        v[2 + 2 * j] = (float)std::cos(M_PI * j * e);
    }
    for (int i = 0; i < 32; ++i) f[i] = host_bf16(v[i]);
}

// features of `rows` points into dst (device); returns once the host
// buffer has been consumed
void upload_features(Ctx& c, qt_lead& L, int rows, const double* En, const int* kidx, uint16_t* dst) {
    std::vector<uint16_t> h((size_t)rows * L.nin);
    for (int r = 0; r < rows; ++r) features(L, En[r], kidx[r], &h[(size_t)r * L.nin]);
    QT_HIP(hipMemcpyAsync(dst, h.data(), h.size() * 2, hipMemcpyHostToDevice, c.stream));
    QT_HIP(hipStreamSynchronize(c.stream));
}

// dst row r = src row idx[r] (rows of `width` 16-bit values, width % 8 == 0)
__global__ void gather_rows_kernel(int rows, int width, const uint16_t* src, const int* idx, uint16_t* dst) {
    const int r = blockIdx.x;
    const uint4* s = (const uint4*)(src + (int64_t)idx[r] * width);
    uint4* d = (uint4*)(dst + (int64_t)r * width);
    for (int e = threadIdx.x; e < width / 8; e += blockDim.x) d[e] = s[e];
}

void forward(Ctx& c, qt_lead& L, int rows) {
    for (int i = 0; i < 4; ++i) {
        mlp_gemm(c.stream, 0, 0, rows, L.dims[i + 1], L.dims[i], L.A[i], L.dims[i], L.wbf[i], L.dims[i + 1], L.Z[i],
                 L.dims[i + 1], L.bias[i]);
        if (i < 3) mlp_gelu_fwd(c.stream, (int64_t)rows * L.dims[i + 1], L.Z[i], L.A[i + 1]);
    }
}

// g0 = U V^H + diag(d) from the output rows, into G (rows members, stride nn)
void assemble(Ctx& c, qt_lead& L, int rows, double2* G) {
    mlp_out_to_lowrank(c.stream, rows, L.n, L.R, L.Z[3], L.W, L.d, L.U, L.V);
    Gemm g{c.stream, nullptr, rows, nullptr};
    g(OP_N, OP_C, L.n, L.n, L.R, op_member(L.U, L.n, (int64_t)L.n * L.R), op_member(L.V, L.n, (int64_t)L.n * L.R), G,
      L.n, L.nn);
    hipLaunchKernelGGL(diag_add_kernel, dim3((L.n + 255) / 256, rows), dim3(256), 0, c.stream, L.n, G, L.nn, L.d);
    QT_HIP(hipGetLastError());
}

// loss of the output rows against the targets L.G; leaves the output-row
// gradient in L.dZw and adds the batch's loss to *L.lossacc
void loss_and_grad(Ctx& c, qt_lead& L, int rows) {
    mlp_out_to_lowrank(c.stream, rows, L.n, L.R, L.Z[3], L.W, L.d, L.U, L.V);
    Gemm g{c.stream, nullptr, rows, nullptr};
    const int64_t nr = (int64_t)L.n * L.R;
    // E = U V^H - G + diag(d)
    g(OP_N, OP_C, L.n, L.n, L.R, op_member(L.U, L.n, nr), op_member(L.V, L.n, nr), L.E, L.n, L.nn, 1.0,
      op_member(L.G, L.n, L.nn), -1.0);
    hipLaunchKernelGGL(diag_add_kernel, dim3((L.n + 255) / 256, rows), dim3(256), 0, c.stream, L.n, L.E, L.nn, L.d);
    QT_HIP(hipGetLastError());
    norm2(c.stream, rows, L.E, nullptr, L.nn, nullptr, L.nE);
    norm2(c.stream, rows, L.G, nullptr, L.nn, nullptr, L.nG);
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, c.stream, rows, L.nE, L.nG, L.scale, L.lossacc);
    QT_HIP(hipGetLastError());
    // dL/dU = scale E V, dL/dV = scale E^H U, dL/dd = scale diag(E) (real and
    // imaginary parts; L = mean |E|^2 / |G|^2)
    g(OP_N, OP_N, L.n, L.R, L.n, op_member(L.E, L.n, L.nn), op_member(L.V, L.n, nr), L.dU, L.n, nr);
    g(OP_C, OP_N, L.n, L.R, L.n, op_member(L.E, L.n, L.nn), op_member(L.U, L.n, nr), L.dV, L.n, nr);
    mlp_lowrank_to_out(c.stream, rows, L.n, L.R, L.E, L.nn, L.dU, L.dV, L.scale, L.dZw, L.W);
}

void backward(Ctx& c, qt_lead& L, int rows) {
    float* dZ = L.dZw;
    for (int i = 3; i >= 0; --i) {
        const int in = L.dims[i], out = L.dims[i + 1];
        mlp_to_bf16(c.stream, (int64_t)rows * out, dZ, L.dZb);
        mlp_gemm(c.stream, 1, 0, in, out, rows, L.A[i], in, L.dZb, out, L.gw[i], out, nullptr);  // A^T dZ
        mlp_colsum(c.stream, rows, out, dZ, L.gb[i]);
        if (i > 0) {
            mlp_gemm(c.stream, 0, 1, rows, in, out, L.dZb, out, L.wbf[i], out, L.dA, in, nullptr);  // dZ W^T
            mlp_gelu_bwd(c.stream, (int64_t)rows * in, L.dA, L.Z[i - 1], L.dZh);
            dZ = L.dZh;
        }
    }
}

void adam(Ctx& c, qt_lead& L) {
    ++L.step;
    for (int i = 0; i < 4; ++i) {
        mlp_adam(c.stream, (int64_t)L.dims[i] * L.dims[i + 1], L.master[i], L.gw[i], L.mw[i], L.vw[i], L.wbf[i],
                 L.o.lr, 0.9f, 0.999f, 1e-8f, L.step);
        mlp_adam(c.stream, L.dims[i + 1], L.bias[i], L.gb[i], L.mb[i], L.vb[i], nullptr, L.o.lr, 0.9f, 0.999f,
                 1e-8f, L.step);
    }
}

// workspace of one lead batch of M members, carved from the context arena
struct LeadWork {
    Work w;
    double2 *g = nullptr, *g0 = nullptr, *gs = nullptr;  // iterate, warm start, comparison
    double* zre = nullptr;                               // energies of the batch
    int* kidx = nullptr;                                 // k-points of the batch
    int *act = nullptr, *act_s = nullptr;                // Newton / doubling flags
    double *n1 = nullptr, *n2 = nullptr, *n3 = nullptr;  // per-member norms, traces
    int *fb_kidx = nullptr, *iters = nullptr;
    double* fb_zre = nullptr;
};

void carve(Ctx& c, qt_lead& L, int M, bool compare, LeadWork& lw) {
    qt_ham h;
    h.nx = 1;
    h.nk = 1;
    h.nm = {L.n};
    h.hoff = {0};
    h.hset = L.nn;
    h.nmax = L.n;
    const int half = (M + 1) / 2;
    const size_t blk = sizeof(double2) * (size_t)M * L.nn + 256;
    c.arena.reserve(work_bytes(h, half, false) + blk * (compare ? 3 : 2) + 16 * (size_t)M * 8 + 8192);
    work_carve(c.arena, lw.w, h, half, false);
    lw.g = c.arena.take<double2>((size_t)M * L.nn);
    lw.g0 = c.arena.take<double2>((size_t)M * L.nn);
    lw.gs = compare ? c.arena.take<double2>((size_t)M * L.nn) : nullptr;
    lw.zre = c.arena.take<double>(M);
    lw.kidx = c.arena.take<int>(M);
    lw.act = c.arena.take<int>(M);
    lw.act_s = c.arena.take<int>(M);
    lw.n1 = c.arena.take<double>(M);
    lw.n2 = c.arena.take<double>(M);
    lw.n3 = c.arena.take<double>(M);
    lw.fb_kidx = c.arena.take<int>(M);
    lw.fb_zre = c.arena.take<double>(M);
    lw.iters = c.arena.take<int>(M);
}

// plain Sancho-Rubio (rgf_mod.f90:482-568) for M members with energies zre
// and k-points kidx (device), G00 into out (stride nn); host iteration counts
void lead_sancho(Ctx& c, qt_lead& L, LeadWork& lw, int M, const double* zre, const int* kidx, double2* out,
                 int* iters_host) {
    if (M <= 0) return;
    Work& w = lw.w;
    copy_blocks(c.stream, L.n, M, w.T[11], L.nn, op_k(L.H10, L.n, L.nn), kidx, 0, false);
    SanchoIO io;
    io.n = L.n;
    io.members = M;
    io.nn = L.nn;
    io.zre = zre;
    io.kidx = kidx;
    io.eta = L.eta;
    io.nmax = L.o.sancho_nmax;
    io.tol = 1e-100;
    io.H00 = op_k(L.H00, L.n, L.nn);
    io.H10 = w.T[11];
    io.G00 = out;
    io.iters_out = lw.iters;
    sancho_run(c, w, io);
    QT_HIP(hipMemcpyAsync(iters_host, lw.iters, sizeof(int) * M, hipMemcpyDeviceToHost, c.stream));
    QT_HIP(hipStreamSynchronize(c.stream));
}

struct Clock {
    hipEvent_t a, b;
    explicit Clock(hipStream_t s) {
        QT_HIP(hipEventCreate(&a));
        QT_HIP(hipEventCreate(&b));
        QT_HIP(hipEventRecord(a, s));
    }
    double stop(hipStream_t s) {
        QT_HIP(hipEventRecord(b, s));
        QT_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        QT_HIP(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return ms;
    }
};

int64_t param_count(const qt_lead& L) {
    int64_t c = 0;
    for (int i = 0; i < 4; ++i) c += (int64_t)L.dims[i] * L.dims[i + 1] + L.dims[i + 1];
    return c;
}

void free_lead(qt_lead* l) {
    for (void* p : l->mem) (void)hipFree(p);
    delete l;
}

}  // namespace

}  // namespace qt

using namespace qt;

extern "C" {

int qt_lead_create(int32_t nm, int32_t nk, const void* const* H00, const void* const* H10, const double* kz,
                   double eta, double emin, double emax, const qt_lead_opts* opts, qt_lead** out) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (nm <= 0 || nm > 1024 || nk <= 0 || !H00 || !H10 || !kz || !opts || !out || !(emax > emin))
        return QT_ERR_INVALID;
    if (opts->rank <= 0 || opts->rank % 8 || opts->hidden <= 0 || opts->hidden % 32 || opts->batch <= 0 ||
        opts->batch % 32 || opts->newton_max < 0 || opts->smith_max < 0 || opts->sancho_nmax <= 0)
        return QT_ERR_INVALID;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    qt_lead* L = new qt_lead();
    try {
        QT_HIP(hipSetDevice(c.device));
        L->n = nm;
        L->nk = nk;
        L->nn = (int64_t)nm * nm;
        L->R = opts->rank;
        L->hid = opts->hidden;
        L->W = (2 * nm + 4 * nm * L->R + 31) / 32 * 32;  // output row padded to a multiple of 32
        L->bmax = std::max(128, opts->batch);
        L->eta = eta;
        L->emin = emin;
        L->emax = emax;
        L->kz.assign(kz, kz + nk);
        L->o = *opts;
        const int dims[5] = {L->nin, L->hid, L->hid, L->hid, L->W};
        std::copy(dims, dims + 5, L->dims);
        L->H00 = dalloc<double2>(L, (size_t)nk * L->nn);
        L->H10 = dalloc<double2>(L, (size_t)nk * L->nn);
        for (int k = 0; k < nk; ++k) {
            QT_HIP(hipMemcpy(L->H00 + k * L->nn, H00[k], sizeof(double2) * L->nn, hipMemcpyHostToDevice));
            QT_HIP(hipMemcpy(L->H10 + k * L->nn, H10[k], sizeof(double2) * L->nn, hipMemcpyHostToDevice));
        }
        // He initialisation of the GELU layers, a small output layer (the
        // warm start begins near zero)
        std::mt19937_64 rng(opts->seed);
        for (int i = 0; i < 4; ++i) {
            const int64_t cnt = (int64_t)dims[i] * dims[i + 1];
            const double sd = i < 3 ? std::sqrt(2.0 / dims[i]) : 0.02 / std::sqrt((double)dims[i]);
            std::normal_distribution<double> nd(0.0, sd);
            std::vector<float> w(cnt);
            std::vector<uint16_t> wb(cnt);
            for (int64_t e = 0; e < cnt; ++e) {
                w[e] = (float)nd(rng);
                wb[e] = host_bf16(w[e]);
            }
            L->master[i] = dalloc<float>(L, cnt);
            L->wbf[i] = dalloc<uint16_t>(L, cnt);
            L->bias[i] = dalloc<float>(L, dims[i + 1]);
            L->mw[i] = dalloc<float>(L, cnt);
            L->vw[i] = dalloc<float>(L, cnt);
            L->gw[i] = dalloc<float>(L, cnt);
            L->mb[i] = dalloc<float>(L, dims[i + 1]);
            L->vb[i] = dalloc<float>(L, dims[i + 1]);
            L->gb[i] = dalloc<float>(L, dims[i + 1]);
            QT_HIP(hipMemcpy(L->master[i], w.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
            QT_HIP(hipMemcpy(L->wbf[i], wb.data(), 2 * cnt, hipMemcpyHostToDevice));
        }
        const int B = L->bmax;
        for (int i = 0; i < 4; ++i) {
            L->A[i] = dalloc<uint16_t>(L, (size_t)B * dims[i]);
            L->Z[i] = dalloc<float>(L, (size_t)B * dims[i + 1]);
        }
        L->dZw = dalloc<float>(L, (size_t)B * L->W);
        L->dZh = dalloc<float>(L, (size_t)B * L->hid);
        L->dA = dalloc<float>(L, (size_t)B * L->hid);
        L->dZb = dalloc<uint16_t>(L, (size_t)B * std::max(L->W, L->hid));
        const size_t nr = (size_t)nm * L->R;
        L->d = dalloc<double2>(L, (size_t)B * nm);
        L->U = dalloc<double2>(L, B * nr);
        L->V = dalloc<double2>(L, B * nr);
        L->dU = dalloc<double2>(L, B * nr);
        L->dV = dalloc<double2>(L, B * nr);
        L->E = dalloc<double2>(L, (size_t)B * L->nn);
        L->G = dalloc<double2>(L, (size_t)B * L->nn);
        L->nE = dalloc<double>(L, B);
        L->nG = dalloc<double>(L, B);
        L->scale = dalloc<double>(L, B);
        L->lossacc = dalloc<double>(L, 1);
        *out = L;
        return QT_OK;
    } catch (const hip_error& e) {
        free_lead(L);
        return status_of(e);
    }
}

int qt_lead_destroy(qt_lead* l) {
    if (!l) return QT_OK;
    if (g_ctx) {
        std::lock_guard<std::mutex> lk(g_ctx->mu);
        (void)hipStreamSynchronize(g_ctx->stream);
    }
    free_lead(l);
    return QT_OK;
}

int qt_lead_param_count(const qt_lead* l, int64_t* count) {
    if (!l || !count) return QT_ERR_INVALID;
    *count = param_count(*l);
    return QT_OK;
}

int qt_lead_get_params(const qt_lead* l, float* params) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!l || !params) return QT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_ctx->mu);
    try {
        QT_HIP(hipStreamSynchronize(g_ctx->stream));
        float* p = params;
        for (int i = 0; i < 4; ++i) {
            const int64_t cnt = (int64_t)l->dims[i] * l->dims[i + 1];
            QT_HIP(hipMemcpy(p, l->master[i], sizeof(float) * cnt, hipMemcpyDeviceToHost));
            p += cnt;
            QT_HIP(hipMemcpy(p, l->bias[i], sizeof(float) * l->dims[i + 1], hipMemcpyDeviceToHost));
            p += l->dims[i + 1];
        }
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_lead_set_params(qt_lead* l, const float* params) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!l || !params) return QT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_ctx->mu);
    try {
        const float* p = params;
        for (int i = 0; i < 4; ++i) {
            const int64_t cnt = (int64_t)l->dims[i] * l->dims[i + 1];
            std::vector<uint16_t> wb(cnt);
            for (int64_t e = 0; e < cnt; ++e) wb[e] = host_bf16(p[e]);
            QT_HIP(hipMemcpy(l->master[i], p, sizeof(float) * cnt, hipMemcpyHostToDevice));
            QT_HIP(hipMemcpy(l->wbf[i], wb.data(), 2 * cnt, hipMemcpyHostToDevice));
            p += cnt;
            QT_HIP(hipMemcpy(l->bias[i], p, sizeof(float) * l->dims[i + 1], hipMemcpyHostToDevice));
            p += l->dims[i + 1];
        }
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_lead_eval(qt_lead* l, int32_t batch, const double* En, const int32_t* kidx, const void* G, float* Y,
                 double* loss, float* grad) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!l || batch <= 0 || batch % 32 || batch > l->bmax || !En || !kidx || !G) return QT_ERR_INVALID;
    for (int i = 0; i < batch; ++i)
        if (kidx[i] < 0 || kidx[i] >= l->nk) return QT_ERR_INVALID;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        qt_lead& L = *l;
        upload_features(c, L, batch, En, kidx, L.A[0]);
        QT_HIP(hipMemcpyAsync(L.G, G, sizeof(double2) * L.nn * batch, hipMemcpyHostToDevice, c.stream));
        forward(c, L, batch);
        if (Y)
            QT_HIP(hipMemcpyAsync(Y, L.Z[3], sizeof(float) * (size_t)batch * L.W, hipMemcpyDeviceToHost, c.stream));
        const double zero = 0.0;
        QT_HIP(hipMemcpyAsync(L.lossacc, &zero, sizeof(double), hipMemcpyHostToDevice, c.stream));
        loss_and_grad(c, L, batch);
        backward(c, L, batch);
        if (loss) QT_HIP(hipMemcpyAsync(loss, L.lossacc, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        QT_HIP(hipStreamSynchronize(c.stream));
        if (grad) {
            float* p = grad;
            for (int i = 0; i < 4; ++i) {
                const int64_t cnt = (int64_t)L.dims[i] * L.dims[i + 1];
                QT_HIP(hipMemcpy(p, L.gw[i], sizeof(float) * cnt, hipMemcpyDeviceToHost));
                p += cnt;
                QT_HIP(hipMemcpy(p, L.gb[i], sizeof(float) * L.dims[i + 1], hipMemcpyDeviceToHost));
                p += L.dims[i + 1];
            }
        }
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

int qt_lead_train(qt_lead* l, int32_t npts, const double* En, const int32_t* kidx, double* loss) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!l || npts <= 0 || !En || !kidx) return QT_ERR_INVALID;
    for (int i = 0; i < npts; ++i)
        if (kidx[i] < 0 || kidx[i] >= l->nk) return QT_ERR_INVALID;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    double2* Gtrain = nullptr;
    uint16_t* feat = nullptr;
    int* order = nullptr;
    try {
        QT_HIP(hipSetDevice(c.device));
        qt_lead& L = *l;
        // targets: plain Sancho-Rubio on the training points
        QT_HIP(hipMalloc((void**)&Gtrain, sizeof(double2) * (size_t)npts * L.nn));
        const int Ms = 128;
        LeadWork lw;
        carve(c, L, Ms, false, lw);
        std::vector<int> it(Ms), unconverged(Ms), valid;
        valid.reserve(npts);
        for (int p0 = 0; p0 < npts; p0 += Ms) {
            const int m = std::min(Ms, npts - p0);
            QT_HIP(hipMemcpyAsync(lw.zre, En + p0, sizeof(double) * m, hipMemcpyHostToDevice, c.stream));
            QT_HIP(hipMemcpyAsync(lw.kidx, kidx + p0, sizeof(int) * m, hipMemcpyHostToDevice, c.stream));
            lead_sancho(c, L, lw, m, lw.zre, lw.kidx, Gtrain + (int64_t)p0 * L.nn, it.data());
            // a member still active after the decimation loop missed the stop test
            QT_HIP(hipMemcpy(unconverged.data(), lw.w.active, sizeof(int) * m, hipMemcpyDeviceToHost));
            for (int b = 0; b < m; ++b)
                if (!unconverged[b]) valid.push_back(p0 + b);
        }
        const int nvalid = (int)valid.size();
        if (nvalid == 0) {
            QT_HIP(hipFree(Gtrain));
            return QT_ERR_INVALID;
        }
        // features of every training point, and minibatches of a seeded
        // permutation of the converged ones, refreshed every epoch (both
        // gathered on the device)
        const int Bt = L.o.batch;
        const int steps = (nvalid + Bt - 1) / Bt;
        QT_HIP(hipMalloc((void**)&feat, sizeof(uint16_t) * (size_t)npts * L.nin));
        QT_HIP(hipMalloc((void**)&order, sizeof(int) * (size_t)steps * Bt));
        for (int p0 = 0; p0 < npts; p0 += 4096)
            upload_features(c, L, std::min(4096, npts - p0), En + p0, kidx + p0, feat + (size_t)p0 * L.nin);
        std::mt19937_64 rng(L.o.seed ^ 0x9e3779b97f4a7c15ull);
        std::vector<int> perm(valid), ord((size_t)steps * Bt);
        for (int ep = 0; ep < L.o.epochs; ++ep) {
            std::shuffle(perm.begin(), perm.end(), rng);
            for (size_t q = 0; q < ord.size(); ++q) ord[q] = perm[q % nvalid];  // the last batch wraps around
            QT_HIP(hipStreamSynchronize(c.stream));
            QT_HIP(hipMemcpy(order, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice));
            const double zero = 0.0;
            QT_HIP(hipMemcpy(L.lossacc, &zero, sizeof(double), hipMemcpyHostToDevice));
            for (int st = 0; st < steps; ++st) {
                const int* idx = order + (size_t)st * Bt;
                hipLaunchKernelGGL(gather_rows_kernel, dim3(Bt), dim3(64), 0, c.stream, Bt, L.nin, feat, idx,
                                   L.A[0]);
                QT_HIP(hipGetLastError());
                copy_blocks(c.stream, L.n, Bt, L.G, L.nn, op_k(Gtrain, L.n, L.nn), idx, 0, false);
                forward(c, L, Bt);
                loss_and_grad(c, L, Bt);
                backward(c, L, Bt);
                adam(c, L);
            }
            double acc = 0.0;
            QT_HIP(hipMemcpyAsync(&acc, L.lossacc, sizeof(double), hipMemcpyDeviceToHost, c.stream));
            QT_HIP(hipStreamSynchronize(c.stream));
            if (loss) loss[ep] = acc / steps;
        }
        QT_HIP(hipFree(Gtrain));
        QT_HIP(hipFree(feat));
        QT_HIP(hipFree(order));
        return QT_OK;
    } catch (const hip_error& e) {
        if (Gtrain) (void)hipFree(Gtrain);
        if (feat) (void)hipFree(feat);
        if (order) (void)hipFree(order);
        return status_of(e);
    }
}

int qt_lead_surface(qt_lead* l, int32_t mode, int32_t flags, int32_t npts, const double* En,
                    const int32_t* kidx, void* const* G00, qt_lead_stat* stats, qt_lead_report* report) {
    if (!g_ctx) return QT_ERR_NOT_INIT;
    if (!l || npts < 0 || (npts > 0 && (!En || !kidx))) return QT_ERR_INVALID;
    if (mode != QT_LEAD_SANCHO && mode != QT_LEAD_NEWTON_MLP && mode != QT_LEAD_NEWTON_BARE &&
        mode != QT_LEAD_NEWTON_GIVEN)
        return QT_ERR_INVALID;
    if (mode == QT_LEAD_NEWTON_GIVEN && npts > 0 && !G00) return QT_ERR_INVALID;
    for (int i = 0; i < npts; ++i)
        if (kidx[i] < 0 || kidx[i] >= l->nk) return QT_ERR_INVALID;
    Ctx& c = *g_ctx;
    std::lock_guard<std::mutex> lk(c.mu);
    try {
        QT_HIP(hipSetDevice(c.device));
        qt_lead& L = *l;
        const bool compare = (flags & QT_LEAD_F_COMPARE) != 0;
        const int Mb = L.bmax;
        LeadWork lw;
        carve(c, L, Mb, compare, lw);
        qt_lead_report rep{};
        const hipStream_t s = c.stream;
        const int n = L.n;
        const int64_t nn = L.nn;
        std::vector<double> h1(Mb), h2(Mb), h3(Mb), res(Mb), prev(Mb), pinc(Mb);
        // host flag arrays outlive every asynchronous upload from them (each
        // is rewritten only after a stream synchronisation)
        std::vector<int> act(Mb), acts(Mb), acc(Mb), nsteps(Mb), ndbl(Mb), fb(Mb), sit(Mb), it(Mb);
        std::vector<double> fbE;
        std::vector<int> fbk, fbm;
        for (int p0 = 0; p0 < npts; p0 += Mb) {
            const int M = std::min(Mb, npts - p0);
            const double* E = En + p0;
            const int* K = kidx + p0;
            QT_HIP(hipMemcpyAsync(lw.zre, E, sizeof(double) * M, hipMemcpyHostToDevice, s));
            QT_HIP(hipMemcpyAsync(lw.kidx, K, sizeof(int) * M, hipMemcpyHostToDevice, s));
            std::fill(nsteps.begin(), nsteps.end(), 0);
            std::fill(ndbl.begin(), ndbl.end(), 0);
            std::fill(fb.begin(), fb.end(), 0);
            std::fill(sit.begin(), sit.end(), 0);
            std::fill(res.begin(), res.end(), 0.0);
            Clock total(s);
            if (mode == QT_LEAD_SANCHO) {
                lead_sancho(c, L, lw, M, lw.zre, lw.kidx, lw.g, it.data());
                for (int b = 0; b < M; ++b) {
                    sit[b] = it[b];
                    fb[b] = 0;
                }
            } else {
                // warm start
                Clock cn(s);
                Work& w = lw.w;
                if (mode == QT_LEAD_NEWTON_MLP) {
                    upload_features(c, L, M, E, K, L.A[0]);
                    forward(c, L, M);
                    assemble(c, L, M, lw.g);
                } else if (mode == QT_LEAD_NEWTON_GIVEN) {
                    const size_t bytes = sizeof(double2) * (size_t)nn;
                    char* stg = c.host_staging(bytes * M);
                    for (int b = 0; b < M; ++b) std::memcpy(stg + bytes * b, G00[p0 + b], bytes);
                    QT_HIP(hipMemcpyAsync(lw.g, stg, bytes * M, hipMemcpyHostToDevice, s));
                    QT_HIP(hipStreamSynchronize(s));
                } else {
                    ShiftArgs sh;  // g0 = (z - H00)^-1, the isolated principal layer
                    sh.n = n;
                    sh.batch = M;
                    sh.out = w.T[0];
                    sh.sout = nn;
                    sh.zre = lw.zre;
                    sh.zim = L.eta;
                    sh.M1 = op_k(L.H00, n, nn);
                    sh.c1 = -1.0;
                    sh.kidx = lw.kidx;
                    shifted(s, sh);
                    inv_batch(c, w, n, M, w.T[0], nn, lw.g, nn, nullptr);
                }
                copy_blocks(s, n, M, lw.g0, nn, op_member(lw.g, n, nn), nullptr, 0, false);
                rep.network_ms += cn.stop(s);
                // Newton
                Clock ct(s);
                double2 *T1 = w.T[0], *T2 = w.T[1], *Mi = w.T[2], *r = w.T[3], *P = w.T[4], *Rm = w.T[5];
                double2 *X1 = w.T[6], *inc = w.T[7], *Ashift = w.T[8];
                std::fill(act.begin(), act.begin() + M, 1);
                std::fill(prev.begin(), prev.end(), HUGE_VAL);
                int nact = M;
                for (int step = 0; step <= L.o.newton_max && nact > 0; ++step) {
                    QT_HIP(hipMemcpyAsync(lw.act, act.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
                    Gemm g{s, &c.timer, M, nullptr};
                    g.active = lw.act;
                    g.flop_batch = nact;
                    g.kidx = lw.kidx;
                    const ZOp h10 = op_k(L.H10, n, nn);
                    // Phi(g) = (z - H00 - H01 g H10)^-1
                    g(OP_N, OP_N, n, n, n, op_member(lw.g, n, nn), h10, T1, n, nn);
                    g(OP_C, OP_N, n, n, n, h10, op_member(T1, n, nn), T2, n, nn);
                    ShiftArgs sh;
                    sh.n = n;
                    sh.batch = M;
                    sh.out = Ashift;
                    sh.sout = nn;
                    sh.zre = lw.zre;
                    sh.zim = L.eta;
                    sh.M1 = op_k(L.H00, n, nn);
                    sh.c1 = -1.0;
                    sh.M2 = op_member(T2, n, nn);
                    sh.c2 = -1.0;
                    sh.kidx = lw.kidx;
                    sh.active = lw.act;
                    shifted(s, sh);
                    inv_batch(c, w, n, M, Ashift, nn, Mi, nn, lw.act, nact);
                    combine(s, nn, M, r, 1.0, Mi, -1.0, lw.g, lw.act);  // r = Phi(g) - g
                    norm2(s, M, r, nullptr, nn, lw.act, lw.n1);
                    norm2(s, M, Mi, nullptr, nn, lw.act, lw.n2);
                    hipLaunchKernelGGL(trace_im_kernel, dim3(M), dim3(256), 0, s, n, Mi, nn, lw.act, lw.n3);
                    QT_HIP(hipGetLastError());
                    QT_HIP(hipMemcpyAsync(h1.data(), lw.n1, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                    QT_HIP(hipMemcpyAsync(h2.data(), lw.n2, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                    QT_HIP(hipMemcpyAsync(h3.data(), lw.n3, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                    QT_HIP(hipStreamSynchronize(s));
                    for (int b = 0; b < M; ++b) {
                        if (!act[b]) continue;
                        const double rv = h2[b] > 0.0 ? std::sqrt(h1[b] / h2[b]) : HUGE_VAL;
                        res[b] = rv;
                        // the surface equation also has advanced roots: only the
                        // retarded one (-Im Tr g > 0, positive spectral weight) is taken
                        if (rv <= L.o.newton_tol && h3[b] < 0.0) {
                            act[b] = 2;  // accept Phi(g)
                        } else if (step == L.o.newton_max || !(rv < prev[b]) || !std::isfinite(rv)) {
                            act[b] = 0;  // no contraction within the budget: fallback
                            fb[b] = 1;
                        }
                        prev[b] = rv;
                    }
                    // accepted members take g = Phi(g)
                    int nacc = 0;
                    for (int b = 0; b < M; ++b) {
                        acc[b] = act[b] == 2;
                        nacc += acc[b];
                    }
                    if (nacc) {
                        QT_HIP(hipMemcpyAsync(lw.act_s, acc.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
                        combine(s, nn, M, lw.g, 1.0, Mi, 0.0, Mi, lw.act_s);
                    }
                    nact = 0;
                    for (int b = 0; b < M; ++b) {
                        if (act[b] == 2) act[b] = 0;
                        nact += act[b];
                    }
                    if (nact == 0) break;
                    // Newton correction: delta - P delta R = r by Smith doubling
                    QT_HIP(hipMemcpyAsync(lw.act, act.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
                    g.flop_batch = nact;
                    g(OP_N, OP_C, n, n, n, op_member(Mi, n, nn), h10, P, n, nn);   // P = Phi H01
                    g(OP_N, OP_N, n, n, n, h10, op_member(Mi, n, nn), Rm, n, nn);  // R = H10 Phi
                    acts = act;
                    std::fill(pinc.begin(), pinc.end(), HUGE_VAL);
                    int nd = nact;
                    for (int j = 0; j < L.o.smith_max && nd > 0; ++j) {
                        QT_HIP(hipMemcpyAsync(lw.act_s, acts.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
                        Gemm gs{s, &c.timer, M, nullptr};
                        gs.active = lw.act_s;
                        gs.flop_batch = nd;
                        gs(OP_N, OP_N, n, n, n, op_member(P, n, nn), op_member(r, n, nn), X1, n, nn);
                        gs(OP_N, OP_N, n, n, n, op_member(X1, n, nn), op_member(Rm, n, nn), inc, n, nn);
                        combine(s, nn, M, r, 1.0, r, 1.0, inc, lw.act_s);  // delta += P delta R
                        norm2(s, M, inc, nullptr, nn, lw.act_s, lw.n1);
                        norm2(s, M, r, nullptr, nn, lw.act_s, lw.n2);
                        gs(OP_N, OP_N, n, n, n, op_member(P, n, nn), op_member(P, n, nn), X1, n, nn);
                        gs(OP_N, OP_N, n, n, n, op_member(Rm, n, nn), op_member(Rm, n, nn), inc, n, nn);
                        std::swap(P, X1);
                        std::swap(Rm, inc);
                        QT_HIP(hipMemcpyAsync(h1.data(), lw.n1, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                        QT_HIP(hipMemcpyAsync(h2.data(), lw.n2, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                        QT_HIP(hipStreamSynchronize(s));
                        nd = 0;
                        for (int b = 0; b < M; ++b) {
                            if (!acts[b]) continue;
                            ++ndbl[b];
                            if (h1[b] <= 1e-28 * h2[b]) {
                                acts[b] = 0;  // increment below 1e-14 of delta
                            } else if (!std::isfinite(h1[b]) || h1[b] > 1e16 * pinc[b]) {
                                // the series blows up (spectral radius of P R above one
                                // at this g): no Newton step, hand the point to Sancho.
                                // Inside a band at small eta the radius is 1 - O(eta), so
                                // increments may grow for a while before they decay
                                acts[b] = 0;
                                act[b] = 0;
                                fb[b] = 1;
                            }
                            if (j == 0) pinc[b] = h1[b];  // first increment: the reference size
                            nd += acts[b];
                        }
                    }
                    QT_HIP(hipMemcpyAsync(lw.act, act.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
                    combine(s, nn, M, lw.g, 1.0, lw.g, 1.0, r, lw.act);  // g += delta
                    for (int b = 0; b < M; ++b) nsteps[b] += act[b];
                }
                rep.newton_ms += ct.stop(s);
                // fallback: plain Sancho-Rubio for the members Newton left
                fbm.clear();
                fbE.clear();
                fbk.clear();
                for (int b = 0; b < M; ++b)
                    if (fb[b]) {
                        fbm.push_back(b);
                        fbE.push_back(E[b]);
                        fbk.push_back(K[b]);
                    }
                if (!fbm.empty()) {
                    Clock cf(s);
                    const int F = (int)fbm.size();
                    QT_HIP(hipMemcpyAsync(lw.fb_zre, fbE.data(), sizeof(double) * F, hipMemcpyHostToDevice, s));
                    QT_HIP(hipMemcpyAsync(lw.fb_kidx, fbk.data(), sizeof(int) * F, hipMemcpyHostToDevice, s));
                    double2* tmp = w.T[12];
                    lead_sancho(c, L, lw, F, lw.fb_zre, lw.fb_kidx, tmp, it.data());
                    for (int q = 0; q < F; ++q) {
                        QT_HIP(hipMemcpyAsync(lw.g + fbm[q] * nn, tmp + q * nn, sizeof(double2) * nn,
                                              hipMemcpyDeviceToDevice, s));
                        sit[fbm[q]] = it[q];
                        res[fbm[q]] = 0.0;
                    }
                    rep.fallback_ms += cf.stop(s);
                }
            }
            rep.total_ms += total.stop(s);
            // warm-start error and the comparison with plain Sancho-Rubio
            std::vector<double> werr(M, 0.0), serr(M, 0.0);
            if (mode != QT_LEAD_SANCHO) {
                norm2(s, M, lw.g0, lw.g, nn, nullptr, lw.n1);
                norm2(s, M, lw.g, nullptr, nn, nullptr, lw.n2);
                QT_HIP(hipMemcpyAsync(h1.data(), lw.n1, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                QT_HIP(hipMemcpyAsync(h2.data(), lw.n2, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                QT_HIP(hipStreamSynchronize(s));
                for (int b = 0; b < M; ++b) werr[b] = std::sqrt(h1[b] / std::max(h2[b], 1e-300));
            }
            if (compare) {
                Clock cc(s);
                lead_sancho(c, L, lw, M, lw.zre, lw.kidx, lw.gs, it.data());
                rep.compare_ms += cc.stop(s);
                norm2(s, M, lw.g, lw.gs, nn, nullptr, lw.n1);
                norm2(s, M, lw.gs, nullptr, nn, nullptr, lw.n2);
                QT_HIP(hipMemcpyAsync(h1.data(), lw.n1, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                QT_HIP(hipMemcpyAsync(h2.data(), lw.n2, sizeof(double) * M, hipMemcpyDeviceToHost, s));
                QT_HIP(hipStreamSynchronize(s));
                for (int b = 0; b < M; ++b) {
                    serr[b] = std::sqrt(h1[b] / std::max(h2[b], 1e-300));
                    if (!sit[b]) sit[b] = it[b];
                }
            }
            if (G00) {
                const size_t bytes = sizeof(double2) * (size_t)nn;
                char* stg = c.host_staging(bytes * M);
                QT_HIP(hipMemcpyAsync(stg, lw.g, bytes * M, hipMemcpyDeviceToHost, s));
                QT_HIP(hipStreamSynchronize(s));
                for (int b = 0; b < M; ++b)
                    if (G00[p0 + b]) std::memcpy(G00[p0 + b], stg + bytes * b, bytes);
            }
            for (int b = 0; b < M; ++b) {
                rep.points += 1;
                rep.fallbacks += fb[b];
                rep.newton_steps += nsteps[b];
                rep.smith_doublings += ndbl[b];
                rep.sancho_iterations += sit[b];
                if (stats) {
                    qt_lead_stat& st = stats[p0 + b];
                    st.newton_steps = nsteps[b];
                    st.smith_doublings = ndbl[b];
                    st.fallback = mode == QT_LEAD_SANCHO ? 1 : fb[b];
                    st.sancho_iters = sit[b];
                    st.residual = res[b];
                    st.warm_err = werr[b];
                    st.sancho_err = serr[b];
                }
            }
        }
        if (report) *report = rep;
        return QT_OK;
    } catch (const hip_error& e) {
        return status_of(e);
    }
}

}  // extern "C"
