// qt_ctx.h -- library state shared by the translation units that run
// batched RGF work: the device context (stream, workspace arena, timers),
// the resident Hamiltonian, the per-batch workspace and the Sancho-Rubio
// driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qt_rgf.h"
#include "qt_internal.h"

struct qt_ham {
    int nx = 0, nk = 0;
    std::vector<int> nm;
    std::vector<int64_t> hoff;  // offset of slab s inside one k set (elements)
    std::vector<int64_t> coff;  // offset of coupling s (0..nx)
    int64_t hset = 0, cset = 0;
    double2* H = nullptr;       // [nk][hset]
    double2* C = nullptr;       // [nk][cset]
    double2* S = nullptr;       // [nk][hset] or null (identity)
    bool uniform = true;
    bool leads_same = false;    // both contacts are the same semi-infinite lead
    int nmax = 0;
};

namespace qt {

constexpr double BOLTZ = 8.61734e-05;  // eV/K, rgf_mod.f90:38

inline double ferm(double a) { return 1.0 / (1.0 + std::exp(a)); }  // rgf_mod.f90:475-479

// ------------------------------------------------------------ device arena
struct Arena {
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void reserve(size_t bytes) {
        if (bytes <= cap) return;
        release();
        QT_HIP(hipMalloc((void**)&base, bytes));
        cap = bytes;
    }
    void reset() { used = 0; }
    template <class T>
    T* take(size_t count) {
        const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        if (used + bytes > cap) throw hip_error("workspace arena exhausted", 4);
        T* p = (T*)(base + used);
        used += bytes;
        return p;
    }
    void release() {
        if (base) (void)hipFree(base);
        base = nullptr;
        cap = used = 0;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int* scratch = nullptr;     // device address of `scratch` on this device
    Timer timer;
    Arena arena;
    qt_timing last{};
    std::mutex mu;
    int* host_count = nullptr;  // pinned
    hipEvent_t ev[4];
    char* staging = nullptr;    // pinned host staging for block outputs (grown on demand)
    size_t staging_cap = 0;
    char* host_staging(size_t bytes) {
        if (bytes > staging_cap) {
            if (staging) QT_HIP(hipHostFree(staging));
            staging = nullptr;
            staging_cap = 0;
            QT_HIP(hipHostMalloc((void**)&staging, bytes));
            staging_cap = bytes;
        }
        return staging;
    }
};

extern Ctx* g_ctx;
extern std::mutex g_init_mu;

// ------------------------------------------------------------ operand helpers
inline ZOp op_member(const double2* p, int64_t ld, int64_t sb) {
    ZOp o;
    o.p = p;
    o.ld = ld;
    o.sb = sb;
    return o;
}
inline ZOp op_k(const double2* p, int64_t ld, int64_t sk) {
    ZOp o;
    o.p = p;
    o.ld = ld;
    o.sk = sk;
    return o;
}

struct Gemm {
    hipStream_t s;
    Timer* t;
    int batch;
    const int* kidx;
    const int* active = nullptr;
    int flop_batch = -1;
    int prec = 0;  // 3: FP64 three-multiplication products (QT_PREC_FP64_3M); 6 or 7: int8 digit-slice
                   // products with that many digits (QT_PREC_MIXED, _MIXED47)
    // C = alpha op(A) op(B) + beta D
    void operator()(int opa, int opb, int M, int N, int K, ZOp A, ZOp B, double2* C, int64_t ldc,
                    int64_t sbc, double alpha = 1.0, ZOp D = ZOp(), double beta = 0.0) const {
        ZGemmArgs g;
        g.M = M;
        g.N = N;
        g.K = K;
        g.batch = batch;
        g.A = A;
        g.B = B;
        g.C = C;
        g.ldc = ldc;
        g.sbc = sbc;
        g.D = D;
        g.alpha = alpha;
        g.beta = D.p ? beta : 0.0;
        g.kidx = kidx;
        g.active = active;
        g.flop_batch = flop_batch;
        g.prec = prec;
        zgemm(s, opa, opb, g, t);
    }
};

// ------------------------------------------------------------ workspace
constexpr int NTEMP = 15;

struct Work {
    int B = 0;               // points in the batch
    int nmax = 0;
    int64_t nn = 0;          // nmax^2
    int64_t ndiag = 0;       // sum(nm)
    double2* Gl = nullptr;   // slab-major: slab s at hoff[s]*B, member stride nm[s]^2
    double2* Gln = nullptr;
    double2* H00 = nullptr;  // Hii + sigma_r S per point (only with sigma), same layout
    double2* SLS = nullptr;  // sigma_lesser S per point (only with sigma), same layout
    double2* T[NTEMP];       // temps, 2B * nn each
    double2* sigF = nullptr; // first-contact self-energy [B][nn]
    double2* sigL = nullptr; // last-contact self-energy  [B][nn]
    int* piv = nullptr;      // 2B * nmax
    double2* ginv = nullptr; // 2B * nmax * kInvPanelMax inversion multipliers
    int* info = nullptr;     // 2B
    double* zre = nullptr;   // 2B energies (the batch twice, for the lead members)
    int* kidx = nullptr;     // 2B
    double* fF = nullptr;    // B  occupation injected at the first contact
    double* fFc = nullptr;   // B  occupation of the first contact's current
    double* fL = nullptr;    // B  occupation of the last contact
    double* nfF = nullptr;   // B  -fF
    double* nfL = nullptr;   // B  -fL
    int* active = nullptr;   // 2B
    int* iters = nullptr;    // 2B
    int* lead_it = nullptr;  // 2B  (first contact, then last contact)
    int* act1 = nullptr;     // 2B  decimation flags of the two lead surfaces
    int* act2 = nullptr;     // 2B
    int* iters2 = nullptr;   // 2B
    int* count = nullptr;    // 1
    double* partial = nullptr;  // 2B * max(64, tiles^2)
    int* infoacc = nullptr;     // B accumulated QT_INFO bits
    double* tr = nullptr;       // B
    double* tre = nullptr;      // B
    double2* diag = nullptr;    // 3 * B * ndiag staging
};

size_t work_bytes(const qt_ham& h, int B, bool sigma);
void work_carve(Arena& a, Work& w, const qt_ham& h, int B, bool sigma);
void inv_batch(Ctx& c, Work& w, int n, int batch, double2* A, int64_t sa, double2* X, int64_t sx,
               const int* active, int flop_batch = -1);
void note_inv_info(Ctx& c, Work& w, int batch, int period);

// ------------------------------------------------------------ Sancho-Rubio
// Members m with z_m = zre[m] + i eta (rgf_mod.f90:508).  H00 addressed
// per member or per k; H10 per member (overwritten).  Outputs G00 (and
// GBB) per member with stride nn.  rgf_mod.f90:482-568.
struct SanchoIO {
    int n = 0, members = 0;
    const double* zre = nullptr;
    const int* kidx = nullptr;
    double eta = 1e-5;
    int nmax = 50;
    double tol = 1e-100;
    bool reset = false;
    ZOp S;           // null = identity
    ZOp H00;
    double2* H10 = nullptr;
    double2* G00 = nullptr;
    double2* GBB = nullptr;
    double2* G00b = nullptr;      // opposite surface of the same lead (run with
                                  // H_01 and H_10 exchanged); null = not wanted
    int64_t nn = 0;
    int* iters_out = nullptr;     // members (device)
    int* iters2_out = nullptr;    // opposite surface
    int* info_bits = nullptr;     // per point accumulator (device), may be null
    int period = 1;               // member m belongs to point m % period
    int64_t total_iters = 0;      // host: sum of iterations over members
};

void sancho_run(Ctx& c, Work& w, SanchoIO& io);

// records the message of a library error for qt_status_string, returns its code
int status_of(const hip_error& e);

}  // namespace qt
