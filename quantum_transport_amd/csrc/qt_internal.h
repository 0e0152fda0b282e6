// qt_internal.h -- shared declarations of the gfx950 RGF library.
//
// Device data layout: column-major complex128 blocks stored as double2
// (x = re, y = im), the layout of Fortran complex(8).  Batched operands are
// addressed as  base + b*sb + kidx[b >> kshift]*sk  so one launch can mix
// per-point buffers (sb != 0) with per-k-point Hamiltonian blocks (sk != 0)
// without pointer arrays.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <stdexcept>
#include <string>

namespace qt {

enum Op { OP_N = 0, OP_T = 1, OP_C = 2 };

struct hip_error : std::runtime_error {
    int code;
    hip_error(const std::string& s, int c) : std::runtime_error(s), code(c) {}
};

#define QT_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess)                                                          \
            throw ::qt::hip_error(std::string(#call) + ": " + hipGetErrorString(_e),   \
                                  _e == hipErrorOutOfMemory ? 4 : 3);                  \
    } while (0)

// Batched operand: element (i, j) of member b is
//   p[b*sb + kidx[b>>kshift]*sk + i + j*ld]
struct ZOp {
    const double2* p = nullptr;
    int64_t ld = 0, sb = 0, sk = 0;
};

struct ZGemmArgs {
    int M = 0, N = 0, K = 0, batch = 0;
    ZOp A, B;
    double2* C = nullptr;
    int64_t ldc = 0, sbc = 0, skc = 0;
    ZOp D;                // optional addend, C = alpha*op(A)op(B) + beta*D
    double alpha = 1.0, beta = 0.0;
    const int* kidx = nullptr;
    int kshift = 0;
    const int* active = nullptr;   // skip members with active[b] == 0
    int skip_lo = 0, skip_hi = 0;  // output columns [skip_lo, skip_hi) not written
    int flop_batch = -1;           // members counted by the timer (-1: batch)
    int prec = 0;                  // 3: FP64 three-multiplication form; 6 or 7: int8 digit-slice product with that
                                   // many digits where the shape allows
};

struct Timer;  // forward

// C = alpha * op(A) op(B) + beta * D, complex128, MFMA f64 (or, with
// prec = 3, three real products per complex one; with prec = 6 or 7 and a
// shape zgemm_i8_fits accepts, the int8 emulation).
void zgemm(hipStream_t s, int opa, int opb, const ZGemmArgs& a, Timer* t);
// Same product with operands rounded to 7 digits - 2 bits (6 digits: 40,
// 7 digits: 47) relative to their row (op(A)) or column (op(B)) maximum, on
// the int8 matrix cores.
bool zgemm_i8_fits(const ZGemmArgs& a);
void zgemm_i8(hipStream_t s, int opa, int opb, const ZGemmArgs& a, int digits);

// Batched in-place-free inversion: Ainv[b] = inv(A[b]); A is destroyed.
// piv: n ints per member, info: one int per member (0 or first zero pivot
// column + 1).
constexpr int kInvPanelMax = 32;  // widest inversion panel (columns)
struct InvArgs {
    int n = 0, batch = 0;
    double2* A = nullptr;       // work matrices, lda = n, stride sa
    int64_t sa = 0;
    double2* X = nullptr;       // output inverse, ldx = n, stride sx
    int64_t sx = 0;
    int* piv = nullptr;         // n per member
    int* info = nullptr;        // 1 per member
    double2* G = nullptr;       // multipliers of the panel steps, stride sg per member
    int64_t sg = 0;             // >= n * kInvPanelMax
    const int* active = nullptr;
    int flop_batch = -1;        // members counted by the timer (-1: batch)
};
void zinv(hipStream_t s, const InvArgs& a, Timer* t);
// The same inversion with every member on stream s: no side streams, no
// events, no timer (nothing is created or allocated on the way).
void zinv_single(hipStream_t s, const InvArgs& a);

// ---- elementwise kernels (qt_elem.hip) ----

// out[b] = z_b*S + c1*M1[b] + c2*M2[b] + c3*M3[b]   (n x n blocks)
// z_b = zre[b] + i*zim ; S == nullptr means identity.
struct ShiftArgs {
    int n = 0, batch = 0;
    double2* out = nullptr;
    int64_t sout = 0;
    const double* zre = nullptr;   // per member (device)
    double zim = 0.0;
    ZOp S;                         // may be null (identity)
    ZOp M1, M2, M3;
    double c1 = 0, c2 = 0, c3 = 0;
    const int* kidx = nullptr;
    int kshift = 0;
    const int* active = nullptr;
};
void shifted(hipStream_t s, const ShiftArgs& a);

// out[b] = a_b*(X[b] - X[b]^H) + c1*Y[b] + c2*Z[b]  (n x n)
// a_b = afac ? afac[b] : aconst.
struct AntiArgs {
    int n = 0, batch = 0;
    double2* out = nullptr;
    int64_t sout = 0;
    ZOp X, Y, Z;
    double c1 = 0, c2 = 0;
    const double* afac = nullptr;
    double aconst = 1.0;
    const int* kidx = nullptr;
    int kshift = 0;
    const int* active = nullptr;
};
void antiherm(hipStream_t s, const AntiArgs& a);

// Contact spectral current
//   cur[b] = Re Tr[ Om (f_b G^> - (f_b - 1) G^<) ],  Om = -(Sig - Sig^H),
//   G^> = G^< + G - G^H   (rgf_mod.f90:191-202, 258-266)
struct CurrentArgs {
    int n = 0, batch = 0;
    ZOp Sig, G, Gl;
    const double* f = nullptr;     // per member
    double* partial = nullptr;     // scratch [batch][tiles]
    double* out = nullptr;         // [batch] (device)
    const int* active = nullptr;
    double scale = 1.0;            // sign convention of the caller
};
void contact_current(hipStream_t s, const CurrentArgs& a);

// diag outputs: gr[b*ld + off + j] = G[j,j], gl = Gl[j,j],
// gg = Gl[j,j] + G[j,j] - conj(G[j,j])
struct DiagArgs {
    int n = 0, batch = 0;
    ZOp G, Gl;
    double2 *gr = nullptr, *gl = nullptr, *gg = nullptr;
    int64_t ld = 0, off = 0;
    const int* active = nullptr;
};
void diag_extract(hipStream_t s, const DiagArgs& a);

// Sancho-Rubio step update (rgf_mod.f90:526-541):
//   H_SS += C; H_BB += C + Ap; err_b = sum |H01n|; converged when err < tol
// With HSS2 set, the same decimation also carries the opposite surface of
// the lead (the run with H_01 and H_10 exchanged): H_SS2 += Ap, stopping on
// sum |H10n|.  `active` is the union used to skip converged members.
struct SanchoUpdArgs {
    int n = 0, batch = 0;
    double2 *HSS = nullptr, *HBB = nullptr, *HSS2 = nullptr;
    const double2 *Cm = nullptr, *Ap = nullptr, *H01n = nullptr, *H10n = nullptr;
    int64_t s = 0;                 // common batch stride of all six
    double* partial = nullptr;     // [batch][2][blocks]
    int* active = nullptr;         // union of the two flags below
    int* active1 = nullptr;        // first surface still iterating
    int* active2 = nullptr;        // second surface still iterating (HSS2)
    int* iters = nullptr;          // iteration count per member (first surface)
    int* iters2 = nullptr;         // second surface
    int iter = 0;
    double tol = 0.0;
    int reset = 0;                 // last iteration with reset_on_fail
    const double2* H00 = nullptr;  // for the reset
    int64_t sh00 = 0, skh00 = 0;
    const int* kidx = nullptr;
};
void sancho_update(hipStream_t s, const SanchoUpdArgs& a);

// Energy-axis correlation (mode 0, corr1d2) or convolution (mode 1, conv1d2)
// of nseries series pairs, series s at X + s*n (qt_fft.hip).  n <= 2048.
void corr_conv(hipStream_t s, int n, int mode, const double2* X, const double2* Y, double2* Z,
               int64_t nseries);
void fft_release();  // frees the twiddle tables
// device table exp(-2 pi i t / L), t < L (built once per L, kept until fft_release)
const double2* fft_twiddles(int L);

// GW assembly on element-major arrays (qt_gw.hip): series (ij, ix, ik) at
// ((ij*nx + ix)*nk + ik)*nen, ij = row + col*nm.  elems: in-band ij to
// assemble (polarisation); pairs: in-band ij with row <= col (self-energy,
// which writes both ij and ji; polarisation takes them instead of elems at
// nk == 1 when given).  Out-of-band outputs are left untouched.
struct GwArgs {
    int nen = 0, nky = 1, nkz = 1, nx = 0, nm = 0;
    double dEn = 0.0, spindeg = 1.0;
    const double2 *Gr = nullptr, *Gl = nullptr, *Gg = nullptr;
    const double2 *Wr = nullptr, *Wl = nullptr, *Wg = nullptr;
    double2 *Pr = nullptr, *Pl = nullptr, *Pg = nullptr;
    double2 *Sr = nullptr, *Sl = nullptr, *Sg = nullptr;
};
void gw_polarization(hipStream_t s, const GwArgs& a, const int* elems, int nelem, const int* pairs = nullptr,
                     int npair = 0);
void gw_selfenergy(hipStream_t s, const GwArgs& a, const int* pairs, int npair);
void gw_mix(hipStream_t s, int64_t n, double alpha, double2* S, const double2* Snew);
// S[t] = G[t] + G[t + shift] for t < n (the two photon-shifted energies)
void eph_sum(hipStream_t s, int64_t n, int64_t shift, const double2* G, double2* S);
// Sr[t] = i Im(Sg[t] - Sl[t]) / 2
void eph_retarded(hipStream_t s, int64_t n, const double2* Sl, const double2* Sg, double2* Sr);
// U, L = the cell blocks (j, j+1) and (j+1, j) of each of the nx slab blocks H
// (nm x nm, cells of nb orbitals), the rest zero
void cell_bands(hipStream_t s, int nm, int nb, int nx, const double2* H, double2* U, double2* L);
// orbital-resolved cell-bond currents of cnt energies (qt_gw_block_current)
// from C1 = U G^<, C2 = L G^< and the slab-bond matrices cur
void cell_current(hipStream_t s, int nb, int ns, int nx, int cnt, const double2* C1, const double2* C2,
                  const double* cur, double* jd);
// acc[i] += sum_e X[i + n e] w[e] (energies in order: deterministic)
void energy_sum(hipStream_t s, int64_t n, int cnt, const double* X, const double* w, double* acc);

// out[b] (real n x n) = Re(X1[b]) - Re(X2[b])
void real_diff(hipStream_t s, int n, int batch, double* out, int64_t sout, ZOp X1, ZOp X2);

// copies: dst[b] = src[b] (n x n), optional conjugate transpose
void copy_blocks(hipStream_t s, int n, int batch, double2* dst, int64_t sdst, ZOp src,
                 const int* kidx, int kshift, bool dagger);

// count active members (device) -> *count
void count_active(hipStream_t s, const int* active, int batch, int* count);

// fill ints
void fill_int(hipStream_t s, int* p, int n, int v);

// ---- multilayer perceptron of the lead warm start (qt_mlp.hip) ----
// BF16 row-major operands, FP32 results:
//   C[m][n] = sum_k opA(m, k) opB(k, n) (+ bias[n]),  K % 32 == 0, ld % 8 == 0;
//   transA = 0: A is M x K (ld lda); 1: A is K x M.  transB = 0: B is K x N; 1: N x K.
void mlp_gemm(hipStream_t s, int transA, int transB, int M, int N, int K, const uint16_t* A, int lda,
              const uint16_t* B, int ldb, float* C, int ldc, const float* bias);
void mlp_gelu_fwd(hipStream_t s, int64_t n, const float* Z, uint16_t* H);            // H = bf16(gelu(Z))
void mlp_gelu_bwd(hipStream_t s, int64_t n, const float* dA, const float* Z, float* dZ);  // dZ = dA gelu'(Z)
void mlp_to_bf16(hipStream_t s, int64_t n, const float* X, uint16_t* Y);
void mlp_colsum(hipStream_t s, int rows, int cols, const float* dZ, float* db);
void mlp_adam(hipStream_t s, int64_t n, float* w, const float* g, float* m, float* v, uint16_t* wbf, float lr,
              float b1, float b2, float eps, int step);
// output row of stride ldy (d re | d im | U re | U im | V re | V im, U(i, k)
// at i R + k, 2n + 4nR values) ->
// complex128 d [b][n], U, V [b][n x R] column-major
void mlp_out_to_lowrank(hipStream_t s, int batch, int n, int R, const float* Y, int ldy, double2* d, double2* U,
                        double2* V);
// scale[b] * (diag(E_b), dU_b, dV_b) -> output gradient rows
void mlp_lowrank_to_out(hipStream_t s, int batch, int n, int R, const double2* E, int64_t sE, const double2* dU,
                        const double2* dV, const double* scale, float* dY, int ldy);

// Per-launch timing accumulator (HIP events on the library stream)
struct Timer {
    bool enabled = false;
    hipEvent_t ev[2][4096];
    int n[2] = {0, 0};          // 0: gemm, 1: inversion
    double flops[2] = {0, 0};
    int64_t launches[2] = {0, 0};
    double ms[2] = {0, 0};
    bool created = false;
    void begin(hipStream_t s, int kind);
    void end(hipStream_t s, int kind, double flops_);
    void flush();               // synchronise and sum pending events
    void reset();
};

}  // namespace qt
