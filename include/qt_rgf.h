/*
 * qt_rgf.h -- C ABI of the MI355X recursive Green's function (RGF) library.
 *
 * This is the drop-in boundary for the reference's RGF path.  Each entry
 * point names the reference routine it replaces; the Fortran modules in
 * quantum_transport_amd/fortran bind these symbols through iso_c_binding
 * and keep the reference subroutine interfaces.
 *
 * Data conventions (binary-compatible with Fortran complex(8) arrays):
 *   - every matrix block is column-major complex128, interleaved re/im,
 *     leading dimension = its row count;
 *   - block lists are arrays of pointers, one per block, the way the
 *     reference passes type_matrix_complex lists;
 *   - slab s is 0-based; H1i[s] is H_{s,s-1} (H1i[0]: left lead coupling,
 *     H1i[nx]: right lead coupling), deviceHam_mod.f90:278-286;
 *   - pointers are host pointers unless QT_F_DEVICE_OUTPUTS says otherwise.
 * Every function returns a qt status code (0 = success).
 * The library is re-entrant per call but serialises calls on one device
 * internally (one HIP stream per device context).
 */
#ifndef QT_RGF_H
#define QT_RGF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
    QT_OK = 0,
    QT_ERR_INVALID = 1,      /* bad argument or shape */
    QT_ERR_NOT_INIT = 2,     /* qt_init() not called */
    QT_ERR_HIP = 3,          /* HIP runtime failure */
    QT_ERR_NOMEM = 4,        /* device memory exhausted */
    QT_ERR_SINGULAR = 5,     /* a block inversion met an exactly zero pivot;
                                that block was zeroed as linalg.f90:50-58 does */
    QT_ERR_UNSUPPORTED = 6,  /* shape / option combination not provided */
    QT_ERR_NO_DEVICE = 7,    /* no HIP device visible */
    QT_ERR_TOO_LARGE = 8     /* request exceeds the fixed scratch of a
                                no-allocation entry point */
};

/* per-point info bits (qt_rgf_outputs.info) */
#define QT_INFO_SINGULAR 0x1   /* some inversion of this point was singular */
#define QT_INFO_LEAD_NMAX 0x2  /* a lead decimation reached nmax (warning,
                                  rgf_mod.f90:549-551) */

/* RGF variants */
#define QT_RGF_FORWARD 0   /* rgf_variableblock_forward,  rgf_mod.f90:46-269 */
#define QT_RGF_BACKWARD 1  /* rgf_variableblock_backward, rgf_mod.f90:272-472 */
#define QT_RGF_GW 2        /* green_RGF_RS, gw_mod.f90:1131-1345: forward sweep with
                              the left contact filled with f(E-mul, TEMPl) and its
                              current taken with f(E-mul, TEMPr), the right contact
                              with f(E-mur, TEMPl) and tr of opposite sign; the
                              right lead couples through H1i[nx] (callers replicate
                              the last coupling) */

/* precision modes */
#define QT_PREC_FP64 0     /* complex128 everywhere */
#define QT_PREC_MIXED 1    /* block products on the int8 matrix cores with operands
                              rounded to 40 bits relative to their row / column
                              maximum (6 base-128 digits); inversions, leads and
                              elementwise work in complex128 (no reference
                              counterpart).  Measured: spectra within ~1e-7 and
                              contact currents within ~1e-6 of QT_PREC_FP64
                              (relative to each quantity's largest value). */
#define QT_PREC_MIXED47 2  /* as QT_PREC_MIXED with 7 digits (47 bits) and 28
                              digit-pair products instead of 21 (no reference
                              counterpart).  Measured at nb = 512 x 100 slabs:
                              spectra within ~1e-9 and contact currents at
                              ~1e-8 of QT_PREC_FP64; on a 24-slab device with kz
                              the contact currents within ~1e-6. */
#define QT_PREC_FP64_3M 3  /* complex128 everywhere, the sweep's block products
                              with three real matrix products per complex one
                              ((Ar + Ai)(Br + Bi) - Ar Br - Ai Bi for the
                              imaginary part) instead of four (no reference
                              counterpart). */

/* option flags (qt_rgf_params.flags) */
#define QT_F_COUPLING_FIX 0x1   /* forward variant: the second pass couples slab
                                   s to s+1 through H1i[s+1] (the reference
                                   uses H1i[s], rgf_mod.f90:215-245, equal for
                                   uniform devices; required for uneven slabs) */
#define QT_F_DEVICE_OUTPUTS 0x2 /* the per-point output arrays of
                                   qt_rgf_outputs are device pointers */
#define QT_F_LEAD_RESET 0x4     /* lead decimation resets H_SS=H_BB=H00 at nmax
                                   (gw_mod.f90:1440-1445) */

/* ---- context --------------------------------------------------------- */

/* Replaces cuda_rgf_init (cuda_rgf_mod.f90:32-44): binds the calling
 * process to HIP device `device`, creates its stream and workspace. */
int qt_init(int device);
/* Replaces cuda_rgf_finish (cuda_rgf_mod.f90:47-51): frees all device
 * memory held by the library. */
int qt_finish(void);
const char* qt_status_string(int status);

/* ---- device-resident Hamiltonian ------------------------------------- */

typedef struct qt_ham qt_ham;

/* Uploads the block lists of nk k-points (devH_build_fromWannierFile
 * output, deviceHam_mod.f90:234-290): Hii[ik*nx + s], H1i[ik*(nx+1) + s],
 * Sii[ik*nx + s] (Sii = NULL means identity blocks).  nm[s] is the size of
 * slab s; H1i[s] is nm[s] x nm[s-1], H1i[0] and H1i[nx] are square. */
int qt_ham_create(int32_t nx, const int32_t* nm, int32_t nk, const void* const* Hii,
                  const void* const* H1i, const void* const* Sii, qt_ham** out);
int qt_ham_destroy(qt_ham* h);

/* ---- batched RGF ----------------------------------------------------- */

typedef struct {
    int32_t variant;    /* QT_RGF_FORWARD | QT_RGF_BACKWARD */
    int32_t precision;  /* QT_PREC_FP64 */
    int32_t flags;      /* QT_F_* */
    int32_t lead_nmax;  /* Sancho-Rubio iteration cap (50, rgf_mod.f90:507) */
    double mul, mur;    /* uniform mul / mur matrices (negf_mod.f90:82-83) */
    double templ, tempr;/* uniform TEMPl / TEMPr matrices (negf_mod.f90:84-85) */
    double lead_eta;    /* Im z in the lead decimation (1e-5, rgf_mod.f90:508) */
    double lead_tol;    /* decimation stop: sum |H01| < tol (1e-100, :488) */
} qt_rgf_params;

typedef struct {
    /* per point (length npts), each may be NULL */
    double* tr;          /* spectral current at the second contact */
    double* tre;         /* spectral current at the first contact */
    int32_t* lead_iters; /* [npts][2] decimation iterations, first/second lead */
    int32_t* info;       /* [npts] QT_INFO_* bits */
    /* diagonals [npts][sum(nm)] complex128 of G^R, G^<, G^> per slab */
    void* gr_diag;
    void* gl_diag;
    void* gg_diag;
    /* full host blocks, [npts*nx] pointers each (point-major), may be NULL:
     * the per-point outputs of rgf_variableblock_{forward,backward} */
    void* const* G_r;
    void* const* G_lesser;
    void* const* G_greater;
    void* const* Jdens;
    void* const* Gl;
    void* const* Gln;
    /* QT_RGF_GW only, [npts*nx] host blocks each, may be NULL:
     * cur: real nm x nm bond-current blocks, Re(H_{l,l+1} G<_{l+1,l}) -
     *      Re(H_{l+1,l} G<_{l,l+1}) (gw_mod.f90:1252-1273);
     * GRi1 / GLi1: G_{l,l+1} and G<_{l,l+1};  GGi1: the reference's
     * Gl^H H^H G^> term (gw_mod.f90:1275-1284) */
    void* const* cur;
    void* const* GRi1;
    void* const* GLi1;
    void* const* GGi1;
} qt_rgf_outputs;

/* Replaces the E x k loop of negf_mod.f90:115-127 around
 * rgf_variableblock_forward / _backward (rgf_mod.f90:46-472) and
 * cuda_rgf_constblocksize (cuda_rgf_mod.f90:53-413): all points
 * (En[ip], k-point kidx[ip]) are solved in batches on the device.
 * sigma_lesser_ph / sigma_r_ph: NULL (zero) or [npts*nx] host block
 * pointers (point-major), the sigma_*_ph(nx) lists of each point. */
int qt_rgf_solve(qt_ham* h, const qt_rgf_params* p, int32_t npts, const double* En,
                 const int32_t* kidx, const void* const* sigma_lesser_ph,
                 const void* const* sigma_r_ph, qt_rgf_outputs* out);

/* ---- lead self-energy ------------------------------------------------ */

/* Batched Sancho-Rubio decimation, replaces sancho (rgf_mod.f90:482-568,
 * cuda_rgf_mod.f90:657-765, gw_mod.f90:1378-1464): for every point ip,
 * z = En[ip] + i*eta; H00[ip], H10[ip] are nm x nm host blocks (pointers may
 * repeat), S00 one block or NULL (identity).  G00/GBB: [npts] host output
 * blocks (GBB may be NULL).  iters/converged: [npts], may be NULL. */
int qt_sancho(int32_t nm, int32_t npts, const double* En, double eta, int32_t nmax,
              double tol, int32_t reset_on_fail, const void* S00, const void* const* H00,
              const void* const* H10, void* const* G00, void* const* GBB, int32_t* iters,
              int32_t* converged);

/* ---- learned lead warm start ------------------------------------------ */

/* No reference counterpart (the reference runs plain decimation, sancho in
 * rgf_mod.f90:482-568, for every point).  A lead (principal layer H00,
 * coupling H10 per kz) gets a small network (E, kz) -> g0 = diag(d) + U V^H,
 * rank R, trained on the device by the library's own kernels on Sancho-Rubio
 * surface Green's functions of the run's first energies; later points start
 * Newton's method on g = (z - H00 - H01 g H10)^-1 (H01 = H10^H, z = E + i eta)
 * from g0, each Newton correction solving its Stein equation by Smith
 * doubling, and fall back to plain Sancho-Rubio when Newton has not reached
 * the residual |(z - H00 - H01 g H10)^-1 - g|_F <= newton_tol |g|_F within
 * newton_max steps. */
typedef struct qt_lead qt_lead;

typedef struct {
    int32_t rank;        /* R, columns of U and V (32) */
    int32_t hidden;      /* width of the 3 hidden GELU layers (256) */
    int32_t batch;       /* training minibatch, multiple of 32 (128) */
    int32_t epochs;      /* passes over the training points (30) */
    float lr;            /* Adam step size (1e-3) */
    int32_t newton_max;  /* Newton steps before the fallback (4) */
    double newton_tol;   /* accepted surface-equation residual (1e-10) */
    int32_t smith_max;   /* doublings per Newton correction (30) */
    int32_t sancho_nmax; /* decimation cap (50, rgf_mod.f90:507) */
    uint64_t seed;       /* weight initialisation and minibatch order */
} qt_lead_opts;

#define QT_LEAD_SANCHO 0        /* plain Sancho-Rubio (the reference's method) */
#define QT_LEAD_NEWTON_MLP 1    /* Newton from the network's g0, Sancho fallback */
#define QT_LEAD_NEWTON_BARE 2   /* Newton from g0 = (z - H00)^-1, Sancho fallback
                                   (the method without the network) */
#define QT_LEAD_NEWTON_GIVEN 3  /* Newton from the caller's blocks G00 (e.g. a
                                   neighbouring energy's result), Sancho fallback */
#define QT_LEAD_F_COMPARE 0x1   /* also run plain Sancho-Rubio and report the
                                   difference (qt_lead_stat.sancho_err) */

typedef struct {
    int32_t newton_steps;    /* Newton corrections applied */
    int32_t smith_doublings; /* doublings over all corrections */
    int32_t fallback;        /* 1: the result is plain Sancho-Rubio's */
    int32_t sancho_iters;    /* decimation steps run for this point (fallback,
                                QT_LEAD_SANCHO, or the comparison) */
    double residual;         /* |Phi(g) - g|_F / |Phi(g)|_F of the returned g
                                (Newton points; 0 for Sancho results) */
    double warm_err;         /* |g0 - g|_F / |g|_F (Newton modes) */
    double sancho_err;       /* |g - g_Sancho|_F / |g_Sancho|_F (QT_LEAD_F_COMPARE) */
} qt_lead_stat;

typedef struct {
    double total_ms;         /* the mode's lead work: network + Newton + fallback */
    double network_ms;       /* features, forward pass, g0 assembly */
    double newton_ms;        /* Newton steps */
    double fallback_ms;      /* Sancho-Rubio on the points Newton left */
    double compare_ms;       /* the comparison run (QT_LEAD_F_COMPARE) */
    int64_t points, fallbacks, newton_steps, smith_doublings, sancho_iterations;
} qt_lead_report;

/* Uploads the lead of nk k-points: H00[ik], H10[ik] (nm x nm host blocks);
 * kz[ik] and the energy window [emin, emax] scale the network's inputs. */
int qt_lead_create(int32_t nm, int32_t nk, const void* const* H00, const void* const* H10, const double* kz,
                   double eta, double emin, double emax, const qt_lead_opts* opts, qt_lead** out);
int qt_lead_destroy(qt_lead* l);
/* Runs plain Sancho-Rubio on the npts training points (En[i], kidx[i]) and
 * trains the network on the results for opts->epochs passes; loss[e] (may be
 * NULL) receives the mean relative squared error of epoch e.  Points whose
 * decimation has not converged within opts->sancho_nmax steps are left out
 * of the training set (their result is no solution of the surface equation);
 * QT_ERR_INVALID when none is left. */
int qt_lead_train(qt_lead* l, int32_t npts, const double* En, const int32_t* kidx, double* loss);
/* Surface Green's functions of npts points by `mode` (QT_LEAD_*); G00 (npts
 * host blocks, may be NULL; required, and read as the warm start, with
 * QT_LEAD_NEWTON_GIVEN), stats (npts, may be NULL), report (may be NULL). */
int qt_lead_surface(qt_lead* l, int32_t mode, int32_t flags, int32_t npts, const double* En,
                    const int32_t* kidx, void* const* G00, qt_lead_stat* stats, qt_lead_report* report);
/* Network parameters as FP32 arrays in layer order (W0, b0, W1, b1, W2, b2,
 * W3, b3; W_l row-major in_l x out_l): the count, and copies out / in (the
 * BF16 weights used by the products are the rounded masters). */
int qt_lead_param_count(const qt_lead* l, int64_t* count);
int qt_lead_get_params(const qt_lead* l, float* params);
int qt_lead_set_params(qt_lead* l, const float* params);
/* One forward and backward pass on `batch` points (a multiple of 32) with
 * host targets G (batch complex128 nm x nm blocks): the network output rows
 * Y (batch x (2 nm + 4 nm R) floats), the loss and the parameter gradients
 * (same order as qt_lead_get_params).  Any of Y, loss, grad may be NULL. */
int qt_lead_eval(qt_lead* l, int32_t batch, const double* En, const int32_t* kidx, const void* G, float* Y,
                 double* loss, float* grad);

/* ---- block primitives ------------------------------------------------ */

/* Batched complex128 product R[b] = alpha*op(A[b])*op(B[b]) + beta*R[b],
 * op in {0: N, 1: T, 2: C}; replaces MUL_C (matrix_c.f90:615-660), whose
 * ZGEMM call is R = op(A) op(B).  Host buffers; member b of X starts at
 * X + b*strideX (complex elements, 0 = shared operand). */
int qt_zgemm_batched(int32_t opa, int32_t opb, int32_t m, int32_t n, int32_t k, int32_t batch,
                     double alpha, const void* A, int64_t lda, int64_t strideA, const void* B,
                     int64_t ldb, int64_t strideB, double beta, void* R, int64_t ldr,
                     int64_t strideR);

/* qt_zgemm_batched with a precision mode (QT_PREC_FP64, _MIXED, _MIXED47 or _FP64_3M).  No
 * reference counterpart: the mixed mode is new.  Shapes with M, N or K below 64
 * or K above 1024 run in FP64 whatever the mode. */
int qt_zgemm_batched_prec(int32_t opa, int32_t opb, int32_t m, int32_t n, int32_t k, int32_t batch,
                          double alpha, const void* A, int64_t lda, int64_t strideA, const void* B,
                          int64_t ldb, int64_t strideB, double beta, void* R, int64_t ldr,
                          int64_t strideR, int32_t precision);

/* Batched in-place inverse of n x n complex128 blocks (stride n*n), replaces
 * invert (linalg.f90:42-62): partial pivoting; a block whose factorisation
 * meets an exactly zero pivot is zeroed and info[b] = that column (1-based),
 * else info[b] = 0.  Returns QT_ERR_SINGULAR when any block failed. */
int qt_invert_batched(int32_t n, int32_t batch, void* A, int32_t* info);

/* Size of the fixed per-device scratch array of qt_invert_batched_noalloc
 * (64 MiB). */
#define QT_NOALLOC_SCRATCH_BYTES (64u << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the inversion of qt_invert_batched on device-resident
 * blocks, with no allocation of any kind.
 *   A:    device, batch blocks of n x n complex128 (stride n*n), destroyed;
 *   X:    device output, inv(A[b]) in the same layout (must not overlap A);
 *   info: device int32[batch] (0, or the 1-based column of the first zero
 *         pivot, that block of X zeroed) or NULL.
 * The pivot lists, the info flags and the panel multipliers live in one
 * fixed array of QT_NOALLOC_SCRATCH_BYTES per device, owned by the library
 * for the life of the process; a call needs
 *   1280 + 4 batch + 4 n batch + 512 n batch  bytes
 * of it (e.g. up to 254 blocks of 512 or 2000 blocks of 64), else
 * QT_ERR_TOO_LARGE and nothing is launched.  Returns after the inversions
 * finished (QT_ERR_SINGULAR when any block was singular).  Every call on a
 * device shares that device's array, which keeps the last call's contents:
 * the library serialises its own calls, so two calls overlapping on one
 * device (e.g. from a second copy of the library in the process) would race
 * on it. */
int qt_invert_batched_noalloc(int32_t n, int32_t batch, void* A, void* X, int32_t* info);

/* ---- energy-axis correlation / convolution ---------------------------- */

/* Batched corr1d2 (fft_mod.f90:58-91, 'fft' branch):
 *   Z_s(i) = sum_e X_s(e) Y_s(e - l),  l = i - n/2  (0-based i; lag 0 at n/2),
 * for nseries series pairs; series s occupies elements [s*n, s*n + n) of
 * X, Y and Z (complex128), i.e. the energy-fastest layout of
 * G_lesser_by_blocks(:, ik, ix, i) in gw_mod.f90:246-255.  n <= 2048.
 * flags: QT_F_DEVICE_OUTPUTS -> X, Y, Z are device pointers, else host.
 * The reference's result for odd n is undefined (non-conformant array
 * assignment); here odd n keeps the same definition. */
int qt_corr1d2_batched(int32_t n, int64_t nseries, const void* X, const void* Y, void* Z,
                       int32_t flags);

/* Batched conv1d2 (fft_mod.f90:129-168, 'fft' branch):
 *   Z_s(e) = sum_{l=-n/2}^{n-1-n/2} X_s(e - l) Y_s(l + n/2),  X_s = 0 outside [0, n),
 * same layout and flags as qt_corr1d2_batched. */
int qt_conv1d2_batched(int32_t n, int64_t nseries, const void* X, const void* Y, void* Z,
                       int32_t flags);

/* ---- GW polarisation / self-energy assembly --------------------------- */

/* Arrays are element-major complex128 of nm*nm*nx*nk*nen samples, the
 * by_blocks view of gw_mod.f90:246-255: series (ij, ix, ik) starts at
 * ((ij*nx + ix)*nk + ik)*nen with ij = row + col*nm (0-based), nk = nky*nkz.
 * Entries with |row - col| > ndiag are written as zero.  G enters as
 * i*Im(G) (gw_mod.f90:444-446).  dEn = En(2) - En(1).  nen <= 2048.
 * flags: QT_F_DEVICE_OUTPUTS -> every array is a device pointer.  With host
 * arrays the slabs are processed in chunks whose device copies stay within
 * 60% of the free device memory, or within QT_WORKSET_BYTES bytes when that
 * environment variable is set. */

/* Replaces the P loop of gw_ephoton_3d_ijs (gw_mod.f90:444-550):
 *   P^<_ij(q) = dE sum_k corr(G^<_ij(k), G^>_ji(k')),
 *   P^>_ij(q) = dE sum_k corr(G^>_ij(k), G^<_ji(k')),
 *   P^R_ij(q) = dE sum_k [corr(G^<_ij(k), conj G^R_ij(k')) + corr(G^R_ij(k), G^<_ji(k'))],
 * dE = -i dEn/(2 pi) spindeg/nk, k' = k - q + nk/2 folded (gw_mod.f90:478-489). */
int qt_gw_polarization(int32_t nen, int32_t nky, int32_t nkz, int32_t nx, int32_t nm, int32_t ndiag,
                       double dEn, double spindeg, const void* Gr, const void* Gl, const void* Gg,
                       void* Pr, void* Pl, void* Pg, int32_t flags);

/* Replaces the Sigma loop of gw_ephoton_3d_ijs (gw_mod.f90:641-743): the
 * new Sigma^R, Sigma^<, Sigma^> from conv1d2 sums over q of G(k') and W(q),
 * dE = i dEn/(2 pi)/nk, Sigma^R = Re Sigma^R + i Im(Sigma^> - Sigma^<)/2,
 * band zeroing and the (S + S^T)/2 symmetrisation. */
int qt_gw_selfenergy(int32_t nen, int32_t nky, int32_t nkz, int32_t nx, int32_t nm, int32_t ndiag,
                     double dEn, const void* Gr, const void* Gl, const void* Gg, const void* Wr,
                     const void* Wl, const void* Wg, void* Sr, void* Sl, void* Sg, int32_t flags);

/* W solve for nw frequencies, replaces green_rgf_calc_w (called at
 * gw_mod.f90:592-599 with Vii(:,:,:,iq), V1i(:,:,:,iq) and the P / W blocks
 * of each energy; the reference does not define it):
 *   W^R = V (I - P^R V)^-1,  W^< = W^R P^< W^R^H,  W^> = W^R P^> W^R^H,
 * diagonal blocks, closed system (no boundary term).  V is Hermitian and
 * block tridiagonal: Vii (nm x nm x nx) holds V_ii, V1i(:,:,i) = V_{i,i-1}
 * for i >= 2 (the H1i convention of green_RGF_RS, gw_mod.f90:1194-1196).
 * P^R, P^<, P^> are block diagonal.  PR/PL/PG/WR/WL/WG: nw consecutive
 * (nm, nm, nx) stacks (the energy-major by_energies layout, gw_mod.f90:246).
 * Returns QT_ERR_SINGULAR when an inverted block was singular (it is zeroed). */
int qt_gw_wsolve_batched(int32_t nm, int32_t nx, int32_t nw, const void* Vii, const void* V1i,
                         const void* PR, const void* PL, const void* PG, void* WR, void* WL, void* WG,
                         int32_t flags);

/* One GW-SCBA iteration with every array on the device: the loop body of
 * gw_ephoton_3d_ijs (gw_mod.f90:372-749) for one k point and nx slabs of equal
 * size nm (h, no overlap):
 *   G^R, G^<, G^> of the nen energies En (host, equally spaced) by the
 *     green_RGF_RS recursion (gw_mod.f90:1131-1345; p: variant QT_RGF_GW, the
 *     contacts' mu / T and the lead options, as for qt_rgf_solve) with
 *     Sigma^R, Sigma^< of S as the scattering self-energies;
 *   P from G (qt_gw_polarization, ndiag, spindeg), W from P and the Coulomb
 *     blocks Vii, V1i (qt_gw_wsolve_batched; device, (nm, nm, nx)),
 *     Sigma_new from G and W (qt_gw_selfenergy), and
 *     S <- S + alpha (Sigma_new - S) (qt_gw_mix).
 * S[3], G[3], P[3]: element-major (nm*nm, nx, 1, nen) complex128 device
 * arrays (R, <, >): S read and updated, G and P written.  tr, tre (host, nen
 * each, may be NULL): the contact currents of the G pass; cur (host, may be
 * NULL): green_RGF_RS's slab-bond current matrices of the G pass, real
 * (nm, nm, nx, nen) (the cur blocks of qt_rgf_outputs).  Needs device
 * scratch of six such arrays.  QT_ERR_SINGULAR when a block inversion met a
 * zero pivot (the iteration still completes). */
int qt_gw_scba_step(qt_ham* h, const qt_rgf_params* p, int32_t nen, const double* En, int32_t ndiag,
                    double spindeg, double alpha, const void* Vii, const void* V1i, void* const* S,
                    void* const* G, void* const* P, double* tr, double* tre, double* cur);

/* e-photon self-energies in the monochromatic assumption, replaces
 * calc_sigma_ephoton_monochromatic_ext (gw_mod.f90:1075-1122):
 *   Sigma^<>(E) = f M_ix [G^<>(E - hw) + G^<>(E + hw)] M_ix,  f = pre_fact intensity / hw^2,
 *   Sigma^R = i Im(Sigma^> - Sigma^<) / 2 (real part zero).
 * Fortran layouts: Mii (nm, nm, length); Gl, Gg (nm, nm, length, nen + 2 nop),
 * the energy axis extended by nop points (hw = nop dE) on each side, so
 * output energy ie reads G at ie and ie + 2 nop; Sr, Sl, Sg (nm, nm, length, nen).
 * hw = 0 is QT_ERR_INVALID.  flags: QT_F_DEVICE_OUTPUTS -> device pointers.
 * Energies are processed in chunks that keep the device working set (host
 * path: G^<, G^> and the three Sigma of a chunk, plus a product scratch) within
 * 60% of the free device memory, or within the byte count in the environment
 * variable QT_WORKSET_BYTES when it is set. */
int qt_gw_sigma_ephoton(int32_t nm, int32_t length, int32_t nen, int32_t nop, double pre_fact,
                        double intensity, double hw, const void* Mii, const void* Gl, const void* Gg,
                        void* Sr, void* Sl, void* Sg, int32_t flags);

/* Cell-bond current spectra and their energy integrals, the role of
 * calc_block_current (called at gw_mod.f90:384 and :1001 with that argument
 * list; the routine itself is not in the reference's sources, so this is
 * the definition: the NEGF bond current between neighbouring cells of nb
 * orbitals, ns cells per slab, nm = nb ns).  Host arrays, Fortran layouts:
 *   Hii (nm, nm, nx) complex: the slab blocks of H;
 *   GL (nm, nm, nx, nen) complex: the G^< diagonal blocks of every energy;
 *   cur (nm, nm, nx, nen) real: green_RGF_RS's slab-bond matrices
 *        Re H_{l,l+1} G^<_{l+1,l} - Re H_{l+1,l} G^<_{l,l+1} (gw_mod.f90:1250-1273);
 *   jdens (nb, nb, nx ns, nen) real out: for bond c = (l, j) (cell j of slab l
 *        to the next cell), j < ns: Re H_{j,j+1} G^<_{j+1,j} - Re H_{j+1,j} G^<_{j,j+1}
 *        (nb x nb products of the cell blocks of Hii(l) and GL(l)); j = ns: the
 *        sum of the ns diagonal nb x nb blocks of cur(l) (zero for l = nx).
 *        The trace of a bond's matrix is the particle current through it;
 *   tot_cur, tot_ecur (nb, nb, nx ns) real out: sum_e jdens(e) w and
 *        sum_e En[e] jdens(e) w with w = dE e0 / (2 pi hbar) e0 spindeg,
 *        dE = En[1] - En[0] (the reference's terminal-current weight,
 *        gw_mod.f90:424-428; 0 for one energy).
 * Energies are processed in chunks sized like qt_gw_sigma_ephoton's. */
int qt_gw_block_current(int32_t nen, int32_t nx, int32_t nb, int32_t ns, const double* En, double spindeg,
                        const void* Hii, const void* GL, const double* cur, double* jdens, double* tot_cur,
                        double* tot_ecur);

/* Collision integrals, the role of calc_block_collision (called at
 * gw_mod.f90:795 and :926 with that argument list; not in the reference's
 * sources, so this is the definition): per slab block and energy
 *   Ispec(E) = Sigma^<(E) G^>(E) - Sigma^>(E) G^<(E)   (nm x nm products),
 * the scattering-in minus scattering-out term of the kinetic equation for
 * G^<, which vanishes at equilibrium (Sigma and G with one Fermi function);
 *   Itot = sum_e Ispec(e) dE spindeg / (2 pi),  dE = En[1] - En[0].
 * Host arrays (nm, nm, nx, nen) complex in, Ispec the same, Itot (nm, nm, nx).
 * Energies are processed in chunks sized like qt_gw_sigma_ephoton's. */
int qt_gw_block_collision(int32_t nen, int32_t nx, int32_t nm, const double* En, double spindeg,
                          const void* Sl, const void* Sg, const void* Gl, const void* Gg, void* Itot,
                          void* Ispec);

/* S += alpha (Snew - S) over n complex128 samples (gw_mod.f90:747-749). */
int qt_gw_mix(int64_t n, double alpha, void* S, const void* Snew, int32_t flags);

/* ---- device memory helpers (outputs kept resident) -------------------- */
int qt_device_alloc(size_t bytes, void** ptr);
int qt_device_free(void* ptr);
int qt_memcpy_h2d(void* dst, const void* src, size_t bytes);
int qt_memcpy_d2h(void* dst, const void* src, size_t bytes);
int qt_synchronize(void);

/* ---- measurement ----------------------------------------------------- */

typedef struct {
    double total_ms;     /* wall time of the last qt_rgf_solve (device events) */
    double lead_ms;      /* lead decimation + self-energies */
    double forward_ms;   /* first pass */
    double backward_ms;  /* second pass + contact currents */
    double gemm_ms;      /* summed duration of the complex GEMM launches */
    double gemm_flops;   /* algorithmic flops of those launches (8 m n k per
                            complex product, active batch members only) */
    double inv_ms;       /* summed duration of the block inversions */
    double inv_flops;    /* 8 n^3 per inverted block */
    int64_t gemm_launches;
    int64_t inv_launches;
    int64_t lead_iterations; /* sum over points and leads of decimation steps */
    int32_t batch;       /* points per device batch */
} qt_timing;

/* When enabled, every GEMM / inversion launch is bracketed by HIP events on
 * the library stream and summed into qt_timing; qt_last_timing returns the
 * sums of the last qt_rgf_solve or qt_sancho call. */
int qt_set_kernel_timing(int enable);
int qt_last_timing(qt_timing* t);

/* ---- device slicing ------------------------------------------------------ */

/* Cuts a device graph into slices with connections only between neighbouring
 * slices (the block-tridiagonal structure of the sweeps), replaces
 * graph_partition::slice (graph_partition.f90:426-541) and the COO-to-graph
 * step its caller needs (convert_fromCOO, :587-595, has an empty body).
 * Host only: needs no qt_init and no device.
 *   npts points 0..npts-1; nedge undirected edges, point pairs
 *   edges[2e], edges[2e+1] (self pairs and repeats ignored) -- for a matrix,
 *   its nonzero off-diagonal entries;
 *   left[nleft], right[nright]: contact points; nright = 0 slices by
 *   breadth-first distance from the left contact (slice_1contact, :426-450),
 *   else by recursive bisection between the contacts (slice_2contacts,
 *   :452-541), a part of at most nmax points (or no more than its outer
 *   contact) kept as one slice.
 * Out: slice_of[npts] = slice of each point (0-based, slices in order from
 * the left contact), *nslices.  QT_ERR_INVALID for bad ids or when the
 * graph is not connected (a point the left contact cannot reach). */
int qt_graph_slice(int32_t npts, int64_t nedge, const int32_t* edges, int32_t nleft, const int32_t* left,
                   int32_t nright, const int32_t* right, int32_t nmax, int32_t* slice_of, int32_t* nslices);

#ifdef __cplusplus
}
#endif
#endif /* QT_RGF_H */
