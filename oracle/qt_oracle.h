/*
 * qt_oracle.h -- CPU FP64 restatement of the reference RGF path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (quantum_transport_amd) never links it and
 * has no CPU fallback.
 *
 * Every routine follows the reference step by step (operation order,
 * constants, quirks); the reference file:line is cited at each routine.
 * Matrices are column-major complex128 (interleaved re/im, the binary layout
 * of Fortran complex(8)).  Block lists mirror type_matrix_complex lists:
 * arrays of pointers, one per block.
 */
#ifndef QT_ORACLE_H
#define QT_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex qto_c;

enum { QTO_OP_N = 0, QTO_OP_T = 1, QTO_OP_C = 2 };

/* C = alpha*op(A)*op(B) + beta*C (BLAS ZGEMM semantics; called from
 * matrix_c.f90:659 and rgf_mod.f90:524-533). */
void qto_zgemm(int opa, int opb, int m, int n, int k, qto_c alpha,
               const qto_c* A, int lda, const qto_c* B, int ldb, qto_c beta,
               qto_c* C, int ldc);

/* LU with partial pivoting (ZGETRF semantics, IZAMAX pivot measure). */
int qto_zgetrf(int n, qto_c* A, int lda, int* ipiv);
/* Inverse from the LU factors (ZGETRI semantics). */
int qto_zgetri(int n, qto_c* A, int lda, const int* ipiv);
/* linalg.f90:42-62: ZGETRF+ZGETRI in place; on failure A is zeroed and the
 * LAPACK info code is returned. */
int qto_invert(qto_c* A, int n);

/* rgf_mod.f90:475-479 */
double qto_ferm(double a);

/* Sancho-Rubio decimation, rgf_mod.f90:482-568 (eta=1e-5, nmax=50,
 * reset_on_fail=0) and gw_mod.f90:1378-1464 (eta=1e-3, nmax=100,
 * reset_on_fail=1).  Returns the number of iterations run; *converged is
 * set to 1 when the |H01| sum fell below tol.  GBB may be NULL. */
int qto_sancho(int nm, double E, double eta, int nmax, double tol,
               int reset_on_fail, const qto_c* S00, const qto_c* H00,
               const qto_c* H10, qto_c* G00, qto_c* GBB, int* converged);

/* Inputs of one energy point (one (E, k) pair) of the block-tridiagonal
 * device.  nm[i] is the size of slab i (0-based); H1i has nx+1 blocks,
 * H1i[i] couples slab i to slab i-1 (H1i[0]: left lead, H1i[nx]: right
 * lead).  Sii / sigma_* may be NULL, meaning identity / zero. */
typedef struct {
    int nx;
    const int* nm;
    double En;
    double mul, mur, templ, tempr;   /* uniform mul/mur/TEMPl/TEMPr matrices */
    const qto_c* const* Hii;
    const qto_c* const* H1i;
    const qto_c* const* Sii;
    const qto_c* const* sigma_lesser_ph;
    const qto_c* const* sigma_r_ph;
    double lead_eta;                 /* 1e-5 in rgf_mod */
    int lead_nmax;                   /* 50 in rgf_mod */
    double lead_tol;                 /* 1e-100 in rgf_mod */
    int coupling_fix;                /* 0: reference indexing; 1: H1i(i+1) in the
                                        second pass of the forward variant */
} qto_point;

/* Outputs: nx blocks each (G_r, G_lesser, G_greater, Jdens, Gl, Gln), any
 * list may be NULL except Gl/Gln which are the caller workspace, as in the
 * reference.  sancho_iters[2] = {first lead, second lead}. */
typedef struct {
    qto_c* const* G_r;
    qto_c* const* G_lesser;
    qto_c* const* G_greater;
    qto_c* const* Jdens;
    qto_c* const* Gl;
    qto_c* const* Gln;
    double tr, tre;
    int sancho_iters[2];
    int info;                        /* first nonzero invert() info, else 0 */
} qto_result;

/* rgf_mod.f90:46-269 */
int qto_rgf_forward(const qto_point* p, qto_result* r);
/* rgf_mod.f90:272-472 */
int qto_rgf_backward(const qto_point* p, qto_result* r);

/* gw_mod.f90:1131-1345 (green_RGF_RS): RGF inside the GW loop on plain
 * arrays.  All blocks nm x nm, column-major, slab s at offset s*nm*nm.
 * Hii, H1i (nx blocks; H1i[s] couples slab s to s-1, H1i[0] the left lead,
 * H1i[nx-1] also the right lead), sigma_* (nx blocks, may be NULL = zero).
 * Outputs ndens/pdens/ldos (nx blocks), cur (nx real blocks), optional
 * GRi1/GLi1/GGi1 (nx blocks, may be NULL).  TEMP[2], mu[2] as in the
 * reference; lead decimation with eta=1e-3, nmax=100 and reset at nmax. */
int qto_green_rgf_rs(const double* TEMP, int nm, int nx, double E, const double* mu,
                     const qto_c* Hii, const qto_c* H1i, const qto_c* sigma_lesser_ph,
                     const qto_c* sigma_r_ph, qto_c* ndens, qto_c* pdens, qto_c* ldos,
                     double* tr, double* tre, double* cur, qto_c* GRi1, qto_c* GLi1,
                     qto_c* GGi1);

/* Batch driver for the CPU baseline: the E x k loop of negf_mod.f90:115-127
 * with one point per OpenMP thread.  Hii/H1i/Sii hold nk sets (set ik at
 * [ik*nx], [ik*(nx+1)], [ik*nx]); kidx[ip] selects the set of point ip.
 * Outputs the reference observables per point: tr, tre and the diagonals of
 * G_r, G_lesser, G_greater (sum(nm) complex per point, may be NULL). */
int qto_rgf_batch(int variant, int nx, const int* nm, int nk,
                  const qto_c* const* Hii, const qto_c* const* H1i,
                  const qto_c* const* Sii, int npts, const double* En,
                  const int* kidx, double mul, double mur, double templ,
                  double tempr, int coupling_fix, int nthreads, double* tr,
                  double* tre, qto_c* gr_diag, qto_c* gl_diag, qto_c* gg_diag,
                  int* sancho_iters);

/* fft_mod.f90:58-91 (corr1d2) and fft_mod.f90:129-168 (conv1d2), computed
 * as the zero-padded linear correlation / convolution that the 'fft'
 * branch evaluates through a length 2n-1 DFT. */
void qto_corr1d2(int n, const qto_c* X, const qto_c* Y, qto_c* Z);
void qto_conv1d2(int n, const qto_c* X, const qto_c* Y, qto_c* Z);

/* GW assembly on element-major arrays (the by_blocks view of
 * gw_mod.f90:246-255): series (ij, ix, ik) of an array A starts at
 * A + ((ij*nx + ix)*nk + ik)*nen, ij = row + col*nm (0-based), nk = nky*nkz.
 * Only |row - col| <= ndiag is assembled; other entries are zero.  G enters
 * as i*Im(G) (gw_mod.f90:444-446).  dEn = En(2) - En(1). */
/* P^<, P^>, P^R, gw_mod.f90:465-519 */
void qto_gw_polarization(int nen, int nky, int nkz, int nx, int nm, int ndiag, double dEn,
                         double spindeg, const qto_c* Gr, const qto_c* Gl, const qto_c* Gg,
                         qto_c* Pr, qto_c* Pl, qto_c* Pg);
/* new Sigma^R, Sigma^<, Sigma^> including the Sigma^R imaginary-part rule,
 * band zeroing and (S + S^T)/2, gw_mod.f90:641-743 */
void qto_gw_selfenergy(int nen, int nky, int nkz, int nx, int nm, int ndiag, double dEn,
                       const qto_c* Gr, const qto_c* Gl, const qto_c* Gg, const qto_c* Wr,
                       const qto_c* Wl, const qto_c* Wg, qto_c* Sr, qto_c* Sl, qto_c* Sg);
/* W solve for one frequency (the routine called at gw_mod.f90:597, which the
 * reference does not define): W^R = V (I - P^R V)^-1, W^<,> = W^R P^<,> W^R^H,
 * diagonal blocks, on (nm,nm,nx) column-major stacks; V1i(:,:,i) = V_{i,i-1}
 * (i >= 2), V Hermitian, P block-diagonal, closed system.  Returns nonzero
 * when a block inversion was singular. */
int qto_gw_wsolve(int nm, int nx, const qto_c* Vii, const qto_c* V1i, const qto_c* PR,
                  const qto_c* PL, const qto_c* PG, qto_c* WR, qto_c* WL, qto_c* WG);
/* S += alpha (Snew - S), gw_mod.f90:747-749 */
void qto_gw_mix(int64_t n, double alpha, qto_c* S, const qto_c* Snew);
/* e-photon self-energies (gw_mod.f90:1075-1122); Fortran layouts: Mii (nm,nm,length),
 * Gl/Gg (nm,nm,length,nen+2nop), Sr/Sl/Sg (nm,nm,length,nen). */
void qto_gw_sigma_ephoton(int nm, int length, int nen, int nop, double pre_fact, double intensity,
                          double hw, const qto_c* Mii, const qto_c* Gl, const qto_c* Gg, qto_c* Sr,
                          qto_c* Sl, qto_c* Sg);

/* Number of threads the batch driver will use, and the BLAS flavour
 * ("plain" restatement or "mkl"). */
const char* qto_blas_flavour(void);

#ifdef __cplusplus
}
#endif
#endif
