/*
 * qt_oracle.c -- CPU FP64 restatement of the reference RGF path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Two flavours are built from this one file:
 *   plain  -- the BLAS/LAPACK routines the reference calls (ZGEMM, ZGETRF,
 *             ZGETRI) are restated below from their published algorithms;
 *   mkl    -- -DQTO_USE_MKL links the vendor ZGEMM/ZGETRF/ZGETRI, as the
 *             reference links -llapack -lblas (Makefile:5-6).  This is the
 *             flavour timed as the CPU baseline.
 * The RGF, Sancho-Rubio, Fermi and correlation code is shared by both.
 */
#include "qt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BOLTZ 8.61734e-05 /* eV/K, rgf_mod.f90:38 */

/* ------------------------------------------------------------------ */
/* BLAS / LAPACK layer                                                 */
/* ------------------------------------------------------------------ */

#ifdef QTO_USE_MKL
extern void zgemm_(const char*, const char*, const int*, const int*, const int*,
                   const void*, const void*, const int*, const void*, const int*,
                   const void*, void*, const int*);
extern void zgetrf_(const int*, const int*, void*, const int*, int*, int*);
extern void zgetri_(const int*, void*, const int*, const int*, void*, const int*,
                    int*);
extern int MKL_Set_Threading_Layer(int);
#define QTO_MKL_THREADING_SEQUENTIAL 1

static const char* op_char(int op) { return op == QTO_OP_N ? "N" : (op == QTO_OP_T ? "T" : "C"); }

void qto_zgemm(int opa, int opb, int m, int n, int k, qto_c alpha, const qto_c* A,
               int lda, const qto_c* B, int ldb, qto_c beta, qto_c* C, int ldc) {
    zgemm_(op_char(opa), op_char(opb), &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C,
           &ldc);
}

int qto_zgetrf(int n, qto_c* A, int lda, int* ipiv) {
    int info = 0;
    zgetrf_(&n, &n, A, &lda, ipiv, &info);
    for (int i = 0; i < n; ++i) ipiv[i] -= 1; /* keep 0-based pivots internally */
    return info;
}

int qto_zgetri(int n, qto_c* A, int lda, const int* ipiv) {
    int info = 0, lwork = n * n;
    int* ip1 = (int*)malloc(sizeof(int) * (size_t)n);
    qto_c* work = (qto_c*)malloc(sizeof(qto_c) * (size_t)lwork);
    for (int i = 0; i < n; ++i) ip1[i] = ipiv[i] + 1;
    zgetri_(&n, A, &lda, ip1, work, &lwork, &info);
    free(work);
    free(ip1);
    return info;
}

const char* qto_blas_flavour(void) { return "mkl"; }

__attribute__((constructor)) static void qto_mkl_sequential(void) {
    MKL_Set_Threading_Layer(QTO_MKL_THREADING_SEQUENTIAL);
}

#else /* plain restatement */

static inline qto_c op_elem(const qto_c* A, int lda, int op, int i, int k) {
    if (op == QTO_OP_N) return A[i + (size_t)k * lda];
    if (op == QTO_OP_T) return A[k + (size_t)i * lda];
    return conj(A[k + (size_t)i * lda]);
}

/* ZGEMM, reference BLAS loop order: column j of C, then the k-sum. */
void qto_zgemm(int opa, int opb, int m, int n, int k, qto_c alpha, const qto_c* A,
               int lda, const qto_c* B, int ldb, qto_c beta, qto_c* C, int ldc) {
    for (int j = 0; j < n; ++j) {
        qto_c* c = C + (size_t)j * ldc;
        if (beta == 0.0) {
            for (int i = 0; i < m; ++i) c[i] = 0.0;
        } else if (beta != 1.0) {
            for (int i = 0; i < m; ++i) c[i] *= beta;
        }
        if (opa == QTO_OP_N) {
            for (int l = 0; l < k; ++l) {
                qto_c t = alpha * op_elem(B, ldb, opb, l, j);
                const qto_c* a = A + (size_t)l * lda;
                for (int i = 0; i < m; ++i) c[i] += t * a[i];
            }
        } else {
            for (int i = 0; i < m; ++i) {
                const qto_c* a = A + (size_t)i * lda;
                qto_c s = 0.0;
                if (opa == QTO_OP_T)
                    for (int l = 0; l < k; ++l) s += a[l] * op_elem(B, ldb, opb, l, j);
                else
                    for (int l = 0; l < k; ++l) s += conj(a[l]) * op_elem(B, ldb, opb, l, j);
                c[i] += alpha * s;
            }
        }
    }
}

static inline double cabs1(qto_c z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* ZGETRF (unblocked ZGETF2 form): pivot = first index of max |re|+|im|
 * (IZAMAX), full-row interchange, column scaling, rank-1 update. */
int qto_zgetrf(int n, qto_c* A, int lda, int* ipiv) {
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = cabs1(A[j + (size_t)j * lda]);
        for (int i = j + 1; i < n; ++i) {
            double v = cabs1(A[i + (size_t)j * lda]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[j] = p;
        if (A[p + (size_t)j * lda] != 0.0) {
            if (p != j)
                for (int c = 0; c < n; ++c) {
                    qto_c t = A[j + (size_t)c * lda];
                    A[j + (size_t)c * lda] = A[p + (size_t)c * lda];
                    A[p + (size_t)c * lda] = t;
                }
            qto_c r = 1.0 / A[j + (size_t)j * lda];
            for (int i = j + 1; i < n; ++i) A[i + (size_t)j * lda] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; ++c) {
            qto_c u = A[j + (size_t)c * lda];
            if (u == 0.0) continue;
            qto_c* col = A + (size_t)c * lda;
            const qto_c* lj = A + (size_t)j * lda;
            for (int i = j + 1; i < n; ++i) col[i] -= lj[i] * u;
        }
    }
    return info;
}

/* ZGETRI: inv(U) in place (ZTRTI2, upper, non-unit), then solve
 * inv(A)*L = inv(U) column by column from the right, then undo the row
 * interchanges as column interchanges in reverse order. */
int qto_zgetri(int n, qto_c* A, int lda, const int* ipiv) {
    for (int j = 0; j < n; ++j)
        if (A[j + (size_t)j * lda] == 0.0) return j + 1;
    for (int j = 0; j < n; ++j) {
        qto_c* cj = A + (size_t)j * lda;
        cj[j] = 1.0 / cj[j];
        qto_c ajj = -cj[j];
        /* x(0:j) := triu(A(0:j,0:j)) * x, x = A(0:j, j) (ZTRMV upper, no-trans) */
        for (int l = 0; l < j; ++l) {
            qto_c t = cj[l];
            if (t == 0.0) continue;
            const qto_c* cl = A + (size_t)l * lda;
            for (int i = 0; i < l; ++i) cj[i] += t * cl[i];
            cj[l] = t * cl[l];
        }
        for (int i = 0; i < j; ++i) cj[i] *= ajj;
    }
    qto_c* work = (qto_c*)malloc(sizeof(qto_c) * (size_t)n);
    for (int j = n - 1; j >= 0; --j) {
        qto_c* cj = A + (size_t)j * lda;
        for (int i = j + 1; i < n; ++i) { work[i] = cj[i]; cj[i] = 0.0; }
        for (int l = j + 1; l < n; ++l) {
            qto_c t = work[l];
            if (t == 0.0) continue;
            const qto_c* cl = A + (size_t)l * lda;
            for (int i = 0; i < n; ++i) cj[i] -= t * cl[i];
        }
    }
    free(work);
    for (int j = n - 2; j >= 0; --j) {
        int jp = ipiv[j];
        if (jp != j)
            for (int i = 0; i < n; ++i) {
                qto_c t = A[i + (size_t)j * lda];
                A[i + (size_t)j * lda] = A[i + (size_t)jp * lda];
                A[i + (size_t)jp * lda] = t;
            }
    }
    return 0;
}

const char* qto_blas_flavour(void) { return "plain"; }
#endif

int qto_invert(qto_c* A, int n) {
    int* ipiv = (int*)malloc(sizeof(int) * (size_t)n);
    int info = qto_zgetrf(n, A, n, ipiv);
    if (info == 0) info = qto_zgetri(n, A, n, ipiv);
    if (info != 0) memset(A, 0, sizeof(qto_c) * (size_t)n * n);
    free(ipiv);
    return info;
}

double qto_ferm(double a) { return 1.0 / (1.0 + exp(a)); }

/* ------------------------------------------------------------------ */
/* small helpers mirroring matrix_c MUL_C / triMUL_C                   */
/* ------------------------------------------------------------------ */

typedef struct { qto_c* m; int r, c; } mat;

static qto_c* zalloc(size_t n) { return (qto_c*)calloc(n ? n : 1, sizeof(qto_c)); }

/* R = op(A)*op(B), matrix_c.f90:615-660; returns the shape of R. */
static void mulc(const qto_c* A, int ra, int ca, int opa, const qto_c* B, int rb, int cb,
                 int opb, qto_c* R, int* rm, int* rn) {
    int m = opa == QTO_OP_N ? ra : ca, k = opa == QTO_OP_N ? ca : ra;
    int n = opb == QTO_OP_N ? cb : rb;
    qto_zgemm(opa, opb, m, n, k, 1.0, A, ra, B, rb, 0.0, R, m);
    if (rm) *rm = m;
    if (rn) *rn = n;
}

/* R = (op(A)*op(B))*op(C), matrix_c.f90:576-613 */
static void trimulc(const qto_c* A, int ra, int ca, int opa, const qto_c* B, int rb,
                    int cb, int opb, const qto_c* C, int rc, int cc, int opc, qto_c* R,
                    qto_c* tmp) {
    int tm, tn;
    mulc(A, ra, ca, opa, B, rb, cb, opb, tmp, &tm, &tn);
    mulc(tmp, tm, tn, QTO_OP_N, C, rc, cc, opc, R, NULL, NULL);
}

static void dagger(const qto_c* A, int r, int c, qto_c* R) { /* R (c x r) = A^H */
    for (int j = 0; j < c; ++j)
        for (int i = 0; i < r; ++i) R[j + (size_t)i * c] = conj(A[i + (size_t)j * r]);
}

/* ------------------------------------------------------------------ */
/* Sancho-Rubio, rgf_mod.f90:482-568 / gw_mod.f90:1378-1464            */
/* ------------------------------------------------------------------ */

int qto_sancho(int nm, double E, double eta, int nmax, double tol, int reset_on_fail,
               const qto_c* S00, const qto_c* H00, const qto_c* H10, qto_c* G00,
               qto_c* GBB, int* converged) {
    size_t nn = (size_t)nm * nm;
    qto_c *A = zalloc(nn), *B = zalloc(nn), *C = zalloc(nn);
    qto_c *H_BB = zalloc(nn), *H_SS = zalloc(nn), *H_01 = zalloc(nn), *H_10 = zalloc(nn);
    qto_c z = E + I * eta;
    memcpy(H_BB, H00, nn * sizeof(qto_c));
    memcpy(H_10, H10, nn * sizeof(qto_c));
    dagger(H_10, nm, nm, H_01);
    memcpy(H_SS, H00, nn * sizeof(qto_c));
    int it, conv = 0;
    for (it = 1; it <= nmax; ++it) {
        for (size_t e = 0; e < nn; ++e) A[e] = z * S00[e] - H_BB[e];
        qto_invert(A, nm);
        qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, A, nm, H_10, nm, 0.0, B, nm);
        qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, H_01, nm, B, nm, 0.0, C, nm);
        for (size_t e = 0; e < nn; ++e) { H_SS[e] += C[e]; H_BB[e] += C[e]; }
        qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, H_10, nm, B, nm, 0.0, C, nm);
        qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, A, nm, H_01, nm, 0.0, B, nm);
        qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, H_10, nm, B, nm, 0.0, A, nm);
        memcpy(H_10, C, nn * sizeof(qto_c));
        for (size_t e = 0; e < nn; ++e) H_BB[e] += A[e];
        qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, H_01, nm, B, nm, 0.0, C, nm);
        memcpy(H_01, C, nn * sizeof(qto_c));
        double error = 0.0;
        for (int k = 0; k < nm; ++k)
            for (int j = 0; j < nm; ++j) {
                qto_c v = C[k + (size_t)j * nm];
                error += sqrt(cimag(v) * cimag(v) + creal(v) * creal(v));
            }
        if (fabs(error) < tol) { conv = 1; break; }
        if (it == nmax && reset_on_fail) {
            memcpy(H_SS, H00, nn * sizeof(qto_c));
            memcpy(H_BB, H00, nn * sizeof(qto_c));
        }
    }
    if (it > nmax) it = nmax;
    for (size_t e = 0; e < nn; ++e) G00[e] = z * S00[e] - H_SS[e];
    qto_invert(G00, nm);
    if (GBB) {
        for (size_t e = 0; e < nn; ++e) GBB[e] = z * S00[e] - H_BB[e];
        qto_invert(GBB, nm);
    }
    free(A); free(B); free(C); free(H_BB); free(H_SS); free(H_01); free(H_10);
    if (converged) *converged = conv;
    return it;
}

/* ------------------------------------------------------------------ */
/* RGF sweeps                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    int M;
    qto_c *A, *B, *C, *T, *H00, *H10, *G00, *sig, *sigL, *GN0, *ident;
} work_t;

static void work_init(work_t* w, int Mmax) {
    size_t nn = (size_t)Mmax * Mmax;
    w->M = Mmax;
    w->A = zalloc(nn); w->B = zalloc(nn); w->C = zalloc(nn); w->T = zalloc(nn);
    w->H00 = zalloc(nn); w->H10 = zalloc(nn); w->G00 = zalloc(nn); w->sig = zalloc(nn);
    w->sigL = zalloc(nn); w->GN0 = zalloc(nn); w->ident = zalloc(nn);
    for (int i = 0; i < Mmax; ++i) w->ident[i + (size_t)i * Mmax] = 1.0;
}

static void work_free(work_t* w) {
    free(w->A); free(w->B); free(w->C); free(w->T); free(w->H00); free(w->H10);
    free(w->G00); free(w->sig); free(w->sigL); free(w->GN0); free(w->ident);
}

/* identity of size M in a buffer laid out with leading dimension M */
static const qto_c* ident_of(work_t* w, int M, qto_c* buf) {
    memset(buf, 0, sizeof(qto_c) * (size_t)M * M);
    for (int i = 0; i < M; ++i) buf[i + (size_t)i * M] = 1.0;
    (void)w;
    return buf;
}

/* H00 = Hii + sigma_r_ph*S  (rgf_mod.f90:77-79) */
static void build_h00(const qto_point* p, int s, int M, const qto_c* S, qto_c* H00, qto_c* tmp) {
    size_t nn = (size_t)M * M;
    if (p->sigma_r_ph && p->sigma_r_ph[s]) {
        mulc(p->sigma_r_ph[s], M, M, QTO_OP_N, S, M, M, QTO_OP_N, tmp, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) H00[e] = p->Hii[s][e] + tmp[e];
    } else {
        for (size_t e = 0; e < nn; ++e) H00[e] = p->Hii[s][e];
    }
}

/* out += sigma_lesser_ph*S  (rgf_mod.f90:99,132) */
static void add_sigl(const qto_point* p, int s, int M, const qto_c* S, qto_c* out, qto_c* tmp) {
    if (p->sigma_lesser_ph && p->sigma_lesser_ph[s]) {
        mulc(p->sigma_lesser_ph[s], M, M, QTO_OP_N, S, M, M, QTO_OP_N, tmp, NULL, NULL);
        for (size_t e = 0; e < (size_t)M * M; ++e) out[e] += tmp[e];
    }
}

static int uniform_sizes(const qto_point* p) {
    for (int s = 1; s < p->nx; ++s)
        if (p->nm[s] != p->nm[0]) return 0;
    return 1;
}

static void note_info(qto_result* r, int info) {
    if (info != 0 && r->info == 0) r->info = info;
}

int qto_rgf_forward(const qto_point* p, qto_result* r) {
    const int nx = p->nx;
    if (nx < 2) return -1;
    if (!p->coupling_fix && !uniform_sizes(p)) return -2;
    int Mmax = 0;
    for (int s = 0; s < nx; ++s) Mmax = p->nm[s] > Mmax ? p->nm[s] : Mmax;
    work_t w;
    work_init(&w, Mmax);
    qto_c* Sbuf = zalloc((size_t)Mmax * Mmax);
    qto_c* tmp2 = zalloc((size_t)Mmax * Mmax);
    const qto_c z = p->En + 0.0 * I; /* rgf_mod.f90:62 */
    r->info = 0;

    /* ---- left contact, rgf_mod.f90:63-104 */
    int s = 0, M = p->nm[0];
    size_t nn = (size_t)M * M;
    const qto_c* S = p->Sii ? p->Sii[0] : ident_of(&w, M, Sbuf);
    build_h00(p, 0, M, S, w.H00, w.B);
    dagger(p->H1i[0], M, M, w.H10);
    r->sancho_iters[0] = qto_sancho(M, p->En, p->lead_eta, p->lead_nmax, p->lead_tol, 0, S,
                                    w.H00, w.H10, w.G00, NULL, NULL);
    trimulc(p->H1i[0], M, M, QTO_OP_N, w.G00, M, M, QTO_OP_N, p->H1i[0], M, M, QTO_OP_C,
            w.sigL, w.T);
    for (size_t e = 0; e < nn; ++e) w.B[e] = z * S[e] - w.H00[e] - w.sigL[e];
    note_info(r, qto_invert(w.B, M));
    memcpy(r->Gl[0], w.B, nn * sizeof(qto_c));
    const double f_first = qto_ferm((p->En - p->mur) / (BOLTZ * p->tempr));
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i)
            w.sig[i + (size_t)j * M] =
                -(w.sigL[i + (size_t)j * M] - conj(w.sigL[j + (size_t)i * M])) * f_first;
    add_sigl(p, 0, M, S, w.sig, w.B);
    trimulc(r->Gl[0], M, M, QTO_OP_N, w.sig, M, M, QTO_OP_N, r->Gl[0], M, M, QTO_OP_C,
            r->Gln[0], w.T);

    /* ---- first pass l -> r, rgf_mod.f90:113-136 */
    for (s = 1; s <= nx - 2; ++s) {
        M = p->nm[s];
        int Mp = p->nm[s - 1];
        nn = (size_t)M * M;
        S = p->Sii ? p->Sii[s] : ident_of(&w, M, Sbuf);
        build_h00(p, s, M, S, w.H00, w.B);
        trimulc(p->H1i[s], M, Mp, QTO_OP_N, r->Gl[s - 1], Mp, Mp, QTO_OP_N, p->H1i[s], M, Mp,
                QTO_OP_C, w.B, w.T);
        for (size_t e = 0; e < nn; ++e) w.A[e] = z * S[e] - w.H00[e] - w.B[e];
        note_info(r, qto_invert(w.A, M));
        memcpy(r->Gl[s], w.A, nn * sizeof(qto_c));
        trimulc(p->H1i[s], M, Mp, QTO_OP_N, r->Gln[s - 1], Mp, Mp, QTO_OP_N, p->H1i[s], M,
                Mp, QTO_OP_C, w.B, w.T);
        add_sigl(p, s, M, S, w.B, w.A);
        trimulc(r->Gl[s], M, M, QTO_OP_N, w.B, M, M, QTO_OP_N, r->Gl[s], M, M, QTO_OP_C,
                r->Gln[s], w.T);
    }

    /* ---- right contact, rgf_mod.f90:142-202 */
    s = nx - 1;
    M = p->nm[s];
    int Mp = p->nm[s - 1];
    nn = (size_t)M * M;
    S = p->Sii ? p->Sii[s] : ident_of(&w, M, Sbuf);
    build_h00(p, s, M, S, w.H00, w.B);
    memcpy(w.H10, p->H1i[nx], nn * sizeof(qto_c));
    r->sancho_iters[1] = qto_sancho(M, p->En, p->lead_eta, p->lead_nmax, p->lead_tol, 0, S,
                                    w.H00, w.H10, w.G00, NULL, NULL);
    qto_c* sigR = tmp2;
    trimulc(w.H10, M, M, QTO_OP_C, w.G00, M, M, QTO_OP_N, w.H10, M, M, QTO_OP_N, sigR, w.T);
    trimulc(p->H1i[s], M, Mp, QTO_OP_N, r->Gl[s - 1], Mp, Mp, QTO_OP_N, p->H1i[s], M, Mp,
            QTO_OP_C, w.B, w.T);
    for (size_t e = 0; e < nn; ++e) w.A[e] = z * S[e] - w.H00[e] - w.B[e] - sigR[e];
    note_info(r, qto_invert(w.A, M));
    qto_c* Gr_last = r->G_r && r->G_r[s] ? r->G_r[s] : w.GN0;
    memcpy(Gr_last, w.A, nn * sizeof(qto_c));
    memcpy(r->Gl[s], w.A, nn * sizeof(qto_c));
    trimulc(p->H1i[s], M, Mp, QTO_OP_N, r->Gln[s - 1], Mp, Mp, QTO_OP_N, p->H1i[s], M, Mp,
            QTO_OP_C, w.B, w.T);
    memset(w.A, 0, nn * sizeof(qto_c));
    add_sigl(p, s, M, S, w.A, w.C);
    const double f_last = qto_ferm((p->En - p->mul) / (BOLTZ * p->templ));
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.sig[e] = -(sigR[e] - conj(sigR[j + (size_t)i * M])) * f_last;
            w.sig[e] = w.sig[e] + w.A[e] + w.B[e];
        }
    /* G_lesser / G_greater of the last slab; they seed the second pass, so
     * they live in the caller's lists (or in scratch when not requested). */
    qto_c** GL = (qto_c**)calloc((size_t)nx, sizeof(qto_c*));
    qto_c** GG = (qto_c**)calloc((size_t)nx, sizeof(qto_c*));
    qto_c** GR = (qto_c**)calloc((size_t)nx, sizeof(qto_c*));
    qto_c** own = (qto_c**)calloc((size_t)3 * nx, sizeof(qto_c*));
    for (int t = 0; t < nx; ++t) {
        size_t bt = (size_t)p->nm[t] * p->nm[t];
        GR[t] = r->G_r && r->G_r[t] ? r->G_r[t] : (own[3 * t] = zalloc(bt));
        GL[t] = r->G_lesser && r->G_lesser[t] ? r->G_lesser[t] : (own[3 * t + 1] = zalloc(bt));
        GG[t] = r->G_greater && r->G_greater[t] ? r->G_greater[t] : (own[3 * t + 2] = zalloc(bt));
    }
    if (GR[s] != Gr_last) memcpy(GR[s], Gr_last, nn * sizeof(qto_c));
    trimulc(GR[s], M, M, QTO_OP_N, w.sig, M, M, QTO_OP_N, GR[s], M, M, QTO_OP_C, GL[s], w.T);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            GG[s][e] = GL[s][e] + (GR[s][e] - conj(GR[s][j + (size_t)i * M]));
        }
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(sigR[e] - conj(sigR[j + (size_t)i * M])) * f_last;
        }
    mulc(w.A, M, M, QTO_OP_N, GG[s], M, M, QTO_OP_N, w.B, NULL, NULL);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(sigR[e] - conj(sigR[j + (size_t)i * M])) * (f_last - 1.0);
        }
    mulc(w.A, M, M, QTO_OP_N, GL[s], M, M, QTO_OP_N, w.C, NULL, NULL);
    double tim = 0.0;
    for (int j = 0; j < M; ++j) tim += creal(w.B[j + (size_t)j * M] - w.C[j + (size_t)j * M]);
    if (r->Jdens && r->Jdens[s])
        for (size_t e = 0; e < nn; ++e) r->Jdens[s][e] = w.B[e] - w.C[e];
    r->tr = tim;

    /* ---- second pass r -> l, rgf_mod.f90:212-250 */
    for (s = nx - 2; s >= 0; --s) {
        M = p->nm[s];
        int Mn = p->nm[s + 1];
        nn = (size_t)M * M;
        /* coupling between slab s and s+1: the reference uses H1i(ii) (the
         * coupling of slab s to s-1), identical for uniform devices */
        const qto_c* W = p->coupling_fix ? p->H1i[s + 1] : p->H1i[s];
        qto_c* Gl = r->Gl[s];
        qto_c* Gln = r->Gln[s];
        /* A = G<(i+1) W Gl^H + G(i+1) W Gln  (Mn x M) */
        trimulc(GL[s + 1], Mn, Mn, QTO_OP_N, W, Mn, M, QTO_OP_N, Gl, M, M, QTO_OP_C, w.A, w.T);
        trimulc(GR[s + 1], Mn, Mn, QTO_OP_N, W, Mn, M, QTO_OP_N, Gln, M, M, QTO_OP_N, w.B, w.T);
        for (size_t e = 0; e < (size_t)Mn * M; ++e) w.A[e] += w.B[e];
        mulc(W, Mn, M, QTO_OP_C, w.A, Mn, M, QTO_OP_N, w.B, NULL, NULL);
        if (r->Jdens && r->Jdens[s])
            for (size_t e = 0; e < nn; ++e) r->Jdens[s][e] = -2.0 * w.B[e];
        /* GN0 = Gl W^H G(i+1);  G = Gl + GN0 W Gl */
        mulc(Gl, M, M, QTO_OP_N, W, Mn, M, QTO_OP_C, w.B, NULL, NULL);
        mulc(w.B, M, Mn, QTO_OP_N, GR[s + 1], Mn, Mn, QTO_OP_N, w.GN0, NULL, NULL);
        mulc(w.GN0, M, Mn, QTO_OP_N, W, Mn, M, QTO_OP_N, w.C, NULL, NULL);
        mulc(w.C, M, M, QTO_OP_N, Gl, M, M, QTO_OP_N, w.A, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GR[s][e] = Gl[e] + w.A[e];
        /* G< = Gln + Gl W^H G<(i+1) W Gl^H */
        mulc(Gl, M, M, QTO_OP_N, W, Mn, M, QTO_OP_C, w.B, NULL, NULL);
        mulc(w.B, M, Mn, QTO_OP_N, GL[s + 1], Mn, Mn, QTO_OP_N, w.C, NULL, NULL);
        mulc(w.C, M, Mn, QTO_OP_N, W, Mn, M, QTO_OP_N, w.A, NULL, NULL);
        mulc(w.A, M, M, QTO_OP_N, Gl, M, M, QTO_OP_C, w.C, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GL[s][e] = Gln[e] + w.C[e];
        /* + GN0 W Gln */
        mulc(w.GN0, M, Mn, QTO_OP_N, W, Mn, M, QTO_OP_N, w.B, NULL, NULL);
        mulc(w.B, M, M, QTO_OP_N, Gln, M, M, QTO_OP_N, w.C, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GL[s][e] += w.C[e];
        /* + Gln W^H GN0^H */
        mulc(Gln, M, M, QTO_OP_N, W, Mn, M, QTO_OP_C, w.B, NULL, NULL);
        mulc(w.B, M, Mn, QTO_OP_N, w.GN0, M, Mn, QTO_OP_C, w.C, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GL[s][e] += w.C[e];
        for (int j = 0; j < M; ++j)
            for (int i = 0; i < M; ++i) {
                size_t e = i + (size_t)j * M;
                GG[s][e] = GL[s][e] + (GR[s][e] - conj(GR[s][j + (size_t)i * M]));
            }
    }

    /* ---- left contact current, rgf_mod.f90:256-266 */
    M = p->nm[0];
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(w.sigL[e] - conj(w.sigL[j + (size_t)i * M])) * f_first;
        }
    mulc(w.A, M, M, QTO_OP_N, GG[0], M, M, QTO_OP_N, w.B, NULL, NULL);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(w.sigL[e] - conj(w.sigL[j + (size_t)i * M])) * (f_first - 1.0);
        }
    mulc(w.A, M, M, QTO_OP_N, GL[0], M, M, QTO_OP_N, w.C, NULL, NULL);
    tim = 0.0;
    for (int j = 0; j < M; ++j) tim += creal(w.B[j + (size_t)j * M] - w.C[j + (size_t)j * M]);
    r->tre = tim;

    for (int t = 0; t < 3 * nx; ++t) free(own[t]);
    free(own); free(GL); free(GG); free(GR);
    free(Sbuf); free(tmp2);
    work_free(&w);
    return 0;
}

int qto_rgf_backward(const qto_point* p, qto_result* r) {
    const int nx = p->nx;
    if (nx < 2) return -1;
    if (!uniform_sizes(p)) return -2;
    const int M = p->nm[0];
    const size_t nn = (size_t)M * M;
    work_t w;
    work_init(&w, M);
    qto_c* Sbuf = zalloc(nn);
    qto_c* sigR = zalloc(nn);
    const qto_c z = p->En + 0.0 * I; /* rgf_mod.f90:287 */
    r->info = 0;
    qto_c** GL = (qto_c**)calloc((size_t)nx, sizeof(qto_c*));
    qto_c** GG = (qto_c**)calloc((size_t)nx, sizeof(qto_c*));
    qto_c** GR = (qto_c**)calloc((size_t)nx, sizeof(qto_c*));
    qto_c** own = (qto_c**)calloc((size_t)3 * nx, sizeof(qto_c*));
    for (int t = 0; t < nx; ++t) {
        GR[t] = r->G_r && r->G_r[t] ? r->G_r[t] : (own[3 * t] = zalloc(nn));
        GL[t] = r->G_lesser && r->G_lesser[t] ? r->G_lesser[t] : (own[3 * t + 1] = zalloc(nn));
        GG[t] = r->G_greater && r->G_greater[t] ? r->G_greater[t] : (own[3 * t + 2] = zalloc(nn));
    }

    /* ---- right contact, rgf_mod.f90:289-325 */
    int s = nx - 1;
    const qto_c* S = p->Sii ? p->Sii[s] : ident_of(&w, M, Sbuf);
    build_h00(p, s, M, S, w.H00, w.B);
    r->sancho_iters[0] = qto_sancho(M, p->En, p->lead_eta, p->lead_nmax, p->lead_tol, 0, S,
                                    w.H00, p->H1i[nx], w.G00, NULL, NULL);
    trimulc(p->H1i[nx], M, M, QTO_OP_N, w.G00, M, M, QTO_OP_N, p->H1i[nx], M, M, QTO_OP_C,
            sigR, w.T);
    for (size_t e = 0; e < nn; ++e) w.B[e] = z * S[e] - w.H00[e] - sigR[e];
    note_info(r, qto_invert(w.B, M));
    memcpy(r->Gl[s], w.B, nn * sizeof(qto_c));
    const double f_first = qto_ferm((p->En - p->mur) / (BOLTZ * p->tempr));
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i)
            w.sig[i + (size_t)j * M] =
                -(sigR[i + (size_t)j * M] - conj(sigR[j + (size_t)i * M])) * f_first;
    add_sigl(p, s, M, S, w.sig, w.B);
    trimulc(r->Gl[s], M, M, QTO_OP_N, w.sig, M, M, QTO_OP_N, r->Gl[s], M, M, QTO_OP_C,
            r->Gln[s], w.T);

    /* ---- first pass r -> l, rgf_mod.f90:330-353 */
    for (s = nx - 2; s >= 1; --s) {
        S = p->Sii ? p->Sii[s] : ident_of(&w, M, Sbuf);
        build_h00(p, s, M, S, w.H00, w.B);
        trimulc(p->H1i[s + 1], M, M, QTO_OP_N, r->Gl[s + 1], M, M, QTO_OP_N, p->H1i[s + 1],
                M, M, QTO_OP_C, w.B, w.T);
        for (size_t e = 0; e < nn; ++e) w.A[e] = z * S[e] - w.H00[e] - w.B[e];
        note_info(r, qto_invert(w.A, M));
        memcpy(r->Gl[s], w.A, nn * sizeof(qto_c));
        trimulc(p->H1i[s + 1], M, M, QTO_OP_N, r->Gln[s + 1], M, M, QTO_OP_N, p->H1i[s + 1],
                M, M, QTO_OP_C, w.B, w.T);
        add_sigl(p, s, M, S, w.B, w.A);
        trimulc(r->Gl[s], M, M, QTO_OP_N, w.B, M, M, QTO_OP_N, r->Gl[s], M, M, QTO_OP_C,
                r->Gln[s], w.T);
    }

    /* ---- left contact, rgf_mod.f90:355-415 */
    s = 0;
    S = p->Sii ? p->Sii[0] : ident_of(&w, M, Sbuf);
    build_h00(p, 0, M, S, w.H00, w.B);
    dagger(p->H1i[0], M, M, w.H10);
    r->sancho_iters[1] = qto_sancho(M, p->En, p->lead_eta, p->lead_nmax, p->lead_tol, 0, S,
                                    w.H00, w.H10, w.G00, NULL, NULL);
    trimulc(p->H1i[0], M, M, QTO_OP_C, w.G00, M, M, QTO_OP_N, p->H1i[0], M, M, QTO_OP_N,
            w.sigL, w.T);
    trimulc(p->H1i[1], M, M, QTO_OP_N, r->Gl[1], M, M, QTO_OP_N, p->H1i[1], M, M, QTO_OP_C,
            w.B, w.T);
    for (size_t e = 0; e < nn; ++e) w.A[e] = z * S[e] - w.H00[e] - w.B[e] - w.sigL[e];
    note_info(r, qto_invert(w.A, M));
    memcpy(GR[0], w.A, nn * sizeof(qto_c));
    memcpy(r->Gl[0], w.A, nn * sizeof(qto_c));
    trimulc(p->H1i[1], M, M, QTO_OP_N, r->Gln[1], M, M, QTO_OP_N, p->H1i[1], M, M, QTO_OP_C,
            w.B, w.T);
    memset(w.A, 0, nn * sizeof(qto_c));
    add_sigl(p, 0, M, S, w.A, w.C);
    const double f_last = qto_ferm((p->En - p->mul) / (BOLTZ * p->templ));
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.sig[e] = -(w.sigL[e] - conj(w.sigL[j + (size_t)i * M])) * f_last;
            w.sig[e] = w.sig[e] + w.A[e] + w.B[e];
        }
    trimulc(GR[0], M, M, QTO_OP_N, w.sig, M, M, QTO_OP_N, GR[0], M, M, QTO_OP_C, GL[0], w.T);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            GG[0][e] = GL[0][e] + (GR[0][e] - conj(GR[0][j + (size_t)i * M]));
        }
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(w.sigL[e] - conj(w.sigL[j + (size_t)i * M])) * f_last;
        }
    mulc(w.A, M, M, QTO_OP_N, GG[0], M, M, QTO_OP_N, w.B, NULL, NULL);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(w.sigL[e] - conj(w.sigL[j + (size_t)i * M])) * (f_last - 1.0);
        }
    mulc(w.A, M, M, QTO_OP_N, GL[0], M, M, QTO_OP_N, w.C, NULL, NULL);
    if (r->Jdens && r->Jdens[0])
        for (size_t e = 0; e < nn; ++e) r->Jdens[0][e] = w.B[e] - w.C[e];
    double tim = 0.0;
    for (int j = 0; j < M; ++j) tim += creal(w.B[j + (size_t)j * M] - w.C[j + (size_t)j * M]);
    r->tre = tim;

    /* ---- second pass l -> r, rgf_mod.f90:420-458 */
    for (s = 1; s <= nx - 1; ++s) {
        const qto_c* W = p->H1i[s];
        qto_c* Gl = r->Gl[s];
        qto_c* Gln = r->Gln[s];
        trimulc(GL[s - 1], M, M, QTO_OP_N, W, M, M, QTO_OP_N, Gl, M, M, QTO_OP_C, w.A, w.T);
        trimulc(GR[s - 1], M, M, QTO_OP_N, W, M, M, QTO_OP_N, Gln, M, M, QTO_OP_N, w.B, w.T);
        for (size_t e = 0; e < nn; ++e) w.A[e] += w.B[e];
        mulc(W, M, M, QTO_OP_C, w.A, M, M, QTO_OP_N, w.B, NULL, NULL);
        if (r->Jdens && r->Jdens[s])
            for (size_t e = 0; e < nn; ++e) r->Jdens[s][e] = -2.0 * w.B[e];
        mulc(Gl, M, M, QTO_OP_N, W, M, M, QTO_OP_C, w.B, NULL, NULL);
        mulc(w.B, M, M, QTO_OP_N, GR[s - 1], M, M, QTO_OP_N, w.GN0, NULL, NULL);
        mulc(w.GN0, M, M, QTO_OP_N, W, M, M, QTO_OP_N, w.C, NULL, NULL);
        mulc(w.C, M, M, QTO_OP_N, Gl, M, M, QTO_OP_N, w.A, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GR[s][e] = Gl[e] + w.A[e];
        mulc(Gl, M, M, QTO_OP_N, W, M, M, QTO_OP_C, w.B, NULL, NULL);
        mulc(w.B, M, M, QTO_OP_N, GL[s - 1], M, M, QTO_OP_N, w.C, NULL, NULL);
        mulc(w.C, M, M, QTO_OP_N, W, M, M, QTO_OP_N, w.A, NULL, NULL);
        mulc(w.A, M, M, QTO_OP_N, Gl, M, M, QTO_OP_C, w.C, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GL[s][e] = Gln[e] + w.C[e];
        mulc(w.GN0, M, M, QTO_OP_N, W, M, M, QTO_OP_N, w.B, NULL, NULL);
        mulc(w.B, M, M, QTO_OP_N, Gln, M, M, QTO_OP_N, w.C, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GL[s][e] += w.C[e];
        mulc(Gln, M, M, QTO_OP_N, W, M, M, QTO_OP_C, w.B, NULL, NULL);
        mulc(w.B, M, M, QTO_OP_N, w.GN0, M, M, QTO_OP_C, w.C, NULL, NULL);
        for (size_t e = 0; e < nn; ++e) GL[s][e] += w.C[e];
        for (int j = 0; j < M; ++j)
            for (int i = 0; i < M; ++i) {
                size_t e = i + (size_t)j * M;
                GG[s][e] = GL[s][e] + (GR[s][e] - conj(GR[s][j + (size_t)i * M]));
            }
    }

    /* ---- right contact current, rgf_mod.f90:459-469 */
    s = nx - 1;
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(sigR[e] - conj(sigR[j + (size_t)i * M])) * f_first;
        }
    mulc(w.A, M, M, QTO_OP_N, GG[s], M, M, QTO_OP_N, w.B, NULL, NULL);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) {
            size_t e = i + (size_t)j * M;
            w.A[e] = -(sigR[e] - conj(sigR[j + (size_t)i * M])) * (f_first - 1.0);
        }
    mulc(w.A, M, M, QTO_OP_N, GL[s], M, M, QTO_OP_N, w.C, NULL, NULL);
    tim = 0.0;
    for (int j = 0; j < M; ++j) tim += creal(w.B[j + (size_t)j * M] - w.C[j + (size_t)j * M]);
    r->tr = tim;

    for (int t = 0; t < 3 * nx; ++t) free(own[t]);
    free(own); free(GL); free(GG); free(GR);
    free(Sbuf); free(sigR);
    work_free(&w);
    return 0;
}

/* ------------------------------------------------------------------ */
/* batch driver (CPU baseline), negf_mod.f90:115-127                   */
/* ------------------------------------------------------------------ */

int qto_rgf_batch(int variant, int nx, const int* nm, int nk, const qto_c* const* Hii,
                  const qto_c* const* H1i, const qto_c* const* Sii, int npts,
                  const double* En, const int* kidx, double mul, double mur, double templ,
                  double tempr, int coupling_fix, int nthreads, double* tr, double* tre,
                  qto_c* gr_diag, qto_c* gl_diag, qto_c* gg_diag, int* sancho_iters) {
    (void)nk;
    size_t ndiag = 0;
    for (int s = 0; s < nx; ++s) ndiag += (size_t)nm[s];
    int status = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int ip = 0; ip < npts; ++ip) {
        int ik = kidx ? kidx[ip] : 0;
        qto_c** Gl = (qto_c**)malloc(sizeof(qto_c*) * (size_t)nx);
        qto_c** Gln = (qto_c**)malloc(sizeof(qto_c*) * (size_t)nx);
        qto_c** GR = (qto_c**)malloc(sizeof(qto_c*) * (size_t)nx);
        qto_c** GL = (qto_c**)malloc(sizeof(qto_c*) * (size_t)nx);
        qto_c** GG = (qto_c**)malloc(sizeof(qto_c*) * (size_t)nx);
        for (int s = 0; s < nx; ++s) {
            size_t b = (size_t)nm[s] * nm[s];
            Gl[s] = zalloc(b); Gln[s] = zalloc(b);
            GR[s] = zalloc(b); GL[s] = zalloc(b); GG[s] = zalloc(b);
        }
        qto_point p;
        memset(&p, 0, sizeof p);
        p.nx = nx; p.nm = nm; p.En = En[ip];
        p.mul = mul; p.mur = mur; p.templ = templ; p.tempr = tempr;
        p.Hii = Hii + (size_t)ik * nx;
        p.H1i = H1i + (size_t)ik * (nx + 1);
        p.Sii = Sii ? Sii + (size_t)ik * nx : NULL;
        p.lead_eta = 1e-5; p.lead_nmax = 50; p.lead_tol = 1e-100;
        p.coupling_fix = coupling_fix;
        qto_result r;
        memset(&r, 0, sizeof r);
        r.G_r = GR; r.G_lesser = GL; r.G_greater = GG; r.Gl = Gl; r.Gln = Gln;
        int st = variant == 0 ? qto_rgf_forward(&p, &r) : qto_rgf_backward(&p, &r);
        if (st != 0) status = st;
        if (tr) tr[ip] = r.tr;
        if (tre) tre[ip] = r.tre;
        if (sancho_iters) { sancho_iters[2 * ip] = r.sancho_iters[0]; sancho_iters[2 * ip + 1] = r.sancho_iters[1]; }
        size_t off = (size_t)ip * ndiag;
        for (int s = 0; s < nx; ++s) {
            for (int j = 0; j < nm[s]; ++j) {
                size_t e = j + (size_t)j * nm[s];
                if (gr_diag) gr_diag[off + j] = GR[s][e];
                if (gl_diag) gl_diag[off + j] = GL[s][e];
                if (gg_diag) gg_diag[off + j] = GG[s][e];
            }
            off += (size_t)nm[s];
            free(Gl[s]); free(Gln[s]); free(GR[s]); free(GL[s]); free(GG[s]);
        }
        free(Gl); free(Gln); free(GR); free(GL); free(GG);
    }
    return status;
}

/* ------------------------------------------------------------------ */
/* energy-axis correlation / convolution, fft_mod.f90:58-168           */
/* ------------------------------------------------------------------ */

/* Z(i) = sum_e X(e) Y(e - l), l = i - n/2 (0-based; the lag-0 element sits
 * at index n/2).  fft_mod.f90:77-89: X zero-padded, Y reversed, length
 * 2n-1 circular product, middle n samples kept. */
void qto_corr1d2(int n, const qto_c* X, const qto_c* Y, qto_c* Z) {
    for (int i = 0; i < n; ++i) {
        int l = i - n / 2;
        qto_c s = 0.0;
        for (int e = 0; e < n; ++e) {
            int q = e - l;
            if (q >= 0 && q < n) s += X[e] * Y[q];
        }
        Z[i] = s;
    }
}

/* Z(e) = sum_{l=-n/2}^{n-1-n/2} X(e - l) Y(l + n/2), X = 0 outside [0, n).
 * fft_mod.f90:153-166: Y rotated so lag 0 is sample 0, X zero-padded.  For
 * even n the upper lag is n/2-1 as in the reference; for odd n (where the
 * reference assignment is non-conformant) every sample of Y is used. */
void qto_conv1d2(int n, const qto_c* X, const qto_c* Y, qto_c* Z) {
    for (int e = 0; e < n; ++e) {
        qto_c s = 0.0;
        for (int l = -n / 2; l <= n - 1 - n / 2; ++l) {
            int q = e - l;
            if (q >= 0 && q < n) s += X[q] * Y[l + n / 2];
        }
        Z[e] = s;
    }
}

/* ------------------------------------------------------------------ */
/* green_RGF_RS, gw_mod.f90:1131-1345                                  */
/* ------------------------------------------------------------------ */

/* C = op(A) op(B), all nm x nm */
static void g2(int opa, int opb, int nm, const qto_c* A, const qto_c* B, qto_c* C) {
    qto_zgemm(opa, opb, nm, nm, nm, 1.0, A, nm, B, nm, 0.0, C, nm);
}

int qto_green_rgf_rs(const double* TEMP, int nm, int nx, double E, const double* mu,
                     const qto_c* Hii, const qto_c* H1i, const qto_c* sigma_lesser_ph,
                     const qto_c* sigma_r_ph, qto_c* ndens, qto_c* pdens, qto_c* ldos,
                     double* tr, double* tre, double* cur, qto_c* GRi1, qto_c* GLi1,
                     qto_c* GGi1) {
    if (nx < 2) return -1;
    const size_t nn = (size_t)nm * nm;
    const double mul = mu[0], mur = mu[1], templ = TEMP[0], tempr = TEMP[1];
    const qto_c z = E + 0.0 * I; /* gw_mod.f90:1155 */
    qto_c *H00 = zalloc(nn), *H10 = zalloc(nn), *A = zalloc(nn), *B = zalloc(nn), *Cm = zalloc(nn);
    qto_c *D = zalloc(nn), *S00 = zalloc(nn), *G00 = zalloc(nn), *GBB = zalloc(nn), *GN0 = zalloc(nn);
    qto_c *Gn = zalloc(nn), *Gp = zalloc(nn), *sig = zalloc(nn), *sigmal = zalloc(nn);
    qto_c *sigmar = zalloc(nn), *Gl = zalloc(nn * nx), *Gln = zalloc(nn * nx);
    qto_c* Glp = zalloc(nn * nx); /* never filled by the reference (stays zero) */
    for (int i = 0; i < nm; ++i) S00[i + (size_t)i * nm] = 1.0;
#define BLK(p, s) ((p) + (size_t)(s) * nn)
#define SR(s) (sigma_r_ph ? BLK(sigma_r_ph, s)[e] : 0.0)
#define SL(s) (sigma_lesser_ph ? BLK(sigma_lesser_ph, s)[e] : 0.0)
    for (size_t e = 0; e < nn * nx; ++e) { ldos[e] = ndens[e] = pdens[e] = 0.0; cur[e] = 0.0; }
    /* left contact :1173-1193 */
    for (size_t e = 0; e < nn; ++e) { H00[e] = BLK(Hii, 0)[e] + SR(0); H10[e] = BLK(H1i, 0)[e]; }
    dagger(H10, nm, nm, D);
    qto_sancho(nm, E, 1e-3, 100, 1e-100, 1, S00, H00, D, G00, GBB, NULL);
    g2(QTO_OP_N, QTO_OP_N, nm, H10, G00, A);
    g2(QTO_OP_N, QTO_OP_C, nm, A, H10, sigmal);
    const double fL_inj = qto_ferm((E - mul) / (BOLTZ * templ));
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            sig[e] = -(sigmal[e] - conj(sigmal[j + (size_t)i * nm])) * fL_inj + SL(0);
        }
    for (size_t e = 0; e < nn; ++e) A[e] = z * S00[e] - H00[e] - sigmal[e];
    qto_invert(A, nm);
    memcpy(BLK(Gl, 0), A, nn * sizeof(qto_c));
    g2(QTO_OP_N, QTO_OP_N, nm, A, sig, B);
    g2(QTO_OP_N, QTO_OP_C, nm, B, A, Cm);
    memcpy(BLK(Gln, 0), Cm, nn * sizeof(qto_c));
    /* first pass :1194-1211 */
    for (int l = 1; l <= nx - 2; ++l) {
        for (size_t e = 0; e < nn; ++e) { H00[e] = BLK(Hii, l)[e] + SR(l); H10[e] = BLK(H1i, l)[e]; }
        g2(QTO_OP_N, QTO_OP_N, nm, H10, BLK(Gl, l - 1), B);
        g2(QTO_OP_N, QTO_OP_C, nm, B, H10, Cm);
        for (size_t e = 0; e < nn; ++e) A[e] = z * S00[e] - H00[e] - Cm[e];
        qto_invert(A, nm);
        memcpy(BLK(Gl, l), A, nn * sizeof(qto_c));
        memcpy(sig, BLK(Gln, l - 1), nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_N, nm, H10, sig, B);
        g2(QTO_OP_N, QTO_OP_C, nm, B, H10, Cm);
        for (size_t e = 0; e < nn; ++e) Cm[e] += SL(l);
        g2(QTO_OP_N, QTO_OP_N, nm, A, Cm, B);
        g2(QTO_OP_N, QTO_OP_C, nm, B, A, Gn);
        memcpy(BLK(Gln, l), Gn, nn * sizeof(qto_c));
    }
    /* right contact :1212-1247 */
    const int L = nx - 1;
    for (size_t e = 0; e < nn; ++e) { H00[e] = BLK(Hii, L)[e] + SR(L); H10[e] = BLK(H1i, L)[e]; }
    qto_sancho(nm, E, 1e-3, 100, 1e-100, 1, S00, H00, H10, G00, GBB, NULL);
    g2(QTO_OP_C, QTO_OP_N, nm, H10, G00, A);
    g2(QTO_OP_N, QTO_OP_N, nm, A, H10, sigmar);
    g2(QTO_OP_N, QTO_OP_N, nm, H10, BLK(Gl, L - 1), B);
    g2(QTO_OP_N, QTO_OP_C, nm, B, H10, Cm);
    for (size_t e = 0; e < nn; ++e) G00[e] = z * S00[e] - H00[e] - sigmar[e] - Cm[e];
    qto_invert(G00, nm);
    memcpy(BLK(ldos, L), G00, nn * sizeof(qto_c));
    memcpy(sig, BLK(Gln, L - 1), nn * sizeof(qto_c));
    g2(QTO_OP_N, QTO_OP_N, nm, H10, sig, B);
    g2(QTO_OP_N, QTO_OP_C, nm, B, H10, Cm);
    const double fR = qto_ferm((E - mur) / (BOLTZ * templ)); /* TEMPl, :1232 */
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            sig[e] = -(sigmar[e] - conj(sigmar[j + (size_t)i * nm])) * fR + Cm[e] + SL(L);
        }
    g2(QTO_OP_N, QTO_OP_N, nm, G00, sig, B);
    g2(QTO_OP_N, QTO_OP_C, nm, B, G00, Gn);
    memcpy(BLK(ndens, L), Gn, nn * sizeof(qto_c));
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            Gp[e] = Gn[e] + (G00[e] - conj(G00[j + (size_t)i * nm]));
        }
    memcpy(BLK(pdens, L), Gp, nn * sizeof(qto_c));
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            A[e] = -(sigmar[e] - conj(sigmar[j + (size_t)i * nm])) * fR;
        }
    g2(QTO_OP_N, QTO_OP_N, nm, A, Gp, B);
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            A[e] = -(sigmar[e] - conj(sigmar[j + (size_t)i * nm])) * (fR - 1.0);
        }
    g2(QTO_OP_N, QTO_OP_N, nm, A, Gn, Cm);
    double tim = 0.0;
    for (int i = 0; i < nm; ++i) tim -= creal(B[i + (size_t)i * nm] - Cm[i + (size_t)i * nm]);
    *tr = tim;
    /* second pass :1250-1330 */
    for (int l = nx - 2; l >= 0; --l) {
        memcpy(H10, BLK(H1i, l), nn * sizeof(qto_c));
        memcpy(A, Gn, nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_C, nm, H10, BLK(Gl, l), B);
        g2(QTO_OP_N, QTO_OP_N, nm, A, B, Cm);
        memcpy(A, BLK(Gln, l), nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_N, nm, H10, A, B);
        g2(QTO_OP_N, QTO_OP_N, nm, G00, B, A);
        for (size_t e = 0; e < nn; ++e) B[e] = Cm[e] + A[e];
        g2(QTO_OP_C, QTO_OP_N, nm, H10, B, A);
        for (size_t e = 0; e < nn; ++e) BLK(cur, l)[e] = creal(A[e]);
        memcpy(A, Gn, nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_C, nm, BLK(Gl, l), H10, B);
        g2(QTO_OP_N, QTO_OP_N, nm, B, A, Cm);
        memcpy(A, BLK(Gln, l), nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_C, nm, A, H10, B);
        g2(QTO_OP_N, QTO_OP_C, nm, B, G00, A);
        for (size_t e = 0; e < nn; ++e) B[e] = Cm[e] + A[e];
        if (GLi1) memcpy(BLK(GLi1, l), B, nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_N, nm, H10, B, A);
        for (size_t e = 0; e < nn; ++e) BLK(cur, l)[e] -= creal(A[e]);
        if (GGi1) {
            memcpy(A, Gp, nn * sizeof(qto_c));
            g2(QTO_OP_C, QTO_OP_C, nm, BLK(Gl, l), H10, B);
            g2(QTO_OP_N, QTO_OP_N, nm, B, A, Cm);
            memcpy(A, BLK(Glp, l), nn * sizeof(qto_c));
            g2(QTO_OP_N, QTO_OP_C, nm, A, H10, B);
            g2(QTO_OP_N, QTO_OP_N, nm, B, G00, A);
            for (size_t e = 0; e < nn; ++e) BLK(GGi1, l)[e] = Cm[e] + A[e];
        }
        memcpy(D, BLK(Gl, l), nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_C, nm, D, H10, B);
        g2(QTO_OP_N, QTO_OP_N, nm, B, G00, GN0);
        if (GRi1) memcpy(BLK(GRi1, l), GN0, nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_N, nm, GN0, H10, A);
        g2(QTO_OP_N, QTO_OP_N, nm, A, D, Cm);
        for (size_t e = 0; e < nn; ++e) G00[e] = BLK(Gl, l)[e] + Cm[e];
        memcpy(BLK(ldos, l), G00, nn * sizeof(qto_c));
        memcpy(A, Gn, nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_C, nm, D, H10, B);
        g2(QTO_OP_N, QTO_OP_N, nm, B, A, Cm);
        g2(QTO_OP_N, QTO_OP_N, nm, Cm, H10, A);
        g2(QTO_OP_N, QTO_OP_C, nm, A, D, Cm);
        for (size_t e = 0; e < nn; ++e) Gn[e] = BLK(Gln, l)[e] + Cm[e];
        memcpy(A, BLK(Gln, l), nn * sizeof(qto_c));
        g2(QTO_OP_N, QTO_OP_N, nm, GN0, H10, B);
        g2(QTO_OP_N, QTO_OP_N, nm, B, A, Cm);
        for (size_t e = 0; e < nn; ++e) Gn[e] += Cm[e];
        g2(QTO_OP_N, QTO_OP_C, nm, A, H10, B);
        g2(QTO_OP_N, QTO_OP_C, nm, B, GN0, Cm);
        for (size_t e = 0; e < nn; ++e) Gn[e] += Cm[e];
        memcpy(BLK(ndens, l), Gn, nn * sizeof(qto_c));
        for (int j = 0; j < nm; ++j)
            for (int i = 0; i < nm; ++i) {
                size_t e = i + (size_t)j * nm;
                BLK(pdens, l)[e] = Gn[e] + (G00[e] - conj(G00[j + (size_t)i * nm]));
            }
    }
    /* left contact current :1332-1341 (TEMPr with mul) */
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            Gp[e] = Gn[e] + (G00[e] - conj(G00[j + (size_t)i * nm]));
        }
    const double fLc = qto_ferm((E - mul) / (BOLTZ * tempr));
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            A[e] = -(sigmal[e] - conj(sigmal[j + (size_t)i * nm])) * fLc;
        }
    g2(QTO_OP_N, QTO_OP_N, nm, A, Gp, B);
    for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i) {
            size_t e = i + (size_t)j * nm;
            A[e] = -(sigmal[e] - conj(sigmal[j + (size_t)i * nm])) * (fLc - 1.0);
        }
    g2(QTO_OP_N, QTO_OP_N, nm, A, Gn, Cm);
    tim = 0.0;
    for (int i = 0; i < nm; ++i) tim += creal(B[i + (size_t)i * nm] - Cm[i + (size_t)i * nm]);
    *tre = tim;
#undef BLK
#undef SR
#undef SL
    free(H00); free(H10); free(A); free(B); free(Cm); free(D); free(S00); free(G00); free(GBB);
    free(GN0); free(Gn); free(Gp); free(sig); free(sigmal); free(sigmar); free(Gl); free(Gln);
    free(Glp);
    return 0;
}

/* ------------------------------------------------------------------ */
/* GW polarisation / self-energy assembly, gw_mod.f90:444-749          */
/* ------------------------------------------------------------------ */

/* the reference's constant (static.f90:15) */
#define QTO_GW_PI 3.14159265359

/* k' paired with (k, q), gw_mod.f90:478-489 and :656-667 (1-based there;
 * 0-based here): k'_z = k_z - q_z + nkz/2 folded into the zone. */
static int kd_index(int ik, int iq, int nky, int nkz) {
    int ikz = ik % nkz + 1, iky = ik / nkz + 1;
    int iqz = iq % nkz + 1, iqy = iq / nkz + 1;
    int ikzd = ikz - iqz + nkz / 2;
    int ikyd = iky - iqy + nky / 2;
    if (ikzd < 1) ikzd += nkz;
    if (ikzd > nkz) ikzd -= nkz;
    if (ikyd < 1) ikyd += nky;
    if (ikyd > nky) ikyd -= nky;
    if (nky == 1) ikyd = 1;
    if (nkz == 1) ikzd = 1;
    return (ikzd - 1) + (ikyd - 1) * nkz;
}

static int in_band(int row, int col, int nm, int ndiag) {
    int l = row - ndiag < 0 ? 0 : row - ndiag;
    int h = row + ndiag > nm - 1 ? nm - 1 : row + ndiag;
    return col >= l && col <= h;
}

/* series (ij, ix, ik) of an element-major array */
#define SER(A, ij, ix, ik) ((A) + ((((size_t)(ij) * nx + (ix)) * nk + (ik)) * (size_t)nen))

/* i * Im(v), the preprocessing of gw_mod.f90:444-446 */
static void imag_part(int n, const qto_c* x, qto_c* y, int conjugate) {
    for (int e = 0; e < n; ++e) y[e] = (conjugate ? -1.0 : 1.0) * cimag(x[e]) * I;
}

void qto_gw_polarization(int nen, int nky, int nkz, int nx, int nm, int ndiag, double dEn,
                         double spindeg, const qto_c* Gr, const qto_c* Gl, const qto_c* Gg,
                         qto_c* Pr, qto_c* Pl, qto_c* Pg) {
    const int nk = nky * nkz;
    const size_t total = (size_t)nm * nm * nx * nk * nen;
    memset(Pr, 0, sizeof(qto_c) * total);
    memset(Pl, 0, sizeof(qto_c) * total);
    memset(Pg, 0, sizeof(qto_c) * total);
    qto_c *a = zalloc(nen), *b = zalloc(nen), *z = zalloc(nen);
    /* dE, gw_mod.f90:516 */
    const qto_c dE = (-dEn / 2.0 / QTO_GW_PI) * I * spindeg / (double)nk;
    for (int col = 0; col < nm; ++col)
        for (int row = 0; row < nm; ++row) {
            if (!in_band(row, col, nm, ndiag)) continue;
            const int ij = row + col * nm, ji = col + row * nm;
            for (int ix = 0; ix < nx; ++ix)
                for (int iq = 0; iq < nk; ++iq)
                    for (int ik = 0; ik < nk; ++ik) {
                        const int ikd = kd_index(ik, iq, nky, nkz);
                        qto_c* pl = SER(Pl, ij, ix, iq);
                        qto_c* pg = SER(Pg, ij, ix, iq);
                        qto_c* pr = SER(Pr, ij, ix, iq);
                        /* P< += corr(G<_ij(k), G>_ji(k')) */
                        imag_part(nen, SER(Gl, ij, ix, ik), a, 0);
                        imag_part(nen, SER(Gg, ji, ix, ikd), b, 0);
                        qto_corr1d2(nen, a, b, z);
                        for (int e = 0; e < nen; ++e) pl[e] += z[e];
                        /* P> += corr(G>_ij(k), G<_ji(k')) */
                        imag_part(nen, SER(Gg, ij, ix, ik), a, 0);
                        imag_part(nen, SER(Gl, ji, ix, ikd), b, 0);
                        qto_corr1d2(nen, a, b, z);
                        for (int e = 0; e < nen; ++e) pg[e] += z[e];
                        /* P^R += corr(G<_ij(k), conj G^R_ij(k')) + corr(G^R_ij(k), G<_ji(k')) */
                        imag_part(nen, SER(Gl, ij, ix, ik), a, 0);
                        imag_part(nen, SER(Gr, ij, ix, ikd), b, 1);
                        qto_corr1d2(nen, a, b, z);
                        for (int e = 0; e < nen; ++e) pr[e] += z[e];
                        imag_part(nen, SER(Gr, ij, ix, ik), a, 0);
                        imag_part(nen, SER(Gl, ji, ix, ikd), b, 0);
                        qto_corr1d2(nen, a, b, z);
                        for (int e = 0; e < nen; ++e) pr[e] += z[e];
                    }
        }
    for (size_t t = 0; t < total; ++t) {
        Pl[t] *= dE;
        Pg[t] *= dE;
        Pr[t] *= dE;
    }
    free(a); free(b); free(z);
}

void qto_gw_selfenergy(int nen, int nky, int nkz, int nx, int nm, int ndiag, double dEn,
                       const qto_c* Gr, const qto_c* Gl, const qto_c* Gg, const qto_c* Wr,
                       const qto_c* Wl, const qto_c* Wg, qto_c* Sr, qto_c* Sl, qto_c* Sg) {
    const int nk = nky * nkz;
    const size_t total = (size_t)nm * nm * nx * nk * nen;
    qto_c *tr = zalloc(total), *tl = zalloc(total), *tg = zalloc(total);
    qto_c *a = zalloc(nen), *z = zalloc(nen);
    for (int col = 0; col < nm; ++col)
        for (int row = 0; row < nm; ++row) {
            if (!in_band(row, col, nm, ndiag)) continue;
            const int ij = row + col * nm;
            for (int ix = 0; ix < nx; ++ix)
                for (int ik = 0; ik < nk; ++ik)
                    for (int iq = 0; iq < nk; ++iq) {
                        const int ikd = kd_index(ik, iq, nky, nkz);
                        qto_c* sl = SER(tl, ij, ix, ik);
                        qto_c* sg = SER(tg, ij, ix, ik);
                        qto_c* sr = SER(tr, ij, ix, ik);
                        /* gw_mod.f90:672-681 */
                        imag_part(nen, SER(Gl, ij, ix, ikd), a, 0);
                        qto_conv1d2(nen, a, SER(Wl, ij, ix, iq), z);
                        for (int e = 0; e < nen; ++e) sl[e] += z[e];
                        qto_conv1d2(nen, a, SER(Wr, ij, ix, iq), z);
                        for (int e = 0; e < nen; ++e) sr[e] += z[e];
                        imag_part(nen, SER(Gg, ij, ix, ikd), a, 0);
                        qto_conv1d2(nen, a, SER(Wg, ij, ix, iq), z);
                        for (int e = 0; e < nen; ++e) sg[e] += z[e];
                        imag_part(nen, SER(Gr, ij, ix, ikd), a, 0);
                        qto_conv1d2(nen, a, SER(Wl, ij, ix, iq), z);
                        for (int e = 0; e < nen; ++e) sr[e] += z[e];
                        qto_conv1d2(nen, a, SER(Wr, ij, ix, iq), z);
                        for (int e = 0; e < nen; ++e) sr[e] += z[e];
                    }
        }
    /* :694-698 */
    const qto_c dE = (dEn / 2.0 / QTO_GW_PI) * I / (double)nk;
    for (size_t t = 0; t < total; ++t) {
        tl[t] *= dE;
        tg[t] *= dE;
        tr[t] *= dE;
        tr[t] = creal(tr[t]) + I * (cimag(tg[t] - tl[t]) / 2.0);
    }
    /* band zeroing (:705-726) leaves out-of-band entries zero; symmetrise
     * (:730-743): S_ij = (S_ij + S_ji) / 2 */
    for (int col = 0; col < nm; ++col)
        for (int row = 0; row < nm; ++row) {
            const int ij = row + col * nm, ji = col + row * nm;
            for (int ix = 0; ix < nx; ++ix)
                for (int ik = 0; ik < nk; ++ik) {
                    const qto_c *r1 = SER(tr, ij, ix, ik), *r2 = SER(tr, ji, ix, ik);
                    const qto_c *l1 = SER(tl, ij, ix, ik), *l2 = SER(tl, ji, ix, ik);
                    const qto_c *g1 = SER(tg, ij, ix, ik), *g2_ = SER(tg, ji, ix, ik);
                    qto_c *ro = SER(Sr, ij, ix, ik), *lo = SER(Sl, ij, ix, ik), *go = SER(Sg, ij, ix, ik);
                    for (int e = 0; e < nen; ++e) {
                        ro[e] = (r1[e] + r2[e]) / 2.0;
                        lo[e] = (l1[e] + l2[e]) / 2.0;
                        go[e] = (g1[e] + g2_[e]) / 2.0;
                    }
                }
        }
    free(tr); free(tl); free(tg); free(a); free(z);
}

void qto_gw_mix(int64_t n, double alpha, qto_c* S, const qto_c* Snew) {
    for (int64_t t = 0; t < n; ++t) S[t] = S[t] + alpha * (Snew[t] - S[t]);
}
#undef SER

/* e-photon self-energies, monochromatic field: calc_sigma_ephoton_monochromatic_ext
 * (gw_mod.f90:1075-1122).  Per energy ie and slab ix (loop order of
 * gw_mod.f90:1092-1116): A = G(ie2 - nop) + G(ie2 + nop) with ie2 = ie + nop
 * (0-based: G at ie and ie + 2 nop), B = M A, A = B M (two ZGEMMs); then the
 * scaling by pre_fact intensity / hw^2 (gw_mod.f90:1119-1120) and
 * Sigma^R = 0 Re + i Im(Sigma^> - Sigma^<) / 2 (gw_mod.f90:1121). */
void qto_gw_sigma_ephoton(int nm, int length, int nen, int nop, double pre_fact, double intensity,
                          double hw, const qto_c* Mii, const qto_c* Gl, const qto_c* Gg, qto_c* Sr,
                          qto_c* Sl, qto_c* Sg) {
    const size_t nn = (size_t)nm * nm;
    qto_c* A = (qto_c*)malloc(nn * sizeof(qto_c));
    qto_c* B = (qto_c*)malloc(nn * sizeof(qto_c));
    for (int ie = 0; ie < nen; ++ie)
        for (int ix = 0; ix < length; ++ix) {
            const qto_c* M = Mii + (size_t)ix * nn;
            const size_t o = ((size_t)ie * length + ix) * nn;
            const size_t lo = ((size_t)ie * length + ix) * nn;
            const size_t hi = ((size_t)(ie + 2 * nop) * length + ix) * nn;
            for (int x = 0; x < 2; ++x) {
                const qto_c* G = x ? Gg : Gl;
                qto_c* S = x ? Sg : Sl;
                for (size_t t = 0; t < nn; ++t) A[t] = G[lo + t] + G[hi + t];
                qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, M, nm, A, nm, 0.0, B, nm);
                qto_zgemm(QTO_OP_N, QTO_OP_N, nm, nm, nm, 1.0, B, nm, M, nm, 0.0, A, nm);
                memcpy(S + o, A, nn * sizeof(qto_c));
            }
        }
    const size_t ns = nn * (size_t)length * nen;
    for (size_t t = 0; t < ns; ++t) {
        Sg[t] = Sg[t] * pre_fact * intensity / (hw * hw);
        Sl[t] = Sl[t] * pre_fact * intensity / (hw * hw);
    }
    for (size_t t = 0; t < ns; ++t) Sr[t] = I * (cimag(Sg[t] - Sl[t]) / 2.0);
    free(A);
    free(B);
}

/* ------------------------------------------------------------------ */
/* W solve (called at gw_mod.f90:597, not defined in the reference)    */
/* ------------------------------------------------------------------ */

/* W^R = V (I - P^R V)^-1 and W^<,> = W^R P^<,> W^R^H for a block-tridiagonal
 * Hermitian V (Vii(:,:,i) = V_ii, V1i(:,:,i) = V_{i,i-1} for i >= 2, the
 * H1i convention of green_RGF_RS, gw_mod.f90:1194-1196) and block-diagonal
 * P (diagonal blocks only, as assembled by the P loop).  With
 * A = I - P^R V (block tridiagonal), G = A^-1 and Z = G P^< G^H:
 *   W^R_ii = sum_j V_ij G_ji,  W^<_ii = sum_{j,k} V_ij Z_jk V_ik^H,
 * j, k in {i-1, i, i+1}.  G and Z come from a left-connected forward sweep
 * and a backward sweep that also yields the distance-2 blocks of Z.  The
 * system is closed (no boundary self-energy for W). */
typedef struct {
    int nm, nx;
    const qto_c *Vii, *V1i;
    qto_c *g, *zl, *zg;     /* left-connected, per slab */
    qto_c *Gd, *Gu, *Gl;    /* G_ii, G_{i,i+1}, G_{i+1,i} */
    qto_c *Z[2][5];         /* Z^<,> diagonals -2..2: Z[.][2+d] slab i holds Z_{i,i+d} */
} wsolve_t;

#define WB(p, i) ((p) + (size_t)(i) * nn)

/* V_ij for |i - j| <= 1, NULL otherwise; tmp receives V1i^H when needed */
static const qto_c* vblock(const wsolve_t* w, int i, int j, qto_c* tmp) {
    const int nm = w->nm;
    const size_t nn = (size_t)nm * nm;
    if (i == j) return WB(w->Vii, i);
    if (i == j + 1) return WB(w->V1i, i);
    if (j == i + 1) {
        dagger(WB(w->V1i, j), nm, nm, tmp);
        return tmp;
    }
    return NULL;
}

/* A_ij = delta_ij I - P^R_i V_ij into out */
static void ablock(const wsolve_t* w, const qto_c* PR, int i, int j, qto_c* out, qto_c* tmp) {
    const int nm = w->nm;
    const size_t nn = (size_t)nm * nm;
    const qto_c* v = vblock(w, i, j, tmp);
    g2(QTO_OP_N, QTO_OP_N, nm, WB(PR, i), v, out);
    for (size_t t = 0; t < nn; ++t) out[t] = -out[t];
    if (i == j)
        for (int d = 0; d < nm; ++d) out[d + (size_t)d * nm] += 1.0;
}

static void zadd(size_t n, qto_c* a, const qto_c* b, qto_c s) {
    for (size_t t = 0; t < n; ++t) a[t] += s * b[t];
}

int qto_gw_wsolve(int nm, int nx, const qto_c* Vii, const qto_c* V1i, const qto_c* PR,
                  const qto_c* PL, const qto_c* PG, qto_c* WR, qto_c* WL, qto_c* WG) {
    const size_t nn = (size_t)nm * nm;
    int status = 0;
    wsolve_t w;
    memset(&w, 0, sizeof w);
    w.nm = nm;
    w.nx = nx;
    w.Vii = Vii;
    w.V1i = V1i;
    w.g = zalloc(nn * nx);
    w.zl = zalloc(nn * nx);
    w.zg = zalloc(nn * nx);
    w.Gd = zalloc(nn * nx);
    w.Gu = zalloc(nn * nx);
    w.Gl = zalloc(nn * nx);
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < 5; ++d) w.Z[s][d] = zalloc(nn * nx);
    qto_c *A1 = zalloc(nn), *A2 = zalloc(nn), *T1 = zalloc(nn), *T2 = zalloc(nn), *T3 = zalloc(nn);
    qto_c *vt = zalloc(nn), *vt2 = zalloc(nn);
    const qto_c* PX[2] = {PL, PG};
    /* forward: g_i = (A_ii - A_{i,i-1} g_{i-1} A_{i-1,i})^-1,
     *          z_i = g_i (P_i + A_{i,i-1} z_{i-1} A_{i,i-1}^H) g_i^H */
    for (int i = 0; i < nx; ++i) {
        ablock(&w, PR, i, i, T1, vt);
        if (i > 0) {
            ablock(&w, PR, i, i - 1, A1, vt);
            ablock(&w, PR, i - 1, i, A2, vt);
            g2(QTO_OP_N, QTO_OP_N, nm, A1, WB(w.g, i - 1), T2);
            g2(QTO_OP_N, QTO_OP_N, nm, T2, A2, T3);
            zadd(nn, T1, T3, -1.0);
        }
        if (qto_invert(T1, nm) != 0) status = 1;
        memcpy(WB(w.g, i), T1, sizeof(qto_c) * nn);
        for (int s = 0; s < 2; ++s) {
            qto_c* z = s ? w.zg : w.zl;
            memcpy(T2, WB(PX[s], i), sizeof(qto_c) * nn);
            if (i > 0) {
                g2(QTO_OP_N, QTO_OP_N, nm, A1, WB(z, i - 1), T3);
                g2(QTO_OP_N, QTO_OP_C, nm, T3, A1, vt2);
                zadd(nn, T2, vt2, 1.0);
            }
            g2(QTO_OP_N, QTO_OP_N, nm, WB(w.g, i), T2, T3);
            g2(QTO_OP_N, QTO_OP_C, nm, T3, WB(w.g, i), WB(z, i));
        }
    }
    /* backward */
    memcpy(WB(w.Gd, nx - 1), WB(w.g, nx - 1), sizeof(qto_c) * nn);
    memcpy(WB(w.Z[0][2], nx - 1), WB(w.zl, nx - 1), sizeof(qto_c) * nn);
    memcpy(WB(w.Z[1][2], nx - 1), WB(w.zg, nx - 1), sizeof(qto_c) * nn);
    for (int i = nx - 2; i >= 0; --i) {
        ablock(&w, PR, i, i + 1, A1, vt);   /* A_{i,i+1} */
        ablock(&w, PR, i + 1, i, A2, vt);   /* A_{i+1,i} */
        const qto_c* gi = WB(w.g, i);
        /* G_{i,i+1} = -g_i A_{i,i+1} G_{i+1,i+1} */
        g2(QTO_OP_N, QTO_OP_N, nm, gi, A1, T1);            /* T1 = g A01 */
        g2(QTO_OP_N, QTO_OP_N, nm, T1, WB(w.Gd, i + 1), WB(w.Gu, i));
        zadd(nn, WB(w.Gu, i), WB(w.Gu, i), -2.0);
        /* G_{i+1,i} = -G_{i+1,i+1} A_{i+1,i} g_i */
        g2(QTO_OP_N, QTO_OP_N, nm, A2, gi, T2);            /* T2 = A10 g */
        g2(QTO_OP_N, QTO_OP_N, nm, WB(w.Gd, i + 1), T2, WB(w.Gl, i));
        zadd(nn, WB(w.Gl, i), WB(w.Gl, i), -2.0);
        /* G_ii = g_i - G_{i,i+1} A_{i+1,i} g_i */
        g2(QTO_OP_N, QTO_OP_N, nm, WB(w.Gu, i), T2, T3);
        memcpy(WB(w.Gd, i), gi, sizeof(qto_c) * nn);
        zadd(nn, WB(w.Gd, i), T3, -1.0);
        for (int s = 0; s < 2; ++s) {
            const qto_c* zi = WB(s ? w.zg : w.zl, i);
            qto_c** Z = w.Z[s];
            /* Z_{i,i+1} = -g A01 Z_{i+1,i+1} - z_i A10^H G_{i+1,i+1}^H */
            g2(QTO_OP_N, QTO_OP_N, nm, T1, WB(Z[2], i + 1), WB(Z[3], i));
            g2(QTO_OP_N, QTO_OP_C, nm, zi, A2, T3);        /* T3 = z A10^H */
            g2(QTO_OP_N, QTO_OP_C, nm, T3, WB(w.Gd, i + 1), vt2);
            zadd(nn, WB(Z[3], i), vt2, 1.0);
            zadd(nn, WB(Z[3], i), WB(Z[3], i), -2.0);
            /* Z_{i+1,i} = -Z_{i+1,i+1} A01^H g^H - G_{i+1,i+1} A10 z_i */
            /* A01^H g^H = (g A01)^H = T1^H */
            g2(QTO_OP_N, QTO_OP_C, nm, WB(Z[2], i + 1), T1, WB(Z[1], i + 1));
            g2(QTO_OP_N, QTO_OP_N, nm, A2, zi, vt);        /* A10 z */
            g2(QTO_OP_N, QTO_OP_N, nm, WB(w.Gd, i + 1), vt, vt2);
            zadd(nn, WB(Z[1], i + 1), vt2, 1.0);
            zadd(nn, WB(Z[1], i + 1), WB(Z[1], i + 1), -2.0);
            if (i + 2 < nx) {
                /* Z_{i,i+2} = -g A01 Z_{i+1,i+2} - z_i A10^H G_{i+2,i+1}^H */
                g2(QTO_OP_N, QTO_OP_N, nm, T1, WB(Z[3], i + 1), WB(Z[4], i));
                g2(QTO_OP_N, QTO_OP_C, nm, T3, WB(w.Gl, i + 1), vt2);
                zadd(nn, WB(Z[4], i), vt2, 1.0);
                zadd(nn, WB(Z[4], i), WB(Z[4], i), -2.0);
                /* Z_{i+2,i} = -Z_{i+2,i+1} A01^H g^H - G_{i+2,i+1} A10 z_i */
                g2(QTO_OP_N, QTO_OP_C, nm, WB(Z[1], i + 2), T1, WB(Z[0], i + 2));
                g2(QTO_OP_N, QTO_OP_N, nm, WB(w.Gl, i + 1), vt, vt2);
                zadd(nn, WB(Z[0], i + 2), vt2, 1.0);
                zadd(nn, WB(Z[0], i + 2), WB(Z[0], i + 2), -2.0);
            }
            /* Z_ii = -g A01 Z_{i+1,i} + z_i M^H,  M = I + g A01 G_{i+1,i+1} A10 */
            g2(QTO_OP_N, QTO_OP_N, nm, T1, WB(Z[1], i + 1), WB(Z[2], i));
            zadd(nn, WB(Z[2], i), WB(Z[2], i), -2.0);
            g2(QTO_OP_N, QTO_OP_N, nm, WB(w.Gu, i), A2, vt);   /* G_{i,i+1} A10 = -g A01 G A10 */
            /* M = I - G_{i,i+1} A10 */
            for (size_t t = 0; t < nn; ++t) vt[t] = -vt[t];
            for (int d = 0; d < nm; ++d) vt[d + (size_t)d * nm] += 1.0;
            g2(QTO_OP_N, QTO_OP_C, nm, zi, vt, vt2);
            zadd(nn, WB(Z[2], i), vt2, 1.0);
        }
    }
    /* W_ii */
    for (int i = 0; i < nx; ++i) {
        qto_c* wr = WB(WR, i);
        memset(wr, 0, sizeof(qto_c) * nn);
        for (int j = i - 1; j <= i + 1; ++j) {
            if (j < 0 || j >= nx) continue;
            const qto_c* v = vblock(&w, i, j, vt);
            const qto_c* Gji = j == i ? WB(w.Gd, i) : (j == i - 1 ? WB(w.Gu, i - 1) : WB(w.Gl, i));
            g2(QTO_OP_N, QTO_OP_N, nm, v, Gji, T1);
            zadd(nn, wr, T1, 1.0);
        }
        for (int s = 0; s < 2; ++s) {
            qto_c* wx = WB(s ? WG : WL, i);
            memset(wx, 0, sizeof(qto_c) * nn);
            for (int k = i - 1; k <= i + 1; ++k) {
                if (k < 0 || k >= nx) continue;
                memset(T2, 0, sizeof(qto_c) * nn);   /* Y_k = sum_j V_ij Z_jk */
                for (int j = i - 1; j <= i + 1; ++j) {
                    if (j < 0 || j >= nx) continue;
                    const qto_c* v = vblock(&w, i, j, vt);
                    const qto_c* zjk = WB(w.Z[s][2 + (k - j)], j);
                    g2(QTO_OP_N, QTO_OP_N, nm, v, zjk, T1);
                    zadd(nn, T2, T1, 1.0);
                }
                const qto_c* vk = vblock(&w, i, k, vt);
                g2(QTO_OP_N, QTO_OP_C, nm, T2, vk, T1);
                zadd(nn, wx, T1, 1.0);
            }
        }
    }
    free(w.g); free(w.zl); free(w.zg); free(w.Gd); free(w.Gu); free(w.Gl);
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < 5; ++d) free(w.Z[s][d]);
    free(A1); free(A2); free(T1); free(T2); free(T3); free(vt); free(vt2);
    return status;
}
#undef WB
