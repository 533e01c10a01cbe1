/*
 * oracle.c -- FP64 CPU restatement of DArray's RBF low-rank-approximation path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
 * of bench.py.  Nothing in darray_optimizer_amd/ or include/ links this file.
 * Declarations and reference citations are in oracle.h.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

static int g_threads = 1;

void orc_set_threads(int nthreads) {
    g_threads = nthreads < 1 ? 1 : nthreads;
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
}

int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------ RBF */

void orc_sqnorms(int n, int d, const double* X, int ldx, double* nrm) {
    /* vecnorm (dkernel.cu:102-114): one running sum per point, features in order. */
    for (int i = 0; i < n; ++i) {
        const double* x = X + (size_t)i * ldx;
        double s = 0.0;
        for (int f = 0; f < d; ++f) s += x[f] * x[f];
        nrm[i] = s;
    }
}

static double dot_d(int d, const double* a, const double* b) {
    double s = 0.0;
#pragma omp simd reduction(+ : s)
    for (int f = 0; f < d; ++f) s += a[f] * b[f];
    return s;
}

/* One K block, rows [i0,i0+mb) x cols [j0,j0+nb), into buf (ld = mb).
 * rbf (dkernel.cu:124-125): K = yi*yj*exp(-gamma*(ni + nj - 2*xij)). */
static void rbf_block(int i0, int mb, int j0, int nb, int d, const double* Xi, int ldxi,
                      const double* Xj, int ldxj, const double* ni, const double* nj,
                      const double* yi, const double* yj, double gamma, int clamp, double* buf) {
    for (int jj = 0; jj < nb; ++jj) {
        const double* xj = Xj + (size_t)(j0 + jj) * ldxj;
        double nrmj = nj[j0 + jj];
        double lj = yj ? yj[j0 + jj] : 1.0;
        for (int ii = 0; ii < mb; ++ii) {
            const double* xi = Xi + (size_t)(i0 + ii) * ldxi;
            double d2 = ni[i0 + ii] + nrmj - 2.0 * dot_d(d, xi, xj);
            if (clamp && d2 < 0.0) d2 = 0.0;
            double li = yi ? yi[i0 + ii] : 1.0;
            buf[ii + (size_t)jj * mb] = li * lj * exp(-gamma * d2);
        }
    }
}

void orc_rbf_tile(int m, int n, int d, const double* Xi, int ldxi, const double* Xj, int ldxj,
                  const double* yi, const double* yj, double gamma, int clamp,
                  double* K, int ldk) {
    double* ni = (double*)malloc(sizeof(double) * (size_t)MAX(m, 1));
    double* nj = (double*)malloc(sizeof(double) * (size_t)MAX(n, 1));
    orc_sqnorms(m, d, Xi, ldxi, ni);
    orc_sqnorms(n, d, Xj, ldxj, nj);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < n; ++j) {
        const double* xj = Xj + (size_t)j * ldxj;
        double lj = yj ? yj[j] : 1.0;
        for (int i = 0; i < m; ++i) {
            double d2 = ni[i] + nj[j] - 2.0 * dot_d(d, Xi + (size_t)i * ldxi, xj);
            if (clamp && d2 < 0.0) d2 = 0.0;
            double li = yi ? yi[i] : 1.0;
            K[i + (size_t)j * ldk] = li * lj * exp(-gamma * d2);
        }
    }
    free(ni);
    free(nj);
}

/* Y[iblk,:] += Kblk * O[jblk,:]; Kblk is mb x nb (ld mb). */
static void block_project(int mb, int nb, int l, const double* Kb, const double* O, int ldo,
                          double* Y, int ldy) {
    for (int c = 0; c < l; ++c) {
        double* y = Y + (size_t)c * ldy;
        const double* o = O + (size_t)c * ldo;
        for (int jj = 0; jj < nb; ++jj) {
            double w = o[jj];
            const double* kc = Kb + (size_t)jj * mb;
#pragma omp simd
            for (int ii = 0; ii < mb; ++ii) y[ii] += kc[ii] * w;
        }
    }
}

void orc_rbf_project(int m, int n, int d, int l, const double* Xi, int ldxi,
                     const double* Xj, int ldxj, const double* yi, const double* yj,
                     double gamma, const double* O, int ldo, double beta, double* Y, int ldy,
                     int chunk) {
    if (chunk <= 0) chunk = 50000; /* NN, dkernel.h:4 */
    double* ni = (double*)malloc(sizeof(double) * (size_t)MAX(m, 1));
    double* nj = (double*)malloc(sizeof(double) * (size_t)MAX(n, 1));
    orc_sqnorms(m, d, Xi, ldxi, ni);
    orc_sqnorms(n, d, Xj, ldxj, nj);
    /* beta on Y first (BLAS semantics: beta == 0 does not read Y). */
    for (int c = 0; c < l; ++c)
        for (int i = 0; i < m; ++i)
            Y[i + (size_t)c * ldy] = (beta == 0.0) ? 0.0 : beta * Y[i + (size_t)c * ldy];
    /* chunk loops of dkernel.h:20-26, then 64 x 256 register/cache blocks inside each chunk. */
    const int MB = 64, NB = 256;
    for (int ic = 0; ic < m; ic += chunk) {
        int rn = MIN(chunk, m - ic);
        for (int jc = 0; jc < n; jc += chunk) {
            int cn = MIN(chunk, n - jc);
            int nblk = (rn + MB - 1) / MB;
#pragma omp parallel
            {
                double* kb = (double*)malloc(sizeof(double) * MB * NB);
#pragma omp for schedule(dynamic, 1)
                for (int b = 0; b < nblk; ++b) {
                    int i0 = ic + b * MB, mb = MIN(MB, ic + rn - i0);
                    for (int j0 = jc; j0 < jc + cn; j0 += NB) {
                        int nb = MIN(NB, jc + cn - j0);
                        rbf_block(i0, mb, j0, nb, d, Xi, ldxi, Xj, ldxj, ni, nj, yi, yj, gamma, 0, kb);
                        block_project(mb, nb, l, kb, O + j0, ldo, Y + i0, ldy);
                    }
                }
                free(kb);
            }
        }
    }
    free(ni);
    free(nj);
}

/* ------------------------------------------------------------ dense */

void orc_gemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc) {
    int at = (ta == 'T' || ta == 't' || ta == 'C' || ta == 'c');
    int bt = (tb == 'T' || tb == 't' || tb == 'C' || tb == 'c');
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j) {
        double* c = C + (size_t)j * ldc;
        for (int i = 0; i < m; ++i) c[i] = (beta == 0.0) ? 0.0 : beta * c[i];
        if (alpha == 0.0 || k == 0) continue;
        if (!at) {
            for (int p = 0; p < k; ++p) {
                double w = alpha * (bt ? B[j + (size_t)p * ldb] : B[p + (size_t)j * ldb]);
                const double* a = A + (size_t)p * lda;
#pragma omp simd
                for (int i = 0; i < m; ++i) c[i] += a[i] * w;
            }
        } else {
            for (int i = 0; i < m; ++i) {
                const double* a = A + (size_t)i * lda;
                double s = 0.0;
                if (!bt) {
                    const double* b = B + (size_t)j * ldb;
#pragma omp simd reduction(+ : s)
                    for (int p = 0; p < k; ++p) s += a[p] * b[p];
                } else {
                    for (int p = 0; p < k; ++p) s += a[p] * B[j + (size_t)p * ldb];
                }
                c[i] += alpha * s;
            }
        }
    }
}

void orc_matrix_multiply_tn(int m, int n, int k, double alpha, const double* A, int lda,
                            const double* B, int ldb, double beta, double* C, int ldc, int bs) {
    /* blas_like.h:73-89: k-blocks of bs, beta -> 1 after the first block. */
    if (bs <= 0) bs = 128;
    if (k == 0) {
        orc_gemm('T', 'N', m, n, 0, alpha, A, lda, B, ldb, beta, C, ldc);
        return;
    }
    for (int kk = 0; kk < k; kk += bs) {
        int ke = MIN(k, kk + bs);
        orc_gemm('T', 'N', m, n, ke - kk, alpha, A + kk, lda, B + kk, ldb, beta, C, ldc);
        beta = 1.0;
    }
}

void orc_syrk_ut(int n, int k, double alpha, const double* A, int lda, double beta,
                 double* C, int ldc) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int j = 0; j < n; ++j) {
        const double* b = A + (size_t)j * lda;
        for (int i = 0; i <= j; ++i) {
            const double* a = A + (size_t)i * lda;
            double s = 0.0;
#pragma omp simd reduction(+ : s)
            for (int p = 0; p < k; ++p) s += a[p] * b[p];
            double c = (beta == 0.0) ? 0.0 : beta * C[i + (size_t)j * ldc];
            C[i + (size_t)j * ldc] = c + alpha * s;
        }
    }
}

int orc_potf2_upper(int n, double* A, int lda) {
    /* LAPACK potf2 "U": U(j,j) = sqrt(A(j,j) - U(0:j,j).U(0:j,j));
     * U(j,k) = (A(j,k) - U(0:j,j).U(0:j,k)) / U(j,j). */
    for (int j = 0; j < n; ++j) {
        double* cj = A + (size_t)j * lda;
        double ajj = cj[j] - dot_d(j, cj, cj);
        if (!(ajj > 0.0)) { /* <= 0 or NaN */
            cj[j] = ajj;
            return j + 1;
        }
        ajj = sqrt(ajj);
        cj[j] = ajj;
        for (int k = j + 1; k < n; ++k) {
            double* ck = A + (size_t)k * lda;
            ck[j] = (ck[j] - dot_d(j, cj, ck)) / ajj;
        }
    }
    return 0;
}

/* Base solve U^T X = alpha B (strsm_("L","U","T",diag), blas_like.h:110). */
static void trsm_lut_base(int unit, int m, int n, const double* A, int lda, double* B, int ldb) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
        double* b = B + (size_t)c * ldb;
        for (int i = 0; i < m; ++i) {
            const double* ai = A + (size_t)i * lda; /* column i: U(0:i, i) */
            double s = b[i] - dot_d(i, ai, b);
            b[i] = unit ? s : s / ai[i];
        }
    }
}

static void trsm_lut_rec(int unit, int m, int n, const double* A, int lda, double* B, int ldb,
                         int bs) {
    if (m <= bs) {
        trsm_lut_base(unit, m, n, A, lda, B, ldb);
        return;
    }
    int m1 = m / 2, m2 = m - m1;
    trsm_lut_rec(unit, m1, n, A, lda, B, ldb, bs);
    /* B2 -= A12^T B1 (matrix_multiply("T","N",-1,1), blas_like.h:121) */
    orc_matrix_multiply_tn(m2, n, m1, -1.0, A + (size_t)m1 * lda, lda, B, ldb, 1.0, B + m1, ldb,
                           128);
    trsm_lut_rec(unit, m2, n, A + m1 + (size_t)m1 * lda, lda, B + m1, ldb, bs);
}

static void scale_mat(int m, int n, double alpha, double* B, int ldb) {
    if (alpha == 1.0) return;
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < m; ++i) B[i + (size_t)c * ldb] *= alpha;
}

void orc_trsm_lut(char diag, int m, int n, double alpha, const double* A, int lda,
                  double* B, int ldb, int bs) {
    if (bs <= 0) bs = 64;
    scale_mat(m, n, alpha, B, ldb);
    trsm_lut_rec(diag == 'U' || diag == 'u', m, n, A, lda, B, ldb, bs);
}

static void trsm_lun_base(int unit, int m, int n, const double* A, int lda, double* B, int ldb) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
        double* b = B + (size_t)c * ldb;
        for (int i = m - 1; i >= 0; --i) {
            double s = b[i];
            for (int k = i + 1; k < m; ++k) s -= A[i + (size_t)k * lda] * b[k];
            b[i] = unit ? s : s / A[i + (size_t)i * lda];
        }
    }
}

static void trsm_lun_rec(int unit, int m, int n, const double* A, int lda, double* B, int ldb,
                         int bs) {
    if (m <= bs) {
        trsm_lun_base(unit, m, n, A, lda, B, ldb);
        return;
    }
    int m1 = m / 2, m2 = m - m1;
    trsm_lun_rec(unit, m2, n, A + m1 + (size_t)m1 * lda, lda, B + m1, ldb, bs);
    /* B1 -= U12 X2 */
    orc_gemm('N', 'N', m1, n, m2, -1.0, A + (size_t)m1 * lda, lda, B + m1, ldb, 1.0, B, ldb);
    trsm_lun_rec(unit, m1, n, A, lda, B, ldb, bs);
}

void orc_trsm_lun(char diag, int m, int n, double alpha, const double* A, int lda,
                  double* B, int ldb, int bs) {
    if (bs <= 0) bs = 64;
    scale_mat(m, n, alpha, B, ldb);
    trsm_lun_rec(diag == 'U' || diag == 'u', m, n, A, lda, B, ldb, bs);
}

void orc_trsm_run(char diag, int m, int n, double alpha, const double* A, int lda,
                  double* B, int ldb) {
    int unit = (diag == 'U' || diag == 'u');
    scale_mat(m, n, alpha, B, ldb);
    /* column j of X: (B(:,j) - sum_{k<j} X(:,k) U(k,j)) / U(j,j) */
    for (int j = 0; j < n; ++j) {
        double* bj = B + (size_t)j * ldb;
        for (int k = 0; k < j; ++k) {
            double u = A[k + (size_t)j * lda];
            const double* xk = B + (size_t)k * ldb;
#pragma omp simd
            for (int i = 0; i < m; ++i) bj[i] -= xk[i] * u;
        }
        if (!unit) {
            double r = A[j + (size_t)j * lda];
            for (int i = 0; i < m; ++i) bj[i] /= r;
        }
    }
}

int orc_potrf_upper(int n, double* A, int lda, int bs) {
    if (bs <= 0) bs = 256;
    for (int i = 0; i < n; i += bs) {
        int b = MIN(bs, n - i);
        double* A11 = A + i + (size_t)i * lda;
        int info = orc_potf2_upper(b, A11, lda);
        if (info) return i + info; /* lapack_like.h:25-28: stop */
        if (i + b < n) {
            double* A12 = A + i + (size_t)(i + b) * lda;
            double* A22 = A + (i + b) + (size_t)(i + b) * lda;
            trsm_lut_base(0, b, n - i - b, A11, lda, A12, lda); /* strsm_, lapack_like.h:36 */
            orc_syrk_ut(n - i - b, b, -1.0, A12, lda, 1.0, A22, lda); /* lapack_like.h:42-44 */
        }
    }
    return 0;
}

int orc_potrf_upper_recursive(int n, double* A, int lda, int bs) {
    if (bs <= 0) bs = 32;
    if (n <= bs) return orc_potf2_upper(n, A, lda);
    int n1 = n / 2, n2 = n - n1;
    int info = orc_potrf_upper_recursive(n1, A, lda, bs);
    if (info) return info;
    double* G12 = A + (size_t)n1 * lda;
    double* A22 = A + n1 + (size_t)n1 * lda;
    orc_trsm_lut('N', n1, n2, 1.0, A, lda, G12, lda, bs);
    /* matrix_multiply("T","N", G12, G12, A22, -1, 1): full A22, both triangles (lapack_like.h:82) */
    orc_matrix_multiply_tn(n2, n2, n1, -1.0, G12, lda, G12, lda, 1.0, A22, lda, 128);
    info = orc_potrf_upper_recursive(n2, A22, lda, bs);
    return info ? n1 + info : 0;
}

/* ----------------------------------------------------- orthogonalisation */

void orc_householder_qr(int m, int n, double* A, int lda, double* R, int ldr) {
    double* tau = (double*)calloc((size_t)MAX(n, 1), sizeof(double));
    /* geqr2: H_j = I - tau v v^T, v(j) = 1, annihilates A(j+1:m, j) (dlarfg). */
    for (int j = 0; j < n; ++j) {
        double* aj = A + (size_t)j * lda;
        double alpha = aj[j];
        double xn = 0.0;
        for (int i = j + 1; i < m; ++i) xn += aj[i] * aj[i];
        xn = sqrt(xn);
        if (xn == 0.0) {
            tau[j] = 0.0;
        } else {
            double beta = -copysign(sqrt(alpha * alpha + xn * xn), alpha);
            tau[j] = (beta - alpha) / beta;
            double sc = 1.0 / (alpha - beta);
            for (int i = j + 1; i < m; ++i) aj[i] *= sc;
            aj[j] = beta;
        }
        /* apply H_j to A(j:m, j+1:n) */
        if (tau[j] != 0.0) {
#pragma omp parallel for schedule(static)
            for (int k = j + 1; k < n; ++k) {
                double* ak = A + (size_t)k * lda;
                double s = ak[j];
                for (int i = j + 1; i < m; ++i) s += aj[i] * ak[i];
                s *= tau[j];
                ak[j] -= s;
                for (int i = j + 1; i < m; ++i) ak[i] -= s * aj[i];
            }
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) R[i + (size_t)j * ldr] = (i <= j) ? A[i + (size_t)j * lda] : 0.0;
    /* explicit Q = H_0 ... H_{n-1} [I; 0] (ormqr "L","N" on the identity, lapack_like.h:331-366) */
    double* Q = (double*)calloc((size_t)m * n, sizeof(double));
    for (int j = 0; j < n; ++j) Q[j + (size_t)j * m] = 1.0;
    for (int j = n - 1; j >= 0; --j) {
        if (tau[j] == 0.0) continue;
        const double* v = A + (size_t)j * lda;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n; ++k) {
            double* qk = Q + (size_t)k * m;
            double s = qk[j];
            for (int i = j + 1; i < m; ++i) s += v[i] * qk[i];
            s *= tau[j];
            qk[j] -= s;
            for (int i = j + 1; i < m; ++i) qk[i] -= s * v[i];
        }
    }
    for (int j = 0; j < n; ++j) memcpy(A + (size_t)j * lda, Q + (size_t)j * m, sizeof(double) * m);
    free(Q);
    free(tau);
}

/* upper(C) = upper(A) * upper(B), all n x n; C may not alias. */
static void trmm_uu(int n, const double* A, int lda, const double* B, int ldb, double* C, int ldc) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            if (i <= j)
                for (int k = i; k <= j; ++k) s += A[i + (size_t)k * lda] * B[k + (size_t)j * ldb];
            C[i + (size_t)j * ldc] = s;
        }
}

int orc_cholqr2(int m, int n, double* Y, int ldy, double* R, int ldr) {
    double* G = (double*)malloc(sizeof(double) * (size_t)n * n);
    double* R1 = (double*)malloc(sizeof(double) * (size_t)n * n);
    int info = 0;
    for (int sweep = 0; sweep < 2; ++sweep) {
        orc_gemm('T', 'N', n, n, m, 1.0, Y, ldy, Y, ldy, 0.0, G, n);
        int inf = orc_potrf_upper(n, G, n, 64);
        if (inf) {
            info = sweep == 0 ? inf : 1000000 + inf;
            break;
        }
        for (int j = 0; j < n; ++j)
            for (int i = j + 1; i < n; ++i) G[i + (size_t)j * n] = 0.0;
        orc_trsm_run('N', m, n, 1.0, G, n, Y, ldy);
        if (sweep == 0)
            memcpy(R1, G, sizeof(double) * (size_t)n * n);
        else
            trmm_uu(n, G, n, R1, n, R, ldr);
    }
    free(G);
    free(R1);
    return info;
}

/* ------------------------------------------------------------- pipeline */

static int orth_inplace(int orth, int n, int l, double* Y, int ldy) {
    double* R = (double*)malloc(sizeof(double) * (size_t)l * l);
    int info = 0;
    if (orth == 0)
        orc_householder_qr(n, l, Y, ldy, R, l);
    else
        info = orc_cholqr2(n, l, Y, ldy, R, l);
    free(R);
    return info;
}

int orc_lra(int n, int d, int l, int q, const double* X, int ldx, const double* y,
            double gamma, const double* Omega, int ldo, int orth, int chunk,
            double* Q, int ldq, double* W, int ldw) {
    /* Y = K*Omega (DArray::LRA, testsvm.cpp:94), orth (testsvm.cpp:105),
     * power iterations, then K*Q (testsvm.cpp:118). */
    orc_rbf_project(n, n, d, l, X, ldx, X, ldx, y, y, gamma, Omega, ldo, 0.0, Q, ldq, chunk);
    int info = orth_inplace(orth, n, l, Q, ldq);
    if (info) return info;
    for (int it = 0; it < q; ++it) {
        orc_rbf_project(n, n, d, l, X, ldx, X, ldx, y, y, gamma, Q, ldq, 0.0, W, ldw, chunk);
        for (int c = 0; c < l; ++c)
            memcpy(Q + (size_t)c * ldq, W + (size_t)c * ldw, sizeof(double) * n);
        info = orth_inplace(orth, n, l, Q, ldq);
        if (info) return info;
    }
    orc_rbf_project(n, n, d, l, X, ldx, X, ldx, y, y, gamma, Q, ldq, 0.0, W, ldw, chunk);
    return 0;
}

void orc_lra_error(int n, int d, int l, const double* X, int ldx, const double* y, double gamma,
                   const double* Q, int ldq, const double* W, int ldw, int chunk, double* err2) {
    (void)chunk;
    const int MB = 64, NB = 256;
    double* nrm = (double*)malloc(sizeof(double) * (size_t)MAX(n, 1));
    orc_sqnorms(n, d, X, ldx, nrm);
    int nblk = (n + MB - 1) / MB;
    double* part = (double*)calloc((size_t)2 * MAX(nblk, 1), sizeof(double));
#pragma omp parallel
    {
        double* kb = (double*)malloc(sizeof(double) * MB * NB);
        double* pb = (double*)malloc(sizeof(double) * MB * NB);
#pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < nblk; ++b) {
            int i0 = b * MB, mb = MIN(MB, n - i0);
            double se = 0.0, sk = 0.0;
            for (int j0 = 0; j0 < n; j0 += NB) {
                int nb = MIN(NB, n - j0);
                rbf_block(i0, mb, j0, nb, d, X, ldx, X, ldx, nrm, nrm, y, y, gamma, 0, kb);
                /* P = Q[iblk,:] * W[jblk,:]^T */
                memset(pb, 0, sizeof(double) * (size_t)mb * nb);
                for (int c = 0; c < l; ++c) {
                    const double* qc = Q + i0 + (size_t)c * ldq;
                    for (int jj = 0; jj < nb; ++jj) {
                        double w = W[j0 + jj + (size_t)c * ldw];
                        double* pc = pb + (size_t)jj * mb;
#pragma omp simd
                        for (int ii = 0; ii < mb; ++ii) pc[ii] += qc[ii] * w;
                    }
                }
                for (size_t e = 0; e < (size_t)mb * nb; ++e) {
                    double kv = kb[e], df = kv - pb[e];
                    se += df * df;
                    sk += kv * kv;
                }
            }
            part[2 * b] = se;
            part[2 * b + 1] = sk;
        }
        free(kb);
        free(pb);
    }
    double se = 0.0, sk = 0.0;
    for (int b = 0; b < nblk; ++b) {
        se += part[2 * b];
        sk += part[2 * b + 1];
    }
    err2[0] = se;
    err2[1] = sk;
    free(part);
    free(nrm);
}

/* ------------------------------------------------------------ Nystrom factor */

/* Round-robin tournament position -> column: position 0 fixed, the others
 * rotate by one per round.  Columns >= n are the padding player (odd n). */
static int rr_player(int pos, int round, int N) {
    return pos == 0 ? 0 : 1 + (pos - 1 + round) % (N - 1);
}

int orc_gesvj(int m, int n, double* A, int lda, double* sva, double* V, int ldv, double tol,
              int max_sweeps) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) V[i + (size_t)j * ldv] = (i == j) ? 1.0 : 0.0;
    const int N = n + (n & 1);
    int sweep = 0, converged = (n < 2);
    for (; sweep < max_sweeps && !converged; ++sweep) {
        int rotated = 0;
        for (int r = 0; r < N - 1; ++r) {
            /* the pairs of one round touch disjoint columns */
#pragma omp parallel for schedule(static) reduction(+ : rotated)
            for (int i = 0; i < N / 2; ++i) {
                int p = rr_player(i, r, N), q = rr_player(N - 1 - i, r, N);
                if (p >= n || q >= n) continue;
                if (p > q) {
                    int t = p;
                    p = q;
                    q = t;
                }
                double* ap = A + (size_t)p * lda;
                double* aq = A + (size_t)q * lda;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int e = 0; e < m; ++e) {
                    al += ap[e] * ap[e];
                    be += aq[e] * aq[e];
                    ga += ap[e] * aq[e];
                }
                if (!(fabs(ga) > tol * sqrt(al * be))) continue;
                double zeta = (be - al) / (2.0 * ga);
                double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int e = 0; e < m; ++e) {
                    double x = ap[e], y = aq[e];
                    ap[e] = c * x - s * y;
                    aq[e] = s * x + c * y;
                }
                double* vp = V + (size_t)p * ldv;
                double* vq = V + (size_t)q * ldv;
                for (int e = 0; e < n; ++e) {
                    double x = vp[e], y = vq[e];
                    vp[e] = c * x - s * y;
                    vq[e] = s * x + c * y;
                }
                rotated += 1;
            }
        }
        converged = (rotated == 0);
    }
    /* singular values = column norms, sorted descending (stable on ties) */
    int* ord = (int*)malloc(sizeof(int) * (size_t)MAX(n, 1));
    double* nv = (double*)malloc(sizeof(double) * (size_t)MAX(n, 1));
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int e = 0; e < m; ++e) s += A[e + (size_t)j * lda] * A[e + (size_t)j * lda];
        nv[j] = sqrt(s);
        ord[j] = j;
    }
    for (int a = 1; a < n; ++a) { /* insertion sort, descending */
        int o = ord[a], b = a - 1;
        while (b >= 0 && nv[ord[b]] < nv[o]) {
            ord[b + 1] = ord[b];
            --b;
        }
        ord[b + 1] = o;
    }
    double* tmpA = (double*)malloc(sizeof(double) * (size_t)MAX(m, 1) * MAX(n, 1));
    double* tmpV = (double*)malloc(sizeof(double) * (size_t)MAX(n, 1) * MAX(n, 1));
    for (int j = 0; j < n; ++j) {
        memcpy(tmpA + (size_t)j * m, A + (size_t)ord[j] * lda, sizeof(double) * m);
        memcpy(tmpV + (size_t)j * n, V + (size_t)ord[j] * ldv, sizeof(double) * n);
        sva[j] = nv[ord[j]];
    }
    for (int j = 0; j < n; ++j) {
        memcpy(A + (size_t)j * lda, tmpA + (size_t)j * m, sizeof(double) * m);
        memcpy(V + (size_t)j * ldv, tmpV + (size_t)j * n, sizeof(double) * n);
    }
    free(tmpA);
    free(tmpV);
    free(nv);
    free(ord);
    return converged ? sweep : -1;
}

int orc_nystrom(int n, int l, int k, const double* Q, int ldq, const double* W, int ldw,
                double* Z, int ldz, double* lam) {
    double* C = (double*)malloc(sizeof(double) * (size_t)MAX(l, 1) * MAX(l, 1));
    double* V = (double*)malloc(sizeof(double) * (size_t)MAX(l, 1) * MAX(l, 1));
    double* sv = (double*)malloc(sizeof(double) * (size_t)MAX(l, 1));
    double* ev = (double*)malloc(sizeof(double) * (size_t)MAX(l, 1));
    int* ord = (int*)malloc(sizeof(int) * (size_t)MAX(l, 1));
    /* C = Q^T W (testsvm.cpp:126-137), symmetrised */
    orc_gemm('T', 'N', l, l, n, 1.0, Q, ldq, W, ldw, 0.0, C, l);
    for (int j = 0; j < l; ++j)
        for (int i = 0; i < j; ++i) {
            double s = 0.5 * (C[i + (size_t)j * l] + C[j + (size_t)i * l]);
            C[i + (size_t)j * l] = C[j + (size_t)i * l] = s;
        }
    double* C0 = (double*)malloc(sizeof(double) * (size_t)MAX(l, 1) * MAX(l, 1));
    memcpy(C0, C, sizeof(double) * (size_t)l * l);
    int st = orc_gesvj(l, l, C, l, sv, V, l, 1e-14, 60);
    /* signed eigenvalues lam_i = v_i . (C0 v_i); C holds C0 V after the sweeps */
    for (int j = 0; j < l; ++j) {
        double s = 0.0;
        for (int e = 0; e < l; ++e) s += V[e + (size_t)j * l] * C[e + (size_t)j * l];
        ev[j] = s;
        ord[j] = j;
    }
    for (int a = 1; a < l; ++a) {
        int o = ord[a], b = a - 1;
        while (b >= 0 && ev[ord[b]] < ev[o]) {
            ord[b + 1] = ord[b];
            --b;
        }
        ord[b + 1] = o;
    }
    /* Vk = V[:, top k] diag(sqrt(max(lam, 0))); Z = Q Vk */
    double* Vk = (double*)malloc(sizeof(double) * (size_t)MAX(l, 1) * MAX(k, 1));
    for (int j = 0; j < k; ++j) {
        double lj = ev[ord[j]], sc = lj > 0.0 ? sqrt(lj) : 0.0;
        if (lam) lam[j] = lj;
        for (int e = 0; e < l; ++e) Vk[e + (size_t)j * l] = V[e + (size_t)ord[j] * l] * sc;
    }
    orc_gemm('N', 'N', n, k, l, 1.0, Q, ldq, Vk, l, 0.0, Z, ldz);
    free(Vk);
    free(C0);
    free(ord);
    free(ev);
    free(sv);
    free(V);
    free(C);
    return st < 0 ? 1 : 0;
}

/* ------------------------------------------------------ barrier SVM solver */

/* u = Z^T x (k) */
static void zt_x(int n, int k, const double* Z, int ldz, const double* x, double* u) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < k; ++c) {
        const double* z = Z + (size_t)c * ldz;
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += z[i] * x[i];
        u[c] = s;
    }
}

/* out = Z u (n) */
static void z_u(int n, int k, const double* Z, int ldz, const double* u, double* out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int c = 0; c < k; ++c) s += Z[i + (size_t)c * ldz] * u[c];
        out[i] = s;
    }
}

static double dot_n(int n, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* sum(log(a + tt s) + log(C - a - tt s)); +inf outside the open box */
static double log_barrier(int n, const double* a, const double* s, double tt, double C) {
    double acc = 0.0;
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        double v = a[i] + (s ? tt * s[i] : 0.0);
        if (!(v > 0.0) || !(C - v > 0.0)) bad = 1;
        else acc += log(v) + log(C - v);
    }
    return bad ? -INFINITY : acc;
}

int orc_svm_barrier(int n, int k, const double* Z, int ldz, const double* y, double C,
                    double cg_tol, int max_outer, int max_newton, int max_cg, double* a,
                    int* stats, double* fval) {
    int plus = 0, minus = 0;
    for (int i = 0; i < n; ++i) {
        if (y[i] > 0) ++plus;
        else ++minus;
    }
    if (plus == 0 || minus == 0) return -1;
    double pval, mval; /* testsvm.cpp:204-211 */
    if (plus > minus) {
        mval = 0.9 * C;
        pval = mval * minus / plus;
    } else {
        pval = 0.9 * C;
        mval = pval * plus / minus;
    }
    for (int i = 0; i < n; ++i) a[i] = y[i] > 0 ? pval : mval;
    double* w = (double*)malloc(sizeof(double) * (size_t)n * 9 + sizeof(double) * (size_t)k * 2);
    double *M = w, *grad = w + n, *D = w + 2 * (size_t)n, *s = w + 3 * (size_t)n;
    double *r = w + 4 * (size_t)n, *z = w + 5 * (size_t)n, *p = w + 6 * (size_t)n;
    double *q = w + 7 * (size_t)n, *Pc = w + 8 * (size_t)n;
    double *u = w + 9 * (size_t)n, *v = u + k;
    for (int i = 0; i < n; ++i) { /* M = diag(Z Z^T), testsvm.cpp:236-243 */
        double m2 = 0.0;
        for (int c = 0; c < k; ++c) m2 += Z[i + (size_t)c * ldz] * Z[i + (size_t)c * ldz];
        M[i] = m2;
    }
    double t = 1.0;
    const double ls_alpha = 0.1, ls_beta = 0.5, mu = 20.0;
    int outer = 0, newton = 0, cgit = 0;
    double sum_a = 0.0, fcur = 0.0;
    for (outer = 1; outer <= max_outer; ++outer) {
        for (int ni = 1; ni <= max_newton; ++ni) {
            /* gradient and barrier Hessian */
            zt_x(n, k, Z, ldz, a, u);
            z_u(n, k, Z, ldz, u, grad);
            double dmax = 0.0, dmin = INFINITY;
            for (int i = 0; i < n; ++i) {
                double ai = a[i], ci = C - ai;
                grad[i] += -1.0 + (-1.0 / ai + 1.0 / ci) / t;
                D[i] = (1.0 / (ai * ai) + 1.0 / (ci * ci)) / t;
                dmax = D[i] > dmax ? D[i] : dmax;
                dmin = D[i] < dmin ? D[i] : dmin;
            }
            if (dmax / dmin > 1e18) break;
            /* pre_conjgrad: (Z Z^T + diag(D)) s = -grad, Jacobi preconditioner */
            for (int i = 0; i < n; ++i) {
                s[i] = 0.0;
                r[i] = -grad[i];
                Pc[i] = M[i] + D[i];
                z[i] = r[i] / Pc[i];
                p[i] = z[i];
            }
            double rs = dot_n(n, r, z), normr = sqrt(dot_n(n, r, r));
            for (int it = 0; normr > cg_tol && it < max_cg; ++it) {
                zt_x(n, k, Z, ldz, p, v);
                z_u(n, k, Z, ldz, v, q);
                for (int i = 0; i < n; ++i) q[i] += D[i] * p[i];
                double al = rs / dot_n(n, p, q);
                for (int i = 0; i < n; ++i) {
                    s[i] += al * p[i];
                    r[i] -= al * q[i];
                    z[i] = r[i] / Pc[i];
                }
                double rs_new = dot_n(n, r, z);
                double be = rs_new / rs;
                for (int i = 0; i < n; ++i) p[i] = z[i] + be * p[i];
                rs = rs_new;
                normr = sqrt(dot_n(n, r, r));
                ++cgit;
            }
            ++newton;
            /* line search on g(a + tt s) = f - log barrier / t; f through Z^T */
            double gs = dot_n(n, grad, s);
            double nd2 = -gs;
            zt_x(n, k, Z, ldz, s, v);
            sum_a = 0.0;
            double sum_s = 0.0;
            for (int i = 0; i < n; ++i) {
                sum_a += a[i];
                sum_s += s[i];
            }
            double uu = dot_n(k, u, u), uv = dot_n(k, u, v), vv = dot_n(k, v, v);
            double g0 = 0.5 * uu - sum_a - log_barrier(n, a, NULL, 0.0, C) / t;
            double tt = 1.0, gnew = g0;
            for (int ls = 0; ls < 64; ++ls) {
                double fn = 0.5 * (uu + 2.0 * tt * uv + tt * tt * vv) - (sum_a + tt * sum_s);
                gnew = fn - log_barrier(n, a, s, tt, C) / t;
                if (gnew <= g0 + ls_alpha * tt * gs) break;
                tt *= ls_beta;
            }
            if (!(gnew <= g0 + ls_alpha * tt * gs)) tt = 0.0, gnew = g0;
            for (int i = 0; i < n; ++i) a[i] += tt * s[i];
            if (nd2 / 2 < 1e-6) break;
            if (ni > 1 && (g0 - gnew) < 1e-9 * fabs(gnew)) break;
        }
        zt_x(n, k, Z, ldz, a, u);
        sum_a = 0.0;
        for (int i = 0; i < n; ++i) sum_a += a[i];
        fcur = 0.5 * dot_n(k, u, u) - sum_a;
        if (2.0 * n / t < fabs(fcur) * 1e-5) break;
        t *= mu;
    }
    if (outer > max_outer) outer = max_outer;
    if (stats) {
        stats[0] = outer;
        stats[1] = newton;
        stats[2] = cgit;
    }
    if (fval) *fval = fcur;
    free(w);
    return 0;
}
