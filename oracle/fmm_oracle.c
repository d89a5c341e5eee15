/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement, in plain C (OpenMP), of the reference's FP64 ABC fast
 * matrix multiplication path and of its plain GEMMs.  Only tests/, the smoke
 * check and the benchmark's cpu_baseline leg may load this library, and only
 * as the checker or the timed CPU baseline; the product path never calls it.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).  FP contraction
 * is off so every a*b and every + is rounded separately, the arithmetic of the
 * reference's vmulpd/vaddpd micro-kernels (no FMA).
 *
 * Functions and what they follow (paths relative to the reference root):
 *   oracle_dgemm_ref   common/dgemm/bl_dgemm_ref.c:27-48   C += A*B, j,i,p loop
 *   oracle_dgemm_blis  common/dgemm/my_dgemm.c:115-210 and
 *                      common/kernels/bl_dgemm_asm_8x4.c:3-341
 *                      per KC=256 block: t = sum_p a*b (sequential, from 0),
 *                      then C = C + t
 *   oracle_peeling     common/dgemm/bl_dgemm_util.c:316-353
 *   oracle_fmm_abc     meta/abc_gen_mix.py:171-251 (packing sums),
 *                      :302-391 (per-product jc/pc/ic loops), :448-472 (top
 *                      level, product order r = 0..R-1, then peeling);
 *                      meta/abc_micro_kernel_gen.py:147-196 (C_j +-= AB) and
 *                      common/kernels/bl_dgemm_asm_8x4_mulstrassen.c:324-383
 *                      (C_j += alpha_j*AB); meta/common_mix.py:124-171,284-310
 *                      (partition and block decoding)
 *   oracle_fmm_ab      meta/ab_gen_mix.py:219-303,341-372 (products into a
 *                      temporary M per KC block, then M_Add)
 *   oracle_fmm_naive   meta/naive_gen_mix.py:60-137,244-361 (explicit S, T,
 *                      M temporaries, bl_dgemm, M_Add)
 *
 * Pinning (tests/test_oracle.py): exact small-integer known-answer tests against
 * numpy integer matmul; bitwise equality of oracle_dgemm_blis and
 * oracle_peeling with the reference's own bl_dgemm / bl_dynamic_peeling
 * compiled from its sources into oracle/_ref; and the per-KC-block epilogue
 * C + alpha*AB against bl_dgemm_asm_8x4_mulstrassen from the same build.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define O_MR 8
#define O_NR 4
#define O_KC 256

/* ------------------------------------------------------------------ GEMMs */

void oracle_dgemm_ref(int m, int n, int k, const double *A, int lda,
                      const double *B, int ldb, double *C, int ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double c = C[(size_t)j * ldc + i];
            for (int p = 0; p < k; ++p)
                c = c + A[(size_t)p * lda + i] * B[(size_t)j * ldb + p];
            C[(size_t)j * ldc + i] = c;
        }
}

/* C[i,j] += sum over KC blocks of (sequential dot over the block).
 * Inner kernel works on an O_MR x O_NR register tile, p sequential, the
 * summation order of bl_dgemm_asm_8x4 for every element. */
static void blis_block(int m, int n, int kb, const double *A, int lda,
                       const double *B, int ldb, double *C, int ldc)
{
    for (int j0 = 0; j0 < n; j0 += O_NR) {
        int nb = n - j0 < O_NR ? n - j0 : O_NR;
        for (int i0 = 0; i0 < m; i0 += O_MR) {
            int mb = m - i0 < O_MR ? m - i0 : O_MR;
            double t[O_NR][O_MR];
            memset(t, 0, sizeof t);
            for (int p = 0; p < kb; ++p) {
                const double *a = A + (size_t)p * lda + i0;
                for (int jj = 0; jj < nb; ++jj) {
                    double b = B[(size_t)(j0 + jj) * ldb + p];
                    for (int ii = 0; ii < mb; ++ii) t[jj][ii] = t[jj][ii] + a[ii] * b;
                }
            }
            for (int jj = 0; jj < nb; ++jj)
                for (int ii = 0; ii < mb; ++ii) {
                    double *c = C + (size_t)(j0 + jj) * ldc + i0 + ii;
                    *c = *c + t[jj][ii];
                }
        }
    }
}

void oracle_dgemm_blis(int m, int n, int k, const double *A, int lda,
                       const double *B, int ldb, double *C, int ldc)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    for (int pc = 0; pc < k; pc += O_KC) {
        int pb = k - pc < O_KC ? k - pc : O_KC;
        int ncb = (n + O_NR - 1) / O_NR;
#pragma omp parallel for schedule(static)
        for (int jb = 0; jb < ncb; ++jb) {
            int j0 = jb * O_NR;
            int nb = n - j0 < O_NR ? n - j0 : O_NR;
            blis_block(m, nb, pb, A + (size_t)pc * lda, lda,
                       B + pc + (size_t)j0 * ldb, ldb, C + (size_t)j0 * ldc, ldc);
        }
    }
}

void oracle_peeling(int m, int n, int k, const double *A, int lda,
                    const double *B, int ldb, double *C, int ldc,
                    int dim1, int dim2, int dim3)
{
    int mr = m % dim1, kr = k % dim2, nr = n % dim3;
    int ms = m - mr, ns = n - nr, ks = k - kr;
    if (kr > 0 && ms > 0 && ns > 0)
        oracle_dgemm_blis(ms, ns, kr, A + (size_t)ks * lda, lda, B + ks, ldb, C, ldc);
    if (nr > 0)
        oracle_dgemm_blis(m, nr, k, A, lda, B + (size_t)ns * ldb, ldb, C + (size_t)ns * ldc, ldc);
    if (mr > 0 && ns > 0)
        oracle_dgemm_blis(mr, ns, k, A + ms, lda, B, ldb, C + ms, ldc);
}

/* ------------------------------------------------------------- FMM plan */

typedef struct { double coef; size_t off; } oterm;  /* element offset of a block */

typedef struct {
    int R;
    int *na, *nb, *nc;        /* per product term counts */
    oterm **ta, **tb, **tc;   /* per product term lists, reference row order */
} oplan;

/* Decode composed row `row` of table which (0:U,1:V,2:W) into grid block
 * (I,J); digits are peeled from the last level (meta/common_mix.py:296-310). */
static void decode_row(int nlev, const int *dims, int which, int row, int *I, int *J)
{
    int a = which == 1 ? 1 : 0, b = which == 0 ? 1 : 2;
    int rb = 1, cb = 1, ii = 0, jj = 0, rest = row;
    for (int l = nlev - 1; l >= 0; --l) {
        int nr = dims[3 * l + a], nc = dims[3 * l + b];
        int d = rest % (nr * nc);
        rest /= nr * nc;
        ii += (d / nc) * rb;
        jj += (d % nc) * cb;
        rb *= nr;
        cb *= nc;
    }
    *I = ii;
    *J = jj;
}

static void build_terms(int nlev, const int *dims, int which, int rows, int R,
                        const double *T, size_t rstride, size_t cstride,
                        int *cnt, oterm **lists)
{
    for (int r = 0; r < R; ++r) {
        int c = 0;
        for (int a = 0; a < rows; ++a) if (T[(size_t)a * R + r] != 0.0) ++c;
        cnt[r] = c;
        lists[r] = (oterm *)malloc(sizeof(oterm) * (c ? c : 1));
        c = 0;
        for (int a = 0; a < rows; ++a) {
            double v = T[(size_t)a * R + r];
            if (v == 0.0) continue;
            int I, J;
            decode_row(nlev, dims, which, a, &I, &J);
            lists[r][c].coef = v;
            lists[r][c].off = (size_t)I * rstride + (size_t)J * cstride;
            ++c;
        }
    }
}

static void free_plan(oplan *p)
{
    for (int r = 0; r < p->R; ++r) { free(p->ta[r]); free(p->tb[r]); free(p->tc[r]); }
    free(p->na); free(p->nb); free(p->nc); free(p->ta); free(p->tb); free(p->tc);
}

static void make_plan(oplan *p, int nlev, const int *dims, int R,
                      const double *U, const double *V, const double *W,
                      int ms, int ks, int ns, int lda, int ldb, int ldc)
{
    int Mx = 1, Kx = 1, Nx = 1;
    for (int l = 0; l < nlev; ++l) { Mx *= dims[3*l]; Kx *= dims[3*l+1]; Nx *= dims[3*l+2]; }
    p->R = R;
    p->na = malloc(sizeof(int) * R); p->nb = malloc(sizeof(int) * R); p->nc = malloc(sizeof(int) * R);
    p->ta = malloc(sizeof(oterm *) * R); p->tb = malloc(sizeof(oterm *) * R); p->tc = malloc(sizeof(oterm *) * R);
    build_terms(nlev, dims, 0, Mx * Kx, R, U, (size_t)ms, (size_t)ks * lda, p->na, p->ta);
    build_terms(nlev, dims, 1, Kx * Nx, R, V, (size_t)ks, (size_t)ns * ldb, p->nb, p->tb);
    build_terms(nlev, dims, 2, Mx * Nx, R, W, (size_t)ms, (size_t)ns * ldc, p->nc, p->tc);
}

static void dims_product(int nlev, const int *dims, int *Mx, int *Kx, int *Nx)
{
    *Mx = *Kx = *Nx = 1;
    for (int l = 0; l < nlev; ++l) { *Mx *= dims[3*l]; *Kx *= dims[3*l+1]; *Nx *= dims[3*l+2]; }
}

/* S[i + p*ms] = sum_t u_t A_t[i + p*lda], terms left to right
 * (meta/abc_gen_mix.py:229-251: first term as is / negated / scaled, then
 * + or - or + (c)*).  Dyadic coefficients make u*a exact, so c*a then add is
 * the reference's rounding for every coefficient. */
static void form_sum(const oterm *t, int nt, const double *X, int ldx,
                     int rows, int cols, double *S, int lds)
{
#pragma omp parallel for schedule(static)
    for (int p = 0; p < cols; ++p)
        for (int i = 0; i < rows; ++i) {
            double s = t[0].coef * X[t[0].off + (size_t)p * ldx + i];
            for (int q = 1; q < nt; ++q) s = s + t[q].coef * X[t[q].off + (size_t)p * ldx + i];
            S[(size_t)p * lds + i] = s;
        }
}

/* One product: for every KC block, AB = S(:,pc) T(pc,:) (sequential p from 0),
 * then every target C_j = C_j + w_j*AB (meta/abc_micro_kernel_gen.py:147-196,
 * bl_dgemm_asm_8x4_mulstrassen.c:324-383). */
static void product_update(int ms, int ns, int ks, const double *S, const double *T,
                           const oterm *tc, int nc, double *C, int ldc)
{
    for (int pc = 0; pc < ks; pc += O_KC) {
        int pb = ks - pc < O_KC ? ks - pc : O_KC;
        int ncb = (ns + O_NR - 1) / O_NR;
#pragma omp parallel for schedule(static)
        for (int jb = 0; jb < ncb; ++jb) {
            int j0 = jb * O_NR;
            int nb = ns - j0 < O_NR ? ns - j0 : O_NR;
            for (int i0 = 0; i0 < ms; i0 += O_MR) {
                int mb = ms - i0 < O_MR ? ms - i0 : O_MR;
                double t[O_NR][O_MR];
                memset(t, 0, sizeof t);
                for (int p = pc; p < pc + pb; ++p) {
                    const double *a = S + (size_t)p * ms + i0;
                    for (int jj = 0; jj < nb; ++jj) {
                        double b = T[(size_t)(j0 + jj) * ks + p];
                        for (int ii = 0; ii < mb; ++ii) t[jj][ii] = t[jj][ii] + a[ii] * b;
                    }
                }
                for (int q = 0; q < nc; ++q) {
                    double w = tc[q].coef;
                    for (int jj = 0; jj < nb; ++jj)
                        for (int ii = 0; ii < mb; ++ii) {
                            double *c = C + tc[q].off + (size_t)(j0 + jj) * ldc + i0 + ii;
                            *c = *c + w * t[jj][ii];
                        }
                }
            }
        }
    }
}

/* Returns 0 on success, -1 on bad arguments, -2 on allocation failure. */
int oracle_fmm_abc(int nlev, const int *dims, int R, const double *U,
                   const double *V, const double *W, int m, int n, int k,
                   const double *A, int lda, const double *B, int ldb,
                   double *C, int ldc)
{
    if (nlev <= 0 || R <= 0 || m < 0 || n < 0 || k < 0) return -1;
    int Mx, Kx, Nx;
    dims_product(nlev, dims, &Mx, &Kx, &Nx);
    int md = m - m % (Mx * O_MR), kd = k - k % Kx, nd = n - n % (Nx * O_NR);
    int ms = md / Mx, ks = kd / Kx, ns = nd / Nx;
    if (ms > 0 && ks > 0 && ns > 0) {
        oplan p;
        make_plan(&p, nlev, dims, R, U, V, W, ms, ks, ns, lda, ldb, ldc);
        double *S = malloc(sizeof(double) * (size_t)ms * ks);
        double *T = malloc(sizeof(double) * (size_t)ks * ns);
        if (!S || !T) { free(S); free(T); free_plan(&p); return -2; }
        for (int r = 0; r < R; ++r) {
            if (p.na[r] == 0 || p.nb[r] == 0 || p.nc[r] == 0) continue;
            form_sum(p.ta[r], p.na[r], A, lda, ms, ks, S, ms);
            form_sum(p.tb[r], p.nb[r], B, ldb, ks, ns, T, ks);
            product_update(ms, ns, ks, S, T, p.tc[r], p.nc[r], C, ldc);
        }
        free(S); free(T);
        free_plan(&p);
    }
    oracle_peeling(m, n, k, A, lda, B, ldb, C, ldc, Mx * O_MR, Kx, Nx * O_NR);
    return 0;
}

/* AB variant: each product's KC-block partial goes to a zeroed temporary M
 * (M = M + AB per block, i.e. the full product), then C_j = C_j + w_j*M. */
int oracle_fmm_ab(int nlev, const int *dims, int R, const double *U,
                  const double *V, const double *W, int m, int n, int k,
                  const double *A, int lda, const double *B, int ldb,
                  double *C, int ldc)
{
    if (nlev <= 0 || R <= 0 || m < 0 || n < 0 || k < 0) return -1;
    int Mx, Kx, Nx;
    dims_product(nlev, dims, &Mx, &Kx, &Nx);
    int md = m - m % (Mx * O_MR), kd = k - k % Kx, nd = n - n % (Nx * O_NR);
    int ms = md / Mx, ks = kd / Kx, ns = nd / Nx;
    if (ms > 0 && ks > 0 && ns > 0) {
        oplan p;
        make_plan(&p, nlev, dims, R, U, V, W, ms, ks, ns, lda, ldb, ldc);
        double *S = malloc(sizeof(double) * (size_t)ms * ks);
        double *T = malloc(sizeof(double) * (size_t)ks * ns);
        double *M = malloc(sizeof(double) * (size_t)ms * ns);
        if (!S || !T || !M) { free(S); free(T); free(M); free_plan(&p); return -2; }
        oterm one = {1.0, 0};
        for (int r = 0; r < R; ++r) {
            if (p.na[r] == 0 || p.nb[r] == 0 || p.nc[r] == 0) continue;
            form_sum(p.ta[r], p.na[r], A, lda, ms, ks, S, ms);
            form_sum(p.tb[r], p.nb[r], B, ldb, ks, ns, T, ks);
            memset(M, 0, sizeof(double) * (size_t)ms * ns);
            product_update(ms, ns, ks, S, T, &one, 1, M, ms);
            for (int q = 0; q < p.nc[r]; ++q) {
                double w = p.tc[r][q].coef;
                double *Cb = C + p.tc[r][q].off;
#pragma omp parallel for schedule(static)
                for (int j = 0; j < ns; ++j)
                    for (int i = 0; i < ms; ++i)
                        Cb[(size_t)j * ldc + i] = Cb[(size_t)j * ldc + i] + w * M[(size_t)j * ms + i];
            }
        }
        free(S); free(T); free(M);
        free_plan(&p);
    }
    oracle_peeling(m, n, k, A, lda, B, ldb, C, ldc, Mx * O_MR, Kx, Nx * O_NR);
    return 0;
}

/* Naive variant: explicit S and T, M = S*T by the plain blocked GEMM into a
 * zeroed temporary, then C_j = C_j + w_j*M. */
int oracle_fmm_naive(int nlev, const int *dims, int R, const double *U,
                     const double *V, const double *W, int m, int n, int k,
                     const double *A, int lda, const double *B, int ldb,
                     double *C, int ldc)
{
    if (nlev <= 0 || R <= 0 || m < 0 || n < 0 || k < 0) return -1;
    int Mx, Kx, Nx;
    dims_product(nlev, dims, &Mx, &Kx, &Nx);
    int md = m - m % (Mx * O_MR), kd = k - k % Kx, nd = n - n % (Nx * O_NR);
    int ms = md / Mx, ks = kd / Kx, ns = nd / Nx;
    if (ms > 0 && ks > 0 && ns > 0) {
        oplan p;
        make_plan(&p, nlev, dims, R, U, V, W, ms, ks, ns, lda, ldb, ldc);
        double *S = malloc(sizeof(double) * (size_t)ms * ks);
        double *T = malloc(sizeof(double) * (size_t)ks * ns);
        double *M = malloc(sizeof(double) * (size_t)ms * ns);
        if (!S || !T || !M) { free(S); free(T); free(M); free_plan(&p); return -2; }
        for (int r = 0; r < R; ++r) {
            if (p.na[r] == 0 || p.nb[r] == 0 || p.nc[r] == 0) continue;
            form_sum(p.ta[r], p.na[r], A, lda, ms, ks, S, ms);
            form_sum(p.tb[r], p.nb[r], B, ldb, ks, ns, T, ks);
            memset(M, 0, sizeof(double) * (size_t)ms * ns);
            oracle_dgemm_blis(ms, ns, ks, S, ms, T, ks, M, ms);
            for (int q = 0; q < p.nc[r]; ++q) {
                double w = p.tc[r][q].coef;
                double *Cb = C + p.tc[r][q].off;
#pragma omp parallel for schedule(static)
                for (int j = 0; j < ns; ++j)
                    for (int i = 0; i < ms; ++i)
                        Cb[(size_t)j * ldc + i] = Cb[(size_t)j * ldc + i] + w * M[(size_t)j * ms + i];
            }
        }
        free(S); free(T); free(M);
        free_plan(&p);
    }
    oracle_peeling(m, n, k, A, lda, B, ldb, C, ldc, Mx * O_MR, Kx, Nx * O_NR);
    return 0;
}

void oracle_set_threads(int nt) { if (nt > 0) omp_set_num_threads(nt); }
int oracle_get_threads(void) { return omp_get_max_threads(); }
