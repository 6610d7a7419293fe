/*
 * bem_oracle.c -- FP64 CPU restatement of the reference's per-RK4-stage BEM
 * path (collocation, linear elements, flat triangles).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, the smoke entry point and
 * the benchmark's CPU-baseline leg.  It is never part of the product path.
 *
 * The loop structure of the assembly follows ColocSim.cpp:21-110 on purpose
 * (OpenMP over triangles, per-thread N x 3 partial columns, a critical scatter,
 * a serial 4-pi pass) so that it doubles as the reference's CPU baseline.
 */
#include "bem_oracle.h"

/* BO_MUTANT selects a planted mistake for the mutation tests (0: none). */
#ifndef BO_MUTANT
#define BO_MUTANT 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double x, y, z; } v3;

static inline v3 V(const double* p) { v3 r = {p[0], p[1], p[2]}; return r; }
static inline v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3 vadd(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline v3 vscale(double s, v3 a) { v3 r = {s * a.x, s * a.y, s * a.z}; return r; }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {   /* vector3d.hpp:75-80 */
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline double vnorm(v3 a) { return sqrt(vdot(a, a)); }

/* ------------------------------------------------------------------------ */
/* quadrature (quadrature.hpp:30-38 gauss_7, :88-96 quadrature_7)            */
/* ------------------------------------------------------------------------ */

static double g7x[7], g7w[7];          /* Gauss-Legendre on (0,1), weights sum to 1 */
static double t7x[7], t7y[7], t7w[7];  /* degree-5 rule on the unit triangle, sum 1/2 */
static int quad_ready = 0;

static void legendre7(double t, double* p, double* dp) {
    double p0 = 1.0, p1 = t;
    for (int n = 2; n <= 7; ++n) {
        double pn = ((2.0 * n - 1.0) * t * p1 - (n - 1.0) * p0) / n;
        p0 = p1; p1 = pn;
    }
    *p = p1;
    *dp = 7.0 * (t * p1 - p0) / (t * t - 1.0);
}

static void init_quadrature(void) {
    if (quad_ready) return;
    /* roots of P_7 by Newton from Chebyshev-like starting points */
    for (int k = 0; k < 7; ++k) {
        double t = cos(M_PI * (k + 0.75) / 7.5);
        for (int it = 0; it < 100; ++it) {
            double p, dp;
            legendre7(t, &p, &dp);
            double dt = p / dp;
            t -= dt;
            if (fabs(dt) < 1e-17) break;
        }
        double p, dp;
        legendre7(t, &p, &dp);
        double w = 2.0 / ((1.0 - t * t) * dp * dp);
        /* ascending on (0,1) */
        g7x[6 - k] = 0.5 * (1.0 + t);
        g7w[6 - k] = 0.5 * w;
    }
    g7x[3] = 0.5;  /* the middle root is exactly 0 */
    /* Dunavant / Strang-Fix degree-5 7-point rule (closed forms) */
    const double s = sqrt(15.0);
    const double a1 = (9.0 + 2.0 * s) / 21.0, b1 = (6.0 - s) / 21.0;
    const double a2 = (9.0 - 2.0 * s) / 21.0, b2 = (6.0 + s) / 21.0;
    const double w0 = 9.0 / 80.0, w1 = (155.0 - s) / 2400.0, w2 = (155.0 + s) / 2400.0;
    const double X[7] = {1.0 / 3.0, a1, b1, b1, a2, b2, b2};
    const double Y[7] = {1.0 / 3.0, b1, a1, b1, b2, a2, b2};
    const double W[7] = {w0, w1, w1, w1, w2, w2, w2};
    for (int q = 0; q < 7; ++q) { t7x[q] = X[q]; t7y[q] = Y[q]; t7w[q] = W[q]; }
    quad_ready = 1;
}

/* 19-point degree-9 triangle rule (Dunavant 1985), listed in the reference as
 * quadrature_19 (quadrature.hpp:126-146): centroid, four orbits
 * (a, b), (b, a), (b, b) and one six-point orbit, in the reference's order;
 * the reference writes each weight as w * 0.5 (the triangle's area). */
static const double q19_orb[5][3] = {   /* a, b, weight */
    {2.06349616025259287e-2, 0.48968251919873701, 3.13347002271398278e-2},
    {0.12582081701412900, 0.43708959149293553, 7.78275410047754301e-2},
    {0.62359292876193562, 0.18820353561903219, 7.96477389272090969e-2},
    {0.91054097321109406, 4.47295133944529688e-2, 2.55776756586981006e-2},
};
static double t19x[19], t19y[19], t19w[19];
static int q19_ready = 0;
static void init_q19(void) {
    if (q19_ready) return;
    int q = 0;
    t19x[q] = 0.33333333333333331; t19y[q] = 0.33333333333333331; t19w[q] = 9.71357962827961025e-2 * 0.5; ++q;
    for (int o = 0; o < 4; ++o) {
        const double a = q19_orb[o][0], b = q19_orb[o][1], w = q19_orb[o][2] * 0.5;
        t19x[q] = a; t19y[q] = b; t19w[q] = w; ++q;
        t19x[q] = b; t19y[q] = a; t19w[q] = w; ++q;
        t19x[q] = b; t19y[q] = b; t19w[q] = w; ++q;
    }
    const double e1 = 0.74119859878449801, e2 = 3.68384120547362581e-2, e3 = 0.22196298916076573;
    const double we = 4.32835393772893970e-2 * 0.5;
    const double ex[6] = {e1, e1, e2, e2, e3, e3}, ey[6] = {e2, e3, e1, e3, e1, e2};
    for (int k = 0; k < 6; ++k) { t19x[q] = ex[k]; t19y[q] = ey[k]; t19w[q] = we; ++q; }
    q19_ready = 1;
}

void bo_quadrature19(double* tx, double* ty, double* tw) {
    init_q19();
    for (int q = 0; q < 19; ++q) { tx[q] = t19x[q]; ty[q] = t19y[q]; tw[q] = t19w[q]; }
}

void bo_quadrature(double* gx, double* gw, double* tx, double* ty, double* tw) {
    init_quadrature();
    for (int q = 0; q < 7; ++q) {
        gx[q] = g7x[q]; gw[q] = g7w[q];
        tx[q] = t7x[q]; ty[q] = t7y[q]; tw[q] = t7w[q];
    }
}

/* ------------------------------------------------------------------------ */
/* panel integrals (Integrator.hpp:116-119, 197-229, 471-513;                */
/*                  Integrator.cpp:12-89, 237-283; Interpolator.hpp:12-43)   */
/* ------------------------------------------------------------------------ */

typedef struct { v3 a, ab, bc, n; double area; } interp_t;

static interp_t make_interp(v3 a, v3 b, v3 c) {  /* Interpolator.hpp:15-21 */
    interp_t I;
    I.a = a; I.ab = vsub(b, a); I.bc = vsub(c, b);
    I.n = vcross(I.ab, I.bc);
    I.area = vnorm(I.n);          /* twice the geometric area */
    I.n = vscale(1.0 / I.area, I.n);
    return I;
}

static inline v3 interp_at(const interp_t* I, double u, double v) {
    return vadd(I->a, vadd(vscale(u, I->ab), vscale(v, I->bc)));
}

/* integrand(z, n) = (1/|z|, -z.n/|z|^3) (Integrator.hpp:116-119) */
static inline void kernel_pair(v3 z, v3 n, double* g, double* h) {
    double inv = 1.0 / vnorm(z);
    *g = inv;
    *h = -vdot(z, n) * inv * inv * inv;
}

/* integrate_disjoint_coloc[_mir] (Integrator.hpp:471-513) */
static void disjoint(v3 x, const interp_t* I, int mirror, double* g3, double* h3) {
    double G[3] = {0, 0, 0}, H[3] = {0, 0, 0};
    for (int q = 0; q < 7; ++q) {
        const double u = t7x[q] + t7y[q], v = t7y[q];  /* transform to the other unit triangle */
        v3 y = interp_at(I, u, v);
        double g, h;
        kernel_pair(vsub(y, x), I->n, &g, &h);
        if (mirror) {                                   /* integrand_coloc_mir, :211-229 */
            v3 ym = y, nm = I->n;
            ym.x = -ym.x; nm.x = -nm.x;
            double g2, h2;
            kernel_pair(vsub(ym, x), nm, &g2, &h2);
            g += g2; h += h2;
        }
        const double phi[3] = {1.0 - u, u - v, v};     /* get_linear_elements, ResultTypes.hpp:111-117 */
        for (int k = 0; k < 3; ++k) {
            G[k] += phi[k] * g * t7w[q];
            H[k] += phi[k] * h * t7w[q];
        }
    }
    for (int k = 0; k < 3; ++k) { g3[k] = G[k] * I->area; h3[k] = H[k] * I->area; }
}

/* integrate_identical_coloc<LinElm> (Integrator.cpp:12-41): x = vertex a */
static void identical(const interp_t* I, double* g3) {
    const double jac = 0.5 * I->area * (M_PI / 4.0);
    double t0 = 0, t1 = 0, t2 = 0;
    for (int p = 0; p < 7; ++p) {
        const double th = g7x[p] * (M_PI / 4.0);
        const double c = cos(th), s = sin(th);
        const double dist = vnorm(vadd(vscale(c, I->ab), vscale(s, I->bc)));
        const double f = jac * g7w[p] / dist;
        const double fac = 1.0 / (c * c);
        t0 += f * c * fac;
        t1 += f * (c - s) * fac;
        t2 += f * s * fac;
    }
    g3[0] = t0; g3[1] = t1; g3[2] = t2;
}

/* one (collocation point, triangle) pair in triangle-vertex order
 * (integrate_Lin_coloc_local[_mir], Integrator.cpp:237-283) */
static void coloc_local(v3 x, v3 A, v3 B, v3 C, int self, int mirror, double* g3, double* h3) {
    if (self < 0) {
        interp_t I = make_interp(A, B, C);
        disjoint(x, &I, mirror, g3, h3);
        return;
    }
    /* cyclic_reorder so that the collocation vertex comes first (Triplet.hpp:55-72) */
    v3 P[3] = {A, B, C};
    v3 a = P[self], b = P[(self + 1) % 3], c = P[(self + 2) % 3];
    interp_t I = make_interp(a, b, c);
    double g[3], h[3] = {0, 0, 0};
    identical(&I, g);
    if (mirror) {  /* integrate_identical_coloc_mir, Integrator.cpp:44-89 */
        v3 am = a, bm = b, cm = c;
        am.x = -am.x; bm.x = -bm.x; cm.x = -cm.x;
        interp_t Im = make_interp(am, bm, cm);
        double gd[3], hd[3];
        disjoint(a, &Im, 0, gd, hd);
        for (int k = 0; k < 3; ++k) { h[k] = -hd[k]; g[k] = gd[k] + g[k]; }
    }
    /* shift back: reordered vertex k is original vertex (k + self) % 3 */
    for (int k = 0; k < 3; ++k) {
        g3[(k + self) % 3] = g[k];
        h3[(k + self) % 3] = h[k];
    }
}

void bo_panel(const double* x, const double* a, const double* b, const double* c,
              int self_vertex, int mirror, double* g3, double* h3) {
    init_quadrature();
    coloc_local(V(x), V(a), V(b), V(c), self_vertex, mirror, g3, h3);
}

/* ColocSim::assemble_matrices (ColocSim.cpp:21-110) */
/* integrate_disjoint_coloc (Integrator.hpp:471-491) under the 19-point rule */
static void disjoint19(v3 x, const interp_t* I, double* g3, double* h3) {
    double G[3] = {0, 0, 0}, H[3] = {0, 0, 0};
    for (int q = 0; q < 19; ++q) {
        const double u = t19x[q] + t19y[q], v = t19y[q];
        v3 y = interp_at(I, u, v);
        double g, h;
        kernel_pair(vsub(y, x), I->n, &g, &h);
        const double phi[3] = {1.0 - u, u - v, v};
        for (int k = 0; k < 3; ++k) {
            G[k] += phi[k] * g * t19w[q];
            H[k] += phi[k] * h * t19w[q];
        }
    }
    for (int k = 0; k < 3; ++k) { g3[k] = G[k] * I->area; h3[k] = H[k] * I->area; }
}

void bo_exterior_panel(const double* y, const double* a, const double* b, const double* c, double* g3, double* h3) {
    init_q19();
    interp_t I = make_interp(V(a), V(b), V(c));
    disjoint19(V(y), &I, g3, h3);
}

/* compute_exterior_pot (LinLinSim.cpp:440-477) with
 * Integrator::get_exterior_potential (Integrator.cpp:287-305): per target a
 * sum over triangles of (-sum G_k psi_k + sum H_k phi_k) / (4 pi) */
void bo_exterior_pot(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, const double* phi,
                     const double* psi, int64_t P, const double* pts, double* out, int nthreads) {
    (void)N;
    init_q19();
    if (nthreads > 0) omp_set_num_threads(nthreads);
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < P; ++i) {
        const v3 y = V(pts + 3 * i);
        double val = 0.0;
        for (int64_t t = 0; t < M; ++t) {
            const uint32_t ia = tri[3 * t], ib = tri[3 * t + 1], ic = tri[3 * t + 2];
            interp_t I = make_interp(V(xyz + 3 * (size_t)ia), V(xyz + 3 * (size_t)ib), V(xyz + 3 * (size_t)ic));
            double g[3], h[3];
            disjoint19(y, &I, g, h);
            double integral = -g[0] * psi[ia] - g[1] * psi[ib] - g[2] * psi[ic]
                              + h[0] * phi[ia] + h[1] * phi[ib] + h[2] * phi[ic];
            integral *= 1.0 / (4.0 * M_PI);
            val += integral;
        }
        out[i] = val;
    }
}

void bo_assemble(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, int mirror,
                 double* G, double* H, int nthreads) {
    init_quadrature();
    memset(G, 0, sizeof(double) * (size_t)N * (size_t)N);
    memset(H, 0, sizeof(double) * (size_t)N * (size_t)N);
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
    {
        double* Gl = (double*)malloc(sizeof(double) * 3 * (size_t)N);
        double* Hl = (double*)malloc(sizeof(double) * 3 * (size_t)N);
#pragma omp for schedule(static)
        for (int64_t j = 0; j < M; ++j) {
            memset(Gl, 0, sizeof(double) * 3 * (size_t)N);
            memset(Hl, 0, sizeof(double) * 3 * (size_t)N);
            const uint32_t t0 = tri[3 * j], t1 = tri[3 * j + 1], t2 = tri[3 * j + 2];
            const v3 A = V(xyz + 3 * (size_t)t0), B = V(xyz + 3 * (size_t)t1), C = V(xyz + 3 * (size_t)t2);
            for (int64_t i = 0; i < N; ++i) {
                int self = -1;
                if ((uint32_t)i == t0) self = 0;
                else if ((uint32_t)i == t1) self = 1;
                else if ((uint32_t)i == t2) self = 2;
                double g3[3], h3[3];
                coloc_local(V(xyz + 3 * i), A, B, C, self, mirror, g3, h3);
                for (int k = 0; k < 3; ++k) { Gl[k * N + i] += g3[k]; Hl[k * N + i] += h3[k]; }
            }
#pragma omp critical
            {
                const uint32_t tk[3] = {t0, t1, t2};
                for (int k = 0; k < 3; ++k) {
                    double* gc = G + (size_t)tk[k] * N;
                    double* hc = H + (size_t)tk[k] * N;
                    for (int64_t i = 0; i < N; ++i) { gc[i] += Gl[k * N + i]; hc[i] += Hl[k * N + i]; }
                }
            }
        }
        free(Gl); free(Hl);
    }
    /* '4-pi rule' (ColocSim.cpp:85-102): every row of H sums to -4 pi */
    for (int64_t i = 0; i < N; ++i) {
        double val = 0.0;
        for (int64_t j = 0; j < N; ++j) val -= H[(size_t)j * N + i];
        H[(size_t)i * N + i] -= (4.0 * M_PI - val);
    }
}

/* ------------------------------------------------------------------------ */
/* dense linear algebra                                                      */
/* ------------------------------------------------------------------------ */

void bo_gemv(int64_t N, const double* A, const double* x, double* y, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const int64_t RB = 512;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t r0 = 0; r0 < N; r0 += RB) {
        const int64_t r1 = r0 + RB < N ? r0 + RB : N;
        double acc[512];
        for (int64_t i = r0; i < r1; ++i) acc[i - r0] = 0.0;
        for (int64_t j = 0; j < N; ++j) {
            const double xj = x[j];
            const double* col = A + (size_t)j * N;
            for (int64_t i = r0; i < r1; ++i) acc[i - r0] += col[i] * xj;
        }
        for (int64_t i = r0; i < r1; ++i) y[i] = acc[i - r0];
    }
}

int64_t bo_lu_factor(int64_t N, double* A, int64_t* piv, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const int64_t NB = 64;
    int64_t info = 0;
#define A_(i, j) A[(size_t)(j) * N + (i)]
    for (int64_t k0 = 0; k0 < N; k0 += NB) {
        const int64_t kb = (N - k0 < NB) ? N - k0 : NB;
        const int64_t k1 = k0 + kb;
        /* panel: unblocked GEPP, first maximum wins ties (PartialPivLU) */
        for (int64_t k = k0; k < k1; ++k) {
            int64_t p = k;
            double best = fabs(A_(k, k));
            for (int64_t i = k + 1; i < N; ++i) {
                const double v = fabs(A_(i, k));
                if (v > best) { best = v; p = i; }
            }
            piv[k] = p;
            if (p != k) {
                for (int64_t j = 0; j < N; ++j) {
                    const double t = A_(k, j); A_(k, j) = A_(p, j); A_(p, j) = t;
                }
            }
            const double d = A_(k, k);
            if (d == 0.0) { if (!info) info = k + 1; continue; }
            const double rd = 1.0 / d;
            for (int64_t i = k + 1; i < N; ++i) A_(i, k) *= rd;
            for (int64_t j = k + 1; j < k1; ++j) {
                const double u = A_(k, j);
                if (u == 0.0) continue;
                for (int64_t i = k + 1; i < N; ++i) A_(i, j) -= A_(i, k) * u;
            }
        }
        if (k1 >= N) break;
        /* U12 = L11^-1 A12 and A22 -= L21 U12, one column of the trailing part per task */
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t j = k1; j < N; ++j) {
            for (int64_t k = k0; k < k1; ++k) {
                const double u = A_(k, j);
                if (u == 0.0) continue;
                for (int64_t i = k + 1; i < k1; ++i) A_(i, j) -= A_(i, k) * u;
            }
            for (int64_t k = k0; k < k1; ++k) {
                const double u = A_(k, j);
                if (u == 0.0) continue;
                const double* lk = &A_(0, k);
                double* aj = &A_(0, j);
                for (int64_t i = k1; i < N; ++i) aj[i] -= lk[i] * u;
            }
        }
    }
#undef A_
    return info;
}

void bo_lu_solve(int64_t N, const double* LU, const int64_t* piv, double* b) {
    for (int64_t k = 0; k < N; ++k) {
        const int64_t p = piv[k];
        if (p != k) { const double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int64_t j = 0; j < N; ++j) {           /* unit lower, column sweep */
        const double bj = b[j];
        const double* col = LU + (size_t)j * N;
        for (int64_t i = j + 1; i < N; ++i) b[i] -= col[i] * bj;
    }
    for (int64_t j = N - 1; j >= 0; --j) {      /* upper, column sweep */
        const double* col = LU + (size_t)j * N;
        b[j] /= col[j];
        const double bj = b[j];
        for (int64_t i = 0; i < j; ++i) b[i] -= col[i] * bj;
    }
}

static double dotv(int64_t n, const double* a, const double* b) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Eigen 3.4 bicgstab() (IterativeLinearSolvers/BiCGSTAB.h) with the
 * DiagonalPreconditioner and a zero initial guess. */
int64_t bo_bicgstab(int64_t N, const double* A, const double* rhs, double* x,
                    double tol, int64_t maxit, double* err, int nthreads) {
    if (tol <= 0.0) tol = 2.220446049250313e-16;
    if (maxit <= 0) maxit = 2 * N;
    double* inv = (double*)malloc(sizeof(double) * N);
    double* r = (double*)malloc(sizeof(double) * N);
    double* r0 = (double*)malloc(sizeof(double) * N);
    double* v = (double*)calloc(N, sizeof(double));
    double* p = (double*)calloc(N, sizeof(double));
    double* y = (double*)malloc(sizeof(double) * N);
    double* z = (double*)malloc(sizeof(double) * N);
    double* s = (double*)malloc(sizeof(double) * N);
    double* t = (double*)malloc(sizeof(double) * N);
    for (int64_t i = 0; i < N; ++i) {
        const double d = A[(size_t)i * N + i];
        inv[i] = (d != 0.0) ? 1.0 / d : 1.0;
        x[i] = 0.0;
    }
    bo_gemv(N, A, x, r, nthreads);
    for (int64_t i = 0; i < N; ++i) { r[i] = rhs[i] - r[i]; r0[i] = r[i]; }
    double r0_sq = dotv(N, r0, r0);
    const double rhs_sq = dotv(N, rhs, rhs);
    int64_t it = 0;
    if (rhs_sq == 0.0) { *err = 0.0; goto done; }
    {
        double rho = 1.0, alpha = 1.0, w = 1.0;
        const double tol2 = tol * tol * rhs_sq;
        const double eps2 = 2.220446049250313e-16 * 2.220446049250313e-16;
        int64_t restarts = 0;
        while (dotv(N, r, r) > tol2 && it < maxit) {
            const double rho_old = rho;
            rho = dotv(N, r0, r);
            if (fabs(rho) < eps2 * r0_sq) {
                bo_gemv(N, A, x, r, nthreads);
                for (int64_t i = 0; i < N; ++i) { r[i] = rhs[i] - r[i]; r0[i] = r[i]; }
                rho = r0_sq = dotv(N, r, r);
                if (restarts++ == 0) it = 0;
            }
            const double beta = (rho / rho_old) * (alpha / w);
            for (int64_t i = 0; i < N; ++i) p[i] = r[i] + beta * (p[i] - w * v[i]);
            for (int64_t i = 0; i < N; ++i) y[i] = inv[i] * p[i];
            bo_gemv(N, A, y, v, nthreads);
            alpha = rho / dotv(N, r0, v);
            for (int64_t i = 0; i < N; ++i) s[i] = r[i] - alpha * v[i];
            for (int64_t i = 0; i < N; ++i) z[i] = inv[i] * s[i];
            bo_gemv(N, A, z, t, nthreads);
            const double tmp = dotv(N, t, t);
            w = (tmp > 0.0) ? dotv(N, t, s) / tmp : 0.0;
            for (int64_t i = 0; i < N; ++i) x[i] += alpha * y[i] + w * z[i];
            for (int64_t i = 0; i < N; ++i) r[i] = s[i] - w * t[i];
            ++it;
        }
        *err = sqrt(dotv(N, r, r) / rhs_sq);
    }
done:
    free(inv); free(r); free(r0); free(v); free(p); free(y); free(z); free(s); free(t);
    return it;
}

/* ------------------------------------------------------------------------ */
/* mesh geometry (Mesh.cpp, LinLinSim.cpp)                                   */
/* ------------------------------------------------------------------------ */

/* generate_triangle_indices (Mesh.cpp:75-84) as CSR in ascending triangle order */
static void incidence(int64_t N, int64_t M, const uint32_t* tri, int64_t** ptr_out, int64_t** idx_out) {
    int64_t* ptr = (int64_t*)calloc(N + 1, sizeof(int64_t));
    for (int64_t j = 0; j < 3 * M; ++j) ptr[tri[j] + 1]++;
    for (int64_t i = 0; i < N; ++i) ptr[i + 1] += ptr[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * N);
    for (int64_t i = 0; i < N; ++i) fill[i] = ptr[i];
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (3 * M + 1));
    for (int64_t j = 0; j < M; ++j)
        for (int k = 0; k < 3; ++k) idx[fill[tri[3 * j + k]]++] = j;
    free(fill);
    *ptr_out = ptr; *idx_out = idx;
}

/* generate_triangle_normals (Mesh.cpp:130-138) */
void bo_triangle_normals(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, double* n) {
    (void)N;
    for (int64_t j = 0; j < M; ++j) {
        v3 a = V(xyz + 3 * (size_t)tri[3 * j]), b = V(xyz + 3 * (size_t)tri[3 * j + 1]),
           c = V(xyz + 3 * (size_t)tri[3 * j + 2]);
        v3 nn = vcross(vsub(b, a), vsub(c, a));
        nn = vscale(1.0 / vnorm(nn), nn);
        n[3 * j] = nn.x; n[3 * j + 1] = nn.y; n[3 * j + 2] = nn.z;
    }
}

/* generate_vertex_normals, Max 1999 weights (Mesh.cpp:141-163) */
void bo_vertex_normals(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, double* n) {
    int64_t *ptr, *idx;
    incidence(N, M, tri, &ptr, &idx);
    for (int64_t i = 0; i < N; ++i) {
        v3 acc = {0, 0, 0};
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
            const uint32_t* t = tri + 3 * idx[e];
            int s = (t[0] == (uint32_t)i) ? 0 : (t[1] == (uint32_t)i) ? 1 : 2;
            v3 a = V(xyz + 3 * (size_t)t[s]), b = V(xyz + 3 * (size_t)t[(s + 1) % 3]),
               c = V(xyz + 3 * (size_t)t[(s + 2) % 3]);
            v3 Bv = vsub(b, a), Cv = vsub(c, a);
            acc = vadd(acc, vscale(1.0 / (vdot(Bv, Bv) * vdot(Cv, Cv)), vcross(Bv, Cv)));
        }
        acc = vscale(1.0 / vnorm(acc), acc);
        n[3 * i] = acc.x; n[3 * i + 1] = acc.y; n[3 * i + 2] = acc.z;
    }
    free(ptr); free(idx);
}

/* generate_tangent_gradients (LinLinSim.cpp:288-305, Dombre 2019) */
void bo_tangent_gradients(int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                          const double* pot, double* grad) {
    (void)N;
    for (int64_t j = 0; j < M; ++j) {
        const uint32_t ia = tri[3 * j], ib = tri[3 * j + 1], ic = tri[3 * j + 2];
        v3 a = V(xyz + 3 * (size_t)ia), b = V(xyz + 3 * (size_t)ib), c = V(xyz + 3 * (size_t)ic);
        v3 ab = vsub(b, a), bc = vsub(c, b);
        v3 n = vcross(ab, bc);
        n = vscale(1.0 / vdot(n, n), n);
        v3 g = vadd(vscale(pot[ic] - pot[ib], vcross(n, ab)), vscale(pot[ia] - pot[ib], vcross(n, bc)));
        grad[3 * j] = g.x; grad[3 * j + 1] = g.y; grad[3 * j + 2] = g.z;
    }
}

/* Eigen::FullPivLU<Matrix<double,6,3>>::compute + solve: complete pivoting
 * (first maximum in column-major order), the solution of the three pivot
 * equations is returned. */
static void fullpiv_solve_6x3(double Mx[6][3], double bv[6], double x[3]) {
    int rowt[3], colt[3], perm[3] = {0, 1, 2};
    int rank = 3;
    double maxpivot = 0.0;
    for (int k = 0; k < 3; ++k) {
        int br = k, bc = k;
        double best = -1.0;
        for (int j = k; j < 3; ++j)
            for (int i = k; i < 6; ++i) {
                const double v = fabs(Mx[i][j]);
                if (v > best) { best = v; br = i; bc = j; }
            }
        if (best == 0.0) {
            rank = k;
            for (int i = k; i < 3; ++i) { rowt[i] = i; colt[i] = i; }
            break;
        }
        if (best > maxpivot) maxpivot = best;
        rowt[k] = br; colt[k] = bc;
        if (br != k)
            for (int j = 0; j < 3; ++j) { double t = Mx[k][j]; Mx[k][j] = Mx[br][j]; Mx[br][j] = t; }
        if (bc != k)
            for (int i = 0; i < 6; ++i) { double t = Mx[i][k]; Mx[i][k] = Mx[i][bc]; Mx[i][bc] = t; }
        for (int i = k + 1; i < 6; ++i) Mx[i][k] /= Mx[k][k];
        if (k < 2)
            for (int j = k + 1; j < 3; ++j)
                for (int i = k + 1; i < 6; ++i) Mx[i][j] -= Mx[i][k] * Mx[k][j];
    }
    /* rank with Eigen's default threshold (epsilon * diagonal size) */
    int nz = 0;
    for (int k = 0; k < rank; ++k)
        if (fabs(Mx[k][k]) > 2.220446049250313e-16 * 3.0 * maxpivot) ++nz;
    for (int k = 0; k < 3; ++k) {
        const int t = perm[k]; perm[k] = perm[colt[k]]; perm[colt[k]] = t;
    }
    double c[6];
    for (int i = 0; i < 6; ++i) c[i] = bv[i];
    for (int k = 0; k < 3; ++k) { const double t = c[k]; c[k] = c[rowt[k]]; c[rowt[k]] = t; }
    for (int i = 1; i < 3; ++i)
        for (int j = 0; j < i; ++j) c[i] -= Mx[i][j] * c[j];
    for (int i = nz - 1; i >= 0; --i) {
        for (int j = i + 1; j < nz; ++j) c[i] -= Mx[i][j] * c[j];
        c[i] /= Mx[i][i];
    }
    for (int i = 0; i < 3; ++i) x[i] = 0.0;
    for (int i = 0; i < nz; ++i) x[perm[i]] = c[i];
}

/* curvatures (Mesh.cpp:256-319, Rusinkiewicz 2004) */
void bo_curvatures(int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                   double* kappa_tri, double* gauss_tri) {
    double* vn = (double*)malloc(sizeof(double) * 3 * N);
    bo_vertex_normals(N, xyz, M, tri, vn);
    for (int64_t j = 0; j < M; ++j) {
        const uint32_t ia = tri[3 * j], ib = tri[3 * j + 1], ic = tri[3 * j + 2];
        v3 a = V(xyz + 3 * (size_t)ia), b = V(xyz + 3 * (size_t)ib), c = V(xyz + 3 * (size_t)ic);
        v3 e0 = vsub(b, a), e1 = vsub(c, b), e2 = vsub(a, c);
        v3 u = vscale(1.0 / vnorm(e0), e0);
        v3 nr = vcross(e0, e1);
        nr = vscale(1.0 / vnorm(nr), nr);
        v3 w = vcross(nr, u);
        double Mx[6][3] = {
            {vdot(e0, u), vdot(e0, w), 0.0}, {0.0, vdot(e0, u), vdot(e0, w)},
            {vdot(e1, u), vdot(e1, w), 0.0}, {0.0, vdot(e1, u), vdot(e1, w)},
            {vdot(e2, u), vdot(e2, w), 0.0}, {0.0, vdot(e2, u), vdot(e2, w)}};
        v3 na = V(vn + 3 * (size_t)ia), nb = V(vn + 3 * (size_t)ib), nc = V(vn + 3 * (size_t)ic);
        double bv[6] = {vdot(vsub(nb, na), u), vdot(vsub(nb, na), w), vdot(vsub(nc, nb), u),
                        vdot(vsub(nc, nb), w), vdot(vsub(na, nc), u), vdot(vsub(na, nc), w)};
        double x[3];
        fullpiv_solve_6x3(Mx, bv, x);
        kappa_tri[j] = 0.5 * (x[0] + x[2]);
        if (gauss_tri) gauss_tri[j] = x[0] * x[2] - x[1] * x[1];
    }
    free(vn);
}

/* LinLinSim::kappa (LinLinSim.cpp:122-146) */
void bo_kappa(int64_t N, const double* xyz, int64_t M, const uint32_t* tri, double* kv) {
    double* kt = (double*)malloc(sizeof(double) * M);
    double* w = (double*)calloc(N, sizeof(double));
    bo_curvatures(N, xyz, M, tri, kt, NULL);
    for (int64_t i = 0; i < N; ++i) kv[i] = 0.0;
    for (int64_t j = 0; j < M; ++j)
        for (int k = 0; k < 3; ++k) { kv[tri[3 * j + k]] += kt[j]; w[tri[3 * j + k]] += 1.0; }
    for (int64_t i = 0; i < N; ++i) kv[i] /= w[i];
    free(kt); free(w);
}

static int64_t uf_find(int64_t* p, int64_t a) {
    while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; }
    return a;
}

int64_t bo_components(int64_t N, int64_t M, const uint32_t* tri, int64_t* comp) {
    int64_t* p = (int64_t*)malloc(sizeof(int64_t) * N);
    for (int64_t i = 0; i < N; ++i) p[i] = i;
    for (int64_t j = 0; j < M; ++j)
        for (int k = 1; k < 3; ++k) {
            int64_t a = uf_find(p, tri[3 * j]), b = uf_find(p, tri[3 * j + k]);
            if (a != b) { if (a < b) p[b] = a; else p[a] = b; }
        }
    int64_t* label = (int64_t*)malloc(sizeof(int64_t) * N);
    for (int64_t i = 0; i < N; ++i) label[i] = -1;
    int64_t n = 0;
    for (int64_t i = 0; i < N; ++i) {
        int64_t r = uf_find(p, i);
        if (label[r] < 0) label[r] = n++;
        comp[i] = label[r];
    }
    free(p); free(label);
    return n;
}

/* volume (Mesh.cpp:206-218) per loose part, triangles in ascending order */
void bo_part_volumes(int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                     const int64_t* comp, int64_t nparts, double* vol) {
    (void)N;
    for (int64_t c = 0; c < nparts; ++c) vol[c] = 0.0;
    for (int64_t j = 0; j < M; ++j) {
        v3 a = V(xyz + 3 * (size_t)tri[3 * j]), b = V(xyz + 3 * (size_t)tri[3 * j + 1]),
           c = V(xyz + 3 * (size_t)tri[3 * j + 2]);
        vol[comp[tri[3 * j]]] += vdot(vcross(a, b), c) / 6.0;
    }
}

/* ------------------------------------------------------------------------ */
/* one stage and the time steppers                                           */
/* ------------------------------------------------------------------------ */

static void solve_dense(const bo_params* P, int64_t N, double* G, const double* b, double* x) {
    if (P->solver == 1) {
        double err;
        bo_bicgstab(N, G, b, x, 0.0, 0, &err, P->nthreads);
    } else {
        int64_t* piv = (int64_t*)malloc(sizeof(int64_t) * N);
        bo_lu_factor(N, G, piv, P->nthreads);
        for (int64_t i = 0; i < N; ++i) x[i] = b[i];
        bo_lu_solve(N, G, piv, x);
        free(piv);
    }
}

static void psi_of(const bo_params* P, int64_t N, const double* xyz, int64_t M,
                   const uint32_t* tri, const double* pot, double* psi) {
    double* G = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* H = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* b = (double*)malloc(sizeof(double) * N);
    bo_assemble(N, xyz, M, tri, P->mirror, G, H, P->nthreads);
    /* ColocSimPin::assemble_matrices (ColocSimPin.cpp:113-117): the columns of
     * the pinned vertices trade places, G(:, j) = -H(:, j), H(:, j) = -G(:, j)
     * (both column-major) */
    for (int64_t j = N - P->n_pin; j < N; ++j) {
        double* g = G + (size_t)j * N;
        double* h = H + (size_t)j * N;
        for (int64_t i = 0; i < N; ++i) {
            const double t = g[i];
            g[i] = -h[i];
            h[i] = -t;
        }
    }
    bo_gemv(N, H, pot, b, P->nthreads);  /* H * phi (LinLinSim.cpp:32) */
    free(H);
    solve_dense(P, N, G, b, psi);
    free(G); free(b);
}

void bo_compute_psi(const bo_params* P, int64_t N, const double* xyz, int64_t M,
                    const uint32_t* tri, const double* phi, double* psi) {
    psi_of(P, N, xyz, M, tri, phi, psi);
}

int64_t bo_pin_boundary(int64_t N, double* xyz, int64_t M, uint32_t* tri, double* phi, double* psi) {
    int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * N);
    int64_t* inv = (int64_t*)malloc(sizeof(int64_t) * N);
    for (int64_t i = 0; i < N; ++i) perm[i] = i;
    int64_t k = N, npin = 0;
    for (int64_t i = 0; i < k; ++i) {
        if (fabs(xyz[3 * i]) < 1e-5) {
            xyz[3 * i] = 0.0;
            const int64_t l = k - 1;
            int64_t tp = perm[i]; perm[i] = perm[l]; perm[l] = tp;
            for (int d = 0; d < 3; ++d) { double t = xyz[3 * i + d]; xyz[3 * i + d] = xyz[3 * l + d]; xyz[3 * l + d] = t; }
            if (phi) { double t = phi[i]; phi[i] = phi[l]; phi[l] = t; }
            if (psi) { double t = psi[i]; psi[i] = psi[l]; psi[l] = t; }
            --i;    /* look at the vertex swapped in */
            --k;
            ++npin;
        }
    }
    for (int64_t i = 0; i < N; ++i) inv[perm[i]] = i;
    for (int64_t e = 0; e < 3 * M; ++e) tri[e] = (uint32_t)inv[tri[e]];
    free(perm); free(inv);
    return npin;
}

double bo_drift_velocity(double c) {
    const double beta = 0.2, c_eq = cos(130.0 / 180.0 * M_PI), c_nl = 0.01;
    return beta * (c_eq - c) + c_nl * (c * c * c) / (c * c - 1.0);
}

/* LinLinSim::position_t (LinLinSim.cpp:26-83) and, with n_pin > 0,
 * ColocSimPin::position_t (ColocSimPin.cpp:139-194): pot's pinned entries
 * take the state's psi for the solve and the solved potential afterwards
 * (in place, as in the reference); the pinned vertices move with the drift
 * law along the wall. */
static void position_impl(const bo_params* P, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                          double* pot, const double* psi_state, double* vel, double* psi_out) {
    double* psi = (double*)malloc(sizeof(double) * N);
    double* tn = (double*)malloc(sizeof(double) * 3 * M);
    double* tg = (double*)malloc(sizeof(double) * 3 * M);
    bo_triangle_normals(N, xyz, M, tri, tn);
    bo_tangent_gradients(N, xyz, M, tri, pot, tg);
    for (int64_t i = N - P->n_pin; i < N; ++i) pot[i] = psi_state[i];
    psi_of(P, N, xyz, M, tri, pot, psi);
    int64_t *ptr, *idx;
    incidence(N, M, tri, &ptr, &idx);
    for (int64_t i = 0; i < N; ++i) {   /* LinLinSim.cpp:48-59 */
        double num = 0.0;
        v3 g = {0, 0, 0};
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
            const int64_t t = idx[e];
            g = vadd(g, vadd(V(tg + 3 * t), vscale(psi[i], V(tn + 3 * t))));
            num += 1.0;
        }
        g = vscale(1.0 / num, g);
        vel[3 * i] = g.x; vel[3 * i + 1] = g.y; vel[3 * i + 2] = g.z;
    }
    if (P->n_pin > 0) {
        double* vn = (double*)malloc(sizeof(double) * 3 * N);
        bo_vertex_normals(N, xyz, M, tri, vn);
        for (int64_t i = N - P->n_pin; i < N; ++i) {
            v3 n = V(vn + 3 * i);
            const double c = n.x;
            n.x = 0.0;
            n = vscale(1.0 / vnorm(n), n);      /* tangential to the wall */
            n = vscale(bo_drift_velocity(c), n);
            vel[3 * i] = n.x; vel[3 * i + 1] = n.y; vel[3 * i + 2] = n.z;
            pot[i] = psi[i];
        }
        free(vn);
    }
    if (psi_out) for (int64_t i = 0; i < N; ++i) psi_out[i] = psi[i];
    free(psi); free(tn); free(tg); free(ptr); free(idx);
}

void bo_position_t(const bo_params* P, int64_t N, const double* xyz, int64_t M,
                   const uint32_t* tri, const double* pot, double* vel, double* psi_out) {
    bo_params Q = *P;
    Q.n_pin = 0;
    double* p = (double*)malloc(sizeof(double) * N);
    memcpy(p, pot, sizeof(double) * N);
    position_impl(&Q, N, xyz, M, tri, p, NULL, vel, psi_out);
    free(p);
}

void bo_pot_t_multi(const bo_params* P, int64_t N, const double* xyz, int64_t M,
                    const uint32_t* tri, const double* vel, double t, double* dpot) {
    int64_t* comp = (int64_t*)malloc(sizeof(int64_t) * N);
    const int64_t np = bo_components(N, M, tri, comp);
    double* vol = (double*)malloc(sizeof(double) * np);
    bo_part_volumes(N, xyz, M, tri, comp, np, vol);
    double* kap = (double*)malloc(sizeof(double) * N);
    bo_kappa(N, xyz, M, tri, kap);
    const double field = P->field_amp * sin(2.0 * M_PI * P->field_freq * t + P->field_phase);
    for (int64_t i = 0; i < N; ++i) {   /* potential_t, Simulation.cpp:73-75 */
        const v3 v = V(vel + 3 * i);
        dpot[i] = 2.0 * P->sigma * kap[i] + 0.5 * vdot(v, v) + P->p_inf -
                  P->epsilon * pow(P->V_0 / vol[comp[i]], P->gamma) + field;
    }
    free(comp); free(vol); free(kap);
}

/* Simulation::get_dt (Simulation.cpp:77-92) */
double bo_get_dt(const bo_params* P, int64_t N, double dp, const double* vel, const double* dpot) {
    double max_pot = 0.0, max_vel = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        max_pot = fmax(fabs(dpot[i]), max_pot);
        const v3 v = V(vel + 3 * i);
        max_vel = fmax(vdot(v, v), max_vel);
    }
#if BO_MUTANT != 4
    max_vel = sqrt(max_vel);
#endif
    double dt = dp / (max_pot + P->dp_balance * max_vel);
    if (P->min_dt > 0.0) dt = fmin(dt, P->min_dt);
    return dt;
}

/* LinLinSim::nopenetration (LinLinSim.cpp:162-190): xf = x0 + dt c, except
 * that a vertex moving towards the wall x = 0 (c.x dt < 0) inside the band
 * x < eps, or entering it within the step, follows an exponential approach
 * instead; its c.x is rewritten to the effective velocity (xf.x - x0.x)/dt. */
void bo_nopenetration(int64_t N, double eps, double dt, const double* x0, double* c, double* xf) {
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < 3; ++d) xf[3 * i + d] = dt * c[3 * i + d] + x0[3 * i + d];
        const double cx = c[3 * i];
        if (!(cx * dt < 0.0)) continue;
        int flag = 0;
        const double xpos = x0[3 * i];
        if (xpos < eps) {
            xf[3 * i] = xpos * exp(cx / xpos * dt);
            flag = 1;
        } else {
            const double te = (eps - xpos) / cx;
            if (fabs(te) < fabs(dt)) {
#if BO_MUTANT == 5
                xf[3 * i] = eps * exp(cx / eps * dt);
#else
                xf[3 * i] = eps * exp(cx / eps * (dt - te));
#endif
                flag = 1;
            }
        }
        if (flag) c[3 * i] = (xf[3 * i] - xpos) / dt;
    }
}

static int any_negative_x(int64_t N, const double* xyz) {  /* test_negative, LinLinSim.cpp:148-159 */
    for (int64_t i = 0; i < N; ++i) if (xyz[3 * i] < 0.0) return 1;
    return 0;
}

int bo_rk4_step(const bo_params* P, int64_t N, double* xyz, int64_t M, const uint32_t* tri,
                double* phi, double* psi, double* time, double dp, int fixdt, double* dt_out) {
    if (P->wall_guards && any_negative_x(N, xyz)) return -1;
    const size_t n3 = 3 * (size_t)N;
    double* x1 = (double*)malloc(sizeof(double) * n3);
    double* xs = (double*)malloc(sizeof(double) * n3);
    double* kx[4];
    double* kp[4];
    double* ps = (double*)malloc(sizeof(double) * N);
    for (int s = 0; s < 4; ++s) {
        kx[s] = (double*)malloc(sizeof(double) * n3);
        kp[s] = (double*)malloc(sizeof(double) * N);
    }
    memcpy(x1, xyz, sizeof(double) * n3);
    const double t0 = *time;
    double dt;
    /* stage 1 */
    /* phi is the stage potential p1, whose pinned entries position_t rewrites */
    position_impl(P, N, x1, M, tri, phi, psi, kx[0], NULL);
    bo_pot_t_multi(P, N, x1, M, tri, kx[0], t0, kp[0]);
    dt = fixdt ? dp : bo_get_dt(P, N, dp, kx[0], kp[0]);
    const double ts[4] = {t0, P->frozen_stage_time ? t0 : t0 + 0.5 * dt,
                          P->frozen_stage_time ? t0 : t0 + 0.5 * dt, P->frozen_stage_time ? t0 : t0 + dt};
    const double cs[3] = {0.5 * dt, 0.5 * dt, dt};
    for (int s = 1; s < 4; ++s) {   /* LinLinSim.cpp:218-234 */
        for (size_t e = 0; e < n3; ++e) xs[e] = cs[s - 1] * kx[s - 1][e] + x1[e];
        for (int64_t i = 0; i < N; ++i) ps[i] = cs[s - 1] * kp[s - 1][i] + phi[i];
        position_impl(P, N, xs, M, tri, ps, psi, kx[s], NULL);
        bo_pot_t_multi(P, N, xs, M, tri, kx[s], ts[s], kp[s]);
    }
    /* average = (k1 + 2 k2 + 2 k3 + k4) / 6 (LinLinSim.cpp:236) */
    double* avg = xs;
#if BO_MUTANT >= 1 && BO_MUTANT <= 3
    /* planted mistakes for the mutation tests: 1 forward Euler (k1 only),
     * 2 midpoint RK2 (k2 only), 3 equal stage weights (k1+k2+k3+k4)/4 */
    static const double wm[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0.25, 0.25, 0.25, 0.25}};
    const double* w = wm[BO_MUTANT - 1];
    for (size_t e = 0; e < n3; ++e)
        avg[e] = w[0] * kx[0][e] + w[1] * kx[1][e] + w[2] * kx[2][e] + w[3] * kx[3][e];
    for (int64_t i = 0; i < N; ++i)
        phi[i] = dt * (w[0] * kp[0][i] + w[1] * kp[1][i] + w[2] * kp[2][i] + w[3] * kp[3][i]) + phi[i];
#else
    for (size_t e = 0; e < n3; ++e)
        avg[e] = (((kx[0][e] + 2.0 * kx[1][e]) + 2.0 * kx[2][e]) + kx[3][e]) * (1.0 / 6.0);
    /* pf (LinLinSim.cpp:238) */
    for (int64_t i = 0; i < N; ++i)
        phi[i] = (dt / 6.0) * (((kp[0][i] + 2.0 * kp[1][i]) + 2.0 * kp[2][i]) + kp[3][i]) + phi[i];
#endif
    /* nopenetration (LinLinSim.cpp:240), applied with the wall guards only */
    if (P->wall_guards) {
        bo_nopenetration(N, P->wall_eps, dt, x1, avg, xyz);
    } else {
        for (size_t e = 0; e < n3; ++e) xyz[e] = dt * avg[e] + x1[e];
    }
    /* psi = average . vertex normal of the new mesh (LinLinSim.cpp:243-249) */
    double* vn = kx[0];
    bo_vertex_normals(N, xyz, M, tri, vn);
    for (int64_t i = 0; i < N; ++i) psi[i] = vdot(V(avg + 3 * i), V(vn + 3 * i));
    *time = t0 + dt;
    if (dt_out) *dt_out = dt;
    free(x1); free(xs); free(ps);
    for (int s = 0; s < 4; ++s) { free(kx[s]); free(kp[s]); }
    if (P->wall_guards && any_negative_x(N, xyz)) return -1;
    return 0;
}

int bo_euler_step(const bo_params* P, int64_t N, double* xyz, int64_t M, const uint32_t* tri,
                  double* phi, const double* psi, double* time, double dp, int fixdt, double* dt_out) {
    const size_t n3 = 3 * (size_t)N;
    double* vel = (double*)malloc(sizeof(double) * n3);
    double* dpot = (double*)malloc(sizeof(double) * N);
    position_impl(P, N, xyz, M, tri, phi, psi, vel, NULL);
    bo_pot_t_multi(P, N, xyz, M, tri, vel, *time, dpot);
    const double dt = fixdt ? dp : bo_get_dt(P, N, dp, vel, dpot);
    for (size_t e = 0; e < n3; ++e) xyz[e] = vel[e] * dt + xyz[e];
    for (int64_t i = 0; i < N; ++i) phi[i] = dt * dpot[i] + phi[i];
    *time += dt;
    if (dt_out) *dt_out = dt;
    free(vel); free(dpot);
    return 0;
}
