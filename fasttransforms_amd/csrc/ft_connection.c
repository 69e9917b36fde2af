/*
 * ft_connection.c -- the 1D connection matrices the harmonic transforms use
 * (src/transforms.c:50-93): Legendre <-> Chebyshev and ultraspherical
 * lambda -> mu, as n x n upper triangular column-major FP64 matrices.
 *
 * The reference's method (banded_source.c): column j of the connection
 * matrix V is an eigenvector of the triangular-banded pencil A V = B V
 * Lambda (A, B bandwidth 2, from the polynomials' differential equations),
 * scaled so that V(j, j) is the ratio of leading coefficients; the other
 * entries come by back substitution, entry (i, j) from entries i+1, i+2
 * (banded_source.c:779-798), all in long double, rounded to double at the
 * end.  Host code, built with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>

#include "ft_internal.h"

/* pi-related constants in long double (as <math.h> gives them under _GNU_SOURCE) */
#ifndef M_PIl
#define M_PIl 3.141592653589793238462643383279502884L
#define M_PI_2l 1.570796326794896619231321691639751442L
#define M_1_PIl 0.318309886183790671537767526745028724L
#define M_2_PIl 0.636619772367581343075535053490057448L
#endif

/* A triangular-banded matrix of bandwidth 2 in long double: entry (i, j),
 * 0 <= j - i <= 2, at d[i + 2 (j + 1)] (banded_source.c:261-275). */
typedef struct {
    int n;
    long double *d;
} tb2;

static int tb_alloc(tb2 *T, int n) {
    T->n = n;
    T->d = calloc((size_t)(3 * (n > 0 ? n : 1)), sizeof(long double));
    return T->d ? FTAMD_OK : FTAMD_ENOMEM;
}
static long double tb_get(const tb2 *T, int i, int j) {
    return (i >= 0 && j >= i && j - i <= 2 && j < T->n) ? T->d[i + 2 * (j + 1)] : 0.0L;
}
static void tb_set(tb2 *T, long double v, int i, int j) {
    if (i >= 0 && j >= i && j - i <= 2 && j < T->n) T->d[i + 2 * (j + 1)] = v;
}

/* Back substitution for the eigenvectors (banded_source.c:779-798): V holds
 * the diagonal on entry. */
static void tb_eigenvectors(const tb2 *A, const tb2 *B, long double *V) {
    const int n = A->n;
    const long double eps = 0x1p-64L;
    for (int j = 1; j < n; j++) {
        const long double lam = tb_get(A, j, j) / tb_get(B, j, j);
        for (int i = j - 1; i >= 0; i--) {
            long double t = 0, kt = 0;
            for (int k = i + 1; k < (i + 3 < n ? i + 3 : n); k++) {
                t += (tb_get(A, i, k) - lam * tb_get(B, i, k)) * V[k + (size_t)j * n];
                kt += (fabsl(tb_get(A, i, k)) + fabsl(lam * tb_get(B, i, k))) * fabsl(V[k + (size_t)j * n]);
            }
            const long double d = lam * tb_get(B, i, i) - tb_get(A, i, i);
            const long double kd = fabsl(lam * tb_get(B, i, i)) + fabsl(tb_get(A, i, i));
            V[i + (size_t)j * n] = (fabsl(d) < 4 * kd * eps || fabsl(t) < 4 * kt * eps) ? 0.0L : t / d;
        }
    }
}

/* Runs the eigenvector recurrence on (A, B) with the diagonal already in Vl
 * and rounds to a new double matrix. */
static double *finish(tb2 *A, tb2 *B, long double *Vl, int n) {
    double *V = NULL;
    if (A->d && B->d && Vl) {
        tb_eigenvectors(A, B, Vl);
        V = malloc(sizeof(double) * (size_t)n * (size_t)(n > 0 ? n : 1));
        if (V)
            for (size_t i = 0; i < (size_t)n * (size_t)n; i++) V[i] = (double)Vl[i];
    }
    free(A->d);
    free(B->d);
    free(Vl);
    return V;
}

/* transforms.c:50-63 with banded_source.c:2024-2060, 2316-2338 */
double *ftamd_plan_leg2cheb(int normleg, int normcheb, int n) {
    tb2 A, B;
    tb_alloc(&A, n);
    tb_alloc(&B, n);
    long double *V = calloc((size_t)n * (size_t)(n > 0 ? n : 1), sizeof(long double));
    if (A.d && B.d) {
        if (n > 1) tb_set(&A, 2, 1, 1);
        for (int i = 2; i < n; i++) {
            tb_set(&A, -i * (i - 1.0L), i - 2, i);
            tb_set(&A, i * (i + 1.0L), i, i);
        }
        if (n > 0) tb_set(&B, normcheb ? sqrtl(2) : 2, 0, 0);
        if (n > 1) tb_set(&B, 1, 1, 1);
        for (int i = 2; i < n; i++) {
            tb_set(&B, -1, i - 2, i);
            tb_set(&B, 1, i, i);
        }
    }
    if (V) {
        const size_t inc = (size_t)n + 1;
        if (normleg) {
            if (n > 0) V[0] = normcheb ? sqrtl(0.5) * tgammal(0.5) : sqrtl(0.5);
            if (n > 1) V[inc] = normcheb ? sqrtl(1.5) * V[0] : sqrtl(1.5);
            for (int i = 2; i < n; i++) V[i * inc] = sqrtl((2 * i + 1) * (2 * i - 1.0L)) * V[(i - 1) * inc] / (2 * i);
        } else {
            if (n > 0) V[0] = normcheb ? tgammal(0.5) : 1;
            if (n > 1) V[inc] = normcheb ? V[0] / sqrtl(2) : 1;
            for (int i = 2; i < n; i++) V[i * inc] = (2 * i - 1) * V[(i - 1) * inc] / (2 * i);
        }
    }
    return finish(&A, &B, V, n);
}

/* transforms.c:65-78 with banded_source.c:2063-2101, 2340-2384 */
double *ftamd_plan_cheb2leg(int normcheb, int normleg, int n) {
    tb2 A, B;
    tb_alloc(&A, n);
    tb_alloc(&B, n);
    long double *V = calloc((size_t)n * (size_t)(n > 0 ? n : 1), sizeof(long double));
    if (A.d && B.d) {
        if (normleg) {
            if (n > 1) tb_set(&A, sqrtl(2.0L / 5), 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&A, -(i + 1) * sqrtl(((i - 1.0L) * i) / ((2 * i - 1.0L) * (2 * i + 1))) * (i + 1), i - 2, i);
                tb_set(&A, i * sqrtl(((i + 1.0L) * (i + 2)) / ((2 * i + 1.0L) * (2 * i + 3))) * i, i, i);
            }
            if (n > 0) tb_set(&B, sqrtl(2.0L / 3), 0, 0);
            if (n > 1) tb_set(&B, sqrtl(2.0L / 5), 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&B, -sqrtl(((i - 1.0L) * i) / ((2 * i - 1.0L) * (2 * i + 1))), i - 2, i);
                tb_set(&B, sqrtl(((i + 1.0L) * (i + 2)) / ((2 * i + 1.0L) * (2 * i + 3))), i, i);
            }
        } else {
            if (n > 1) tb_set(&A, 1.0L / 3, 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&A, -(i + 1) / (2 * i + 1.0L) * (i + 1), i - 2, i);
                tb_set(&A, i / (2 * i + 1.0L) * i, i, i);
            }
            if (n > 0) tb_set(&B, 1, 0, 0);
            if (n > 1) tb_set(&B, 1.0L / 3, 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&B, -1 / (2 * i + 1.0L), i - 2, i);
                tb_set(&B, 1 / (2 * i + 1.0L), i, i);
            }
        }
    }
    if (V) {
        const size_t inc = (size_t)n + 1;
        if (normcheb) {
            if (normleg) {
                if (n > 0) V[0] = sqrtl(2) / tgammal(0.5);
                if (n > 1) V[inc] = V[0] / sqrtl(1.5);
                for (int i = 2; i < n; i++) V[i * inc] = (2 * i) / sqrtl((2 * i + 1) * (2 * i - 1.0L)) * V[(i - 1) * inc];
            } else {
                if (n > 0) V[0] = 1 / tgammal(0.5);
                if (n > 1) V[inc] = sqrtl(2) * V[0];
                for (int i = 2; i < n; i++) V[i * inc] = (2 * i) * V[(i - 1) * inc] / (2 * i - 1);
            }
        } else {
            if (normleg) {
                if (n > 0) V[0] = sqrtl(2);
                if (n > 1) V[inc] = 1 / sqrtl(1.5);
                for (int i = 2; i < n; i++) V[i * inc] = (2 * i) / sqrtl((2 * i + 1) * (2 * i - 1.0L)) * V[(i - 1) * inc];
            } else {
                if (n > 0) V[0] = 1;
                if (n > 1) V[inc] = 1;
                for (int i = 2; i < n; i++) V[i * inc] = (2 * i) * V[(i - 1) * inc] / (2 * i - 1);
            }
        }
    }
    return finish(&A, &B, V, n);
}

/* transforms.c:80-93 with banded_source.c:2102-2131, 2386-2430 */
double *ftamd_plan_ultra2ultra(int norm1, int norm2, int n, double lambda_, double mu_) {
    const long double lambda = lambda_, mu = mu_;
    tb2 A, B;
    tb_alloc(&A, n);
    tb_alloc(&B, n);
    long double *V = calloc((size_t)n * (size_t)(n > 0 ? n : 1), sizeof(long double));
    if (A.d && B.d) {
        if (norm2) {
            if (n > 1) tb_set(&A, (1 + 2 * lambda) * copysignl(sqrtl((2 * mu + 1) / (2 * mu + 4)), mu), 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&A, -(i + 2 * mu) * (i + 2 * (mu - lambda)) * copysignl(sqrtl((i * (i - 1.0L)) / (4 * (i + mu) * (i + mu - 1))), mu), i - 2, i);
                tb_set(&A, i * (i + 2 * lambda) * copysignl(sqrtl(((i + 2 * mu) * (i + 2 * mu + 1)) / (4 * (i + mu) * (i + mu + 1))), mu), i, i);
            }
            if (n > 0) tb_set(&B, sqrtl((2 * mu + 1) / (2 * mu + 2)), 0, 0);
            if (n > 1) tb_set(&B, copysignl(sqrtl((2 * mu + 1) / (2 * mu + 4)), mu), 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&B, -copysignl(sqrtl((i * (i - 1.0L)) / (4 * (i + mu) * (i + mu - 1))), mu), i - 2, i);
                tb_set(&B, copysignl(sqrtl(((i + 2 * mu) * (i + 2 * mu + 1)) / (4 * (i + mu) * (i + mu + 1))), mu), i, i);
            }
        } else {
            if (n > 1) tb_set(&A, (1 + 2 * lambda) * mu / (1 + mu), 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&A, -(i + 2 * mu) * (i + 2 * (mu - lambda)) * mu / (i + mu), i - 2, i);
                tb_set(&A, i * (i + 2 * lambda) * mu / (i + mu), i, i);
            }
            if (n > 0) tb_set(&B, 1, 0, 0);
            if (n > 1) tb_set(&B, mu / (1 + mu), 1, 1);
            for (int i = 2; i < n; i++) {
                tb_set(&B, -mu / (i + mu), i - 2, i);
                tb_set(&B, mu / (i + mu), i, i);
            }
        }
    }
    if (V) {
        const size_t inc = (size_t)n + 1;
        if (norm1) {
            if (norm2) {
                if (n > 0)
                    V[0] = sqrtl(tgammal(lambda + 1) / (tgammal(0.5) * tgammal(lambda + 0.5))) *
                           sqrtl(tgammal(0.5) * tgammal(mu + 0.5) / tgammal(mu + 1));
                for (int i = 1; i < n; i++)
                    V[i * inc] = sqrtl((i - 1 + mu) / i * (i - 1 + 2 * mu) / (i + mu)) *
                                 sqrtl(i / (i - 1 + lambda) * (i + lambda) / (i - 1 + 2 * lambda)) * (i - 1 + lambda) /
                                 (i - 1 + mu) * V[(i - 1) * inc];
            } else {
                if (n > 0) V[0] = sqrtl(tgammal(lambda + 1) / (tgammal(0.5) * tgammal(lambda + 0.5)));
                for (int i = 1; i < n; i++)
                    V[i * inc] = sqrtl(i / (i - 1 + lambda) * (i + lambda) / (i - 1 + 2 * lambda)) * (i - 1 + lambda) /
                                 (i - 1 + mu) * V[(i - 1) * inc];
            }
        } else {
            if (norm2) {
                if (n > 0) V[0] = sqrtl(tgammal(0.5) * tgammal(mu + 0.5) / tgammal(mu + 1));
                for (int i = 1; i < n; i++)
                    V[i * inc] = sqrtl((i - 1 + mu) / i * (i - 1 + 2 * mu) / (i + mu)) * (i - 1 + lambda) / (i - 1 + mu) *
                                 V[(i - 1) * inc];
            } else {
                if (n > 0) V[0] = 1;
                for (int i = 1; i < n; i++) V[i * inc] = (i - 1 + lambda) / (i - 1 + mu) * V[(i - 1) * inc];
            }
        }
    }
    return finish(&A, &B, V, n);
}

/* --------------------------------------------------------------- Jacobi -- */

/* Entries of the reference's banded Jacobi operators (banded_source.c:
 * 1774-1905), long double, normalized (norm) or not.  Derivative D^order of
 * P^{(a,b)}: entry (j-order, j).  Multiplication x P^{(a,b)}: (j+d, j) for
 * d = -1, 0, 1.  Raising P^{(a,b)} -> P^{(a+1,b+1)}: (j-2+e, j) for e = 0, 1,
 * 2.  Lowering (1-x^2) P^{(a+1,b+1)} -> P^{(a,b)}: (j+e, j), e = 0, 1, 2. */
static long double jac_derivative(int norm, int order, int j, long double a, long double b) {
    long double v = 1;
    for (int k = 0; k < order; k++) v *= norm ? (j - k) * (j + a + b + k + 1) : (j + a + b + k + 1) / 2;
    return norm ? sqrtl(v) : v;
}

static long double jac_multiplication(int norm, int d, int j, long double a, long double b) {
    if (norm) {
        if (d == -1)
            return j == 1 ? 2 * sqrtl(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 2) * (a + b + 3)))
                          : 2 * sqrtl((j * (j + a) * (j + b) * (j + a + b)) /
                                      ((2 * j + a + b - 1) * (2 * j + a + b) * (2 * j + a + b) * (2 * j + a + b + 1)));
        if (d == 0) return j == 0 ? (b - a) / (a + b + 2) : ((b - a) * (a + b)) / ((2 * j + a + b) * (2 * j + a + b + 2));
        return j == 0 ? 2 * sqrtl(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 2) * (a + b + 3)))
                      : 2 * sqrtl(((j + 1) * (j + a + 1) * (j + b + 1) * (j + a + b + 1)) /
                                  ((2 * j + a + b + 1) * (2 * j + a + b + 2) * (2 * j + a + b + 2) * (2 * j + a + b + 3)));
    }
    if (d == -1) return (2 * (j + a) * (j + b)) / ((2 * j + a + b) * (2 * j + a + b + 1));
    if (d == 0) return j == 0 ? (b - a) / (a + b + 2) : ((b - a) * (a + b)) / ((2 * j + a + b) * (2 * j + a + b + 2));
    return j == 0 ? 2 / (a + b + 2) : (2 * (j + 1) * (j + a + b + 1)) / ((2 * j + a + b + 1) * (2 * j + a + b + 2));
}

static long double jac_raising(int norm, int e, int j, long double a, long double b) {
    if (norm) {
        if (e == 0)
            return -2 * sqrtl(((j - 1) * j * (j + a) * (j + b)) /
                              ((2 * j + a + b - 1) * (2 * j + a + b) * (2 * j + a + b) * (2 * j + a + b + 1)));
        if (e == 1) return 2 * (a - b) * sqrtl(j * (j + a + b + 1)) / ((2 * j + a + b) * (2 * j + a + b + 2));
        return j == 0 ? 2 * sqrtl(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 3)))
                      : 2 * sqrtl(((j + a + 1) * (j + b + 1) * (j + a + b + 1) * (j + a + b + 2)) /
                                  ((2 * j + a + b + 1) * (2 * j + a + b + 2) * (2 * j + a + b + 2) * (2 * j + a + b + 3)));
    }
    if (e == 0) return -((j + a) * (j + b)) / ((2 * j + a + b) * (2 * j + a + b + 1));
    if (e == 1) return ((a - b) * (j + a + b + 1)) / ((2 * j + a + b) * (2 * j + a + b + 2));
    return j == 0 ? 1 : ((j + a + b + 1) * (j + a + b + 2)) / ((2 * j + a + b + 1) * (2 * j + a + b + 2));
}

static long double jac_lowering(int norm, int e, int j, long double a, long double b) {
    if (norm) {
        if (e == 0)
            return j == 0 ? 2 * sqrtl(((a + 1) * (b + 1)) / ((a + b + 2) * (a + b + 3)))
                          : 2 * sqrtl(((j + a + 1) * (j + b + 1) * (j + a + b + 1) * (j + a + b + 2)) /
                                      ((2 * j + a + b + 1) * (2 * j + a + b + 2) * (2 * j + a + b + 2) * (2 * j + a + b + 3)));
        if (e == 1) return 2 * (a - b) * sqrtl((j + 1) * (j + a + b + 2)) / ((2 * j + a + b + 2) * (2 * j + a + b + 4));
        return -2 * sqrtl(((j + 1) * (j + 2) * (j + a + 2) * (j + b + 2)) /
                          ((2 * j + a + b + 3) * (2 * j + a + b + 4) * (2 * j + a + b + 4) * (2 * j + a + b + 5)));
    }
    if (e == 0) return (4 * (j + a + 1) * (j + b + 1)) / ((2 * j + a + b + 2) * (2 * j + a + b + 3));
    if (e == 1) return (4 * (a - b) * (j + 1)) / ((2 * j + a + b + 2) * (2 * j + a + b + 4));
    return -(4 * (j + 1) * (j + 2)) / ((2 * j + a + b + 3) * (2 * j + a + b + 4));
}

/* transforms.c:95-108 with banded_source.c:2133-2172 (diagonal) and
 * 2432-2455: A = -L1 D2 + (alpha-beta) D1 + (alpha+beta+2) M1 D1 in the
 * P^{(gamma+1, delta+1)} basis (D1, D2 derivatives of P^{(gamma,delta)}, L1
 * lowering and M1 multiplication at (gamma+1, delta+1)), B the raising
 * operator of P^{(gamma, delta)}; every product has a single term, so the
 * entries are formed in the reference's order. */
double *ftamd_plan_jac2jac(int norm1, int norm2, int n, double alpha_, double beta_, double gamma_, double delta_) {
    const long double al = alpha_, be = beta_, ga = gamma_, de = delta_;
    tb2 A, B;
    tb_alloc(&A, n);
    tb_alloc(&B, n);
    long double *V = calloc((size_t)n * (size_t)(n > 0 ? n : 1), sizeof(long double));
    if (A.d && B.d) {
        for (int j = 0; j < n; j++)
            for (int i = j - 2 < 0 ? 0 : j - 2; i <= j; i++) {
                long double ab = 0;
                if (j >= 2) ab = jac_lowering(norm2, i - (j - 2), j - 2, ga + 1, de + 1) * jac_derivative(norm2, 2, j, ga, de);
                long double v = -1 * ab + 0 * 0.0L;
                v = 1 * v + (al - be) * (i == j - 1 ? jac_derivative(norm2, 1, j, ga, de) : 0.0L);
                ab = 0;
                if (j >= 1) ab = jac_multiplication(norm2, i - (j - 1), j - 1, ga + 1, de + 1) * jac_derivative(norm2, 1, j, ga, de);
                v = (al + be + 2) * ab + 1 * v;
                tb_set(&A, v, i, j);
                tb_set(&B, jac_raising(norm2, i - (j - 2), j, ga, de), i, j);
            }
    }
    if (V) {
        const size_t inc = (size_t)n + 1;
        for (int i = 0; i < n; i++) {
            long double d;
            if (norm1 && norm2) {
                if (i == 0)
                    d = sqrtl(tgammal(al + be + 2) / (powl(2, al + be + 1) * tgammal(al + 1) * tgammal(be + 1)) *
                              (powl(2, ga + de + 1) * tgammal(ga + 1) * tgammal(de + 1)) / tgammal(ga + de + 2));
                else if (i == 1)
                    d = (al + be + 2) / (ga + de + 2) * sqrtl((al + be + 3) / ((al + 1) * (be + 1)) * ((ga + 1) * (de + 1)) / (ga + de + 3)) * V[0];
                else
                    d = (2 * i + al + be) / (2 * i + ga + de) *
                        sqrtl(((2 * i + al + be - 1) * (2 * i + al + be + 1)) / ((i + al) * (i + be) * (i + al + be)) *
                              ((i + ga) * (i + de) * (i + ga + de)) / ((2 * i + ga + de - 1) * (2 * i + ga + de + 1))) *
                        V[(i - 1) * inc];
            } else if (norm1) {
                if (i == 0) d = sqrtl(tgammal(al + be + 2) / (powl(2, al + be + 1) * tgammal(al + 1) * tgammal(be + 1)));
                else if (i == 1) d = (al + be + 2) / (ga + de + 2) * sqrtl((al + be + 3) / ((al + 1) * (be + 1))) * V[0];
                else
                    d = sqrtl(((2 * i + al + be - 1) * (2 * i + al + be) * (2 * i + al + be) * (2 * i + al + be + 1)) /
                              (i * (i + al) * (i + be) * (i + al + be))) *
                        (i * (i + ga + de)) / ((2 * i + ga + de - 1) * (2 * i + ga + de)) * V[(i - 1) * inc];
            } else if (norm2) {
                if (i == 0) d = sqrtl((powl(2, ga + de + 1) * tgammal(ga + 1) * tgammal(de + 1)) / tgammal(ga + de + 2));
                else if (i == 1) d = (al + be + 2) / (ga + de + 2) * sqrtl(((ga + 1) * (de + 1)) / (ga + de + 3)) * V[0];
                else
                    d = ((2 * i + al + be - 1) * (2 * i + al + be)) / (i * (i + al + be)) *
                        sqrtl((i * (i + ga) * (i + de) * (i + ga + de)) /
                              ((2 * i + ga + de - 1) * (2 * i + ga + de) * (2 * i + ga + de) * (2 * i + ga + de + 1))) *
                        V[(i - 1) * inc];
            } else {
                if (i == 0) d = 1;
                else if (i == 1) d = (al + be + 2) / (ga + de + 2);
                else
                    d = ((2 * i + al + be - 1) * (2 * i + al + be) * (i + ga + de)) /
                        ((i + al + be) * (2 * i + ga + de - 1) * (2 * i + ga + de)) * V[(i - 1) * inc];
            }
            V[i * inc] = d;
        }
    }
    return finish(&A, &B, V, n);
}

/* transforms.c:157-169: rows scaled by sqrt(1/pi), sqrt(2/pi) (rounded to
 * double first) for the unnormalized Chebyshev target. */
double *ftamd_plan_jac2cheb(int normjac, int normcheb, int n, double alpha, double beta) {
    double *V = ftamd_plan_jac2jac(normjac, 1, n, alpha, beta, -0.5, -0.5);
    if (V && !normcheb)
        for (int j = 0; j < n; j++)
            for (int i = 0; i <= j; i++) {
                const double sclrow = i ? (double)sqrtl(M_2_PIl) : (double)sqrtl(M_1_PIl);
                V[i + (size_t)j * n] *= sclrow;
            }
    return V;
}

/* transforms.c:171-183: columns scaled by sqrt(pi), sqrt(pi/2). */
double *ftamd_plan_cheb2jac(int normcheb, int normjac, int n, double alpha, double beta) {
    double *V = ftamd_plan_jac2jac(1, normjac, n, -0.5, -0.5, alpha, beta);
    if (V && !normcheb)
        for (int j = 0; j < n; j++) {
            const double sclcol = j ? (double)sqrtl(M_PI_2l) : (double)sqrtl(M_PIl);
            for (int i = 0; i <= j; i++) V[i + (size_t)j * n] *= sclcol;
        }
    return V;
}

/* transforms.c:185-197 (with :141-155): ultraspherical(lambda) -> Chebyshev
 * through Jacobi(lambda-1/2, lambda-1/2) -> Jacobi(-1/2, -1/2); for the
 * unnormalized Chebyshev side the entries of equal parity (the others are
 * zero) scaled by the rows' sqrt(1/pi), sqrt(2/pi); for the unnormalized
 * ultraspherical side the columns by the ratio of the two normalizations. */
double *ftamd_plan_ultra2cheb(int normultra, int normcheb, int n, double lambda) {
    double *V = ftamd_plan_jac2jac(normultra, 1, n, lambda - 0.5, lambda - 0.5, -0.5, -0.5);
    if (!V) return NULL;
    if (!normultra) {
        double sclcol = 1;
        for (int j = 0; j < n; j++) {
            if (j > 0) sclcol = (2 * lambda + j - 1) / (lambda + j - 0.5) * sclcol;
            for (int i = 0; i <= j; i++) V[i + (size_t)j * n] *= sclcol;
        }
    }
    if (!normcheb)
        for (int j = 0; j < n; j++)
            for (int i = j; i >= 0; i -= 2) {
                const double sclrow = i ? (double)sqrtl(M_2_PIl) : (double)sqrtl(M_1_PIl);
                V[i + (size_t)j * n] *= sclrow;
            }
    return V;
}

/* transforms.c:199-211 (with :125-139): the inverse direction. */
double *ftamd_plan_cheb2ultra(int normcheb, int normultra, int n, double lambda) {
    double *V = ftamd_plan_jac2jac(1, normultra, n, -0.5, -0.5, lambda - 0.5, lambda - 0.5);
    if (!V) return NULL;
    if (!normultra) {
        double sclrow = 1;
        double *row = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        if (!row) {
            free(V);
            return NULL;
        }
        for (int i = 0; i < n; i++) {
            if (i > 0) sclrow = (lambda + i - 0.5) / (2 * lambda + i - 1) * sclrow;
            row[i] = sclrow;
        }
        for (int j = 0; j < n; j++)
            for (int i = 0; i <= j; i++) V[i + (size_t)j * n] *= row[i];
        free(row);
    }
    if (!normcheb)
        for (int j = 0; j < n; j++) {
            const double sclcol = j ? (double)sqrtl(M_PI_2l) : (double)sqrtl(M_PIl);
            for (int i = j; i >= 0; i -= 2) V[i + (size_t)j * n] *= sclcol;
        }
    return V;
}
