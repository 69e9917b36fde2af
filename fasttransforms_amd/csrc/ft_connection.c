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
