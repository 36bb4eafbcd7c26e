/*
 * Drives the FP64 oracle (oracle/joseph_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer on the CPU: forward and backprojection in SET
 * and ADD mode on small parallel and cone geometries (with detector
 * supersampling, a tilted detector and rays along every axis), checking the
 * adjoint identity <A x, y> = <x, A^T y>. Exits non-zero on a failed check;
 * the sanitizers abort on the first error they find.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct {
    int32_t kind;
    int32_t n_angles, det_rows, det_cols;
    const double *vectors;
    int32_t nx, ny, nz;
    double vmin[3], vmax[3];
    int32_t det_supersampling;
} oracle_geom;

void oracle_fp(const oracle_geom *g, const double *vol, double *sino, int additive);
void oracle_bp(const oracle_geom *g, double *vol, const double *sino, int additive);
void oracle_axes(const oracle_geom *g, int32_t *axes);

static double frand(uint64_t *s)
{
    *s = *s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(*s >> 11) / 9007199254740992.0;
}

/* Angles rotating about the axis `ax` (0 = x, 1 = y, 2 = z), detector at the
 * origin, optional tilt of the detector's v axis. */
static void vectors(int kind, int na, int ax, double tilt, double *vec)
{
    for (int a = 0; a < na; ++a) {
        const double th = 2.0 * M_PI * a / na;
        double r[3] = {0, 0, 0}, u[3] = {0, 0, 0}, v[3] = {0, 0, 0};
        const int p = (ax + 1) % 3, q = (ax + 2) % 3;
        r[p] = cos(th);
        r[q] = sin(th);
        u[p] = -sin(th) * 0.03;
        u[q] = cos(th) * 0.03;
        v[ax] = 0.03 * cos(tilt);
        v[p] = 0.03 * sin(tilt);
        double *w = vec + 12 * a;
        for (int c = 0; c < 3; ++c) {
            w[c] = kind ? -3.0 * r[c] : r[c]; /* source (cone) or ray direction */
            w[3 + c] = 0.0;
            w[6 + c] = u[c];
            w[9 + c] = v[c];
        }
    }
}

static int check(int kind, int ax, double tilt, int ss, int nx, int ny, int nz, int rows, int cols, int na)
{
    double *vec = malloc(sizeof(double) * 12 * na);
    vectors(kind, na, ax, tilt, vec);
    oracle_geom g = {kind, na, rows, cols, vec, nx, ny, nz, {-0.5, -0.4, -0.45}, {0.5, 0.45, 0.4}, ss};
    const int64_t nv = (int64_t)nx * ny * nz, ns = (int64_t)rows * na * cols;
    double *x = malloc(sizeof(double) * nv), *bx = malloc(sizeof(double) * nv);
    double *y = malloc(sizeof(double) * ns), *ax_ = malloc(sizeof(double) * ns);
    int32_t *axes = malloc(sizeof(int32_t) * na);
    uint64_t s = 12345u + (uint64_t)(kind * 7 + ax * 3 + ss);
    for (int64_t i = 0; i < nv; ++i) x[i] = frand(&s);
    for (int64_t i = 0; i < ns; ++i) y[i] = frand(&s);
    oracle_axes(&g, axes);
    oracle_fp(&g, x, ax_, 0);
    oracle_fp(&g, x, ax_, 1); /* ADD: 2 A x */
    oracle_bp(&g, bx, y, 0);
    double l = 0, r = 0, n1 = 0, n2 = 0;
    for (int64_t i = 0; i < ns; ++i) {
        l += 0.5 * ax_[i] * y[i];
        n1 += 0.25 * ax_[i] * ax_[i];
    }
    for (int64_t i = 0; i < nv; ++i) r += x[i] * bx[i];
    for (int64_t i = 0; i < ns; ++i) n2 += y[i] * y[i];
    const double rel = fabs(l - r) / (sqrt(n1) * sqrt(n2));
    printf("kind %d axis %d tilt %.2f ss %d: adjoint %.2e\n", kind, ax, tilt, ss, rel);
    free(vec);
    free(x);
    free(bx);
    free(y);
    free(ax_);
    free(axes);
    return !(rel < 1e-12 && n1 > 0);
}

int main(void)
{
    int bad = 0;
    for (int kind = 0; kind < 2; ++kind)
        for (int ax = 0; ax < 3; ++ax) bad += check(kind, ax, 0.0, 1, 9, 8, 7, 10, 12, 7);
    bad += check(1, 2, 0.3, 1, 9, 8, 7, 10, 12, 7);
    bad += check(0, 2, 0.3, 2, 9, 8, 7, 10, 12, 5);
    bad += check(1, 1, 0.0, 3, 6, 7, 8, 7, 9, 4);
    printf(bad ? "oracle checks FAILED\n" : "oracle checks passed\n");
    return bad ? 1 : 0;
}
