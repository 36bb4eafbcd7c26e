/*
 * FP64 CPU oracle for the tomographic projection path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP kernels are
 * compared against, and the CPU baseline that the benchmark times beside
 * them. No product code path may link, load or call it.
 *
 * What it restates. tomosipo hands every projection to ASTRA through
 * astra.experimental.direct_FPBP3D (tomosipo/astra.py:147-153) with a
 * "cuda3d" projector built from the geometry dictionaries
 * (tomosipo/astra.py:80-98, geometry/volume.py:286-301,
 * geometry/cone_vec.py:174-191, geometry/parallel_vec.py:175-192).
 * ASTRA itself is not part of the reference; its published algorithm for
 * that projector is Joseph's ray-driven method (P. M. Joseph, "An improved
 * algorithm for reprojecting rays through pixel images", IEEE TMI 1(3),
 * 1982) in vector geometry, restated here:
 *
 *   Frames. Volume shape (nz, ny, nx), window [min, max] per axis, voxel size
 *   h = (max - min) / n. Grid coordinate g = (X - min) / h - 1/2 so voxel
 *   centres sit on integers. Detector pixel (i, j) (row i, column j) has its
 *   centre at det + (j + 1/2 - cols/2) u + (i + 1/2 - rows/2) v; with detector
 *   supersampling k the sub-ray (p, q) of that pixel starts at
 *   det + (j + (q + 1/2)/k - cols/2) u + (i + (p + 1/2)/k - rows/2) v.
 *   Sinogram layout y[i][a][j] (tomosipo links/base.py:21-31), volume layout
 *   x[z][y][x].
 *
 *   Dominant axis per angle: e = (det - src) / h (cone) or ray / h (parallel),
 *   m = argmax |e_c|, ties resolved in the order x, y, z.
 *
 *   Forward projection of a ray with point P and direction D (grid units):
 *   for every slice n = 0 .. N_m - 1, Q = P + ((n - P_m) / D_m) D, the sample
 *   is the bilinear interpolation of slice n at (Q_c1, Q_c2) with zero outside
 *   the volume; the ray value is l * sum(samples) with
 *   l = |h * D| / |D_m| (world length per slice step). A pixel's value is the
 *   mean over its k*k sub-rays. SET mode writes, ADD mode accumulates
 *   (astra.py:132-135).
 *
 *   Backprojection is defined as the exact transpose of that matrix:
 *   x[k] += sum over rays and slices of l * w_bilinear * y / k^2.
 *
 * Parallelism: FP over (angle, row); BP per angle over slices n, which is
 * race-free because a sample in slice n only touches voxels with k_m = n.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t kind; /* 0 = parallel3d_vec, 1 = cone_vec */
    int32_t n_angles, det_rows, det_cols;
    const double *vectors; /* n_angles x 12: [ray|src, det, u, v], each (x, y, z) */
    int32_t nx, ny, nz;
    double vmin[3], vmax[3]; /* (x, y, z) */
    int32_t det_supersampling;
} oracle_geom;

typedef struct {
    int m, c1, c2;                 /* dominant axis and the two in-slice axes (0=x,1=y,2=z) */
    double S[3];                   /* source (cone) in grid coordinates */
    double E[3];                   /* ray direction (parallel) in grid units */
    double P0[3], U[3], V[3];      /* pixel (0,0) corner-relative origin and steps, grid units */
    int cone;
} angle_setup;

static void setup_angle(const oracle_geom *g, int a, const double h[3], angle_setup *s)
{
    const double *vec = g->vectors + 12 * (int64_t)a;
    const double n3[3] = {g->nx, g->ny, g->nz};
    s->cone = g->kind == 1;
    double e[3];
    for (int c = 0; c < 3; ++c) {
        double src = vec[c], det = vec[3 + c], u = vec[6 + c], v = vec[9 + c];
        s->U[c] = u / h[c];
        s->V[c] = v / h[c];
        /* grid coordinate of the detector centre, then shift to the corner of pixel (0,0) */
        double dg = (det - g->vmin[c]) / h[c] - 0.5;
        s->P0[c] = dg - 0.5 * g->det_cols * s->U[c] - 0.5 * g->det_rows * s->V[c];
        if (s->cone) {
            s->S[c] = (src - g->vmin[c]) / h[c] - 0.5;
            e[c] = (det - src) / h[c];
            s->E[c] = 0.0;
        } else {
            s->E[c] = src / h[c];
            e[c] = s->E[c];
            s->S[c] = 0.0;
        }
    }
    (void)n3;
    double ax = fabs(e[0]), ay = fabs(e[1]), az = fabs(e[2]);
    if (ax >= ay && ax >= az) s->m = 0;
    else if (ay >= az) s->m = 1;
    else s->m = 2;
    s->c1 = s->m == 0 ? 1 : 0;
    s->c2 = s->m == 2 ? 1 : 2;
}

/* Sub-ray point P and direction D for detector position (fi, fj) measured
 * from the corner of pixel (0,0) in pixel units. */
static void ray_of(const angle_setup *s, double fi, double fj, double P[3], double D[3])
{
    for (int c = 0; c < 3; ++c) {
        P[c] = s->P0[c] + fj * s->U[c] + fi * s->V[c];
        D[c] = s->cone ? P[c] - s->S[c] : s->E[c];
    }
}

static inline int64_t vidx(const oracle_geom *g, int64_t x, int64_t y, int64_t z)
{
    return (z * g->ny + y) * g->nx + x;
}

static void voxel_sizes(const oracle_geom *g, double h[3], int64_t n[3])
{
    n[0] = g->nx; n[1] = g->ny; n[2] = g->nz;
    for (int c = 0; c < 3; ++c) h[c] = (g->vmax[c] - g->vmin[c]) / (double)n[c];
}

/* Value of one (sub-)ray. */
static double fp_ray(const oracle_geom *g, const angle_setup *s, const double h[3],
                     const int64_t n[3], const double *vol, const double P[3], const double D[3])
{
    const int m = s->m, c1 = s->c1, c2 = s->c2;
    if (D[m] == 0.0) return 0.0;
    double len = sqrt(h[0] * h[0] * D[0] * D[0] + h[1] * h[1] * D[1] * D[1] +
                      h[2] * h[2] * D[2] * D[2]) / fabs(D[m]);
    double sum = 0.0;
    for (int64_t sl = 0; sl < n[m]; ++sl) {
        double t = ((double)sl - P[m]) / D[m];
        double q1 = P[c1] + t * D[c1];
        double q2 = P[c2] + t * D[c2];
        double f1 = floor(q1), f2 = floor(q2);
        int64_t k1 = (int64_t)f1, k2 = (int64_t)f2;
        double w1 = q1 - f1, w2 = q2 - f2;
        for (int d1 = 0; d1 < 2; ++d1) {
            int64_t i1 = k1 + d1;
            if (i1 < 0 || i1 >= n[c1]) continue;
            double a1 = d1 ? w1 : 1.0 - w1;
            for (int d2 = 0; d2 < 2; ++d2) {
                int64_t i2 = k2 + d2;
                if (i2 < 0 || i2 >= n[c2]) continue;
                double a2 = d2 ? w2 : 1.0 - w2;
                int64_t idx3[3];
                idx3[m] = sl; idx3[c1] = i1; idx3[c2] = i2;
                sum += a1 * a2 * vol[vidx(g, idx3[0], idx3[1], idx3[2])];
            }
        }
    }
    return len * sum;
}

void oracle_fp(const oracle_geom *g, const double *vol, double *sino, int additive)
{
    double h[3];
    int64_t n[3];
    voxel_sizes(g, h, n);
    const int k = g->det_supersampling > 0 ? g->det_supersampling : 1;
    const int64_t rows = g->det_rows, cols = g->det_cols, na = g->n_angles;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t a = 0; a < na; ++a) {
        for (int64_t i = 0; i < rows; ++i) {
            angle_setup s;
            setup_angle(g, (int)a, h, &s);
            for (int64_t j = 0; j < cols; ++j) {
                double acc = 0.0;
                for (int p = 0; p < k; ++p)
                    for (int q = 0; q < k; ++q) {
                        double P[3], D[3];
                        ray_of(&s, (double)i + (p + 0.5) / k, (double)j + (q + 0.5) / k, P, D);
                        acc += fp_ray(g, &s, h, n, vol, P, D);
                    }
                acc /= (double)(k * k);
                int64_t o = (i * na + a) * cols + j;
                sino[o] = additive ? sino[o] + acc : acc;
            }
        }
    }
}

void oracle_bp(const oracle_geom *g, double *vol, const double *sino, int additive)
{
    double h[3];
    int64_t n[3];
    voxel_sizes(g, h, n);
    const int k = g->det_supersampling > 0 ? g->det_supersampling : 1;
    const int64_t rows = g->det_rows, cols = g->det_cols, na = g->n_angles;
    const int64_t nvox = n[0] * n[1] * n[2];
    const int64_t nsub = rows * cols * k * k;
    /* per angle: (P, D, l * y / k^2) of every sub-ray */
    double *rays = (double *)malloc(sizeof(double) * 7 * (size_t)nsub);
    if (!rays) return;
    if (!additive) memset(vol, 0, sizeof(double) * (size_t)nvox);
    for (int64_t a = 0; a < na; ++a) {
        angle_setup s;
        setup_angle(g, (int)a, h, &s);
        const int m = s.m, c1 = s.c1, c2 = s.c2;
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < nsub; ++r) {
            int64_t i = r / (cols * k * k), rem = r % (cols * k * k);
            int64_t j = rem / (k * k), pq = rem % (k * k);
            int p = (int)(pq / k), q = (int)(pq % k);
            double *R = rays + 7 * r;
            ray_of(&s, (double)i + (p + 0.5) / k, (double)j + (q + 0.5) / k, R, R + 3);
            double *D = R + 3;
            double yv = sino[(i * na + a) * cols + j] / (double)(k * k);
            if (D[m] == 0.0) { R[6] = 0.0; continue; }
            double len = sqrt(h[0] * h[0] * D[0] * D[0] + h[1] * h[1] * D[1] * D[1] +
                              h[2] * h[2] * D[2] * D[2]) / fabs(D[m]);
            R[6] = len * yv;
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t sl = 0; sl < n[m]; ++sl) {
            for (int64_t r = 0; r < nsub; ++r) {
                const double *P = rays + 7 * r, *D = P + 3;
                double wy = P[6];
                if (wy == 0.0) continue;
                double t = ((double)sl - P[m]) / D[m];
                double q1 = P[c1] + t * D[c1];
                double q2 = P[c2] + t * D[c2];
                double f1 = floor(q1), f2 = floor(q2);
                int64_t k1 = (int64_t)f1, k2 = (int64_t)f2;
                double w1 = q1 - f1, w2 = q2 - f2;
                for (int d1 = 0; d1 < 2; ++d1) {
                    int64_t i1 = k1 + d1;
                    if (i1 < 0 || i1 >= n[c1]) continue;
                    double a1 = d1 ? w1 : 1.0 - w1;
                    for (int d2 = 0; d2 < 2; ++d2) {
                        int64_t i2 = k2 + d2;
                        if (i2 < 0 || i2 >= n[c2]) continue;
                        double a2 = d2 ? w2 : 1.0 - w2;
                        int64_t idx3[3];
                        idx3[m] = sl; idx3[c1] = i1; idx3[c2] = i2;
                        vol[vidx(g, idx3[0], idx3[1], idx3[2])] += a1 * a2 * wy;
                    }
                }
            }
        }
    }
    free(rays);
}

/* Dominant axis of every angle (0 = x, 1 = y, 2 = z), for tests. */
void oracle_axes(const oracle_geom *g, int32_t *axes)
{
    double h[3];
    int64_t n[3];
    voxel_sizes(g, h, n);
    for (int a = 0; a < g->n_angles; ++a) {
        angle_setup s;
        setup_angle(g, a, h, &s);
        axes[a] = s.m;
    }
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
