/*
 * Device geometries: the `pumas_medium_cb` contract (include/pumas.h:478-511)
 * written as data plus one locate function.
 *
 * locate() answers the two questions a medium callback answers: the medium
 * at the current position (an index, -1 for outside / NULL) and, when asked,
 * a proposed step along the propagation direction (reverse to the momentum
 * in backward mode, include/pumas.h:504-507) with its checking mode.
 *
 * This header is plain C99 so that the same definitions serve the device
 * engine (hipcc) and host-side callbacks with identical arithmetic.
 */
#ifndef PUMAS_GEOMETRY_H
#define PUMAS_GEOMETRY_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define GEO_FN __host__ __device__ static inline
#else
#define GEO_FN static inline
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum { GEO_INFINITE = 0, GEO_SLABS = 1, GEO_TERRAIN = 2, GEO_VOXELS = 3 };
enum { GEO_STEP_CHECK = 0, GEO_STEP_RAW = 1 };

/* Step offset for raw steps, so that a step ends inside the next medium. */
#define GEO_STEP_EPSILON 1E-07
/* Direction cosines below this are treated as parallel to a plane. */
#define GEO_PARALLEL 1.1920928955078125E-07
/* Proposed step when no boundary lies ahead, in m. */
#define GEO_STEP_FREE 1E+03

/* Non uniform local properties of a media table entry
 * (enum pumas_gpu_locals): kind 0 is uniform. */
enum { GEO_LOCALS_UNIFORM = 0, GEO_LOCALS_EXPONENTIAL = 1 };
struct pg_locals {
        int kind;
        double scale, z_ref, cos_min;
};

struct pg_geometry {
        int kind;
        int step_type;
        /* Media table: material index and density (<= 0: material default),
         * and the local properties of each entry (NULL: all uniform). */
        int n_media;
        const int * material;
        const double * density;
        const struct pg_locals * locals;
        /* Slabs. */
        int n_layers;
        const double * z;
        /* Terrain / voxels. */
        int nx, ny, nz;
        double x0, y0, z0, dx, dy, dz;
        const float * height;
        const float * cell_density;
        const uint8_t * label;
        double z_top;
};

/* Distance along a line to leave [lo, hi) on one axis, or GEO_STEP_FREE. */
GEO_FN double geo_exit_distance(double x, double u, double lo, double hi)
{
        if (u > GEO_PARALLEL) return (hi - x) / u;
        if (u < -GEO_PARALLEL) return (lo - x) / u;
        return GEO_STEP_FREE;
}

/* Bilinear terrain height at (x, y); the point must lie inside the grid. */
GEO_FN double geo_terrain_height(
    const struct pg_geometry * g, double x, double y, int ix, int iy)
{
        const double hx = (x - (g->x0 + ix * g->dx)) / g->dx;
        const double hy = (y - (g->y0 + iy * g->dy)) / g->dy;
        const float * h = g->height + (int64_t)iy * g->nx + ix;
        const double h00 = h[0], h10 = h[1];
        const double h01 = h[g->nx], h11 = h[g->nx + 1];
        return (h00 * (1. - hx) + h10 * hx) * (1. - hy) +
            (h01 * (1. - hx) + h11 * hx) * hy;
}

/*
 * Distance along pos + s u to the bilinear surface of cell (ix, iy): the
 * smallest root s > 0 of z(s) - h(x(s), y(s)), a quadratic in s, or
 * GEO_STEP_FREE when the line does not meet the surface patch's extension.
 * In cell coordinates h = a + b hx + c hy + d hx hy (the bilinear form of
 * geo_terrain_height), and hx, hy are linear in s. Roots past the cell walls
 * are cut by the caller with the wall distances.
 */
GEO_FN double geo_terrain_crossing(const struct pg_geometry * g,
    const double * pos, const double * u, int ix, int iy)
{
        const float * hp = g->height + (int64_t)iy * g->nx + ix;
        const double h00 = hp[0], h10 = hp[1];
        const double h01 = hp[g->nx], h11 = hp[g->nx + 1];
        const double b = h10 - h00, c = h01 - h00;
        const double d = (h11 - h10) - (h01 - h00);
        const double hx = (pos[0] - (g->x0 + ix * g->dx)) / g->dx;
        const double hy = (pos[1] - (g->y0 + iy * g->dy)) / g->dy;
        const double ax = u[0] / g->dx, ay = u[1] / g->dy;
        /* f(s) = A s^2 + B s + C. */
        const double A = -d * ax * ay;
        const double B = u[2] - (b + d * hy) * ax - (c + d * hx) * ay;
        const double C = pos[2] - (h00 + b * hx + c * hy + d * hx * hy);
        double r1 = -1., r2 = -1.;
        if (A == 0.) {
                if (B != 0.) r1 = -C / B;
        } else {
                /* The roots as q / A and C / q, without cancellation. */
                const double disc = B * B - 4. * A * C;
                if (disc < 0.) return GEO_STEP_FREE;
                const double q = -0.5 * (B + ((B >= 0.) ? sqrt(disc) :
                                                          -sqrt(disc)));
                r1 = q / A;
                if (q != 0.) r2 = C / q;
        }
        double s = GEO_STEP_FREE;
        if ((r1 > 0.) && (r1 < s)) s = r1;
        if ((r2 > 0.) && (r2 < s)) s = r2;
        return s;
}

/*
 * Locate the medium at `pos` and optionally propose a step along `sgn * dir`.
 * Returns the checking mode of the proposed step (GEO_STEP_*).
 */
GEO_FN int geo_locate(const struct pg_geometry * g, const double * pos,
    const double * dir, double sgn, int * medium, double * step)
{
        switch (g->kind) {
        case GEO_INFINITE:
                *medium = 0;
                if (step != 0) *step = 0.;
                return g->step_type;

        case GEO_SLABS: {
                const double z = pos[2];
                int k = -1;
                int i;
                for (i = 0; i < g->n_layers; i++) {
                        if ((z >= g->z[i]) && (z < g->z[i + 1])) {
                                k = i;
                                break;
                        }
                }
                *medium = k;
                if (step != 0) {
                        if (k < 0) {
                                *step = -1.;
                        } else {
                                double s = geo_exit_distance(
                                    z, sgn * dir[2], g->z[k], g->z[k + 1]);
                                if (g->step_type == GEO_STEP_RAW)
                                        s += GEO_STEP_EPSILON;
                                *step = s;
                        }
                }
                return g->step_type;
        }

        case GEO_TERRAIN: {
                /* Cells of the heightmap: ix in [0, nx - 2], iy in
                 * [0, ny - 2]. Rock below the surface (medium 1 + cell),
                 * air above up to z_top (medium 0), outside elsewhere. */
                const double fx = (pos[0] - g->x0) / g->dx;
                const double fy = (pos[1] - g->y0) / g->dy;
                int ix = (int)fx, iy = (int)fy;
                if ((fx < 0.) || (fy < 0.) || (ix > g->nx - 2) ||
                    (iy > g->ny - 2) || (pos[2] < g->z0) ||
                    (pos[2] >= g->z_top)) {
                        *medium = -1;
                        if (step != 0) *step = -1.;
                        return GEO_STEP_CHECK;
                }
                const double h = geo_terrain_height(g, pos[0], pos[1], ix, iy);
                const int cell = iy * (g->nx - 1) + ix;
                *medium = (pos[2] < h) ? 1 + cell : 0;
                if (step != 0) {
                        /* The nearest of the cell walls, the top (air) or
                         * bottom (rock) plane and the exact crossing of the
                         * cell's bilinear surface. */
                        const double u[3] = { sgn * dir[0], sgn * dir[1],
                                sgn * dir[2] };
                        const double xl = g->x0 + ix * g->dx;
                        const double yl = g->y0 + iy * g->dy;
                        double s = geo_exit_distance(
                            pos[0], u[0], xl, xl + g->dx);
                        const double sy = geo_exit_distance(
                            pos[1], u[1], yl, yl + g->dy);
                        if (sy < s) s = sy;
                        double sz = GEO_STEP_FREE;
                        if ((*medium == 0) && (u[2] > GEO_PARALLEL))
                                sz = (g->z_top - pos[2]) / u[2];
                        else if ((*medium > 0) && (u[2] < -GEO_PARALLEL))
                                sz = (g->z0 - pos[2]) / u[2];
                        if (sz < s) s = sz;
                        const double sh =
                            geo_terrain_crossing(g, pos, u, ix, iy);
                        if (sh < s) s = sh;
                        if (s < GEO_STEP_EPSILON) s = GEO_STEP_EPSILON;
                        *step = s;
                }
                return GEO_STEP_CHECK;
        }

        case GEO_VOXELS: {
                const double fx = (pos[0] - g->x0) / g->dx;
                const double fy = (pos[1] - g->y0) / g->dy;
                const double fz = (pos[2] - g->z0) / g->dz;
                int ix = (int)fx, iy = (int)fy, iz = (int)fz;
                if ((fx < 0.) || (fy < 0.) || (fz < 0.) || (ix >= g->nx) ||
                    (iy >= g->ny) || (iz >= g->nz)) {
                        *medium = -1;
                        if (step != 0) *step = -1.;
                        return g->step_type;
                }
                *medium = 1 + (iz * g->ny + iy) * g->nx + ix;
                if (step != 0) {
                        const double xl = g->x0 + ix * g->dx;
                        const double yl = g->y0 + iy * g->dy;
                        const double zl = g->z0 + iz * g->dz;
                        double s = geo_exit_distance(
                            pos[0], sgn * dir[0], xl, xl + g->dx);
                        const double sy = geo_exit_distance(
                            pos[1], sgn * dir[1], yl, yl + g->dy);
                        if (sy < s) s = sy;
                        const double sz = geo_exit_distance(
                            pos[2], sgn * dir[2], zl, zl + g->dz);
                        if (sz < s) s = sz;
                        if (g->step_type == GEO_STEP_RAW)
                                s += GEO_STEP_EPSILON;
                        *step = s;
                }
                return g->step_type;
        }
        }
        *medium = -1;
        if (step != 0) *step = -1.;
        return GEO_STEP_CHECK;
}

/* The media table entry whose local properties apply to a medium, or -1
 * when the medium is uniform by construction (terrain ground columns and
 * voxels, whose densities are per cell). */
GEO_FN int geo_locals_entry(const struct pg_geometry * g, int medium)
{
        if (g->locals == 0) return -1;
        switch (g->kind) {
        case GEO_TERRAIN:
                return (medium == 0) ? 0 : -1;
        case GEO_VOXELS:
                return -1;
        default:
                return medium;
        }
}

/* Material index and density of a medium (density <= 0: default). */
GEO_FN void geo_medium_properties(const struct pg_geometry * g, int medium,
    int * material, double * density)
{
        switch (g->kind) {
        case GEO_TERRAIN:
                if (medium == 0) {
                        *material = g->material[0];
                        *density = g->density[0];
                } else {
                        *material = g->material[1];
                        *density = g->cell_density[medium - 1];
                }
                return;
        case GEO_VOXELS: {
                const int v = medium - 1;
                const int label = g->label[v];
                *material = g->material[label];
                *density = g->cell_density[v];
                return;
        }
        default:
                *material = g->material[medium];
                *density = g->density[medium];
                return;
        }
}

#ifdef __cplusplus
}
#endif

#endif
