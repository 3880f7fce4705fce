/*
 * Test infrastructure: the device geometries (pumas_amd/csrc/geometry.h)
 * built for the host, located on arrays of points, for comparison with an
 * independent classification of the same descriptions
 * (tests/test_geometry_pin.py).
 */
#include <stdint.h>
#include <stdlib.h>

#include "pumas_gpu.h"
#include "geometry.h"

/* Medium and proposed step at each of n points, along sgn * dir. */
int geo_check_locate(const struct pumas_gpu_geometry_desc * d,
    const double * pos, const double * dir, double sgn, int64_t n,
    int * medium, double * step, int * step_type)
{
        struct pg_geometry g = { 0 };
        int * material = malloc((d->n_media > 0 ? d->n_media : 1) *
            sizeof(int));
        double * density = malloc((d->n_media > 0 ? d->n_media : 1) *
            sizeof(double));
        if ((material == NULL) || (density == NULL)) return -1;
        for (int i = 0; i < d->n_media; i++) {
                material[i] = d->media[i].material;
                density[i] = d->media[i].density;
        }
        g.kind = d->kind;
        g.step_type = d->step_type;
        g.n_media = d->n_media;
        g.material = material;
        g.density = density;
        g.n_layers = d->n_layers;
        g.z = d->z;
        g.nx = d->nx;
        g.ny = d->ny;
        g.nz = d->nz;
        g.x0 = d->x0;
        g.y0 = d->y0;
        g.z0 = d->z0;
        g.dx = d->dx;
        g.dy = d->dy;
        g.dz = d->dz;
        g.height = d->height;
        g.cell_density = d->density;
        g.label = d->label;
        g.z_top = d->z_top;
        for (int64_t i = 0; i < n; i++)
                step_type[i] = geo_locate(&g, pos + 3 * i, dir + 3 * i, sgn,
                    medium + i, step + i);
        free(material);
        free(density);
        return 0;
}
