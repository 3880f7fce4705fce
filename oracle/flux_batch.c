/*
 * Test infrastructure: array front end of the reference's example flux
 * models (examples/pumas/flux.c, compiled from the reference's source by
 * oracle/Makefile next to this file). Used as the oracle of the on-device
 * flux tallies.
 */
#include <stddef.h>

#include "flux.h"

/* kind 1: flux_gaisser, kind 2: flux_gccly. */
int ref_flux_batch(int kind, size_t n, const double * cos_theta,
    const double * kinetic, const double * charge, double * out)
{
        for (size_t i = 0; i < n; i++) {
                if (kind == 1)
                        out[i] = flux_gaisser(cos_theta[i], kinetic[i],
                            charge[i]);
                else if (kind == 2)
                        out[i] = flux_gccly(cos_theta[i], kinetic[i],
                            charge[i]);
                else
                        return -1;
        }
        return 0;
}
