/*
 * ref_shim.c -- harness around the compiled reference (test infrastructure).
 *
 * Exposes plain-C entry points of the reference library for ctypes, where the
 * reference API takes callbacks or structs that ctypes cannot express:
 * solve_aberth with dk_log_det_ms as dlogf (ms_core.c:320-328), the Domain
 * struct, and the scalar eval_total_wave over a grid (ms_core.c:629-640).
 * Built by oracle/ref.mk into oracle/_ref/libmsref_shim.so.
 */
#include <stdlib.h>
#include <stdio.h>
#include "dcomplex_type.h"
#include "medium_type.h"
#include "domain_type.h"
#include "incident_wave_type.h"
#include "ms_matrix_util.h"
#include "root_finder.h"
#include "incident_wave_util.h"

void init_medium(Medium* med, int d, int n, ScatteringModel* scmodel, Shape shape, double ratio);

static Medium* shim_med;
static dcomplex shim_dlogf(dcomplex k) { return dk_log_det_ms(shim_med, k); }

void shim_aberth(Medium* med, double xmin, double xmax, double ymin, double ymax,
                 int* nz, dcomplex* z, int nexpsup, int maxit, double toler) {
    Domain dom = {xmin, xmax, ymin, ymax};
    shim_med = med;
    solve_aberth(shim_dlogf, &dom, nz, z, nexpsup, maxit, toler, 0);
}

void shim_wave_grid(Medium* med, int source, double kre, double kim, const dcomplex* a,
                    int nx, int ny, double xmin, double xmax, double ymin, double ymax, dcomplex* psi) {
    IncidentWave iw;
    iw.source = (IncidentWaveSource)source;
    iw.k = kre + I * kim;
    double dx = (xmax - xmin) / (nx - 1), dy = (ymax - ymin) / (ny - 1);
    double* pos = (double*)calloc(med->d, sizeof(double));
    for (int l = 0; l < nx * ny; l++) {
        pos[0] = xmin + (l % nx) * dx;
        if (med->d > 1) pos[1] = ymax - (l / nx) * dy;
        psi[l] = eval_total_wave(&iw, med, (dcomplex*)a, pos);
    }
    free(pos);
}
