/*
 * ref_ext_shim.c -- second harness around the compiled reference (test
 * infrastructure), linked against every object of the MSModel core rather than
 * the Foldy-Lax subset of libmsref.so.  It exposes, with plain-C signatures for
 * ctypes:
 *   - the k-point generators of the complex map, the complex cut and the log cut
 *     (init_complexmap + get_arg_cmap, complex_map_util.c:27-31, 53-;
 *     init_complexcut + get_arg_ccut, complex_cut_util.c:32-36, 59-;
 *     init_logcut + get_arg_lcut, log_cut_util.c:35-37, 72-101);
 *   - the whole [eigenstate_list] driver run_eigenstate_list_serial (ms_core.c:
 *     406-430: Halton starts, the root finder of the chosen method -- Aberth on
 *     det M, or damped Muller on the smallest eigenvalue or on 1/Tr M^-1 -- then
 *     compute_eigenstates), returning the roots, eigenvalues and eigenvectors that
 *     save_eigenstate writes (eigenstate_util.c:202-213).
 * Built by oracle/ref.mk into oracle/_ref/libmsref_ext.so.
 */
#include <stdlib.h>
#include <string.h>
#include "dcomplex_type.h"
#include "medium_type.h"
#include "domain_type.h"
#include "complex_map_type.h"
#include "complex_cut_type.h"
#include "log_cut_type.h"
#include "eigenstate_list_type.h"
#include "complex_map_util.h"
#include "complex_cut_util.h"
#include "log_cut_util.h"
#include "eigenstate_list_util.h"

void run_eigenstate_list_serial(EigenstateList* eigls);
/* defined in complex_cut_util.c and log_cut_util.c, not declared in their headers */
dcomplex get_arg_ccut(ComplexCut* ccut, int l);
dcomplex get_arg_lcut(LogCut* lcut, int i);

void shim_kgrid_map(int nx, int ny, double xmin, double xmax, double ymin, double ymax, dcomplex* k) {
    ComplexMap cm;
    init_complexmap(&cm, nx, ny, xmin, xmax, ymin, ymax, 1, "grid");
    for (int l = 0; l < nx * ny; l++) k[l] = get_arg_cmap(&cm, l);
    del_complexmap(&cm);
}

/* 3 (ns + 2) points, triplet-major */
void shim_kgrid_cut(int ns, double zr0, double zi0, double zr1, double zi1, dcomplex* k) {
    ComplexCut cc;
    init_complexcut(&cc, ns, zr0 + I * zi0, zr1 + I * zi1, 1, "grid");
    for (int l = 0; l < cc.ntot; l++) k[l] = get_arg_ccut(&cc, l);
    del_complexcut(&cc);
}

void shim_kgrid_logcut(int ns, double rek, double imkmin, double imkmax, dcomplex* k) {
    LogCut lc;
    init_logcut(&lc, ns, rek, imkmin, imkmax, 1, "grid");
    for (int l = 0; l < lc.ntot; l++) k[l] = get_arg_lcut(&lc, l);
    del_logcut(&lc);
}

/* method: 0 determinant, 1 mineigval, 2 invtraceinv (EigenstateMethod).  Writes at most
 * ntarget states: kroot[i], mu[i], eigv[i * n + j]; returns the number of states. */
int shim_eigenstate_list(Medium* med, int method, int ntarget, double xmin, double xmax, double ymin, double ymax,
                         int maxit, double toler, dcomplex* kroot, dcomplex* mu, dcomplex* eigv) {
    EigenstateList el;
    memset(&el, 0, sizeof el);
    Domain* dom = (Domain*)malloc(sizeof(Domain));
    dom->xmin = xmin; dom->xmax = xmax; dom->ymin = ymin; dom->ymax = ymax;
    el.med = med;
    el.seed = 0;
    el.method = (EigenstateMethod)method;
    el.ntarget = ntarget;
    el.dom = dom;
    el.maxit = maxit;
    el.toler = toler;
    el.verb = 0;
    run_eigenstate_list_serial(&el);
    const int n = med->n, ns = el.nstate;
    for (int i = 0; i < ns; i++) {
        kroot[i] = el.eigtab[i]->kroot;
        mu[i] = el.eigtab[i]->mu;
        memcpy(eigv + (size_t)i * n, el.eigtab[i]->eigv, sizeof(dcomplex) * n);
    }
    del_eigenstate_list(&el);     /* frees dom too */
    return ns;
}
