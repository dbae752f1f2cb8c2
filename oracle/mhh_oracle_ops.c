/*
 * mhh_oracle_ops.c -- CPU FP64 restatement of the per-operator loops.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * follows one reference loop nest; the expression trees keep the reference
 * order of operations so that the FP64 results agree to rounding.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "mhh_oracle.h"

#define DSMALL 1.e-9      /* include/constants.h:97 */
#define KAPPA  0.4        /* include/constants.h:28 */
#define GRAV   9.81       /* include/constants.h:29 */

static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double sq(double a) { return a*a; }

/* Stencil helpers: include/finite_difference.h:37-156. */
static inline double i2(double a, double b) { return 0.5*(a + b); }
static inline double i4ws(double a, double b, double c, double d)
{ return (7./12.)*(b+c) - (1./12.)*(a+d); }
static inline double i3ws(double a, double b, double c, double d)
{ return (3./12.)*(c-b) - (1./12.)*(d-a); }
static inline double i6ws(double a, double b, double c, double d, double e, double f)
{ return (37./60.)*(c+d) - (8./60.)*(b+e) + (1./60.)*(a+f); }
static inline double i5ws(double a, double b, double c, double d, double e, double f)
{ return (10./60.)*(d-c) - (5./60.)*(e-b) + (1./60.)*(f-a); }

void mo_dims_init(mo_dims* d, int itot, int jtot, int ktot,
                  int igc, int jgc, int kgc,
                  double xsize, double ysize, double zsize)
{
    /* src/grid.cxx:106-193, serial: npx = npy = 1 */
    d->itot = itot; d->jtot = jtot; d->ktot = ktot;
    d->imax = itot; d->jmax = jtot; d->kmax = ktot;
    d->igc = igc; d->jgc = jgc; d->kgc = kgc;
    d->icells = itot + 2*igc;
    d->jcells = jtot + 2*jgc;
    d->kcells = ktot + 2*kgc;
    d->ijcells = d->icells*d->jcells;
    d->ncells = d->ijcells*d->kcells;
    d->istart = igc; d->iend = itot + igc;
    d->jstart = jgc; d->jend = jtot + jgc;
    d->kstart = kgc; d->kend = ktot + kgc;
    d->xsize = xsize; d->ysize = ysize; d->zsize = zsize;
    d->dx = xsize / itot;
    d->dy = ysize / jtot;
}

/* ------------------------------------------------------------------ grid */
void mo_grid_calculate(const mo_dims* d, double* z, double* zh, double* dz,
                       double* dzh, double* dzi, double* dzhi)
{
    /* src/grid.cxx:275-305: 2nd-order ghost and half levels.  z holds the
       interior levels at [kstart, kend) on entry. */
    const int ks = d->kstart, ke = d->kend, kc = d->kcells;
    z[ks-1] = -z[ks];
    z[ke]   = 2.*d->zsize - z[ke-1];

    for (int k=ks+1; k<ke; ++k)
        zh[k] = 0.5*(z[k-1] + z[k]);
    zh[ks] = 0.;
    zh[ke] = d->zsize;

    for (int k=1; k<kc; ++k)
    {
        dzh[k]  = z[k] - z[k-1];
        dzhi[k] = 1./dzh[k];
    }
    dzh [ks-1] = dzh [ks+1];
    dzhi[ks-1] = dzhi[ks+1];

    for (int k=1; k<kc-1; ++k)
    {
        dz[k]  = zh[k+1] - zh[k];
        dzi[k] = 1./dz[k];
    }
    dz [ks-1] = dz [ks];
    dzi[ks-1] = dzi[ks];
    dz [ke]   = dz [ke-1];
    dzi[ke]   = dzi[ke-1];
}

/* ---------------------------------------------------------------- fields */
void mo_srand(unsigned int seed) { srand(seed); }
int  mo_rand(void) { return rand(); }

void mo_randomize(const mo_dims* d, double* data, const double* z,
                  double rndamp, double rndz, double rndexp)
{
    /* src/fields.cxx:1035-1079 (the caller seeds once with rndseed+2). */
    const int jj = d->icells, kk = d->ijcells;
    int kendrnd = d->kstart;
    while (z[kendrnd] < rndz)
        ++kendrnd;

    for (int k=d->kstart; k<kendrnd; ++k)
    {
        const double rndfac = pow((rndz - z[k])/rndz, rndexp);
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
                data[i + j*jj + k*kk] = rndfac * rndamp * ((double)rand() / (double)RAND_MAX - 0.5);
    }
}

void mo_add_mean_prof(const mo_dims* d, double* data, const double* prof, double offset)
{
    /* src/fields.cxx:1082-1101: prof holds ktot values starting at kstart. */
    const int jj = d->icells, kk = d->ijcells;
    #pragma omp parallel for
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
                data[i + j*jj + k*kk] += prof[k - d->kstart] - offset;
}

/* -------------------------------------------------------------- boundary */
void mo_boundary_cyclic(const mo_dims* d, double* a, int edge)
{
    /* src/boundary_cyclic.cxx:369-443 (serial). */
    const int jj = d->icells, kk = d->ijcells;
    if (edge == MO_EDGE_EW || edge == MO_EDGE_BOTH)
    {
        #pragma omp parallel for
        for (int k=0; k<d->kcells; ++k)
            for (int j=0; j<d->jcells; ++j)
                for (int i=0; i<d->igc; ++i)
                    a[i + j*jj + k*kk] = a[d->iend - d->igc + i + j*jj + k*kk];
        #pragma omp parallel for
        for (int k=0; k<d->kcells; ++k)
            for (int j=0; j<d->jcells; ++j)
                for (int i=0; i<d->igc; ++i)
                    a[i + d->iend + j*jj + k*kk] = a[i + d->istart + j*jj + k*kk];
    }
    if (edge == MO_EDGE_NS || edge == MO_EDGE_BOTH)
    {
        if (d->jtot > 1)
        {
            for (int k=0; k<d->kcells; ++k)
                for (int j=0; j<d->jgc; ++j)
                    for (int i=0; i<d->icells; ++i)
                        a[i + j*jj + k*kk] = a[i + (d->jend - d->jgc + j)*jj + k*kk];
            for (int k=0; k<d->kcells; ++k)
                for (int j=0; j<d->jgc; ++j)
                    for (int i=0; i<d->icells; ++i)
                        a[i + (j + d->jend)*jj + k*kk] = a[i + (j + d->jstart)*jj + k*kk];
        }
        else
        {
            for (int k=d->kstart; k<d->kend; ++k)
                for (int j=0; j<d->jgc; ++j)
                    for (int i=0; i<d->icells; ++i)
                    {
                        const double ref = a[i + d->jstart*jj + k*kk];
                        a[i + j*jj + k*kk] = ref;
                        a[i + (d->jend + j)*jj + k*kk] = ref;
                    }
        }
    }
}

void mo_set_bc(const mo_dims* d, double* a, double* agrad, double* aflux,
               int sw, double aval, double visc, double offset, int set_flux_grad)
{
    /* src/boundary.cxx:51-91 over the full ijcells plane. */
    const int n = d->ijcells;
    if (sw == MO_BC_DIRICHLET)
        for (int ij=0; ij<n; ++ij) a[ij] = aval - offset;
    else if (sw == MO_BC_NEUMANN)
        for (int ij=0; ij<n; ++ij) { agrad[ij] = aval; if (set_flux_grad) aflux[ij] = -aval*visc; }
    else if (sw == MO_BC_FLUX)
        for (int ij=0; ij<n; ++ij) { aflux[ij] = aval; if (set_flux_grad) agrad[ij] = -aval/visc; }
}

void mo_ghost_cells_bot_2nd(const mo_dims* d, double* a, const double* dzh, int bctype,
                            const double* abot, const double* agradbot)
{
    /* src/boundary.cxx:700-729 */
    const int kk = d->ijcells, ks = d->kstart;
    if (bctype == MO_BC_DIRICHLET)
        for (int ij=0; ij<d->ijcells; ++ij)
            a[ij + (ks-1)*kk] = 2.*abot[ij] - a[ij + ks*kk];
    else if (bctype == MO_BC_NEUMANN || bctype == MO_BC_FLUX)
        for (int ij=0; ij<d->ijcells; ++ij)
            a[ij + (ks-1)*kk] = -agradbot[ij]*dzh[ks] + a[ij + ks*kk];
}

void mo_ghost_cells_top_2nd(const mo_dims* d, double* a, const double* dzh, int bctype,
                            double* atop, const double* agradtop)
{
    /* src/boundary.cxx:731-775 */
    const int kk = d->ijcells, ke = d->kend;
    if (bctype == MO_BC_DIRICHLET || bctype == MO_BC_OFF)
    {
        if (bctype == MO_BC_OFF)
            for (int ij=0; ij<d->ijcells; ++ij)
                atop[ij] = (3./2.)*a[ij + (ke-1)*kk] - (1./2.)*a[ij + (ke-2)*kk];
        for (int ij=0; ij<d->ijcells; ++ij)
            a[ij + ke*kk] = 2.*atop[ij] - a[ij + (ke-1)*kk];
    }
    else if (bctype == MO_BC_NEUMANN || bctype == MO_BC_FLUX)
        for (int ij=0; ij<d->ijcells; ++ij)
            a[ij + ke*kk] = agradtop[ij]*dzh[ke] + a[ij + (ke-1)*kk];
}

/* ----------------------------------------------------------------- smag2 */
static inline double strain2_core(const double* u, const double* v, const double* w,
                                  int ijk, int ii, int jj, int kk,
                                  double dxi, double dyi, double dzik,
                                  double dzhik, double dzhikp)
{
    /* Interior expression of include/diff_kernels.h:103-136. */
    return 2.*(
           + sq((u[ijk+ii]-u[ijk])*dxi)
           + sq((v[ijk+jj]-v[ijk])*dyi)
           + sq((w[ijk+kk]-w[ijk])*dzik)
           + 0.125*sq((u[ijk      ]-u[ijk   -jj])*dyi  + (v[ijk      ]-v[ijk-ii   ])*dxi)
           + 0.125*sq((u[ijk+ii   ]-u[ijk+ii-jj])*dyi  + (v[ijk+ii   ]-v[ijk      ])*dxi)
           + 0.125*sq((u[ijk   +jj]-u[ijk      ])*dyi  + (v[ijk   +jj]-v[ijk-ii+jj])*dxi)
           + 0.125*sq((u[ijk+ii+jj]-u[ijk+ii   ])*dyi  + (v[ijk+ii+jj]-v[ijk   +jj])*dxi)
           + 0.125*sq((u[ijk      ]-u[ijk   -kk])*dzhik  + (w[ijk      ]-w[ijk-ii   ])*dxi)
           + 0.125*sq((u[ijk+ii   ]-u[ijk+ii-kk])*dzhik  + (w[ijk+ii   ]-w[ijk      ])*dxi)
           + 0.125*sq((u[ijk   +kk]-u[ijk      ])*dzhikp + (w[ijk   +kk]-w[ijk-ii+kk])*dxi)
           + 0.125*sq((u[ijk+ii+kk]-u[ijk+ii   ])*dzhikp + (w[ijk+ii+kk]-w[ijk   +kk])*dxi)
           + 0.125*sq((v[ijk      ]-v[ijk   -kk])*dzhik  + (w[ijk      ]-w[ijk-jj   ])*dyi)
           + 0.125*sq((v[ijk+jj   ]-v[ijk+jj-kk])*dzhik  + (w[ijk+jj   ]-w[ijk      ])*dyi)
           + 0.125*sq((v[ijk   +kk]-v[ijk      ])*dzhikp + (w[ijk   +kk]-w[ijk-jj+kk])*dyi)
           + 0.125*sq((v[ijk+jj+kk]-v[ijk+jj   ])*dzhikp + (w[ijk+jj+kk]-w[ijk   +kk])*dyi) );
}

void mo_calc_strain2(const mo_dims* d, int surface_model, double* strain2,
                     const double* u, const double* v, const double* w,
                     const double* ugradbot, const double* vgradbot,
                     const double* z, const double* dzi, const double* dzhi)
{
    /* include/diff_kernels.h:34-142; surface_model != 0 is Surface_model::Enabled. */
    (void)z;
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int k_offset = surface_model ? 1 : 0;

    if (surface_model)
    {
        const int k = d->kstart;
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ij = i + j*jj;
                const int ijk = i + j*jj + k*kk;
                strain2[ijk] = 2.*(
                        + sq((u[ijk+ii]-u[ijk])*dxi)
                        + sq((v[ijk+jj]-v[ijk])*dyi)
                        + sq((w[ijk+kk]-w[ijk])*dzi[k])
                        + 0.125*sq((u[ijk      ]-u[ijk   -jj])*dyi  + (v[ijk      ]-v[ijk-ii   ])*dxi)
                        + 0.125*sq((u[ijk+ii   ]-u[ijk+ii-jj])*dyi  + (v[ijk+ii   ]-v[ijk      ])*dxi)
                        + 0.125*sq((u[ijk   +jj]-u[ijk      ])*dyi  + (v[ijk   +jj]-v[ijk-ii+jj])*dxi)
                        + 0.125*sq((u[ijk+ii+jj]-u[ijk+ii   ])*dyi  + (v[ijk+ii+jj]-v[ijk   +jj])*dxi)
                        + 0.5*sq(ugradbot[ij])
                        + 0.125*sq((w[ijk      ]-w[ijk-ii   ])*dxi)
                        + 0.125*sq((w[ijk+ii   ]-w[ijk      ])*dxi)
                        + 0.125*sq((w[ijk   +kk]-w[ijk-ii+kk])*dxi)
                        + 0.125*sq((w[ijk+ii+kk]-w[ijk   +kk])*dxi)
                        + 0.5*sq(vgradbot[ij])
                        + 0.125*sq((w[ijk      ]-w[ijk-jj   ])*dyi)
                        + 0.125*sq((w[ijk+jj   ]-w[ijk      ])*dyi)
                        + 0.125*sq((w[ijk   +kk]-w[ijk-jj+kk])*dyi)
                        + 0.125*sq((w[ijk+jj+kk]-w[ijk   +kk])*dyi) );
                strain2[ijk] += DSMALL;
            }
    }

    #pragma omp parallel for
    for (int k=d->kstart+k_offset; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                strain2[ijk] = strain2_core(u, v, w, ijk, ii, jj, kk, dxi, dyi, dzi[k], dzhi[k], dzhi[k+1]);
                strain2[ijk] += DSMALL;
            }
}

void mo_calc_N2(const mo_dims* d, double* N2, const double* th, const double* dzi,
                const double* thref)
{
    /* src/thermo_dry.cxx:66-79 */
    const int jj = d->icells, kk = d->ijcells;
    #pragma omp parallel for
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                N2[ijk] = GRAV/thref[k]*0.5*(th[ijk+kk] - th[ijk-kk])*dzi[k];
            }
}

static void mirror_evisc_walls(const mo_dims* d, double* evisc)
{
    /* src/diff_smag2.cxx:193-205: mirror over both walls, full planes. */
    const int kk = d->ijcells, kb = d->kstart, kt = d->kend-1;
    for (int ij=0; ij<d->ijcells; ++ij)
    {
        evisc[ij + (kb-1)*kk] = evisc[ij + kb*kk];
        evisc[ij + (kt+1)*kk] = evisc[ij + kt*kk];
    }
}

void mo_calc_evisc(const mo_dims* d, int surface_model, int sw_mason, double* evisc,
                   const double* N2, const double* bgradbot, const double* z,
                   const double* dz, const double* z0m, double cs, double tPr)
{
    /* src/diff_smag2.cxx:148-269; evisc holds S^2 on entry. */
    const int jj = d->icells, kk = d->ijcells;
    const double dx = d->dx, dy = d->dy;

    if (!surface_model)
    {
        #pragma omp parallel for
        for (int k=d->kstart; k<d->kend; ++k)
        {
            const double mlen = cs*pow(dx*dy*dz[k], 1./3.);
            const double fac = sq(mlen);
            for (int j=d->jstart; j<d->jend; ++j)
                for (int i=d->istart; i<d->iend; ++i)
                {
                    const int ijk = i + j*jj + k*kk;
                    double RitPrratio = N2[ijk] / evisc[ijk] / tPr;
                    RitPrratio = fmin2(RitPrratio, 1.-DSMALL);
                    evisc[ijk] = fac * sqrt(evisc[ijk]) * sqrt(1.-RitPrratio);
                }
        }
        mirror_evisc_walls(d, evisc);
    }
    else
    {
        const double n_mason = 2.;
        {
            const int k = d->kstart;
            const double mlen0 = cs*pow(dx*dy*dz[k], 1./3.);
            for (int j=d->jstart; j<d->jend; ++j)
                for (int i=d->istart; i<d->iend; ++i)
                {
                    const int ij = i + j*jj;
                    const int ijk = i + j*jj + k*kk;
                    double RitPrratio = bgradbot[ij] / evisc[ijk] / tPr;
                    RitPrratio = fmin2(RitPrratio, 1.-DSMALL);
                    double mlen = mlen0;
                    if (sw_mason)
                        mlen = pow(1./(1./pow(mlen0, n_mason) + 1./(pow(KAPPA*(z[k]+z0m[ij]), n_mason))), 1./n_mason);
                    evisc[ijk] = sq(mlen) * sqrt(evisc[ijk]) * sqrt(1.-RitPrratio);
                }
        }
        #pragma omp parallel for
        for (int k=d->kstart+1; k<d->kend; ++k)
        {
            const double mlen0 = cs*pow(dx*dy*dz[k], 1./3.);
            for (int j=d->jstart; j<d->jend; ++j)
                for (int i=d->istart; i<d->iend; ++i)
                {
                    const int ij = i + j*jj;
                    const int ijk = i + j*jj + k*kk;
                    double RitPrratio = N2[ijk] / evisc[ijk] / tPr;
                    RitPrratio = fmin2(RitPrratio, 1.-DSMALL);
                    double mlen = mlen0;
                    if (sw_mason)
                        mlen = pow(1./(1./pow(mlen0, n_mason) + 1./(pow(KAPPA*(z[k]+z0m[ij]), n_mason))), 1./n_mason);
                    evisc[ijk] = sq(mlen) * sqrt(evisc[ijk]) * sqrt(1.-RitPrratio);
                }
        }
    }
    mo_boundary_cyclic(d, evisc, MO_EDGE_BOTH);
}

void mo_calc_evisc_neutral(const mo_dims* d, int surface_model, int sw_mason, double* evisc,
                           const double* u, const double* v,
                           const double* z, const double* dz, const double* dzhi,
                           const double* z0m, double cs, double visc)
{
    /* src/diff_smag2.cxx:47-146; evisc holds S^2 on entry. */
    const int jj = d->icells, kk = d->ijcells;
    const double dx = d->dx, dy = d->dy, zsize = d->zsize;
    const double n_mason = 1.;
    const double A_vandriest = 26.;
    const int ks = d->kstart, ke = d->kend;

    if (!surface_model)
    {
        #pragma omp parallel for
        for (int k=ks; k<ke; ++k)
        {
            const double mlen_smag = cs*pow(dx*dy*dz[k], 1./3.);
            for (int j=d->jstart; j<d->jend; ++j)
                for (int i=d->istart; i<d->iend; ++i)
                {
                    const int ijk_bot = i + j*jj + ks*kk;
                    const int ijk_top = i + j*jj + ke*kk;
                    const double u_tau_bot = pow(
                            sq(visc*(u[ijk_bot] - u[ijk_bot-kk])*dzhi[ks])
                          + sq(visc*(v[ijk_bot] - v[ijk_bot-kk])*dzhi[ks]), 0.25);
                    const double u_tau_top = pow(
                            sq(visc*(u[ijk_top] - u[ijk_top-kk])*dzhi[ke])
                          + sq(visc*(v[ijk_top] - v[ijk_top-kk])*dzhi[ke]), 0.25);
                    const double fac_bot = 1. - exp(-(        z[k] *u_tau_bot) / (A_vandriest*visc));
                    const double fac_top = 1. - exp(-((zsize-z[k])*u_tau_top) / (A_vandriest*visc));
                    const double fac = fmin2(fac_bot, fac_top);
                    const int ijk = i + j*jj + k*kk;
                    evisc[ijk] = sq(fac * mlen_smag) * sqrt(evisc[ijk]);
                }
        }
        mirror_evisc_walls(d, evisc);
    }
    else
    {
        #pragma omp parallel for
        for (int k=ks; k<ke; ++k)
        {
            const double mlen0 = cs*pow(dx*dy*dz[k], 1./3.);
            for (int j=d->jstart; j<d->jend; ++j)
                for (int i=d->istart; i<d->iend; ++i)
                {
                    const int ij = i + j*jj;
                    const int ijk = i + j*jj + k*kk;
                    double mlen = mlen0;
                    if (sw_mason)
                        mlen = pow(1./(1./pow(mlen0, n_mason) + 1./(pow(KAPPA*(z[k]+z0m[ij]), n_mason))), 1./n_mason);
                    evisc[ijk] = sq(mlen) * sqrt(evisc[ijk]);
                }
        }
    }
    mo_boundary_cyclic(d, evisc, MO_EDGE_BOTH);
}

/* -------------------------------------------------------- smag2 diffusion */
void mo_diff_u(const mo_dims* d, int surface_model, double* ut,
               const double* u, const double* v, const double* w,
               const double* dzi, const double* dzhi, const double* evisc,
               const double* fluxbot, const double* fluxtop,
               const double* rhoref, const double* rhorefh, double visc)
{
    /* include/diff_kernels.h:144-240 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int k_offset = surface_model ? 1 : 0;
    const int ks = d->kstart, ke = d->kend;

    if (surface_model)
    {
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ij = i + j*jj;
                const int ijk = i + j*jj + ks*kk;
                const double evisce = evisc[ijk   ] + visc;
                const double eviscw = evisc[ijk-ii] + visc;
                const double eviscn = 0.25*(evisc[ijk-ii   ] + evisc[ijk   ] + evisc[ijk-ii+jj] + evisc[ijk+jj]) + visc;
                const double eviscs = 0.25*(evisc[ijk-ii-jj] + evisc[ijk-jj] + evisc[ijk-ii   ] + evisc[ijk   ]) + visc;
                const double evisct = 0.25*(evisc[ijk-ii   ] + evisc[ijk   ] + evisc[ijk-ii+kk] + evisc[ijk+kk]) + visc;
                ut[ijk] +=
                         + ( evisce*(u[ijk+ii]-u[ijk   ])*dxi
                           - eviscw*(u[ijk   ]-u[ijk-ii])*dxi ) * 2.*dxi
                         + ( eviscn*((u[ijk+jj]-u[ijk   ])*dyi + (v[ijk+jj]-v[ijk-ii+jj])*dxi)
                           - eviscs*((u[ijk   ]-u[ijk-jj])*dyi + (v[ijk   ]-v[ijk-ii   ])*dxi) ) * dyi
                         + ( rhorefh[ks+1] * evisct*((u[ijk+kk]-u[ijk   ])* dzhi[ks+1] + (w[ijk+kk]-w[ijk-ii+kk])*dxi)
                           + rhorefh[ks  ] * fluxbot[ij] ) / rhoref[ks] * dzi[ks];
            }
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ij = i + j*jj;
                const int ijk = i + j*jj + (ke-1)*kk;
                const double evisce = evisc[ijk   ] + visc;
                const double eviscw = evisc[ijk-ii] + visc;
                const double eviscn = 0.25*(evisc[ijk-ii   ] + evisc[ijk   ] + evisc[ijk-ii+jj] + evisc[ijk+jj]) + visc;
                const double eviscs = 0.25*(evisc[ijk-ii-jj] + evisc[ijk-jj] + evisc[ijk-ii   ] + evisc[ijk   ]) + visc;
                const double eviscb = 0.25*(evisc[ijk-ii-kk] + evisc[ijk-kk] + evisc[ijk-ii   ] + evisc[ijk   ]) + visc;
                ut[ijk] +=
                         + ( evisce*(u[ijk+ii]-u[ijk   ])*dxi
                           - eviscw*(u[ijk   ]-u[ijk-ii])*dxi ) * 2.*dxi
                         + ( eviscn*((u[ijk+jj]-u[ijk   ])*dyi  + (v[ijk+jj]-v[ijk-ii+jj])*dxi)
                           - eviscs*((u[ijk   ]-u[ijk-jj])*dyi  + (v[ijk   ]-v[ijk-ii   ])*dxi) ) * dyi
                         + (- rhorefh[ke  ] * fluxtop[ij]
                            - rhorefh[ke-1] * eviscb*((u[ijk   ]-u[ijk-kk])* dzhi[ke-1] + (w[ijk   ]-w[ijk-ii   ])*dxi) ) / rhoref[ke-1] * dzi[ke-1];
            }
    }

    #pragma omp parallel for
    for (int k=ks+k_offset; k<ke-k_offset; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                const double evisce = evisc[ijk   ] + visc;
                const double eviscw = evisc[ijk-ii] + visc;
                const double eviscn = 0.25*(evisc[ijk-ii   ] + evisc[ijk   ] + evisc[ijk-ii+jj] + evisc[ijk+jj]) + visc;
                const double eviscs = 0.25*(evisc[ijk-ii-jj] + evisc[ijk-jj] + evisc[ijk-ii   ] + evisc[ijk   ]) + visc;
                const double evisct = 0.25*(evisc[ijk-ii   ] + evisc[ijk   ] + evisc[ijk-ii+kk] + evisc[ijk+kk]) + visc;
                const double eviscb = 0.25*(evisc[ijk-ii-kk] + evisc[ijk-kk] + evisc[ijk-ii   ] + evisc[ijk   ]) + visc;
                ut[ijk] +=
                         + ( evisce*(u[ijk+ii]-u[ijk   ])*dxi
                           - eviscw*(u[ijk   ]-u[ijk-ii])*dxi ) * 2.*dxi
                         + ( eviscn*((u[ijk+jj]-u[ijk   ])*dyi  + (v[ijk+jj]-v[ijk-ii+jj])*dxi)
                           - eviscs*((u[ijk   ]-u[ijk-jj])*dyi  + (v[ijk   ]-v[ijk-ii   ])*dxi) ) * dyi
                         + ( rhorefh[k+1] * evisct*((u[ijk+kk]-u[ijk   ])* dzhi[k+1] + (w[ijk+kk]-w[ijk-ii+kk])*dxi)
                           - rhorefh[k  ] * eviscb*((u[ijk   ]-u[ijk-kk])* dzhi[k  ] + (w[ijk   ]-w[ijk-ii   ])*dxi) ) / rhoref[k] * dzi[k];
            }
}

void mo_diff_v(const mo_dims* d, int surface_model, double* vt,
               const double* u, const double* v, const double* w,
               const double* dzi, const double* dzhi, const double* evisc,
               const double* fluxbot, const double* fluxtop,
               const double* rhoref, const double* rhorefh, double visc)
{
    /* include/diff_kernels.h:242-340 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int k_offset = surface_model ? 1 : 0;
    const int ks = d->kstart, ke = d->kend;

    if (surface_model)
    {
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ij = i + j*jj;
                const int ijk = i + j*jj + ks*kk;
                const double evisce = 0.25*(evisc[ijk   -jj] + evisc[ijk   ] + evisc[ijk+ii-jj] + evisc[ijk+ii]) + visc;
                const double eviscw = 0.25*(evisc[ijk-ii-jj] + evisc[ijk-ii] + evisc[ijk   -jj] + evisc[ijk   ]) + visc;
                const double eviscn = evisc[ijk   ] + visc;
                const double eviscs = evisc[ijk-jj] + visc;
                const double evisct = 0.25*(evisc[ijk   -jj] + evisc[ijk   ] + evisc[ijk+kk-jj] + evisc[ijk+kk]) + visc;
                vt[ijk] +=
                         + ( evisce*((v[ijk+ii]-v[ijk   ])*dxi + (u[ijk+ii]-u[ijk+ii-jj])*dyi)
                           - eviscw*((v[ijk   ]-v[ijk-ii])*dxi + (u[ijk   ]-u[ijk   -jj])*dyi) ) * dxi
                         + ( eviscn*(v[ijk+jj]-v[ijk   ])*dyi
                           - eviscs*(v[ijk   ]-v[ijk-jj])*dyi ) * 2.*dyi
                         + ( rhorefh[ks+1] * evisct*((v[ijk+kk]-v[ijk   ])*dzhi[ks+1] + (w[ijk+kk]-w[ijk-jj+kk])*dyi)
                           + rhorefh[ks  ] * fluxbot[ij] ) / rhoref[ks] * dzi[ks];
            }
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ij = i + j*jj;
                const int ijk = i + j*jj + (ke-1)*kk;
                const double evisce = 0.25*(evisc[ijk   -jj] + evisc[ijk   ] + evisc[ijk+ii-jj] + evisc[ijk+ii]) + visc;
                const double eviscw = 0.25*(evisc[ijk-ii-jj] + evisc[ijk-ii] + evisc[ijk   -jj] + evisc[ijk   ]) + visc;
                const double eviscn = evisc[ijk   ] + visc;
                const double eviscs = evisc[ijk-jj] + visc;
                const double eviscb = 0.25*(evisc[ijk-kk-jj] + evisc[ijk-kk] + evisc[ijk   -jj] + evisc[ijk   ]) + visc;
                vt[ijk] +=
                         + ( evisce*((v[ijk+ii]-v[ijk   ])*dxi + (u[ijk+ii]-u[ijk+ii-jj])*dyi)
                           - eviscw*((v[ijk   ]-v[ijk-ii])*dxi + (u[ijk   ]-u[ijk   -jj])*dyi) ) * dxi
                         + ( eviscn*(v[ijk+jj]-v[ijk   ])*dyi
                           - eviscs*(v[ijk   ]-v[ijk-jj])*dyi ) * 2.*dyi
                         + (- rhorefh[ke  ] * fluxtop[ij]
                            - rhorefh[ke-1] * eviscb*((v[ijk   ]-v[ijk-kk])*dzhi[ke-1] + (w[ijk   ]-w[ijk-jj   ])*dyi) ) / rhoref[ke-1] * dzi[ke-1];
            }
    }

    #pragma omp parallel for
    for (int k=ks+k_offset; k<ke-k_offset; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                const double evisce = 0.25*(evisc[ijk   -jj] + evisc[ijk   ] + evisc[ijk+ii-jj] + evisc[ijk+ii]) + visc;
                const double eviscw = 0.25*(evisc[ijk-ii-jj] + evisc[ijk-ii] + evisc[ijk   -jj] + evisc[ijk   ]) + visc;
                const double eviscn = evisc[ijk   ] + visc;
                const double eviscs = evisc[ijk-jj] + visc;
                const double evisct = 0.25*(evisc[ijk   -jj] + evisc[ijk   ] + evisc[ijk+kk-jj] + evisc[ijk+kk]) + visc;
                const double eviscb = 0.25*(evisc[ijk-kk-jj] + evisc[ijk-kk] + evisc[ijk   -jj] + evisc[ijk   ]) + visc;
                vt[ijk] +=
                         + ( evisce*((v[ijk+ii]-v[ijk   ])*dxi + (u[ijk+ii]-u[ijk+ii-jj])*dyi)
                           - eviscw*((v[ijk   ]-v[ijk-ii])*dxi + (u[ijk   ]-u[ijk   -jj])*dyi) ) * dxi
                         + ( eviscn*(v[ijk+jj]-v[ijk   ])*dyi
                           - eviscs*(v[ijk   ]-v[ijk-jj])*dyi ) * 2.*dyi
                         + ( rhorefh[k+1] * evisct*((v[ijk+kk]-v[ijk   ])*dzhi[k+1] + (w[ijk+kk]-w[ijk-jj+kk])*dyi)
                           - rhorefh[k  ] * eviscb*((v[ijk   ]-v[ijk-kk])*dzhi[k  ] + (w[ijk   ]-w[ijk-jj   ])*dyi) ) / rhoref[k] * dzi[k];
            }
}

void mo_diff_w(const mo_dims* d, double* wt,
               const double* u, const double* v, const double* w,
               const double* dzi, const double* dzhi, const double* evisc,
               const double* rhoref, const double* rhorefh, double visc)
{
    /* include/diff_kernels.h:342-390 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    #pragma omp parallel for
    for (int k=d->kstart+1; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                const double evisce = 0.25*(evisc[ijk   -kk] + evisc[ijk   ] + evisc[ijk+ii-kk] + evisc[ijk+ii]) + visc;
                const double eviscw = 0.25*(evisc[ijk-ii-kk] + evisc[ijk-ii] + evisc[ijk   -kk] + evisc[ijk   ]) + visc;
                const double eviscn = 0.25*(evisc[ijk   -kk] + evisc[ijk   ] + evisc[ijk+jj-kk] + evisc[ijk+jj]) + visc;
                const double eviscs = 0.25*(evisc[ijk-jj-kk] + evisc[ijk-jj] + evisc[ijk   -kk] + evisc[ijk   ]) + visc;
                const double evisct = evisc[ijk   ] + visc;
                const double eviscb = evisc[ijk-kk] + visc;
                wt[ijk] +=
                         + ( evisce*((w[ijk+ii]-w[ijk   ])*dxi + (u[ijk+ii]-u[ijk+ii-kk])*dzhi[k])
                           - eviscw*((w[ijk   ]-w[ijk-ii])*dxi + (u[ijk   ]-u[ijk   -kk])*dzhi[k]) ) * dxi
                         + ( eviscn*((w[ijk+jj]-w[ijk   ])*dyi + (v[ijk+jj]-v[ijk+jj-kk])*dzhi[k])
                           - eviscs*((w[ijk   ]-w[ijk-jj])*dyi + (v[ijk   ]-v[ijk   -kk])*dzhi[k]) ) * dyi
                         + ( rhoref[k  ] * evisct*(w[ijk+kk]-w[ijk   ])*dzi[k  ]
                           - rhoref[k-1] * eviscb*(w[ijk   ]-w[ijk-kk])*dzi[k-1] ) / rhorefh[k] * 2.*dzhi[k];
            }
}

void mo_diff_c(const mo_dims* d, int surface_model, double* at, const double* a,
               const double* dzi, const double* dzhi, const double* evisc,
               const double* fluxbot, const double* fluxtop,
               const double* rhoref, const double* rhorefh, double tPr, double visc)
{
    /* include/diff_kernels.h:392-484 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxidxi = 1./(d->dx*d->dx), dyidyi = 1./(d->dy*d->dy);
    const double tPr_i = 1./tPr;
    const int k_offset = surface_model ? 1 : 0;
    const int ks = d->kstart, ke = d->kend;

    if (surface_model)
    {
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ij = i + j*jj;
                const int ijk = i + j*jj + ks*kk;
                const double evisce = 0.5*(evisc[ijk   ]+evisc[ijk+ii]) * tPr_i + visc;
                const double eviscw = 0.5*(evisc[ijk-ii]+evisc[ijk   ]) * tPr_i + visc;
                const double eviscn = 0.5*(evisc[ijk   ]+evisc[ijk+jj]) * tPr_i + visc;
                const double eviscs = 0.5*(evisc[ijk-jj]+evisc[ijk   ]) * tPr_i + visc;
                const double evisct = 0.5*(evisc[ijk   ]+evisc[ijk+kk]) * tPr_i + visc;
                at[ijk] +=
                         + ( evisce*(a[ijk+ii]-a[ijk   ])
                           - eviscw*(a[ijk   ]-a[ijk-ii]) ) * dxidxi
                         + ( eviscn*(a[ijk+jj]-a[ijk   ])
                           - eviscs*(a[ijk   ]-a[ijk-jj]) ) * dyidyi
                         + ( rhorefh[ks+1] * evisct*(a[ijk+kk]-a[ijk   ])*dzhi[ks+1]
                           + rhorefh[ks  ] * fluxbot[ij] ) / rhoref[ks] * dzi[ks];
            }
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ij = i + j*jj;
                const int ijk = i + j*jj + (ke-1)*kk;
                const double evisce = 0.5*(evisc[ijk   ]+evisc[ijk+ii]) * tPr_i + visc;
                const double eviscw = 0.5*(evisc[ijk-ii]+evisc[ijk   ]) * tPr_i + visc;
                const double eviscn = 0.5*(evisc[ijk   ]+evisc[ijk+jj]) * tPr_i + visc;
                const double eviscs = 0.5*(evisc[ijk-jj]+evisc[ijk   ]) * tPr_i + visc;
                const double eviscb = 0.5*(evisc[ijk-kk]+evisc[ijk   ]) * tPr_i + visc;
                at[ijk] +=
                         + ( evisce*(a[ijk+ii]-a[ijk   ])
                           - eviscw*(a[ijk   ]-a[ijk-ii]) ) * dxidxi
                         + ( eviscn*(a[ijk+jj]-a[ijk   ])
                           - eviscs*(a[ijk   ]-a[ijk-jj]) ) * dyidyi
                         + (-rhorefh[ke  ] * fluxtop[ij]
                           - rhorefh[ke-1] * eviscb*(a[ijk   ]-a[ijk-kk])*dzhi[ke-1] ) / rhoref[ke-1] * dzi[ke-1];
            }
    }

    #pragma omp parallel for
    for (int k=ks+k_offset; k<ke-k_offset; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                const double evisce = 0.5*(evisc[ijk   ]+evisc[ijk+ii]) * tPr_i + visc;
                const double eviscw = 0.5*(evisc[ijk-ii]+evisc[ijk   ]) * tPr_i + visc;
                const double eviscn = 0.5*(evisc[ijk   ]+evisc[ijk+jj]) * tPr_i + visc;
                const double eviscs = 0.5*(evisc[ijk-jj]+evisc[ijk   ]) * tPr_i + visc;
                const double evisct = 0.5*(evisc[ijk   ]+evisc[ijk+kk]) * tPr_i + visc;
                const double eviscb = 0.5*(evisc[ijk-kk]+evisc[ijk   ]) * tPr_i + visc;
                at[ijk] +=
                         + ( evisce*(a[ijk+ii]-a[ijk   ])
                           - eviscw*(a[ijk   ]-a[ijk-ii]) ) * dxidxi
                         + ( eviscn*(a[ijk+jj]-a[ijk   ])
                           - eviscs*(a[ijk   ]-a[ijk-jj]) ) * dyidyi
                         + ( rhorefh[k+1] * evisct*(a[ijk+kk]-a[ijk   ])*dzhi[k+1]
                           - rhorefh[k  ] * eviscb*(a[ijk   ]-a[ijk-kk])*dzhi[k]  ) / rhoref[k] * dzi[k];
            }
}

double mo_calc_dnmul(const mo_dims* d, const double* evisc, const double* dzi, double tPr)
{
    /* include/diff_kernels.h:486-511 */
    const int jj = d->icells, kk = d->ijcells;
    const double dxidxi = 1./(d->dx*d->dx), dyidyi = 1./(d->dy*d->dy);
    const double tPrfac_i = 1./fmin2(1., tPr);
    double dnmul = 0;
    #pragma omp parallel for reduction(max:dnmul)
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                dnmul = fmax2(dnmul, fabs(evisc[ijk]*tPrfac_i*(dxidxi + dyidyi + dzi[k]*dzi[k])));
            }
    return dnmul;
}

/* ------------------------------------------------------------ advec_2i5 */
double mo_advec_2i5_cfl(const mo_dims* d, const double* u, const double* v, const double* w,
                        const double* dzi, double dt)
{
    /* src/advec_2i5.cxx:59-148 */
    const int ii1 = 1, ii2 = 2, ii3 = 3;
    const int jj1 = d->icells, jj2 = 2*jj1, jj3 = 3*jj1;
    const int kk1 = d->ijcells, kk2 = 2*kk1, kk3 = 3*kk1;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int ks = d->kstart, ke = d->kend;
    double cfl = 0;

    #pragma omp parallel for reduction(max:cfl)
    for (int k=ks; k<ke; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                double wpart;
                if (k == ks || k == ke-1)
                    wpart = fabs(i2(w[ijk], w[ijk+kk1]))*dzi[k];
                else if (k == ks+1 || k == ke-2)
                    wpart = fabs(i4ws(w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2]))*dzi[k];
                else
                    wpart = fabs(i6ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2], w[ijk+kk3]))*dzi[k];
                cfl = fmax2(cfl, fabs(i6ws(u[ijk-ii2], u[ijk-ii1], u[ijk], u[ijk+ii1], u[ijk+ii2], u[ijk+ii3]))*dxi
                               + fabs(i6ws(v[ijk-jj2], v[ijk-jj1], v[ijk], v[ijk+jj1], v[ijk+jj2], v[ijk+jj3]))*dyi
                               + wpart);
            }
    return cfl*dt;
}

/* Advective flux-divergence terms along one horizontal direction
   (src/advec_2i5.cxx:195-213): fp/fm are the advecting velocities at the
   upper/lower faces, s is the advected field and st its stride.  The two
   returned terms are the centred part and the upwind part, kept separate so
   that the caller sums all four horizontal terms left to right, as the
   reference expression does. */
typedef struct { double c, u; } hterms;
static inline hterms hflux(double fp, double fm, const double* s, int ijk, int st, double di)
{
    hterms t;
    t.c = - ( fp * i6ws(s[ijk-2*st], s[ijk-st], s[ijk], s[ijk+st], s[ijk+2*st], s[ijk+3*st])
            - fm * i6ws(s[ijk-3*st], s[ijk-2*st], s[ijk-st], s[ijk], s[ijk+st], s[ijk+2*st]) ) * di;
    t.u =   ( fabs(fp) * i5ws(s[ijk-2*st], s[ijk-st], s[ijk], s[ijk+st], s[ijk+2*st], s[ijk+3*st])
            - fabs(fm) * i5ws(s[ijk-3*st], s[ijk-2*st], s[ijk-st], s[ijk], s[ijk+st], s[ijk+2*st]) ) * di;
    return t;
}

static inline double hsum(hterms x, hterms y) { return ((x.c + x.u) + y.c) + y.u; }

/* Vertical flux divergence of one field with the 2i5 wall-reduced order
   (src/advec_2i5.cxx:216-299 for u/v, 648-727 for s): wp/wm are the
   velocities at the top/bottom faces (already interpolated where needed),
   rp/rm the face densities, rc the cell density and dzc the inverse cell
   height.  lvl encodes the level class: 0=ks, 1=ks+1, 2=ks+2, 3=interior,
   4=ke-3, 5=ke-2, 6=ke-1. */
static inline double vflux_c(int lvl, double wp, double wm, double rp, double rm, double rc, double dzc,
                             const double* s, int ijk, int kk)
{
    const int kk1 = kk, kk2 = 2*kk, kk3 = 3*kk;
    switch (lvl)
    {
        case 0:
            return - ( rp * wp * i2(s[ijk], s[ijk+kk1]) ) / rc * dzc;
        case 1:
            return - ( rp * wp * i4ws(s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2])
                     - rm * wm * i2(s[ijk-kk1], s[ijk]) ) / rc * dzc
                   + ( rp * fabs(wp) * i3ws(s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2]) ) / rc * dzc;
        case 2:
            return - ( rp * wp * i6ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2], s[ijk+kk3])
                     - rm * wm * i4ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1]) ) / rc * dzc
                   + ( rp * fabs(wp) * i5ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2], s[ijk+kk3])
                     - rm * fabs(wm) * i3ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1]) ) / rc * dzc;
        case 3:
            return - ( rp * wp * i6ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2], s[ijk+kk3])
                     - rm * wm * i6ws(s[ijk-kk3], s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2]) ) / rc * dzc
                   + ( rp * fabs(wp) * i5ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2], s[ijk+kk3])
                     - rm * fabs(wm) * i5ws(s[ijk-kk3], s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2]) ) / rc * dzc;
        case 4:
            return - ( rp * wp * i4ws(s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2])
                     - rm * wm * i6ws(s[ijk-kk3], s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2]) ) / rc * dzc
                   + ( rp * fabs(wp) * i3ws(s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2])
                     - rm * fabs(wm) * i5ws(s[ijk-kk3], s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2]) ) / rc * dzc;
        case 5:
            return - ( rp * wp * i2(s[ijk], s[ijk+kk1])
                     - rm * wm * i4ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1]) ) / rc * dzc
                   - ( rm * fabs(wm) * i3ws(s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1]) ) / rc * dzc;
        default:
            return - ( -rm * wm * i2(s[ijk-kk1], s[ijk]) ) / rc * dzc;
    }
}

static inline int level_class(int k, int ks, int ke)
{
    if (k == ks) return 0;
    if (k == ks+1) return 1;
    if (k == ks+2) return 2;
    if (k == ke-1) return 6;
    if (k == ke-2) return 5;
    if (k == ke-3) return 4;
    return 3;
}

void mo_advec_2i5_u(const mo_dims* d, double* ut, const double* u, const double* v,
                    const double* w, const double* dzi, const double* rhoref, const double* rhorefh)
{
    /* src/advec_2i5.cxx:150-300: horizontal pass, then vertical pass. */
    const int ii1 = 1, jj1 = d->icells, kk1 = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int ks = d->kstart, ke = d->kend;

    #pragma omp parallel for
    for (int k=ks; k<ke; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                ut[ijk] += hsum(hflux(i2(u[ijk], u[ijk+ii1]), i2(u[ijk-ii1], u[ijk]), u, ijk, ii1, dxi),
                              hflux(i2(v[ijk-ii1+jj1], v[ijk+jj1]), i2(v[ijk-ii1], v[ijk]), u, ijk, jj1, dyi));
            }

    #pragma omp parallel for
    for (int k=ks; k<ke; ++k)
    {
        const int lvl = level_class(k, ks, ke);
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                ut[ijk] += vflux_c(lvl, i2(w[ijk-ii1+kk1], w[ijk+kk1]), i2(w[ijk-ii1], w[ijk]),
                                   rhorefh[k+1], rhorefh[k], rhoref[k], dzi[k], u, ijk, kk1);
            }
    }
}

void mo_advec_2i5_v(const mo_dims* d, double* vt, const double* u, const double* v,
                    const double* w, const double* dzi, const double* rhoref, const double* rhorefh)
{
    /* src/advec_2i5.cxx:302-450 */
    const int ii1 = 1, jj1 = d->icells, kk1 = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int ks = d->kstart, ke = d->kend;

    #pragma omp parallel for
    for (int k=ks; k<ke; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                vt[ijk] += hsum(hflux(i2(u[ijk+ii1-jj1], u[ijk+ii1]), i2(u[ijk-jj1], u[ijk]), v, ijk, ii1, dxi),
                              hflux(i2(v[ijk], v[ijk+jj1]), i2(v[ijk-jj1], v[ijk]), v, ijk, jj1, dyi));
            }

    #pragma omp parallel for
    for (int k=ks; k<ke; ++k)
    {
        const int lvl = level_class(k, ks, ke);
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                vt[ijk] += vflux_c(lvl, i2(w[ijk-jj1+kk1], w[ijk+kk1]), i2(w[ijk-jj1], w[ijk]),
                                   rhorefh[k+1], rhorefh[k], rhoref[k], dzi[k], v, ijk, kk1);
            }
    }
}

void mo_advec_2i5_w(const mo_dims* d, double* wt, const double* u, const double* v,
                    const double* w, const double* dzhi, const double* rhoref, const double* rhorefh)
{
    /* src/advec_2i5.cxx:452-579 */
    const int ii1 = 1, jj1 = d->icells, kk1 = d->ijcells, kk2 = 2*kk1, kk3 = 3*kk1;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int ks = d->kstart, ke = d->kend;

    #pragma omp parallel for
    for (int k=ks+1; k<ke; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                wt[ijk] += hsum(hflux(i2(u[ijk+ii1-kk1], u[ijk+ii1]), i2(u[ijk-kk1], u[ijk]), w, ijk, ii1, dxi),
                              hflux(i2(v[ijk+jj1-kk1], v[ijk+jj1]), i2(v[ijk-kk1], v[ijk]), w, ijk, jj1, dyi));
            }

    #pragma omp parallel for
    for (int k=ks+1; k<ke; ++k)
    {
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                const double wp = i2(w[ijk], w[ijk+kk1]);
                const double wm = i2(w[ijk-kk1], w[ijk]);
                const double rp = rhoref[k], rm = rhoref[k-1], rc = rhorefh[k], dzc = dzhi[k];
                double t;
                if (k == ks+1)
                    t = - ( rp * wp * i4ws(w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2])
                          - rm * wm * i2(w[ijk-kk1], w[ijk]) ) / rc * dzc
                        + ( rp * fabs(wp) * i3ws(w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2]) ) / rc * dzc;
                else if (k == ks+2)
                    t = - ( rp * wp * i6ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2], w[ijk+kk3])
                          - rm * wm * i4ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1]) ) / rc * dzc
                        + ( rp * fabs(wp) * i5ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2], w[ijk+kk3])
                          - rm * fabs(wm) * i3ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1]) ) / rc * dzc;
                else if (k == ke-2)
                    t = - ( rp * wp * i4ws(w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2])
                          - rm * wm * i6ws(w[ijk-kk3], w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2]) ) / rc * dzc
                        + ( rp * fabs(wp) * i3ws(w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2])
                          - rm * fabs(wm) * i5ws(w[ijk-kk3], w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2]) ) / rc * dzc;
                else if (k == ke-1)
                    t = - ( rp * wp * i2(w[ijk], w[ijk+kk1])
                          - rm * wm * i4ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1]) ) / rc * dzc
                        - ( rm * fabs(wm) * i3ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1]) ) / rc * dzc;
                else
                    t = - ( rp * wp * i6ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2], w[ijk+kk3])
                          - rm * wm * i6ws(w[ijk-kk3], w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2]) ) / rc * dzc
                        + ( rp * fabs(wp) * i5ws(w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2], w[ijk+kk3])
                          - rm * fabs(wm) * i5ws(w[ijk-kk3], w[ijk-kk2], w[ijk-kk1], w[ijk], w[ijk+kk1], w[ijk+kk2]) ) / rc * dzc;
                wt[ijk] += t;
            }
    }
}

void mo_advec_2i5_s(const mo_dims* d, double* st, const double* s, const double* u,
                    const double* v, const double* w, const double* dzi,
                    const double* rhoref, const double* rhorefh)
{
    /* src/advec_2i5.cxx:581-728 */
    const int ii1 = 1, jj1 = d->icells, kk1 = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int ks = d->kstart, ke = d->kend;

    #pragma omp parallel for
    for (int k=ks; k<ke; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                st[ijk] += hsum(hflux(u[ijk+ii1], u[ijk], s, ijk, ii1, dxi),
                              hflux(v[ijk+jj1], v[ijk], s, ijk, jj1, dyi));
            }

    #pragma omp parallel for
    for (int k=ks; k<ke; ++k)
    {
        const int lvl = level_class(k, ks, ke);
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                st[ijk] += vflux_c(lvl, w[ijk+kk1], w[ijk], rhorefh[k+1], rhorefh[k], rhoref[k], dzi[k], s, ijk, kk1);
            }
    }
}

/* -------------------------------------------------------------- advec_2 */
/* Koren (1993) limited face flux, include/advec_monotonic.h:29-94: the
   general form, and the bottom/top variants that take the plain upwind value
   for flow out of / into the wall-adjacent face. */
static inline double koren_phi(double num, double dsm)
{
    const double eps = DBL_EPSILON;
    const double denom = copysign(1., dsm) * fmax2(fabs(dsm), eps);
    const double two_r = 2. * num / denom;
    return fmax2(0., fmin2(two_r, fmin2((1./3.)*(1. + two_r), 2.)));
}
static inline double flux_lim(double u, double sm2, double sm1, double sp1, double sp2)
{
    if (u >= 0.)
    {
        const double phi = koren_phi(sp1 - sm1, sm1 - sm2);
        return u*(sm1 + 0.5*phi*(sm1 - sm2));
    }
    const double phi = koren_phi(sm1 - sp1, sp1 - sp2);
    return u*(sp1 + 0.5*phi*(sp1 - sp2));
}
static inline double flux_lim_bot(double u, double sm2, double sm1, double sp1, double sp2)
{
    if (u >= 0.)
        return u*sm1;
    const double phi = koren_phi(sm1 - sp1, sp1 - sp2);
    return u*(sp1 + 0.5*phi*(sp1 - sp2));
}
static inline double flux_lim_top(double u, double sm2, double sm1, double sp1, double sp2)
{
    if (u >= 0.)
    {
        const double phi = koren_phi(sp1 - sm1, sm1 - sm2);
        return u*(sm1 + 0.5*phi*(sm1 - sm2));
    }
    return u*sp1;
}

void mo_advec_2i5_s_lim(const mo_dims* d, double* st, const double* s, const double* u,
                        const double* v, const double* w, const double* dzi,
                        const double* rhoref, const double* rhorefh)
{
    /* include/advec_monotonic.h:97-201: interior levels, then kstart,
       kstart+1, kend-2, kend-1 with the wall variants of the vertical flux. */
    const int ii1 = 1, ii2 = 2, jj1 = d->icells, jj2 = 2*d->icells, kk1 = d->ijcells, kk2 = 2*d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    const int ks = d->kstart, ke = d->kend;
    #pragma omp parallel for
    for (int k=ks; k<ke; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj1 + k*kk1;
                const double fx = ( flux_lim(u[ijk+ii1], s[ijk-ii1], s[ijk    ], s[ijk+ii1], s[ijk+ii2])
                                  - flux_lim(u[ijk    ], s[ijk-ii2], s[ijk-ii1], s[ijk    ], s[ijk+ii1]) );
                const double fy = ( flux_lim(v[ijk+jj1], s[ijk-jj1], s[ijk    ], s[ijk+jj1], s[ijk+jj2])
                                  - flux_lim(v[ijk    ], s[ijk-jj2], s[ijk-jj1], s[ijk    ], s[ijk+jj1]) );
                double fz;
                if (k == ks)
                    fz = ( rhorefh[k+1] * flux_lim_bot(w[ijk+kk1], s[ijk-kk1], s[ijk], s[ijk+kk1], s[ijk+kk2]) );
                else if (k == ks+1)
                    fz = ( rhorefh[k+1] * flux_lim    (w[ijk+kk1], s[ijk-kk1], s[ijk    ], s[ijk+kk1], s[ijk+kk2])
                         - rhorefh[k  ] * flux_lim_bot(w[ijk    ], s[ijk-kk2], s[ijk-kk1], s[ijk    ], s[ijk+kk1]) );
                else if (k == ke-2)
                    fz = ( rhorefh[k+1] * flux_lim_top(w[ijk+kk1], s[ijk-kk1], s[ijk    ], s[ijk+kk1], s[ijk+kk2])
                         - rhorefh[k  ] * flux_lim    (w[ijk    ], s[ijk-kk2], s[ijk-kk1], s[ijk    ], s[ijk+kk1]) );
                else if (k == ke-1)
                    fz = ( - rhorefh[k] * flux_lim_top(w[ijk], s[ijk-kk2], s[ijk-kk1], s[ijk], s[ijk+kk1]) );
                else
                    fz = ( rhorefh[k+1] * flux_lim(w[ijk+kk1], s[ijk-kk1], s[ijk    ], s[ijk+kk1], s[ijk+kk2])
                         - rhorefh[k  ] * flux_lim(w[ijk    ], s[ijk-kk2], s[ijk-kk1], s[ijk    ], s[ijk+kk1]) );
                st[ijk] += - fx * dxi - fy * dyi - fz / rhoref[k] * dzi[k];
            }
}

double mo_advec_2_cfl(const mo_dims* d, const double* u, const double* v, const double* w,
                      const double* dzi, double dt)
{
    /* src/advec_2.cxx:50-80 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    double cfl = 0;
    #pragma omp parallel for reduction(max:cfl)
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                cfl = fmax2(cfl, fabs(i2(u[ijk], u[ijk+ii]))*dxi + fabs(i2(v[ijk], v[ijk+jj]))*dyi
                               + fabs(i2(w[ijk], w[ijk+kk]))*dzi[k]);
            }
    return cfl*dt;
}

void mo_advec_2_u(const mo_dims* d, double* ut, const double* u, const double* v,
                  const double* w, const double* dzi, const double* rhoref, const double* rhorefh)
{
    /* src/advec_2.cxx:82-108 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    #pragma omp parallel for
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                ut[ijk] +=
                         - ( i2(u[ijk   ], u[ijk+ii]) * i2(u[ijk   ], u[ijk+ii])
                           - i2(u[ijk-ii], u[ijk   ]) * i2(u[ijk-ii], u[ijk   ]) ) * dxi
                         - ( i2(v[ijk-ii+jj], v[ijk+jj]) * i2(u[ijk   ], u[ijk+jj])
                           - i2(v[ijk-ii   ], v[ijk   ]) * i2(u[ijk-jj], u[ijk   ]) ) * dyi
                         - ( rhorefh[k+1] * i2(w[ijk-ii+kk], w[ijk+kk]) * i2(u[ijk   ], u[ijk+kk])
                           - rhorefh[k  ] * i2(w[ijk-ii   ], w[ijk   ]) * i2(u[ijk-kk], u[ijk   ]) ) / rhoref[k] * dzi[k];
            }
}

void mo_advec_2_v(const mo_dims* d, double* vt, const double* u, const double* v,
                  const double* w, const double* dzi, const double* rhoref, const double* rhorefh)
{
    /* src/advec_2.cxx:110-136 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    #pragma omp parallel for
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                vt[ijk] +=
                         - ( i2(u[ijk+ii-jj], u[ijk+ii]) * i2(v[ijk   ], v[ijk+ii])
                           - i2(u[ijk   -jj], u[ijk   ]) * i2(v[ijk-ii], v[ijk   ]) ) * dxi
                         - ( i2(v[ijk   ], v[ijk+jj]) * i2(v[ijk   ], v[ijk+jj])
                           - i2(v[ijk-jj], v[ijk   ]) * i2(v[ijk-jj], v[ijk   ]) ) * dyi
                         - ( rhorefh[k+1] * i2(w[ijk-jj+kk], w[ijk+kk]) * i2(v[ijk   ], v[ijk+kk])
                           - rhorefh[k  ] * i2(w[ijk-jj   ], w[ijk   ]) * i2(v[ijk-kk], v[ijk   ]) ) / rhoref[k] * dzi[k];
            }
}

void mo_advec_2_w(const mo_dims* d, double* wt, const double* u, const double* v,
                  const double* w, const double* dzhi, const double* rhoref, const double* rhorefh)
{
    /* src/advec_2.cxx:138-164 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    #pragma omp parallel for
    for (int k=d->kstart+1; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                wt[ijk] +=
                         - ( i2(u[ijk+ii-kk], u[ijk+ii]) * i2(w[ijk   ], w[ijk+ii])
                           - i2(u[ijk   -kk], u[ijk   ]) * i2(w[ijk-ii], w[ijk   ]) ) * dxi
                         - ( i2(v[ijk+jj-kk], v[ijk+jj]) * i2(w[ijk   ], w[ijk+jj])
                           - i2(v[ijk   -kk], v[ijk   ]) * i2(w[ijk-jj], w[ijk   ]) ) * dyi
                         - ( rhoref[k  ] * i2(w[ijk   ], w[ijk+kk]) * i2(w[ijk   ], w[ijk+kk])
                           - rhoref[k-1] * i2(w[ijk-kk], w[ijk   ]) * i2(w[ijk-kk], w[ijk   ]) ) / rhorefh[k] * dzhi[k];
            }
}

void mo_advec_2_s(const mo_dims* d, double* st, const double* s, const double* u,
                  const double* v, const double* w, const double* dzi,
                  const double* rhoref, const double* rhorefh)
{
    /* src/advec_2.cxx:166-202 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxi = 1./d->dx, dyi = 1./d->dy;
    #pragma omp parallel for
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                st[ijk] +=
                         - ( u[ijk+ii] * i2(s[ijk   ], s[ijk+ii])
                           - u[ijk   ] * i2(s[ijk-ii], s[ijk   ]) ) * dxi
                         - ( v[ijk+jj] * i2(s[ijk   ], s[ijk+jj])
                           - v[ijk   ] * i2(s[ijk-jj], s[ijk   ]) ) * dyi
                         - ( rhorefh[k+1] * w[ijk+kk] * i2(s[ijk   ], s[ijk+kk])
                           - rhorefh[k  ] * w[ijk   ] * i2(s[ijk-kk], s[ijk   ]) ) / rhoref[k] * dzi[k];
            }
}

/* --------------------------------------------------------------- diff_2 */
void mo_diff_2_c(const mo_dims* d, double* at, const double* a, double visc,
                 const double* dzi, const double* dzhi)
{
    /* src/diff_2.cxx:39-61 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxidxi = 1/(d->dx*d->dx), dyidyi = 1/(d->dy*d->dy);
    #pragma omp parallel for
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                at[ijk] += visc * (
                        + ( (a[ijk+ii] - a[ijk   ])
                          - (a[ijk   ] - a[ijk-ii]) ) * dxidxi
                        + ( (a[ijk+jj] - a[ijk   ])
                          - (a[ijk   ] - a[ijk-jj]) ) * dyidyi
                        + ( (a[ijk+kk] - a[ijk   ]) * dzhi[k+1]
                          - (a[ijk   ] - a[ijk-kk]) * dzhi[k]   ) * dzi[k] );
            }
}

void mo_diff_2_w(const mo_dims* d, double* wt, const double* w, double visc,
                 const double* dzi, const double* dzhi)
{
    /* src/diff_2.cxx:63-85 */
    const int ii = 1, jj = d->icells, kk = d->ijcells;
    const double dxidxi = 1/(d->dx*d->dx), dyidyi = 1/(d->dy*d->dy);
    #pragma omp parallel for
    for (int k=d->kstart+1; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                wt[ijk] += visc * (
                        + ( (w[ijk+ii] - w[ijk   ])
                          - (w[ijk   ] - w[ijk-ii]) ) * dxidxi
                        + ( (w[ijk+jj] - w[ijk   ])
                          - (w[ijk   ] - w[ijk-jj]) ) * dyidyi
                        + ( (w[ijk+kk] - w[ijk   ]) * dzi[k]
                          - (w[ijk   ] - w[ijk-kk]) * dzi[k-1] ) * dzhi[k] );
            }
}

/* -------------------------------------------------------------- thermo */
void mo_buoyancy_tend_2nd(const mo_dims* d, double* wt, const double* th, const double* threfh)
{
    /* src/thermo_dry.cxx:163-180 */
    const int jj = d->icells, kk = d->ijcells;
    #pragma omp parallel for
    for (int k=d->kstart+1; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                wt[ijk] += GRAV/threfh[k] * (i2(th[ijk-kk], th[ijk]) - threfh[k]);
            }
}

/* --------------------------------------------------------------- rk3 */
void mo_rk3(const mo_dims* d, double* a, double* at, int substep, double dt)
{
    /* src/timeloop.cxx:251-286 */
    static const double cA[] = {0., -5./9., -153./128.};
    static const double cB[] = {1./3., 15./16., 8./15.};
    const int jj = d->icells, kk = d->ijcells;
    #pragma omp parallel for
    for (int k=d->kstart; k<d->kend; ++k)
        for (int j=d->jstart; j<d->jend; ++j)
            for (int i=d->istart; i<d->iend; ++i)
            {
                const int ijk = i + j*jj + k*kk;
                a[ijk] += cB[substep]*dt*at[ijk];
            }
    const int substepn = (substep+1) % 3;
    if (substepn == 0)
    {
        for (int n=0; n<d->ncells; ++n)
            at[n] = 0.;
    }
    else
    {
        #pragma omp parallel for
        for (int k=d->kstart; k<d->kend; ++k)
            for (int j=d->jstart; j<d->jend; ++j)
                for (int i=d->istart; i<d->iend; ++i)
                    at[i + j*jj + k*kk] *= cA[substepn];
    }
}

double mo_rk3_subdt(double dt, int substep)
{
    /* src/timeloop.cxx:288-293 */
    static const double cB[] = {1./3., 15./16., 8./15.};
    return cB[substep]*dt;
}

unsigned long mo_convert_to_itime(double t)
{
    /* include/timeloop.h:142-146 */
    return (unsigned long)(1e9 * t + 0.5);
}
