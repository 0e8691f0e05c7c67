/*
 * fwi_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A CPU restatement, written for this repository, of the FWILab2d C v1.2.6
 * acoustic FWI hot path.  Every function names the reference lines it follows
 * (paths relative to FWILab2d/FWILab2dCLangV1.2.6/).  Fields are flat
 * column-major arrays, index ix*Nz + iz (z fastest), which is the memory order
 * of the reference's alloc2float(Nz, Nx) blocks.
 *
 * REAL selects the build: float reproduces the reference's own FP32 fields with
 * its FP64 accumulators and FP64 source term; double is the FP64 oracle.
 *
 * Pinning (see tests/): stencil coefficients and their consistency sums,
 * polynomial and brute-force FP64 stencil checks, MEAL closed forms, checkpoint
 * plan vs an independent schedule simulation and vs the store-everything
 * gradient, Ricker/fpeak against the reference's shipped wavelet file, closed
 * forms for the optimisation helpers, numpy FFT for the trace filters.
 */
#include "fwi_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#define IS_F64 (sizeof(real) == sizeof(double))
#define MST 5                        /* propagation.h:21 */
#define PI_D 3.141592653589793       /* Seismic Unix PI */
#define SQR(x) ((x) * (x))

#define EXPORT __attribute__((visibility("default")))

#if defined(ORACLE_F64)
#define RPOW pow
#define REXP exp
#define RLOG log
#define RLOG10 log10
#define RSQRT sqrt
#define RCOS cos
#else
#define RPOW powf
#define REXP expf
#define RLOG logf
#define RLOG10 log10f
#define RSQRT sqrtf
#define RCOS cosf
#endif

static void die(const char *msg)
{
    fprintf(stderr, "oracle: %s\n", msg);
    exit(1);
}

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    if (!p)
        die("out of memory");
    return p;
}

EXPORT int oracle_is_f64(void) { return IS_F64 ? 1 : 0; }

/* ------------------------------------------------------------------------- */
/* Verbose debug files (propagation.c:419-441, 597-601, 717-723, 846-852,
 * 917-943, 996-1005; write_bin, dataio.c:277-304).  The engine's file names
 * use a shot index the reference never assigns (propagation.c:399): the
 * oracle passes the shot loop's index (vb_shot), and 0 for a direct engine call. */
static int vb_level = 0, vb_iter = 0, vb_shot = 0;
static int vb_shot_interval = 20, vb_snap_interval = 1000;

static int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v ? atoi(v) : dflt;
}

static void vb_write(const char *name, const real *x, size_t n)
{
    FILE *fp = fopen(name, "wb");
    if (!fp)
        die("write_bin: cannot open file");
    for (size_t i = 0; i < n; i++)
    {
        float v = (float)x[i];
        if (fwrite(&v, sizeof(float), 1, fp) != 1)
            die("write_bin: short write");
    }
    fclose(fp);
}

static void vb_write_f(const char *name, const float *x, size_t n)
{
    FILE *fp = fopen(name, "wb");
    if (!fp)
        die("write_bin: cannot open file");
    if (fwrite(x, sizeof(float), n, fp) != n)
        die("write_bin: short write");
    fclose(fp);
}

/* the engine's snapshot condition: verbose > 1, every shotInterval-th shot,
 * every snapInterval-th step */
static int vb_snap(int it)
{
    return vb_level > 1 && vb_shot % vb_shot_interval == 0 && it % vb_snap_interval == 0;
}

static void vb_field(const char *what, int it, const real *p, int Nz, int Nx, float dz, float dx)
{
    char name[512];
    snprintf(name, sizeof(name), "%s_iter%03d_ishot%04d_it%05d_%d_%d_%.2fm_%.2fm.bin", what,
             vb_iter, vb_shot, it, Nz, Nx, dz, dx);
    vb_write(name, p, (size_t)Nz * Nx);
}

/* ------------------------------------------------------------------------- */
/* Stencil: propagation.c:13-74.  Same operand order as the reference line 66-71. */
static void fd_sweep(real *restrict p2, const real *restrict p1, const real *restrict p0,
                     const real *invrho, const real *invrhox, const real *invrhoz,
                     const real *K, const real *coef1, const real *coef2, const real *coef3,
                     int Nz, real invdx, real invdz, real invdxdx, real invdzdz, int ixs,
                     int ixe, int izs, int ize)
{
    const real C1 = (real)0.77793115f, C2 = (real)-0.17388691f, C3 = (real)0.02338713f;
    const real D0 = (real)-2.8215452f, D1 = (real)1.57663238f, D2 = (real)-0.18347238f,
               D3 = (real)0.01761260f;
#pragma omp parallel for schedule(static)
    for (int ix = ixs; ix <= ixe; ++ix)
    {
        const real *c = p1 + (size_t)ix * Nz;
        const real *l1 = c - Nz, *l2 = c - 2 * Nz, *l3 = c - 3 * Nz;
        const real *r1 = c + Nz, *r2 = c + 2 * Nz, *r3 = c + 3 * Nz;
        size_t o = (size_t)ix * Nz;
        for (int iz = izs; iz <= ize; ++iz)
        {
            real dpdz = (C1 * (c[iz + 1] - c[iz - 1]) + C2 * (c[iz + 2] - c[iz - 2]) +
                         C3 * (c[iz + 3] - c[iz - 3])) * invdz;
            real dpdx = (C1 * (r1[iz] - l1[iz]) + C2 * (r2[iz] - l2[iz]) +
                         C3 * (r3[iz] - l3[iz])) * invdx;
            real dpdzdz = (D0 * c[iz] + D1 * (c[iz + 1] + c[iz - 1]) +
                           D2 * (c[iz + 2] + c[iz - 2]) + D3 * (c[iz + 3] + c[iz - 3])) *
                          invdzdz;
            real dpdxdx = (D0 * c[iz] + D1 * (r1[iz] + l1[iz]) + D2 * (r2[iz] + l2[iz]) +
                           D3 * (r3[iz] + l3[iz])) * invdxdx;
            p2[o + iz] = coef1[o + iz] * c[iz] + coef2[o + iz] * p0[o + iz] +
                         coef3[o + iz] * K[o + iz] *
                             (invrhox[o + iz] * dpdx + invrhoz[o + iz] * dpdz +
                              invrho[o + iz] * (dpdxdx + dpdzdz));
        }
    }
}

/* Free surface: propagation.c:76-101 (all Nx columns, rows 0..pml). */
static void free_surface(real *p2, int Nx, int Nz, int pml)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Nx; ix++)
    {
        real *c = p2 + (size_t)ix * Nz;
        c[pml] = (real)0;
        for (int iz = 0; iz <= pml; iz++)
            c[pml - iz] = -c[pml + iz];
    }
}

/* ------------------------------------------------------------------------- */
/* Set-up.  pad_model2d common.c:110-176 (edge replication). */
static void pad_model(const float *mod, int nz, int nx, real *modp, int nl)
{
    int Nz = nz + 2 * nl, Nx = nx + 2 * nl;
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++)
            modp[(size_t)(ix + nl) * Nz + iz + nl] = (real)mod[(size_t)ix * nz + iz];
    for (int ix = 0; ix < Nx; ix++)
        for (int iz = 0; iz < nl; iz++)
            modp[(size_t)ix * Nz + iz] = modp[(size_t)ix * Nz + nl];
    for (int ix = 0; ix < Nx; ix++)
        for (int iz = nz + nl; iz < Nz; iz++)
            modp[(size_t)ix * Nz + iz] = modp[(size_t)ix * Nz + nz + nl - 1];
    for (int ix = 0; ix < nl; ix++)
        for (int iz = 0; iz < Nz; iz++)
            modp[(size_t)ix * Nz + iz] = modp[(size_t)nl * Nz + iz];
    for (int ix = nx + nl; ix < Nx; ix++)
        for (int iz = 0; iz < Nz; iz++)
            modp[(size_t)ix * Nz + iz] = modp[(size_t)(nx + nl - 1) * Nz + iz];
}

/* absorbWeight propagation.c:173-190: w = 2^(d^alpha) - 1 through exp/log. */
static real absorb_weight(real alpha, int l, int nlayer)
{
    real d = ((real)0.1f * (real)l) / ((real)0.1f * (real)nlayer);
    return REXP(RLOG((real)2.0f) * RPOW(d, alpha)) - (real)1.0f;
}

/* mealCoefs2d propagation.c:192-354 with vs == NULL (Vs = vp/sqrt 3). */
static void meal(real *D, real *Beta, const real *vp, real fpeak, real dz, real dx, int NZ, int NX,
                 int nlayer)
{
    const real pd = (real)2;
    const real Rcoef = RPOW((real)10.0f, (real)-1.0f * ((RLOG10((real)nlayer) - (real)1.0f) /
                                                        (RLOG10((real)2.0f))) -
                                             (real)3.0f - (real)0.5f * pd);
    const real PPWfd = (real)10.0f;
    const real thx = (real)nlayer * dx, thz = (real)nlayer * dz;
    const real alpha = (real)2.5f;
    size_t Np = (size_t)NX * NZ;
    real *Vs = xcalloc(Np, sizeof(real)), *Dx = xcalloc(Np, sizeof(real));
    real *Dz = xcalloc(Np, sizeof(real)), *Bx = xcalloc(Np, sizeof(real));
    real *Bz = xcalloc(Np, sizeof(real));
    for (size_t i = 0; i < Np; i++)
    {
        Vs[i] = vp[i] / RSQRT((real)3.0f);
        D[i] = 0;
        Beta[i] = 1;
        Bx[i] = 1;
        Bz[i] = 1;
    }
    const real lR = RLOG(Rcoef);
    for (int ix = 0; ix < NX; ix++)
        for (int l = 1; l <= nlayer; l++)
        {
            real w = absorb_weight(alpha, l, nlayer);
            int iz = nlayer - l;
            size_t i = (size_t)ix * NZ + iz;
            real d0 = (real)-1.0f * ((real)1.0f + pd) * vp[i] * lR / ((real)2.0f * thz);
            real ppw = Vs[i] / (dz * fpeak);
            real bz = (real)2.0f * ppw / PPWfd;
            Dz[i] = d0 * w;
            Bz[i] = (real)1.0f + (bz - (real)1.0f) * w;
            iz = NZ - nlayer - 1 + l;
            i = (size_t)ix * NZ + iz;
            d0 = (real)-1.0f * ((real)1.0f + pd) * vp[i] * lR / ((real)2.0f * thz);
            ppw = Vs[i] / (dz * fpeak);
            bz = (real)2.0f * ppw / PPWfd;
            Dz[i] = d0 * w;
            Bz[i] = (real)1.0f + (bz - (real)1.0f) * w;
        }
    for (int l = 1; l <= nlayer; l++)
    {
        real w = absorb_weight(alpha, l, nlayer);
        for (int iz = 0; iz < NZ; iz++)
        {
            int ix = nlayer - l;
            size_t i = (size_t)ix * NZ + iz;
            real d0 = (real)-1.0f * ((real)1.0f + pd) * vp[i] * lR / ((real)2.0f * thx);
            real ppw = Vs[i] / (dx * fpeak);
            real bx = (real)2.0f * ppw / PPWfd;
            Dx[i] = d0 * w;
            Bx[i] = (real)1.0f + (bx - (real)1.0f) * w;
            ix = NX - nlayer - 1 + l;
            i = (size_t)ix * NZ + iz;
            d0 = (real)-1.0f * ((real)1.0f + pd) * vp[i] * lR / ((real)2.0f * thx);
            ppw = Vs[i] / (dx * fpeak);
            bx = (real)2.0f * ppw / PPWfd;
            Dx[i] = d0 * w;
            Bx[i] = (real)1.0f + (bx - (real)1.0f) * w;
        }
    }
    /* powf(x, 2.0f) is x*x exactly */
    for (size_t i = 0; i < Np; i++)
    {
        D[i] = RSQRT(Dx[i] * Dx[i] + Dz[i] * Dz[i]);
        Beta[i] = RSQRT((Bx[i] - (real)1.0f) * (Bx[i] - (real)1.0f) +
                        (Bz[i] - (real)1.0f) * (Bz[i] - (real)1.0f)) +
                  (real)1.0f;
    }
    free(Vs);
    free(Dx);
    free(Dz);
    free(Bx);
    free(Bz);
}

/* acoustic_iso2d_rgfd_parameter propagation.c:103-171 with calculate_gradient2d
 * common.c:416-479 as the fallback in the outer 6 cells. */
static void rgfd(const real *vp, const real *rho, int NX, int NZ, real dx, real dz, real *K,
                 real *invrho, real *invrhox, real *invrhoz)
{
    const real C1 = (real)0.91067892f, C2 = (real)-0.34187892f, C3 = (real)0.13833962f;
    const real C4 = (real)-0.04880710f, C5 = (real)0.01302148f, C6 = (real)-0.00199047f;
    const real invdx = (real)1.0f / dx, invdz = (real)1.0f / dz;
    size_t Np = (size_t)NX * NZ;
    for (size_t i = 0; i < Np; i++)
    {
        K[i] = rho[i] * vp[i] * vp[i];
        invrho[i] = (real)1.0f / rho[i];
    }
#define F(ix, iz) invrho[(size_t)(ix) * NZ + (iz)]
    for (int ix = 0; ix < NX; ix++)
        for (int iz = 0; iz < NZ; iz++)
        {
            real v;
            if (ix == 0)
                v = (F(ix + 1, iz) - F(ix, iz)) / dx;
            else if (ix == NX - 1)
                v = (F(ix, iz) - F(ix - 1, iz)) / dx;
            else
                v = (F(ix + 1, iz) - F(ix - 1, iz)) / ((real)2.0f * dx);
            invrhox[(size_t)ix * NZ + iz] = v;
            if (iz == 0)
                v = (F(ix, iz + 1) - F(ix, iz)) / dz;
            else if (iz == NZ - 1)
                v = (F(ix, iz) - F(ix, iz - 1)) / dz;
            else
                v = (F(ix, iz + 1) - F(ix, iz - 1)) / ((real)2.0f * dz);
            invrhoz[(size_t)ix * NZ + iz] = v;
        }
    for (int ix = 6; ix < NX - 6; ix++)
        for (int iz = 6; iz < NZ - 6; iz++)
        {
            invrhox[(size_t)ix * NZ + iz] =
                (C1 * (F(ix + 1, iz) - F(ix - 1, iz)) + C2 * (F(ix + 2, iz) - F(ix - 2, iz)) +
                 C3 * (F(ix + 3, iz) - F(ix - 3, iz)) + C4 * (F(ix + 4, iz) - F(ix - 4, iz)) +
                 C5 * (F(ix + 5, iz) - F(ix - 5, iz)) + C6 * (F(ix + 6, iz) - F(ix - 6, iz))) *
                invdx;
            invrhoz[(size_t)ix * NZ + iz] =
                (C1 * (F(ix, iz + 1) - F(ix, iz - 1)) + C2 * (F(ix, iz + 2) - F(ix, iz - 2)) +
                 C3 * (F(ix, iz + 3) - F(ix, iz - 3)) + C4 * (F(ix, iz + 4) - F(ix, iz - 4)) +
                 C5 * (F(ix, iz + 5) - F(ix, iz - 5)) + C6 * (F(ix, iz + 6) - F(ix, iz - 6))) *
                invdz;
        }
#undef F
}

typedef struct
{
    int Nx, Nz;
    real *vp, *rho, *D, *Beta, *c1, *c2, *c3, *K, *ir, *irx, *irz;
} Coefs;

/* Engine set-up, propagation.c:508-532.  The coefficient expressions of lines
 * 525-527 are evaluated in double from FP32 products (dt*dt*D*D etc. are float
 * products in C before the promotion). */
static void make_coefs(Coefs *c, const float *vpm, const float *rhom, int nx, int nz, int pml,
                       float dx, float dz, float dt, float fpeak)
{
    int Nx = nx + 2 * pml, Nz = nz + 2 * pml;
    size_t Np = (size_t)Nx * Nz;
    c->Nx = Nx;
    c->Nz = Nz;
    real **all[] = {&c->vp, &c->rho, &c->D, &c->Beta, &c->c1, &c->c2,
                    &c->c3, &c->K,   &c->ir, &c->irx, &c->irz};
    for (int i = 0; i < 11; i++)
        *all[i] = xcalloc(Np, sizeof(real));
    pad_model(vpm, nz, nx, c->vp, pml);
    pad_model(rhom, nz, nx, c->rho, pml);
    meal(c->D, c->Beta, c->vp, (real)fpeak, (real)dz, (real)dx, Nz, Nx, pml);
    const real rdt = (real)dt;
    for (size_t i = 0; i < Np; i++)
    {
        c->D[i] /= c->Beta[i];
        c->vp[i] /= c->Beta[i];
        c->rho[i] *= c->Beta[i];
        if (IS_F64)
        {
            double D = (double)c->D[i], t = (double)dt;
            c->c1[i] = (real)((2.0 - t * t * D * D) / (1.0 + t * D));
            c->c2[i] = (real)(-1.0 * (1.0 - t * D) / (1.0 + t * D));
            c->c3[i] = (real)((t * t) / (1.0 + t * D));
        }
        else
        {
            real D = c->D[i];
            real dt2D2 = rdt * rdt * D * D; /* float products */
            real dtD = rdt * D;
            real dt2 = rdt * rdt;
            c->c1[i] = (real)((2.0 - (double)dt2D2) / (1.0 + (double)dtD));
            c->c2[i] = (real)(-1.0 * (1.0 - (double)dtD) / (1.0 + (double)dtD));
            c->c3[i] = (real)((double)dt2 / (1.0 + (double)dtD));
        }
    }
    rgfd(c->vp, c->rho, Nx, Nz, (real)dx, (real)dz, c->K, c->ir, c->irx, c->irz);
}

static void free_coefs(Coefs *c)
{
    real *all[] = {c->vp, c->rho, c->D, c->Beta, c->c1, c->c2, c->c3, c->K, c->ir, c->irx, c->irz};
    for (int i = 0; i < 11; i++)
        free(all[i]);
}

/* ------------------------------------------------------------------------- */
/* Checkpoint store: checkpoint.c:29-265.  p[it] -> slot; pairs (kd-1, kd) get
 * their own slots, the d-2 steps in between share one ring that restarts after
 * every pair.  fullstore keeps u(it) for every it % MST == 0 instead. */
typedef struct
{
    int nt, d, nCkpt, ntPoint, fullstore;
    int *status;
    int *slot;
    real *storage;
    size_t Np;
} Ckpt;

static void ckpt_plan(int nt, int *d, int *nc, int *ntp)
{
    int dd = (int)roundf(sqrtf(2.0f * (float)(nt - 1)));
    if (dd < 3)
        dd = 3;
    int n = 0;
    for (int it = 0; it < nt; it++)
        if (it % dd == 0 || (it + 1) % dd == 0)
            n++;
    *d = dd;
    *nc = n;
    *ntp = n + dd - 2;
}

static void ckpt_alloc(Ckpt *c, int Nz, int Nx, int nt, int fullstore)
{
    c->nt = nt;
    c->fullstore = fullstore;
    c->Np = (size_t)Nz * Nx;
    c->status = xcalloc(nt, sizeof(int));
    c->slot = xcalloc(nt, sizeof(int));
    if (fullstore)
    {
        c->d = 0;
        c->ntPoint = nt / MST + 1;
        for (int it = 0; it < nt; it++)
            c->slot[it] = (it % MST == 0) ? it / MST : -1;
    }
    else
    {
        ckpt_plan(nt, &c->d, &c->nCkpt, &c->ntPoint);
        int ichk = 0, icur = c->nCkpt;
        for (int it = 0; it < nt; it++)
        {
            if (it % c->d == 0 || (it + 1) % c->d == 0)
            {
                c->slot[it] = ichk++;
                icur = c->nCkpt;
            }
            else
                c->slot[it] = icur++;
        }
    }
    c->storage = xcalloc((size_t)c->ntPoint * c->Np, sizeof(real));
}

static void ckpt_free(Ckpt *c)
{
    free(c->status);
    free(c->slot);
    free(c->storage);
}

static real *ckpt_field(Ckpt *c, int it) { return c->storage + (size_t)c->slot[it] * c->Np; }

/* save_acoustic2dCheckPoint checkpoint.c:130-159 */
static void ckpt_save_pair(Ckpt *c, const real *p0, const real *p1, int it)
{
    if (c->fullstore)
    {
        if (it % MST == 0)
            memcpy(ckpt_field(c, it), p1, c->Np * sizeof(real));
        return;
    }
    if (it >= 0 && it < c->nt && it % c->d == 0)
    {
        if (it - 1 >= 0)
        {
            memcpy(ckpt_field(c, it - 1), p0, c->Np * sizeof(real));
            c->status[it - 1] = 1;
        }
        memcpy(ckpt_field(c, it), p1, c->Np * sizeof(real));
        c->status[it] = 1;
    }
}

/* get_acoustic2dCheckPointStatus checkpoint.c:236-265 */
static int ckpt_status(const Ckpt *c, int it)
{
    if (c->status[it] == 1)
        return -1;
    int s = it;
    while (s >= 0 && c->status[s] == 0)
        s--;
    return s;
}

EXPORT void oracle_checkpoint_plan(int nt, int *d, int *nc, int *ntp, int *nrec)
{
    ckpt_plan(nt, d, nc, ntp);
    /* replay the status logic of the reverse loop with no fields */
    int *status = xcalloc(nt, sizeof(int));
    for (int it = 0; it < nt; it++)
        if (it % *d == 0)
        {
            if (it - 1 >= 0)
                status[it - 1] = 1;
            status[it] = 1;
        }
    int r = 0;
    for (int it = nt - 1; it >= 0; --it)
    {
        if (status[it] == 1)
            continue;
        int s = it;
        while (s >= 0 && status[s] == 0)
            s--;
        for (int itt = s; itt < it; itt++)
        {
            status[itt + 1] = 1;
            r++;
        }
    }
    free(status);
    *nrec = r;
}

/* calculate_memory_cost_checkpoint checkpoint.c:267-299 */
EXPORT float oracle_memory_cost_checkpoint(int nt, int NX, int NZ)
{
    int d, nc, ntp;
    ckpt_plan(nt, &d, &nc, &ntp);
    /* the int product times FSIZE, a size_t (checkpoint.c:296) */
    return 1.0f * (float)((size_t)(ntp * NZ * NX) * sizeof(float)) / powf(1024.0f, 3);
}

/* ------------------------------------------------------------------------- */
/* Trace helpers from common.c */

/* taper_traces common.c:796-895 */
EXPORT void oracle_taper_traces(float **tr, int nt, int ntr, int Lt, int Lx)
{
    if (Lt > 0)
    {
        if (Lt > nt / 3)
            Lt = nt / 3;
        float *cf = xcalloc(Lt + 1, sizeof(float));
        for (int i = 0; i <= Lt; i++)
            cf[i] = 0.5f * (1.0f + cosf((float)PI_D * (float)i / (float)Lt));
        for (int itr = 0; itr < ntr; itr++)
            for (int i = 0; i <= Lt; i++)
                tr[itr][i] *= cf[Lt - i];
        for (int itr = 0; itr < ntr; itr++)
            for (int i = 0; i <= Lt; i++)
                tr[itr][nt - Lt - 1 + i] *= cf[i];
        free(cf);
    }
    if (Lx > 0)
    {
        if (Lx > ntr / 3)
            Lx = ntr / 3;
        float *cf = xcalloc(Lx + 1, sizeof(float));
        for (int i = 0; i <= Lx; i++)
            cf[i] = 0.5f * (1.0f + cosf((float)PI_D * (float)i / (float)Lx));
        for (int i = 0; i <= Lx; i++)
            for (int it = 0; it < nt; it++)
                tr[i][it] *= cf[Lx - i];
        for (int i = 0; i <= Lx; i++)
            for (int it = 0; it < nt; it++)
                tr[ntr - Lx - 1 + i][it] *= cf[i];
        free(cf);
    }
}

/* Plain DFT pieces standing in for FFTW (single precision in the reference,
 * double here).  O(n^2): small cases only. */
static void dft_r2c(const float *x, int n, double *re, double *im)
{
    for (int k = 0; k <= n / 2; k++)
    {
        double sr = 0, si = 0;
        for (int j = 0; j < n; j++)
        {
            long m = ((long)k * j) % n;
            double a = -2.0 * PI_D * (double)m / n;
            sr += x[j] * cos(a);
            si += x[j] * sin(a);
        }
        re[k] = sr;
        im[k] = si;
    }
}

/* FFTW c2r (unnormalised): imaginary parts of bin 0 and of bin n/2 (n even)
 * are ignored. */
static void dft_c2r(const double *re, const double *im, int n, double *x)
{
    for (int j = 0; j < n; j++)
    {
        double s = re[0];
        int kmax = (n % 2 == 0) ? n / 2 - 1 : (n - 1) / 2;
        for (int k = 1; k <= kmax; k++)
        {
            long m = ((long)k * j) % n;
            double a = 2.0 * PI_D * (double)m / n;
            s += 2.0 * (re[k] * cos(a) - im[k] * sin(a));
        }
        if (n % 2 == 0)
            s += re[n / 2] * ((j % 2) ? -1.0 : 1.0);
        x[j] = s;
    }
}

static void window_filter(float **tr, float **out, int nt, int ntr, const float *win)
{
    double *re = xcalloc(nt / 2 + 1, sizeof(double)), *im = xcalloc(nt / 2 + 1, sizeof(double));
    double *x = xcalloc(nt, sizeof(double));
    for (int itr = 0; itr < ntr; itr++)
    {
        dft_r2c(tr[itr], nt, re, im);
        for (int k = 0; k < nt / 2 + 1; k++)
        {
            re[k] *= win[k];
            im[k] *= win[k];
        }
        dft_c2r(re, im, nt, x);
        for (int j = 0; j < nt; j++)
            out[itr][j] = (float)(x[j] / nt);
    }
    free(re);
    free(im);
    free(x);
}

/* apply_traces_lowpass common.c:584-687 (window of length nt, first nt/2+1 used) */
EXPORT void oracle_lowpass_traces(float **tr, float **out, int nt, int ntr, float fpass,
                                  float fcut, float dt)
{
    if (fcut <= fpass)
        die("fcut must be greater than fpass");
    float df = 1.0f / ((float)nt * dt), L = fcut - fpass;
    float *win = xcalloc(nt, sizeof(float));
    for (int i = 0; i < nt; i++)
    {
        float f = (float)i * df;
        if (f < fpass)
            win[i] = 1.0f;
        else if (f >= fpass && f <= fcut)
            win[i] = 0.5f * (1.0f + cosf((float)PI_D * (f - fpass) / L));
        else
            win[i] = 0.0f;
    }
    window_filter(tr, out, nt, ntr, win);
    free(win);
}

/* apply_traces_bandpass common.c:689-794; its cosf arguments are formed in
 * double (PI is a double constant there). */
EXPORT void oracle_bandpass_traces(float **tr, float **out, int nt, int ntr, float fpass,
                                   float fcut, float dt)
{
    if (fcut <= fpass)
        die("fcut must be greater than fpass");
    float df = 1.0f / ((float)nt * dt), lo = fpass, hi = fcut - fpass;
    float *win = xcalloc(nt, sizeof(float));
    for (int i = 0; i < nt; i++)
    {
        float f = (float)i * df;
        if (f < fpass)
            win[i] = 0.5f * (1.0f + cosf((float)(PI_D * (fpass - f) / lo)));
        else if (f >= fpass && f <= fcut)
            win[i] = 0.5f * (1.0f + cosf((float)(PI_D * (f - fpass) / hi)));
        else
            win[i] = 0.0f;
    }
    window_filter(tr, out, nt, ntr, win);
    free(win);
}

/* interpolate_traces common.c:482-581: zero-stuff by r, rectangular window of
 * gain r in the spectrum, inverse, /ns.  r <= 1 copies. */
EXPORT void oracle_interpolate_traces(float **tr, int ntr, int nt, int r, float **trs)
{
    if (r <= 1)
    {
        for (int itr = 0; itr < ntr; itr++)
            memcpy(trs[itr], tr[itr], sizeof(float) * nt);
        return;
    }
    int ns = (nt - 1) * r + 1, keep = ns / (2 * r);
    float *win = xcalloc(ns, sizeof(float));
    for (int k = 0; k < keep; k++)
        win[k] = (float)r;
    for (int k = ns - keep + 1; k < ns; k++)
        win[k] = (float)r;
    double *re = xcalloc(ns, sizeof(double)), *im = xcalloc(ns, sizeof(double));
    for (int itr = 0; itr < ntr; itr++)
    {
        for (int k = 0; k < ns; k++)
        {
            double sr = 0, si = 0;
            for (int it = 0; it < nt; it++)
            {
                long m = ((long)k * it * r) % ns;
                double a = -2.0 * PI_D * (double)m / ns;
                sr += tr[itr][it] * cos(a);
                si += tr[itr][it] * sin(a);
            }
            re[k] = sr * win[k];
            im[k] = si * win[k];
        }
        for (int j = 0; j < ns; j++)
        {
            double s = 0;
            for (int k = 0; k < ns; k++)
            {
                if (win[k] == 0.0f)
                    continue;
                long m = ((long)k * j) % ns;
                double a = 2.0 * PI_D * (double)m / ns;
                s += re[k] * cos(a) - im[k] * sin(a);
            }
            trs[itr][j] = (float)(s / ns);
        }
    }
    free(win);
    free(re);
    free(im);
}

/* calculate_peak_frequency common.c:363-413: first bin of maximum |rfft|. */
EXPORT float oracle_peak_frequency(const float *s, int nt, float dt)
{
    float df = 1.0f / ((float)nt * dt), fpeak = 0.0f, maxamp = -1.0f;
    double *re = xcalloc(nt / 2 + 1, sizeof(double)), *im = xcalloc(nt / 2 + 1, sizeof(double));
    dft_r2c(s, nt, re, im);
    for (int i = 0; i <= nt / 2; i++)
    {
        float amp = (float)sqrt(re[i] * re[i] + im[i] * im[i]);
        if (amp > maxamp)
        {
            maxamp = amp;
            fpeak = (float)i * df;
        }
    }
    free(re);
    free(im);
    return fpeak;
}

/* ------------------------------------------------------------------------- */
/* optimization.c */

/* compute_adjoint_source2d optimization.c:300-419.  vp is the FULL model and
 * igx is window-local (the reference's own indexing, lines 352-354). */
EXPORT void oracle_adjoint_source(OracleRecord2D *pre, OracleRecord2D *obs, float **vp,
                                  float freq, int filter, OracleRecord2D *adj)
{
    int nt = adj->ns, ntr = adj->ntr;
    for (int itr = 0; itr < ntr; itr++)
        for (int it = 0; it < nt; it++)
            adj->traces_p[itr][it] = pre->traces_p[itr][it] - obs->traces_p[itr][it];
    double vsum = 0.0;
    for (int itr = 0; itr < ntr; itr++)
        vsum += vp[adj->igx[itr]][adj->igz[itr]];
    float vavg = (float)(vsum / ntr);
    if (freq == 0.0f)
        freq = adj->fpeak;
    float kpass = freq / vavg;
    int Lt = (int)roundf((2.0f / freq) / adj->dt) + 1;
    int Lx = (int)roundf((2.0f / kpass) / adj->dtr) + 1;
    if (filter == 0 || filter == 1)
    {
        float **tt = xcalloc(nt, sizeof(float *));
        float *blk = xcalloc((size_t)nt * ntr, sizeof(float));
        for (int i = 0; i < nt; i++)
            tt[i] = blk + (size_t)i * ntr;
        void (*flt)(float **, float **, int, int, float, float, float) =
            filter == 0 ? oracle_lowpass_traces : oracle_bandpass_traces;
        oracle_taper_traces(adj->traces_p, nt, ntr, Lt, 0);
        flt(adj->traces_p, adj->traces_p, nt, ntr, freq, 2.0f * freq, adj->dt);
        oracle_taper_traces(adj->traces_p, nt, ntr, 0, Lx);
        for (int itr = 0; itr < ntr; itr++)
            for (int it = 0; it < nt; it++)
                tt[it][itr] = adj->traces_p[itr][it];
        flt(tt, tt, ntr, nt, kpass, 2.0f * kpass, adj->dtr);
        for (int it = 0; it < nt; it++)
            for (int itr = 0; itr < ntr; itr++)
                adj->traces_p[itr][it] = tt[it][itr];
        free(blk);
        free(tt);
    }
    oracle_taper_traces(adj->traces_p, nt, ntr, Lt, Lx);
}

/* calculate_objective_value optimization.c:99-131: float squares, double sum */
EXPORT float oracle_objective(OracleRecord2D *rec, int Nshot)
{
    double fun = 0.0;
    for (int s = 0; s < Nshot; s++)
        for (int itr = 0; itr < rec[s].ntr; itr++)
            for (int it = 0; it < rec[s].ns; it++)
            {
                float v = rec[s].traces_p[itr][it];
                fun += v * v;
            }
    fun *= 0.5;
    return (float)fun;
}

/* calculate_step_length optimization.c:134-175 */
EXPORT float oracle_step_length(OracleRecord2D *a, OracleRecord2D *b, int Nshot)
{
    double num = 0.0, den = 0.0;
    for (int s = 0; s < Nshot; s++)
        for (int itr = 0; itr < a[s].ntr; itr++)
            for (int it = 0; it < a[s].ns; it++)
            {
                float r0 = a[s].traces_p[itr][it];
                float delta = r0 - b[s].traces_p[itr][it];
                num += r0 * delta;
                den += delta * delta;
            }
    if (den == 0.0)
        return 0.0f;
    return (float)(num / den);
}

/* compute_model_perturbation optimization.c:178-298 */
EXPORT float oracle_model_perturbation(float fraction, const float *slow, const float *g,
                                       float *out, int len)
{
    int hist[1001];
    float Max = 0.0f, Min = 0.0f;
    memset(hist, 0, sizeof(hist));
    for (int i = 0; i < len; i++)
    {
        if (g[i] > Max)
            Max = g[i];
        if (g[i] < Min)
            Min = g[i];
    }
    float interval = (float)((Max - Min) / 1000.0);
    int pos = 0;
    for (int i = 0; i < len; i++)
    {
        pos = (int)((g[i] - Min) / interval + 0.5);
        hist[pos]++;
    }
    int peak = 0;
    for (int i = 0; i < 1001; i++)
        if (hist[i] > peak)
        {
            peak = hist[i];
            pos = i;
        }
    int hi = pos, lo = pos;
    const float cap = 0.99f;
    float acc = (float)hist[hi] / (float)len;
    while (acc < cap)
    {
        int moved = 0;
        if (hi < 1000)
        {
            hi++;
            acc += (float)hist[hi] / (float)len;
            moved = 1;
        }
        if (lo > 0)
        {
            lo--;
            acc += (float)hist[lo] / (float)len;
            moved = 1;
        }
        if (!moved)
            break; /* the reference would spin here; never reached when the mass sums to 1 */
    }
    float a = Min + (float)lo * interval, b = Min + (float)hi * interval;
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    float scalar = (a > b) ? a : b;
    double avg = 0;
    for (int i = 0; i < len; i++)
        avg += slow[i];
    scalar = (float)((avg / len * (double)fraction) / (double)scalar);
    for (int i = 0; i < len; i++)
        out[i] = -scalar * g[i];
    return scalar;
}

/* apply_gradient_shot_precondition optimization.c:58-97 */
EXPORT void oracle_shot_precondition(float *grad, float *precond, int n)
{
    double sum = 0.0;
    for (int i = 0; i < n; i++)
        sum += (double)precond[i];
    float avg = (float)(sum / n);
    for (int i = 0; i < n; i++)
    {
        float f = 1.0f + precond[i] / (7.0f * avg);
        grad[i] = grad[i] / f;
        precond[i] = precond[i] / f;
    }
}

/* apply_gradient_precondition optimization.c:13-56 */
EXPORT void oracle_global_precondition(float *grad, const float *precond, int n, float alpha)
{
    if (alpha <= 0.0f)
        alpha = 0.005f;
    double e = 0.0;
    for (int i = 0; i < n; i++)
        e += (double)precond[i] * (double)precond[i];
    float rms = (float)sqrt(e / n);
    for (int i = 0; i < n; i++)
        grad[i] = grad[i] / (precond[i] + alpha * rms);
}

/* ------------------------------------------------------------------------- */
/* common.c model helpers used by the FWI driver (acoustic2d_fwi.c:418-465) */

/* smooth2D common.c:927-1040: cos^2 window (double cos, float square), float
 * products summed in double, / (sum w + FLT_MIN). */
EXPORT void oracle_smooth2d(float **in, float **out, int nx, int nz, int N, int iszdirect)
{
    float *w = xcalloc(2 * N + 1, sizeof(float));
    for (int i = 0; i < 2 * N + 1; i++)
    {
        w[i] = (float)cos(PI_D * (i - N) / (2.0 * N));
        w[i] *= w[i];
    }
    int n = (iszdirect == 1) ? nz : nx;
    for (int j = 0; j < n; j++)
    {
        int ws = (j >= N) ? 0 : N - j;
        int ds = (j >= N) ? j - N : 0;
        int we = (n - 1 - j < N) ? 2 * N - (N - (n - 1 - j)) : 2 * N;
        int m = (iszdirect == 1) ? nx : nz;
        for (int k = 0; k < m; k++)
        {
            double a1 = 0.0, a2 = 0.0;
            for (int i = ws, i2 = ds; i <= we; i++, i2++)
            {
                float v = (iszdirect == 1) ? in[k][i2] : in[i2][k];
                a1 += w[i] * v;
                a2 += w[i];
            }
            if (iszdirect == 1)
                out[k][j] = (float)(a1 / (a2 + FLT_MIN));
            else
                out[j][k] = (float)(a1 / (a2 + FLT_MIN));
        }
    }
    free(w);
}

/* vp2density common.c:298-330 */
EXPORT void oracle_vp2density(const float *vp, float *rho, int n)
{
    for (int i = 0; i < n; i++)
    {
        if (vp[i] <= 0.0f)
            die("all input P-wave velocities must be positive");
        rho[i] = 310.0f * powf(vp[i], 0.25f);
    }
}

/* clip_model common.c:218-246 */
EXPORT void oracle_clip(float *m, int n, float lo, float hi)
{
    for (int i = 0; i < n; i++)
    {
        if (m[i] > hi)
            m[i] = hi;
        else if (m[i] < lo)
            m[i] = lo;
    }
}

/* v2slow / slow2v common.c:248-296 */
EXPORT void oracle_v2slow(const float *v, float *s, long n)
{
    for (long i = 0; i < n; i++)
        s[i] = (v[i] != 0.0f) ? 1.0f / v[i] : 0.0f;
}

EXPORT void oracle_slow2v(const float *s, float *v, long n)
{
    for (long i = 0; i < n; i++)
        v[i] = (s[i] != 0.0f) ? 1.0f / s[i] : 0.0f;
}

/* set_waterlayer_value common.c:898-924 */
EXPORT void oracle_set_water(float **m, float value, int wd, int nx, int nz)
{
    (void)nz;
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < wd; iz++)
            m[ix][iz] = value;
}

/* ------------------------------------------------------------------------- */
/* Engine: iso_acoustic2d_propagation_engine, propagation.c:356-775.
 * vpm/rhom: window models, contiguous [nx][nz].  mode 0 forward, 1 forward with
 * checkpoint saves and pseudo-Hessian (written to out[nx][nz] when
 * precondition==1), 2 adjoint with reconstruction and imaging (out = gradient). */
static void engine(OracleRecord2D *rec, Ckpt *ck, const float *vpm, const float *rhom, int nx,
                   int nz, float dx, float dz, float dt, int nt, int pml, int fs, int precondition,
                   int mode, float *out)
{
    const int N = 3;
    Coefs c;
    if (vb_level > 1)
    {
        /* lines 423-441 */
        vb_shot_interval = env_int("shotInterval", 20);
        vb_snap_interval = env_int("snapInterval", 1000);
        if (vb_shot_interval <= 0 || vb_snap_interval <= 0)
            die("shotInterval and snapInterval must be positive");
    }
    make_coefs(&c, vpm, rhom, nx, nz, pml, dx, dz, dt, rec->fpeak);
    const int Nx = c.Nx, Nz = c.Nz;
    const size_t Np = (size_t)Nx * Nz;
    const real rdt = (real)dt;
    const real invdx = (real)1.0f / (real)dx, invdz = (real)1.0f / (real)dz;
    const real invdxdx = (real)1.0f / ((real)dx * (real)dx);
    const real invdzdz = (real)1.0f / ((real)dz * (real)dz);
    const real invdtdt = (real)1.0f / (rdt * rdt);
    const int isx = rec->isx + pml, isz = rec->isz + pml;
    const int ntr = rec->ntr, sr = rec->sampleRate;
    const int ixs = N, ixe = Nx - N - 1, izs = (fs == 0) ? N : pml, ize = Nz - N - 1;
    const size_t isrc = (size_t)isx * Nz + isz;

    real *p0 = xcalloc(Np, sizeof(real)), *p1 = xcalloc(Np, sizeof(real));
    real *p2 = xcalloc(Np, sizeof(real)), *tmp;
    double *acc = NULL;
    if ((mode == 1 && precondition == 1) || mode == 2)
        acc = xcalloc((size_t)nx * nz, sizeof(double));

#define SWEEP(a, b, d)                                                                          \
    fd_sweep(a, b, d, c.ir, c.irx, c.irz, c.K, c.c1, c.c2, c.c3, Nz, invdx, invdz, invdxdx,      \
             invdzdz, ixs, ixe, izs, ize)

    if (mode == 0 || mode == 1)
    {
        for (int it = 0; it < nt; ++it)
        {
            if (mode == 1)
                ckpt_save_pair(ck, p0, p1, it);
            SWEEP(p2, p1, p0);
            /* line 558: source term in double */
            if (IS_F64)
                p2[isrc] += 1.0 * rec->src[it] * c.K[isrc] * (double)dt * (double)dt *
                            (double)invdx * (double)invdz;
            else
                p2[isrc] = (real)((double)p2[isrc] + 1.0 * rec->src[it] * (double)c.K[isrc] *
                                                         (double)dt * (double)dt *
                                                         (double)invdx * (double)invdz);
            if (fs == 1)
                free_surface(p2, Nx, Nz, pml);
            if (it % sr == 0)
            {
                int k = it / sr;
                for (int itr = 0; itr < ntr; ++itr)
                {
                    size_t i = (size_t)(rec->igx[itr] + pml) * Nz + rec->igz[itr] + pml;
                    rec->traces_p[itr][k] = (float)p1[i];
                }
            }
            if (mode == 1 && it % MST == 0 && precondition == 1)
            {
#pragma omp parallel for schedule(static)
                for (int ix = pml; ix < pml + nx; ix++)
                    for (int iz = pml; iz < pml + nz; iz++)
                    {
                        size_t i = (size_t)ix * Nz + iz;
                        real v = (p2[i] - (real)2.0f * p1[i] + p0[i]) * invdtdt;
                        acc[(size_t)(ix - pml) * nz + iz - pml] += (double)(v * v);
                    }
            }
            /* lines 597-601 */
            if (vb_snap(it))
                vb_field("wavefield_forward", it, p1, Nz, Nx, dz, dx);
            tmp = p0;
            p0 = p1;
            p1 = p2;
            p2 = tmp;
        }
        if (mode == 1 && precondition == 1)
        {
            /* line 619: 4 / SQR(rho*vp) with the square taken in float */
            for (size_t i = 0; i < (size_t)nx * nz; i++)
            {
                real rv = (real)rhom[i] * (real)vpm[i];
                acc[i] = 4.0 * acc[i] / ((double)(rv * rv));
                out[i] = (float)acc[i];
            }
        }
    }

    if (mode == 2)
    {
        real *f0 = xcalloc(Np, sizeof(real)), *f1 = xcalloc(Np, sizeof(real));
        real *f2 = xcalloc(Np, sizeof(real));
        /* interpolate_traces (propagation.c:635); tail beyond (ns-1)*sr is zero */
        float **adj = xcalloc(ntr, sizeof(float *));
        float *adjblk = xcalloc((size_t)ntr * nt, sizeof(float));
        for (int i = 0; i < ntr; i++)
            adj[i] = adjblk + (size_t)i * nt;
        oracle_interpolate_traces(rec->traces_p, ntr, rec->ns, sr, adj);
        /* the content of the reference's p1_fwd, which its verbose snapshot
         * writes: zeros (line 498), the last recomputed field, or u(it) copied
         * in for imaging (line 709) */
        real *pf = (vb_level > 1) ? xcalloc(Np, sizeof(real)) : NULL;

        for (int it = nt - 1; it >= 0; --it)
        {
            if (!ck->fullstore)
            {
                int s = ckpt_status(ck, it);
                if (s != -1)
                {
                    /* get_acoustic2dCheckPoint: loads only when s % d == 0 */
                    if (s % ck->d == 0)
                    {
                        if (s - 1 >= 0)
                            memcpy(f0, ckpt_field(ck, s - 1), Np * sizeof(real));
                        else
                            memset(f0, 0, Np * sizeof(real));
                        memcpy(f1, ckpt_field(ck, s), Np * sizeof(real));
                    }
                    for (int itt = s; itt < it; itt++)
                    {
                        SWEEP(f2, f1, f0);
                        if (fs == 1)
                            free_surface(f2, Nx, Nz, pml);
                        /* line 662: source term in float, after the free surface */
                        f2[isrc] += (real)rec->src[itt] * c.K[isrc] * rdt * rdt * invdx * invdz;
                        memcpy(ckpt_field(ck, itt + 1), f2, Np * sizeof(real));
                        ck->status[itt + 1] = 1;
                        tmp = f0;
                        f0 = f1;
                        f1 = f2;
                        f2 = tmp;
                    }
                    if (pf)
                        memcpy(pf, f1, Np * sizeof(real));
                }
            }
            for (int itr = 0; itr < ntr; ++itr)
            {
                size_t i = (size_t)(rec->igx[itr] + pml) * Nz + rec->igz[itr] + pml;
                p1[i] -= (real)adj[itr][it] * c.K[i] * rdt * rdt * invdx * invdz;
            }
            SWEEP(p2, p1, p0);
            if (fs == 1)
                free_surface(p2, Nx, Nz, pml);
            if (it % MST == 0)
            {
                const real *u = ckpt_field(ck, it);
                if (pf)
                    memcpy(pf, u, Np * sizeof(real));
#pragma omp parallel for schedule(static)
                for (int ix = pml; ix < pml + nx; ++ix)
                    for (int iz = pml; iz < pml + nz; ++iz)
                    {
                        size_t i = (size_t)ix * Nz + iz;
                        acc[(size_t)(ix - pml) * nz + iz - pml] +=
                            (double)((p2[i] - (real)2.0f * p1[i] + p0[i]) * invdtdt * u[i]);
                    }
            }
            /* lines 717-723 */
            if (vb_snap(it))
            {
                vb_field("wavefield_forward_reconstruction", it, pf, Nz, Nx, dz, dx);
                vb_field("wavefield_backward", it, p1, Nz, Nx, dz, dx);
            }
            tmp = p0;
            p0 = p1;
            p1 = p2;
            p2 = tmp;
        }
        /* line 740 */
        for (size_t i = 0; i < (size_t)nx * nz; i++)
        {
            acc[i] = 2.0 * acc[i] / ((double)((real)rhom[i] * (real)vpm[i]));
            out[i] = (float)acc[i];
        }
        free(f0);
        free(f1);
        free(f2);
        free(pf);
        free(adj);
        free(adjblk);
    }
#undef SWEEP
    free(p0);
    free(p1);
    free(p2);
    free(acc);
    free_coefs(&c);
}

/* iso_acoustic2d_propagation, propagation.c:777-1036 */
static void propagation(float **vp, float **rho, OracleRecord2D *obs, int Nshot, int nz, int nx,
                        float dx, float dz, float dt, int nt, int pml, int fs, float freq,
                        float filter, int precondition, int problem, OracleRecord2D *pre,
                        float **grad, int fullstore)
{
    /* lines 846-852 */
    vb_shot_interval = env_int("shotInterval", 20);
    if (vb_level > 0 && vb_shot_interval <= 0)
        die("shotInterval must be positive");
    char name[512];
    if (problem == 0 || problem == 2)
    {
        for (int s = 0; s < Nshot; s++)
        {
            int ix0 = obs[s].ix_min, nxx = obs[s].ix_max - ix0 + 1;
            vb_shot = s;
            engine(&pre[s], NULL, vp[0] + (size_t)ix0 * nz, rho[0] + (size_t)ix0 * nz, nxx, nz,
                   dx, dz, dt, nt, pml, fs, precondition, 0, NULL);
            if (problem == 2)
                oracle_adjoint_source(&pre[s], &obs[s], vp, freq, (int)filter, &pre[s]);
        }
        return;
    }
    if (problem != 1)
        return;
    size_t n = (size_t)nx * nz;
    double *gsum = xcalloc(n, sizeof(double));
    double *psum = (precondition == 1) ? xcalloc(n, sizeof(double)) : NULL;
    for (int s = 0; s < Nshot; s++)
    {
        int ix0 = obs[s].ix_min, nxx = obs[s].ix_max - ix0 + 1;
        int Nx = nxx + 2 * pml, Nz = nz + 2 * pml;
        Ckpt ck;
        ckpt_alloc(&ck, Nz, Nx, nt, fullstore);
        float *gs = xcalloc((size_t)nxx * nz, sizeof(float));
        float *ps = (precondition == 1) ? xcalloc((size_t)nxx * nz, sizeof(float)) : NULL;
        const float *vw = vp[0] + (size_t)ix0 * nz, *rw = rho[0] + (size_t)ix0 * nz;
        const int out = vb_level > 0 && s % vb_shot_interval == 0;
        const size_t nrec = (size_t)pre[s].ns * pre[s].ntr;
        vb_shot = s;
        engine(&pre[s], &ck, vw, rw, nxx, nz, dx, dz, dt, nt, pml, fs, precondition, 1, ps);
        /* lines 917-922 */
        if (out)
        {
            snprintf(name, sizeof(name), "predicted_iter%03d_shot%05d_%d_%d_%.2fms_%.2fm.bin",
                     vb_iter, s + 1, pre[s].ns, pre[s].ntr, pre[s].dt * 1e3, pre[s].dtr);
            vb_write_f(name, pre[s].traces_p[0], nrec);
        }
        oracle_adjoint_source(&pre[s], &obs[s], vp, freq, (int)filter, &pre[s]);
        if (out)
        {
            snprintf(name, sizeof(name), "adjsource_iter%03d_shot%05d_%d_%d_%.2fms_%.2fm.bin",
                     vb_iter, s + 1, pre[s].ns, pre[s].ntr, pre[s].dt * 1e3, pre[s].dtr);
            vb_write_f(name, pre[s].traces_p[0], nrec);
        }
        engine(&pre[s], &ck, vw, rw, nxx, nz, dx, dz, dt, nt, pml, fs, precondition, 2, gs);
        /* lines 939-944 */
        if (out)
        {
            snprintf(name, sizeof(name), "gradient_raw_iter%03d_shot%05d_%d_%d_%.2fm_%.2fm.bin",
                     vb_iter, s + 1, nz, nxx, dz, dx);
            vb_write_f(name, gs, (size_t)nz * nxx);
        }
        if (precondition == 1)
            oracle_shot_precondition(gs, ps, nxx * nz);
        for (size_t i = 0; i < (size_t)nxx * nz; i++)
            gsum[(size_t)ix0 * nz + i] += gs[i];
        if (precondition == 1)
            for (size_t i = 0; i < (size_t)nxx * nz; i++)
                psum[(size_t)ix0 * nz + i] += ps[i];
        ckpt_free(&ck);
        free(gs);
        free(ps);
    }
    for (size_t i = 0; i < n; i++)
        grad[0][i] = (float)gsum[i];
    /* lines 996-1005 */
    if (vb_level > 0)
    {
        snprintf(name, sizeof(name), "gradient_raw_iter%03d_%d_%d_%.2fm_%.2fm.bin", vb_iter, nz, nx,
                 dz, dx);
        vb_write_f(name, grad[0], n);
    }
    if (precondition == 1)
    {
        float *pf = xcalloc(n, sizeof(float));
        for (size_t i = 0; i < n; i++)
            pf[i] = (float)psum[i];
        if (vb_level > 0)
        {
            snprintf(name, sizeof(name), "precond_iter%03d_%d_%d_%.2fm_%.2fm.bin", vb_iter, nz, nx,
                     dz, dx);
            vb_write_f(name, pf, n);
        }
        oracle_global_precondition(grad[0], pf, (int)n, 0.005f);
        free(pf);
    }
    free(gsum);
    free(psum);
}

EXPORT void oracle_propagation(float **vp, float **rho, OracleRecord2D *obs, int Nshot, int nz,
                               int nx, float dx, float dz, float dt, int nt, int pmlThick,
                               int freesurface, float freq, float filter, int precondition,
                               int problem, OracleRecord2D *pre, float **grad)
{
    propagation(vp, rho, obs, Nshot, nz, nx, dx, dz, dt, nt, pmlThick, freesurface, freq, filter,
                precondition, problem, pre, grad, 0);
}

/* oracle_propagation with the reference's iter and verbose arguments: the
 * debug files of the verbose levels in the working directory. */
EXPORT void oracle_propagation_verbose(float **vp, float **rho, OracleRecord2D *obs, int Nshot,
                                       int nz, int nx, float dx, float dz, float dt, int nt,
                                       int pmlThick, int freesurface, float freq, float filter,
                                       int precondition, int problem, OracleRecord2D *pre,
                                       float **grad, int iter, int verbose)
{
    vb_level = verbose;
    vb_iter = iter;
    propagation(vp, rho, obs, Nshot, nz, nx, dx, dz, dt, nt, pmlThick, freesurface, freq, filter,
                precondition, problem, pre, grad, 0);
    vb_level = 0;
    vb_iter = 0;
    vb_shot = 0;
}

EXPORT void oracle_gradient_fullstore(float **vp, float **rho, OracleRecord2D *obs, int Nshot,
                                      int nz, int nx, float dx, float dz, float dt, int nt,
                                      int pmlThick, int freesurface, float freq, float filter,
                                      int precondition, OracleRecord2D *pre, float **grad)
{
    propagation(vp, rho, obs, Nshot, nz, nx, dx, dz, dt, nt, pmlThick, freesurface, freq, filter,
                precondition, 1, pre, grad, 1);
}

/* ------------------------------------------------------------------------- */
/* Float-array entry points for unit tests of single pieces. */
EXPORT void oracle_fd_step(float *p2, const float *p1, const float *p0, const float *invrho,
                           const float *invrhox, const float *invrhoz, const float *K,
                           const float *coef1, const float *coef2, const float *coef3, int Nx,
                           int Nz, float invdx, float invdz, float invdxdx, float invdzdz, int ixs,
                           int ixe, int izs, int ize)
{
    size_t Np = (size_t)Nx * Nz;
    const float *src[] = {p1, p0, invrho, invrhox, invrhoz, K, coef1, coef2, coef3};
    real *buf[10];
    for (int k = 0; k < 9; k++)
    {
        buf[k] = xcalloc(Np, sizeof(real));
        for (size_t i = 0; i < Np; i++)
            buf[k][i] = (real)src[k][i];
    }
    buf[9] = xcalloc(Np, sizeof(real));
    for (size_t i = 0; i < Np; i++)
        buf[9][i] = (real)p2[i];
    fd_sweep(buf[9], buf[0], buf[1], buf[2], buf[3], buf[4], buf[5], buf[6], buf[7], buf[8], Nz,
             (real)invdx, (real)invdz, (real)invdxdx, (real)invdzdz, ixs, ixe, izs, ize);
    for (size_t i = 0; i < Np; i++)
        p2[i] = (float)buf[9][i];
    for (int k = 0; k < 10; k++)
        free(buf[k]);
}

EXPORT void oracle_freesurface(float *p2, int Nx, int Nz, int pml)
{
    size_t Np = (size_t)Nx * Nz;
    real *b = xcalloc(Np, sizeof(real));
    for (size_t i = 0; i < Np; i++)
        b[i] = (real)p2[i];
    free_surface(b, Nx, Nz, pml);
    for (size_t i = 0; i < Np; i++)
        p2[i] = (float)b[i];
    free(b);
}

EXPORT void oracle_pad_model(const float *mod, int nz, int nx, float *modp, int nl)
{
    size_t Np = (size_t)(nx + 2 * nl) * (nz + 2 * nl);
    real *b = xcalloc(Np, sizeof(real));
    pad_model(mod, nz, nx, b, nl);
    for (size_t i = 0; i < Np; i++)
        modp[i] = (float)b[i];
    free(b);
}

EXPORT void oracle_meal_coefs(float *D, float *Beta, const float *vp, float fpeak, float dz,
                              float dx, int NZ, int NX, int nlayer)
{
    size_t Np = (size_t)NX * NZ;
    real *v = xcalloc(Np, sizeof(real)), *d = xcalloc(Np, sizeof(real));
    real *b = xcalloc(Np, sizeof(real));
    for (size_t i = 0; i < Np; i++)
        v[i] = (real)vp[i];
    meal(d, b, v, (real)fpeak, (real)dz, (real)dx, NZ, NX, nlayer);
    for (size_t i = 0; i < Np; i++)
    {
        D[i] = (float)d[i];
        Beta[i] = (float)b[i];
    }
    free(v);
    free(d);
    free(b);
}

EXPORT void oracle_setup_coefs(const float *vpm, const float *rhom, int nx, int nz, int pml,
                               float dx, float dz, float dt, float fpeak, float *D, float *Beta,
                               float *coef1, float *coef2, float *coef3, float *K,
                               float *invrho, float *invrhox, float *invrhoz)
{
    Coefs c;
    make_coefs(&c, vpm, rhom, nx, nz, pml, dx, dz, dt, fpeak);
    size_t Np = (size_t)c.Nx * c.Nz;
    float *dst[] = {D, Beta, coef1, coef2, coef3, K, invrho, invrhox, invrhoz};
    real *srcs[] = {c.D, c.Beta, c.c1, c.c2, c.c3, c.K, c.ir, c.irx, c.irz};
    for (int k = 0; k < 9; k++)
        if (dst[k])
            for (size_t i = 0; i < Np; i++)
                dst[k][i] = (float)srcs[k][i];
    free_coefs(&c);
}
