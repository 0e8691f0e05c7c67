/* acoustic2d_modeling: forward modeling of all shots of an acquisition on the
 * GPU through libfwiedu's iso_acoustic2d_propagation (problem 0), the flow of
 * the reference driver (main/acoustic2d_modeling.c:110-388): models and
 * wavelet from float32 files, Gardner density or a density file, water rows,
 * the geometry from a JSON file or from parameters, grid geometry, modeling,
 * the SU gather "shots_p_<Nshot>_<ns>_<traces>_<dt>ms.su".
 *
 * Parameters (key=value): nx nz nt dx dz dt srcfile vpfile [rhofile]
 * [Gardner=1] [freesurface=0] [pmlThick=30] [waterdepth=0] [vp_water=1500]
 * [rho_water=1000] [sampleRate=1] [prefix=shots_p] [verbose=0], and the geometry:
 * jsonfile=, or, without it, the spread of the reference's parameters (in
 * cells; :191-246) [Nshot=99999] [fsx=1] [ds=1] [sdepth=2] [offsmin=0]
 * [offsmax=99999] [dtr=1] [gdepth=2].
 * verbose > 0 also writes the models and the wavelet (vp_*, rho_*, wavelet_*
 * .bin); verbose > 1 the library's wavefield files (see fwiedu.h).
 */
#include <math.h>
#include <time.h>

#include "../../include/fwiedu.h"
#include "args.h"

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

int main(int argc, char **argv)
{
    args_init(argc, argv);
    const int nx = arg_int("nx", 0, 1), nz = arg_int("nz", 0, 1), nt = arg_int("nt", 0, 1);
    const float dx = arg_float("dx", 0, 1), dz = arg_float("dz", 0, 1), dt = arg_float("dt", 0, 1);
    const int freesurface = arg_int("freesurface", 0, 0), pml = arg_int("pmlThick", 30, 0);
    const int gardner = arg_int("Gardner", 1, 0), wd = arg_int("waterdepth", 0, 0);
    const float vpw = arg_float("vp_water", 1500.0f, 0), rhow = arg_float("rho_water", 1000.0f, 0);
    const int sampleRate = arg_int("sampleRate", 1, 0);
    const char *srcfile = arg_string("srcfile", 1), *vpfile = arg_string("vpfile", 1);
    const char *rhofile = arg_string("rhofile", 0);
    const char *jsonfile = arg_string("jsonfile", 0);
    const char *prefix = arg_string("prefix", 0);
    const int verbose = arg_int("verbose", 0, 0);
    if (!gardner && !rhofile) die("didn't specify data file of density model:", "rhofile=");

    float *x = (float *)malloc(sizeof(float) * nx), *z = (float *)malloc(sizeof(float) * nz);
    for (int i = 0; i < nx; ++i) x[i] = (float)i * dx;
    for (int i = 0; i < nz; ++i) z[i] = (float)i * dz;
    float *src = (float *)malloc(sizeof(float) * nt);
    float **vp = alloc2(nz, nx), **rho = alloc2(nz, nx);
    read_floats(srcfile, src, (size_t)nt);
    read_floats(vpfile, vp[0], (size_t)nx * nz);
    if (gardner)
        for (long i = 0; i < (long)nx * nz; ++i) rho[0][i] = 310.0f * powf(vp[0][i], 0.25f);
    else
        read_floats(rhofile, rho[0], (size_t)nx * nz);
    set_water(vp, vpw, wd, nx);
    set_water(rho, rhow, wd, nx);
    if (verbose > 0) write_inputs("", vp, rho, src, nz, nx, dz, dx, nt, dt);

    int Nshot;
    Record2D *obs;
    if (jsonfile) {
        Nshot = fwiedu_json_shot_count(jsonfile);
        obs = (Record2D *)calloc(Nshot, sizeof(Record2D));
        fwiedu_observe2d_from_json(obs, jsonfile, Nshot, x, nx, z, nz, dt, nt, src, sampleRate);
    } else {
        /* the reference's defaults, in cells, then in metres (:196-246) */
        const int nshot_req = arg_int("Nshot", 99999, 0);
        const float fsx = arg_float("fsx", 1.0f, 0) * dx, ds = arg_float("ds", 1.0f, 0) * dx;
        const float sdepth = arg_float("sdepth", 2.0f, 0) * dz;
        const float offsmin = arg_float("offsmin", 0.0f, 0) * dx;
        const float offsmax = arg_float("offsmax", 99999.0f, 0) * dx;
        const float dtr = arg_float("dtr", 1.0f, 0) * dx, gdepth = arg_float("gdepth", 2.0f, 0) * dz;
        float *sx, *sz;
        Nshot = generate_shot_positions(x, nx, z, nz, fsx, ds, sdepth, nshot_req, &sx, &sz);
        obs = (Record2D *)calloc(Nshot, sizeof(Record2D));
        init_observe2d(obs, Nshot, x, nx, z, nz, dt, nt, sx, sz, src, offsmin, offsmax, dtr, gdepth,
                       sampleRate);
        free(sx);
        free(sz);
    }
    Record2D *pre = (Record2D *)calloc(Nshot, sizeof(Record2D));
    create_grid_observe2d(obs, Nshot, dx, dz, nx);
    for (int s = 0; s < Nshot; ++s) {
        pre[s] = obs[s];  /* header copy: geometry pointers shared, fresh traces */
        pre[s].traces_p = alloc2(obs[s].ns, obs[s].ntr);
    }
    fprintf(stdout, "%s\nModeling %d shots, %d x %d, nt = %d ...\n", fwiedu_version(), Nshot, nz,
            nx, nt);
    fflush(stdout);
    const double t0 = now();
    iso_acoustic2d_propagation(vp, rho, obs, Nshot, nz, nx, dx, dz, dt, nt, pml, freesurface, 0, 0,
                               0, 0, 0, verbose, pre, NULL);
    write_acoustic2d_traces((char *)(prefix ? prefix : "shots_p"), pre, Nshot);
    fprintf(stdout, "Modeling takes %.3f s\n", now() - t0);
    return 0;
}
