/* acoustic2d_fwi: multiscale FWI on the GPU, the flow of the reference driver
 * (main/acoustic2d_fwi.c:129-499) with each iteration on the device
 * (fwiedu_session_iterate: gradient, smoothing, perturbation, line-search
 * trial, step length, model update), then the final modeling.
 *
 * Parameters (key=value): nx nz nt dx dz dt srcfile vpfile recordfile Nshot
 * iteration=.. freq=.. filter=.. rx=.. rz=.. (one value per band)
 * [rhofile] [Gardner=1] [freesurface=0] [pmlThick=30] [precondition=1]
 * [waterdepth=0] [vp_water=1500] [rho_water=1000] [vpmin=0.001]
 * [vpmax=10000] [rhomin=0.001] [rhomax=10000] [verbose=0].
 * Outputs: grad_/vp_/rho_iterNNN_*.bin per iteration, objective_function_N.bin,
 * vp_/rho_inversion_*.bin and the predicted gather shots_p_predicted_*.su;
 * with verbose > 0 also vp_init_*, rho_init_*, wavelet_*.bin and the
 * library's per-shot files of each gradient (see fwiedu.h).
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
    const int precondition = arg_int("precondition", 1, 0), verbose = arg_int("verbose", 0, 0);
    const float vpw = arg_float("vp_water", 1500.0f, 0), rhow = arg_float("rho_water", 1000.0f, 0);
    const float vpmin = arg_float("vpmin", 0.001f, 0), vpmax = arg_float("vpmax", 10000.0f, 0);
    const float rhomin = arg_float("rhomin", 0.001f, 0), rhomax = arg_float("rhomax", 10000.0f, 0);
    const char *srcfile = arg_string("srcfile", 1), *vpfile = arg_string("vpfile", 1);
    const char *rhofile = arg_string("rhofile", 0), *recordfile = arg_string("recordfile", 1);
    const int Nshot = arg_int("Nshot", 0, 1);
    const int nb = arg_count("iteration");
    if (nb == 0) die("length of iteration is 0", NULL);
    if (arg_count("freq") != nb) die("length of freq != length of iteration", NULL);
    if (arg_count("filter") != nb) die("length of iteration != length of filter", NULL);
    if (arg_count("rx") != nb) die("length of iteration != length of rx", NULL);
    if (arg_count("rz") != nb) die("length of iteration != length of rz", NULL);
    if (!gardner && !rhofile) die("didn't specify data file of density model:", "rhofile=");
    int *iteration = (int *)malloc(sizeof(int) * nb), *filter = (int *)malloc(sizeof(int) * nb);
    int *rx = (int *)malloc(sizeof(int) * nb), *rz = (int *)malloc(sizeof(int) * nb);
    float *freq = (float *)malloc(sizeof(float) * nb);
    arg_ints("iteration", iteration, nb);
    arg_ints("filter", filter, nb);
    arg_ints("rx", rx, nb);
    arg_ints("rz", rz, nb);
    arg_floats("freq", freq, nb);
    int iterNum = 0;
    for (int k = 0; k < nb; ++k) iterNum += iteration[k];

    const long n = (long)nx * nz;
    float *src = (float *)malloc(sizeof(float) * nt);
    float **vp = alloc2(nz, nx), **rho = alloc2(nz, nx), **grad = alloc2(nz, nx);
    read_floats(srcfile, src, (size_t)nt);
    read_floats(vpfile, vp[0], (size_t)n);
    if (gardner)
        for (long i = 0; i < n; ++i) rho[0][i] = 310.0f * powf(vp[0][i], 0.25f);
    else
        read_floats(rhofile, rho[0], (size_t)n);
    set_water(vp, vpw, wd, nx);
    set_water(rho, rhow, wd, nx);

    /* input_record2d (dataio.c:505-560): gathers, wavelet and its fpeak */
    if (get_shots_number(recordfile) != Nshot)
        die("the number of input shots does not match the gather file", recordfile);
    Record2D *obs = (Record2D *)calloc(Nshot, sizeof(Record2D));
    read_gather2d_from_sufile(obs, recordfile, Nshot, nt);
    const float fpeak = calculate_peak_frequency(src, nt, dt);
    for (int s = 0; s < Nshot; ++s) {
        obs[s].src = src;
        obs[s].fpeak = fpeak;
    }
    if (verbose > 0) write_inputs("_init", vp, rho, src, nz, nx, dz, dx, nt, dt);
    create_grid_observe2d(obs, Nshot, dx, dz, nx);

    fprintf(stdout, "%s\nInversion starts: %d shots, %d x %d, nt = %d, %d iterations\n",
            fwiedu_version(), Nshot, nz, nx, nt, iterNum);
    fflush(stdout);
    void *sess = fwiedu_session_create(vp, rho, obs, Nshot, nz, nx, dx, dz, dt, nt, pml,
                                       freesurface, freq[0], (float)filter[0], precondition);
    float *fun1 = (float *)calloc(iterNum, sizeof(float));
    char name[512];
    int iter = 0;
    for (int k = 0; k < nb; ++k)
        for (int i = 0; i < iteration[k]; ++i) {
            const double t0 = now();
            FwieduIteration it;
            memset(&it, 0, sizeof(it));
            it.freq = freq[k];
            it.filter = (float)filter[k];
            it.rz = rz[k];
            it.rx = rx[k];
            it.waterdepth = wd;
            it.vp_water = vpw;
            it.rho_water = rhow;
            it.vpmin = vpmin;
            it.vpmax = vpmax;
            it.rhomin = rhomin;
            it.rhomax = rhomax;
            it.gardner = gardner ? 1 : 0;
            it.fraction = 0.01f;
            it.iter = iter + 1;
            it.verbose = verbose;
            fwiedu_session_iterate(sess, &it, vp, rho, grad);
            fun1[iter] = it.fun1;
            snprintf(name, sizeof(name), "grad_iter%03d_%05d_%05d_%.2fm_%.2fm.bin", iter + 1, nz, nx,
                     dz, dx);
            write_floats(name, grad[0], (size_t)n);
            snprintf(name, sizeof(name), "vp_iter%03d_%05d_%05d_%.2fm_%.2fm.bin", iter + 1, nz, nx,
                     dz, dx);
            write_floats(name, vp[0], (size_t)n);
            snprintf(name, sizeof(name), "rho_iter%03d_%05d_%05d_%.2fm_%.2fm.bin", iter + 1, nz, nx,
                     dz, dx);
            write_floats(name, rho[0], (size_t)n);
            fprintf(stdout,
                    "Iteration %03d: freq = %4.1fHz, steplength = %e, fun1 = %e, fun2 = %e, takes %f "
                    "seconds.\n",
                    iter + 1, freq[k], it.steplen, it.fun1, it.fun2, now() - t0);
            fflush(stdout);
            ++iter;
        }
    fwiedu_session_destroy(sess);

    /* final predicted data */
    Record2D *pre = (Record2D *)calloc(Nshot, sizeof(Record2D));
    for (int s = 0; s < Nshot; ++s) {
        pre[s] = obs[s];
        pre[s].traces_p = alloc2(obs[s].ns, obs[s].ntr);
    }
    iso_acoustic2d_propagation(vp, rho, obs, Nshot, nz, nx, dx, dz, dt, nt, pml, freesurface, 0, 0,
                               0, 0, 0, verbose, pre, NULL);
    write_acoustic2d_traces("shots_p_predicted", pre, Nshot);
    snprintf(name, sizeof(name), "objective_function_%03d.bin", iterNum);
    write_floats(name, fun1, (size_t)iterNum);
    snprintf(name, sizeof(name), "vp_inversion_%05d_%05d_%.2fm_%.2fm.bin", nz, nx, dz, dx);
    write_floats(name, vp[0], (size_t)n);
    snprintf(name, sizeof(name), "rho_inversion_%05d_%05d_%.2fm_%.2fm.bin", nz, nx, dz, dx);
    write_floats(name, rho[0], (size_t)n);
    return 0;
}
