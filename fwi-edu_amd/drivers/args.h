/* Command-line parameters of the drivers, in Seismic Unix's key=value form
 * (the reference drivers read theirs with SU's getpar family).  Lists are
 * comma separated ("freq=3,5,8"). */
#ifndef FWIEDU_DRIVER_ARGS_H
#define FWIEDU_DRIVER_ARGS_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int g_argc;
static char **g_argv;

static inline void args_init(int argc, char **argv)
{
    g_argc = argc;
    g_argv = argv;
}

/* the value text of key, or NULL; the last occurrence wins */
static inline const char *arg_text(const char *key)
{
    const size_t n = strlen(key);
    const char *v = NULL;
    for (int i = 1; i < g_argc; ++i)
        if (strncmp(g_argv[i], key, n) == 0 && g_argv[i][n] == '=') v = g_argv[i] + n + 1;
    return v;
}

static inline void die(const char *msg, const char *key)
{
    fprintf(stderr, "<error>: %s %s\n", msg, key ? key : "");
    exit(1);
}

static inline int arg_int(const char *key, int def, int required)
{
    const char *v = arg_text(key);
    if (!v) {
        if (required) die("must specify", key);
        return def;
    }
    return atoi(v);
}

static inline float arg_float(const char *key, float def, int required)
{
    const char *v = arg_text(key);
    if (!v) {
        if (required) die("must specify", key);
        return def;
    }
    return (float)atof(v);
}

static inline const char *arg_string(const char *key, int required)
{
    const char *v = arg_text(key);
    if ((!v || !v[0]) && required) die("didn't specify", key);
    return (v && v[0]) ? v : NULL;
}

/* number of comma-separated values of key (0 when absent) */
static inline int arg_count(const char *key)
{
    const char *v = arg_text(key);
    if (!v || !v[0]) return 0;
    int n = 1;
    for (; *v; ++v) n += (*v == ',');
    return n;
}

static inline void arg_floats(const char *key, float *out, int n)
{
    const char *v = arg_text(key);
    for (int i = 0; i < n; ++i) {
        out[i] = (float)atof(v);
        v = strchr(v, ',');
        if (!v) break;
        ++v;
    }
}

static inline void arg_ints(const char *key, int *out, int n)
{
    const char *v = arg_text(key);
    for (int i = 0; i < n; ++i) {
        out[i] = atoi(v);
        v = strchr(v, ',');
        if (!v) break;
        ++v;
    }
}

/* an [n2][n1] float array as a row table over one block (SU alloc2float) */
static inline float **alloc2(int n1, int n2)
{
    float **p = (float **)malloc(sizeof(float *) * (size_t)n2);
    float *b = (float *)calloc((size_t)n1 * n2, sizeof(float));
    if (!p || !b) die("out of memory", NULL);
    for (int i = 0; i < n2; ++i) p[i] = b + (size_t)i * n1;
    return p;
}

static inline void read_floats(const char *path, float *dst, size_t n)
{
    FILE *f = fopen(path, "rb");
    if (!f) die("cannot open", path);
    if (fread(dst, sizeof(float), n, f) != n) die("short read from", path);
    fclose(f);
}

static inline void write_floats(const char *path, const float *src, size_t n)
{
    FILE *f = fopen(path, "wb");
    if (!f) die("cannot create", path);
    if (fwrite(src, sizeof(float), n, f) != n) die("short write to", path);
    fclose(f);
}

/* The model and wavelet files both drivers write at verbose > 0
 * (acoustic2d_modeling.c:327-336, acoustic2d_fwi.c:370-381); tag is "" or
 * "_init". */
static inline void write_inputs(const char *tag, float **vp, float **rho, const float *src,
                                int nz, int nx, float dz, float dx, int nt, float dt)
{
    char name[512];
    snprintf(name, sizeof(name), "vp%s_%d_%d_%.2fm_%.2fm.bin", tag, nz, nx, dz, dx);
    write_floats(name, vp[0], (size_t)nz * nx);
    snprintf(name, sizeof(name), "rho%s_%d_%d_%.2fm_%.2fm.bin", tag, nz, nx, dz, dx);
    write_floats(name, rho[0], (size_t)nz * nx);
    snprintf(name, sizeof(name), "wavelet_%d_%.2fms.bin", nt, dt * 1000);
    write_floats(name, src, (size_t)nt);
}

/* set_waterlayer_value (common.c:898-924): rows iz < wd of every column */
static inline void set_water(float **m, float value, int wd, int nx)
{
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < wd; ++iz) m[ix][iz] = value;
}

#endif
