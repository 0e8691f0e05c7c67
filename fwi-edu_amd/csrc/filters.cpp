// FFT work on traces on the device: the band filters apply_traces_lowpass /
// apply_traces_bandpass (common.c:584-794) as used by compute_adjoint_source2d
// with filter 0 / 1 (optimization.c:368-418), and interpolate_traces
// (common.c:482-581) for adjoint sources recorded at sampleRate > 1.
//
// The reference transforms each trace with FFTW single precision (r2c), scales
// the first n/2+1 bins by a cosine window, transforms back (c2r) and divides by
// n.  Here the batch of traces goes through one hipFFT R2C / C2R plan pair (FP32,
// unnormalised like FFTW), the window and the division are kernels, and the
// windows are evaluated on the host with the reference's float expressions.
#include <hipfft/hipfft.h>

#include "engine.hpp"

namespace fwiedu {

#define FFTCHECK(x)                                                                   \
    do {                                                                              \
        hipfftResult r_ = (x);                                                        \
        if (r_ != HIPFFT_SUCCESS) fail("<error>: hipFFT error %d at %s:%d", (int)r_,  \
                                       __FILE__, __LINE__);                          \
    } while (0)

// Window of apply_traces_lowpass (kind 0, common.c:615-640) or
// apply_traces_bandpass (kind 1, common.c:726-752) for n samples at step dt.
std::vector<float> filter_window(int kind, int n, float fpass, float fcut, float dt)
{
    if (!(fcut > fpass)) fail("<error>: fcut must be greater than fpass");
    const double PI = 3.141592653589793;
    const float df = 1.0f / ((float)n * dt);
    std::vector<float> w(n, 0.0f);
    for (int i = 0; i < n; ++i) {
        const float f = (float)i * df;
        if (kind == 0) {
            const float L = fcut - fpass;
            if (f < fpass)
                w[i] = 1.0f;
            else if (f <= fcut)
                w[i] = 0.5f * (1.0f + cosf((float)PI * (f - fpass) / L));
        } else {
            // the band-pass forms its cosine argument in double (PI is double)
            if (f < fpass)
                w[i] = 0.5f * (1.0f + cosf((float)(PI * (fpass - f) / fpass)));
            else if (f <= fcut)
                w[i] = 0.5f * (1.0f + cosf((float)(PI * (f - fpass) / (fcut - fpass))));
        }
    }
    return w;
}

static std::pair<void *, void *> plans_for(Ctx &c, int n, int batch)
{
    auto key = std::make_pair(n, batch);
    auto it = c.fft_plans.find(key);
    if (it != c.fft_plans.end()) return it->second;
    hipfftHandle f = nullptr, b = nullptr;
    FFTCHECK(hipfftPlan1d(&f, n, HIPFFT_R2C, batch));
    FFTCHECK(hipfftPlan1d(&b, n, HIPFFT_C2R, batch));
    auto p = std::make_pair((void *)f, (void *)b);
    c.fft_plans[key] = p;
    return p;
}

// x: [batch][n] float on the device, filtered in place on the library stream.
void filter_traces_device(int kind, float *x, int n, int batch, float fpass, float fcut, float dt)
{
    Ctx &c = ctx();
    if (n <= 0 || batch <= 0) return;
    const int nb = n / 2 + 1;
    std::vector<float> w = filter_window(kind, n, fpass, fcut, dt);
    float *wd = c.fftwin.as<float>(nb + 64);
    h2d(wd, w.data(), nb * sizeof(float));
    float2 *spec = c.fftspec.as<float2>((size_t)batch * nb + 64);
    auto p = plans_for(c, n, batch);
    hipfftHandle f = (hipfftHandle)p.first, b = (hipfftHandle)p.second;
    FFTCHECK(hipfftSetStream(f, c.st));
    FFTCHECK(hipfftSetStream(b, c.st));
    FFTCHECK(hipfftExecR2C(f, (hipfftReal *)x, (hipfftComplex *)spec));
    launch_spec_window(spec, (long)batch * nb, nb, wd, c.st);
    FFTCHECK(hipfftExecC2R(b, (hipfftComplex *)spec, (hipfftReal *)x));
    launch_div(x, (long)batch * n, (float)n, c.st);
}

// C2C plans are keyed (-n, batch) beside the (n, batch) R2C / C2R pairs.
static hipfftHandle c2c_plan(Ctx &c, int n, int batch)
{
    auto key = std::make_pair(-n, batch);
    auto it = c.fft_plans.find(key);
    if (it != c.fft_plans.end()) return (hipfftHandle)it->second.first;
    hipfftHandle p = nullptr;
    FFTCHECK(hipfftPlan1d(&p, n, HIPFFT_C2C, batch));
    c.fft_plans[key] = std::make_pair((void *)p, (void *)nullptr);
    return p;
}

void interpolate_traces_device(const float *in, int ntr, int ns, int r, float *out, int nt_out)
{
    Ctx &c = ctx();
    if (ntr <= 0 || ns <= 0) fail("<error>: invalid input (ntr=%d, nt=%d) in function interpolate_traces!", ntr, ns);
    if (r <= 1) {
        // common.c:511-516 copies (the caller has ns == nt_out)
        if (ns > nt_out) fail("<error>: interpolate_traces: %d samples into %d", ns, nt_out);
        HIPCHECK(hipMemcpy2DAsync(out, (size_t)nt_out * 4, in, (size_t)ns * 4, (size_t)ns * 4, ntr,
                                  hipMemcpyDeviceToDevice, c.st));
        if (ns < nt_out)
            HIPCHECK(hipMemset2DAsync(out + ns, (size_t)nt_out * 4, 0, (size_t)(nt_out - ns) * 4,
                                      ntr, c.st));
        return;
    }
    const int M = (ns - 1) * r + 1;
    if (M > nt_out) fail("<error>: interpolate_traces: %d samples into %d", M, nt_out);
    float2 *buf = c.fftspec.as<float2>((size_t)ntr * M + 64);
    launch_zero_stuff(in, ntr, ns, r, buf, M, c.st);
    hipfftHandle p = c2c_plan(c, M, ntr);
    FFTCHECK(hipfftSetStream(p, c.st));
    FFTCHECK(hipfftExecC2C(p, (hipfftComplex *)buf, (hipfftComplex *)buf, HIPFFT_FORWARD));
    launch_interp_window(buf, ntr, M, r, c.st);
    FFTCHECK(hipfftExecC2C(p, (hipfftComplex *)buf, (hipfftComplex *)buf, HIPFFT_BACKWARD));
    launch_interp_out(buf, ntr, M, out, nt_out, c.st);
}

void release_fft(Ctx &c)
{
    for (auto &e : c.fft_plans) {
        if (e.second.first) (void)hipfftDestroy((hipfftHandle)e.second.first);
        if (e.second.second) (void)hipfftDestroy((hipfftHandle)e.second.second);
    }
    c.fft_plans.clear();
}

}  // namespace fwiedu
