// The HBM streaming rate of this GPU for the stencil's traffic pattern: c =
// a + b over float32 arrays (two streams read, one written, each element
// touched once) at array sizes inside and beyond the 256 MB Infinity Cache,
// with dword or 16-byte accesses per lane and the default or the
// non-temporal cache policy, and c = a (copy) for comparison.  It prints one
// JSON object: per array size the best add rate over the variants and launch
// sizes, the best copy rate, and the rate of each variant.  Rates count the
// bytes the kernel must move (12 B per add element, 8 B per copy element).
//
//   hbm_stream [reps]
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#define CK(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "hbm_stream: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            exit(1);                                                                  \
        }                                                                             \
    } while (0)

template <bool NT>
__global__ __launch_bounds__(256) void add_dword(const float *a, const float *b, float *c, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        if (NT)
            __builtin_nontemporal_store(
                __builtin_nontemporal_load(a + i) + __builtin_nontemporal_load(b + i), c + i);
        else
            c[i] = a[i] + b[i];
    }
}

__global__ __launch_bounds__(256) void add_float4(const float4 *a, const float4 *b, float4 *c,
                                                  long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const float4 x = a[i], y = b[i];
        c[i] = make_float4(x.x + y.x, x.y + y.y, x.z + y.z, x.w + y.w);
    }
}

__global__ __launch_bounds__(256) void copy_float4(const float4 *a, float4 *c, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        c[i] = a[i];
}

int main(int argc, char **argv)
{
    const int reps = argc > 1 ? atoi(argv[1]) : 3;
    const long sizes_mib[] = {64, 256, 700, 2048};
    const int grids[] = {2048, 32768};
    const char *names[] = {"add_dword", "add_dword_nontemporal", "add_float4", "copy_float4"};
    const long maxn = 2048L << 18;  // floats per array (2 GiB)
    float *a, *b, *c;
    CK(hipMalloc(&a, maxn * 4));
    CK(hipMalloc(&b, maxn * 4));
    CK(hipMalloc(&c, maxn * 4));
    CK(hipMemset(a, 0, maxn * 4));
    CK(hipMemset(b, 0, maxn * 4));
    CK(hipMemset(c, 0, maxn * 4));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    printf("{\"pattern\": \"c = a + b over float32 arrays (12 B per element), copy c = a (8 B)\", "
           "\"array_mib\": [");
    for (int s = 0; s < 4; ++s) printf("%s%ld", s ? ", " : "", sizes_mib[s]);
    printf("], \"sizes\": [");
    for (int s = 0; s < 4; ++s) {
        const long n = sizes_mib[s] << 18;  // floats in the array
        double best[4] = {0, 0, 0, 0};
        for (int gi = 0; gi < 2; ++gi) {
            const int grid = grids[gi];
            for (int rep = 0; rep < reps; ++rep) {
                for (int k = 0; k < 4; ++k) {
                    const int iters = 10;
                    auto run = [&]() {
                        if (k == 0)
                            add_dword<false><<<grid, 256>>>(a, b, c, n);
                        else if (k == 1)
                            add_dword<true><<<grid, 256>>>(a, b, c, n);
                        else if (k == 2)
                            add_float4<<<grid, 256>>>((const float4 *)a, (const float4 *)b,
                                                      (float4 *)c, n / 4);
                        else
                            copy_float4<<<grid, 256>>>((const float4 *)a, (float4 *)c, n / 4);
                    };
                    run();
                    CK(hipGetLastError());
                    CK(hipDeviceSynchronize());
                    CK(hipEventRecord(e0));
                    for (int i = 0; i < iters; ++i) run();
                    CK(hipEventRecord(e1));
                    CK(hipEventSynchronize(e1));
                    float ms = 0;
                    CK(hipEventElapsedTime(&ms, e0, e1));
                    const double bytes = (k == 3 ? 8.0 : 12.0) * n;
                    const double tbs = bytes / (ms / iters * 1e-3) / 1e12;
                    if (tbs > best[k]) best[k] = tbs;
                }
            }
        }
        const double add = best[0] > best[1] ? (best[0] > best[2] ? best[0] : best[2])
                                             : (best[1] > best[2] ? best[1] : best[2]);
        printf("%s{\"array_mib\": %ld, \"add_tb_s\": %.3f, \"copy_tb_s\": %.3f, \"variants_tb_s\": {",
               s ? ", " : "", sizes_mib[s], add, best[3]);
        for (int k = 0; k < 4; ++k) printf("%s\"%s\": %.3f", k ? ", " : "", names[k], best[k]);
        printf("}}");
    }
    printf("]}\n");
    CK(hipFree(a));
    CK(hipFree(b));
    CK(hipFree(c));
    return 0;
}
