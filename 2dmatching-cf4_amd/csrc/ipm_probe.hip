// Bandwidth calibration for the benchmark (not part of the particle path):
// a pure streaming pass with the particle pass's access pattern -- eight
// arrays of doubles read, seven written back, 120 bytes per element, one
// element per thread -- and a plain one-array device copy.  What these reach
// on a device is the practical ceiling of an HBM-bound pass with that
// pattern; the particle pass does the same traffic plus its push, gather,
// deposit and selection work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

namespace {

struct Arrays8 {
    double *f[8];
};

__global__ void __launch_bounds__(256) stream8_kernel(Arrays8 a, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = a.f[k][i];
#pragma unroll
    for (int k = 0; k < 7; ++k) a.f[k][i] = v[k] * 0.5 + v[7];
}

// the same stream with non-temporal loads and stores (as the particle pass
// issues them): they bypass the L2's normal allocation and measure faster
__global__ void __launch_bounds__(256) stream8_nt_kernel(Arrays8 a, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = __builtin_nontemporal_load(&a.f[k][i]);
#pragma unroll
    for (int k = 0; k < 7; ++k) __builtin_nontemporal_store(v[k] * 0.5 + v[7], &a.f[k][i]);
}

__global__ void __launch_bounds__(256) copy_kernel(const double *__restrict__ src, double *__restrict__ dst, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

// f: eight device arrays of n doubles each (contents are overwritten).
// Runs one untimed pass, then `reps` timed passes of each kernel on `stream`;
// ms[0] = median stream8 pass, ms[1] = median copy of f[7] into f[6],
// ms[2] = median non-temporal stream8 pass.
// Returns 0, or the HIP error code.
extern "C" int ipm_probe_stream(double *const *f, long long n, int reps, void *stream, double *ms)
{
    if (!f || !ms || n <= 0 || reps <= 0) return (int)hipErrorInvalidValue;
    for (int k = 0; k < 8; ++k)
        if (!f[k]) return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    Arrays8 a;
    for (int k = 0; k < 8; ++k) a.f[k] = f[k];
    const long long blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffLL) return (int)hipErrorInvalidValue;
    hipEvent_t e0, e1;
    hipError_t rc = hipEventCreate(&e0);
    if (rc != hipSuccess) return (int)rc;
    if ((rc = hipEventCreate(&e1)) != hipSuccess) { (void)hipEventDestroy(e0); return (int)rc; }
    std::vector<float> t8, t1, tn;
    for (int r = 0; r <= reps && rc == hipSuccess; ++r) {
        float x = 0.0f, y = 0.0f, z = 0.0f;
        rc = hipEventRecord(e0, s);
        if (rc == hipSuccess) hipLaunchKernelGGL(stream8_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, n);
        if (rc == hipSuccess) rc = hipGetLastError();
        if (rc == hipSuccess) rc = hipEventRecord(e1, s);
        if (rc == hipSuccess) rc = hipEventSynchronize(e1);
        if (rc == hipSuccess) rc = hipEventElapsedTime(&x, e0, e1);
        if (rc == hipSuccess) rc = hipEventRecord(e0, s);
        if (rc == hipSuccess) hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, f[7], f[6], n);
        if (rc == hipSuccess) rc = hipGetLastError();
        if (rc == hipSuccess) rc = hipEventRecord(e1, s);
        if (rc == hipSuccess) rc = hipEventSynchronize(e1);
        if (rc == hipSuccess) rc = hipEventElapsedTime(&y, e0, e1);
        if (rc == hipSuccess) rc = hipEventRecord(e0, s);
        if (rc == hipSuccess) hipLaunchKernelGGL(stream8_nt_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, n);
        if (rc == hipSuccess) rc = hipGetLastError();
        if (rc == hipSuccess) rc = hipEventRecord(e1, s);
        if (rc == hipSuccess) rc = hipEventSynchronize(e1);
        if (rc == hipSuccess) rc = hipEventElapsedTime(&z, e0, e1);
        if (r > 0) { t8.push_back(x); t1.push_back(y); tn.push_back(z); }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != hipSuccess) return (int)rc;
    std::sort(t8.begin(), t8.end());
    std::sort(t1.begin(), t1.end());
    std::sort(tn.begin(), tn.end());
    ms[0] = t8[t8.size() / 2];
    ms[1] = t1[t1.size() / 2];
    ms[2] = tn[tn.size() / 2];
    return 0;
}
