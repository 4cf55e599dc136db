// lb3d_bwprobe.hip -- the practical HBM ceiling of the card, for the benchmark:
// a 16-byte-per-lane grid-stride copy, a read-only sweep and a write-only
// sweep over buffers far larger than the caches, and a copy in the population
// layout (one thread per site, 38 or 57 planes), timed with HIP events.  Not
// part of the timestep library; bench.py loads it as liblb3d_bwprobe.so.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// every thread writes its partial sum, so the reads cannot be dropped
__global__ __launch_bounds__(256) void k_read(const float4 *__restrict__ a, float *__restrict__ out, size_t n)
{
    float s = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float4 v = a[i];
        s += (v.x + v.y) + (v.z + v.w);
    }
    out[blockIdx.x * (size_t)blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_write(float4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = make_float4(1.f, 2.f, 3.f, 4.f);
}

// the population layout's access pattern: one thread per site reads its K
// values from K planes ns apart and writes them to K planes (a step's
// population traffic without the arithmetic or the neighbour offsets)
template <int K>
__global__ __launch_bounds__(256) void k_planes(const float *__restrict__ a, float *__restrict__ b, uint32_t ns)
{
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= ns) return;
    float v[K];
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = a[(size_t)k * ns + s];
#pragma unroll
    for (int k = 0; k < K; k++) b[(size_t)k * ns + s] = v[k] * 1.0001f;
}

}  // namespace

// Rate in GB/s (bytes read and written) of the plane copy k_planes over
// `bytes` bytes per buffer, for planes = 38 (binary FP32 populations) or 57
// (ternary).  Returns 0, or a HIP error code.
extern "C" int lb3d_bw_probe_planes(size_t bytes, int planes, int reps, double *out)
{
    if ((planes != 38 && planes != 57) || reps < 1 || !out) return (int)hipErrorInvalidValue;
    const uint32_t ns = (uint32_t)(bytes / sizeof(float) / (size_t)planes);
    if (ns < 256) return (int)hipErrorInvalidValue;
    const size_t n = (size_t)planes * ns;  // floats per buffer, <= bytes / 4
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t rc = hipMalloc(&a, n * sizeof(float));
    if (rc == hipSuccess) rc = hipMalloc(&b, n * sizeof(float));
    if (rc == hipSuccess) rc = hipMemset(a, 0, n * sizeof(float));
    if (rc == hipSuccess) rc = hipEventCreate(&e0);
    if (rc == hipSuccess) rc = hipEventCreate(&e1);
    if (rc == hipSuccess) {
        const unsigned nb = (ns + 255u) / 256u;
        auto go = [&]() {
            if (planes == 57)
                k_planes<57><<<nb, 256>>>((const float *)a, (float *)b, ns);
            else
                k_planes<38><<<nb, 256>>>((const float *)a, (float *)b, ns);
        };
        go();
        rc = hipEventRecord(e0, 0);
        for (int r = 0; r < reps && rc == hipSuccess; r++) go();
        if (rc == hipSuccess) rc = hipEventRecord(e1, 0);
        if (rc == hipSuccess) rc = hipEventSynchronize(e1);
        float ms = 0.f;
        if (rc == hipSuccess) rc = hipEventElapsedTime(&ms, e0, e1);
        if (rc == hipSuccess) rc = hipGetLastError();
        if (rc == hipSuccess) *out = 2.0 * (double)n * sizeof(float) * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc == hipSuccess ? 0 : (int)rc;
}

// Best rate in GB/s over a few grid sizes of each sweep of `bytes` bytes per
// buffer (copy counts the bytes read and written): out[0] copy, out[1] read,
// out[2] write.  Returns 0, or a HIP error code.
extern "C" int lb3d_bw_probe(size_t bytes, int reps, double *out)
{
    const int grids[] = {1024, 2048, 4096, 65536};
    // the read sweep stores one float per thread of the largest grid into b
    if (bytes < (size_t)65536 * 256 * sizeof(float) || reps < 1 || !out) return (int)hipErrorInvalidValue;
    const size_t n4 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t rc = hipMalloc(&a, n4 * 16);
    if (rc == hipSuccess) rc = hipMalloc(&b, n4 * 16);
    if (rc == hipSuccess) rc = hipMemset(a, 0, n4 * 16);
    if (rc == hipSuccess) rc = hipEventCreate(&e0);
    if (rc == hipSuccess) rc = hipEventCreate(&e1);
    for (int kind = 0; kind < 3 && rc == hipSuccess; kind++) {
        double best = 0.0;
        for (int grid : grids) {
            // the read sweep's partial sums (grid x 256 floats) go to b
            auto go = [&]() {
                if (kind == 0)
                    k_copy<<<grid, 256>>>((const float4 *)a, (float4 *)b, n4);
                else if (kind == 1)
                    k_read<<<grid, 256>>>((const float4 *)a, (float *)b, n4);
                else
                    k_write<<<grid, 256>>>((float4 *)b, n4);
            };
            go();
            rc = hipEventRecord(e0, 0);
            for (int r = 0; r < reps && rc == hipSuccess; r++) go();
            if (rc == hipSuccess) rc = hipEventRecord(e1, 0);
            if (rc == hipSuccess) rc = hipEventSynchronize(e1);
            float ms = 0.f;
            if (rc == hipSuccess) rc = hipEventElapsedTime(&ms, e0, e1);
            if (rc == hipSuccess) rc = hipGetLastError();
            if (rc != hipSuccess) break;
            double moved = (double)(n4 * 16) * (kind == 0 ? 2.0 : 1.0) * reps;
            double gbs = moved / (ms * 1e-3) / 1e9;
            if (gbs > best) best = gbs;
        }
        out[kind] = best;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc == hipSuccess ? 0 : (int)rc;
}
