// hbm_probe.hip - achievable HBM bandwidth of plain streaming kernels on one
// gfx950 GPU, as the practical ceiling the stage kernel is compared with.
//
// Patterns (each stream 504*504*256 doubles = 520 MB, larger than the 256 MB
// last-level cache):
//   copy      b[i] = a[i]                      16-byte lanes (copy2) or 8 (copy1)
//   read      sum of a[i] (result kept)        16-byte lanes
//   write     b[i] = const                     16-byte lanes
//   multi     nin input and nout output streams, 8-byte lanes, one element per
//             stream per thread: the stream mix of the RK stages
//             (stage 1: 5 in / 10 out, stage 2: 10 / 10, stage 3: 10 / 5).
// Each case: one untimed launch, then 10 launches timed with HIP events.
// Output: one line per case, "name  ms  GB/s" (bytes moved / time).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                          \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            std::exit(1);                                                              \
        }                                                                              \
    } while (0)

namespace {

__global__ void k_copy2(const double2* __restrict__ a, double2* __restrict__ b, long n2) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2;
         i += (long)gridDim.x * blockDim.x)
        b[i] = a[i];
}
__global__ void k_copy1(const double* __restrict__ a, double* __restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        b[i] = a[i];
}
__global__ void k_read2(const double2* __restrict__ a, double* __restrict__ out, long n2) {
    double s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2;
         i += (long)gridDim.x * blockDim.x) {
        const double2 v = a[i];
        s += v.x + v.y;
    }
    if (s == 1234.5) out[0] = s;  // keeps the loads; never true for the zero-filled input
}
__global__ void k_write2(double2* __restrict__ b, long n2) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2;
         i += (long)gridDim.x * blockDim.x)
        b[i] = make_double2(1.0, 2.0);
}
struct Streams {
    const double* in[10];
    double* out[10];
};
__global__ void k_multi(Streams p, int nin, int nout, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int f = 0; f < nin; ++f) s += p.in[f][i];
        for (int f = 0; f < nout; ++f) p.out[f][i] = s + f;
    }
}

}  // namespace

int main() {
    const long n = 504L * 504 * 256;  // doubles per stream
    std::vector<double*> buf(20);
    for (auto& b : buf) {
        CK(hipMalloc(&b, n * 8));
        CK(hipMemset(b, 0, n * 8));
    }
    double* out;
    CK(hipMalloc(&out, 64));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    auto timeit = [&](const char* name, double bytes, auto launch) {
        launch();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(e0));
        const int reps = 10;
        for (int r = 0; r < reps; ++r) launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        std::printf("%-28s %8.3f ms  %7.1f GB/s\n", name, ms / reps,
                    bytes / (ms / reps * 1e-3) / 1e9);
        std::fflush(stdout);
    };
    const long n2 = n / 2;
    char nm[64];
    for (int grid : {1024, 4096, 16384}) {
        std::snprintf(nm, sizeof nm, "copy2 grid %d", grid);
        timeit(nm, 16.0 * n, [&] {
            k_copy2<<<grid, 256>>>((const double2*)buf[0], (double2*)buf[1], n2);
        });
        std::snprintf(nm, sizeof nm, "copy1 grid %d", grid);
        timeit(nm, 16.0 * n, [&] { k_copy1<<<grid, 256>>>(buf[0], buf[1], n); });
        std::snprintf(nm, sizeof nm, "read2 grid %d", grid);
        timeit(nm, 8.0 * n, [&] { k_read2<<<grid, 256>>>((const double2*)buf[0], out, n2); });
        std::snprintf(nm, sizeof nm, "write2 grid %d", grid);
        timeit(nm, 8.0 * n, [&] { k_write2<<<grid, 256>>>((double2*)buf[1], n2); });
    }
    Streams p;
    for (int f = 0; f < 10; ++f) {
        p.in[f] = buf[f];
        p.out[f] = buf[10 + f];
    }
    const int mix[][2] = {{5, 10}, {10, 10}, {10, 5}, {10, 0}, {1, 10}};
    for (const auto& m : mix)
        for (int grid : {2048, 8192}) {
            std::snprintf(nm, sizeof nm, "multi in %d out %d grid %d", m[0], m[1], grid);
            timeit(nm, 8.0 * n * (m[0] + m[1]), [&] { k_multi<<<grid, 256>>>(p, m[0], m[1], n); });
        }
    for (auto& b : buf) CK(hipFree(b));
    CK(hipFree(out));
    std::printf("done\n");
    return 0;
}
