// probe.hip -- the HBM rates a hand-written streaming kernel reaches on this
// GPU: copy (read + write), read only and write only, over buffers far
// larger than the caches.  They set the achievable ceiling that the step's
// kernels are compared with (beside the 8 TB/s vendor peak), and the copy is
// the known-size kernel that calibrates the HBM byte counters.  And the
// vector FMA issue rate in FP32 and FP64, which bounds the stencil kernels'
// arithmetic in the two precision forms.
#include "ps3d_internal.h"

#include <algorithm>

using namespace ps3d;

namespace {

// a 16-byte vector of the compiler's own (the nontemporal builtins take it)
typedef double vec2 __attribute__((ext_vector_type(2)));

constexpr int PT = 256;     // threads per work-group
constexpr int UNR = 4;      // 16-byte vectors per thread per pass

// each work-group moves PT * UNR vectors per pass, grid-strided
template <bool NT>
__global__ __launch_bounds__(PT) void k_copy(const vec2 *__restrict__ src, vec2 *__restrict__ dst, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * PT * UNR;
    for (int64_t base = (int64_t)blockIdx.x * PT * UNR + threadIdx.x; base < n; base += stride) {
        vec2 v[UNR];
#pragma unroll
        for (int q = 0; q < UNR; ++q) {
            const int64_t i = base + (int64_t)q * PT;
            if (i < n) v[q] = NT ? __builtin_nontemporal_load(src + i) : src[i];
        }
#pragma unroll
        for (int q = 0; q < UNR; ++q) {
            const int64_t i = base + (int64_t)q * PT;
            if (i < n) {
                if (NT) __builtin_nontemporal_store(v[q], dst + i);
                else dst[i] = v[q];
            }
        }
    }
}

template <bool NT>
__global__ __launch_bounds__(PT) void k_read(const vec2 *__restrict__ src, double *__restrict__ sink, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * PT * UNR;
    double acc = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * PT * UNR + threadIdx.x; base < n; base += stride) {
#pragma unroll
        for (int q = 0; q < UNR; ++q) {
            const int64_t i = base + (int64_t)q * PT;
            if (i < n) {
                const vec2 v = NT ? __builtin_nontemporal_load(src + i) : src[i];
                acc += v.x + v.y;
            }
        }
    }
    // one value per thread that is never zero in practice keeps the loads
    if (acc == 1.2345e300) sink[(int64_t)blockIdx.x * PT + threadIdx.x] = acc;
}

template <bool NT>
__global__ __launch_bounds__(PT) void k_write(vec2 *__restrict__ dst, int64_t n, double val)
{
    const int64_t stride = (int64_t)gridDim.x * PT * UNR;
    const vec2 v = {val, val};
    for (int64_t base = (int64_t)blockIdx.x * PT * UNR + threadIdx.x; base < n; base += stride) {
#pragma unroll
        for (int q = 0; q < UNR; ++q) {
            const int64_t i = base + (int64_t)q * PT;
            if (i < n) {
                if (NT) __builtin_nontemporal_store(v, dst + i);
                else dst[i] = v;
            }
        }
    }
}

// vector issue rate: FMI independent FMA chains per lane, so the SIMDs issue
// back to back; T = float is compiled, like every kernel here, without the
// packed FP32 instructions
constexpr int FMI = 8;
template <class T>
__global__ __launch_bounds__(PT) void k_fma(T *__restrict__ sink, T a, T b, int iters)
{
    T x[FMI];
#pragma unroll
    for (int q = 0; q < FMI; ++q) x[q] = (T)(threadIdx.x + q);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int q = 0; q < FMI; ++q) {
            if constexpr (sizeof(T) == 4) x[q] = __builtin_fmaf(x[q], a, b);
            else x[q] = __builtin_fma(x[q], a, b);
        }
    }
    T s = T(0);
#pragma unroll
    for (int q = 0; q < FMI; ++q) s += x[q];
    sink[(int64_t)blockIdx.x * PT + threadIdx.x] = s;
}

// best of reps launches, in seconds
template <class F>
double best_time(hipStream_t s, int reps, F launch)
{
    hipEvent_t e0, e1;
    PS3D_HIP_CHECK(hipEventCreate(&e0));
    PS3D_HIP_CHECK(hipEventCreate(&e1));
    double best = 1e30;
    for (int r = 0; r < reps + 1; ++r) {   // the first launch is a warm-up
        PS3D_HIP_CHECK(hipEventRecord(e0, s));
        launch();
        PS3D_HIP_CHECK(hipGetLastError());
        PS3D_HIP_CHECK(hipEventRecord(e1, s));
        PS3D_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        PS3D_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0) best = std::min(best, 1e-3 * (double)ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

extern "C" int ps3d_hbm_probe(int64_t bytes, int reps, int grid, double *out)
{
    return guard([&] {
        PS3D_REQUIRE(out, "null output");
        PS3D_REQUIRE(bytes >= (1 << 20) && bytes % 16 == 0, "bytes: a multiple of 16, at least 1 MiB");
        PS3D_REQUIRE(reps >= 1 && reps <= 1000, "reps: 1 .. 1000");
        int dev = 0, ncu = 0;
        PS3D_HIP_CHECK(hipGetDevice(&dev));
        PS3D_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int64_t n = bytes / 16;
        const int64_t full = (n + PT * UNR - 1) / (PT * UNR);
        // default: 8 work-groups per CU, grid-strided (enough waves in flight to
        // cover HBM latency); grid > 0 sets the count (capped at one pass)
        const int nb = (int)std::min<int64_t>(full, grid > 0 ? grid : 8 * std::max(ncu, 1));
        vec2 *a = nullptr, *b = nullptr;
        double *sink = nullptr;
        hipStream_t s = nullptr;
        PS3D_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            PS3D_HIP_CHECK(hipMalloc(&a, (size_t)bytes));
            PS3D_HIP_CHECK(hipMalloc(&b, (size_t)bytes));
            PS3D_HIP_CHECK(hipMalloc(&sink, sizeof(double) * (size_t)nb * PT));
            hipLaunchKernelGGL(k_write<false>, dim3(nb), dim3(PT), 0, s, a, n, 1.0);
            PS3D_HIP_CHECK(hipGetLastError());
            const double B = (double)bytes;
            const double tc = best_time(s, reps, [&] { hipLaunchKernelGGL(k_copy<false>, dim3(nb), dim3(PT), 0, s, a, b, n); });
            const double tcn = best_time(s, reps, [&] { hipLaunchKernelGGL(k_copy<true>, dim3(nb), dim3(PT), 0, s, a, b, n); });
            const double tr = best_time(s, reps, [&] { hipLaunchKernelGGL(k_read<false>, dim3(nb), dim3(PT), 0, s, a, sink, n); });
            const double trn = best_time(s, reps, [&] { hipLaunchKernelGGL(k_read<true>, dim3(nb), dim3(PT), 0, s, a, sink, n); });
            const double tw = best_time(s, reps, [&] { hipLaunchKernelGGL(k_write<false>, dim3(nb), dim3(PT), 0, s, b, n, 2.0); });
            const double twn = best_time(s, reps, [&] { hipLaunchKernelGGL(k_write<true>, dim3(nb), dim3(PT), 0, s, b, n, 2.0); });
            // GB/s: copy counts read + write bytes
            out[0] = 2.0 * B / tc * 1e-9;
            out[1] = 2.0 * B / tcn * 1e-9;
            out[2] = B / tr * 1e-9;
            out[3] = B / trn * 1e-9;
            out[4] = B / tw * 1e-9;
            out[5] = B / twn * 1e-9;
            out[6] = (double)nb;
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(a);
            (void)hipFree(b);
            (void)hipFree(sink);
            (void)hipStreamDestroy(s);
            throw;
        }
        PS3D_HIP_CHECK(hipStreamSynchronize(s));
        (void)hipFree(a);
        (void)hipFree(b);
        (void)hipFree(sink);
        (void)hipStreamDestroy(s);
    });
}

extern "C" int ps3d_fma_probe(int reps, double *out)
{
    return guard([&] {
        PS3D_REQUIRE(out, "null output");
        PS3D_REQUIRE(reps >= 1 && reps <= 1000, "reps: 1 .. 1000");
        int dev = 0, ncu = 0;
        PS3D_HIP_CHECK(hipGetDevice(&dev));
        PS3D_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int nb = 32 * std::max(ncu, 1), iters = 4096;
        double *sink = nullptr;
        hipStream_t s = nullptr;
        PS3D_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            PS3D_HIP_CHECK(hipMalloc(&sink, sizeof(double) * (size_t)nb * PT));
            const double flops = 2.0 * FMI * iters * (double)nb * PT;
            const double t32 = best_time(s, reps, [&] {
                hipLaunchKernelGGL(k_fma<float>, dim3(nb), dim3(PT), 0, s, reinterpret_cast<float *>(sink), 0.999f,
                                   0.001f, iters);
            });
            const double t64 = best_time(s, reps, [&] {
                hipLaunchKernelGGL(k_fma<double>, dim3(nb), dim3(PT), 0, s, sink, 0.999, 0.001, iters);
            });
            out[0] = flops / t32 * 1e-12;
            out[1] = flops / t64 * 1e-12;
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(sink);
            (void)hipStreamDestroy(s);
            throw;
        }
        PS3D_HIP_CHECK(hipStreamSynchronize(s));
        (void)hipFree(sink);
        (void)hipStreamDestroy(s);
    });
}
