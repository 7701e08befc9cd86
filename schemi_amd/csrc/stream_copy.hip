// stream_copy.hip -- a plain device-to-device streaming copy, the bandwidth
// calibration for the stage kernels' roofline figures and for the rocprofv3
// byte counters (FETCH_SIZE / WRITE_SIZE).  Not part of the step's C ABI:
// bench.py loads it as libschemi_copy.so.
//
// One pass: each workgroup of 256 lanes moves 1024 16-byte words (four
// non-temporal loads in flight per lane, then four non-temporal stores), one
// workgroup per 16 KiB of the buffer.  On MI355X this streams 6.6 TB/s
// (read + write) over 1 GiB, against 4.9-5.0 TB/s for a grid-stride loop
// over 8-32 workgroups per CU and 5.1 TB/s for the runtime's own copy.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// a native 16-byte vector (the non-temporal builtins take clang vectors)
typedef unsigned int word4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy_kernel(const word4 *__restrict__ src,
                                                   word4 *__restrict__ dst, int64_t n)
{
    const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    word4 w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (base + 256 * j < n)
            w[j] = __builtin_nontemporal_load(&src[base + 256 * j]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (base + 256 * j < n)
            __builtin_nontemporal_store(w[j], &dst[base + 256 * j]);
}

}  // namespace

// Copies `bytes` (a multiple of 16; both pointers 16-byte aligned device
// memory) from src to dst `reps` times on a stream of its own and stores the
// fastest copy's duration (HIP events on that stream) in *best_ms.  Returns
// 0, or -1 for bad arguments, or the HIP error code.
extern "C" int schemi_stream_copy(const void *src, void *dst, int64_t bytes, int reps,
                                  float *best_ms)
{
    if (!src || !dst || !best_ms || bytes <= 0 || bytes % 16 || reps < 1 ||
        (reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16)
        return -1;
    if ((bytes / 16 + 1023) / 1024 > 0x7fffffff)
        return -1;
    hipError_t e;
    hipStream_t s;
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess)
        return (int)e;
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return (int)e;
    }
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        (void)hipStreamDestroy(s);
        return (int)e;
    }
    const int64_t n = bytes / 16;
    const unsigned grid = (unsigned)((n + 1023) / 1024);
    float best = 0.f;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        e = hipEventRecord(e0, s);
        if (e != hipSuccess)
            break;
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, s,
                           static_cast<const word4 *>(src), static_cast<word4 *>(dst), n);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipEventRecord(e1, s);
        if (e == hipSuccess)
            e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
            (r == 0 || ms < best))
            best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    *best_ms = best;
    return (int)e;
}
