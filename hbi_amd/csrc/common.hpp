// common.hpp -- error handling, device buffers and small block primitives.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdint>
#include <cstdio>
#include <exception>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hbi.h"

namespace hbi {

void set_error(const std::string &msg);

struct Error {
    int code;
    std::string msg;
};

#define HBI_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::hbi::Error{HBI_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};          \
    } while (0)

#define HBI_REQUIRE(cond, msg)                                                                          \
    do {                                                                                                \
        if (!(cond)) throw ::hbi::Error{HBI_E_ARG, (msg)};                                              \
    } while (0)

// run an ABI body, turning exceptions into status codes + hbi_last_error()
template <class F>
int guarded(F &&f)
{
    try {
        f();
        return HBI_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception &e) {
        set_error(e.what());
        return HBI_E_ARG;
    }
}

// device allocation owned by the library
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count)
    {
        release();
        n = count;
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{HBI_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed"};
        }
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Waits for the copy: a pageable host source may be read after
    // hipMemcpyAsync returns, so it must outlive the copy; callers often pass
    // temporaries.
    void upload(const T *h, size_t count, hipStream_t s)
    {
        if (count > n) alloc(count);
        if (count) {
            HBI_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
            HBI_HIP(hipStreamSynchronize(s));
        }
    }
    void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void download(T *h, size_t count, hipStream_t s) const
    {
        if (count) HBI_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

// Host -> device copy of a small host object (e.g. a control block on the
// stack), complete on return.
inline void to_device_sync(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    HBI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HBI_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ blocks --
constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// block-wide sum / max for blockDim.x = NT (multiple of 64); every thread gets the result.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / kWave; i++) s += lds[i];
    return s;
}
template <int NT>
__device__ __forceinline__ double block_max(double v, double *lds)
{
    v = wave_max(v);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) lds[w] = v;
    __syncthreads();
    double s = lds[0];
    for (int i = 1; i < NT / kWave; i++) s = fmax(s, lds[i]);
    return s;
}

// (value, index) argmax / argmin, ties to the lower index; NaN values never win
template <int NT, bool kMax>
__device__ __forceinline__ void block_argext(double &val, int &idx, double *ldsv, int *ldsi)
{
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(val, o, kWave);
        const int i2 = __shfl_xor(idx, o, kWave);
        const bool better = kMax ? (v2 > val || (v2 == val && i2 < idx)) : (v2 < val || (v2 == val && i2 < idx));
        if (i2 >= 0 && (idx < 0 || better)) { val = v2; idx = i2; }
    }
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    __syncthreads();
    if (l == 0) { ldsv[w] = val; ldsi[w] = idx; }
    __syncthreads();
    val = ldsv[0];
    idx = ldsi[0];
    for (int i = 1; i < NT / kWave; i++) {
        const double v2 = ldsv[i];
        const int i2 = ldsi[i];
        const bool better = kMax ? (v2 > val || (v2 == val && i2 < idx)) : (v2 < val || (v2 == val && i2 < idx));
        if (i2 >= 0 && (idx < 0 || better)) { val = v2; idx = i2; }
    }
}

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__host__ __device__ inline uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

}  // namespace hbi
