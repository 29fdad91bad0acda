// hip_util.hpp -- host helpers shared by the C-ABI sources: HIP status checks
// that throw, and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace axg {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                                   \
    do {                                                                                              \
        hipError_t err__ = (x);                                                                       \
        if (err__ != hipSuccess)                                                                      \
            throw HipError(std::string("HIP || ") + #x + ": " + hipGetErrorString(err__));            \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    void alloc(size_t count) {
        free();
        n = count;
        if (count) HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) HIPCHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
    void zero(hipStream_t s) {
        if (n) HIPCHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
};

}  // namespace axg
