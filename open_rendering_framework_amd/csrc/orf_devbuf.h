// orf_devbuf.h -- owning device allocation used by the host runtime.
#pragma once

#include <vector>

#include "orf_common.h"

namespace orf {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t n)
    {
        if (n <= bytes && p) return;
        release();
        if (n == 0) n = 16;
        ORF_HIP_CHECK(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T>
    void upload(const std::vector<T>& v)
    {
        ensure(v.size() * sizeof(T));
        if (!v.empty()) ORF_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    template <class T>
    T* as() const
    {
        return (T*)p;
    }
};

}  // namespace orf
