// Owning device allocation (hipMalloc / hipFree), move-free and copy-free.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace nuto_amd
{

struct DevBuf
{
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf()
    {
        Release();
    }
    void Release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t Alloc(size_t n)
    {
        Release();
        hipError_t e = hipMalloc(&p, n ? n : 16);
        if (e != hipSuccess)
            p = nullptr;
        else
            bytes = n;
        return e;
    }
    // upload of n bytes on `stream` (allocates), complete when this returns
    hipError_t Upload(const void* src, size_t n, hipStream_t stream)
    {
        hipError_t e = Alloc(n);
        if (e == hipSuccess && n)
            e = hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(stream);
        return e;
    }
    template <class T>
    T* As() const
    {
        return static_cast<T*>(p);
    }
};

}  // namespace nuto_amd
