// HIP runtime error checking and device allocation helpers for host code.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace sc {

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw std::runtime_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));      \
    } while (0)

template <class T>
static void dev_alloc(T*& p, size_t n)
{
    if (p) HIP_OK(hipFree(p));
    p = nullptr;
    if (n) HIP_OK(hipMalloc((void**)&p, n * sizeof(T)));
}

template <class T>
static void dev_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace sc
