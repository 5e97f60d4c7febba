// Host-side vocabulary of the DEMSolver / WCSPHSolver facade: pi(),
// symMatrix, HostDeviceArray1D and status checking.  HIP's own double3 /
// int3 types and their arithmetic operators are used as they are.
//
// Reference: DEMSPH/kernel/myUtility/myVec.h:13 (pi), myMat.h:15-33
// (symMatrix, make_symMatrix), myHostDeviceArray.h:89-286
// (HostDeviceArray1D).  The names and member functions are the reference's
// so that application code written against it compiles unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../demsph.h"

constexpr double pi() { return 3.14159265358979323846; }

struct symMatrix {
    double xx, yy, zz;
    double xy, xz, yz;
    symMatrix() : xx(0.0), yy(0.0), zz(0.0), xy(0.0), xz(0.0), yz(0.0) {}
    symMatrix(double a, double b, double c, double d, double e, double f) : xx(a), yy(b), zz(c), xy(d), xz(e), yz(f)
    {
    }
};

inline symMatrix make_symMatrix(double xx, double yy, double zz, double xy, double xz, double yz)
{
    return symMatrix(xx, yy, zz, xy, xz, yz);
}

namespace demsph {

// Throws std::runtime_error carrying the library's message for a non-OK status.
inline void check(int status)
{
    if (status != DEMSPH_OK) throw std::runtime_error(demsph_last_error());
}

inline void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace demsph

// A host vector with a device mirror (the hook inputs of the applications,
// e.g. a per-ball external force uploaded at every hook call).  Device memory
// is released synchronously in the destructor.  Copies are not allowed;
// moves are.
template <class T>
struct HostDeviceArray1D {
    std::vector<T> h_data;
    size_t d_size{0};
    T* d_ptr{nullptr};

    HostDeviceArray1D() = default;
    ~HostDeviceArray1D() { releaseDeviceSync(); }
    HostDeviceArray1D(const HostDeviceArray1D&) = delete;
    HostDeviceArray1D& operator=(const HostDeviceArray1D&) = delete;
    HostDeviceArray1D(HostDeviceArray1D&& o) noexcept : h_data(std::move(o.h_data)), d_size(o.d_size), d_ptr(o.d_ptr)
    {
        o.d_ptr = nullptr;
        o.d_size = 0;
    }
    HostDeviceArray1D& operator=(HostDeviceArray1D&& o) noexcept
    {
        if (this != &o) {
            releaseDeviceSync();
            h_data = std::move(o.h_data);
            d_size = o.d_size;
            d_ptr = o.d_ptr;
            o.d_ptr = nullptr;
            o.d_size = 0;
        }
        return *this;
    }

    size_t hostSize() const { return h_data.size(); }
    size_t deviceSize() const { return d_size; }

    void releaseDevice(hipStream_t stream)
    {
        if (d_ptr) {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(d_ptr);
        }
        d_ptr = nullptr;
        d_size = 0;
    }
    void releaseDeviceSync()
    {
        if (d_ptr) {
            (void)hipDeviceSynchronize();
            (void)hipFree(d_ptr);
        }
        d_ptr = nullptr;
        d_size = 0;
    }
    void allocateDevice(size_t n, hipStream_t stream, bool zeroFill = true)
    {
        if (n != d_size) {
            releaseDevice(stream);
            if (n) demsph::check_hip(hipMalloc(&d_ptr, n * sizeof(T)), "hipMalloc");
            d_size = n;
        }
        if (zeroFill && n) demsph::check_hip(hipMemsetAsync(d_ptr, 0, n * sizeof(T), stream), "hipMemsetAsync");
    }

    void pushHost(const T& v) { h_data.push_back(v); }
    void insertHost(size_t i, const T& v) { h_data.insert(h_data.begin() + (std::ptrdiff_t)i, v); }
    void eraseHost(size_t i) { h_data.erase(h_data.begin() + (std::ptrdiff_t)i); }
    void clearHost() { h_data.clear(); }
    void reserveHost(size_t n) { h_data.reserve(n); }
    void resizeHost(size_t n) { h_data.resize(n); }
    const std::vector<T>& hostRef() const { return h_data; }
    void setHost(const std::vector<T>& v) { h_data = v; }

    void copyHostToDevice(hipStream_t stream)
    {
        allocateDevice(h_data.size(), stream, false);
        if (!h_data.empty())
            demsph::check_hip(hipMemcpyAsync(d_ptr, h_data.data(), h_data.size() * sizeof(T), hipMemcpyHostToDevice,
                                             stream),
                              "hipMemcpyAsync");
    }
    void copyDeviceToHost(hipStream_t stream)
    {
        h_data.resize(d_size);
        if (d_size) {
            demsph::check_hip(hipMemcpyAsync(h_data.data(), d_ptr, d_size * sizeof(T), hipMemcpyDeviceToHost, stream),
                              "hipMemcpyAsync");
            demsph::check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
        }
    }
    std::vector<T> getHostCopy()
    {
        std::vector<T> h(d_size);
        if (d_size) {
            demsph::check_hip(hipMemcpy(h.data(), d_ptr, d_size * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        }
        return h;
    }
};
