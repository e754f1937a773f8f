// Library runtime: the one HIP stream of the process, error handling in the reference's wording
// (src/portable.hpp:43-49: "CUDA error in <what>: <msg>" becomes "HIP error in <what>: <msg>"), and a device
// buffer that is zero-filled on allocation like the reference's Buf<T> (src/portable.hpp:104-165).
#pragma once
#include <hip/hip_runtime.h>
#include <stdexcept>
#include <string>
#include <utility>

namespace g3d {

hipStream_t stream();
void set_error(const std::string& msg);
const char* last_error();
void hip_check(hipError_t e, const char* what);
inline void launch_check(const char* what) { hip_check(hipGetLastError(), what); }

template <class F> int guard(F&& f) {
    try { f(); return 0; }
    catch (const std::exception& e) { set_error(e.what()); return 1; }
}

inline unsigned blocks_for(long n, int t) { return unsigned((n + t - 1) / t); }

template <class T> class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void resize(size_t n) {
        release();
        n_ = n;
        if (!n_) return;
        hip_check(hipMalloc(reinterpret_cast<void**>(&p_), n_ * sizeof(T)), "hipMalloc");
        zero();
    }
    void zero() { if (n_) hip_check(hipMemsetAsync(p_, 0, n_ * sizeof(T), stream()), "hipMemsetAsync"); }
    T* data() { return p_; }
    const T* data() const { return p_; }
    size_t size() const { return n_; }
    void to_host(T* h, size_t off, size_t cnt) const {
        if (!cnt) return;
        hip_check(hipMemcpyAsync(h, p_ + off, cnt * sizeof(T), hipMemcpyDeviceToHost, stream()), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(stream()), "synchronize");
    }
    void from_host(const T* h, size_t off, size_t cnt) {
        if (!cnt) return;
        hip_check(hipMemcpyAsync(p_ + off, h, cnt * sizeof(T), hipMemcpyHostToDevice, stream()), "hipMemcpy H2D");
        hip_check(hipStreamSynchronize(stream()), "synchronize");
    }
    void swap(DevBuf& o) noexcept { std::swap(p_, o.p_); std::swap(n_, o.n_); }
private:
    void release() {
        if (p_) { (void)hipStreamSynchronize(stream()); (void)hipFree(p_); }
        p_ = nullptr; n_ = 0;
    }
    T* p_ = nullptr;
    size_t n_ = 0;
};

}  // namespace g3d
