// Batched transverse FFT (rows, columns) and the runtime/memory entry points of the C ABI.
#include <cstdlib>

#include "g3d_fft.hpp"
#include "g3d_fft_core.hpp"
#include "g3d_runtime.hpp"
#include <cmath>
#include <complex>
#include <cstdio>
#include <mutex>
#include <vector>

namespace g3d {

// ------------------------------------------------------------------ runtime
namespace {
thread_local std::string g_err;
hipStream_t g_stream = nullptr;
std::once_flag g_once;
}
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
hipStream_t stream() {
    std::call_once(g_once, [] {
        // default priority (a stream of another priority is never created)
        hip_check(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate");
    });
    return g_stream;
}

// ------------------------------------------------------------------ FFT kernels
namespace {
using namespace fftcore;
constexpr int kThreads = 256;
constexpr int kTile = 2048;          // complex values per workgroup tile (32 KiB of LDS)

// C: cx (FP64) or cxf (FP32); S: its real type
template <class C, class S>
__global__ void __launch_bounds__(kThreads) k_fft_rows(C* data, long nrows, int n, int logn, int rpw, const cx* tw,
                                                       int inv, S scale) {
    extern __shared__ __align__(16) unsigned char smraw[];
    C* sm = reinterpret_cast<C*>(smraw);
    const int sp = padded(n);
    C* buf = sm;
    C* tws = sm + size_t(rpw) * sp;
    const long row0 = long(blockIdx.x) * rpw;
    const long left = nrows - row0;
    const int nr = left < rpw ? int(left) : rpw;
    load_twiddles(tws, tw, n >> 1, inv);
    C* g = data + row0 * n;
    const int tot = nr * n;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e >> logn, i = e & (n - 1);
        buf[r * sp + lpos(bitrev(i, logn))] = g[e];
    }
    __syncthreads();
    lds_fft(buf, n, logn, nr, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e >> logn, i = e & (n - 1);
        const C v = buf[r * sp + lpos(i)];
        g[e] = scale == S(1) ? v : scale * v;
    }
}

// columns c0 .. c0+nc-1 of plane p; LDS holds column c at [c * padded(ny)]
template <class C, class S>
__global__ void __launch_bounds__(kThreads) k_fft_cols(C* data, int nx, int ny, int logny, int nc, const cx* tw,
                                                       int inv, S scale) {
    extern __shared__ __align__(16) unsigned char smraw[];
    C* sm = reinterpret_cast<C*>(smraw);
    const int sp = padded(ny);
    C* buf = sm;
    C* tws = sm + size_t(nc) * sp;
    const int tiles = nx / nc;
    const long p = blockIdx.x / tiles;
    const int c0 = (blockIdx.x % tiles) * nc;
    C* g = data + p * long(nx) * ny + c0;
    load_twiddles(tws, tw, ny >> 1, inv);
    const int tot = nc * ny;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int j = e / nc, c = e - j * nc;
        buf[c * sp + lpos(bitrev(j, logny))] = g[long(j) * nx + c];
    }
    __syncthreads();
    lds_fft(buf, ny, logny, nc, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int j = e / nc, c = e - j * nc;
        const C v = buf[c * sp + lpos(j)];
        g[long(j) * nx + c] = scale == S(1) ? v : scale * v;
    }
}

int ilog2(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }
// a workgroup of 64 ... 256 lanes: one lane per group of eight elements of the tile
int threads_for(int groups) { int t = 64; while (t < groups && t < kThreads) t <<= 1; return t; }
bool pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

cx* twiddles(int n) {
    std::vector<cx> h(n / 2);
    for (int k = 0; k < n / 2; ++k) {
        const std::complex<double> w = std::polar(1.0, -2.0 * M_PI * k / n);
        h[k] = cx{w.real(), w.imag()};
    }
    cx* d = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void**>(&d), h.size() * sizeof(cx)), "hipMalloc");
    hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(cx), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return d;
}
}  // namespace

void fft_plan_init(FftPlan& p, int nx, int ny, bool two_d) {
    if (!pow2(nx) || nx > 1024 || (two_d && (!pow2(ny) || ny > 1024)) || ny < 1)
        throw std::runtime_error("g3d FFT: sizes must be powers of two in 2 ... 1024 (nx " + std::to_string(nx) +
                                 ", ny " + std::to_string(ny) + ")");
    p.nx = nx; p.ny = ny; p.two_d = two_d; p.lx = ilog2(nx); p.ly = two_d ? ilog2(ny) : 0;
    p.twx = twiddles(nx);
    p.twy = two_d ? (ny == nx ? p.twx : twiddles(ny)) : nullptr;
}
void fft_plan_free(FftPlan& p) {
    if (p.twy && p.twy != p.twx) (void)hipFree(p.twy);
    if (p.twx) (void)hipFree(p.twx);
    p.twx = p.twy = nullptr;
}

template <class C, class S>
void fft_rows_t(const FftPlan& p, C* data, long nrows, bool inverse, S scale) {
    if (nrows <= 0) return;
    const int n = p.nx;
    int rpw = kTile / n;
    if (rpw < 1) rpw = 1;
    while (rpw > 1 && (nrows + rpw - 1) / rpw < 512) rpw >>= 1;
    const size_t lds = (size_t(rpw) * (n + n / 8) + n / 2) * sizeof(C);
    const int thr = threads_for(rpw * (n / 8 > 0 ? n / 8 : 1));
    hipLaunchKernelGGL((k_fft_rows<C, S>), dim3(blocks_for(nrows, rpw)), dim3(thr), lds, stream(), data, nrows, n, p.lx,
                       rpw, p.twx, int(inverse), scale);
    launch_check("fft rows");
}
template <class C, class S>
void fft_cols_t(const FftPlan& p, C* data, int count, bool inverse, S scale) {
    if (count <= 0) return;
    static const int forced = [] { const char* e = getenv("G3D_FFT_COLS_NC"); return e ? atoi(e) : 0; }();
    int nc = kTile / p.ny;
    if (nc > p.nx) nc = p.nx;
    while (nc > 4 && long(count) * (p.nx / nc) < 512) nc >>= 1;   // >= 4 columns: 64-byte row segments
    if (nc < 1) nc = 1;
    if (forced > 0 && forced <= p.nx && p.nx % forced == 0) nc = forced;
    const size_t lds = (size_t(nc) * (p.ny + p.ny / 8) + p.ny / 2) * sizeof(C);
    const long blocks = long(count) * (p.nx / nc);
    const int thr = threads_for(nc * (p.ny / 8 > 0 ? p.ny / 8 : 1));
    hipLaunchKernelGGL((k_fft_cols<C, S>), dim3(unsigned(blocks)), dim3(thr), lds, stream(), data, p.nx, p.ny, p.ly, nc,
                       p.twy, int(inverse), scale);
    launch_check("fft columns");
}
void fft_rows(const FftPlan& p, cx* data, long nrows, bool inverse, double scale) {
    fft_rows_t(p, data, nrows, inverse, scale);
}
void fft_cols(const FftPlan& p, cx* data, int count, bool inverse, double scale) {
    fft_cols_t(p, data, count, inverse, scale);
}
void fft_run(const FftPlan& p, cx* data, int count, bool inverse) {
    if (count <= 0) return;
    const long rows = long(count) * p.ny;
    if (!p.two_d) { fft_rows(p, data, rows, inverse, inverse ? 1.0 / p.nx : 1.0); return; }
    fft_rows(p, data, rows, inverse, 1.0);
    fft_cols(p, data, count, inverse, inverse ? 1.0 / (double(p.nx) * p.ny) : 1.0);
}
// the same transforms in FP32 (the same butterflies and twiddles, rounded to float)
void fft_run_f(const FftPlan& p, cxf* data, int count, bool inverse) {
    if (count <= 0) return;
    const long rows = long(count) * p.ny;
    if (!p.two_d) { fft_rows_t(p, data, rows, inverse, inverse ? float(1.0 / p.nx) : 1.0f); return; }
    fft_rows_t(p, data, rows, inverse, 1.0f);
    fft_cols_t(p, data, count, inverse, inverse ? float(1.0 / (double(p.nx) * p.ny)) : 1.0f);
}

}  // namespace g3d

// ------------------------------------------------------------------ C ABI: runtime, memory, FFT

extern "C" {
int g3d_init(int device) {
    return g3d::guard([&] {
        g3d::hip_check(hipSetDevice(device), "hipSetDevice");
        (void)g3d::stream();
    });
}
const char* g3d_last_error(void) { return g3d::last_error(); }
const char* g3d_status_string(int code) {
    switch (code) {
    case 0: return "ok";
    case 1: return "error (see g3d_last_error)";
    case 2: return "size mismatch";
    case 3: return "unknown name";
    default: return "unknown status";
    }
}
int g3d_sync(void) { return g3d::guard([] { g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize"); }); }
int g3d_malloc(void** p, size_t bytes) {
    return g3d::guard([&] {
        *p = nullptr;
        if (!bytes) return;
        g3d::hip_check(hipMalloc(p, bytes), "hipMalloc");
        g3d::hip_check(hipMemsetAsync(*p, 0, bytes, g3d::stream()), "hipMemset");
    });
}
int g3d_free(void* p) {
    return g3d::guard([&] {
        if (!p) return;
        g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize");
        g3d::hip_check(hipFree(p), "hipFree");
    });
}
int g3d_memset0(void* p, size_t bytes) {
    return g3d::guard([&] { if (bytes) g3d::hip_check(hipMemsetAsync(p, 0, bytes, g3d::stream()), "hipMemset"); });
}
static int copy(void* d, const void* s, size_t b, hipMemcpyKind k, const char* what) {
    return g3d::guard([&] {
        if (!b) return;
        g3d::hip_check(hipMemcpyAsync(d, s, b, k, g3d::stream()), what);
        g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize");
    });
}
int g3d_memcpy_h2d(void* d, const void* s, size_t b) { return copy(d, s, b, hipMemcpyHostToDevice, "hipMemcpy H2D"); }
int g3d_memcpy_d2h(void* d, const void* s, size_t b) { return copy(d, s, b, hipMemcpyDeviceToHost, "hipMemcpy D2H"); }
int g3d_memcpy_d2d(void* d, const void* s, size_t b) { return copy(d, s, b, hipMemcpyDeviceToDevice, "hipMemcpy D2D"); }

int g3d_fft_plan(int nx, int ny, int two_d, g3d_fft** plan) {
    return g3d::guard([&] {
        auto* f = new g3d_fft;
        try { g3d::fft_plan_init(f->p, nx, ny, two_d != 0); }
        catch (...) { delete f; throw; }
        *plan = f;
    });
}
int g3d_fft_exec(g3d_fft* plan, g3d_cpx* data, int count, int inverse) {
    return g3d::guard([&] { g3d::fft_run(plan->p, data, count, inverse != 0); });
}
int g3d_fft_destroy(g3d_fft* plan) {
    return g3d::guard([&] { if (plan) { g3d::fft_plan_free(plan->p); delete plan; } });
}
}
