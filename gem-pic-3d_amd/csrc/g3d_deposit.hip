// Beam deposit in the reference's summation order (src/sim_particles.cpp:56-107).
//
// The reference adds every particle's TSC contributions into jw one particle after the other; with FP64 addition
// the order fixes the last bits, and the start field amplifies them (the relaxation's long step D = 100 dz and the
// near-singular co-moving modes turn 1e-16 into ~1e-12).  Global atomics would sum in schedule order.  Here the
// deposit is a sort instead:
//   1. one lane per particle writes its contribution records: record r = ip R + slot (R = 9 on the Fourier y grid,
//      18 on the free-hand grid: 9 node cells with j_x, j_z and 9 half-node cells with j_y), the cell key
//      (group, half-cell, point) and the values; cells a particle does not touch get the key ~0;
//   2. a stable radix sort of (key, r) groups each cell's records in particle order;
//   3. the values are copied into that order (one lane per record), so that
//   4. one lane per cell (the head of its run of keys) finds the end of its run by bisection; a short run it sums
//      itself from contiguous memory in that order, from zero, as the reference does, and stores the sums into jw;
//   5. a long run (a cell at the core of a dense bunch collects thousands of records) goes to a list, and one wave
//      per long run sums it: the wave loads 64 records at a time with coalesced loads into LDS (the next chunk's
//      loads in flight meanwhile) and one lane adds them in record order (the additions are the same sequence; a
//      lane alone on a long run waited for its loads every few records).
// The result is the reference's single-thread jw, bit for bit.
#include <rocprim/device/device_radix_sort.hpp>

#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 256;
constexpr unsigned long long kNone = ~0ull;
// key of a record: cell = hw nk + point in the low cb bits (cb: the bits of nh nk), the half-node group in bit cb;
// the sort looks at bits 0 .. cb+1 only, where the empty key ~0 has all ones and so follows every cell

__global__ void __launch_bounds__(kT) k_dep_emit(DepositArgs c, Parts P, long np, int R, int cb,
                                                 unsigned long long* key, unsigned* rec, double* val) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    const long r0 = ip * R;
    for (int s = 0; s < R; ++s) { key[r0 + s] = kNone; rec[r0 + s] = unsigned(r0 + s); }
    if (P.alive[ip] != 1) return;
    int m; double fz;
    c.zax.locate(P.zeta[ip], m, fz);
    const int hw = m - (c.a - 2);
    if (hw < 0 || hw >= c.nh || m < 0 || m > c.N - 2) return;
    const double px = P.px[ip], py = P.py[ip], pz = P.pz[ip];
    const double g = sqrt(1 + px * px + py * py + pz * pz);
    const double vx = px / g, vy = py / g, vz = pz / g;
    const double f = P.q[ip] * P.w[ip] / (c.hcg[m + 2] * c.dx);
    int ix[3], jn[3], jh[3]; double wx[3], wn[3], wh[3];
    tsc_periodic(P.x[ip], c.dx, c.nx, ix, wx);
    const unsigned long long base = (unsigned long long)hw * (unsigned long long)c.nk;
    double* V = val + 3 * r0;
    if (c.yf) {
        tsc_periodic(P.y[ip], c.dy, c.ny, jn, wn);
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) {
                const int s = 3 * v + u;
                const double w = f * wx[u] * wn[v] * c.ia[0];
                key[r0 + s] = base + (unsigned long long)(long(jn[v]) * c.nx + ix[u]);
                V[3 * s] = w * vx / vz; V[3 * s + 1] = w * vy / vz; V[3 * s + 2] = w;
            }
        return;
    }
    tsc_freehand(c.yax, P.y[ip], jn, wn, jh, wh);
    for (int v = 0; v < 3; ++v) {
        const bool nod = jn[v] >= 1 && jn[v] <= c.ny - 1, hal = jh[v] >= 0 && jh[v] <= c.ny - 1;
        for (int u = 0; u < 3; ++u) {
            const int s = 3 * v + u;
            if (nod) {
                const double w = f * wx[u] * wn[v] * c.ia[jn[v]];
                key[r0 + s] = base + (unsigned long long)(long(jn[v]) * c.nx + ix[u]);
                V[3 * s] = w * vx / vz; V[3 * s + 1] = 0; V[3 * s + 2] = w;
            }
            if (hal) {
                key[r0 + 9 + s] = (1ull << cb) | (base + (unsigned long long)(long(jh[v]) * c.nx + ix[u]));
                V[3 * (9 + s)] = 0; V[3 * (9 + s) + 2] = 0;
                V[3 * (9 + s) + 1] = f * wx[u] * wh[v] * c.id[jh[v]] * vy / vz;
            }
        }
    }
}

// the values of the sorted records in sorted order
__global__ void __launch_bounds__(kT) k_dep_order(const unsigned long long* key, const unsigned* rec, const double* val,
                                                  long n, double* vs) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || key[t] == kNone) return;
    const double* v = val + 3 * long(rec[t]);
    vs[3 * t] = v[0]; vs[3 * t + 1] = v[1]; vs[3 * t + 2] = v[2];
}

constexpr int kShortRun = 32;                      // longer runs are summed by a wave each (k_dep_sum_long)

// the sums of one cell into jw
__device__ __forceinline__ void dep_store(unsigned long long k, int cb, double s0, double s1, double s2, double* J,
                                          long nh, long nk, int yf) {
    const bool half = (k >> cb) & 1;
    const long cell = long(k & ((1ull << cb) - 1));
    const long hw = cell / nk, p = cell - hw * nk;
    if (half) { J[(1 * nh + hw) * nk + p] = s1; return; }
    J[(0 * nh + hw) * nk + p] = s0;
    if (yf) J[(1 * nh + hw) * nk + p] = s1;
    J[(2 * nh + hw) * nk + p] = s2;
}

// one lane per sorted record; the head of each run of equal keys sums a short run in record (= particle) order and
// lists a long one
__global__ void __launch_bounds__(kT) k_dep_sum(const unsigned long long* key, const double* vs, long n, double* J,
                                                long nh, long nk, int yf, int cb, long* runs, int* nruns, int cap) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = key[i];
    if (k == kNone || (i > 0 && key[i - 1] == k)) return;
    long lo = i + 1, hi = n;                        // the run ends at the first record with another key
    while (lo < hi) {
        const long mid = lo + ((hi - lo) >> 1);
        if (key[mid] == k) lo = mid + 1; else hi = mid;
    }
    if (lo - i > kShortRun) {
        const int r = atomicAdd(nruns, 1);
        if (r < cap) { runs[2 * r] = i; runs[2 * r + 1] = lo; return; }
    }
    // in blocks of kU records, the next block's loads in flight during this block's additions (the additions
    // themselves stay in record order)
    constexpr int kU = 8;
    double s0 = 0, s1 = 0, s2 = 0;
    long t = i;
    double a[3 * kU], b[3 * kU];
    if (t + kU <= lo)
        for (int u = 0; u < 3 * kU; ++u) a[u] = vs[3 * t + u];
    while (t + kU <= lo) {
        const long tn = t + kU;
        const bool more = tn + kU <= lo;
        if (more)
            for (int u = 0; u < 3 * kU; ++u) b[u] = vs[3 * tn + u];
        for (int u = 0; u < kU; ++u) { s0 += a[3 * u]; s1 += a[3 * u + 1]; s2 += a[3 * u + 2]; }
        t = tn;
        if (more)
            for (int u = 0; u < 3 * kU; ++u) a[u] = b[u];
    }
    for (; t < lo; ++t) { s0 += vs[3 * t]; s1 += vs[3 * t + 1]; s2 += vs[3 * t + 2]; }
    dep_store(k, cb, s0, s1, s2, J, nh, nk, yf);
}

// one wave per listed long run (persistent: wave w takes runs w, w + waves, ...; every wave leaves when the list is
// done): chunks of 64 records through LDS, lane 0 adds them in record order from zero
constexpr int kRunWaves = 4;                        // waves per workgroup
__global__ void __launch_bounds__(64 * kRunWaves) k_dep_sum_long(const unsigned long long* key, const double* vs,
                                                                 const long* runs, const int* nruns, int cap,
                                                                 double* J, long nh, long nk, int yf, int cb) {
    __shared__ double buf[kRunWaves][2][3 * 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long waves = long(gridDim.x) * kRunWaves;
    const int nr = min(*nruns, cap);
    for (long r = long(blockIdx.x) * kRunWaves + wv; r < nr; r += waves) {
        const long i = runs[2 * r], lo = runs[2 * r + 1];
        double s0 = 0, s1 = 0, s2 = 0;
        int cur = 0;
        // the first chunk into LDS
        for (int c = lane; c < 3 * 64; c += 64) {
            const long t = i + c / 3;
            buf[wv][0][c] = t < lo ? vs[3 * i + c] : 0.0;
        }
        __builtin_amdgcn_wave_barrier();
        for (long c0 = i; c0 < lo; c0 += 64) {
            // the next chunk's loads in flight while lane 0 adds this one
            double nx[3];
            const long n0 = c0 + 64;
            for (int q = 0; q < 3; ++q) {
                const int c = lane + 64 * q;
                nx[q] = n0 + c / 3 < lo ? vs[3 * n0 + c] : 0.0;
            }
            if (lane == 0) {
                const int m = int(min(long(64), lo - c0));
                const double* b = buf[wv][cur];
                int u = 0;
                // 16 records at a time: their 48 values into registers with one LDS round trip, then the additions
                for (; u + 16 <= m; u += 16) {
                    double w[48];
#pragma unroll
                    for (int q = 0; q < 48; ++q) w[q] = b[3 * u + q];
#pragma unroll
                    for (int q = 0; q < 16; ++q) { s0 += w[3 * q]; s1 += w[3 * q + 1]; s2 += w[3 * q + 2]; }
                }
                for (; u < m; ++u) { s0 += b[3 * u]; s1 += b[3 * u + 1]; s2 += b[3 * u + 2]; }
            }
            for (int q = 0; q < 3; ++q) buf[wv][cur ^ 1][lane + 64 * q] = nx[q];
            cur ^= 1;
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) dep_store(key[i], cb, s0, s1, s2, J, nh, nk, yf);
    }
}
}  // namespace

void bunch_deposit(const DepositArgs& a, Parts p, long np, DepositScratch& s) {
    if (np <= 0) return;
    const int R = a.yf ? 9 : 18;
    const long n = np * R;
    if (n > long(0xffffffffu)) throw std::runtime_error("beam deposit: too many contribution records");
    int cb = 1;                                     // the bits of the cells hw nk + point
    while ((1ull << cb) < (unsigned long long)a.nh * (unsigned long long)a.nk) ++cb;
    if (cb > 60) throw std::runtime_error("beam deposit: grid too large");
    if (s.key0.size() < size_t(n)) {
        const size_t cap = size_t(n) + size_t(n) / 4;
        s.key0.resize(cap); s.key1.resize(cap); s.rec0.resize(cap); s.rec1.resize(cap); s.val.resize(3 * cap);
        s.vs.resize(3 * cap);
    }
    hipLaunchKernelGGL(k_dep_emit, dim3(blocks_for(np, kT)), dim3(kT), 0, stream(), a, p, np, R, cb, s.key0.data(),
                       s.rec0.data(), s.val.data());
    launch_check("beam deposit records");
    // bits 0 .. cb+1: the cell, the group bit, and above every valid key the low bits of the empty key ~0
    const unsigned end_bit = unsigned(cb) + 2;
    size_t bytes = 0;
    hip_check(rocprim::radix_sort_pairs(nullptr, bytes, s.key0.data(), s.key1.data(), s.rec0.data(), s.rec1.data(),
                                        n, 0, end_bit, stream()), "beam deposit sort size");
    if (s.tmp.size() < bytes) s.tmp.resize(bytes + bytes / 4 + 256);
    bytes = s.tmp.size();
    hip_check(rocprim::radix_sort_pairs(s.tmp.data(), bytes, s.key0.data(), s.key1.data(), s.rec0.data(),
                                        s.rec1.data(), n, 0, end_bit, stream()), "beam deposit sort");
    hipLaunchKernelGGL(k_dep_order, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), s.key1.data(), s.rec1.data(),
                       s.val.data(), n, s.vs.data());
    // long runs: at most one per kShortRun + 1 records
    const int cap = int(n / (kShortRun + 1) + 1);
    if (s.runs.size() < size_t(2 * cap)) s.runs.resize(size_t(2 * cap));
    if (s.nruns.size() < 1) s.nruns.resize(1);
    hip_check(hipMemsetAsync(s.nruns.data(), 0, sizeof(int), stream()), "hipMemsetAsync");
    hipLaunchKernelGGL(k_dep_sum, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), s.key1.data(), s.vs.data(), n, a.J,
                       long(a.nh), a.nk, a.yf, cb, s.runs.data(), s.nruns.data(), cap);
    hipLaunchKernelGGL(k_dep_sum_long, dim3(1024), dim3(64 * kRunWaves), 0, stream(), s.key1.data(), s.vs.data(),
                       s.runs.data(), s.nruns.data(), cap, a.J, long(a.nh), a.nk, a.yf, cb);
    launch_check("beam deposit sums");
}

}  // namespace g3d
