// March kernels: every mode from the front slice to the rear (one lane per mode, slice n+1 carried in registers,
// the explicit inputs of the next slice prefetched while the current one is computed), and one slice.
#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {

__device__ __forceinline__ Mode6 load6(const MarchView& c, long o) {
    return Mode6{c.Rx[o], c.Ry[o], c.Ez[o], c.Bz[o], c.Lx[o], c.Ly[o]};
}
__device__ __forceinline__ void store6(const MarchView& c, long o, const Mode6& m) {
    c.Rx[o] = m.Rx; c.Ry[o] = m.Ry; c.Ez[o] = m.Ez; c.Bz[o] = m.Bz; c.Lx[o] = m.Lx; c.Ly[o] = m.Ly;
}

// workgroups of one wave for the march of every mode, so that a small grid (64 x 64 modes: 64 waves) spreads
// over as many compute units
constexpr int kMarchT = 64;
// slices hi-1 ... lo of every mode (src/sim_fields.cpp:7-31, Fourier y), one lane per mode
__global__ void __launch_bounds__(kMarchT) k_march_all(MarchView c, int lo, int hi) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.nk) return;
    Mode6 up{};
    if (hi < c.N) up = load6(c, long(hi - c.a) * c.nk + i);
    // the explicit part and the currents of the next slice are loaded while this slice is marched (the march of
    // a mode is a chain over the slices: each load on it would add its latency to every slice)
    long o = long(hi - 1 - c.a) * c.nk + i;
    Mode6 old{c.Rx[o], c.Ry[o], c.Ez[o], c.Bz[o], {}, {}};
    Cur9 cur = load_cur(c, i, hi - 1);
    for (int n = hi - 1; n >= lo; --n) {
        Mode6 nxt{};
        Cur9 curn{};
        const long on = o - c.nk;
        if (n > lo) {
            nxt.Rx = c.Rx[on]; nxt.Ry = c.Ry[on]; nxt.Ez = c.Ez[on]; nxt.Bz = c.Bz[on];
            curn = load_cur(c, i, n - 1);
        }
        up = march_mode(c, i, n, old, up, cur);
        store6(c, o, up);
        old = nxt;
        cur = curn;
        o = on;
    }
}

// one slice n; bk* optional: restore = 0 saves slice n before the update, restore = 1 reads it from there
__global__ void __launch_bounds__(256) k_march_slice(MarchView c, int n, cx* bRx, cx* bRy, cx* bEz, cx* bBz,
                                                     int restore) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.nk) return;
    const long o = long(n - c.a) * c.nk + i;
    Mode6 old;
    if (restore && bRx) { old.Rx = bRx[i]; old.Ry = bRy[i]; old.Ez = bEz[i]; old.Bz = bBz[i]; }
    else {
        old.Rx = c.Rx[o]; old.Ry = c.Ry[o]; old.Ez = c.Ez[o]; old.Bz = c.Bz[o];
        if (bRx) { bRx[i] = old.Rx; bRy[i] = old.Ry; bEz[i] = old.Ez; bBz[i] = old.Bz; }
    }
    Mode6 up{};
    if (n < c.N - 1) up = load6(c, o + c.nk);
    store6(c, o, march_mode(c, i, n, old, up));
}

MarchView view_of(const g3d_march_ctx& m) {
    MarchView v;
    v.Rx = m.Rx; v.Ry = m.Ry; v.Ez = m.Ez; v.Bz = m.Bz; v.Lx = m.Lx; v.Ly = m.Ly;
    v.j0x = m.j0x; v.j0y = m.j0y; v.jkx = m.jkx; v.jky = m.jky; v.jkz = m.jkz;
    v.kx = m.kx; v.ky = m.ky; v.hcg = m.hcg; v.wlo = m.wlo; v.whi = m.whi;
    v.D = m.D; v.a = m.a; v.N = m.N; v.nx = m.nx; v.nk = m.nk; v.ny = m.ny;
    return v;
}
void require_one_gpu(const g3d_march_ctx& m) {
    if (m.a != 0 || m.b != m.N) throw std::runtime_error("g3d march: one GPU owns all slices (a = 0, b = N)");
    if (!m.ky && (!m.scp || !m.iD || !m.iA || m.ny < 4))
        throw std::runtime_error("g3d march: the free-hand y march needs iD, iA, scp (2 nk doubles) and ny >= 4");
}
LineView lines_of(const g3d_march_ctx& m) {
    LineView L;
    L.mv = view_of(m); L.iD = m.iD; L.iA = m.iA; L.scp = m.scp; L.ny = m.ny; L.piv = nullptr;
    return L;
}
}  // namespace

void march_all(const MarchView& v, int lo, int hi) {
    if (hi <= lo) return;
    hipLaunchKernelGGL(k_march_all, dim3(blocks_for(v.nk, kMarchT)), dim3(kMarchT), 0, stream(), v, lo, hi);
    launch_check("march");
}
void march_one(const MarchView& v, int n, cx* bRx, cx* bRy, cx* bEz, cx* bBz, int restore) {
    hipLaunchKernelGGL(k_march_slice, dim3(blocks_for(v.nk, 256)), dim3(256), 0, stream(), v, n, bRx, bRy, bEz, bBz,
                       restore);
    launch_check("march slice");
}

}  // namespace g3d

extern "C" {
int g3d_march(const g3d_march_ctx* c, int lo, int hi) {
    return g3d::guard([&] {
        g3d::require_one_gpu(*c);
        if (c->ky) g3d::march_all(g3d::view_of(*c), lo, hi);
        else g3d::march_lines(g3d::lines_of(*c), lo, hi);
    });
}
int g3d_march_slice(const g3d_march_ctx* c, int n, int restore) {
    return g3d::guard([&] {
        g3d::require_one_gpu(*c);
        if (c->ky) {
            g3d::march_one(g3d::view_of(*c), n, c->bkRx, c->bkRy, c->bkEz, c->bkBz, restore);
            return;
        }
        if (c->bkRx) {                     // the backup of slice n, then the lines of slice n
            const long o = long(n - c->a) * c->nk;
            g3d::cx* sl[4] = {c->Rx + o, c->Ry + o, c->Ez + o, c->Bz + o};
            g3d::cx* bk[4] = {c->bkRx, c->bkRy, c->bkEz, c->bkBz};
            for (int k = 0; k < 4; ++k) {
                if (restore) g3d::copy_planes(sl[k], bk[k], c->nk);
                else g3d::copy_planes(bk[k], sl[k], c->nk);
            }
        }
        g3d::march_lines(g3d::lines_of(*c), n, n + 1);
    });
}
}
