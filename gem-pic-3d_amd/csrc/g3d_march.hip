// March kernels: every mode from the front slice to the rear (one lane per mode, slice n+1 carried in registers,
// the explicit inputs of the next slice prefetched while the current one is computed), and one slice.
#include <cstdlib>
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
    int ixx, j;
    mode_xy(c, i, ixx, j);
    // the wavenumbers once and the slice weights with the slice's data, ahead of it: a load issued inside the
    // chain would wait, in the in-order memory counter, for the stores of the slice before it
    const double kxv = c.kx[ixx], kyv = c.ky[j];
    Mode6 up{};
    if (hi < c.N) up = load6(c, long(hi - c.a) * c.nk + i);
    // the explicit part and the currents of the next slice are loaded while this slice is marched (the march of
    // a mode is a chain over the slices: each load on it would add its latency to every slice)
    long o = long(hi - 1 - c.a) * c.nk + i;
    Mode6 old{c.Rx[o], c.Ry[o], c.Ez[o], c.Bz[o], {}, {}};
    Cur9 cur = load_cur(c, i, hi - 1);
    SliceW w = slice_w(c, hi - 1);
    for (int n = hi - 1; n >= lo; --n) {
        Mode6 nxt{};
        Cur9 curn{};
        SliceW wn{};
        const long on = o - c.nk;
        if (n > lo) {
            nxt.Rx = c.Rx[on]; nxt.Ry = c.Ry[on]; nxt.Ez = c.Ez[on]; nxt.Bz = c.Bz[on];
            curn = load_cur(c, i, n - 1);
            wn = slice_w(c, n - 1);
        }
        up = march_mode_k(c, kxv, kyv, w, n, old, up, cur);
        store6(c, o, up);
        old = nxt;
        cur = curn;
        w = wn;
        o = on;
    }
}

// The same march with its inputs staged in LDS, for grids of few modes, where the chain of one wave per 64 modes
// is the whole kernel: wave 0 of a workgroup marches 64 modes slice by slice, reading only LDS and storing the
// results, while waves 1-3 load the next block of kStageB slices into the other of two LDS buffers (so the loads
// are issued a block ahead and never wait behind the marching wave's stores).  Per slice and mode the block holds
// R_x, R_y, E_z, B_z and j0_x, j0_y, jk_x, jk_y of plane n and jk_z of plane n + 1 (the currents of plane n + 1
// that slice n also reads are the previous slice's, kept in registers), and per slice wlo, whi and the step h.
constexpr int kStageB = 3;
constexpr int kStageComp = 9;
constexpr int kStageT = 256;
constexpr size_t kStageLds = size_t(2) * kStageB * kStageComp * 64 * sizeof(cx) + size_t(2) * kStageB * 3 * 8;
static_assert(kStageLds <= 65536, "march staging: one workgroup's LDS");
constexpr long kStageMaxModes = 16384;

template <int COMP>
__device__ __forceinline__ cx stage_src(const MarchView& c, long o) {
    if constexpr (COMP == 0) return c.Rx[o];
    else if constexpr (COMP == 1) return c.Ry[o];
    else if constexpr (COMP == 2) return c.Ez[o];
    else if constexpr (COMP == 3) return c.Bz[o];
    else if constexpr (COMP == 4) return c.j0x[o];
    else if constexpr (COMP == 5) return c.j0y[o];
    else if constexpr (COMP == 6) return c.jkx[o];
    else if constexpr (COMP == 7) return c.jky[o];
    else return c.jkz[o + c.nk];
}

// what producer wave 1 + W stages of block b: components 3 W ... 3 W + 2 of every slice (slices below lo and modes
// beyond nk load a valid element that is never read) and, in wave 1, the weights of the block's slices
struct StageRegs {
    cx v[kStageB][3];
    double w[3];
};
template <int W>
__device__ __forceinline__ void stage_load(const MarchView& c, int b, int lo, int hi, long ic, int lane,
                                           StageRegs& r) {
#pragma unroll
    for (int s = 0; s < kStageB; ++s) {
        const int n = max(hi - 1 - b * kStageB - s, lo);
        const long o = long(n - c.a) * c.nk + ic;
        r.v[s][0] = stage_src<3 * W>(c, o);
        r.v[s][1] = stage_src<3 * W + 1>(c, o);
        r.v[s][2] = stage_src<3 * W + 2>(c, o);
    }
    if (W == 0) {
        const int n = max(hi - 1 - b * kStageB - min(lane, kStageB - 1), lo);
        const SliceW w = slice_w(c, n);
        r.w[0] = w.wl; r.w[1] = w.wu; r.w[2] = w.h;
    }
}
template <int W>
__device__ __forceinline__ void stage_write(cx* stg, double* wts, int b, int lane, const StageRegs& r) {
    const int buf = b & 1;
#pragma unroll
    for (int s = 0; s < kStageB; ++s)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) stg[((buf * kStageB + s) * kStageComp + 3 * W + cc) * 64 + lane] = r.v[s][cc];
    if (W == 0 && lane < kStageB) {
        double* d = wts + (buf * kStageB + lane) * 3;
        d[0] = r.w[0]; d[1] = r.w[1]; d[2] = r.w[2];
    }
}
// the producer loop of wave 1 + W: block b + 1 is written in the interval of block b from registers loaded in the
// interval before, so each block's loads have two march intervals to arrive
template <int W>
__device__ __forceinline__ void producer(const MarchView& c, cx* stg, double* wts, int nblk, int lo, int hi, long ic,
                                         int lane) {
    StageRegs r;
    stage_load<W>(c, 0, lo, hi, ic, lane, r);
    stage_write<W>(stg, wts, 0, lane, r);
    if (nblk > 1) stage_load<W>(c, 1, lo, hi, ic, lane, r);
    for (int b = 0; b < nblk; ++b) {
        __syncthreads();                 // buffer b & 1 staged; the other one read by the march of block b - 1
        if (b + 1 < nblk) stage_write<W>(stg, wts, b + 1, lane, r);
        if (b + 2 < nblk) stage_load<W>(c, b + 2, lo, hi, ic, lane, r);
    }
}

// slices hi-1 ... lo of every mode, 64 modes per workgroup of kStageT lanes (blocks of kStageB slices: block b
// holds slices hi-1-b kStageB ... down to lo)
__global__ void __launch_bounds__(kStageT) k_march_staged(MarchView c, int lo, int hi) {
    extern __shared__ __align__(16) unsigned char stg_raw[];
    cx* stg = reinterpret_cast<cx*>(stg_raw);
    double* wts = reinterpret_cast<double*>(stg + 2 * kStageB * kStageComp * 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long i = long(blockIdx.x) * 64 + lane;
    const bool on = i < c.nk;
    const long ic = on ? i : c.nk - 1;
    const int nblk = (hi - lo + kStageB - 1) / kStageB;
    if (wave > 0) {
        // the producers' own loops: kept apart from the marching wave's, whose registers must never look like
        // the destinations of loads in flight (a wait on them would also wait for the march's stores)
        if (wave == 1) producer<0>(c, stg, wts, nblk, lo, hi, ic, lane);
        else if (wave == 2) producer<1>(c, stg, wts, nblk, lo, hi, ic, lane);
        else producer<2>(c, stg, wts, nblk, lo, hi, ic, lane);
        return;
    }
    Mode6 up{};
    cx u0x{}, u0y{}, ukx{}, uky{};       // j0_x, j0_y, jk_x, jk_y of plane n + 1
    double kxv = 0, kyv = 0;
    if (on) {
        int ixx, j;
        mode_xy(c, i, ixx, j);
        kxv = c.kx[ixx]; kyv = c.ky[j];
        if (hi < c.N) up = load6(c, long(hi - c.a) * c.nk + i);
        const long ou = long(hi - c.a) * c.nk + i;
        u0x = c.j0x[ou]; u0y = c.j0y[ou]; ukx = c.jkx[ou]; uky = c.jky[ou];
    }
    __builtin_amdgcn_s_waitcnt(0);       // these loads complete here, before the chain
    for (int b = 0; b < nblk; ++b) {
        __syncthreads();
        const int buf = b & 1;
        for (int s = 0; s < kStageB; ++s) {
            const int n = hi - 1 - b * kStageB - s;
            if (n < lo) break;
            const cx* q = stg + (buf * kStageB + s) * kStageComp * 64 + lane;
            const double* d = wts + (buf * kStageB + s) * 3;
            const SliceW w{d[0], d[1], d[2]};
            const Mode6 old{q[0], q[64], q[128], q[192], {}, {}};
            const cx dx = q[256], dy = q[320], dkx = q[384], dky = q[448], ukz = q[512];
            const Cur9 J{dx, u0x, dy, u0y, dkx, ukx, dky, uky, ukz};
            if (on) {
                up = march_mode_k(c, kxv, kyv, w, n, old, up, J);
                store6(c, long(n - c.a) * c.nk + i, up);
            }
            u0x = dx; u0y = dy; ukx = dkx; uky = dky;
        }
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
    // the staged march up to one workgroup per compute unit (128 x 128 modes); G3D_MARCH_STAGED=0 turns it off
    const char* e = getenv("G3D_MARCH_STAGED");
    if ((e ? atoi(e) : 1) && v.nk <= kStageMaxModes)
        hipLaunchKernelGGL(k_march_staged, dim3(blocks_for(v.nk, 64)), dim3(kStageT), kStageLds, stream(), v, lo,
                           hi);
    else
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
