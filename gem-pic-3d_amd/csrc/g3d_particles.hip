// Particle kernels: the RK4 push in z of the beam list with its gathers and the analytic laser, the beam deposit,
// the streaming half-cell pass (two Boris rotations, the deposit, trapping, ADK ionization), and the list tools.
// One lane per particle over structure-of-arrays streams.
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 256;

// ---------------------------------------------------------------- streaming particles (src/sim_stream.hpp)
// LDS windows of a tile, WX x WY nodes from node (x0, y0) (periodic); null for untiled particles:
//   lds   the 5 current planes of the deposit (double)
//   fld   the fields of slices fl and fl+1 for the gather (float, [slice][component][WY][WX]), loaded once per
//         workgroup instead of 108 scattered loads per particle
struct Win {
    double* lds;
    const float* fld;
    int x0, y0, WX, WY, fl;
};

// E, B from the tile's field window (Fourier y): the same weights and the same order of the sums as gather(), so
// the values are identical; false when the stencil or the slice pair is not in the window (then the caller
// gathers from F)
__device__ __forceinline__ bool gather_window(const FieldView& c, const Win& w, double x, double y, double zeta,
                                              double E[3], double B[3]) {
    if (!c.self) return false;
    int n; double fz;
    c.zax.locate(zeta, n, fz);
    if (n == c.a - 1 && fz > 1 - 1e-9) { n = c.a; fz = 0; }
    if (n - c.a != w.fl) return false;
    int ix[3], jn[3]; double wx[3], wn[3];
    tsc_periodic(x, c.dx, c.nx, ix, wx);
    tsc_periodic(y, c.dy, c.ny, jn, wn);
    const int u0 = wrap_index(ix[0] - w.x0, c.nx), v0 = wrap_index(jn[0] - w.y0, c.ny);
    if (u0 + 2 >= w.WX || v0 + 2 >= w.WY) return false;
    const double gz = 1 - fz;
    int k[9]; double wt[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) { k[3 * j + i] = (v0 + j) * w.WX + u0 + i; wt[3 * j + i] = wx[i] * wn[j]; }
    const int pl = w.WX * w.WY;
    double v[6];
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        const float* q0 = w.fld + comp * pl;
        const float* q1 = w.fld + (6 + comp) * pl;
        double s = 0;
#pragma unroll
        for (int t = 0; t < 9; ++t) s += wt[t] * (gz * double(q0[k[t]]) + fz * double(q1[k[t]]));
        v[comp] = s;
    }
    E[0] = v[0]; E[1] = v[1]; E[2] = v[2]; B[0] = v[3]; B[1] = v[4]; B[2] = v[5];
    return true;
}

// E, B from the tile's field window on the free-hand y grid (window rows y0 .. y0+WY-1 are node indices; the
// half-nodes share them): the rows beyond the walls skipped and the same weights and order of the sums as
// gather_t<0>, so the values are identical; false when a row or column of the stencil or the slice pair is not in
// the window (then the caller gathers from F)
__device__ __forceinline__ bool gather_window_fh(const FieldView& c, const Win& w, double x, double y, double zeta,
                                                 double E[3], double B[3]) {
    if (!c.self) return false;
    int n; double fz;
    c.zax.locate(zeta, n, fz);
    if (n == c.a - 1 && fz > 1 - 1e-9) { n = c.a; fz = 0; }
    if (n - c.a != w.fl) return false;
    int ix[3], jn[3], jh[3]; double wx[3], wn[3], wh[3];
    tsc_periodic(x, c.dx, c.nx, ix, wx);
    const int u0 = wrap_index(ix[0] - w.x0, c.nx);
    if (u0 + 2 >= w.WX) return false;
    tsc_freehand(c.yax, y, jn, wn, jh, wh);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        if (jn[j] >= 0 && jn[j] < c.ny && (jn[j] < w.y0 || jn[j] >= w.y0 + w.WY)) return false;
        if (jh[j] >= 0 && jh[j] < c.ny && (jh[j] < w.y0 || jh[j] >= w.y0 + w.WY)) return false;
    }
    const double gz = 1 - fz;
    const int pl = w.WX * w.WY;
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        const float* q0 = w.fld + comp * pl;
        const float* q1 = w.fld + (6 + comp) * pl;
        const bool hf = half_comp(comp);
        double v = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int jj = hf ? jh[j] : jn[j];
            const double wy = hf ? wh[j] : wn[j];
            if (jj < 0 || jj >= c.ny) continue;
            const int r = (jj - w.y0) * w.WX + u0;
#pragma unroll
            for (int i = 0; i < 3; ++i) v += wx[i] * wy * (gz * double(q0[r + i]) + fz * double(q1[r + i]));
        }
        if (comp < 3) E[comp] = v; else B[comp - 3] = v;
    }
    return true;
}

// ---------------------------------------------------------------- RK4 in z (src/sim_kernels.hpp:73-149)
struct St { double x, y, zeta, px, py, pz; };

// E, B at a beam particle: from the field window of its workgroup when the stencil and the slice pair lie in it (the
// same values: gather_window sums as gather does), else from F
__device__ __forceinline__ bool rhs(const PushArgs& c, const Win* win, const St& Y, double z, double qm, St& d) {
    const double g = sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
    if (Y.pz < 0.3 * g) return false;
    const double gm = (1 + Y.px * Y.px + Y.py * Y.py) / (g + Y.pz);
    double E[3], B[3], Ee[3], Be[3];
    const double xw = wrap_mod(Y.x, c.fv.Lx), yw = c.fv.yf ? wrap_mod(Y.y, c.fv.Ly) : Y.y;
    if (!(win && gather_window(c.fv, *win, xw, yw, Y.zeta, E, B))) gather(c.fv, xw, yw, Y.zeta, E, B);
    c.laser.field(Y.x + c.X0, z, z - (Y.zeta - c.zc), Ee, Be);
    for (int k = 0; k < 3; ++k) { E[k] += Ee[k]; B[k] += Be[k]; }
    const double ip = 1 / Y.pz;
    d.x = Y.px * ip; d.y = Y.py * ip; d.zeta = -gm * ip;
    d.px = qm * (g * E[0] + Y.py * B[2] - Y.pz * B[1]) * ip;
    d.py = qm * (g * E[1] + Y.pz * B[0] - Y.px * B[2]) * ip;
    d.pz = qm * (g * E[2] + Y.px * B[1] - Y.py * B[0]) * ip;
    return true;
}
__device__ __forceinline__ St axpy(const St& a, double h, const St& b) {
    return St{a.x + h * b.x, a.y + h * b.y, a.zeta + h * b.zeta, a.px + h * b.px, a.py + h * b.py, a.pz + h * b.pz};
}

// beam particle ip from s0 (or from its hand-over point) to s1 (src/sim_kernels.hpp:98-149)
__device__ __forceinline__ void push_beam_one(const PushArgs& c, const Win* win, Parts P, long ip, double z0) {
    const int alive = P.alive[ip];
    if (alive != 1 && alive != 3) return;
    double z = alive == 3 ? c.s1 - P.rem[ip] : z0;
    if (c.freeze && is_bunch(P.sp[ip])) { P.z[ip] = c.s1; return; }
    St Y{P.x[ip], P.y[ip], P.zeta[ip], P.px[ip], P.py[ip], P.pz[ip]};
    const double qm = P.qm[ip];
    auto store = [&](const St& s) {
        P.x[ip] = s.x; P.y[ip] = s.y; P.zeta[ip] = s.zeta; P.px[ip] = s.px; P.py[ip] = s.py; P.pz[ip] = s.pz;
    };
    while (z < c.s1 - 1e-12 * fabs(c.s1)) {
        if (Y.zeta < c.zhand) { store(Y); P.alive[ip] = 2; P.rem[ip] = c.s1 - z; return; }
        if (c.vz_min > 0 && Y.pz < c.vz_min * sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz)) {
            store(Y); P.z[ip] = z; P.alive[ip] = 4;
            return;
        }
        double ds = c.dz_push < c.s1 - z ? c.dz_push : c.s1 - z;
        {
            const double g = sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
            const double lz = 0.5 * c.hmin * Y.pz * (g + Y.pz) / (1 + Y.px * Y.px + Y.py * Y.py);
            if (lz < ds) ds = (c.s1 - z) / ceil((c.s1 - z) / lz);
        }
        if (c.laser.on) {
            const double g = sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
            const double hl = g - Y.px * c.laser.kx - Y.pz * c.laser.kz;
            const double lim = c.push_phase * Y.pz / (c.laser.kL * (hl > 1e-12 ? hl : 1e-12));
            if (lim < ds) ds = (c.s1 - z) / ceil((c.s1 - z) / lim);
        }
        St k1, k2, k3, k4;
        const double hd = 0.5 * ds;
        if (!(rhs(c, win, Y, z, qm, k1) && rhs(c, win, axpy(Y, hd, k1), z + hd, qm, k2) &&
              rhs(c, win, axpy(Y, hd, k2), z + hd, qm, k3) && rhs(c, win, axpy(Y, ds, k3), z + ds, qm, k4))) {
            P.alive[ip] = 0;
            return;
        }
        const double w = ds / 6;
        Y.x += w * (k1.x + 2 * k2.x + 2 * k3.x + k4.x);
        Y.y += w * (k1.y + 2 * k2.y + 2 * k3.y + k4.y);
        Y.zeta += w * (k1.zeta + 2 * k2.zeta + 2 * k3.zeta + k4.zeta);
        Y.px += w * (k1.px + 2 * k2.px + 2 * k3.px + k4.px);
        Y.py += w * (k1.py + 2 * k2.py + 2 * k3.py + k4.py);
        Y.pz += w * (k1.pz + 2 * k2.pz + 2 * k3.pz + k4.pz);
        z += ds;
    }
    store(Y);
    P.z[ip] = c.s1;
    int al = (Y.zeta < 0 || Y.zeta >= c.zmax || Y.y < c.ylo || Y.y > c.yhi) ? 0 : 1;
    if (al == 1 && Y.zeta < c.zhand) { al = 2; P.rem[ip] = 0; }
    P.alive[ip] = al;
}

__global__ void __launch_bounds__(kT) k_bunch_push(PushArgs c, Parts P, long np, double z0) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    push_beam_one(c, nullptr, P, ip, z0);
}

// ---- the beam push with field windows (Fourier y): the particles binned by (slice pair, transverse tile of kBT x kBT
// cells) through an index list (the list itself keeps its order), a workgroup per chunk of at most kBChunk
// particles of a bin with the bin's window of both slices in LDS (a chain of four dependent gathers per RK4 sub-step
// then waits on LDS instead of memory)
constexpr int kBT = 8, kBM = 1, kBW = kBT + 3 + 2 * kBM, kBChunk = 1024;
constexpr unsigned long long kBNone = ~0ull;

// the bin of every particle that moves (alive 1 or 3, not frozen); the bin nbin (no window) for one outside the
// slices of F; ~0 for the others (a frozen bunch particle only takes the new z here, as the plain push does)
__global__ void __launch_bounds__(kT) k_bpush_keys(PushArgs c, Parts P, long np, int ntx, int nty, long nbin,
                                                   unsigned long long* key, unsigned* idx) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= np) return;
    idx[ip] = unsigned(ip);
    key[ip] = kBNone;
    const int alive = P.alive[ip];
    if (alive != 1 && alive != 3) return;
    if (c.freeze && is_bunch(P.sp[ip])) { P.z[ip] = c.s1; return; }
    const FieldView& F = c.fv;
    int n; double fz;
    F.zax.locate(P.zeta[ip], n, fz);
    if (n == F.a - 1 && fz > 1 - 1e-9) { n = F.a; fz = 0; }
    const int fl = n - F.a;
    if (fl < 0 || fl >= F.ns + 1) { key[ip] = (unsigned long long)nbin; return; }
    int ix = int(floor(wrap_mod(P.x[ip], F.Lx) / F.dx)), iy = int(floor(wrap_mod(P.y[ip], F.Ly) / F.dy));
    ix = ix < 0 ? 0 : (ix >= F.nx ? F.nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= F.ny ? F.ny - 1 : iy);
    key[ip] = (unsigned long long)((long(fl) * nty + iy / kBT) * ntx + ix / kBT);
}
// the chunks of the bins: the head of each run of equal keys lists its run in chunks of at most kBChunk
__global__ void __launch_bounds__(kT) k_bpush_chunks(const unsigned long long* key, long n, long* chunk, int* nchunk,
                                                     int cap) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = key[i];
    if (k == kBNone || (i > 0 && key[i - 1] == k)) return;
    long lo = i + 1, hi = n;
    while (lo < hi) {
        const long mid = lo + ((hi - lo) >> 1);
        if (key[mid] == k) lo = mid + 1; else hi = mid;
    }
    const int m = int((lo - i + kBChunk - 1) / kBChunk);
    const int r = atomicAdd(nchunk, m);
    for (int q = 0; q < m && r + q < cap; ++q) {
        chunk[2 * (r + q)] = i + long(q) * kBChunk;
        chunk[2 * (r + q) + 1] = min(lo, i + long(q + 1) * kBChunk);
    }
}
// persistent: workgroup g takes the chunks g, g + grid, ...; every workgroup leaves when the list is done
__global__ void __launch_bounds__(kT) k_bunch_push_win(PushArgs c, Parts P, const unsigned long long* key,
                                                       const unsigned* idx, const long* chunk, const int* nchunk,
                                                       int cap, int ntx, int nty, long nbin, double z0) {
    __shared__ float win[12 * kBW * kBW];
    const FieldView& F = c.fv;
    const int nc = min(*nchunk, cap);
    for (int q = blockIdx.x; q < nc; q += gridDim.x) {
        const long lo = chunk[2 * q], hi = chunk[2 * q + 1];
        const unsigned long long k = key[lo];
        if (k == (unsigned long long)nbin) {         // outside the slices of F: no window
            for (long t = lo + threadIdx.x; t < hi; t += blockDim.x) push_beam_one(c, nullptr, P, long(idx[t]), z0);
            continue;
        }
        const int tx = int(k % (unsigned long long)ntx), ty = int((k / ntx) % (unsigned long long)nty);
        const int fl = int(k / ((unsigned long long)ntx * nty));
        const int x0 = tx * kBT - 1 - kBM, y0 = ty * kBT - 1 - kBM;
        constexpr int pl = kBW * kBW;
        const long cs = long(F.ns + 2) * F.nk;
        __syncthreads();                             // the previous chunk's particles are done with the window
        for (int r = threadIdx.x; r < pl; r += blockDim.x) {
            const int v = r / kBW, u = r - v * kBW;
            const float* src = F.F + long(fl) * F.nk + long(wrap_index(y0 + v, F.ny)) * F.nx + wrap_index(x0 + u, F.nx);
#pragma unroll
            for (int comp = 0; comp < 6; ++comp) {
                win[comp * pl + r] = src[comp * cs];
                win[(6 + comp) * pl + r] = src[comp * cs + F.nk];
            }
        }
        __syncthreads();
        const Win w{nullptr, win, x0, y0, kBW, kBW, fl};
        for (long t = lo + threadIdx.x; t < hi; t += blockDim.x) push_beam_one(c, &w, P, long(idx[t]), z0);
    }
}

// TSC deposit of f (dt/Delta) v/(h dx dy) (src/sim_stream.hpp:33-64) into plane comp of the half-cell current:
// the tile's LDS window when the whole 3 x 3 stencil lies inside it, else jsr with global atomics.  Zero
// contributions are skipped (adding +-0 to a sum that starts at +0 never changes it).
// a deposit into the current planes: float atomics in the FP32-mixed form (c.Jf), else FP64
__device__ __forceinline__ void jadd(const StreamArgs& c, long k, double v) {
    if (c.Jf) atomicAdd(&c.Jf[k], float(v));
    else atomicAdd(&c.J[k], v);
}

template <int YF>
__device__ __forceinline__ void sdeposit(const StreamArgs& c, const Win* win, double x, double y, double f, double vx,
                                         double vy, double vz) {
    const FieldView& P = c.fv;
    int ix[3], jn[3], jh[3]; double wx[3], wn[3], wh[3];
    tsc_periodic(x, P.dx, P.nx, ix, wx);
    const long nk = P.nk;
    const int cq = f < 0 ? 3 : 4;
    const bool mv = vx != 0 || vy != 0 || vz != 0;
    if (YF == 1) {
        tsc_periodic(y, P.dy, P.ny, jn, wn);
        double w[9];
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int u = 0; u < 3; ++u) w[3 * v + u] = f * wx[u] * wn[v] * c.ia[0];
        int u0 = 0, v0 = 0;
        bool in = false;
        if (win) {
            u0 = wrap_index(ix[0] - win->x0, P.nx);
            v0 = wrap_index(jn[0] - win->y0, P.ny);
            in = u0 + 2 < win->WX && v0 + 2 < win->WY;
        }
        if (in) {
            const long wpl = long(win->WX) * win->WY;
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    double* L = win->lds + long(v0 + v) * win->WX + (u0 + u);
                    const double ww = w[3 * v + u];
                    if (mv) {
                        atomicAdd(L, ww * vx);
                        atomicAdd(L + wpl, ww * vy);
                        atomicAdd(L + 2 * wpl, ww * vz);
                    }
                    atomicAdd(L + cq * wpl, ww);
                }
            return;
        }
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const long k = long(jn[v]) * P.nx + ix[u];
                const double ww = w[3 * v + u];
                if (mv) {
                    jadd(c, k, ww * vx);
                    jadd(c, nk + k, ww * vy);
                    jadd(c, 2 * nk + k, ww * vz);
                }
                jadd(c, cq * nk + k, ww);
            }
        return;
    }
    tsc_freehand(P.yax, y, jn, wn, jh, wh);
    if (win) {
        // the window's rows are node (and half-node) indices from y0, not wrapped: the walls bound the grid
        const int u0 = wrap_index(ix[0] - win->x0, P.nx), vn = jn[0] - win->y0, vh = jh[0] - win->y0;
        if (u0 + 2 < win->WX && vn >= 0 && vn + 2 < win->WY && vh >= 0 && vh + 2 < win->WY) {
            const long wpl = long(win->WX) * win->WY;
            for (int v = 0; v < 3; ++v) {
                const bool nod = jn[v] >= 1 && jn[v] <= P.ny - 1, hal = jh[v] >= 0 && jh[v] <= P.ny - 1;
                for (int u = 0; u < 3; ++u) {
                    if (nod) {
                        double* L = win->lds + long(vn + v) * win->WX + (u0 + u);
                        const double w = f * wx[u] * wn[v] * c.ia[jn[v]];
                        if (mv) { atomicAdd(L, w * vx); atomicAdd(L + 2 * wpl, w * vz); }
                        atomicAdd(L + cq * wpl, w);
                    }
                    if (hal && mv) {
                        double* L = win->lds + long(vh + v) * win->WX + (u0 + u);
                        atomicAdd(L + wpl, f * wx[u] * wh[v] * c.id[jh[v]] * vy);
                    }
                }
            }
            return;
        }
    }
    for (int v = 0; v < 3; ++v) {
        const bool nod = jn[v] >= 1 && jn[v] <= P.ny - 1, hal = jh[v] >= 0 && jh[v] <= P.ny - 1;
        for (int u = 0; u < 3; ++u) {
            if (nod) {
                const long k = long(jn[v]) * P.nx + ix[u];
                const double w = f * wx[u] * wn[v] * c.ia[jn[v]];
                if (mv) { jadd(c, k, w * vx); jadd(c, 2 * nk + k, w * vz); }
                jadd(c, cq * nk + k, w);
            }
            if (hal && mv) {
                const long k = long(jh[v]) * P.nx + ix[u];
                jadd(c, nk + k, f * wx[u] * wh[v] * c.id[jh[v]] * vy);
            }
        }
    }
}

// E, B at (x, y, zeta) (x, y wrapped as the grid wants): from the tile's field window when it holds the stencil,
// else from F; the values are the same either way
template <int YF>
__device__ __forceinline__ void gather_stream(const FieldView& P, const Win* win, double x, double y, double zeta,
                                              double E[3], double B[3]) {
    bool got = false;
    if (win && win->fld) {
        if constexpr (YF == 1) got = gather_window(P, *win, x, y, zeta, E, B);
        else got = gather_window_fh(P, *win, x, y, zeta, E, B);
    }
    if (!got) gather_t<YF>(P, x, y, zeta, E, B);
}

// |E| of the grid and the analytic laser (src/sim_stream.hpp:67-74)
template <int YF>
__device__ __forceinline__ double field_abs(const StreamArgs& c, const Win* win, double x, double y, double z,
                                            double zeta) {
    double E[3], B[3], Ee[3], Be[3];
    gather_stream<YF>(c.fv, win, wrap_mod(x, c.fv.Lx), YF ? wrap_mod(y, c.fv.Ly) : y, zeta, E, B);
    c.laser.field(x + c.X0, z, z - (zeta - c.zc), Ee, Be);
    for (int k = 0; k < 3; ++k) E[k] += Ee[k];
    return sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
}

struct PState {
    double x, y, zeta, px, py, pz, w, z, q, qm;
    int alive, sp;
    unsigned long long id;
};

// Monte Carlo ADK (src/sim_stream.hpp:80-101): sequential levels with exponential waiting times; every event
// clones the ion as an electron into the born list
__device__ __forceinline__ void ionize(const StreamArgs& c, PState& p, double E1, double E0, double dt) {
    const IonDev I = c.ions[p.sp - SP_ION];
    if (!I.ionizing || (E1 < I.Emin && E0 < I.Emin)) return;
    const uint64_t base = key_particle(c.key, p.id);
    int Z = int(floor(p.q + 0.5));
    const int Z0 = Z;
    unsigned draw = 0;
    double tleft = dt;
    while (Z < I.Zmax) {
        const AdkLevel L = c.adk[I.off + Z];
        const double w = 0.5 * (adk_rate(L, E1) + adk_rate(L, E0));
        if (!(w > 0)) break;
        const double tw = -log(1.0 - uniform01(base + draw++)) / w;
        if (tw > tleft) break;
        tleft -= tw; ++Z;
        const int j = atomicAdd(c.bcnt, 1);
        if (j < c.bcap) {
            const Parts& b = c.born;
            b.x[j] = p.x; b.y[j] = p.y; b.zeta[j] = p.zeta; b.px[j] = p.px; b.py[j] = p.py; b.pz[j] = p.pz;
            b.w[j] = p.w; b.z[j] = p.z; b.q[j] = -1; b.qm[j] = -1; b.rem[j] = 0;
            b.alive[j] = 1; b.sp[j] = SP_FREED + (p.sp - SP_ION); b.id[j] = p.id;
        }
    }
    if (Z != Z0) { p.q = Z; p.qm = Z / I.mass; }
}

// one particle through half-cell n (src/sim_stream.hpp:104-171)
// YF: the grid (1 Fourier y, 0 free-hand y); ION: the list may hold ions; LASER: the deck has the analytic laser
// (without it the laser's field is zero and its evaluation would only cost registers)
template <int YF, bool ION, bool LASER>
__device__ __forceinline__ void stream_one(const StreamArgs& c, const Win* win, PState& p) {
    const int n = c.n;
    const double zn = c.zs[n], zn1 = c.zs[n + 1], h = zn1 - zn;
    if (p.alive == 6) {
        if (p.zeta > zn) p.alive = 1;
        else return;
    }
    if (p.alive != 1) return;
    const double hl = p.zeta >= zn1 - 1e-9 * h ? h : (p.zeta - zn > 0 ? p.zeta - zn : 0.0);
    const FieldView& P = c.fv;
    const bool ion = ION && is_ion(p.sp);
    if (ion && !c.ions[p.sp - SP_ION].mobile) {
        const double f = p.q * p.w * hl / (c.Delta * h * P.dx);
        if (f != 0) sdeposit<YF>(c, win, p.x, p.y, f, 0, 0, 0);
        const double E1 = field_abs<YF>(c, win, p.x, p.y, p.z, p.zeta), E0 = field_abs<YF>(c, win, p.x, p.y, p.z, zn);
        p.zeta = zn;
        ionize(c, p, E1, E0, hl);
        return;
    }
    const double zeta0 = p.zeta;
    const double g0 = sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz);
    const double omv0 = (1 + p.px * p.px + p.py * p.py) / (g0 * (g0 + p.pz));
    const double dt0 = hl / omv0;
    const double xm = p.x + 0.5 * dt0 * p.px / g0, ym = p.y + 0.5 * dt0 * p.py / g0;
    const double zm = p.z + 0.5 * dt0 * p.pz / g0, zetam = p.zeta - 0.5 * hl;
    double E[3], B[3], Ee[3], Be[3];
    const double xg = wrap_mod(xm, P.Lx), yg = YF ? wrap_mod(ym, P.Ly) : ym;
    gather_stream<YF>(P, win, xg, yg, zetam, E, B);
    if (LASER) {
        c.laser.field(xm + c.X0, zm, zm - (zetam - c.zc), Ee, Be);
        for (int k = 0; k < 3; ++k) { E[k] += Ee[k]; B[k] += Be[k]; }
    }
    double ux = p.px, uy = p.py, uz = p.pz;
    boris(ux, uy, uz, E, B, p.qm, dt0);
    double g1 = sqrt(1 + ux * ux + uy * uy + uz * uz);
    double dt = 2 * hl / (omv0 + (1 + ux * ux + uy * uy) / (g1 * (g1 + uz)));
    ux = p.px; uy = p.py; uz = p.pz;
    boris(ux, uy, uz, E, B, p.qm, dt);
    g1 = sqrt(1 + ux * ux + uy * uy + uz * uz);
    double vx = 0.5 * (p.px / g0 + ux / g1), vy = 0.5 * (p.py / g0 + uy / g1), vz = 0.5 * (p.pz / g0 + uz / g1);
    double zeta_end = zn;
    bool trapped = false;
    if (p.z + vz * dt >= c.sTrap) {
        const double vzc = vz > 1e-3 ? vz : 1e-3;
        const double dtp = (c.sTrap - p.z) / vzc > 0 ? (c.sTrap - p.z) / vzc : 0.0;
        ux = p.px; uy = p.py; uz = p.pz;
        boris(ux, uy, uz, E, B, p.qm, dtp);
        g1 = sqrt(1 + ux * ux + uy * uy + uz * uz);
        vx = 0.5 * (p.px / g0 + ux / g1); vy = 0.5 * (p.py / g0 + uy / g1); vz = 0.5 * (p.pz / g0 + uz / g1);
        dt = dtp;
        zeta_end = p.zeta - (1 - vz) * dtp;
        trapped = true;
    }
    const double xo = p.x, yo = p.y;
    p.x += vx * dt; p.y += vy * dt;
    p.z = trapped ? c.s1 : p.z + vz * dt;
    p.px = ux; p.py = uy; p.pz = uz;
    p.zeta = zeta_end;
    const double f = p.q * p.w * dt / (c.Delta * h * P.dx);
    if (f != 0) sdeposit<YF>(c, win, 0.5 * (xo + p.x), 0.5 * (yo + p.y), f, vx, vy, vz);
    if (!YF && (p.y < c.ylo || p.y > c.yhi)) { p.alive = 0; return; }
    if (ion && !trapped && c.ions[p.sp - SP_ION].ionizing)
        ionize(c, p, field_abs<YF>(c, win, xm, ym, zm, zeta0), field_abs<YF>(c, win, xm, ym, zm, zn), dt);
    if (trapped) {
        p.alive = 5;
        const double vze = uz / g1;
        if (vze >= c.vz_beam) sdeposit<YF>(c, win, p.x, p.y, p.q * p.w / (h * P.dx * vze), ux / g1, uy / g1, vze);
    }
}

__device__ __forceinline__ PState load_pstate(const Parts& S, long i) {
    return PState{S.x[i], S.y[i], S.zeta[i], S.px[i], S.py[i], S.pz[i], S.w[i], S.z[i], S.q[i], S.qm[i], S.alive[i],
                  S.sp[i], S.id[i]};
}

// p: the state of list entry i (loaded by the caller); the pushed state goes to D[i]
template <int YF, bool ION, bool LASER>
__device__ __forceinline__ void stream_index(const StreamArgs& c, const Win* win, PState p, Parts D, long i) {
    stream_one<YF, ION, LASER>(c, win, p);
    if (c.skip_state) return;
    D.x[i] = p.x; D.y[i] = p.y; D.zeta[i] = p.zeta; D.px[i] = p.px; D.py[i] = p.py; D.pz[i] = p.pz;
    D.z[i] = p.z; D.q[i] = p.q; D.qm[i] = p.qm; D.alive[i] = p.alive;
}

// workgroups 0 .. ntile-1: the particles of one tile with the LDS window, written to the tile's slab at the end;
// the following workgroups: the untiled particles from off[ntile] on, one per lane, global atomics
template <int YF, bool ION, bool LASER>
__global__ void __launch_bounds__(kStreamBlockMax) k_stream_halfcell(StreamArgs c, Parts S, Parts D, long n,
                                                                     TileGeom tg) {
    extern __shared__ double win_lds[];
    const long b = blockIdx.x;
    if (b < tg.ntile) {
        // free-hand y (every gather reads F): the tiles of one XCD a band of consecutive tile rows, so that
        // neighbouring tiles meet in one L2 (workgroup b runs on XCD b mod 8)
        const long t = (tg.fh && (tg.ntile & 7) == 0) ? (b & 7) * (tg.ntile >> 3) + (b >> 3) : b;
        // the particles binned into this tile (tg.idx: indices into the list; null: the contiguous load ranges).
        // The range, this lane's first particle and its state first: their dependent loads overlap those of the
        // field window.
        const long lo = tg.off[t], hi = tg.off[t + 1];
        long k = lo + threadIdx.x;
        long i = k < hi ? (tg.idx ? long(tg.idx[k]) : k) : 0;
        PState p{};
        if (k < hi) p = load_pstate(S, i);
        const int pl = tg.WX * tg.WY;
        const long wn = 5L * pl;
        const int tx = int(t % tg.ntx), ty = int(t / tg.ntx);
        const int x0 = tx * tg.TX - 1 - tg.M, y0 = ty * tg.TY - 1 - tg.M;
        const FieldView& F = c.fv;
        const int fl = c.n - F.a;
        float* fw = reinterpret_cast<float*>(win_lds + wn);
        for (long r = threadIdx.x; r < wn; r += blockDim.x) win_lds[r] = 0;
        const bool fwin = tg.fwin && F.self;
        // slices fl, fl + 1 of the 6 components: [slice][comp][v][u]; one node per lane, 12 loads.  Free-hand y: the
        // rows beyond the walls are never read (zero)
        if (fwin)
            for (int r = threadIdx.x; r < pl; r += blockDim.x) {
                const int v = r / tg.WX, u = r - v * tg.WX;
                const int gx = wrap_index(x0 + u, F.nx), gy = YF ? wrap_index(y0 + v, F.ny) : y0 + v;
                const bool in = YF || (gy >= 0 && gy < F.ny);
                const float* src = F.F + long(fl) * F.nk + long(in ? gy : 0) * F.nx + gx;
                const long cs = long(F.ns + 2) * F.nk;
#pragma unroll
                for (int comp = 0; comp < 6; ++comp) {
                    fw[comp * pl + r] = in ? src[comp * cs] : 0.0f;
                    fw[(6 + comp) * pl + r] = in ? src[comp * cs + F.nk] : 0.0f;
                }
            }
        __syncthreads();
        Win w{win_lds, fwin ? fw : nullptr, x0, y0, tg.WX, tg.WY, fl};
        // free-hand y: the y locates of this tile's particles search the window's rows first
        StreamArgs cl = c;
        if (YF == 0 && !c.fv.yax.uniform && c.fv.yax.n >= 3) {
            const int nn = c.fv.yax.n;
            const int lo_r = y0 < 0 ? 0 : (y0 > nn - 2 ? nn - 2 : y0);
            const int hi_r = y0 + tg.WY > nn - 1 ? nn - 1 : (y0 + tg.WY < lo_r + 1 ? lo_r + 1 : y0 + tg.WY);
            cl.fv.yax.hlo = lo_r; cl.fv.yax.hhi = hi_r;
        }
        const double zn = c.zs[c.n];
        unsigned cnt = 0;
        while (k < hi) {
            cnt += pushed_through(p.alive, p.zeta, zn);
            stream_index<YF, ION, LASER>(cl, &w, p, D, i);
            k += blockDim.x;
            if (k < hi) {
                i = tg.idx ? long(tg.idx[k]) : k;
                p = load_pstate(S, i);
            }
        }
        __syncthreads();
        if (YF == 1 && !c.Jf) {
            double* sl = tg.slab + t * wn;
            for (long r = threadIdx.x; r < wn; r += blockDim.x) sl[r] = win_lds[r];
        } else                  // into the current planes (free-hand y: rows outside the walls hold nothing)
            for (long r = threadIdx.x; r < wn; r += blockDim.x) {
                const double v = win_lds[r];
                if (v == 0) continue;
                const int cc = int(r / pl), rr = int(r - long(cc) * pl), vv = rr / tg.WX, u = rr - vv * tg.WX;
                const int gy = YF ? wrap_index(y0 + vv, F.ny) : y0 + vv;
                if (gy < 0 || gy >= F.ny) continue;
                jadd(c, cc * F.nk + long(gy) * F.nx + wrap_index(x0 + u, F.nx), v);
            }
        if (c.npush) add_pushes(c.npush, cnt);
        return;
    }
    const long i = tg.n_tiled + (b - tg.ntile) * blockDim.x + threadIdx.x;
    unsigned cnt = 0;
    if (i < n) {
        const PState p = load_pstate(S, i);
        cnt = pushed_through(p.alive, p.zeta, c.zs[c.n]);
        stream_index<YF, ION, LASER>(c, nullptr, p, D, i);
    }
    if (c.npush) add_pushes(c.npush, cnt);
}

// ---------------------------------------------------------------- binning of the streaming list by tile
// The list keeps its order in memory; a tile's workgroup visits the particles whose current cell lies in the tile
// (idx[off[t] .. off[t+1])), so that particles that have moved away from their loading tile still deposit into
// and gather from LDS windows.  Only live particles (alive 1, waiting 6) are binned: the others are final for the
// rest of the step, and their state is copied into the second buffer here so that both buffers agree.
__device__ __forceinline__ int tile_of(const TileGeom& tg, double x, double y, double dx, double dy, int nx, int ny) {
    int ix = int(floor(x / dx)) % nx, iy;
    if (ix < 0) ix += nx;
    if (tg.fh) {                // free-hand y: the cell of the node table, clamped to the grid
        double f;
        tg.fy.locate(y, iy, f);
        iy = iy < 0 ? 0 : (iy >= ny ? ny - 1 : iy);
    } else {
        iy = int(floor(y / dy)) % ny;
        if (iy < 0) iy += ny;
    }
    return (iy / tg.TY) * tg.ntx + ix / tg.TX;
}
__global__ void __launch_bounds__(kT) k_bin_count(Parts S, Parts D, long n, TileGeom tg, double dx, double dy,
                                                  int nx, int ny, int* cnt) {
    extern __shared__ int hist[];
    for (int t = threadIdx.x; t < tg.ntile; t += blockDim.x) hist[t] = 0;
    __syncthreads();
    for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += long(gridDim.x) * blockDim.x) {
        const int al = S.alive[i];
        if (al == 1 || al == 6) atomicAdd(&hist[tile_of(tg, S.x[i], S.y[i], dx, dy, nx, ny)], 1);
        else {
            D.x[i] = S.x[i]; D.y[i] = S.y[i]; D.zeta[i] = S.zeta[i]; D.px[i] = S.px[i]; D.py[i] = S.py[i];
            D.pz[i] = S.pz[i]; D.z[i] = S.z[i]; D.q[i] = S.q[i]; D.qm[i] = S.qm[i]; D.alive[i] = al;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < tg.ntile; t += blockDim.x)
        if (hist[t]) atomicAdd(&cnt[t], hist[t]);
}
// exclusive scan of the tile counts into off[0 .. ntile] and the scatter cursors (one workgroup: a chunk of tiles
// per lane, then a Hillis-Steele scan of the 1024 chunk sums)
__global__ void __launch_bounds__(1024) k_bin_scan(const int* cnt, long* off, int* cur, long ntile) {
    __shared__ long part[2][1024];
    const long per = (ntile + blockDim.x - 1) / blockDim.x;
    const long t0 = threadIdx.x * per, t1 = t0 + per < ntile ? t0 + per : ntile;
    long s = 0;
    for (long t = t0; t < t1; ++t) s += cnt[t];
    int src = 0;
    part[0][threadIdx.x] = s;
    __syncthreads();
    for (unsigned d = 1; d < blockDim.x; d <<= 1) {
        const long v = part[src][threadIdx.x] + (threadIdx.x >= d ? part[src][threadIdx.x - d] : 0);
        part[1 - src][threadIdx.x] = v;
        src = 1 - src;
        __syncthreads();
    }
    long a = part[src][threadIdx.x] - s;             // exclusive prefix of this lane's chunk
    for (long t = t0; t < t1; ++t) { off[t] = a; cur[t] = int(a); a += cnt[t]; }
    if (threadIdx.x == blockDim.x - 1) off[ntile] = part[src][threadIdx.x];
}
// one lane per particle: its rank among the workgroup's particles of the same tile (LDS atomics), one global
// reservation per tile the workgroup touches, then the index into its slot
__global__ void __launch_bounds__(kT) k_bin_scatter(Parts S, long n, TileGeom tg, double dx, double dy, int nx,
                                                    int ny, int* cur, int* idx) {
    extern __shared__ int lcnt[];                    // [ntile] counts, then [ntile] reserved bases
    int* lbase = lcnt + tg.ntile;
    for (int t = threadIdx.x; t < tg.ntile; t += blockDim.x) lcnt[t] = 0;
    __syncthreads();
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    int t = -1, r = 0;
    if (i < n) {
        const int al = S.alive[i];
        if (al == 1 || al == 6) {
            t = tile_of(tg, S.x[i], S.y[i], dx, dy, nx, ny);
            r = atomicAdd(&lcnt[t], 1);
        }
    }
    __syncthreads();
    for (int u = threadIdx.x; u < tg.ntile; u += blockDim.x)
        if (lcnt[u]) lbase[u] = atomicAdd(&cur[u], lcnt[u]);
    __syncthreads();
    if (t >= 0) idx[lbase[t] + r] = int(i);
}
__global__ void k_append_born(Parts b, const int* bcnt, Parts p0, Parts p1, long base, long cap) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long cnt = *bcnt;
    if (i >= cnt || base + i >= cap) return;
    const long d = base + i;
    for (Parts* t : {&p0, &p1}) {
        t->x[d] = b.x[i]; t->y[d] = b.y[i]; t->zeta[d] = b.zeta[i]; t->px[d] = b.px[i]; t->py[d] = b.py[i];
        t->pz[d] = b.pz[i]; t->z[d] = b.z[i]; t->q[d] = b.q[i]; t->qm[d] = b.qm[i]; t->alive[d] = b.alive[i];
    }
    p0.w[d] = b.w[i]; p0.sp[d] = b.sp[i]; p0.id[d] = b.id[i]; p0.rem[d] = 0;
}

// ---------------------------------------------------------------- list tools
__global__ void k_centroid_partial(Parts P, long np, double X0, double Y0, double* part) {
    __shared__ double s[3][kT];
    double a = 0, b = 0, w = 0;
    for (long i = long(blockIdx.x) * kT + threadIdx.x; i < np; i += long(gridDim.x) * kT) {
        if (P.alive[i] != 1 || !is_bunch(P.sp[i])) continue;
        a += P.w[i] * (P.x[i] + X0); b += P.w[i] * (P.y[i] + Y0); w += P.w[i];
    }
    s[0][threadIdx.x] = a; s[1][threadIdx.x] = b; s[2][threadIdx.x] = w;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o)
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = s[k][0];
}
constexpr int kCentroidBlocks = 256;
// the partial sums staged in LDS by the whole workgroup, then added in block order by one lane (the loads off its
// chain of additions; the same sums in the same order)
__global__ void k_centroid_final(const double* part, int nb, double* out) {
    __shared__ double s[3 * kCentroidBlocks];
    for (int k = threadIdx.x; k < 3 * nb; k += blockDim.x) s[k] = part[k];
    __syncthreads();
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0, b = 0, w = 0;
    for (int k = 0; k < nb; ++k) { a += s[3 * k]; b += s[3 * k + 1]; w += s[3 * k + 2]; }
    out[0] = a; out[1] = b; out[2] = w;
}

__device__ void to_aos(const Parts& p, long i, g3d_particle& o) {
    o.x = p.x[i]; o.y = p.y[i]; o.zeta = p.zeta[i]; o.px = p.px[i]; o.py = p.py[i]; o.pz = p.pz[i]; o.w = p.w[i];
    o.rem = p.rem[i]; o.z = p.z[i]; o.q = p.q[i]; o.qm = p.qm[i]; o.alive = p.alive[i]; o.sp = p.sp[i]; o.id = p.id[i];
}
__global__ void k_collect(Parts p, long n, int code, g3d_particle* out, long* idx, int* count, long cap) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || p.alive[i] != code) return;
    const int j = atomicAdd(count, 1);
    if (j >= cap) return;
    to_aos(p, i, out[j]);
    idx[j] = i;
}
__global__ void k_count(Parts p, long n, int code, int* count) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool hit = i < n && p.alive[i] == code;
    const unsigned long long m = __ballot(hit);
    if (m && (threadIdx.x & 63) == 0) atomicAdd(count, int(__popcll(m)));
}
__global__ void k_from_aos(const g3d_particle* in, Parts p, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const g3d_particle o = in[i];
    p.x[i] = o.x; p.y[i] = o.y; p.zeta[i] = o.zeta; p.px[i] = o.px; p.py[i] = o.py; p.pz[i] = o.pz; p.w[i] = o.w;
    p.rem[i] = o.rem; p.z[i] = o.z; p.q[i] = o.q; p.qm[i] = o.qm; p.alive[i] = o.alive; p.sp[i] = o.sp; p.id[i] = o.id;
}
__global__ void k_to_aos(Parts p, g3d_particle* out, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    to_aos(p, i, out[i]);
}
__global__ void k_copy(Parts s, Parts d, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d.x[i] = s.x[i]; d.y[i] = s.y[i]; d.zeta[i] = s.zeta[i]; d.px[i] = s.px[i]; d.py[i] = s.py[i]; d.pz[i] = s.pz[i];
    d.z[i] = s.z[i]; d.q[i] = s.q[i]; d.qm[i] = s.qm[i]; d.alive[i] = s.alive[i];
}
__global__ void k_shift(Parts p, long n, double sx, double sy) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p.x[i] -= sx; p.y[i] -= sy;
}
}  // namespace

void bunch_push(const PushArgs& a, Parts p, long np, double s0) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_bunch_push, dim3(blocks_for(np, kT)), dim3(kT), 0, stream(), a, p, np, s0);
    launch_check("bunch push");
}
void bunch_push_windowed(const PushArgs& a, Parts p, long np, double s0, BpushScratch& s) {
    if (np <= 0) return;
    const FieldView& F = a.fv;
    if (!F.yf || !F.self || F.nx < 2 * kBT || F.ny < 2 * kBT || np > long(0xffffffffu)) {
        bunch_push(a, p, np, s0);
        return;
    }
    const int ntx = (F.nx + kBT - 1) / kBT, nty = (F.ny + kBT - 1) / kBT;
    const long nbin = long(F.ns + 1) * nty * ntx;
    int eb = 1;                                      // bits of the bins and the bin nbin, one more for ~0
    while ((1ll << eb) <= nbin) ++eb;
    ++eb;
    if (s.key0.size() < size_t(np)) {
        const size_t cap = size_t(np) + size_t(np) / 4;
        s.key0.resize(cap); s.key1.resize(cap); s.idx0.resize(cap); s.idx1.resize(cap);
    }
    const int cap = int(np / 1 + 1);                 // chunks: at most one per particle
    if (s.chunk.size() < size_t(2 * cap)) s.chunk.resize(size_t(2 * cap));
    if (s.nchunk.size() < 1) s.nchunk.resize(1);
    hipLaunchKernelGGL(k_bpush_keys, dim3(blocks_for(np, kT)), dim3(kT), 0, stream(), a, p, np, ntx, nty, nbin,
                       s.key0.data(), s.idx0.data());
    size_t bytes = 0;
    hip_check(rocprim::radix_sort_pairs(nullptr, bytes, s.key0.data(), s.key1.data(), s.idx0.data(), s.idx1.data(),
                                        np, 0, unsigned(eb), stream()), "beam push sort size");
    if (s.tmp.size() < bytes) s.tmp.resize(bytes + bytes / 4 + 256);
    bytes = s.tmp.size();
    hip_check(rocprim::radix_sort_pairs(s.tmp.data(), bytes, s.key0.data(), s.key1.data(), s.idx0.data(),
                                        s.idx1.data(), np, 0, unsigned(eb), stream()), "beam push sort");
    hip_check(hipMemsetAsync(s.nchunk.data(), 0, sizeof(int), stream()), "hipMemsetAsync");
    hipLaunchKernelGGL(k_bpush_chunks, dim3(blocks_for(np, kT)), dim3(kT), 0, stream(), s.key1.data(), np,
                       s.chunk.data(), s.nchunk.data(), cap);
    hipLaunchKernelGGL(k_bunch_push_win, dim3(2048), dim3(kT), 0, stream(), a, p, s.key1.data(), s.idx1.data(),
                       s.chunk.data(), s.nchunk.data(), cap, ntx, nty, nbin, s0);
    launch_check("bunch push (windows)");
}
void stream_halfcell(const StreamArgs& a, Parts src, Parts dst, long nstrm, const TileGeom& tg,
                     const LaunchTimer& lt) {
    if (nstrm <= 0) return;
    // tile workgroups of tg.block lanes, then the untiled tail in workgroups of the same size
    const int bs = tg.ntile ? tg.block : kT;
    if (bs < 64 || bs > kStreamBlockMax || bs % 64) throw std::runtime_error("stream half-cell: bad block size");
    const long tail_blocks = nstrm > tg.n_tiled ? long(blocks_for(nstrm - tg.n_tiled, bs)) : 0;
    // the deposit window (5 double planes) and the field window (12 float planes)
    const size_t lds = tg.ntile ? size_t(tg.WX) * tg.WY * (5 * sizeof(double) + (tg.fwin ? 12 * sizeof(float) : 0)) : 0;
    const dim3 grid(unsigned(tg.ntile + tail_blocks)), blk(bs);
    if (a.npush && long(grid.x) > a.npush_cap) throw std::runtime_error("stream half-cell: push counter too small");
    const bool ion = a.ions != nullptr;
    const unsigned sh = unsigned(lds);
    auto* k = a.laser.on
        ? (a.fv.yf ? (ion ? k_stream_halfcell<1, true, true> : k_stream_halfcell<1, false, true>)
                   : (ion ? k_stream_halfcell<0, true, true> : k_stream_halfcell<0, false, true>))
        : (a.fv.yf ? (ion ? k_stream_halfcell<1, true, false> : k_stream_halfcell<1, false, false>)
                   : (ion ? k_stream_halfcell<0, true, false> : k_stream_halfcell<0, false, false>));
    launch_timed_kernel(lt, k, grid, blk, sh, a, src, dst, nstrm, tg);
    launch_check("stream half-cell");
}
// one workgroup: the further chunks of every tile, listed through an LDS counter
__global__ void __launch_bounds__(1024) k_tile_chunks(const long* off, long ntile, int chunk, int cap, int* xtile,
                                                      long* xstart, int* xcount) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (long t = threadIdx.x; t < ntile; t += blockDim.x)
        for (long s = off[t] + chunk; s < off[t + 1]; s += chunk) {
            const int e = atomicAdd(&cnt, 1);
            if (e < cap) { xtile[e] = int(t); xstart[e] = s; }
        }
    __syncthreads();
    if (threadIdx.x == 0) xcount[0] = cnt < cap ? cnt : cap;
}
void tile_chunks(const TileGeom& tg, int* xtile, long* xstart, int* xcount) {
    if (tg.ntile <= 0 || tg.chunk <= 0) return;
    hipLaunchKernelGGL(k_tile_chunks, dim3(1), dim3(1024), 0, stream(), tg.off, tg.ntile, tg.chunk, tg.max_extra,
                       xtile, xstart, xcount);
    launch_check("tile chunks");
}
void bin_tiles(Parts S, Parts D, long n, const TileGeom& tg, double dx, double dy, int nx, int ny, int* cnt,
               int* cur, long* off, int* idx) {
    if (n <= 0 || tg.ntile <= 0) return;
    if (tg.ntile > 16384) throw std::runtime_error("bin tiles: more than 16384 tiles");
    hip_check(hipMemsetAsync(cnt, 0, sizeof(int) * size_t(tg.ntile), stream()), "hipMemsetAsync");
    const unsigned blocks = unsigned(std::min<long>(1024, long(blocks_for(n, kT))));
    hipLaunchKernelGGL(k_bin_count, dim3(blocks), dim3(kT), sizeof(int) * size_t(tg.ntile), stream(), S, D, n, tg, dx,
                       dy, nx, ny, cnt);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, stream(), cnt, off, cur, tg.ntile);
    hipLaunchKernelGGL(k_bin_scatter, dim3(blocks_for(n, kT)), dim3(kT), 2 * sizeof(int) * size_t(tg.ntile), stream(),
                       S, n, tg, dx, dy, nx, ny, cur, idx);
    launch_check("bin tiles");
}
void append_born(Parts born, const int* bcnt, Parts p0, Parts p1, long base, long cap) {
    if (cap <= base) return;
    hipLaunchKernelGGL(k_append_born, dim3(blocks_for(cap - base, kT)), dim3(kT), 0, stream(), born, bcnt, p0, p1, base,
                       cap);
    launch_check("append born");
}
void bunch_centroid(Parts p, long np, double X0, double Y0, double* out3) {
    static double* part = nullptr;
    const int nb = kCentroidBlocks;
    if (!part) hip_check(hipMalloc(reinterpret_cast<void**>(&part), 3 * nb * sizeof(double)), "hipMalloc");
    hipLaunchKernelGGL(k_centroid_partial, dim3(nb), dim3(kT), 0, stream(), p, np, X0, Y0, part);
    hipLaunchKernelGGL(k_centroid_final, dim3(1), dim3(64), 0, stream(), part, nb, out3);
    launch_check("centroid");
}
void collect_alive(Parts p, long n, int code, g3d_particle* out, long* idx, int* count, long cap) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_collect, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), p, n, code, out, idx, count, cap);
    launch_check("collect");
}
void count_alive(Parts p, long n, int code, int* count) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_count, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), p, n, code, count);
    launch_check("count");
}
void parts_from_aos(const g3d_particle* in, Parts p, long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_from_aos, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), in, p, n);
    launch_check("particles from AoS");
}
void parts_to_aos(Parts p, g3d_particle* out, long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_aos, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), p, out, n);
    launch_check("particles to AoS");
}
// the fresh plasma lattice of a step, generated in tile order (see LatticeArgs); both list buffers
__global__ void k_lattice(LatticeArgs L, Parts A, Parts B, long n) {
    const long g = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int pp = L.px * L.py;
    const long t = g / L.per_tile, r = g - t * L.per_tile;
    const long cpt = L.per_tile / pp;                // cells of a tile: the sub-position outermost
    const int sub = int(r / cpt);
    const long c = r - sub * cpt;
    const int jl = int(c / L.TX), il = int(c - long(jl) * L.TX);
    const int b2 = sub / L.px, a2 = sub - b2 * L.px;
    const int tx = int(t % L.ntx), ty = int(t / L.ntx);
    const int j = ty * L.TY + jl, i = tx * L.TX + il;
    // the host loader's formulas and operation order (src/sim_plasma.cpp:62-64)
    const double y0 = j * L.dy, hy = L.dy;
    const double x = (i + (a2 + 0.5) / L.px) * L.dx;
    const double y = y0 + (b2 + 0.5) / L.py * hy;
    const double X = x + L.X0, Y = y + L.Y0, r2 = X * X + Y * Y;
    const double gch = L.parabolic ? 1 + L.channel_depth * r2 / (L.r_channel * L.r_channel) : 1.0;
    const double dens = L.f * gch > 0 ? L.n0 * L.f * gch : 0.0;
    const double w = dens * L.dx * hy * L.Delta / (L.px * L.py);
    A.x[g] = x; A.y[g] = y; A.zeta[g] = L.zf; A.px[g] = 0; A.py[g] = 0; A.pz[g] = 0; A.w[g] = w; A.rem[g] = 0;
    A.z[g] = L.s0; A.q[g] = -1; A.qm[g] = -1; A.alive[g] = 1; A.sp[g] = SP_PLASMA;
    A.id[g] = L.id0 + (((unsigned long long)j * L.nx + i) * L.py + b2) * L.px + a2;
    B.x[g] = x; B.y[g] = y; B.zeta[g] = L.zf; B.px[g] = 0; B.py[g] = 0; B.pz[g] = 0; B.z[g] = L.s0; B.q[g] = -1;
    B.qm[g] = -1; B.alive[g] = 1;
}
void plasma_lattice(const LatticeArgs& L, Parts A, Parts B, long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_lattice, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), L, A, B, n);
    launch_check("plasma lattice");
}
void parts_copy(Parts s, Parts d, long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), s, d, n);
    launch_check("particles copy");
}
void parts_shift(Parts p, long n, double sx, double sy) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_shift, dim3(blocks_for(n, kT)), dim3(kT), 0, stream(), p, n, sx, sy);
    launch_check("particles shift");
}

}  // namespace g3d
