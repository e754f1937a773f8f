// The FP32-mixed form of the plasma pass (Fourier y, electron streams): the deposit, the transforms and the
// per-mode implicit update run in FP32 on the vector ALUs.  FP64 is kept for
//   - the particle coordinates x, y, zeta, z and their accumulation along s (the increments are FP32),
//   - the trapezoidal s-integration of R (R and its explicit part R + D (F - J)),
//   - the running currents jk, j0, js0 and the prediction bookkeeping (the charge totals),
// and the spectra stay FP64 in HBM, so the FP64 and mixed passes share every array outside the pass.
//
//   k_stream_mixed      one half-cell pass of the streaming electrons (src/sim_stream.hpp:104-171): gather and
//                       deposit through the tile's LDS windows in FP32, FP32 Boris rotations
//   k_rows_fwd_mixed    jsr + slabs (float) -> x transforms in FP32 -> pair unpacked per row
//   k_cols_mixed        y transforms of three planes in FP32 (forward or inverse)
//   k_modes_mixed       filters, bookkeeping (FP64), march with the FP32 implicit solve, packed planes
//   k_rows_inv_mixed    x inverse transforms in FP32 and E, B of the slice into F
#include <cstdlib>
#include "g3d_fft_core.hpp"
#include "g3d_pass.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
using namespace fftcore;
constexpr int kT = 256;
// the FP32 arithmetic of this file may contract a * b + c into one fused operation (the FP64 parity build keeps
// contraction off everywhere else)
#pragma clang fp contract(fast)

// FP32 reciprocal and square root of the hardware (v_rcp_f32, v_sqrt_f32: 1 ulp), for the FP32 physics
__device__ __forceinline__ float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_f(float x) { return __builtin_amdgcn_sqrtf(x); }

// 1/dx, 1/dy and the first inverse area factor, once per lane
struct InvF {
    double ix, iy;
    float ia0;
};

// TSC weights in FP32 on a periodic uniform axis: the cell from the FP64 index coordinate x / d (x * (1/d): the
// cell may differ from the FP64 pass where x / d lies within an ulp of a half-integer), the fraction in FP32.
// Returns the middle node's index, not wrapped (the window path wraps only its offset; tsc_nodes wraps all three).
__device__ __forceinline__ int tsc_f(double x, double inv, float w[3]) {
    const double xi = x * inv;
    const int i0 = int(floor(xi + 0.5));
    const float f = float(xi - i0);
    const float a = 0.5f - f, b = 0.5f + f;
    w[0] = 0.5f * a * a;
    w[1] = 0.75f - f * f;
    w[2] = 0.5f * b * b;
    return i0;
}
__device__ __forceinline__ void tsc_nodes(int i0, int n, int idx[3]) {
    idx[0] = wrap_index(i0 - 1, n); idx[1] = wrap_index(i0, n); idx[2] = wrap_index(i0 + 1, n);
}

__device__ __forceinline__ void boris_f(float& ux, float& uy, float& uz, const float E[3], const float B[3], float qm,
                                        float dt) {
    const float qh = 0.5f * qm * dt;
    const float ax = ux + qh * E[0], ay = uy + qh * E[1], az = uz + qh * E[2];
    const float qg = qh * rcp_f(sqrt_f(1.0f + ax * ax + ay * ay + az * az));
    const float tx = qg * B[0], ty = qg * B[1], tz = qg * B[2];
    const float t2 = tx * tx + ty * ty + tz * tz;
    const float px = ax + (ay * tz - az * ty), py = ay + (az * tx - ax * tz), pz = az + (ax * ty - ay * tx);
    const float sf = 2.0f * rcp_f(1.0f + t2);
    const float bx = ax + sf * (py * tz - pz * ty), by = ay + sf * (pz * tx - px * tz), bz = az + sf * (px * ty - py * tx);
    ux = bx + qh * E[0]; uy = by + qh * E[1]; uz = bz + qh * E[2];
}

// LDS windows of a tile, node-major so that one node is one address: the field window first (12 floats per node:
// E, B at slice fl then fl + 1, three 16-byte reads) and the current window after it (5 doubles per node: j_x, j_y,
// j_z, rho-, rho+; FP64 sums of the FP32 contributions, since the FP32 LDS atomic add serialises on the
// same-address conflicts of a tile's particles and the FP64 one does not)
constexpr int kFwNode = 12, kDepNode = 5;
struct WinF {
    double* lds;              // kDepNode doubles per node
    const float4* fld;        // 3 float4 per node, or null
    int x0, y0, WX, WY, fl;
};

// E, B at a point: the tile's field window when the stencil and the slice pair are in it, else F
__device__ __forceinline__ void gather_f(const FieldView& c, const InvF& iv, const WinF* w, double x, double y,
                                         double zeta, float E[3], float B[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) { E[k] = 0; B[k] = 0; }
    if (!c.self) return;
    int n; double fz;
    c.zax.locate(zeta, n, fz);
    if (n == c.a - 1 && fz > 1 - 1e-9) { n = c.a; fz = 0; }
    const int fl = n - c.a;
    if (fl < 0 || fl > c.ns) return;
    float wx[3], wn[3];
    const int i0 = tsc_f(x, iv.ix, wx), j0 = tsc_f(y, iv.iy, wn);
    const float f1 = float(fz), f0 = 1.0f - f1;
    float wt[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) wt[3 * j + i] = wx[i] * wn[j];
    float v[6];
    if (w && w->fld && fl == w->fl) {
        const int u0 = wrap_index(i0 - 1 - w->x0, c.nx), v0 = wrap_index(j0 - 1 - w->y0, c.ny);
        if (u0 + 2 < w->WX && v0 + 2 < w->WY) {
            float s0[6] = {0, 0, 0, 0, 0, 0}, s1[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float4* q = w->fld + 3 * ((v0 + j) * w->WX + u0 + i);
                    const float4 a = q[0], b = q[1], d = q[2];
                    const float t = wt[3 * j + i];
                    s0[0] += t * a.x; s0[1] += t * a.y; s0[2] += t * a.z; s0[3] += t * a.w;
                    s0[4] += t * b.x; s0[5] += t * b.y;
                    s1[0] += t * b.z; s1[1] += t * b.w;
                    s1[2] += t * d.x; s1[3] += t * d.y; s1[4] += t * d.z; s1[5] += t * d.w;
                }
            E[0] = f0 * s0[0] + f1 * s1[0]; E[1] = f0 * s0[1] + f1 * s1[1]; E[2] = f0 * s0[2] + f1 * s1[2];
            B[0] = f0 * s0[3] + f1 * s1[3]; B[1] = f0 * s0[4] + f1 * s1[4]; B[2] = f0 * s0[5] + f1 * s1[5];
            return;
        }
    }
    int ix[3], jn[3];
    tsc_nodes(i0, c.nx, ix);
    tsc_nodes(j0, c.ny, jn);
    const long cs = long(c.ns + 2) * c.nk;
    const float* p0 = c.F + long(fl) * c.nk;
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        const float* q0 = p0 + comp * cs;
        const float* q1 = q0 + c.nk;
        float s = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const long k = long(jn[j]) * c.nx + ix[i];
                s += wt[3 * j + i] * (f0 * q0[k] + f1 * q1[k]);
            }
        v[comp] = s;
    }
    E[0] = v[0]; E[1] = v[1]; E[2] = v[2]; B[0] = v[3]; B[1] = v[4]; B[2] = v[5];
}

// TSC deposit of f v into the FP32 current planes (window or jsr), as the FP64 deposit; the charge density planes
// (rho-, rho+) only when rho (their one reader, the density lines, takes the last pass's)
__device__ __forceinline__ void sdeposit_f(const StreamArgs& c, const InvF& iv, float* J, const WinF* win, double x,
                                           double y, float f, float vx, float vy, float vz, bool rho) {
    const FieldView& P = c.fv;
    float wx[3], wn[3];
    const int i0 = tsc_f(x, iv.ix, wx), j0 = tsc_f(y, iv.iy, wn);
    const long nk = P.nk;
    const int cq = f < 0 ? 3 : 4;
    const bool mv = vx != 0 || vy != 0 || vz != 0;
    const float fi = f * iv.ia0;
    if (win) {
        const int u0 = wrap_index(i0 - 1 - win->x0, P.nx), v0 = wrap_index(j0 - 1 - win->y0, P.ny);
        if (u0 + 2 < win->WX && v0 + 2 < win->WY) {
            // the products in FP64 from the FP32 weights and velocity: the LDS sums are FP64, and an FP64 multiply
            // issues as fast as a conversion
            double dx[3], dy[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) { dx[k] = double(fi * wx[k]); dy[k] = double(wn[k]); }
            const double dvx = vx, dvy = vy, dvz = vz;
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    double* L = win->lds + kDepNode * ((v0 + v) * win->WX + (u0 + u));
                    const double ww = dx[u] * dy[v];
                    if (mv) {
                        atomicAdd(L, ww * dvx);
                        atomicAdd(L + 1, ww * dvy);
                        atomicAdd(L + 2, ww * dvz);
                    }
                    if (rho) atomicAdd(L + cq, ww);
                }
            return;
        }
    }
    int ix[3], jn[3];
    tsc_nodes(i0, P.nx, ix);
    tsc_nodes(j0, P.ny, jn);
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const long k = long(jn[v]) * P.nx + ix[u];
            const float ww = fi * wx[u] * wn[v];
            if (mv) {
                atomicAdd(&J[k], ww * vx);
                atomicAdd(&J[nk + k], ww * vy);
                atomicAdd(&J[2 * nk + k], ww * vz);
            }
            if (rho) atomicAdd(&J[cq * nk + k], ww);
        }
}

// the state of a streaming electron, loaded ahead of its push (FP64 in HBM)
struct MState {
    double x, y, zeta, z, px, py, pz, q, qm, w;
    int alive;
};
__device__ __forceinline__ MState load_mstate(const Parts& S, long i) {
    return MState{S.x[i], S.y[i], S.zeta[i], S.z[i], S.px[i], S.py[i], S.pz[i], S.q[i], S.qm[i], S.w[i], S.alive[i]};
}

// the values of a launch that every electron uses: the half-cell's ends and length, the FP32 factor
// 1 / (Delta h dx) of the deposit, and whether the density planes are deposited
struct HalfCellF {
    double zn, zn1, h;
    float iq;
    bool rho;
};
__device__ __forceinline__ HalfCellF half_cell(const StreamArgs& c, int rho) {
    const double zn = c.zs[c.n], zn1 = c.zs[c.n + 1], h = zn1 - zn;
    return HalfCellF{zn, zn1, h, rcp_f(float(c.Delta * h * c.fv.dx)), rho != 0};
}

// one electron through half-cell n (src/sim_stream.hpp:104-171), the physics of the FP64 pass in FP32; s: its state
// at the start (loaded by the caller), written to D[i].  The TSC positions are not wrapped into the box first: the
// node indices are (wrap_index), and the FP32 fraction does not see the difference.
// LASER: the deck has the analytic laser (its evaluation would otherwise set the register budget of every launch)
template <bool LASER>
__device__ __forceinline__ void stream_one_mixed(const StreamArgs& c, const HalfCellF& hc, const InvF& iv, float* J,
                                                 const WinF* win, const MState& s, Parts D, long i) {
    const double zn = hc.zn, zn1 = hc.zn1, h = hc.h;
    double x = s.x, y = s.y, zeta = s.zeta, z = s.z;
    const double pxd = s.px, pyd = s.py, pzd = s.pz;
    const double q = s.q, qmd = s.qm, w = s.w;
    int alive = s.alive;
    auto store = [&](double ux, double uy, double uz) {
        D.x[i] = x; D.y[i] = y; D.zeta[i] = zeta; D.px[i] = ux; D.py[i] = uy; D.pz[i] = uz;
        D.z[i] = z; D.q[i] = q; D.qm[i] = qmd; D.alive[i] = alive;
    };
    if (alive == 6) {
        if (zeta > zn) alive = 1;
        else { store(pxd, pyd, pzd); return; }
    }
    if (alive != 1) { store(pxd, pyd, pzd); return; }
    const FieldView& P = c.fv;
    const double hl = zeta >= zn1 - 1e-9 * h ? h : (zeta - zn > 0 ? zeta - zn : 0.0);
    const float hf = float(hl), qm = float(qmd);
    const float px = float(pxd), py = float(pyd), pz = float(pzd);
    const float g0 = sqrt_f(1 + px * px + py * py + pz * pz), ig0 = rcp_f(g0);
    const float omv0 = (1 + px * px + py * py) * rcp_f(g0 * (g0 + pz));
    const float dt0 = hf * rcp_f(omv0), hs = 0.5f * dt0 * ig0;
    const double xm = x + double(hs * px), ym = y + double(hs * py);
    const double zm = z + double(hs * pz), zetam = zeta - 0.5 * hl;
    float E[3], B[3];
    gather_f(P, iv, win, xm, ym, zetam, E, B);
    if (LASER && c.laser.on) {
        double Ee[3], Be[3];
        c.laser.field(xm + c.X0, zm, zm - (zetam - c.zc), Ee, Be);
        for (int k = 0; k < 3; ++k) { E[k] += float(Ee[k]); B[k] += float(Be[k]); }
    }
    float ux = px, uy = py, uz = pz;
    boris_f(ux, uy, uz, E, B, qm, dt0);
    float g1 = sqrt_f(1 + ux * ux + uy * uy + uz * uz);
    float dt = 2 * hf * rcp_f(omv0 + (1 + ux * ux + uy * uy) * rcp_f(g1 * (g1 + uz)));
    ux = px; uy = py; uz = pz;
    boris_f(ux, uy, uz, E, B, qm, dt);
    g1 = sqrt_f(1 + ux * ux + uy * uy + uz * uz);
    float ig1 = rcp_f(g1);
    float vx = 0.5f * (px * ig0 + ux * ig1), vy = 0.5f * (py * ig0 + uy * ig1), vz = 0.5f * (pz * ig0 + uz * ig1);
    double zeta_end = zn;
    bool trapped = false;
    if (z + double(vz * dt) >= c.sTrap) {
        const float vzc = vz > 1e-3f ? vz : 1e-3f;
        const float dtp = float((c.sTrap - z) / vzc) > 0 ? float((c.sTrap - z) / vzc) : 0.0f;
        ux = px; uy = py; uz = pz;
        boris_f(ux, uy, uz, E, B, qm, dtp);
        g1 = sqrt_f(1 + ux * ux + uy * uy + uz * uz);
        ig1 = rcp_f(g1);
        vx = 0.5f * (px * ig0 + ux * ig1); vy = 0.5f * (py * ig0 + uy * ig1); vz = 0.5f * (pz * ig0 + uz * ig1);
        dt = dtp;
        zeta_end = zeta - double((1 - vz) * dtp);
        trapped = true;
    }
    const double xo = x, yo = y;
    x += double(vx * dt); y += double(vy * dt);
    z = trapped ? c.s1 : z + double(vz * dt);
    zeta = zeta_end;
    // q w dt / (Delta h dx), the cell factors in FP32
    const float qw = float(q * w) * hc.iq;
    const float f = qw * dt;
    if (f != 0) sdeposit_f(c, iv, J, win, 0.5 * (xo + x), 0.5 * (yo + y), f, vx, vy, vz, hc.rho);
    if (trapped) {
        alive = 5;
        const float vze = uz * ig1;
        if (vze >= float(c.vz_beam)) sdeposit_f(c, iv, J, win, x, y, qw * rcp_f(vze), ux * ig1, uy * ig1, vze, hc.rho);
    }
    store(ux, uy, uz);
}

// WPE: waves per SIMD the register allocation must allow (0: the compiler's choice)
template <int WPE, bool LASER>
__global__ void __launch_bounds__(kStreamBlockMax) __attribute__((amdgpu_waves_per_eu(WPE ? WPE : 1)))
k_stream_mixed(StreamArgs c, float* J, float* slab, Parts S, Parts D, long n, TileGeom tg, int rho) {
    extern __shared__ float4 winf4[];
    const long b = blockIdx.x;
    const long nt = tg.ntile + tg.max_extra;
    if (b < nt) {
        // workgroups b < max_extra: the further chunks of dense tiles (first, so that they run beside the first
        // chunks), added to J with atomics; then tile b - max_extra's first chunk, written to its slab
        long t, lo, hi;
        const bool extra = b < tg.max_extra;
        if (!extra) { t = b - tg.max_extra; lo = tg.off[t]; hi = tg.off[t + 1]; }
        else {
            if (b >= tg.xcount[0]) return;
            t = tg.xtile[b]; lo = tg.xstart[b]; hi = tg.off[t + 1];
        }
        if (tg.chunk > 0 && hi - lo > tg.chunk) hi = lo + tg.chunk;
        // the range, this lane's first particle and its state first: their dependent loads overlap the loads of
        // the field window
        long k = lo + threadIdx.x;
        long i = k < hi ? (tg.idx ? long(tg.idx[k]) : k) : 0;
        MState st{};
        if (k < hi) st = load_mstate(S, i);
        const int pl = tg.WX * tg.WY;
        const int wn = kDepNode * pl;
        const int tx = int(t % tg.ntx), ty = int(t / tg.ntx);
        const int x0 = tx * tg.TX - 1 - tg.M, y0 = ty * tg.TY - 1 - tg.M;
        const FieldView& F = c.fv;
        const int fl = c.n - F.a;
        double* dep = reinterpret_cast<double*>(winf4 + 3 * pl);
        for (int r = threadIdx.x; r < wn; r += blockDim.x) dep[r] = 0;
        if (F.self)      // one node per lane: its 12 values from the 6 component planes of slices fl, fl + 1
            for (int r = threadIdx.x; r < pl; r += blockDim.x) {
                const int v = r / tg.WX, u = r - v * tg.WX;
                const int gx = wrap_index(x0 + u, F.nx), gy = wrap_index(y0 + v, F.ny);
                const float* src = F.F + long(fl) * F.nk + long(gy) * F.nx + gx;
                const long cs = long(F.ns + 2) * F.nk;
                float q[kFwNode];
#pragma unroll
                for (int comp = 0; comp < 6; ++comp) {
                    q[comp] = src[comp * cs];
                    q[6 + comp] = src[comp * cs + F.nk];
                }
                winf4[3 * r] = make_float4(q[0], q[1], q[2], q[3]);
                winf4[3 * r + 1] = make_float4(q[4], q[5], q[6], q[7]);
                winf4[3 * r + 2] = make_float4(q[8], q[9], q[10], q[11]);
            }
        __syncthreads();
        WinF w{dep, F.self ? winf4 : nullptr, x0, y0, tg.WX, tg.WY, fl};
        const InvF iv{1.0 / F.dx, 1.0 / F.dy, float(c.ia[0])};
        const HalfCellF hc = half_cell(c, rho);
        while (k < hi) {
            stream_one_mixed<LASER>(c, hc, iv, J, &w, st, D, i);
            k += blockDim.x;
            if (k < hi) {
                i = tg.idx ? long(tg.idx[k]) : k;
                st = load_mstate(S, i);
            }
        }
        __syncthreads();
        if (!extra && !tg.aflush) {
            float* sl = slab + t * long(wn);
            for (int r = threadIdx.x; r < wn; r += blockDim.x) sl[r] = float(dep[r]);
        } else
            for (int r = threadIdx.x; r < wn; r += blockDim.x) {
                const double v = dep[r];
                if (v == 0) continue;
                const int node = r / kDepNode, cc = r - node * kDepNode, vv = node / tg.WX, u = node - vv * tg.WX;
                const long g = long(wrap_index(y0 + vv, F.ny)) * F.nx + wrap_index(x0 + u, F.nx);
                atomicAdd(&J[cc * F.nk + g], float(v));
            }
        return;
    }
    const long i = tg.n_tiled + (b - nt) * blockDim.x + threadIdx.x;
    if (i < n)
        stream_one_mixed<LASER>(c, half_cell(c, rho), InvF{1.0 / c.fv.dx, 1.0 / c.fv.dy, float(c.ia[0])}, J, nullptr,
                                load_mstate(S, i), D, i);
}

// ---------------------------------------------------------------- rows forward (FP32)
// tr: X column-major (element (kx, j) at kx ny + j), for the fused column kernel
__global__ void __launch_bounds__(kT) k_rows_fwd_mixed(float* J, const float* slab, TileGeom tg, int nx, int ny, int lx,
                                                       int rpw, const cx* twx, cxf* X, double* rl, double* rli, int i0,
                                                       int j0, int tr, int xcd) {
    extern __shared__ cxf smf[];
    const int sp = padded(nx);
    cxf* buf = smf;
    cxf* tws = smf + size_t(2 * rpw) * sp;
    load_twiddles(tws, twx, nx >> 1, false);
    const long nk = long(nx) * ny;
    const int row0 = row_block(blockIdx.x, gridDim.x, xcd) * rpw;
    const int tot = rpw * nx;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, ixx = e - r * nx, j = row0 + r;
        const long i = long(j) * nx + ixx;
        float v[5];
        for (int cc = 0; cc < 5; ++cc) v[cc] = J[cc * nk + i];
        if (tg.ntile) {
            const int tpx = ixx / tg.TX, tpy = j / tg.TY;
            const long wpl = long(tg.WX) * tg.WY;
            for (int dy = -1; dy <= 1; ++dy) {
                if ((dy == 1 && tg.nty <= 2) || (dy != 0 && tg.nty == 1)) continue;
                const int ty = wrap_index(tpy + dy, tg.nty);
                int vv = j - (ty * tg.TY - 1 - tg.M);
                vv = wrap_index(vv, ny);
                if (vv >= tg.WY) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dx == 1 && tg.ntx <= 2) || (dx != 0 && tg.ntx == 1)) continue;
                    const int tx = wrap_index(tpx + dx, tg.ntx);
                    int uu = ixx - (tx * tg.TX - 1 - tg.M);
                    uu = wrap_index(uu, nx);
                    if (uu >= tg.WX) continue;
                    const float* sl = slab + ((long(ty) * tg.ntx + tx) * wpl + long(vv) * tg.WX + uu) * kDepNode;
                    for (int cc = 0; cc < 5; ++cc) v[cc] += sl[cc];
                }
            }
        }
        if (rl) {
            if (j == j0) rl[ixx] = v[3];
            if (ixx == i0) rl[nx + j] = v[3];
            if (rli) {
                if (j == j0) rli[ixx] = v[4];
                if (ixx == i0) rli[nx + j] = v[4];
            }
        }
        for (int cc = 0; cc < 5; ++cc) J[cc * nk + i] = 0;
        const int b = lpos(bitrev(ixx, lx));
        buf[(2 * r) * sp + b] = mkf(v[0], v[1]);
        buf[(2 * r + 1) * sp + b] = mkf(v[2], 0);
    }
    __syncthreads();
    lds_fft(buf, nx, lx, 2 * rpw, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, kx = e - r * nx, j = row0 + r;
        const long i = tr ? long(kx) * ny + j : long(j) * nx + kx;
        const cxf z = buf[(2 * r) * sp + lpos(kx)], zm = conjg(buf[(2 * r) * sp + lpos((nx - kx) & (nx - 1))]);
        const cxf d = z - zm;
        X[i] = 0.5f * (z + zm);
        X[nk + i] = 0.5f * mkf(d.im, -d.re);
        X[2 * nk + i] = buf[(2 * r + 1) * sp + lpos(kx)];
    }
}

// ---------------------------------------------------------------- columns (FP32), nc columns of 3 planes
__global__ void __launch_bounds__(kT) k_cols_mixed(cxf* data, int nx, int ny, int ly, int nc, const cx* tw, int inv) {
    extern __shared__ cxf smf[];
    const int sp = padded(ny);
    cxf* buf = smf;
    cxf* tws = smf + size_t(nc) * sp;
    const int tiles = nx / nc;
    const long p = blockIdx.x / tiles;
    const int c0 = (blockIdx.x % tiles) * nc;
    cxf* g = data + p * long(nx) * ny + c0;
    load_twiddles(tws, tw, ny >> 1, inv);
    const int tot = nc * ny;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int j = e / nc, c = e - j * nc;
        buf[c * sp + lpos(bitrev(j, ly))] = g[long(j) * nx + c];
    }
    __syncthreads();
    lds_fft(buf, ny, ly, nc, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int j = e / nc, c = e - j * nc;
        g[long(j) * nx + c] = buf[c * sp + lpos(j)];
    }
}

// ---------------------------------------------------------------- the per-mode work (bookkeeping FP64)
__global__ void __launch_bounds__(kT) k_modes_mixed(ModeArgs m, const cxf* X, cxf* T, int kind) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nk = m.mv.nk;
    if (i >= nk) return;
    int kx, j;
    mode_xy(m.mv, i, kx, j);
    cx u[3] = {mk(0), mk(0), mk(0)};
    if (kind != 1) {
        const double k2 = m.kxs[kx] * m.kxs[kx] + m.kys[j] * m.kys[j];
        const float f = float((kx == 0 && j == 0 && !m.keep0) ? 0.0 : m.filt[kx] * m.filty[j] / (1 + m.a2 * k2));
        for (int k = 0; k < 3; ++k) u[k] = to_d(f * X[k * nk + i]);
    }
    cx t[3];
    mode_work<true>(m, i, u, t);
    for (int k = 0; k < 3; ++k) T[k * nk + i] = to_f(t[k]);
}

// ---------------------------------------------------------------- rows inverse (FP32) and F of the slice
__global__ void __launch_bounds__(kT) k_rows_inv_mixed(const cxf* T, float* Fn, long fpl, int nx, int ny, int lx,
                                                       int rpw, const cx* twx, float scale, int tr, int xcd) {
    extern __shared__ cxf smf[];
    const int sp = padded(nx);
    cxf* buf = smf;
    cxf* tws = smf + size_t(3 * rpw) * sp;
    load_twiddles(tws, twx, nx >> 1, true);
    const long nk = long(nx) * ny;
    const int row0 = row_block(blockIdx.x, gridDim.x, xcd) * rpw;
    const int tot = rpw * nx;
    for (int e = threadIdx.x; e < 3 * tot; e += blockDim.x) {
        const int k = e / tot, rem = e - k * tot, r = rem / nx, ixx = rem - r * nx;
        const long i = tr ? long(ixx) * ny + row0 + r : long(row0 + r) * nx + ixx;
        buf[(3 * r + k) * sp + lpos(bitrev(ixx, lx))] = T[k * nk + i];
    }
    __syncthreads();
    lds_fft(buf, nx, lx, 3 * rpw, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, ixx = e - r * nx;
        const long i = long(row0 + r) * nx + ixx;
        const cxf R = scale * buf[(3 * r) * sp + lpos(ixx)], L = scale * buf[(3 * r + 1) * sp + lpos(ixx)];
        const cxf Z = scale * buf[(3 * r + 2) * sp + lpos(ixx)];
        Fn[0 * fpl + i] = 0.5f * (R.re + L.re);
        Fn[1 * fpl + i] = 0.5f * (R.im + L.im);
        Fn[3 * fpl + i] = 0.5f * (L.im - R.im);
        Fn[4 * fpl + i] = 0.5f * (R.re - L.re);
        Fn[2 * fpl + i] = Z.re;
        Fn[5 * fpl + i] = Z.im;
    }
}

// ---------------------------------------------------------------- columns + modes + inverse columns, fused
// One workgroup per k_x column (ny lanes, one mode each), the modes column-major (i = kx ny + j): the column's three
// planes of X in LDS, y-transforms, the per-mode work of every mode of the column, the inverse y-transforms of the
// three packed planes, T written back.  The per-mode arrays of a column are contiguous in this layout.
// BS: the largest ny of the instance (one lane per mode); BS = 256 leaves each lane the registers to keep its
// mode's loads in flight together (there is one wave per SIMD at 256 x 256 modes)
constexpr int kColMax = 1024;
template <int BS>
__global__ void __launch_bounds__(BS) k_colmodes_mixed(ModeArgs m, const cxf* X, cxf* T, int ny, int ly,
                                                       const cx* twy, int kind) {
    extern __shared__ cxf smf[];
    const int sp = padded(ny);
    cxf* buf = smf;
    cxf* twf = smf + 3 * sp;
    cxf* twi = twf + (ny >> 1);
    load_twiddles(twf, twy, ny >> 1, false);
    load_twiddles(twi, twy, ny >> 1, true);
    const long nk = m.mv.nk;
    const int kx = blockIdx.x;
    const long c0 = long(kx) * ny;
    const int jj = threadIdx.x;                      // this lane's mode (kx, jj)
    const long i = c0 + jj;
    ModeLoad L;
    if constexpr (BS <= 256) L = mode_load<true>(m, i);     // the mode's state, in flight during the transforms
    if (kind != 1)
        for (int e = jj; e < 3 * ny; e += blockDim.x) {
            const int p = e / ny, j = e - p * ny;
            buf[p * sp + lpos(bitrev(j, ly))] = X[p * nk + c0 + j];
        }
    __syncthreads();
    if (kind != 1) lds_fft(buf, ny, ly, 3, sp, twf);
    cx u[3] = {mk(0), mk(0), mk(0)};
    if (kind != 1) {
        const double k2 = m.kxs[kx] * m.kxs[kx] + m.kys[jj] * m.kys[jj];
        const float f = float((kx == 0 && jj == 0 && !m.keep0) ? 0.0 : m.filt[kx] * m.filty[jj] / (1 + m.a2 * k2));
        for (int k = 0; k < 3; ++k) u[k] = to_d(f * buf[k * sp + lpos(jj)]);
    }
    cx t[3];
    if constexpr (BS <= 256) mode_compute<true>(m, i, L, u, t);
    else mode_work<true>(m, i, u, t);          // (more lanes per workgroup: no registers to hold the state ahead)
    __syncthreads();                                 // every lane has read its X before T overwrites the buffer
    const int b = lpos(bitrev(jj, ly));
    for (int k = 0; k < 3; ++k) buf[k * sp + b] = to_f(t[k]);
    __syncthreads();
    lds_fft(buf, ny, ly, 3, sp, twi);
    for (int k = 0; k < 3; ++k) T[k * nk + i] = buf[k * sp + lpos(jj)];
}

int lanes(int work) { int t = 64; while (t < work && t < kT) t <<= 1; return t; }
}  // namespace

void stream_mixed(const StreamArgs& a, float* J, float* slab, Parts src, Parts dst, long nstrm, const TileGeom& tg,
                  const LaunchTimer& lt, bool rho) {
    if (nstrm <= 0) return;
    const int bs = tg.ntile ? tg.block : kT;
    if (bs < 64 || bs > kStreamBlockMax || bs % 64) throw std::runtime_error("stream half-cell (mixed): bad block size");
    const long tail_blocks = nstrm > tg.n_tiled ? long(blocks_for(nstrm - tg.n_tiled, bs)) : 0;
    const size_t lds = tg.ntile ? size_t(tg.WX) * tg.WY * (kDepNode * sizeof(double) + kFwNode * sizeof(float)) : 0;
    const dim3 grid(unsigned(tg.ntile + tg.max_extra + tail_blocks)), blk(bs);
    auto* k = a.laser.on ? (tg.wpe == 4 ? k_stream_mixed<4, true> : k_stream_mixed<0, true>)
                         : (tg.wpe == 4 ? k_stream_mixed<4, false> : k_stream_mixed<0, false>);
    launch_timed_kernel(lt, k, grid, blk, unsigned(lds), a, J, slab, src, dst, nstrm, tg, int(rho));
    launch_check("stream half-cell (mixed)");
}
void pass_mixed(const ModeArgs& m, const FftPlan& p, float* J, const float* slab, const TileGeom& tg, cxf* X, cxf* T,
                float* Fn, long fpl, double* rl, double* rli, int i0, int j0, int kind, int rpw) {
    const long nk = long(p.nx) * p.ny;
    const int cm = m.mv.cm;
    if (cm && (p.ny > kColMax || p.ny < 2)) throw std::runtime_error("plasma pass: column kernel needs 2 <= ny <= 1024");
    if (kind != 1) {
        const size_t lds = (size_t(2 * rpw) * (p.nx + p.nx / 8) + p.nx / 2) * sizeof(cxf);
        TileGeom tr = tg;
        if (tg.aflush) tr.ntile = 0;      // the windows are already in J
        hipLaunchKernelGGL(k_rows_fwd_mixed, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), lds, stream(), J, slab, tr, p.nx,
                           p.ny, p.lx, rpw, p.twx, X, rl, rli, i0, j0, cm, cm && xcd_rows());
    }
    const size_t li = (size_t(3 * rpw) * (p.nx + p.nx / 8) + p.nx / 2) * sizeof(cxf);
    if (cm) {       // modes column-major: one fused launch between the row transforms
        const size_t lm = (size_t(3) * (p.ny + p.ny / 8) + p.ny) * sizeof(cxf);
        auto* kc = p.ny <= 256 ? k_colmodes_mixed<256> : k_colmodes_mixed<kColMax>;
        hipLaunchKernelGGL(kc, dim3(unsigned(p.nx)), dim3(unsigned(p.ny)), lm, stream(), m, X, T, p.ny, p.ly, p.twy,
                           kind);
        hipLaunchKernelGGL(k_rows_inv_mixed, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), li, stream(), T, Fn, fpl, p.nx,
                           p.ny, p.lx, rpw, p.twx, float(1.0 / (double(p.nx) * p.ny)), 1, xcd_rows());
        launch_check("plasma pass (mixed, column-major modes)");
        return;
    }
    int nc = 8;                        // 64-byte row segments (measured best at 256 x 256 against 4 and 16)
    if (nc > p.nx) nc = p.nx;
    static const int forced = [] { const char* e = getenv("G3D_FFT_COLS_NC"); return e ? atoi(e) : 0; }();
    if (forced > 0 && forced <= p.nx && p.nx % forced == 0) nc = forced;
    const size_t lc = (size_t(nc) * (p.ny + p.ny / 8) + p.ny / 2) * sizeof(cxf);
    const int tc = lanes(nc * (p.ny / 8 > 0 ? p.ny / 8 : 1));
    if (kind != 1)
        hipLaunchKernelGGL(k_cols_mixed, dim3(unsigned(3 * (p.nx / nc))), dim3(tc), lc, stream(), X, p.nx, p.ny, p.ly, nc,
                           p.twy, 0);
    hipLaunchKernelGGL(k_modes_mixed, dim3(blocks_for(nk, kT)), dim3(kT), 0, stream(), m, X, T, kind);
    hipLaunchKernelGGL(k_cols_mixed, dim3(unsigned(3 * (p.nx / nc))), dim3(tc), lc, stream(), T, p.nx, p.ny, p.ly, nc,
                       p.twy, 1);
    hipLaunchKernelGGL(k_rows_inv_mixed, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), li, stream(), T, Fn, fpl, p.nx, p.ny,
                       p.lx, rpw, p.twx, float(1.0 / (double(p.nx) * p.ny)), 0, 0);
    launch_check("plasma pass (mixed)");
}

}  // namespace g3d
