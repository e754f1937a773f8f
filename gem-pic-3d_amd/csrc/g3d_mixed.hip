// The FP32-mixed form of the plasma pass (Fourier y, electron streams): the deposit into the current planes, the
// transforms and the per-mode implicit update run in FP32 on the vector ALUs.  FP64 is kept for
//   - the push of the streaming electrons (gather, Boris rotations, positions and momenta and their accumulation
//     along s; see stream_one_mixed) and the sums of the tile's current window,
//   - the trapezoidal s-integration of R (R and its explicit part R + D (F - J)),
//   - the running currents jk, j0, js0 and the prediction bookkeeping (the charge totals),
// and the spectra stay FP64 in HBM, so the FP64 and mixed passes share every array outside the pass.
//
//   k_stream_mixed      one half-cell pass of the streaming electrons (src/sim_stream.hpp:104-171): gather from
//                       the tile's LDS field window, FP64 Boris rotations, deposit into the tile's current window,
//                       added to the FP32 current planes
//   k_rows_fwd_mixed    current planes (float) -> x transforms in FP32 -> pair unpacked per row
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

// The push of the streaming electrons runs in FP64 arithmetic in this form too: FP32 rounding in the gather and
// the Boris step perturbs every trajectory by about 1e-7 per half-cell, and behind the wake, where electron orbits
// cross, the march amplifies such perturbations by a few hundred (measured at 256 x 256 x 1024: a per-slice field
// error of 1.7e-4 with an FP32 push against 6.5e-5 with an FP64 push, both with the FP32 field pass).  Without
// packed FP32 the FP64 and FP32 vector rates of gfx950 are equal; what FP64 adds is the conversions of the float
// fields (in the gather, from a float window: cheaper than the twice as many LDS reads of a double window) and its
// division and square root, which this form takes from the hardware
// estimates (v_rcp_f64, v_rsq_f64) refined by two Newton steps each (the relative error of the estimate is squared
// by each step) instead of the IEEE sequences of the plain FP64 form.
__device__ __forceinline__ double rcp_d(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double rsq_d(double x) {       // 1 / sqrt(x), x > 0
    double r = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = fma(-h * r, r, 0.5);
    r = fma(r, e, r);
    e = fma(-h * r, r, 0.5);
    return fma(r, e, r);
}

// 1/dx and 1/dy, once per lane
struct InvD {
    double ix, iy;
};

// TSC weights on a periodic uniform axis from the index coordinate x / d (taken as x * (1/d): the cell may differ
// from the plain FP64 form where x / d lies within an ulp of a half-integer).  Returns the middle node's index, not
// wrapped (the window path wraps only its offset; tsc_nodes wraps all three).
__device__ __forceinline__ int tsc_d(double x, double inv, double w[3]) {
    const double xi = x * inv;
    const int i0 = int(floor(xi + 0.5));
    const double f = xi - i0;
    const double a = 0.5 - f, b = 0.5 + f;
    w[0] = 0.5 * a * a;
    w[1] = 0.75 - f * f;
    w[2] = 0.5 * b * b;
    return i0;
}
__device__ __forceinline__ void tsc_nodes(int i0, int n, int idx[3]) {
    idx[0] = wrap_index(i0 - 1, n); idx[1] = wrap_index(i0, n); idx[2] = wrap_index(i0 + 1, n);
}

// the relativistic Boris step of src/sim_stream.hpp:8-18 with the fast reciprocal and inverse square root
__device__ __forceinline__ void boris_d(double& ux, double& uy, double& uz, const double E[3], const double B[3],
                                        double qm, double dt) {
    const double qh = 0.5 * qm * dt;
    const double ax = ux + qh * E[0], ay = uy + qh * E[1], az = uz + qh * E[2];
    const double qg = qh * rsq_d(1.0 + ax * ax + ay * ay + az * az);
    const double tx = qg * B[0], ty = qg * B[1], tz = qg * B[2];
    const double t2 = tx * tx + ty * ty + tz * tz;
    const double px = ax + (ay * tz - az * ty), py = ay + (az * tx - ax * tz), pz = az + (ax * ty - ay * tx);
    const double sf = 2.0 * rcp_d(1.0 + t2);
    const double bx = ax + sf * (py * tz - pz * ty), by = ay + sf * (pz * tx - px * tz), bz = az + sf * (px * ty - py * tx);
    ux = bx + qh * E[0]; uy = by + qh * E[1]; uz = bz + qh * E[2];
}

// LDS windows of a tile, node-major so that one node is one address: the field window first (12 floats per node:
// E, B at slice fl then fl + 1, the float fields as stored, widened to double in the gather; three 16-byte reads
// per stencil node) and the current window after it (5 doubles per node: j_x, j_y, j_z, rho-, rho+).  23 x 15
// nodes (16 x 8 tile, margin 2) take 30.4 KB.
constexpr int kFwNode = 12, kDepNode = 5;
struct WinD {
    double* lds;              // kDepNode doubles per node
    const float4* fld;        // 3 float4 per node, or null
    int x0, y0, WX, WY, fl;
};

// E, B at a point: the tile's field window when the stencil and the slice pair are in it, else F
__device__ __forceinline__ void gather_d(const FieldView& c, const InvD& iv, const WinD* w, double x, double y,
                                         double zeta, double E[3], double B[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) { E[k] = 0; B[k] = 0; }
    if (!c.self) return;
    int n; double fz;
    c.zax.locate(zeta, n, fz);
    if (n == c.a - 1 && fz > 1 - 1e-9) { n = c.a; fz = 0; }
    const int fl = n - c.a;
    if (fl < 0 || fl > c.ns) return;
    double wx[3], wn[3];
    const int i0 = tsc_d(x, iv.ix, wx), j0 = tsc_d(y, iv.iy, wn);
    const double f1 = fz, f0 = 1.0 - fz;
    double wt[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) wt[3 * j + i] = wx[i] * wn[j];
    if (w && w->fld && fl == w->fl) {
        const int u0 = wrap_index(i0 - 1 - w->x0, c.nx), v0 = wrap_index(j0 - 1 - w->y0, c.ny);
        if (u0 + 2 < w->WX && v0 + 2 < w->WY) {
            double s0[6] = {0, 0, 0, 0, 0, 0}, s1[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float4* q = w->fld + 3 * ((v0 + j) * w->WX + u0 + i);
                    const float4 a = q[0], b = q[1], d = q[2];
                    const double t = wt[3 * j + i];
                    s0[0] += t * double(a.x); s0[1] += t * double(a.y); s0[2] += t * double(a.z);
                    s0[3] += t * double(a.w); s0[4] += t * double(b.x); s0[5] += t * double(b.y);
                    s1[0] += t * double(b.z); s1[1] += t * double(b.w); s1[2] += t * double(d.x);
                    s1[3] += t * double(d.y); s1[4] += t * double(d.z); s1[5] += t * double(d.w);
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
    double v[6];
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        const float* q0 = p0 + comp * cs;
        const float* q1 = q0 + c.nk;
        double sum = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const long k = long(jn[j]) * c.nx + ix[i];
                sum += wt[3 * j + i] * (f0 * double(q0[k]) + f1 * double(q1[k]));
            }
        v[comp] = sum;
    }
    E[0] = v[0]; E[1] = v[1]; E[2] = v[2]; B[0] = v[3]; B[1] = v[4]; B[2] = v[5];
}

// TSC deposit of f v into the current window (FP64 sums) or, outside it, into the FP32 current planes with float
// atomics (src/sim_stream.hpp:33-48); f includes the inverse area factor.  The charge density planes (rho-, rho+)
// only when rho (their one reader, the density lines, takes the last pass's)
__device__ __forceinline__ void sdeposit_d(const StreamArgs& c, const InvD& iv, float* J, const WinD* win, double x,
                                           double y, double f, double vx, double vy, double vz, bool rho) {
    const FieldView& P = c.fv;
    double wx[3], wn[3];
    const int i0 = tsc_d(x, iv.ix, wx), j0 = tsc_d(y, iv.iy, wn);
    const long nk = P.nk;
    const int cq = f < 0 ? 3 : 4;
    const bool mv = vx != 0 || vy != 0 || vz != 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) wx[k] *= f;
    if (win) {
        const int u0 = wrap_index(i0 - 1 - win->x0, P.nx), v0 = wrap_index(j0 - 1 - win->y0, P.ny);
        if (u0 + 2 < win->WX && v0 + 2 < win->WY) {
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    double* L = win->lds + kDepNode * ((v0 + v) * win->WX + (u0 + u));
                    const double ww = wx[u] * wn[v];
                    if (mv) {
                        atomicAdd(L, ww * vx);
                        atomicAdd(L + 1, ww * vy);
                        atomicAdd(L + 2, ww * vz);
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
            const double ww = wx[u] * wn[v];
            if (mv) {
                atomicAdd(&J[k], float(ww * vx));
                atomicAdd(&J[nk + k], float(ww * vy));
                atomicAdd(&J[2 * nk + k], float(ww * vz));
            }
            if (rho) atomicAdd(&J[cq * nk + k], float(ww));
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

// the values of a launch that every electron uses: the half-cell's ends and length, the deposit factor
// ia / (Delta h dx) of the streaming current and ia / (h dx) of a trapped electron's beam current, and whether the
// density planes are deposited
struct HalfCellD {
    double zn, zn1, h, iq, iqb;
    bool rho;
};
__device__ __forceinline__ HalfCellD half_cell(const StreamArgs& c, int rho) {
    const double zn = c.zs[c.n], zn1 = c.zs[c.n + 1], h = zn1 - zn;
    return HalfCellD{zn, zn1, h, c.ia[0] / (c.Delta * h * c.fv.dx), c.ia[0] / (h * c.fv.dx), rho != 0};
}

// one electron through half-cell n (src/sim_stream.hpp:104-171); s: its state at the start (loaded by the caller),
// written to D[i].  The TSC positions are not wrapped into the box first: the node indices are (wrap_index).
// LASER: the deck has the analytic laser (its evaluation would otherwise set the register budget of every launch)
template <bool LASER>
__device__ __forceinline__ void stream_one_mixed(const StreamArgs& c, const HalfCellD& hc, const InvD& iv, float* J,
                                                 const WinD* win, const MState& s, Parts D, long i) {
    const double zn = hc.zn, zn1 = hc.zn1, h = hc.h;
    double x = s.x, y = s.y, zeta = s.zeta, z = s.z;
    const double px = s.px, py = s.py, pz = s.pz;
    const double q = s.q, qm = s.qm, w = s.w;
    int alive = s.alive;
    // q and qm are not written: this form streams electrons without ionization, so no push changes them, and every
    // loader (the lattice, the host list, the joining particles) writes the same values into both state buffers
    auto store = [&](double ux, double uy, double uz) {
        if (c.skip_state) return;
        D.x[i] = x; D.y[i] = y; D.zeta[i] = zeta; D.px[i] = ux; D.py[i] = uy; D.pz[i] = uz;
        D.z[i] = z; D.alive[i] = alive;
    };
    if (alive == 6) {
        if (zeta > zn) alive = 1;
        else { store(px, py, pz); return; }
    }
    if (alive != 1) { store(px, py, pz); return; }
    const FieldView& P = c.fv;
    const double hl = zeta >= zn1 - 1e-9 * h ? h : (zeta - zn > 0 ? zeta - zn : 0.0);
    const double a0 = 1 + px * px + py * py, s0 = a0 + pz * pz;
    const double ig0 = rsq_d(s0), g0 = s0 * ig0;
    const double omv0 = a0 * rcp_d(g0 * (g0 + pz));
    const double dt0 = hl * rcp_d(omv0), hs = 0.5 * dt0 * ig0;
    const double xm = x + hs * px, ym = y + hs * py, zm = z + hs * pz, zetam = zeta - 0.5 * hl;
    double E[3], B[3];
    gather_d(P, iv, win, xm, ym, zetam, E, B);
    if (LASER && c.laser.on) {
        double Ee[3], Be[3];
        c.laser.field(xm + c.X0, zm, zm - (zetam - c.zc), Ee, Be);
        for (int k = 0; k < 3; ++k) { E[k] += Ee[k]; B[k] += Be[k]; }
    }
    double ux = px, uy = py, uz = pz;
    boris_d(ux, uy, uz, E, B, qm, dt0);
    double a1 = 1 + ux * ux + uy * uy, s1 = a1 + uz * uz;
    double ig1 = rsq_d(s1), g1 = s1 * ig1;
    double dt = 2 * hl * rcp_d(omv0 + a1 * rcp_d(g1 * (g1 + uz)));
    ux = px; uy = py; uz = pz;
    boris_d(ux, uy, uz, E, B, qm, dt);
    ig1 = rsq_d(1 + ux * ux + uy * uy + uz * uz);
    double vx = 0.5 * (px * ig0 + ux * ig1), vy = 0.5 * (py * ig0 + uy * ig1), vz = 0.5 * (pz * ig0 + uz * ig1);
    double zeta_end = zn;
    bool trapped = false;
    if (z + vz * dt >= c.sTrap) {
        const double vzc = vz > 1e-3 ? vz : 1e-3;
        const double dtp = (c.sTrap - z) * rcp_d(vzc) > 0 ? (c.sTrap - z) * rcp_d(vzc) : 0.0;
        ux = px; uy = py; uz = pz;
        boris_d(ux, uy, uz, E, B, qm, dtp);
        ig1 = rsq_d(1 + ux * ux + uy * uy + uz * uz);
        vx = 0.5 * (px * ig0 + ux * ig1); vy = 0.5 * (py * ig0 + uy * ig1); vz = 0.5 * (pz * ig0 + uz * ig1);
        dt = dtp;
        zeta_end = zeta - (1 - vz) * dtp;
        trapped = true;
    }
    const double xo = x, yo = y;
    x += vx * dt; y += vy * dt;
    z = trapped ? c.s1 : z + vz * dt;
    zeta = zeta_end;
    // the streaming current q w dt / (Delta h dx) of the half-cell
    const double qw = q * w;
    const double f = qw * hc.iq * dt;
    if (f != 0) sdeposit_d(c, iv, J, win, 0.5 * (xo + x), 0.5 * (yo + y), f, vx, vy, vz, hc.rho);
    if (trapped) {
        alive = 5;
        const double vze = uz * ig1;
        // joins the beam list: its beam current q w / (h dx v_z) of level s1 from now on
        if (vze >= c.vz_beam) sdeposit_d(c, iv, J, win, x, y, qw * hc.iqb * rcp_d(vze), ux * ig1, uy * ig1, vze, hc.rho);
    }
    store(ux, uy, uz);
}

// WPE: waves per SIMD the register allocation must allow (0: the compiler's choice)
template <int WPE, bool LASER>
__global__ void __launch_bounds__(kStreamBlockMax) __attribute__((amdgpu_waves_per_eu(WPE ? WPE : 1)))
k_stream_mixed(StreamArgs c, float* J, float* slab, Parts S, Parts D, long n, TileGeom tg, int rho) {
    extern __shared__ double2 wind2[];
    const long b = blockIdx.x;
    const long nt = tg.ntile + tg.max_extra;
    if (b < nt) {
        // workgroups b < max_extra: the further chunks of dense tiles (first, so that they run beside the first
        // chunks), added to J with atomics; then tile b - max_extra's first chunk, written to its slab
        long t, lo, hi;
        const bool extra = b < tg.max_extra;
        if (!extra) {
            t = b - tg.max_extra;
            if (tg.xcd) t = row_block(int(t), int(tg.ntile), 1);
            lo = tg.off[t]; hi = tg.off[t + 1];
        }
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
        float4* fw = reinterpret_cast<float4*>(wind2);
        double* dep = reinterpret_cast<double*>(fw + 3 * pl);
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
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    fw[3 * r + k] = make_float4(q[4 * k], q[4 * k + 1], q[4 * k + 2], q[4 * k + 3]);
            }
        __syncthreads();
        WinD w{dep, F.self ? fw : nullptr, x0, y0, tg.WX, tg.WY, fl};
        const InvD iv{1.0 / F.dx, 1.0 / F.dy};
        const HalfCellD hc = half_cell(c, rho);
        unsigned cnt = 0;
        while (k < hi) {
            cnt += pushed_through(st.alive, st.zeta, hc.zn);
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
        if (c.npush) add_pushes(c.npush, cnt);
        return;
    }
    const long i = tg.n_tiled + (b - nt) * blockDim.x + threadIdx.x;
    unsigned cnt = 0;
    if (i < n) {
        const HalfCellD hc = half_cell(c, rho);
        const MState st = load_mstate(S, i);
        cnt = pushed_through(st.alive, st.zeta, hc.zn);
        stream_one_mixed<LASER>(c, hc, InvD{1.0 / c.fv.dx, 1.0 / c.fv.dy}, J, nullptr, st, D, i);
    }
    if (c.npush) add_pushes(c.npush, cnt);
}

// the spectra of the FP32 forward transforms are stored in float (cxf) or, for the FP64 line pass of the free-hand
// grid, in double (cx)
__device__ __forceinline__ void put(cxf* d, cxf v) { *d = v; }
__device__ __forceinline__ void put(cx* d, cxf v) { *d = to_d(v); }

// ---------------------------------------------------------------- rows forward (FP32)
// tr: X column-major (element (kx, j) at kx ny + j), for the fused column kernel
template <class XT>
__global__ void __launch_bounds__(kT) k_rows_fwd_mixed(float* J, const float* slab, TileGeom tg, int nx, int ny, int lx,
                                                       int rpw, const cx* twx, XT* X, double* rl, double* rli, int i0,
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
        put(X + i, 0.5f * (z + zm));
        put(X + nk + i, 0.5f * mkf(d.im, -d.re));
        put(X + 2 * nk + i, buf[(2 * r + 1) * sp + lpos(kx)]);
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
    if (a.npush && long(grid.x) > a.npush_cap) throw std::runtime_error("stream half-cell (mixed): push counter too small");
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
        hipLaunchKernelGGL(k_rows_fwd_mixed<cxf>, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), lds, stream(), J, slab, tr, p.nx,
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

// free-hand y: the x transforms of the current planes in FP32 into the double spectra X of the line pass (the lines
// and the inverse transforms stay FP64)
void rows_fwd_mixed_fh(const FftPlan& p, float* J, cx* X, double* rl, double* rli, int i0, int j0, int rpw) {
    TileGeom tr{};
    const size_t lds = (size_t(2 * rpw) * (p.nx + p.nx / 8) + p.nx / 2) * sizeof(cxf);
    hipLaunchKernelGGL(k_rows_fwd_mixed<cx>, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), lds, stream(), J,
                       static_cast<const float*>(nullptr), tr, p.nx, p.ny, p.lx, rpw, p.twx, X, rl, rli, i0, j0, 0, 0);
    launch_check("plasma rows forward (mixed, free-hand)");
}

}  // namespace g3d
