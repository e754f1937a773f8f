// The streaming push of one electron through a half-cell, forward and reverse, on a periodic Fourier-y grid with
// the fields of the two bounding slices given in FP64: the differentiable piece of the plasma pass.
//
// Forward: the operations of stream_one (g3d_particles.hip; src/sim_stream.hpp:104-171) for an electron with no
// ionization and no analytic laser, in the same order, with the gather and the deposit of the Fourier-y path.
// Reverse: the same code run on dual numbers (value + 13 partial derivatives: the start state x, y, ζ, z, p_x,
// p_y, p_z and the gathered E, B), so that one lane gets the gradient of
//     S = <adjoint of the pushed state, pushed state> + <adjoint of J, this electron's deposit>
// with respect to its start state and its gathered fields; the latter go back to the field planes through the
// transpose of the gather (the TSC and ζ weights of the midpoint).  The position dependence of the gathered fields
// and of the deposit weights enters through the derivatives of the TSC weights.  Branches (entry of a not yet
// streaming electron, trapping, the clamp of v_z, the sign of the charge) follow the values, as in the forward
// push, and are held fixed.  No counterpart of the reverse in the reference.
#include "g3d_dual.hpp"
#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 256;
constexpr int kNd = 13;             // x y ζ z px py pz | Ex Ey Ez Bx By Bz

using Dn = Dual<kNd>;

// a periodic coordinate moved into [0, L): the same shift on the value and no change of its derivatives
__device__ __forceinline__ double twrap(double x, double L) { return wrap_mod(x, L); }
__device__ __forceinline__ Dn twrap(const Dn& x, double L) { return x + (wrap_mod(x.v, L) - x.v); }
using D1 = Dual<1>;                 // one direction: the tangent-linear push
__device__ __forceinline__ D1 dt1(double v, double d) { D1 r; r.v = v; r.d[0] = d; return r; }
__device__ __forceinline__ D1 twrap(const D1& x, double L) { return x + (wrap_mod(x.v, L) - x.v); }

// TSC weights of x on a periodic uniform axis (tsc_periodic: the index coordinate x / d) and their derivatives
// with respect to x
__device__ __forceinline__ void tsc_wd(double x, double d, int n, int idx[3], double w[3], double dw[3]) {
    tsc_periodic(x, d, n, idx, w);
    const double xi = x / d;
    const double f = xi - floor(xi + 0.5);
    dw[0] = -(0.5 - f) / d;
    dw[1] = -2 * f / d;
    dw[2] = (0.5 + f) / d;
}

// ---------------------------------------------------------------- the push on T = double or Dn
template <typename T>
struct SP { T x, y, zeta, z, px, py, pz; };

template <typename T>
__device__ __forceinline__ void tboris(T& ux, T& uy, T& uz, const T E[3], const T B[3], double qm, const T& dt) {
    const T qh = 0.5 * qm * dt;
    const T ax = ux + qh * E[0], ay = uy + qh * E[1], az = uz + qh * E[2];
    const T gm = tsqrt(1.0 + ax * ax + ay * ay + az * az);
    const T tx = qh * B[0] / gm, ty = qh * B[1] / gm, tz = qh * B[2] / gm;
    const T t2 = tx * tx + ty * ty + tz * tz;
    const T px = ax + (ay * tz - az * ty), py = ay + (az * tx - ax * tz), pz = az + (ax * ty - ay * tx);
    const T sf = 2.0 / (1.0 + t2);
    const T bx = ax + sf * (py * tz - pz * ty), by = ay + sf * (pz * tx - px * tz), bz = az + sf * (px * ty - py * tx);
    ux = bx + qh * E[0]; uy = by + qh * E[1]; uz = bz + qh * E[2];
}

// gather(xg, yg, zetam, E, B) and deposit(x, y, f, vx, vy, vz) are supplied by the caller
template <typename T, class Gather, class Deposit>
__device__ __forceinline__ void tpush(const g3d_spush_ctx& c, SP<T>& p, double q, double qm, double w, int& alive,
                                      Gather&& gather, Deposit&& deposit) {
    const double zn = c.zn, zn1 = c.zn1, h = zn1 - zn;
    const double Lx = c.nx * c.dx, Ly = c.ny * c.dy;
    if (alive == 6) {
        if (val(p.zeta) > zn) alive = 1;
        else return;
    }
    if (alive != 1) return;
    const T hl = val(p.zeta) >= zn1 - 1e-9 * h ? cst<T>(h) : (val(p.zeta) - zn > 0 ? p.zeta - zn : cst<T>(0));
    const T g0 = tsqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz);
    const T omv0 = (1 + p.px * p.px + p.py * p.py) / (g0 * (g0 + p.pz));
    const T dt0 = hl / omv0;
    const T xm = p.x + 0.5 * dt0 * p.px / g0, ym = p.y + 0.5 * dt0 * p.py / g0;
    const T zetam = p.zeta - 0.5 * hl;
    T E[3], B[3];
    gather(twrap(xm, Lx), twrap(ym, Ly), zetam, E, B);
    T ux = p.px, uy = p.py, uz = p.pz;
    tboris(ux, uy, uz, E, B, qm, dt0);
    T g1 = tsqrt(1 + ux * ux + uy * uy + uz * uz);
    T dt = 2 * hl / (omv0 + (1 + ux * ux + uy * uy) / (g1 * (g1 + uz)));
    ux = p.px; uy = p.py; uz = p.pz;
    tboris(ux, uy, uz, E, B, qm, dt);
    g1 = tsqrt(1 + ux * ux + uy * uy + uz * uz);
    T vx = 0.5 * (p.px / g0 + ux / g1), vy = 0.5 * (p.py / g0 + uy / g1), vz = 0.5 * (p.pz / g0 + uz / g1);
    T zeta_end = cst<T>(zn);
    bool trapped = false;
    if (val(p.z + vz * dt) >= c.sTrap) {
        const T vzc = val(vz) > 1e-3 ? vz : cst<T>(1e-3);
        const T dtq = (c.sTrap - p.z) / vzc;
        const T dtp = val(dtq) > 0 ? dtq : cst<T>(0);
        ux = p.px; uy = p.py; uz = p.pz;
        tboris(ux, uy, uz, E, B, qm, dtp);
        g1 = tsqrt(1 + ux * ux + uy * uy + uz * uz);
        vx = 0.5 * (p.px / g0 + ux / g1); vy = 0.5 * (p.py / g0 + uy / g1); vz = 0.5 * (p.pz / g0 + uz / g1);
        dt = dtp;
        zeta_end = p.zeta - (1 - vz) * dtp;
        trapped = true;
    }
    const T xo = p.x, yo = p.y;
    p.x = p.x + vx * dt; p.y = p.y + vy * dt;
    p.z = trapped ? cst<T>(c.s1) : p.z + vz * dt;
    p.px = ux; p.py = uy; p.pz = uz;
    p.zeta = zeta_end;
    const T f = q * w * dt / (c.Delta * h * c.dx);
    if (val(f) != 0) deposit(twrap(0.5 * (xo + p.x), Lx), twrap(0.5 * (yo + p.y), Ly), f, vx, vy, vz);
    if (trapped) {
        alive = 5;
        const T vze = uz / g1;
        if (val(vze) >= c.vz_beam)
            deposit(twrap(p.x, Lx), twrap(p.y, Ly), q * w / (h * c.dx * vze), ux / g1, uy / g1, vze);
    }
}

// the gather of the Fourier-y path (gather_t) on the two given slices: nodes, weights and the ζ weights
struct GatherPoint {
    long k[9];
    double w[9], wdx[9], wdy[9], gz, fz;
};
__device__ __forceinline__ GatherPoint gather_point(const g3d_spush_ctx& c, double x, double y, double zeta) {
    GatherPoint g;
    int ix[3], jn[3]; double wx[3], wn[3], dwx[3], dwn[3];
    tsc_wd(x, c.dx, c.nx, ix, wx, dwx);
    tsc_wd(y, c.dy, c.ny, jn, wn, dwn);
    g.fz = (zeta - c.zeta0) * c.ihz - c.n;
    g.gz = 1 - g.fz;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            g.k[3 * j + i] = long(jn[j]) * c.nx + ix[i];
            g.w[3 * j + i] = wx[i] * wn[j];
            g.wdx[3 * j + i] = dwx[i] * wn[j];
            g.wdy[3 * j + i] = wx[i] * dwn[j];
        }
    return g;
}
// component comp at the point (as gather_t), and its derivatives in x, y, ζ
__device__ __forceinline__ double gather_comp(const g3d_spush_ctx& c, const GatherPoint& g, int comp, double* gx,
                                              double* gy, double* gzeta) {
    const long nk = long(c.nx) * c.ny;
    const double* q0 = c.F0 + comp * nk;
    const double* q1 = c.F1 + comp * nk;
    double s = 0, sx = 0, sy = 0, sz = 0;
    for (int t = 0; t < 9; ++t) {
        const double a = q0[g.k[t]], b = q1[g.k[t]];
        const double m = g.gz * a + g.fz * b;
        s += g.w[t] * m;
        sx += g.wdx[t] * m;
        sy += g.wdy[t] * m;
        sz += g.w[t] * (b - a);
    }
    if (gx) { *gx = sx; *gy = sy; *gzeta = sz * c.ihz; }
    return s;
}

__global__ void __launch_bounds__(kT) k_spush(g3d_spush_ctx c) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.np) return;
    const long nk = long(c.nx) * c.ny;
    SP<double> p{c.x[i], c.y[i], c.zeta[i], c.z[i], c.px[i], c.py[i], c.pz[i]};
    int alive = c.alive[i];
    auto gather = [&](double xg, double yg, double zm, double E[3], double B[3]) {
        const GatherPoint g = gather_point(c, xg, yg, zm);
        double v[6];
        for (int comp = 0; comp < 6; ++comp) v[comp] = gather_comp(c, g, comp, nullptr, nullptr, nullptr);
        E[0] = v[0]; E[1] = v[1]; E[2] = v[2]; B[0] = v[3]; B[1] = v[4]; B[2] = v[5];
    };
    auto deposit = [&](double x, double y, double f, double vx, double vy, double vz) {
        int ix[3], jn[3]; double wx[3], wn[3];
        tsc_periodic(x, c.dx, c.nx, ix, wx);
        tsc_periodic(y, c.dy, c.ny, jn, wn);
        const int cq = f < 0 ? 3 : 4;
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) {
                const long k = long(jn[v]) * c.nx + ix[u];
                const double ww = f * wx[u] * wn[v] * c.ia;
                atomicAdd(&c.J[k], ww * vx);
                atomicAdd(&c.J[nk + k], ww * vy);
                atomicAdd(&c.J[2 * nk + k], ww * vz);
                atomicAdd(&c.J[cq * nk + k], ww);
            }
    };
    tpush<double>(c, p, c.q[i], c.qm[i], c.w[i], alive, gather, deposit);
    c.x[i] = p.x; c.y[i] = p.y; c.zeta[i] = p.zeta; c.z[i] = p.z; c.px[i] = p.px; c.py[i] = p.py; c.pz[i] = p.pz;
    c.alive[i] = alive;
}

__global__ void __launch_bounds__(kT) k_spush_vjp(g3d_spush_ctx c, g3d_spush_adj a) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.np) return;
    const long nk = long(c.nx) * c.ny;
    SP<Dn> p{dvar<kNd>(c.x[i], 0), dvar<kNd>(c.y[i], 1), dvar<kNd>(c.zeta[i], 2), dvar<kNd>(c.z[i], 3),
             dvar<kNd>(c.px[i], 4), dvar<kNd>(c.py[i], 5), dvar<kNd>(c.pz[i], 6)};
    int alive = c.alive[i];
    Dn S = dconst<kNd>(0);
    GatherPoint gp{};
    bool gathered = false;
    auto gather = [&](const Dn& xg, const Dn& yg, const Dn& zm, Dn E[3], Dn B[3]) {
        gp = gather_point(c, xg.v, yg.v, zm.v);
        gathered = true;
        for (int comp = 0; comp < 6; ++comp) {
            double gx, gy, gzeta;
            const double s = gather_comp(c, gp, comp, &gx, &gy, &gzeta);
            Dn r = dvar<kNd>(s, 7 + comp);
            for (int k = 0; k < 7; ++k) r.d[k] = gx * xg.d[k] + gy * yg.d[k] + gzeta * zm.d[k];
            if (comp < 3) E[comp] = r; else B[comp - 3] = r;
        }
    };
    auto deposit = [&](const Dn& x, const Dn& y, const Dn& f, const Dn& vx, const Dn& vy, const Dn& vz) {
        int ix[3], jn[3]; double wx[3], wn[3], dwx[3], dwn[3];
        tsc_wd(x.v, c.dx, c.nx, ix, wx, dwx);
        tsc_wd(y.v, c.dy, c.ny, jn, wn, dwn);
        const int cq = f.v < 0 ? 3 : 4;
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) {
                const long k = long(jn[v]) * c.nx + ix[u];
                Dn W = dconst<kNd>(wx[u] * wn[v]);
                for (int t = 0; t < kNd; ++t) W.d[t] = dwx[u] * wn[v] * x.d[t] + wx[u] * dwn[v] * y.d[t];
                const Dn ww = f * W * c.ia;
                const double* bJ = a.bJ;
                S = S + ww * (bJ[k] * vx + bJ[nk + k] * vy + bJ[2 * nk + k] * vz + bJ[cq * nk + k]);
            }
    };
    tpush<Dn>(c, p, c.q[i], c.qm[i], c.w[i], alive, gather, deposit);
    S = S + a.bx[i] * p.x + a.by[i] * p.y + a.bzeta[i] * p.zeta + a.bz[i] * p.z + a.bpx[i] * p.px +
        a.bpy[i] * p.py + a.bpz[i] * p.pz;
    a.gx[i] = S.d[0]; a.gy[i] = S.d[1]; a.gzeta[i] = S.d[2]; a.gz[i] = S.d[3];
    a.gpx[i] = S.d[4]; a.gpy[i] = S.d[5]; a.gpz[i] = S.d[6];
    if (!gathered) return;
    for (int comp = 0; comp < 6; ++comp) {          // the transpose of the gather
        const double e = S.d[7 + comp];
        if (e == 0) continue;
        for (int t = 0; t < 9; ++t) {
            atomicAdd(&a.gF0[comp * nk + gp.k[t]], e * gp.w[t] * gp.gz);
            atomicAdd(&a.gF1[comp * nk + gp.k[t]], e * gp.w[t] * gp.fz);
        }
    }
}

// The tangent-linear push: the primal state and fields as k_spush (the state arrays are read, not written), the
// tangents of the state (st: x y zeta z px py pz, [7][np]) and of the two field planes (dF0, dF1) carried through
// it on dual numbers of one direction; the tangent of the deposit is added to dJ, and with `write` the pushed
// tangents replace st.

__global__ void __launch_bounds__(kT) k_spush_tlm(g3d_spush_ctx c, const double* dF0, const double* dF1, double* st,
                                                  double* dJ, int write) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.np) return;
    const long nk = long(c.nx) * c.ny, np = c.np;
    SP<D1> p{dt1(c.x[i], st[i]), dt1(c.y[i], st[np + i]), dt1(c.zeta[i], st[2 * np + i]), dt1(c.z[i], st[3 * np + i]),
             dt1(c.px[i], st[4 * np + i]), dt1(c.py[i], st[5 * np + i]), dt1(c.pz[i], st[6 * np + i])};
    int alive = c.alive[i];
    auto gather = [&](const D1& xg, const D1& yg, const D1& zm, D1 E[3], D1 B[3]) {
        const GatherPoint g = gather_point(c, xg.v, yg.v, zm.v);
        for (int comp = 0; comp < 6; ++comp) {
            double gx, gy, gzeta;
            const double s = gather_comp(c, g, comp, &gx, &gy, &gzeta);
            double ds = gx * xg.d[0] + gy * yg.d[0] + gzeta * zm.d[0];
            for (int t = 0; t < 9; ++t) ds += g.w[t] * (g.gz * dF0[comp * nk + g.k[t]] + g.fz * dF1[comp * nk + g.k[t]]);
            if (comp < 3) E[comp] = dt1(s, ds); else B[comp - 3] = dt1(s, ds);
        }
    };
    auto deposit = [&](const D1& x, const D1& y, const D1& f, const D1& vx, const D1& vy, const D1& vz) {
        int ix[3], jn[3]; double wx[3], wn[3], dwx[3], dwn[3];
        tsc_wd(x.v, c.dx, c.nx, ix, wx, dwx);
        tsc_wd(y.v, c.dy, c.ny, jn, wn, dwn);
        const int cq = f.v < 0 ? 3 : 4;
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) {
                const long k = long(jn[v]) * c.nx + ix[u];
                const D1 W = dt1(wx[u] * wn[v], dwx[u] * wn[v] * x.d[0] + wx[u] * dwn[v] * y.d[0]);
                const D1 ww = f * W * c.ia;
                atomicAdd(&dJ[k], (ww * vx).d[0]);
                atomicAdd(&dJ[nk + k], (ww * vy).d[0]);
                atomicAdd(&dJ[2 * nk + k], (ww * vz).d[0]);
                atomicAdd(&dJ[cq * nk + k], ww.d[0]);
            }
    };
    tpush<D1>(c, p, c.q[i], c.qm[i], c.w[i], alive, gather, deposit);
    if (write) {
        st[i] = p.x.d[0]; st[np + i] = p.y.d[0]; st[2 * np + i] = p.zeta.d[0]; st[3 * np + i] = p.z.d[0];
        st[4 * np + i] = p.px.d[0]; st[5 * np + i] = p.py.d[0]; st[6 * np + i] = p.pz.d[0];
    }
}
}  // namespace

void stream_push_tlm(const g3d_spush_ctx& c, const double* dF0, const double* dF1, double* st, double* dJ, int write) {
    if (c.np <= 0) return;
    hipLaunchKernelGGL(k_spush_tlm, dim3(blocks_for(c.np, kT)), dim3(kT), 0, stream(), c, dF0, dF1, st, dJ, write);
    launch_check("stream push tangent");
}
}  // namespace g3d

using namespace g3d;
extern "C" {
int g3d_stream_push(const g3d_spush_ctx* c) {
    return g3d::guard([&] {
        if (!c || c->nx < 1 || c->ny < 1 || c->np < 0) throw std::invalid_argument("stream push: bad context");
        if (c->np == 0) return;
        hipLaunchKernelGGL(k_spush, dim3(blocks_for(c->np, kT)), dim3(kT), 0, stream(), *c);
        launch_check("stream push");
    });
}
int g3d_stream_push_vjp(const g3d_spush_ctx* c, const g3d_spush_adj* g) {
    return g3d::guard([&] {
        if (!c || !g || c->nx < 1 || c->ny < 1 || c->np < 0)
            throw std::invalid_argument("stream push vjp: bad context");
        if (c->np == 0) return;
        hipLaunchKernelGGL(k_spush_vjp, dim3(blocks_for(c->np, kT)), dim3(kT), 0, stream(), *c, *g);
        launch_check("stream push vjp");
    });
}
}
