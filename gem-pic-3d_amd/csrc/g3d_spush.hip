// The streaming push of one electron through a half-cell, forward and reverse, on a periodic Fourier-y grid with
// the fields of the two bounding slices given in FP64: the differentiable piece of the plasma pass.
//
// Forward: the operations of stream_one (g3d_particles.hip; src/sim_stream.hpp:104-171) for an electron with no
// ionization and no analytic laser, in the same order, with the gather and the deposit of the Fourier-y path.
// Reverse: the forward once with its intermediates, then every statement transposed in reverse order (the Boris
// rotations through boris_rev), so that one lane gets the gradient of
//     S = <adjoint of the pushed state, pushed state> + <adjoint of J, this electron's deposit>
// with respect to its start state and its gathered fields; the latter go back to the field planes through the
// transpose of the gather (the TSC and ζ weights of the midpoint).  The tangent-linear push runs the same code on
// dual numbers of one direction.  The position dependence of the gathered fields
// and of the deposit weights enters through the derivatives of the TSC weights.  Branches (entry of a not yet
// streaming electron, trapping, the clamp of v_z, the sign of the charge) follow the values, as in the forward
// push, and are held fixed.  No counterpart of the reverse in the reference.
#include "g3d_boris_rev.hpp"
#include "g3d_dual.hpp"
#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 256;

// a periodic coordinate moved into [0, L): the same shift on the value and no change of its derivatives
__device__ __forceinline__ double twrap(double x, double L) { return wrap_mod(x, L); }
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

// The TSC weights of a point, their x and y derivatives and nodes (tsc_wd on both axes)
struct Tsc2 {
    long k[9];
    double w[9], wdx[9], wdy[9];
};
__device__ __forceinline__ Tsc2 tsc2(const g3d_spush_ctx& c, double x, double y) {
    Tsc2 t;
    int ix[3], jn[3]; double wx[3], wn[3], dwx[3], dwn[3];
    tsc_wd(x, c.dx, c.nx, ix, wx, dwx);
    tsc_wd(y, c.dy, c.ny, jn, wn, dwn);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            t.k[3 * j + i] = long(jn[j]) * c.nx + ix[i];
            t.w[3 * j + i] = wx[i] * wn[j];
            t.wdx[3 * j + i] = dwx[i] * wn[j];
            t.wdy[3 * j + i] = wx[i] * dwn[j];
        }
    return t;
}
// the reverse of a deposit of F v at (xd, yd) into J (F the whole factor with the inverse area): adds the adjoints
// of F, v and the point
__device__ __forceinline__ void deposit_rev(const g3d_spush_ctx& c, const double* bJ, double xd, double yd, double F,
                                            const double v[3], double& Fb, double vb[3], double& xdb, double& ydb) {
    const long nk = long(c.nx) * c.ny;
    const Tsc2 t = tsc2(c, xd, yd);
    const int cq = F < 0 ? 3 : 4;
    double sw = 0, sx = 0, sy = 0, sv[3] = {0, 0, 0};
    for (int q = 0; q < 9; ++q) {
        const long k = t.k[q];
        const double b0 = bJ[k], b1 = bJ[nk + k], b2 = bJ[2 * nk + k];
        const double G = b0 * v[0] + b1 * v[1] + b2 * v[2] + bJ[cq * nk + k];
        sw += t.w[q] * G;
        sx += t.wdx[q] * G;
        sy += t.wdy[q] * G;
        sv[0] += t.w[q] * b0; sv[1] += t.w[q] * b1; sv[2] += t.w[q] * b2;
    }
    Fb += sw;
    for (int q = 0; q < 3; ++q) vb[q] += F * sv[q];
    xdb += F * sx;
    ydb += F * sy;
}

// The reverse of the push (T = double in tpush), written by hand: the forward once with its intermediates kept, then
// every statement transposed in reverse order; the Boris rotations through boris_rev.  The branch decisions and
// the skipped deposit of a zero factor are the forward's.  Gives the same numbers as the forward-mode reverse on
// 13-direction dual numbers, in registers (that one spilled 828 bytes per lane).
__global__ void __launch_bounds__(kT) k_spush_vjp(g3d_spush_ctx c, g3d_spush_adj a) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.np) return;
    const long nk = long(c.nx) * c.ny;
    const double x = c.x[i], y = c.y[i], zeta = c.zeta[i], z = c.z[i];
    const double p[3] = {c.px[i], c.py[i], c.pz[i]};
    const double q = c.q[i], qm = c.qm[i], w = c.w[i];
    double xb = a.bx[i], yb = a.by[i], zetab = a.bzeta[i], zb = a.bz[i];
    double pb[3] = {a.bpx[i], a.bpy[i], a.bpz[i]};
    auto out = [&](double gx, double gy, double gzeta, double gz, const double gp[3]) {
        a.gx[i] = gx; a.gy[i] = gy; a.gzeta[i] = gzeta; a.gz[i] = gz; a.gpx[i] = gp[0]; a.gpy[i] = gp[1];
        a.gpz[i] = gp[2];
    };
    int alive = c.alive[i];
    const double zn = c.zn, zn1 = c.zn1, h = zn1 - zn, Lx = c.nx * c.dx, Ly = c.ny * c.dy;
    if (alive == 6 && zeta > zn) alive = 1;
    if (alive != 1) { out(xb, yb, zetab, zb, pb); return; }   // not pushed: the identity
    // ---- forward, keeping the intermediates
    const int hmode = zeta >= zn1 - 1e-9 * h ? 0 : (zeta - zn > 0 ? 1 : 2);
    const double hl = hmode == 0 ? h : (hmode == 1 ? zeta - zn : 0.0);
    const double g0 = sqrt(1 + p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const double A0 = 1 + p[0] * p[0] + p[1] * p[1], D0 = g0 * (g0 + p[2]);
    const double omv0 = A0 / D0;
    const double dt0 = hl / omv0;
    const double xm = x + 0.5 * dt0 * p[0] / g0, ym = y + 0.5 * dt0 * p[1] / g0;
    const double zetam = zeta - 0.5 * hl;
    double E[3], B[3];
    {
        const GatherPoint gp = gather_point(c, twrap(xm, Lx), twrap(ym, Ly), zetam);
        for (int comp = 0; comp < 6; ++comp) {
            const double v = gather_comp(c, gp, comp, nullptr, nullptr, nullptr);
            if (comp < 3) E[comp] = v; else B[comp - 3] = v;
        }
    }
    double u1[3] = {p[0], p[1], p[2]};
    boris(u1[0], u1[1], u1[2], E, B, qm, dt0);
    const double g1 = sqrt(1 + u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    const double A1 = 1 + u1[0] * u1[0] + u1[1] * u1[1], D1 = g1 * (g1 + u1[2]);
    const double omv1 = A1 / D1;
    const double ssum = omv0 + omv1;
    const double dt = 2 * hl / ssum;
    double u2[3] = {p[0], p[1], p[2]};
    boris(u2[0], u2[1], u2[2], E, B, qm, dt);
    const double g2 = sqrt(1 + u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
    double v2[3];
    for (int k = 0; k < 3; ++k) v2[k] = 0.5 * (p[k] / g0 + u2[k] / g2);
    const bool trapped = z + v2[2] * dt >= c.sTrap;
    double uF[3], gF, vF[3], DT, dtq = 0, vzc = 0;
    bool vz_clamped = false, dtp_clamped = false;
    if (trapped) {
        vz_clamped = !(v2[2] > 1e-3);
        vzc = vz_clamped ? 1e-3 : v2[2];
        dtq = (c.sTrap - z) / vzc;
        dtp_clamped = !(dtq > 0);
        DT = dtp_clamped ? 0.0 : dtq;
        for (int k = 0; k < 3; ++k) uF[k] = p[k];
        boris(uF[0], uF[1], uF[2], E, B, qm, DT);
        gF = sqrt(1 + uF[0] * uF[0] + uF[1] * uF[1] + uF[2] * uF[2]);
        for (int k = 0; k < 3; ++k) vF[k] = 0.5 * (p[k] / g0 + uF[k] / gF);
    } else {
        for (int k = 0; k < 3; ++k) { uF[k] = u2[k]; vF[k] = v2[k]; }
        gF = g2;
        DT = dt;
    }
    const double x1 = x + vF[0] * DT, y1 = y + vF[1] * DT;
    const double f = q * w * DT / (c.Delta * h * c.dx);
    // ---- reverse
    double uFb[3] = {pb[0], pb[1], pb[2]};           // p1 = uF
    double vFb[3] = {0, 0, 0}, DTb = 0, gFb = 0, x1b = xb, y1b = yb;
    double pbv[3] = {0, 0, 0};                        // the adjoint of the start momentum
    double xs = 0, ys = 0, zs_ = 0, zetas = 0;       // the adjoints of the start position
    if (trapped) {
        const double vze = uF[2] / gF;
        if (vze >= c.vz_beam) {                       // the beam current of the trapped electron at (x1, y1)
            const double C = q * w * c.ia / (h * c.dx);
            const double F2 = C / vze;
            const double v3[3] = {uF[0] / gF, uF[1] / gF, vze};
            double F2b = 0, v3b[3] = {0, 0, 0}, xdb = 0, ydb = 0;
            deposit_rev(c, a.bJ, twrap(x1, Lx), twrap(y1, Ly), F2, v3, F2b, v3b, xdb, ydb);
            x1b += xdb; y1b += ydb;
            double vzeb = v3b[2] - F2b * C / (vze * vze);
            uFb[0] += v3b[0] / gF; uFb[1] += v3b[1] / gF; uFb[2] += vzeb / gF;
            gFb -= (v3b[0] * uF[0] + v3b[1] * uF[1] + vzeb * uF[2]) / (gF * gF);
        }
    }
    if (f != 0) {                                     // the current of the half-cell at the midpoint
        const double F = f * c.ia;
        double Fb = 0, xdb = 0, ydb = 0;
        deposit_rev(c, a.bJ, twrap(0.5 * (x + x1), Lx), twrap(0.5 * (y + y1), Ly), F, vF, Fb, vFb, xdb, ydb);
        xs += 0.5 * xdb; x1b += 0.5 * xdb;
        ys += 0.5 * ydb; y1b += 0.5 * ydb;
        DTb += Fb * c.ia * q * w / (c.Delta * h * c.dx);
    }
    // zeta1: trapped: zeta - (1 - vF_z) DT, else zn; z1: trapped: s1, else z + vF_z DT
    if (trapped) {
        zetas += zetab;
        vFb[2] += zetab * DT;
        DTb -= zetab * (1 - vF[2]);
    } else {
        zs_ += zb;
        vFb[2] += zb * DT;
        DTb += zb * vF[2];
    }
    // x1 = x + vF_x DT, y1 likewise
    vFb[0] += x1b * DT; vFb[1] += y1b * DT; DTb += x1b * vF[0] + y1b * vF[1];
    xs += x1b; ys += y1b;
    // vF = (p / g0 + uF / gF) / 2
    double g0b = 0;
    for (int k = 0; k < 3; ++k) {
        pbv[k] += 0.5 * vFb[k] / g0;
        g0b -= 0.5 * vFb[k] * p[k] / (g0 * g0);
        uFb[k] += 0.5 * vFb[k] / gF;
        gFb -= 0.5 * vFb[k] * uF[k] / (gF * gF);
    }
    for (int k = 0; k < 3; ++k) uFb[k] += gFb * uF[k] / gF;
    double Eb[3] = {0, 0, 0}, Bb[3] = {0, 0, 0}, qmb = 0, dtb = 0;
    double ab[3];
    if (trapped) {
        double DTbb = 0;                              // the adjoint of dtp through its own rotation
        boris_rev(p, E, B, qm, DT, uFb, ab, Eb, Bb, qmb, DTbb);
        for (int k = 0; k < 3; ++k) pbv[k] += ab[k];
        DTb += DTbb;
        // dtp = max(dtq, 0), dtq = (sTrap - z) / vzc, vzc = max(v2_z, 1e-3)
        const double dtqb = dtp_clamped ? 0.0 : DTb;
        zs_ -= dtqb / vzc;
        const double vzcb = -dtqb * dtq / vzc;
        double v2b[3] = {0, 0, vz_clamped ? 0.0 : vzcb};
        // v2 = (p / g0 + u2 / g2) / 2
        double u2b[3], g2b = 0;
        for (int k = 0; k < 3; ++k) {
            pbv[k] += 0.5 * v2b[k] / g0;
            g0b -= 0.5 * v2b[k] * p[k] / (g0 * g0);
            u2b[k] = 0.5 * v2b[k] / g2;
            g2b -= 0.5 * v2b[k] * u2[k] / (g2 * g2);
        }
        for (int k = 0; k < 3; ++k) u2b[k] += g2b * u2[k] / g2;
        boris_rev(p, E, B, qm, dt, u2b, ab, Eb, Bb, qmb, dtb);
        for (int k = 0; k < 3; ++k) pbv[k] += ab[k];
    } else {
        dtb += DTb;
        boris_rev(p, E, B, qm, dt, uFb, ab, Eb, Bb, qmb, dtb);
        for (int k = 0; k < 3; ++k) pbv[k] += ab[k];
    }
    // dt = 2 hl / (omv0 + omv1)
    double hlb = dtb * 2 / ssum;
    const double sb = -dtb * dt / ssum;
    double omv0b = sb;
    const double omv1b = sb;
    // omv1 = A1 / D1, A1 = 1 + u1x^2 + u1y^2, D1 = g1 (g1 + u1z)
    {
        const double A1b = omv1b / D1, D1b = -omv1b * omv1 / D1;
        double u1b[3] = {2 * u1[0] * A1b, 2 * u1[1] * A1b, D1b * g1};
        const double g1b = D1b * (2 * g1 + u1[2]);
        for (int k = 0; k < 3; ++k) u1b[k] += g1b * u1[k] / g1;
        double dt0b = 0;
        boris_rev(p, E, B, qm, dt0, u1b, ab, Eb, Bb, qmb, dt0b);
        for (int k = 0; k < 3; ++k) pbv[k] += ab[k];
        // the gather at (xm, ym, zetam), its point formed again: the field values' adjoints go to the planes, their
        // position derivatives on
        const GatherPoint gp = gather_point(c, twrap(xm, Lx), twrap(ym, Ly), zetam);
        double xmb = 0, ymb = 0, zetamb = 0;
        for (int comp = 0; comp < 6; ++comp) {
            const double e = comp < 3 ? Eb[comp] : Bb[comp - 3];
            double gxc, gyc, gzc;
            gather_comp(c, gp, comp, &gxc, &gyc, &gzc);
            xmb += e * gxc; ymb += e * gyc; zetamb += e * gzc;
            if (e == 0) continue;
            for (int t = 0; t < 9; ++t) {
                atomicAdd(&a.gF0[comp * nk + gp.k[t]], e * gp.w[t] * gp.gz);
                atomicAdd(&a.gF1[comp * nk + gp.k[t]], e * gp.w[t] * gp.fz);
            }
        }
        // zetam = zeta - hl / 2; xm = x + dt0 px / (2 g0); ym likewise
        zetas += zetamb;
        hlb -= 0.5 * zetamb;
        xs += xmb; ys += ymb;
        dt0b += 0.5 * (xmb * p[0] + ymb * p[1]) / g0;
        pbv[0] += 0.5 * xmb * dt0 / g0;
        pbv[1] += 0.5 * ymb * dt0 / g0;
        g0b -= 0.5 * (xmb * p[0] + ymb * p[1]) * dt0 / (g0 * g0);
        // dt0 = hl / omv0
        hlb += dt0b / omv0;
        omv0b -= dt0b * dt0 / omv0;
    }
    // omv0 = A0 / D0, A0 = 1 + px^2 + py^2, D0 = g0 (g0 + pz)
    {
        const double A0b = omv0b / D0, D0b = -omv0b * omv0 / D0;
        pbv[0] += 2 * p[0] * A0b; pbv[1] += 2 * p[1] * A0b; pbv[2] += D0b * g0;
        g0b += D0b * (2 * g0 + p[2]);
    }
    for (int k = 0; k < 3; ++k) pbv[k] += g0b * p[k] / g0;
    if (hmode == 1) zetas += hlb;                      // hl = zeta - zn
    out(xs, ys, zetas, zs_, pbv);
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
