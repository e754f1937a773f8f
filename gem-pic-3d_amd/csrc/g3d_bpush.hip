// The RK4 push in z of beam particles, forward and reverse, with the fields of the slices given in FP64: the
// differentiable piece of the bunch push.
//
// Forward: the operations of k_bunch_push (g3d_particles.hip; src/sim_kernels.hpp:75-149) in the same order: the
// sub-step length from the slice spacing and the laser phase, four right-hand sides per sub-step, each with the
// gather (TSC in x and y, linear in ζ) and the analytic plane wave, the hand-over, slow and lost particle codes.
// Reverse: one lane per particle runs the sub-steps backwards.  For sub-step j it recomputes the state at the start
// of the sub-step from the start of the push (no storage; the pushes take few sub-steps), then runs the sub-step on
// dual numbers (the 6 state values and the E, B of its four stages as the inputs), which gives the adjoint of the
// state before the sub-step from the adjoint after it, and the adjoints of the four gathered fields; those go back
// to the field planes through the transpose of each stage's gather.  Sub-step lengths and every branch are the
// forward push's and are held fixed.  No counterpart of the reverse in the reference.
#include "g3d_dual.hpp"
#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 64;                 // the reverse lane holds large dual-number state: small workgroups
constexpr int kNs = 6;                 // x y ζ px py pz
constexpr int kNb = kNs + 24;          // and E, B of the four stages of a sub-step
using Db = Dual<kNb>;

template <typename T>
struct St6 { T x, y, zeta, px, py, pz; };
template <typename T>
G3D_HD St6<T> axpy(const St6<T>& a, double h, const St6<T>& b) {
    return St6<T>{a.x + h * b.x, a.y + h * b.y, a.zeta + h * b.zeta, a.px + h * b.px, a.py + h * b.py, a.pz + h * b.pz};
}
G3D_HD double twrap(double x, double L) { return wrap_mod(x, L); }
G3D_HD Db twrap(const Db& x, double L) { return x + (wrap_mod(x.v, L) - x.v); }
using B1 = Dual<1>;                 // one direction: the tangent-linear push
G3D_HD B1 bt1(double v, double d) { B1 r; r.v = v; r.d[0] = d; return r; }
G3D_HD B1 twrap(const B1& x, double L) { return x + (wrap_mod(x.v, L) - x.v); }

// the gather point of (x, y, ζ) (gather_t on a uniform ζ axis with the first slice at index 0): nodes, TSC weights
// and their x, y derivatives, the slice pair and its ζ weights; ok = false where the point has no slice pair
struct GP3 {
    long k[9];
    double w[9], wdx[9], wdy[9], gz, fz;
    int fl;
    bool ok;
};
G3D_HD void tsc_wd(double x, double d, int n, int idx[3], double w[3], double dw[3]) {
    tsc_periodic(x, d, n, idx, w);
    const double xi = x / d;
    const double f = xi - floor(xi + 0.5);
    dw[0] = -(0.5 - f) / d;
    dw[1] = -2 * f / d;
    dw[2] = (0.5 + f) / d;
}
G3D_HD GP3 gp3(const g3d_bpush_ctx& c, double x, double y, double zeta) {
    GP3 g{};
    const double t = (zeta - c.zeta0) * c.ihz;
    int n = int(floor(t));
    double fz = t - n;
    if (n == -1 && fz > 1 - 1e-9) { n = 0; fz = 0; }
    g.ok = n >= 0 && n <= c.nsl - 2;
    if (!g.ok) return g;
    g.fl = n; g.fz = fz; g.gz = 1 - fz;
    int ix[3], jn[3]; double wx[3], wn[3], dwx[3], dwn[3];
    tsc_wd(x, c.dx, c.nx, ix, wx, dwx);
    tsc_wd(y, c.dy, c.ny, jn, wn, dwn);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            g.k[3 * j + i] = long(jn[j]) * c.nx + ix[i];
            g.w[3 * j + i] = wx[i] * wn[j];
            g.wdx[3 * j + i] = dwx[i] * wn[j];
            g.wdy[3 * j + i] = wx[i] * dwn[j];
        }
    return g;
}
// component comp at the gather point, and (if gx) its x, y, ζ derivatives
G3D_HD double gp3_comp(const g3d_bpush_ctx& c, const GP3& g, int comp, double* gx, double* gy, double* gzeta) {
    if (!g.ok) { if (gx) *gx = *gy = *gzeta = 0; return 0; }
    const long nk = long(c.nx) * c.ny;
    const double* q0 = c.F + (long(comp) * c.nsl + g.fl) * nk;
    const double* q1 = q0 + nk;
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

// the analytic plane wave (PlaneWave::field) on T
template <typename T>
G3D_HD void tlaser(const g3d_bpush_ctx& c, const T& x, double z, const T& t, T E[3], T B[3]) {
    for (int k = 0; k < 3; ++k) { E[k] = cst<T>(0); B[k] = cst<T>(0); }
    if (!c.laser_on) return;
    const T xi = t - c.kx * x - c.kz * z - c.eta_c;
    if (fabs(val(xi)) > 5 * c.tau) return;
    const double tt = c.tau * c.tau;
    const T g = texp(-1.3862943611198906 * xi * xi / tt);
    const T gp = -2.7725887222397811 * xi / tt * g;
    const T ph = c.kL * xi + c.cep, cs = tcos(ph), sn = tsin(ph);
    const T d1 = c.a0 * (gp * cs - c.kL * g * sn), d2 = c.a0 * (gp * sn + c.kL * g * cs);
    if (c.laser_pol == 0) E[1] = -d1;
    else {
        E[0] = -d1 * c.kz; E[2] = d1 * c.kx;
        if (c.laser_pol == 2) E[1] = -d2;
    }
    B[0] = -c.kz * E[1]; B[1] = c.kz * E[0] - c.kx * E[2]; B[2] = c.kx * E[1];
}

// the right-hand side of stage s (rhs of k_bunch_push); fld(s, x, y, ζ, E, B) gathers
template <typename T, class Fld>
G3D_HD bool trhs(const g3d_bpush_ctx& c, int s, const St6<T>& Y, double z, double qm, St6<T>& d, Fld& fld) {
    const T g = tsqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
    if (val(Y.pz) < 0.3 * val(g)) return false;
    const T gm = (1 + Y.px * Y.px + Y.py * Y.py) / (g + Y.pz);
    T E[3], B[3], Ee[3], Be[3];
    fld(s, twrap(Y.x, c.nx * c.dx), twrap(Y.y, c.ny * c.dy), Y.zeta, E, B);
    tlaser(c, Y.x + c.X0, z, z - (Y.zeta - c.zc), Ee, Be);
    for (int k = 0; k < 3; ++k) { E[k] = E[k] + Ee[k]; B[k] = B[k] + Be[k]; }
    const T ip = 1 / Y.pz;
    d.x = Y.px * ip; d.y = Y.py * ip; d.zeta = -gm * ip;
    d.px = qm * (g * E[0] + Y.py * B[2] - Y.pz * B[1]) * ip;
    d.py = qm * (g * E[1] + Y.pz * B[0] - Y.px * B[2]) * ip;
    d.pz = qm * (g * E[2] + Y.px * B[1] - Y.py * B[0]) * ip;
    return true;
}
template <typename T, class Fld>
G3D_HD bool tsub(const g3d_bpush_ctx& c, St6<T>& Y, double z, double ds, double qm, Fld& fld) {
    St6<T> k1, k2, k3, k4;
    const double hd = 0.5 * ds;
    if (!(trhs(c, 0, Y, z, qm, k1, fld) && trhs(c, 1, axpy(Y, hd, k1), z + hd, qm, k2, fld) &&
          trhs(c, 2, axpy(Y, hd, k2), z + hd, qm, k3, fld) && trhs(c, 3, axpy(Y, ds, k3), z + ds, qm, k4, fld)))
        return false;
    const double w = ds / 6;
    Y.x = Y.x + w * (k1.x + 2 * k2.x + 2 * k3.x + k4.x);
    Y.y = Y.y + w * (k1.y + 2 * k2.y + 2 * k3.y + k4.y);
    Y.zeta = Y.zeta + w * (k1.zeta + 2 * k2.zeta + 2 * k3.zeta + k4.zeta);
    Y.px = Y.px + w * (k1.px + 2 * k2.px + 2 * k3.px + k4.px);
    Y.py = Y.py + w * (k1.py + 2 * k2.py + 2 * k3.py + k4.py);
    Y.pz = Y.pz + w * (k1.pz + 2 * k2.pz + 2 * k3.pz + k4.pz);
    return true;
}
// the sub-step length at (Y, z)
G3D_HD double sub_len(const g3d_bpush_ctx& c, const St6<double>& Y, double z) {
    double ds = c.dz_push < c.s1 - z ? c.dz_push : c.s1 - z;
    {
        const double g = sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
        const double lz = 0.5 * c.hmin * Y.pz * (g + Y.pz) / (1 + Y.px * Y.px + Y.py * Y.py);
        if (lz < ds) ds = (c.s1 - z) / ceil((c.s1 - z) / lz);
    }
    if (c.laser_on) {
        const double g = sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
        const double hl = g - Y.px * c.kx - Y.pz * c.kz;
        const double lim = c.push_phase * Y.pz / (c.kL * (hl > 1e-12 ? hl : 1e-12));
        if (lim < ds) ds = (c.s1 - z) / ceil((c.s1 - z) / lim);
    }
    return ds;
}
// how the push of a particle ends: after `steps` sub-steps, by the end of the step (0), a hand-over (2), a slow
// particle (4) or a failed right-hand side (-1)
struct PushEnd { int steps, how; double z; };
template <class Fld>
G3D_HD PushEnd run_primal(const g3d_bpush_ctx& c, St6<double>& Y, double z, double qm, int max_steps, Fld& fld) {
    int k = 0;
    while (z < c.s1 - 1e-12 * fabs(c.s1)) {
        if (k == max_steps) return PushEnd{k, 0, z};
        if (Y.zeta < c.zhand) return PushEnd{k, 2, z};
        if (c.vz_min > 0 && Y.pz < c.vz_min * sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz))
            return PushEnd{k, 4, z};
        const double ds = sub_len(c, Y, z);
        if (!tsub(c, Y, z, ds, qm, fld)) return PushEnd{k, -1, z};
        z += ds;
        ++k;
    }
    return PushEnd{k, 0, z};
}
struct PrimalFld {
    const g3d_bpush_ctx* c;
    G3D_HD void operator()(int, double x, double y, double zeta, double E[3], double B[3]) const {
        const GP3 g = gp3(*c, x, y, zeta);
        for (int k = 0; k < 3; ++k) {
            E[k] = gp3_comp(*c, g, k, nullptr, nullptr, nullptr);
            B[k] = gp3_comp(*c, g, 3 + k, nullptr, nullptr, nullptr);
        }
    }
};

// particle i, forward
G3D_HD void bpush_one(const g3d_bpush_ctx& c, long i) {
    const int alive = c.alive[i];
    if (alive != 1 && alive != 3) return;
    const double z0 = alive == 3 ? c.s1 - c.rem[i] : c.s0;
    St6<double> Y{c.x[i], c.y[i], c.zeta[i], c.px[i], c.py[i], c.pz[i]};
    PrimalFld fld{&c};
    const PushEnd e = run_primal(c, Y, z0, c.qm[i], -1, fld);
    if (e.how == -1) { c.alive[i] = 0; return; }
    c.x[i] = Y.x; c.y[i] = Y.y; c.zeta[i] = Y.zeta; c.px[i] = Y.px; c.py[i] = Y.py; c.pz[i] = Y.pz;
    if (e.how == 2) { c.alive[i] = 2; c.rem[i] = c.s1 - e.z; return; }
    if (e.how == 4) { c.z[i] = e.z; c.alive[i] = 4; return; }
    c.z[i] = c.s1;
    int al = (Y.zeta < 0 || Y.zeta >= c.zmax || Y.y < c.ylo || Y.y > c.yhi) ? 0 : 1;
    if (al == 1 && Y.zeta < c.zhand) { al = 2; c.rem[i] = 0; }
    c.alive[i] = al;
}

// particle i, reverse; add(p, v) adds v to *p (an atomic add on the device)
template <class Add>
G3D_HD void bpush_vjp_one(const g3d_bpush_ctx& c, const g3d_bpush_adj& a, long i, Add&& add) {
    double ab[kNs] = {a.bx[i], a.by[i], a.bzeta[i], a.bpx[i], a.bpy[i], a.bpz[i]};
    auto out = [&] {
        a.gx[i] = ab[0]; a.gy[i] = ab[1]; a.gzeta[i] = ab[2]; a.gpx[i] = ab[3]; a.gpy[i] = ab[4]; a.gpz[i] = ab[5];
    };
    const int alive = c.alive[i];
    if (alive != 1 && alive != 3) { out(); return; }
    const double z0 = alive == 3 ? c.s1 - c.rem[i] : c.s0;
    const double qm = c.qm[i];
    const St6<double> Y0{c.x[i], c.y[i], c.zeta[i], c.px[i], c.py[i], c.pz[i]};
    PrimalFld pf{&c};
    St6<double> Y = Y0;
    const PushEnd e = run_primal(c, Y, z0, qm, -1, pf);
    if (e.how == -1) { out(); return; }           // the failed push keeps the start state
    const long nk = long(c.nx) * c.ny;
    for (int j = e.steps - 1; j >= 0; --j) {
        Y = Y0;
        const PushEnd ej = run_primal(c, Y, z0, qm, j, pf);   // the state at the start of sub-step j
        const double ds = sub_len(c, Y, ej.z);
        St6<Db> Yd{dvar<kNb>(Y.x, 0), dvar<kNb>(Y.y, 1), dvar<kNb>(Y.zeta, 2), dvar<kNb>(Y.px, 3), dvar<kNb>(Y.py, 4),
                   dvar<kNb>(Y.pz, 5)};
        GP3 gps[4];
        auto fld = [&](int s, const Db& x, const Db& y, const Db& zeta, Db E[3], Db B[3]) {
            gps[s] = gp3(c, x.v, y.v, zeta.v);
            for (int comp = 0; comp < 6; ++comp) {
                double gx, gy, gzeta;
                const double v = gp3_comp(c, gps[s], comp, &gx, &gy, &gzeta);
                // the position of a later stage depends on the fields of the earlier ones too: the chain runs
                // over every input, then the stage's own field value is an input of its own
                Db r = dconst<kNb>(v);
                for (int k = 0; k < kNb; ++k) r.d[k] = gx * x.d[k] + gy * y.d[k] + gzeta * zeta.d[k];
                r.d[kNs + 6 * s + comp] += 1;
                if (comp < 3) E[comp] = r; else B[comp - 3] = r;
            }
        };
        tsub(c, Yd, ej.z, ds, qm, fld);
        const Db S = ab[0] * Yd.x + ab[1] * Yd.y + ab[2] * Yd.zeta + ab[3] * Yd.px + ab[4] * Yd.py + ab[5] * Yd.pz;
        for (int k = 0; k < kNs; ++k) ab[k] = S.d[k];
        for (int s = 0; s < 4; ++s) {               // the transpose of the stage gathers
            if (!gps[s].ok) continue;
            for (int comp = 0; comp < 6; ++comp) {
                const double g = S.d[kNs + 6 * s + comp];
                if (g == 0) continue;
                double* p0 = a.gF + (long(comp) * c.nsl + gps[s].fl) * nk;
                for (int t = 0; t < 9; ++t) {
                    add(&p0[gps[s].k[t]], g * gps[s].w[t] * gps[s].gz);
                    add(&p0[nk + gps[s].k[t]], g * gps[s].w[t] * gps[s].fz);
                }
            }
        }
    }
    out();
}

// particle i, tangent-linear: the primal push decides the sub-steps (held fixed); the sub-steps then run again on
// dual numbers of one direction from the tangent of the start state, each stage's gather adding the tangent of the
// fields (dF, the layout of F) at its point; the pushed tangents replace dst ([6][stride]: x y zeta px py pz)
G3D_HD void bpush_tlm_one(const g3d_bpush_ctx& c, const double* dF, double* dst, long stride, long i) {
    const int alive = c.alive[i];
    if (alive != 1 && alive != 3) return;
    const double z0 = alive == 3 ? c.s1 - c.rem[i] : c.s0;
    const double qm = c.qm[i];
    const St6<double> Y0{c.x[i], c.y[i], c.zeta[i], c.px[i], c.py[i], c.pz[i]};
    PrimalFld pf{&c};
    St6<double> Y = Y0;
    const PushEnd e = run_primal(c, Y, z0, qm, -1, pf);
    if (e.how == -1) return;                        // the failed push keeps the start state
    const long nk = long(c.nx) * c.ny;
    St6<B1> Yd{bt1(Y0.x, dst[i]), bt1(Y0.y, dst[stride + i]), bt1(Y0.zeta, dst[2 * stride + i]),
               bt1(Y0.px, dst[3 * stride + i]), bt1(Y0.py, dst[4 * stride + i]), bt1(Y0.pz, dst[5 * stride + i])};
    auto fld = [&](int, const B1& x, const B1& y, const B1& zeta, B1 E[3], B1 B[3]) {
        const GP3 g = gp3(c, x.v, y.v, zeta.v);
        for (int comp = 0; comp < 6; ++comp) {
            double gx, gy, gzeta;
            const double v = gp3_comp(c, g, comp, &gx, &gy, &gzeta);
            double d = gx * x.d[0] + gy * y.d[0] + gzeta * zeta.d[0];
            if (g.ok) {
                const double* q0 = dF + (long(comp) * c.nsl + g.fl) * nk;
                for (int t = 0; t < 9; ++t) d += g.w[t] * (g.gz * q0[g.k[t]] + g.fz * q0[nk + g.k[t]]);
            }
            if (comp < 3) E[comp] = bt1(v, d); else B[comp - 3] = bt1(v, d);
        }
    };
    double z = z0;
    for (int j = 0; j < e.steps; ++j) {
        const St6<double> Yp{Yd.x.v, Yd.y.v, Yd.zeta.v, Yd.px.v, Yd.py.v, Yd.pz.v};
        const double ds = sub_len(c, Yp, z);
        tsub(c, Yd, z, ds, qm, fld);
        z += ds;
    }
    dst[i] = Yd.x.d[0]; dst[stride + i] = Yd.y.d[0]; dst[2 * stride + i] = Yd.zeta.d[0];
    dst[3 * stride + i] = Yd.px.d[0]; dst[4 * stride + i] = Yd.py.d[0]; dst[5 * stride + i] = Yd.pz.d[0];
}
__global__ void __launch_bounds__(kT) k_bpush_tlm(g3d_bpush_ctx c, const double* dF, double* dst, long stride) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < c.np) bpush_tlm_one(c, dF, dst, stride, i);
}

__global__ void __launch_bounds__(kT) k_bpush(g3d_bpush_ctx c) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < c.np) bpush_one(c, i);
}
__global__ void __launch_bounds__(kT) k_bpush_vjp(g3d_bpush_ctx c, g3d_bpush_adj a) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < c.np) bpush_vjp_one(c, a, i, [](double* p, double v) { atomicAdd(p, v); });
}
}  // namespace

void bunch_push_tlm(const g3d_bpush_ctx& c, const double* dF, double* dst, long stride) {
    if (c.np <= 0) return;
    hipLaunchKernelGGL(k_bpush_tlm, dim3(blocks_for(c.np, kT)), dim3(kT), 0, stream(), c, dF, dst, stride);
    launch_check("bunch push tangent");
}
}  // namespace g3d

using namespace g3d;
extern "C" {
int g3d_bunch_push_piece(const g3d_bpush_ctx* c) {
    return g3d::guard([&] {
        if (!c || c->nx < 1 || c->ny < 1 || c->nsl < 2 || c->np < 0)
            throw std::invalid_argument("bunch push: bad context");
        if (c->np == 0) return;
        hipLaunchKernelGGL(k_bpush, dim3(blocks_for(c->np, kT)), dim3(kT), 0, stream(), *c);
        launch_check("bunch push");
    });
}
int g3d_bunch_push_vjp(const g3d_bpush_ctx* c, const g3d_bpush_adj* g) {
    return g3d::guard([&] {
        if (!c || !g || c->nx < 1 || c->ny < 1 || c->nsl < 2 || c->np < 0)
            throw std::invalid_argument("bunch push vjp: bad context");
        if (c->np == 0) return;
        hipLaunchKernelGGL(k_bpush_vjp, dim3(blocks_for(c->np, kT)), dim3(kT), 0, stream(), *c, *g);
        launch_check("bunch push vjp");
    });
}
}
