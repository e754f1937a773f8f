// Shared device/host helpers of the gfx950 march: complex arithmetic in the reference's evaluation order, the
// TSC weights, the axis lookup, the field gather, the analytic plane-wave laser and the Boris rotation.
//
// Every formula here is evaluated in the same operand order as the reference's device code, and the library is
// compiled with -ffp-contract=off, so that the FP64 kernels round like the reference's g++ -O3 x86-64 build
// (which emits no FMA).  That is what makes cell indices, weights and trapping decisions bit-identical.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../include/g3d.h"

#define G3D_HD __host__ __device__ __forceinline__

namespace g3d {

using cx = g3d_cpx;
G3D_HD cx mk(double r, double i = 0) { return cx{r, i}; }
G3D_HD cx operator+(cx a, cx b) { return cx{a.re + b.re, a.im + b.im}; }
G3D_HD cx operator-(cx a, cx b) { return cx{a.re - b.re, a.im - b.im}; }
G3D_HD cx operator-(cx a) { return cx{-a.re, -a.im}; }
// (a.re b.re - a.im b.im) + (a.re b.im + a.im b.re) i, as src/portable.hpp:30 and std::complex
G3D_HD cx operator*(cx a, cx b) { return cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
G3D_HD cx operator*(double s, cx a) { return cx{s * a.re, s * a.im}; }
G3D_HD cx operator*(cx a, double s) { return cx{s * a.re, s * a.im}; }
G3D_HD cx operator/(cx a, double s) { return cx{a.re / s, a.im / s}; }
G3D_HD cx conjg(cx a) { return cx{a.re, -a.im}; }
G3D_HD cx ik(double k) { return cx{0, k}; }                       // i k
G3D_HD cx times_minus_i(cx z) { return cx{z.im, -z.re}; }         // -i z

// single-precision complex of the FP32-mixed form, with the same operand order
struct cxf { float re, im; };
G3D_HD cxf mkf(float r, float i = 0) { return cxf{r, i}; }
G3D_HD cxf operator+(cxf a, cxf b) { return cxf{a.re + b.re, a.im + b.im}; }
G3D_HD cxf operator-(cxf a, cxf b) { return cxf{a.re - b.re, a.im - b.im}; }
G3D_HD cxf operator-(cxf a) { return cxf{-a.re, -a.im}; }
G3D_HD cxf operator*(cxf a, cxf b) { return cxf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
G3D_HD cxf operator*(float s, cxf a) { return cxf{s * a.re, s * a.im}; }
G3D_HD cxf operator*(cxf a, float s) { return cxf{s * a.re, s * a.im}; }
G3D_HD cxf operator/(cxf a, float s) { return cxf{a.re / s, a.im / s}; }
G3D_HD cxf conjg(cxf a) { return cxf{a.re, -a.im}; }
G3D_HD cxf to_f(cx a) { return cxf{float(a.re), float(a.im)}; }
G3D_HD cx to_d(cxf a) { return cx{a.re, a.im}; }
// a twiddle of the double table in the precision of the transform
template <class C> G3D_HD C twiddle_as(cx w);
template <> G3D_HD cx twiddle_as<cx>(cx w) { return w; }
template <> G3D_HD cxf twiddle_as<cxf>(cx w) { return to_f(w); }

// ------------------------------------------------------------------ grids
// Quadratic spline (TSC) weights about the nearest integer point of xi (src/sim_kernels.hpp:21-26).
G3D_HD void tsc_w(double xi, int idx[3], double w[3]) {
    const int i0 = int(floor(xi + 0.5));
    const double f = xi - i0;
    const double a = 0.5 - f, b = 0.5 + f;
    w[0] = 0.5 * a * a;
    w[1] = 0.75 - f * f;
    w[2] = 0.5 * b * b;
    idx[0] = i0 - 1; idx[1] = i0; idx[2] = i0 + 1;
}
// i mod n in [0, n) for n > 0, as ((i % n) + n) % n: a compare and an add when i lies in [-n, 2n) (every stencil
// node of a wrapped coordinate), the integer division only outside
G3D_HD int wrap_index(int i, int n) {
    if (i >= 0 && i < n) return i;
    if (i < 0 && i >= -n) return i + n;
    if (i >= n && i < 2 * n) return i - n;
    const int m = i % n;
    return m < 0 ? m + n : m;
}
// periodic uniform axis with nodes at i*d (src/sim_kernels.hpp:28-31)
G3D_HD void tsc_periodic(double x, double d, int n, int idx[3], double w[3]) {
    tsc_w(x / d, idx, w);
    for (int k = 0; k < 3; ++k) idx[k] = wrap_index(idx[k], n);
}

// Node table X[0..n-1] with the direct path for uniform grids (src/grid.hpp:35-46): X[i] <= x < X[i+1],
// end cells continued outside.
struct Axis {
    const double* X;
    int n, uniform;
    double x0, ih;
    // search hint (0 <= hlo < hhi <= n-1, or hhi = 0: none): when X[hlo] <= x < X[hhi] the bisection starts from
    // [hlo, hhi]; the cell with X[i] <= x < X[i+1] is unique, so the result is the same
    int hlo = 0, hhi = 0;
    G3D_HD void locate(double x, int& i, double& f) const {
        if (uniform) {
            const double t = (x - x0) * ih;
            i = int(floor(t)); f = t - i;
            return;
        }
        if (x < X[0]) {
            const double t = (x - X[0]) / (X[1] - X[0]);
            i = int(floor(t)); f = t - i;
            return;
        }
        if (x >= X[n - 1]) {
            const double t = (x - X[n - 1]) / (X[n - 1] - X[n - 2]);
            const int k = int(floor(t));
            i = n - 1 + k; f = t - k;
            return;
        }
        int lo = 0, hi = n - 1;
        if (hhi > hlo && X[hlo] <= x && x < X[hhi]) { lo = hlo; hi = hhi; }
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (X[mid] <= x) lo = mid; else hi = mid;
        }
        i = lo; f = (x - X[lo]) / (X[lo + 1] - X[lo]);
    }
};
// free-hand y: nodes at the integers of the index coordinate, half-nodes at the half integers
// (src/sim_kernels.hpp:34-40)
G3D_HD void tsc_freehand(const Axis& ax, double y, int jn[3], double wn[3], int jh[3], double wh[3]) {
    int i; double f;
    ax.locate(y, i, f);
    const double xi = i + f;
    tsc_w(xi, jn, wn);
    tsc_w(xi - 0.5, jh, wh);
}
G3D_HD bool half_comp(int comp) { return comp == 1 || comp == 3 || comp == 5; }   // Ey Bx Bz (sim_internal.hpp:21)
// fmod(fmod(x, L) + L, L) of the reference (src/sim_kernels.hpp:80), exactly, without the library fmod for
// |x| < L: then fmod(x, L) = x, and t = x + L lies in [0, 2L], where fmod(t, L) is t, t - L or t - 2L, each
// difference exact (Sterbenz).
G3D_HD double wrap_mod(double x, double L) {
    const double r = (x > -L && x < L) ? x : fmod(x, L);
    const double t = r + L;
    return t < L ? t : (t < 2 * L ? t - L : t - 2 * L);
}

// ------------------------------------------------------------------ analytic plane wave (src/sim.hpp:191-210)
struct PlaneWave {
    int on, pol;
    double a0, kL, tau, cep, eta_c, kx, kz;
    G3D_HD void field(double x, double z, double t, double E[3], double B[3]) const {
        E[0] = E[1] = E[2] = B[0] = B[1] = B[2] = 0;
        if (!on) return;
        const double xi = t - kx * x - kz * z - eta_c;
        if (fabs(xi) > 5 * tau) return;
        const double tt = tau * tau;
        const double g = exp(-1.3862943611198906 * xi * xi / tt);
        const double gp = -2.7725887222397811 * xi / tt * g;
        const double ph = kL * xi + cep, c = cos(ph), s = sin(ph);
        const double d1 = a0 * (gp * c - kL * g * s), d2 = a0 * (gp * s + kL * g * c);
        if (pol == 0) E[1] = -d1;
        else {
            E[0] = -d1 * kz; E[2] = d1 * kx;
            if (pol == 2) E[1] = -d2;
        }
        B[0] = -kz * E[1]; B[1] = kz * E[0] - kx * E[2]; B[2] = kx * E[1];
    }
};

// ------------------------------------------------------------------ fields of the slices
// Everything the gather needs (the field part of the reference's PushCtx, src/sim_kernels.hpp:10-17).
struct FieldView {
    const float* F;
    int a, ns, nx, ny, yf, self;
    long nk;
    double dx, dy, Lx, Ly;
    Axis zax, yax;
};

// E, B at (x, y, zeta): TSC in x and y, linear in zeta between slices (src/sim_kernels.hpp:44-72).  The sums run
// over y then x in the reference's order.
// YF: 1 Fourier y, 0 free-hand y, -1 decided at run time from c.yf
template <int YF>
G3D_HD void gather_t(const FieldView& c, double x, double y, double zeta, double E[3], double B[3]) {
    E[0] = E[1] = E[2] = B[0] = B[1] = B[2] = 0;
    if (!c.self) return;
    int n; double fz;
    c.zax.locate(zeta, n, fz);
    if (n == c.a - 1 && fz > 1 - 1e-9) { n = c.a; fz = 0; }
    const int fl = n - c.a;
    if (fl < 0 || fl > c.ns) return;
    int ix[3], jn[3], jh[3]; double wx[3], wn[3], wh[3];
    tsc_periodic(x, c.dx, c.nx, ix, wx);
    const double gz = 1 - fz;
    if (YF == 1 || (YF < 0 && c.yf)) {
        // Fourier y: all components on the periodic nodes; the nine offsets and weights are shared by the six
        // components and every array stays in registers
        tsc_periodic(y, c.dy, c.ny, jn, wn);
        long k[9]; double w[9];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) { k[3 * j + i] = long(jn[j]) * c.nx + ix[i]; w[3 * j + i] = wx[i] * wn[j]; }
        const long cs = long(c.ns + 2) * c.nk;
        const float* p0 = c.F + long(fl) * c.nk;
        double v[6];
#pragma unroll
        for (int comp = 0; comp < 6; ++comp) {
            const float* q0 = p0 + comp * cs;
            const float* q1 = q0 + c.nk;
            double s = 0;
#pragma unroll
            for (int t = 0; t < 9; ++t) s += w[t] * (gz * double(q0[k[t]]) + fz * double(q1[k[t]]));
            v[comp] = s;
        }
        E[0] = v[0]; E[1] = v[1]; E[2] = v[2]; B[0] = v[3]; B[1] = v[4]; B[2] = v[5];
        return;
    }
    if (YF == 1) return;
    tsc_freehand(c.yax, y, jn, wn, jh, wh);
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        const float* p0 = c.F + (size_t(comp) * (c.ns + 2) + fl) * c.nk;
        const float* p1 = p0 + c.nk;
        const bool hf = half_comp(comp);
        double v = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int jj = hf ? jh[j] : jn[j];
            const double wy = hf ? wh[j] : wn[j];
            if (jj < 0 || jj >= c.ny) continue;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const long k = long(jj) * c.nx + ix[i];
                v += wx[i] * wy * (gz * double(p0[k]) + fz * double(p1[k]));
            }
        }
        if (comp < 3) E[comp] = v; else B[comp - 3] = v;
    }
}
G3D_HD void gather(const FieldView& c, double x, double y, double zeta, double E[3], double B[3]) {
    gather_t<-1>(c, x, y, zeta, E, B);
}

// relativistic Boris step of the momentum u over dt (src/sim_stream.hpp:8-18)
G3D_HD void boris(double& ux, double& uy, double& uz, const double E[3], const double B[3], double qm, double dt) {
    const double qh = 0.5 * qm * dt;
    const double ax = ux + qh * E[0], ay = uy + qh * E[1], az = uz + qh * E[2];
    const double gm = sqrt(1.0 + ax * ax + ay * ay + az * az);
    const double tx = qh * B[0] / gm, ty = qh * B[1] / gm, tz = qh * B[2] / gm;
    const double t2 = tx * tx + ty * ty + tz * tz;
    const double px = ax + (ay * tz - az * ty), py = ay + (az * tx - ax * tz), pz = az + (ax * ty - ay * tx);
    const double sf = 2.0 / (1.0 + t2);
    const double bx = ax + sf * (py * tz - pz * ty), by = ay + sf * (pz * tx - px * tz), bz = az + sf * (px * ty - py * tx);
    ux = bx + qh * E[0]; uy = by + qh * E[1]; uz = bz + qh * E[2];
}

// ------------------------------------------------------------------ counter-based random numbers
// SplitMix64 finaliser and the key chain of the ionization draws (src/ionization.hpp:37-46)
G3D_HD uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
G3D_HD double uniform01(uint64_t key) { return double(mix64(key) >> 11) * (1.0 / 9007199254740992.0); }
G3D_HD uint64_t key_step(uint64_t seed, long step) { return mix64(seed * 0x9E3779B97F4A7C15ULL + uint64_t(step)); }
G3D_HD uint64_t key_slice(uint64_t ks, int n, int pass) { return mix64(ks ^ ((uint64_t(n) << 8) | uint64_t(pass))); }
G3D_HD uint64_t key_particle(uint64_t kl, uint64_t id) { return mix64(kl ^ mix64(id)); }

// ADK level in code units (src/ionization.hpp:25-30)
struct AdkLevel { double logpre, power, expo, Epeak, Emin; };
struct IonDev { double mass; int mobile, ionizing, Zmax, off; double Emin; };
G3D_HD double adk_rate(const AdkLevel& L, double E) {
    if (!(E >= L.Emin) || !(E > 0)) return 0.0;
    const double Ec = E < L.Epeak ? E : L.Epeak;
    return exp(L.logpre - L.power * log(Ec) + L.expo / Ec);
}

// species codes (src/sim.hpp:29-34)
enum { SP_BUNCH = 0, SP_TRAPPED = 1, SP_PLASMA = 2, SP_ION = 3, SP_FREED = 50, SP_INJECTED = 100 };
G3D_HD bool is_bunch(int sp) { return sp <= 0; }
G3D_HD bool is_ion(int sp) { return sp >= SP_ION && sp < SP_FREED; }
G3D_HD bool is_freed(int sp) { return sp >= SP_FREED && sp < SP_INJECTED; }
G3D_HD int trapped_species(int sp) { return is_freed(sp) ? sp - SP_FREED + SP_INJECTED : sp >= SP_INJECTED ? sp : SP_TRAPPED; }

}  // namespace g3d
