// Per-mode march of the Fourier-y spectra (src/sim_kernels.hpp:169-210) as a device function shared by the
// whole-window march, the one-slice march and the fused plasma mode kernel.
#pragma once
#include "g3d_common.hpp"

namespace g3d {

struct MarchView {
    cx *Rx, *Ry, *Ez, *Bz, *Lx, *Ly;
    const cx *j0x, *j0y, *jkx, *jky, *jkz;
    const double *kx, *ky, *hcg, *wlo, *whi;
    double D;
    int a, N, nx;
    long nk;
};

// values of one mode at one slice: R_x R_y E_z B_z L_x L_y
struct Mode6 { cx Rx, Ry, Ez, Bz, Lx, Ly; };

// Slice n of level s1 for mode i.  `old` holds R_x, R_y, E_z, B_z of slice n at level s0 (the explicit part),
// `up` the six values of slice n+1 at level s1 (unused at the front slice n = N-1).  Returns slice n at s1.
// The operand order is that of the reference so that the FP64 results round identically.
G3D_HD Mode6 march_mode(const MarchView& c, long i, int n, const Mode6& old, const Mode6& up) {
    const long o = long(n - c.a) * c.nk + i;
    const int ixx = int(i % c.nx), j = int(i / c.nx);
    const double kxv = c.kx[ixx], kyv = c.ky[j];
    const cx a = ik(kxv), b = ik(kyv);
    const double k2 = kxv * kxv + kyv * kyv, D = c.D, wl = c.wlo[n], wu = c.whi[n];
    const long od = o, ou = o + c.nk;                  // half-cells n-1 and n
    const cx Jx = wl * (c.j0x[od] + c.jkx[od]) + wu * (c.j0x[ou] + c.jkx[ou]);
    const cx Jy = wl * (c.j0y[od] + c.jky[od]) + wu * (c.j0y[ou] + c.jky[ou]);
    const cx Rsx = old.Rx + D * (a * old.Ez + b * old.Bz) - D * Jx;
    const cx Rsy = old.Ry + D * (b * old.Ez - a * old.Bz) - D * Jy;
    Mode6 r;
    if (n == c.N - 1) {                                // front of the window: L = E_z = B_z = 0
        r.Rx = Rsx; r.Ry = Rsy; r.Ez = r.Bz = r.Lx = r.Ly = mk(0);
        return r;
    }
    const double h = c.hcg[n + 2];
    const double qh = 0.25 * h, hh = 0.5 * h;
    const cx upx = c.jkx[ou], upy = c.jky[ou], upz = c.jkz[ou];
    const cx Lsx = up.Lx - qh * (a * up.Ez - b * up.Bz) - hh * upx;
    const cx Lsy = up.Ly - qh * (a * up.Bz + b * up.Ez) - hh * upy;
    const cx Cp = 0.5 * (a * (up.Ry + up.Ly) - b * (up.Rx + up.Lx));
    const cx Qp = -0.5 * (a * (up.Rx - up.Lx)) + 0.5 * (b * (up.Ly - up.Ry));
    const cx SB = a * Rsy + a * Lsy - b * Rsx - b * Lsx;
    const cx SE = -(a * Rsx) + a * Lsx + b * Lsy - b * Rsy;
    const double den = 1 + qh * (D + qh) * k2;
    const cx Bzn = (up.Bz - hh * Cp - qh * SB) / den;
    const cx Ezn = (up.Ez - hh * Qp - qh * SE - h * upz) / den;
    r.Ez = Ezn; r.Bz = Bzn;
    r.Rx = Rsx + D * (a * Ezn + b * Bzn);
    r.Ry = Rsy + D * (b * Ezn - a * Bzn);
    r.Lx = Lsx - qh * (a * Ezn - b * Bzn);
    r.Ly = Lsy - qh * (a * Bzn + b * Ezn);
    return r;
}

// launches (g3d_march.hip): slices hi-1 ... lo of every mode; one slice n with the optional pass backup
void march_all(const MarchView& v, int lo, int hi);
void march_one(const MarchView& v, int n, cx* bRx, cx* bRy, cx* bEz, cx* bBz, int restore);

}  // namespace g3d
