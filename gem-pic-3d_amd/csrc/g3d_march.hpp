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
    int ny = 0;
    int cm = 0;          // 1: modes column-major (i = ix ny + j, the plasma pass's layout), else i = j nx + ix
};

// (ix, j) of mode i
G3D_HD void mode_xy(const MarchView& c, long i, int& ixx, int& j) {
    if (c.cm) { ixx = int(i / c.ny); j = int(i - long(ixx) * c.ny); }
    else { ixx = int(i % c.nx); j = int(i / c.nx); }
}

// values of one mode at one slice: R_x R_y E_z B_z L_x L_y
struct Mode6 { cx Rx, Ry, Ez, Bz, Lx, Ly; };

// the currents of mode i that the update of slice n reads: j0 and jk of half-cell n-1 (d) and n (u)
struct Cur9 { cx j0x_d, j0x_u, j0y_d, j0y_u, jkx_d, jkx_u, jky_d, jky_u, jkz_u; };
G3D_HD Cur9 load_cur(const MarchView& c, long i, int n) {
    const long od = long(n - c.a) * c.nk + i, ou = od + c.nk;
    return Cur9{c.j0x[od], c.j0x[ou], c.j0y[od], c.j0y[ou], c.jkx[od], c.jkx[ou], c.jky[od], c.jky[ou], c.jkz[ou]};
}

// Slice n of level s1 for mode i.  `old` holds R_x, R_y, E_z, B_z of slice n at level s0 (the explicit part),
// `up` the six values of slice n+1 at level s1 (unused at the front slice n = N-1), J the currents.  Returns
// slice n at s1.  The operand order is that of the reference so that the FP64 results round identically.
// march_mode_k: the same with the mode's wavenumbers and the slice's weights given (a chain over the slices of one
// mode loads the wavenumbers once and each slice's weights ahead of it)
struct SliceW { double wl, wu, h; };
G3D_HD SliceW slice_w(const MarchView& c, int n) {
    return SliceW{c.wlo[n], c.whi[n], n < c.N - 1 ? c.hcg[n + 2] : 0.0};
}
G3D_HD Mode6 march_mode_k(const MarchView& c, double kxv, double kyv, const SliceW& w, int n, const Mode6& old,
                          const Mode6& up, const Cur9& J) {
    const cx a = ik(kxv), b = ik(kyv);
    const double k2 = kxv * kxv + kyv * kyv, D = c.D, wl = w.wl, wu = w.wu;
    const cx Jx = wl * (J.j0x_d + J.jkx_d) + wu * (J.j0x_u + J.jkx_u);
    const cx Jy = wl * (J.j0y_d + J.jky_d) + wu * (J.j0y_u + J.jky_u);
    const cx Rsx = old.Rx + D * (a * old.Ez + b * old.Bz) - D * Jx;
    const cx Rsy = old.Ry + D * (b * old.Ez - a * old.Bz) - D * Jy;
    Mode6 r;
    if (n == c.N - 1) {                                // front of the window: L = E_z = B_z = 0
        r.Rx = Rsx; r.Ry = Rsy; r.Ez = r.Bz = r.Lx = r.Ly = mk(0);
        return r;
    }
    const double h = w.h;
    const double qh = 0.25 * h, hh = 0.5 * h;
    const cx upx = J.jkx_u, upy = J.jky_u, upz = J.jkz_u;
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
G3D_HD Mode6 march_mode(const MarchView& c, long i, int n, const Mode6& old, const Mode6& up, const Cur9& J) {
    int ixx, j;
    mode_xy(c, i, ixx, j);
    return march_mode_k(c, c.kx[ixx], c.ky[j], slice_w(c, n), n, old, up, J);
}

G3D_HD Mode6 march_mode(const MarchView& c, long i, int n, const Mode6& old, const Mode6& up) {
    return march_mode(c, i, n, old, up, load_cur(c, i, n));
}

// The same update in the FP32-mixed form: the explicit part R_s = R + D (F - J) and the trapezoidal update of R
// stay FP64 (the s-integration of R); the implicit solve for E_z, B_z and the L update run in FP32.
G3D_HD Mode6 march_mode_mixed(const MarchView& c, long i, int n, const Mode6& old, const Mode6& up, const Cur9& J) {
    int ixx, j;
    mode_xy(c, i, ixx, j);
    const double kxv = c.kx[ixx], kyv = c.ky[j];
    const cx a = ik(kxv), b = ik(kyv);
    const double D = c.D, wl = c.wlo[n], wu = c.whi[n];
    const cx Jx = wl * (J.j0x_d + J.jkx_d) + wu * (J.j0x_u + J.jkx_u);
    const cx Jy = wl * (J.j0y_d + J.jky_d) + wu * (J.j0y_u + J.jky_u);
    const cx Rsx = old.Rx + D * (a * old.Ez + b * old.Bz) - D * Jx;
    const cx Rsy = old.Ry + D * (b * old.Ez - a * old.Bz) - D * Jy;
    Mode6 r;
    if (n == c.N - 1) {
        r.Rx = Rsx; r.Ry = Rsy; r.Ez = r.Bz = r.Lx = r.Ly = mk(0);
        return r;
    }
    const float h = float(c.hcg[n + 2]), qh = 0.25f * h, hh = 0.5f * h, Df = float(D);
    const cxf af = mkf(0, float(kxv)), bf = mkf(0, float(kyv));
    const float k2 = float(kxv * kxv + kyv * kyv);
    const cxf uRx = to_f(up.Rx), uRy = to_f(up.Ry), uEz = to_f(up.Ez), uBz = to_f(up.Bz), uLx = to_f(up.Lx),
              uLy = to_f(up.Ly);
    const cxf sRx = to_f(Rsx), sRy = to_f(Rsy);
    const cxf Lsx = uLx - qh * (af * uEz - bf * uBz) - hh * to_f(J.jkx_u);
    const cxf Lsy = uLy - qh * (af * uBz + bf * uEz) - hh * to_f(J.jky_u);
    const cxf Cp = 0.5f * (af * (uRy + uLy) - bf * (uRx + uLx));
    const cxf Qp = -0.5f * (af * (uRx - uLx)) + 0.5f * (bf * (uLy - uRy));
    const cxf SB = af * sRy + af * Lsy - bf * sRx - bf * Lsx;
    const cxf SE = -(af * sRx) + af * Lsx + bf * Lsy - bf * sRy;
    const float den = 1 + qh * (Df + qh) * k2;
    const cxf Bzn = (uBz - hh * Cp - qh * SB) / den;
    const cxf Ezn = (uEz - hh * Qp - qh * SE - h * to_f(J.jkz_u)) / den;
    r.Ez = to_d(Ezn); r.Bz = to_d(Bzn);
    r.Rx = Rsx + D * (a * r.Ez + b * r.Bz);
    r.Ry = Rsy + D * (b * r.Ez - a * r.Bz);
    r.Lx = to_d(Lsx - qh * (af * Ezn - bf * Bzn));
    r.Ly = to_d(Lsy - qh * (af * Bzn + bf * Ezn));
    return r;
}

// launches (g3d_march.hip): slices hi-1 ... lo of every mode; one slice n with the optional pass backup
void march_all(const MarchView& v, int lo, int hi);
void march_one(const MarchView& v, int n, cx* bRx, cx* bRy, cx* bEz, cx* bBz, int restore);

}  // namespace g3d
