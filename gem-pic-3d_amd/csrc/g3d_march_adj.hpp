// The transposed per-mode march (the adjoint of march_mode), shared by the adjoint of the whole-window march
// (g3d_adjoint.hip) and the reverse of a plasma pass (g3d_grad.hip).  No counterpart in the reference.
//
// The per-mode march of slice n (march_mode) is complex-linear in its inputs: the level-s0 values of slice n
// (R_x, R_y, E_z, B_z), the level-s1 values of slice n+1 (six), and the currents of half-cells n-1 and n.  Its
// coefficients are real (D, h/4, h/2, h, 1/den, the zeta weights) or the pure imaginary a = i k_x, b = i k_y.  For
// the real inner product <u, v> = Re sum conj(u) v, the adjoint of v = A u is u' = A^H v', so every statement of
// the forward update is transposed in reverse order with conj(a) = -a, conj(b) = -b.
#pragma once
#include "g3d_march.hpp"

namespace g3d {

struct AdjView {
    cx *Rx, *Ry, *Ez, *Bz, *Lx, *Ly;          // in: adjoints of the outputs; out: adjoints of the level-s0 inputs
    cx* up;                                   // out: 6 planes, adjoint of slice hi (hi < N), or null
    cx *j0x, *j0y, *jkx, *jky, *jkz;          // out (+=): adjoints of the currents, planes n - a
    const double *kx, *ky, *hcg, *wlo, *whi;
    double D;
    int a, N, nx;
    long nk;
};

// adjoints of one slice's inputs given the adjoint r of its outputs; o: old (R_x, R_y, E_z, B_z), u: slice n+1,
// jd, ju: currents of half-cells n-1 (d) and n (u) as (j0x, j0y, jkx, jky[, jkz])
struct CurAdj { cx j0x_d, j0x_u, j0y_d, j0y_u, jkx_d, jkx_u, jky_d, jky_u, jkz_u; };

__device__ __forceinline__ void march_mode_adj(const AdjView& c, long i, int n, const Mode6& r, Mode6& o, Mode6& u,
                                               CurAdj& J) {
    const int ixx = int(i % c.nx), j = int(i / c.nx);
    const double kxv = c.kx[ixx], kyv = c.ky[j];
    const cx ac = ik(-kxv), bc = ik(-kyv);                 // conj(a), conj(b)
    const double k2 = kxv * kxv + kyv * kyv, D = c.D, wl = c.wlo[n], wu = c.whi[n];
    cx Rsx = r.Rx, Rsy = r.Ry;
    u = Mode6{};
    J = CurAdj{};
    if (n != c.N - 1) {
        const double h = c.hcg[n + 2];
        const double qh = 0.25 * h, hh = 0.5 * h;
        const double den = 1 + qh * (D + qh) * k2;
        cx Lsx = r.Lx, Lsy = r.Ly;
        // r.Rx = Rsx + D (a Ezn + b Bzn), r.Ry = Rsy + D (b Ezn - a Bzn),
        // r.Lx = Lsx - qh (a Ezn - b Bzn), r.Ly = Lsy - qh (a Bzn + b Ezn), r.Ez = Ezn, r.Bz = Bzn
        const cx Ezb = r.Ez + D * (ac * r.Rx) + D * (bc * r.Ry) - qh * (ac * r.Lx) - qh * (bc * r.Ly);
        const cx Bzb = r.Bz + D * (bc * r.Rx) - D * (ac * r.Ry) + qh * (bc * r.Lx) - qh * (ac * r.Ly);
        // Ezn = (up.Ez - hh Qp - qh SE - h upz) / den,  Bzn = (up.Bz - hh Cp - qh SB) / den
        const cx te = Ezb / den, tb = Bzb / den;
        u.Ez = te; u.Bz = tb;
        const cx Qpb = -hh * te, SEb = -qh * te, Cpb = -hh * tb, SBb = -qh * tb;
        J.jkz_u = -h * te;
        // SE = -a Rsx + a Lsx + b Lsy - b Rsy,  SB = a Rsy + a Lsy - b Rsx - b Lsx
        Rsx = Rsx - ac * SEb - bc * SBb;
        Rsy = Rsy - bc * SEb + ac * SBb;
        Lsx = Lsx + ac * SEb - bc * SBb;
        Lsy = Lsy + bc * SEb + ac * SBb;
        // Qp = -0.5 a (up.Rx - up.Lx) + 0.5 b (up.Ly - up.Ry),  Cp = 0.5 (a (up.Ry + up.Ly) - b (up.Rx + up.Lx))
        const cx qa = 0.5 * (ac * Qpb), qb = 0.5 * (bc * Qpb), ca = 0.5 * (ac * Cpb), cb = 0.5 * (bc * Cpb);
        u.Rx = -qa - cb;
        u.Lx = qa - cb;
        u.Ly = qb + ca;
        u.Ry = -qb + ca;
        // Lsx = up.Lx - qh (a up.Ez - b up.Bz) - hh jkx_u,  Lsy = up.Ly - qh (a up.Bz + b up.Ez) - hh jky_u
        u.Lx = u.Lx + Lsx;
        u.Ly = u.Ly + Lsy;
        u.Ez = u.Ez - qh * (ac * Lsx) - qh * (bc * Lsy);
        u.Bz = u.Bz + qh * (bc * Lsx) - qh * (ac * Lsy);
        J.jkx_u = -hh * Lsx;
        J.jky_u = -hh * Lsy;
    }
    // Rsx = old.Rx + D (a old.Ez + b old.Bz) - D Jx,  Rsy = old.Ry + D (b old.Ez - a old.Bz) - D Jy
    o.Rx = Rsx;
    o.Ry = Rsy;
    o.Ez = D * (ac * Rsx) + D * (bc * Rsy);
    o.Bz = D * (bc * Rsx) - D * (ac * Rsy);
    o.Lx = o.Ly = mk(0);
    const cx Jxb = -D * Rsx, Jyb = -D * Rsy;
    // Jx = wl (j0x_d + jkx_d) + wu (j0x_u + jkx_u), Jy likewise
    J.j0x_d = wl * Jxb; J.jkx_d = wl * Jxb; J.j0x_u = wu * Jxb; J.jkx_u = J.jkx_u + wu * Jxb;
    J.j0y_d = wl * Jyb; J.jky_d = wl * Jyb; J.j0y_u = wu * Jyb; J.jky_u = J.jky_u + wu * Jyb;
}

}  // namespace g3d
