// Adjoint (reverse-mode) kernels of the march, for gradients through it.  No counterpart in the reference.
// Here: the transposed per-mode update (below), the TSC deposit with its transpose (the gather) and the gather's
// position derivatives, and the reverse of the Boris rotation.
//
// The per-mode march of slice n (march_mode) is complex-linear in its inputs: the level-s0 values of slice n
// (R_x, R_y, E_z, B_z), the level-s1 values of slice n+1 (six), and the currents of half-cells n-1 and n.  Its
// coefficients are real (D, h/4, h/2, h, 1/den, the ζ weights) or the pure imaginary a = i k_x, b = i k_y.  For the
// real inner product <u, v> = Re Σ conj(u) v, the adjoint of v = A u is u' = A^H v', so every statement of the
// forward update is transposed in reverse order with conj(a) = -a, conj(b) = -b.  The march over slices runs front
// to rear (n = hi-1 ... lo, slice n reading slice n+1); its adjoint runs rear to front, carrying the adjoint of the
// slice n+1 values from the update of slice n to that of slice n+1.
#include "g3d_boris_rev.hpp"
#include "g3d_kernels.hpp"
#include "g3d_march_adj.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {

// slices lo ... hi-1 of every mode, rear to front: the transpose of k_march_all
__global__ void __launch_bounds__(256) k_march_adj(AdjView c, int lo, int hi) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.nk) return;
    Mode6 carry{};                                  // adjoint of slice n's outputs from slice n-1's use of them
    cx d0x = mk(0), d0y = mk(0), dkx = mk(0), dky = mk(0), dkz = mk(0);   // half-cell n-1 (plane n - a) so far
    for (int n = lo; n < hi; ++n) {
        const long o = long(n - c.a) * c.nk + i;
        const Mode6 r{c.Rx[o] + carry.Rx, c.Ry[o] + carry.Ry, c.Ez[o] + carry.Ez, c.Bz[o] + carry.Bz,
                      c.Lx[o] + carry.Lx, c.Ly[o] + carry.Ly};
        Mode6 old, up;
        CurAdj J;
        march_mode_adj(c, i, n, r, old, up, J);
        c.Rx[o] = old.Rx; c.Ry[o] = old.Ry; c.Ez[o] = old.Ez; c.Bz[o] = old.Bz; c.Lx[o] = old.Lx; c.Ly[o] = old.Ly;
        // plane n - a is half-cell n-1: the "u" currents of slice n-1 and the "d" currents of slice n
        c.j0x[o] = c.j0x[o] + d0x + J.j0x_d;
        c.j0y[o] = c.j0y[o] + d0y + J.j0y_d;
        c.jkx[o] = c.jkx[o] + dkx + J.jkx_d;
        c.jky[o] = c.jky[o] + dky + J.jky_d;
        c.jkz[o] = c.jkz[o] + dkz;
        d0x = J.j0x_u; d0y = J.j0y_u; dkx = J.jkx_u; dky = J.jky_u; dkz = J.jkz_u;
        carry = up;
    }
    const long o = long(hi - c.a) * c.nk + i;        // plane hi - a: half-cell hi-1
    c.j0x[o] = c.j0x[o] + d0x;
    c.j0y[o] = c.j0y[o] + d0y;
    c.jkx[o] = c.jkx[o] + dkx;
    c.jky[o] = c.jky[o] + dky;
    c.jkz[o] = c.jkz[o] + dkz;
    if (c.up && hi < c.N) {
        const long nk = c.nk;
        c.up[i] = carry.Rx; c.up[nk + i] = carry.Ry; c.up[2 * nk + i] = carry.Ez;
        c.up[3 * nk + i] = carry.Bz; c.up[4 * nk + i] = carry.Lx; c.up[5 * nk + i] = carry.Ly;
    }
}

// ---------------------------------------------------------------- TSC deposit and its transpose
__global__ void __launch_bounds__(256) k_tsc_deposit(g3d_tsc_ctx c, const double* q, double* plane) {
    const long p = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= c.np) return;
    int ix[3], jn[3]; double wx[3], wn[3];
    tsc_periodic(c.x[p], c.dx, c.nx, ix, wx);
    tsc_periodic(c.y[p], c.dy, c.ny, jn, wn);
    for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) atomicAdd(&plane[long(jn[v]) * c.nx + ix[u]], wx[u] * wn[v] * q[p]);
}
// d/dx of the TSC weights of x (index coordinate xi = x/d, f = xi - round(xi)): the derivative of
// 1/2 (1/2 - f)^2, 3/4 - f^2, 1/2 (1/2 + f)^2, times 1/d
__device__ __forceinline__ void tsc_dweights(double x, double d, double dw[3]) {
    const double xi = x / d;
    const double f = xi - floor(xi + 0.5);
    dw[0] = -(0.5 - f) / d;
    dw[1] = -2 * f / d;
    dw[2] = (0.5 + f) / d;
}
__global__ void __launch_bounds__(256) k_tsc_gather(g3d_tsc_ctx c, const double* plane, double* out, double* gx,
                                                    double* gy) {
    const long p = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= c.np) return;
    int ix[3], jn[3]; double wx[3], wn[3], dx[3], dy[3];
    tsc_periodic(c.x[p], c.dx, c.nx, ix, wx);
    tsc_periodic(c.y[p], c.dy, c.ny, jn, wn);
    double s = 0, sx = 0, sy = 0;
    const bool grad = gx && gy;
    if (grad) { tsc_dweights(c.x[p], c.dx, dx); tsc_dweights(c.y[p], c.dy, dy); }
    for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) {
            const double P = plane[long(jn[v]) * c.nx + ix[u]];
            s += wx[u] * wn[v] * P;
            if (grad) { sx += dx[u] * wn[v] * P; sy += wx[u] * dy[v] * P; }
        }
    out[p] = s;
    if (grad) { gx[p] = sx; gy[p] = sy; }
}

// ---------------------------------------------------------------- Boris rotation and its reverse
__global__ void __launch_bounds__(256) k_boris(g3d_boris_args a) {
    const long p = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    const double E[3] = {a.Ex[p], a.Ey[p], a.Ez[p]}, B[3] = {a.Bx[p], a.By[p], a.Bz[p]};
    double ux = a.ux[p], uy = a.uy[p], uz = a.uz[p];
    boris(ux, uy, uz, E, B, a.qm[p], a.dt[p]);
    a.ux[p] = ux; a.uy[p] = uy; a.uz[p] = uz;
}
// reverse mode of boris() (g3d_common.hpp): boris_rev (g3d_boris_rev.hpp)
__global__ void __launch_bounds__(256) k_boris_vjp(g3d_boris_args a, g3d_boris_adj g) {
    const long p = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= a.np) return;
    const double E[3] = {a.Ex[p], a.Ey[p], a.Ez[p]}, B[3] = {a.Bx[p], a.By[p], a.Bz[p]};
    const double u[3] = {a.ux[p], a.uy[p], a.uz[p]};
    const double ub[3] = {g.ubx[p], g.uby[p], g.ubz[p]};
    double ab[3], Eb[3] = {0, 0, 0}, Bb[3] = {0, 0, 0}, qmb = 0, dtb = 0;
    boris_rev(u, E, B, a.qm[p], a.dt[p], ub, ab, Eb, Bb, qmb, dtb);
    g.gux[p] = ab[0]; g.guy[p] = ab[1]; g.guz[p] = ab[2];
    g.gEx[p] = Eb[0]; g.gEy[p] = Eb[1]; g.gEz[p] = Eb[2];
    g.gBx[p] = Bb[0]; g.gBy[p] = Bb[1]; g.gBz[p] = Bb[2];
    g.gqm[p] = qmb;
    g.gdt[p] = dtb;
}

}  // namespace
}  // namespace g3d

extern "C" {
int g3d_tsc_deposit(const g3d_tsc_ctx* c, const double* q, double* plane) {
    return g3d::guard([&] {
        if (!c || !c->x || !c->y || !q || !plane || c->nx < 1 || c->ny < 1 || c->np < 0)
            throw std::invalid_argument("tsc deposit: bad context");
        if (c->np == 0) return;
        hipLaunchKernelGGL(g3d::k_tsc_deposit, dim3(g3d::blocks_for(c->np, 256)), dim3(256), 0, g3d::stream(), *c, q,
                           plane);
        g3d::launch_check("tsc deposit");
    });
}
int g3d_tsc_gather(const g3d_tsc_ctx* c, const double* plane, double* out, double* gx, double* gy) {
    return g3d::guard([&] {
        if (!c || !c->x || !c->y || !plane || !out || c->nx < 1 || c->ny < 1 || c->np < 0 || (!gx) != (!gy))
            throw std::invalid_argument("tsc gather: bad context");
        if (c->np == 0) return;
        hipLaunchKernelGGL(g3d::k_tsc_gather, dim3(g3d::blocks_for(c->np, 256)), dim3(256), 0, g3d::stream(), *c, plane,
                           out, gx, gy);
        g3d::launch_check("tsc gather");
    });
}
int g3d_boris(const g3d_boris_args* a) {
    return g3d::guard([&] {
        if (!a || a->np < 0) throw std::invalid_argument("boris: bad arguments");
        if (a->np == 0) return;
        hipLaunchKernelGGL(g3d::k_boris, dim3(g3d::blocks_for(a->np, 256)), dim3(256), 0, g3d::stream(), *a);
        g3d::launch_check("boris");
    });
}
int g3d_boris_vjp(const g3d_boris_args* a, const g3d_boris_adj* g) {
    return g3d::guard([&] {
        if (!a || !g || a->np < 0) throw std::invalid_argument("boris vjp: bad arguments");
        if (a->np == 0) return;
        hipLaunchKernelGGL(g3d::k_boris_vjp, dim3(g3d::blocks_for(a->np, 256)), dim3(256), 0, g3d::stream(), *a, *g);
        g3d::launch_check("boris vjp");
    });
}
int g3d_march_adjoint(const g3d_march_adj_ctx* m, int lo, int hi) {
    return g3d::guard([&] {
        if (!m || !m->Rx || !m->Ry || !m->Ez || !m->Bz || !m->Lx || !m->Ly || !m->j0x || !m->j0y || !m->jkx ||
            !m->jky || !m->jkz || !m->kx || !m->ky || !m->hcg || !m->wlo || !m->whi)
            throw std::invalid_argument("march adjoint: null pointer in the context");
        if (m->nx < 1 || m->ny < 1 || m->nk != long(m->nx) * m->ny)
            throw std::invalid_argument("march adjoint: nk must be nx * ny");
        if (lo < m->a || hi > m->N || lo >= hi)
            throw std::invalid_argument("march adjoint: need a <= lo < hi <= N");
        g3d::AdjView v;
        v.Rx = m->Rx; v.Ry = m->Ry; v.Ez = m->Ez; v.Bz = m->Bz; v.Lx = m->Lx; v.Ly = m->Ly; v.up = m->up;
        v.j0x = m->j0x; v.j0y = m->j0y; v.jkx = m->jkx; v.jky = m->jky; v.jkz = m->jkz;
        v.kx = m->kx; v.ky = m->ky; v.hcg = m->hcg; v.wlo = m->wlo; v.whi = m->whi;
        v.D = m->D; v.a = m->a; v.N = m->N; v.nx = m->nx; v.nk = m->nk;
        hipLaunchKernelGGL(g3d::k_march_adj, dim3(g3d::blocks_for(v.nk, 256)), dim3(256), 0, g3d::stream(), v, lo,
                           hi);
        g3d::launch_check("march adjoint");
    });
}
}
