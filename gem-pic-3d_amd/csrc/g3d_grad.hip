// The reverse of the field side of a march step (no counterpart in the reference): the transposes of the real
// fields of a pass, of the per-mode bookkeeping and march of a plasma pass, of the sponge with the real fields of a
// step, and of the beam currents.  Each forward operation is linear in the spectra for fixed particles, so each
// reverse applies the transpose for the real inner product <u, v> = Re sum conj(u) v: a complex-linear map with
// coefficient c has the adjoint conj(c); z -> conj(z) maps an adjoint l to conj(l); the forward transform
// sum_r x e^{-ikr} has the unnormalised inverse as its adjoint.  One lane per point of a plane, as the forward
// kernels (g3d_fields.hip, g3d_pass.hpp).  Fourier y, the reference's mode layout i = j nx + ix.
#include "g3d_grad.hpp"
#include "g3d_march_adj.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 256;

__device__ __forceinline__ cx negi(cx z) { return mk(z.im, -z.re); }       // -i z
__device__ __forceinline__ cx posi(cx z) { return mk(-z.im, z.re); }       // i z
__device__ __forceinline__ long mirror_of(long i, int nx, int ny) {
    const int ixx = int(i % nx), j = int(i / nx);
    return long((ny - j) % ny) * nx + (nx - ixx) % nx;
}

// E_x = (R_x + L_x)/2, E_y = (R_y + L_y)/2, B_x = (L_y - R_y)/2, B_y = (R_x - L_x)/2, E_z, B_z from the inverse
// transforms of R_x + i R_y, L_x + i L_y, E_z + i B_z: the transpose
__global__ void k_toF_adj(const double* lF, long fpl, long nk, int count, cx* lT, long ts) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(count) * nk) return;
    const double e0 = lF[idx], e1 = lF[fpl + idx], e2 = lF[2 * fpl + idx], e3 = lF[3 * fpl + idx],
                 e4 = lF[4 * fpl + idx], e5 = lF[5 * fpl + idx];
    lT[idx] = mk(0.5 * (e0 + e4), 0.5 * (e1 - e3));
    lT[ts + idx] = mk(0.5 * (e0 - e4), 0.5 * (e1 + e3));
    lT[2 * ts + idx] = mk(e2, e5);
}

// One pass of half-cell n reversed (mode_compute of g3d_pass.hpp, FP64): the outputs' adjoints are read from the
// locations the pass wrote (and those cleared), then the inputs' adjoints are added to the locations it read.
// Forward, for a pass that is not the front slice:
//   sp = first ? pin : sin;  si0 = pin (first)
//   first, n >= 1:   q = jk(n-1) + js0(n-1);  pin = js0(n-1)
//   jk(n) = jk(n) + u - sp;  sin = u
//   n >= 1, extrap:  d0 = u - js0(n);  t = js0(n-1) + d0 + wex (d0 - dl);  q = q + t - pin;  pin = t
//   jk(n-1) = q (when set);  slice n = march(slice n at s0, slice n+1, j0, jk);  T = packed slice n
//   last:  R(n+1) -= qcorr (u - si0) (x, y);  dl = u - js0(n);  js0(n) = u
// The front slice (top): q = jk(N-2) + js0(N-2) - pin;  pin = js0(N-2);  jk(N-2) = q;  slice N-1 = march(...)
__global__ void __launch_bounds__(kT) k_pass_adj(PassAdj c) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nk = c.nk;
    if (i >= nk) return;
    AdjView av{};
    av.kx = c.kx; av.ky = c.ky; av.hcg = c.hcg; av.wlo = c.wlo; av.whi = c.whi; av.D = c.D; av.a = c.a; av.N = c.N;
    av.nx = c.nx; av.nk = nk;
    const int n = c.n;
    const long o = long(n - c.a) * nk + i, on = o + nk;
    const long hq = long(n) * nk + i, hp = hq + nk;
    cx* ljk[3] = {c.ljkx, c.ljky, c.ljkz};
    cx* ljs0[3] = {c.ljs0x, c.ljs0y, c.ljs0z};
    const cx t0 = c.sc * c.lT[i], t1 = c.sc * c.lT[nk + i], t2 = c.sc * c.lT[2 * nk + i];
    Mode6 r{};
    if (c.last || !c.bk) {                          // slice n written: its adjoint is consumed here
        r = Mode6{c.lRx[o], c.lRy[o], c.lEz[o], c.lBz[o], c.lLx[o], c.lLy[o]};
        c.lRx[o] = c.lRy[o] = c.lEz[o] = c.lBz[o] = c.lLx[o] = c.lLy[o] = mk(0);
    }
    r.Rx = r.Rx + t0; r.Ry = r.Ry + negi(t0);
    r.Lx = r.Lx + t1; r.Ly = r.Ly + negi(t1);
    r.Ez = r.Ez + t2; r.Bz = r.Bz + negi(t2);
    Mode6 lo, lu;
    CurAdj J;
    march_mode_adj(av, i, n, r, lo, lu, J);
    c.lRx[o] = c.lRx[o] + lo.Rx; c.lRy[o] = c.lRy[o] + lo.Ry; c.lEz[o] = c.lEz[o] + lo.Ez; c.lBz[o] = c.lBz[o] + lo.Bz;
    c.lj0x[o] = c.lj0x[o] + J.j0x_d; c.lj0y[o] = c.lj0y[o] + J.j0y_d;
    c.lj0x[on] = c.lj0x[on] + J.j0x_u; c.lj0y[on] = c.lj0y[on] + J.j0y_u;
    if (c.top) {
        const cx lq[3] = {ljk[0][hq] + J.jkx_d, ljk[1][hq] + J.jky_d, ljk[2][hq]};
        const cx lpp[3] = {J.jkx_u, J.jky_u, J.jkz_u};
        for (int k = 0; k < 3; ++k) {
            const cx lpin = c.lpin[k * nk + i];
            ljk[k][hq] = lq[k];                     // q = jk + js0 - pin
            ljs0[k][hq] = ljs0[k][hq] + lq[k] + lpin;
            c.lpin[k * nk + i] = mk(0) - lq[k];
            ljk[k][hp] = ljk[k][hp] + lpp[k];
        }
        return;
    }
    // R(n+1) = up - qcorr (u - si0) on the last pass: the adjoint of the corrected values, read before the march's
    // adjoint of up is added (up carries it unchanged)
    const cx ronx = c.last ? c.lRx[on] : mk(0), rony = c.last ? c.lRy[on] : mk(0);
    if (n < c.N - 1) {
        c.lRx[on] = c.lRx[on] + lu.Rx; c.lRy[on] = c.lRy[on] + lu.Ry; c.lEz[on] = c.lEz[on] + lu.Ez;
        c.lBz[on] = c.lBz[on] + lu.Bz; c.lLx[on] = c.lLx[on] + lu.Lx; c.lLy[on] = c.lLy[on] + lu.Ly;
    }
    const bool ext = n >= 1 && c.extrap, fb = c.first && n >= 1, pin_set = fb || ext;
    cx lq[3], lpp[3], lu_[3], lpin[3], lsi0o[3], ljs0q[3], ljs0p[3], ldl0[3], lpin0[3], lsin0[3];
    for (int k = 0; k < 3; ++k) {
        const long kk = k * nk + i;
        lpp[k] = ljk[k][hp]; ljk[k][hp] = mk(0);
        lu_[k] = c.lsin[kk]; c.lsin[kk] = mk(0);                   // sin = u
        if (pin_set) { lq[k] = ljk[k][hq]; ljk[k][hq] = mk(0); lpin[k] = c.lpin[kk]; c.lpin[kk] = mk(0); }
        else { lq[k] = mk(0); lpin[k] = mk(0); }
        if (c.first) { lsi0o[k] = c.lsi0[kk]; c.lsi0[kk] = mk(0); }
        else lsi0o[k] = mk(0);
        ljs0q[k] = ljs0p[k] = ldl0[k] = lpin0[k] = lsin0[k] = mk(0);
    }
    lq[0] = lq[0] + J.jkx_d; lq[1] = lq[1] + J.jky_d;
    lpp[0] = lpp[0] + J.jkx_u; lpp[1] = lpp[1] + J.jky_u; lpp[2] = lpp[2] + J.jkz_u;
    cx lsi0v[2] = {mk(0), mk(0)};
    if (c.last) {
        lu_[0] = lu_[0] - c.qcorr * ronx; lu_[1] = lu_[1] - c.qcorr * rony;
        lsi0v[0] = c.qcorr * ronx; lsi0v[1] = c.qcorr * rony;
        for (int k = 0; k < 3; ++k) {
            const long kk = k * nk + i;
            const cx ld = c.ldl[kk], ls = ljs0[k][hp];
            c.ldl[kk] = mk(0); ljs0[k][hp] = mk(0);
            lu_[k] = lu_[k] + ld + ls;                // dl = u - js0(n), js0(n) = u
            ljs0p[k] = ljs0p[k] - ld;
        }
    }
    for (int k = 0; k < 3; ++k) {
        cx lpc = lpin[k];                             // adjoint of pin after the pass
        if (ext) {
            const cx lt = lpc + lq[k];                // pin = t, q = q' + t - pin'
            lpc = mk(0) - lq[k];                      // adjoint of pin'
            ljs0q[k] = ljs0q[k] + lt;
            const cx ld0 = (1 + c.wex) * lt;          // t = js0(n-1) + (1 + wex) d0 - wex dl
            ldl0[k] = ldl0[k] - c.wex * lt;
            lu_[k] = lu_[k] + ld0;                    // d0 = u - js0(n)
            ljs0p[k] = ljs0p[k] - ld0;
        }
        cx ljkq = lq[k];
        if (fb) ljs0q[k] = ljs0q[k] + lq[k] + lpc;    // q' = jk(n-1) + js0(n-1), pin' = js0(n-1)
        else lpin0[k] = lpin0[k] + lpc;               // pin' = pin
        ljk[k][hq] = ljk[k][hq] + ljkq;
        ljk[k][hp] = ljk[k][hp] + lpp[k];             // jk(n) = jk(n) + u - sp
        lu_[k] = lu_[k] + lpp[k];
        if (c.first) lpin0[k] = lpin0[k] - lpp[k] + lsi0o[k];
        else lsin0[k] = lsin0[k] - lpp[k];
    }
    if (c.last) {
        if (c.first) { lpin0[0] = lpin0[0] + lsi0v[0]; lpin0[1] = lpin0[1] + lsi0v[1]; }
        else { c.lsi0[i] = c.lsi0[i] + lsi0v[0]; c.lsi0[nk + i] = c.lsi0[nk + i] + lsi0v[1]; }
    }
    for (int k = 0; k < 3; ++k) {
        const long kk = k * nk + i;
        ljs0[k][hq] = ljs0[k][hq] + ljs0q[k];
        ljs0[k][hp] = ljs0[k][hp] + ljs0p[k];
        c.ldl[kk] = c.ldl[kk] + ldl0[k];
        c.lpin[kk] = c.lpin[kk] + lpin0[k];
        if (!c.first) c.lsin[kk] = c.lsin[kk] + lsin0[k];
    }
    // u = f X: the filters and the smoothing of the streaming current
    const int kx = int(i % c.nx), j = int(i / c.nx);
    const double k2 = c.kxs[kx] * c.kxs[kx] + c.kys[j] * c.kys[j];
    const double f = (kx == 0 && j == 0 && !c.keep0) ? 0.0 : c.filt[kx] * c.filty[j] / (1 + c.a2 * k2);
    for (int k = 0; k < 3; ++k) c.lU[k * nk + i] = f * lu_[k];
}

__global__ void k_real_part(const cx* X, double* J, double scale, long nk) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    for (int k = 0; k < 3; ++k) J[k * nk + i] = scale * X[k * nk + i].re;
    J[3 * nk + i] = 0;
    J[4 * nk + i] = 0;
}
__global__ void k_f2d(const float* src, long ss, double* dst, long ds, long n, int count) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(count) * n) return;
    const long c = idx / n, i = idx - c * n;
    dst[c * ds + i] = double(src[c * ss + i]);
}
__global__ void k_addp(double* y, long ds, const double* x, long ss, long n, int count) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(count) * n) return;
    const long c = idx / n, i = idx - c * n;
    y[c * ds + i] += x[c * ss + i];
}

// ---------------------------------------------------------------- sponge and real fields
// packed planes T_k = (R, L, E_z/B_z) -> adjoints; with the sponge R was rewritten as X - opb i k psi (X, Y the
// unpacked sponge spectra), so its adjoint goes to X, Y and the bound field psi
__global__ void k_pack3_adj(SpongeAdj c, const cx* T, long ts, double sc, cx* X, cx* Y, cx* P) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ts) return;
    const long i = idx % c.g.nk;
    const cx a0 = sc * T[idx], a1 = sc * T[ts + idx], a2 = sc * T[2 * ts + idx];
    c.lLx[idx] = c.lLx[idx] + a1; c.lLy[idx] = c.lLy[idx] + negi(a1);
    c.lEz[idx] = c.lEz[idx] + a2; c.lBz[idx] = c.lBz[idx] + negi(a2);
    if (!c.sponge) {
        c.lRx[idx] = c.lRx[idx] + a0; c.lRy[idx] = c.lRy[idx] + negi(a0);
        return;
    }
    const cx lr = c.lRx[idx] + a0, lq = c.lRy[idx] + negi(a0);
    c.lRx[idx] = mk(0); c.lRy[idx] = mk(0);
    X[idx] = lr; Y[idx] = lq;
    const int ixx = int(i % c.g.nx), j = int(i / c.g.nx);
    P[idx] = c.opb * (ik(c.g.kx[ixx]) * lr) + c.opb * (ik(c.g.ky[j]) * lq);
}
// X = (z + conj z(-k))/2, Y = (z - conj z(-k))/(2i): the transpose into the adjoint of z
__global__ void k_unpack_adj(int nx, int ny, long nk, const cx* X, const cx* Y, cx* S, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % nk, m = idx - i + mirror_of(i, nx, ny);
    S[idx] = 0.5 * (X[idx] + conjg(X[m])) + posi(0.5 * (Y[idx] + conjg(Y[m])));
}
// S1 = (R_x - R_bx) + i (R_y - R_by), R_b = -opb i k psi
__global__ void k_sponge_in_adj(SpongeAdj c, const cx* S, cx* P, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % c.g.nk;
    const int ixx = int(i % c.g.nx), j = int(i / c.g.nx);
    const cx lu = S[idx], lv = negi(S[idx]);
    c.lRx[idx] = c.lRx[idx] + lu;
    c.lRy[idx] = c.lRy[idx] + lv;
    P[idx] = P[idx] - c.opb * (ik(c.g.kx[ixx]) * lu) - c.opb * (ik(c.g.ky[j]) * lv);
}
// psi(l) = (rho(l) + g2 (cp rho(l+1) - (cm + cp) rho(l) + cm rho(l-1)) / k^2) / k^2 with
// rho(n') = wl[n'] jz(n') + wu[n'] jz(n'+1) (half-cell planes): the transpose, gathered per plane h of jkz
__global__ void k_jkz_adj(SpongeAdj c, const cx* P, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const SpecGeom& g = c.g;
    const long nk = g.nk, i = idx % nk;
    const int h = int(idx / nk);
    const int ixx = int(i % g.nx), j = int(i / g.nx);
    const double kxv = g.kx[ixx], kyv = g.ky[j], k2 = kxv * kxv + kyv * kyv;
    if (!(k2 > 0)) return;
    auto coef = [&](int l, int dl) {
        const int n = g.a + l;
        const double hm = c.hc[n + 1], hpp = c.hc[n + 2];
        const double cm = 2 / ((hm + hpp) * hm), cp = 2 / ((hm + hpp) * hpp);
        if (dl == 0) return 1 / k2 - c.g2 * (cm + cp) / (k2 * k2);
        return c.g2 * (dl > 0 ? cp : cm) / (k2 * k2);
    };
    cx s = mk(0);
    for (int dl = -1; dl <= 1; ++dl) {
        // plane h as h0 = l + dl of rho(n' = a + l + dl) with weight wl[n']
        int l = h - dl, nn = g.a + h;
        if (l >= 0 && l < g.ns && nn >= 0 && nn <= g.N - 1) s = s + (coef(l, dl) * c.wl[nn]) * P[long(l) * nk + i];
        // plane h as h1 = l + dl + 1 with weight wu[n'], n' = a + h - 1
        l = h - 1 - dl; nn = g.a + h - 1;
        if (l >= 0 && l < g.ns && nn >= 0 && nn <= g.N - 1) s = s + (coef(l, dl) * c.wu[nn]) * P[long(l) * nk + i];
    }
    c.ljkz[idx] = c.ljkz[idx] + s;
}

// ---------------------------------------------------------------- beam currents
__global__ void k_cur_unpack_adj(CurrentsAdj c, cx* X, cx* Y, cx* Z, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % c.g.nk;
    const int ixx = int(i % c.g.nx), j = int(i / c.g.nx);
    const double f = (ixx == 0 && j == 0 && !c.keep0) ? 0.0 : c.g.filt[ixx] * c.g.filty[j];
    X[idx] = f * c.ljkx[idx]; Y[idx] = f * c.ljky[idx]; Z[idx] = f * c.ljkz[idx];
    c.ljkx[idx] = c.ljky[idx] = c.ljkz[idx] = mk(0);
}
// jk plane hk (half-cell a - 1 + hk) = the zeta filter of jw planes hk, hk + 1, hk + 2: the transpose per jw plane
__global__ void k_zfilter_adj(CurrentsAdj c, const cx* S, const cx* Z, double* ljw, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const SpecGeom& g = c.g;
    const long nk = g.nk, i = idx % nk;
    const int p = int(idx / nk), nh = g.ns + 3;
    double v[3] = {0, 0, 0};
    for (int d = 0; d <= 2; ++d) {                  // plane p is lo (d = 0), mid (1) or hi (2) of hk = p - d
        const int hk = p - d, hc_ = g.a - 1 + hk;
        if (hk < 0 || hk > g.ns || hc_ < 0 || hc_ > g.N - 2) continue;
        double w;
        if (c.zf && c.zu) w = d == 1 ? 0.5 : 0.25;
        else if (c.zf) w = (d == 0 ? 0.25 * c.hcg[hc_ + 1] : d == 1 ? 0.5 * c.hcg[hc_ + 2] : 0.25 * c.hcg[hc_ + 3]) / c.hcg[hc_ + 2];
        else w = d == 1 ? 1.0 : 0.0;
        const long q = long(hk) * nk + i;
        v[0] += w * S[q].re; v[1] += w * S[q].im; v[2] += w * Z[q].re;
    }
    for (int comp = 0; comp < 3; ++comp) ljw[(long(comp) * nh + p) * nk + i] = v[comp];
}
// one beam particle's TSC deposit (k_dep_emit, Fourier y): jx = W px / pz, jy = W py / pz, jz = W with
// W = q w wx(x) wy(y) ia / (h dx) in its half-cell; the transpose in x, y, p_x, p_y, p_z
__global__ void k_dep_adj(CurrentsAdj c, Parts P, long np, const double* ljw, double* lx, double* ly, double* lpx,
                          double* lpy, double* lpz) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= np || P.alive[ip] != 1) return;
    const DepositArgs& d = c.d;
    int m; double fz;
    d.zax.locate(P.zeta[ip], m, fz);
    const int hw = m - (d.a - 2);
    if (hw < 0 || hw >= d.nh || m < 0 || m > d.N - 2) return;
    const double px = P.px[ip], py = P.py[ip], pz = P.pz[ip];
    const double f = P.q[ip] * P.w[ip] / (d.hcg[m + 2] * d.dx) * d.ia[0];
    int ix[3], jn[3]; double wx[3], wn[3], dwx[3], dwn[3];
    tsc_periodic(P.x[ip], d.dx, d.nx, ix, wx);
    tsc_periodic(P.y[ip], d.dy, d.ny, jn, wn);
    const double xi = P.x[ip] / d.dx, fx = xi - floor(xi + 0.5);
    const double yi = P.y[ip] / d.dy, fy = yi - floor(yi + 0.5);
    dwx[0] = -(0.5 - fx) / d.dx; dwx[1] = -2 * fx / d.dx; dwx[2] = (0.5 + fx) / d.dx;
    dwn[0] = -(0.5 - fy) / d.dy; dwn[1] = -2 * fy / d.dy; dwn[2] = (0.5 + fy) / d.dy;
    const long nk = d.nk, pl = long(d.nh) * nk, base = long(hw) * nk;
    double gx = 0, gy = 0, sx = 0, sy = 0;
    for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) {
            const long k = base + long(jn[v]) * d.nx + ix[u];
            const double a0 = ljw[k], a1 = ljw[pl + k], a2 = ljw[2 * pl + k];
            const double gsum = a0 * px / pz + a1 * py / pz + a2;
            gx += dwx[u] * wn[v] * gsum;
            gy += wx[u] * dwn[v] * gsum;
            const double W = wx[u] * wn[v];
            sx += W * a0;
            sy += W * a1;
        }
    lx[ip] += f * gx;
    ly[ip] += f * gy;
    lpx[ip] += f * sx / pz;
    lpy[ip] += f * sy / pz;
    lpz[ip] -= f * (sx * px + sy * py) / (pz * pz);
}

// the same deposit's tangent (the tangent of zeta moves no current: held fixed)
__global__ void k_dep_tlm(CurrentsAdj c, Parts P, long np, const double* tx, const double* ty, const double* tpx,
                          const double* tpy, const double* tpz, double* djw) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= np || P.alive[ip] != 1) return;
    const DepositArgs& d = c.d;
    int m; double fz;
    d.zax.locate(P.zeta[ip], m, fz);
    const int hw = m - (d.a - 2);
    if (hw < 0 || hw >= d.nh || m < 0 || m > d.N - 2) return;
    const double px = P.px[ip], py = P.py[ip], pz = P.pz[ip];
    const double f = P.q[ip] * P.w[ip] / (d.hcg[m + 2] * d.dx) * d.ia[0];
    int ix[3], jn[3]; double wx[3], wn[3];
    tsc_periodic(P.x[ip], d.dx, d.nx, ix, wx);
    tsc_periodic(P.y[ip], d.dy, d.ny, jn, wn);
    const double xi = P.x[ip] / d.dx, fx = xi - floor(xi + 0.5);
    const double yi = P.y[ip] / d.dy, fy = yi - floor(yi + 0.5);
    const double dwx[3] = {-(0.5 - fx) / d.dx, -2 * fx / d.dx, (0.5 + fx) / d.dx};
    const double dwn[3] = {-(0.5 - fy) / d.dy, -2 * fy / d.dy, (0.5 + fy) / d.dy};
    const double ax = px / pz, ay = py / pz;
    const double dax = tpx[ip] / pz - px * tpz[ip] / (pz * pz), day = tpy[ip] / pz - py * tpz[ip] / (pz * pz);
    const long nk = d.nk, pl = long(d.nh) * nk, base = long(hw) * nk;
    for (int v = 0; v < 3; ++v)
        for (int u = 0; u < 3; ++u) {
            const long k = base + long(jn[v]) * d.nx + ix[u];
            const double W = f * wx[u] * wn[v];
            const double dW = f * (dwx[u] * wn[v] * tx[ip] + wx[u] * dwn[v] * ty[ip]);
            atomicAdd(&djw[k], dW * ax + W * dax);
            atomicAdd(&djw[pl + k], dW * ay + W * day);
            atomicAdd(&djw[2 * pl + k], dW);
        }
}
__global__ void k_to_F_d(const cx* T, long ts, double* F, long fpl, long nk, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const cx R = T[idx], L = T[ts + idx], Z = T[2 * ts + idx];
    F[0 * fpl + idx] = 0.5 * (R.re + L.re);
    F[1 * fpl + idx] = 0.5 * (R.im + L.im);
    F[3 * fpl + idx] = 0.5 * (L.im - R.im);
    F[4 * fpl + idx] = 0.5 * (R.re - L.re);
    F[2 * fpl + idx] = Z.re;
    F[5 * fpl + idx] = Z.im;
    (void)nk;
}

__global__ void k_seed_pz(Parts P, long np, int sp0, double inv, double* lpz) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= np) return;
    lpz[i] = (P.alive[i] == 1 && P.sp[i] == sp0) ? inv : 0.0;
}

__global__ void k_laser_dot(SpecGeom g, const cx* T, const cx* lRx, const cx* lRy, long tot, double* out) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    double v = 0;
    int l = 0;
    if (idx < tot) {
        const long i = idx % g.nk;
        l = int(idx / g.nk);
        const cx z = T[idx], zm = conjg(T[idx - i + mirror_of(i, g.nx, g.ny)]);
        const cx X = 0.5 * (z + zm), dd = z - zm, Y = 0.5 * mk(dd.im, -dd.re);
        v = lRx[idx].re * X.re + lRx[idx].im * X.im + lRy[idx].re * Y.re + lRy[idx].im * Y.im;
    }
    if (g.nk % 64 == 0) {                           // a wave stays within one slice: reduce it first
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if ((threadIdx.x & 63) == 0 && idx < tot) atomicAdd(out + l, v);
    } else if (idx < tot) atomicAdd(out + l, v);
}
// a reverse pass's inputs: the two float slices (6 planes fpl apart) into double planes (nk apart), the field
// adjoints the streaming reverse scatters into cleared
__global__ void k_pass_prep(const float* f0, const float* f1, long fpl, double* F0d, double* F1d, double* gF0,
                            double* gF1, long nk) {
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= 6 * nk) return;
    const long c = e / nk, i = e - c * nk;
    F0d[e] = double(f0[c * fpl + i]);
    F1d[e] = double(f1[c * fpl + i]);
    gF0[e] = 0;
    gF1[e] = 0;
}
// a reverse pass's sums: acc += gP (m values); dst0 += gF0 (planes s0 apart; skipped when null); dst1 += gF1
__global__ void k_pass_sums(double* acc, const double* gP, long m, double* dst0, long s0, const double* gF0,
                            double* dst1, long s1, const double* gF1, long nk) {
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m) acc[e] += gP[e];
    if (e >= 6 * nk) return;
    const long c = e / nk, i = e - c * nk;
    if (dst0) dst0[c * s0 + i] += gF0[e];
    dst1[c * s1 + i] += gF1[e];
}

// a tangent-linear pass's inputs: the two float slices into double planes as k_pass_prep; the tangent fields of
// the slice (t0, when given) and of slice n + 1 (t1), planes fpl apart, into t0d, t1d (0 + x, as the clear and sum
// they replace); the tangent current planes J (nJ values) cleared
__global__ void k_tangent_prep(const float* f0, const float* f1, long fpl, double* F0d, double* F1d, const double* t0,
                               double* t0d, const double* t1, double* t1d, double* J, long nJ, long nk) {
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < nJ) J[e] = 0;
    if (e >= 6 * nk) return;
    const long c = e / nk, i = e - c * nk;
    F0d[e] = double(f0[c * fpl + i]);
    F1d[e] = double(f1[c * fpl + i]);
    if (t0) t0d[e] = 0.0 + t0[c * fpl + i];
    t1d[e] = 0.0 + t1[c * fpl + i];
}

struct SpanSet {
    CopySpan s[kMaxSpans];
};
// copy y of the set (blockIdx.y), its words over the x blocks
__global__ void __launch_bounds__(256) k_copy_spans(SpanSet c) {
    const CopySpan sp = c.s[blockIdx.y];
    const unsigned* src = static_cast<const unsigned*>(sp.src);
    unsigned* dst = static_cast<unsigned*>(sp.dst);
    for (long i = long(blockIdx.x) * blockDim.x + threadIdx.x; i < sp.words; i += long(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

#define G3D_GLAUNCH(kern, n, ...)                                                                            \
    do {                                                                                                     \
        const long n_ = (n);                                                                                 \
        if (n_ > 0) {                                                                                        \
            hipLaunchKernelGGL(kern, dim3(blocks_for(n_, kT)), dim3(kT), 0, stream(), __VA_ARGS__);          \
            launch_check(#kern);                                                                             \
        }                                                                                                    \
    } while (0)

// the unnormalised inverse transform (the adjoint of fft_run's forward)
static void fft_inverse_unscaled(const FftPlan& p, cx* X, int count) {
    fft_rows(p, X, long(count) * p.ny, true, 1.0);
    fft_cols(p, X, count, true, 1.0);
}

void fields_to_F_adj(const double* lF, long fpl, long nk, int count, cx* lT, long ts) {
    G3D_GLAUNCH(k_toF_adj, long(count) * nk, lF, fpl, nk, count, lT, ts);
}
void pass_mode_adj(const PassAdj& c) { G3D_GLAUNCH(k_pass_adj, c.nk, c); }
void real_part_planes(const cx* X, double* J, double scale, long nk) { G3D_GLAUNCH(k_real_part, nk, X, J, scale, nk); }
void planes_to_double(const float* src, long ss, double* dst, long ds, long n, int count) {
    G3D_GLAUNCH(k_f2d, long(count) * n, src, ss, dst, ds, n, count);
}
void add_planes(double* y, long ds, const double* x, long ss, long n, int count) {
    G3D_GLAUNCH(k_addp, long(count) * n, y, ds, x, ss, n, count);
}
void add_values(double* y, const double* x, long n) { G3D_GLAUNCH(k_addp, n, y, n, x, n, n, 1); }

void pass_prep(const float* f0, const float* f1, long fpl, double* F0d, double* F1d, double* gF0, double* gF1,
               long nk) {
    G3D_GLAUNCH(k_pass_prep, 6 * nk, f0, f1, fpl, F0d, F1d, gF0, gF1, nk);
}
void pass_sums(double* acc, const double* gP, long m, double* dst0, long s0, const double* gF0, double* dst1, long s1,
               const double* gF1, long nk) {
    G3D_GLAUNCH(k_pass_sums, std::max(m, 6 * nk), acc, gP, m, dst0, s0, gF0, dst1, s1, gF1, nk);
}

void tangent_prep(const float* f0, const float* f1, long fpl, double* F0d, double* F1d, const double* t0, double* t0d,
                  const double* t1, double* t1d, double* J, long nJ, long nk) {
    G3D_GLAUNCH(k_tangent_prep, std::max(nJ, 6 * nk), f0, f1, fpl, F0d, F1d, t0, t0d, t1, t1d, J, nJ, nk);
}

void copy_spans(const CopySpan* spans, int count) {
    if (count <= 0) return;
    if (count > kMaxSpans) throw std::runtime_error("copy_spans: at most 8 copies per launch");
    SpanSet c{};
    long most = 0;
    for (int k = 0; k < count; ++k) { c.s[k] = spans[k]; most = std::max(most, spans[k].words); }
    if (most <= 0) return;
    const unsigned bx = std::min<unsigned>(blocks_for(most, 256), 2048u);
    hipLaunchKernelGGL(k_copy_spans, dim3(bx, unsigned(count)), dim3(256), 0, stream(), c);
    launch_check("copy spans");
}

void sponge_real_fields_adj(const SpongeAdj& c, const FftPlan& p, double* lF, long fpl, cx* T, cx* X, cx* Y, cx* P) {
    const SpecGeom& g = c.g;
    const long ts = long(g.ns) * g.nk;
    fields_to_F_adj(lF, fpl, g.nk, g.ns, T, ts);
    fft_run(p, T, 3 * g.ns, false);
    G3D_GLAUNCH(k_pack3_adj, ts, c, T, ts, 1.0 / (double(g.nx) * g.ny), X, Y, P);
    if (!c.sponge) return;
    // S1 -> inverse transform, mask, forward transform: with a real diagonal mask its own transpose
    G3D_GLAUNCH(k_unpack_adj, ts, g.nx, g.ny, g.nk, X, Y, T, ts);
    fft_run(p, T, g.ns, true);
    sponge_mask(g, T, c.mx, c.mn, c.mh, g.ns);
    fft_run(p, T, g.ns, false);
    G3D_GLAUNCH(k_sponge_in_adj, ts, c, T, P, ts);
    G3D_GLAUNCH(k_jkz_adj, long(g.ns + 1) * g.nk, c, P, long(g.ns + 1) * g.nk);
}

void currents_adj(const CurrentsAdj& c, const FftPlan& p, Parts P, long np, double* lx, double* ly, double* lpx,
                  double* lpy, double* lpz, cx* X, cx* Y, cx* Z, cx* S, double* ljw) {
    const SpecGeom& g = c.g;
    const long tot = long(g.ns + 1) * g.nk;
    G3D_GLAUNCH(k_cur_unpack_adj, tot, c, X, Y, Z, tot);
    G3D_GLAUNCH(k_unpack_adj, tot, g.nx, g.ny, g.nk, X, Y, S, tot);
    fft_inverse_unscaled(p, S, g.ns + 1);
    fft_inverse_unscaled(p, Z, g.ns + 1);
    G3D_GLAUNCH(k_zfilter_adj, long(g.ns + 3) * g.nk, c, S, Z, ljw, long(g.ns + 3) * g.nk);
    G3D_GLAUNCH(k_dep_adj, np, c, P, np, ljw, lx, ly, lpx, lpy, lpz);
}

void deposit_tlm(const CurrentsAdj& c, Parts P, long np, const double* dx, const double* dy, const double* dpx,
                 const double* dpy, const double* dpz, double* djw) {
    G3D_GLAUNCH(k_dep_tlm, np, c, P, np, dx, dy, dpx, dpy, dpz, djw);
}
void fields_to_F_double(const cx* T, long ts, double* F, long fpl, long nk, int count) {
    G3D_GLAUNCH(k_to_F_d, long(count) * nk, T, ts, F, fpl, nk, long(count) * nk);
}
void seed_mean_pz(Parts P, long np, int sp0, double inv_count, double* lpz) {
    G3D_GLAUNCH(k_seed_pz, np, P, np, sp0, inv_count, lpz);
}
void laser_projection(const SpecGeom& g, const cx* T, const cx* lRx, const cx* lRy, int count, double* out) {
    G3D_GLAUNCH(k_laser_dot, long(count) * g.nk, g, T, lRx, lRy, long(count) * g.nk, out);
}

}  // namespace g3d
