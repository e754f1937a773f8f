// Spectral kernels of the Fourier-y grid around the batched FFT: the zeta filter and unpacking of the beam
// currents, the sponge with the bound field, the real-space fields, the grid laser, the co-moving start field.
// One lane per point of a plane; every kernel is one pass over its planes.
#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 256;

__device__ __forceinline__ long mirror(long i, int nx, int ny) {
    const int ixx = int(i % nx), j = int(i / nx);
    return long((ny - j) % ny) * nx + (nx - ixx) % nx;
}
// X = (z + conj z(-k))/2, Y = (z - conj z(-k))/(2i) of a packed pair (src/sim_particles.cpp:153-155)
__device__ __forceinline__ void unpack_pair(cx z, cx zm, cx& X, cx& Y) {
    X = 0.5 * (z + zm);
    const cx d = z - zm;
    Y = 0.5 * mk(d.im, -d.re);
}

// ---------------------------------------------------------------- currents (src/sim_particles.cpp:120-158)
__global__ void k_zfilter(SpecGeom g, const double* J, const double* hc, int zf, int zu, cx* S1, cx* jz) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long tot = long(g.ns + 1) * g.nk;
    if (idx >= tot) return;
    const int nh = g.ns + 3;
    const int hk = int(idx / g.nk);
    const long i = idx % g.nk;
    const int hc_ = g.a - 1 + hk, hw = hk + 1;
    double v[3] = {0, 0, 0};
    if (hc_ >= 0 && hc_ <= g.N - 2)
        for (int comp = 0; comp < 3; ++comp) {
            const double* Jc = J + size_t(comp) * nh * g.nk;
            const double lo = Jc[(hw - 1) * g.nk + i], mid = Jc[hw * g.nk + i], hi = Jc[(hw + 1) * g.nk + i];
            if (zf && zu) v[comp] = 0.25 * lo + 0.5 * mid + 0.25 * hi;
            else v[comp] = zf ? (0.25 * hc[hc_ + 1] * lo + 0.5 * hc[hc_ + 2] * mid + 0.25 * hc[hc_ + 3] * hi) / hc[hc_ + 2]
                              : mid;
        }
    S1[idx] = mk(v[0], v[1]);
    jz[idx] = mk(v[2], 0);
}
__global__ void k_cur_unpack(SpecGeom g, const cx* S1, cx* jx, cx* jy, cx* jz, int keep0) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(g.ns + 1) * g.nk) return;
    const long i = idx % g.nk, base = idx - i;
    const int ixx = int(i % g.nx), j = int(i / g.nx);
    cx X, Y;
    unpack_pair(S1[idx], conjg(S1[base + mirror(i, g.nx, g.ny)]), X, Y);
    const double f = (ixx == 0 && j == 0 && !keep0) ? 0.0 : g.filt[ixx] * g.filty[j];
    jx[idx] = f * X; jy[idx] = f * Y; jz[idx] = f * jz[idx];
}

// ---------------------------------------------------------------- sponge (src/sim_fields.cpp:163-268)
__global__ void k_sponge_pack(SpecGeom g, const cx* rx, const cx* ry, const cx* jz, const double* hc, const double* wl,
                              const double* wu, double g2, double opb, cx* s1, cx* s2) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= long(g.ns) * g.nk) return;
    const long i = idx % g.nk;
    const int ixx = int(i % g.nx), j = int(i / g.nx), l = int(idx / g.nk), n = g.a + l;
    const long ic = g.cm ? long(ixx) * g.ny + j : i;       // the mode in the layout of R and jk
    // charge of slice n + dl from the filtered half-cell currents (the ghost half-cells of one GPU are zero)
    auto rho_at = [&](int dl) -> cx {
        const int nn = n + dl;
        if (nn < 0 || nn > g.N - 1) return mk(0);
        const int h0 = l + dl, h1 = l + dl + 1;
        const cx a = (h0 < 0 || h0 > g.ns) ? mk(0) : jz[long(h0) * g.nk + ic];
        const cx b = (h1 < 0 || h1 > g.ns) ? mk(0) : jz[long(h1) * g.nk + ic];
        return wl[nn] * a + wu[nn] * b;
    };
    const double hm = hc[n + 1], hp = hc[n + 2];
    const double cm = 2 / ((hm + hp) * hm), cp = 2 / ((hm + hp) * hp);
    const double kxv = g.kx[ixx], kyv = g.ky[j];
    const double k2 = kxv * kxv + kyv * kyv;
    cx psi = mk(0);
    if (k2 > 0) {
        const cx r0 = rho_at(0);
        const cx d2 = g2 * (cp * rho_at(1) - (cm + cp) * r0 + cm * rho_at(-1));
        psi = (r0 + d2 / k2) / k2;
    }
    s2[idx] = psi;
    const cx bx = -opb * (ik(kxv) * psi), by = -opb * (ik(kyv) * psi);
    const long f = long(l) * g.nk + ic;
    const cx u = rx[f] - bx, v = ry[f] - by;
    s1[idx] = mk(u.re - v.im, u.im + v.re);
}
__global__ void k_sponge_mask(SpecGeom g, cx* s1, const double* mx, const double* mn, const double* mh, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % g.nk;
    const int j = int(i / g.nx), ixx = int(i % g.nx);
    s1[idx] = mk(mx[ixx] * mn[j] * s1[idx].re, mx[ixx] * mh[j] * s1[idx].im);
}
// optional sponge unpack (+ R_b), then the three packed planes R_x + i R_y, L_x + i L_y, E_z + i B_z
__global__ void k_pack3(SpecGeom g, const cx* s1, const cx* s2, double opb, cx* rx, cx* ry, const cx* lx, const cx* ly,
                        const cx* ez, const cx* bz, cx* T, long ts, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % g.nk, base = idx - i;
    const int ixx = int(i % g.nx), j = int(i / g.nx);
    const long f = g.cm ? base + long(ixx) * g.ny + j : idx;     // the mode in the layout of R, L, E_z, B_z
    cx r = rx[f], q = ry[f];
    if (s1) {
        cx X, Y;
        unpack_pair(s1[idx], conjg(s1[base + mirror(i, g.nx, g.ny)]), X, Y);
        r = X - opb * (ik(g.kx[ixx]) * s2[idx]);
        q = Y - opb * (ik(g.ky[j]) * s2[idx]);
        rx[f] = r; ry[f] = q;
    }
    const cx l = lx[f], m = ly[f], e = ez[f], b = bz[f];
    T[idx] = mk(r.re - q.im, r.im + q.re);
    T[ts + idx] = mk(l.re - m.im, l.im + m.re);
    T[2 * ts + idx] = mk(e.re - b.im, e.im + b.re);
}
// The same two kernels with R, L, E_z, B_z and jk column-major (mode ix ny + j) and the FFT scratch row-major:
// one workgroup per tile of 16 x 16 modes of a plane, the column-major arrays staged through LDS so that both
// layouts are read and written along their rows.  Lane t: (ix0 + t / 16, j0 + t % 16) for the column-major
// accesses, (ix0 + t % 16, j0 + t / 16) for the row-major ones; LDS tiles are [ix - ix0][j - j0].
constexpr int kTT = 16;
__global__ void __launch_bounds__(256) k_sponge_pack_cm(SpecGeom g, const cx* rx, const cx* ry, const cx* jz,
                                                        const double* hc, const double* wl, const double* wu,
                                                        double g2, double opb, cx* s1, cx* s2) {
    __shared__ cx trx[kTT][kTT + 1], tr_y[kTT][kTT + 1], tjz[4][kTT][kTT + 1];
    const int ntx = g.nx / kTT, nty = g.ny / kTT;
    const long per = long(ntx) * nty;
    const int l = int(blockIdx.x / per), t = int(blockIdx.x - l * per), ix0 = (t % ntx) * kTT, j0 = (t / ntx) * kTT;
    const int a = threadIdx.x / kTT, b = threadIdx.x % kTT;
    {
        const long ic = long(ix0 + a) * g.ny + j0 + b;
        trx[a][b] = rx[long(l) * g.nk + ic];
        tr_y[a][b] = ry[long(l) * g.nk + ic];
        for (int q = 0; q < 4; ++q) {                 // half-cell planes l - 1 ... l + 2
            const int h = l - 1 + q;
            tjz[q][a][b] = (h < 0 || h > g.ns) ? mk(0) : jz[long(h) * g.nk + ic];
        }
    }
    __syncthreads();
    const int ixx = ix0 + b, j = j0 + a, n = g.a + l;
    const long idx = long(l) * g.nk + long(j) * g.nx + ixx;
    auto rho_at = [&](int dl) -> cx {
        const int nn = n + dl;
        if (nn < 0 || nn > g.N - 1) return mk(0);
        return wl[nn] * tjz[dl + 1][b][a] + wu[nn] * tjz[dl + 2][b][a];
    };
    const double hm = hc[n + 1], hp = hc[n + 2];
    const double cm = 2 / ((hm + hp) * hm), cp = 2 / ((hm + hp) * hp);
    const double kxv = g.kx[ixx], kyv = g.ky[j];
    const double k2 = kxv * kxv + kyv * kyv;
    cx psi = mk(0);
    if (k2 > 0) {
        const cx r0 = rho_at(0);
        const cx d2 = g2 * (cp * rho_at(1) - (cm + cp) * r0 + cm * rho_at(-1));
        psi = (r0 + d2 / k2) / k2;
    }
    s2[idx] = psi;
    const cx bx = -opb * (ik(kxv) * psi), by = -opb * (ik(kyv) * psi);
    const cx u = trx[b][a] - bx, v = tr_y[b][a] - by;
    s1[idx] = mk(u.re - v.im, u.im + v.re);
}
// TO: cx, or cxf for the FP32 transforms of the FP32-mixed form
__device__ __forceinline__ void put(cx* T, long i, cx v) { T[i] = v; }
__device__ __forceinline__ void put(cxf* T, long i, cx v) { T[i] = to_f(v); }
template <class TO>
__global__ void __launch_bounds__(256) k_pack3_cm(SpecGeom g, const cx* s1, const cx* s2, double opb, cx* rx, cx* ry,
                                                  const cx* lx, const cx* ly, const cx* ez, const cx* bz, TO* T,
                                                  long ts) {
    __shared__ cx tile[6][kTT][kTT + 1];
    const int ntx = g.nx / kTT, nty = g.ny / kTT;
    const long per = long(ntx) * nty;
    const int l = int(blockIdx.x / per), t = int(blockIdx.x - l * per), ix0 = (t % ntx) * kTT, j0 = (t / ntx) * kTT;
    const int a = threadIdx.x / kTT, b = threadIdx.x % kTT;
    const long base = long(l) * g.nk;
    const long ic = base + long(ix0 + a) * g.ny + j0 + b;
    tile[0][a][b] = rx[ic]; tile[1][a][b] = ry[ic]; tile[2][a][b] = lx[ic];
    tile[3][a][b] = ly[ic]; tile[4][a][b] = ez[ic]; tile[5][a][b] = bz[ic];
    __syncthreads();
    const int ixx = ix0 + b, j = j0 + a;
    const long i = long(j) * g.nx + ixx, idx = base + i;
    cx r = tile[0][b][a], q = tile[1][b][a];
    if (s1) {
        cx X, Y;
        unpack_pair(s1[idx], conjg(s1[base + mirror(i, g.nx, g.ny)]), X, Y);
        r = X - opb * (ik(g.kx[ixx]) * s2[idx]);
        q = Y - opb * (ik(g.ky[j]) * s2[idx]);
    }
    const cx lv = tile[2][b][a], m = tile[3][b][a], e = tile[4][b][a], bv = tile[5][b][a];
    put(T, idx, mk(r.re - q.im, r.im + q.re));
    put(T, ts + idx, mk(lv.re - m.im, lv.im + m.re));
    put(T, 2 * ts + idx, mk(e.re - bv.im, e.im + bv.re));
    if (!s1) return;
    __syncthreads();
    tile[0][b][a] = r; tile[1][b][a] = q;
    __syncthreads();
    rx[ic] = tile[0][a][b]; ry[ic] = tile[1][a][b];
}
// E_x = (R_x + L_x)/2, E_y = (R_y + L_y)/2, B_x = (L_y - R_y)/2, B_y = (R_x - L_x)/2, E_z, B_z
// (src/sim_fields.cpp:280-293) into float F
__global__ void k_to_F_f(const cxf* T, long ts, float* F, long fpl, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const cxf R = T[idx], L = T[ts + idx], Z = T[2 * ts + idx];
    F[0 * fpl + idx] = 0.5f * (R.re + L.re);
    F[1 * fpl + idx] = 0.5f * (R.im + L.im);
    F[3 * fpl + idx] = 0.5f * (L.im - R.im);
    F[4 * fpl + idx] = 0.5f * (R.re - L.re);
    F[2 * fpl + idx] = Z.re;
    F[5 * fpl + idx] = Z.im;
}
__global__ void k_to_F(const cx* T, long ts, float* F, long fpl, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const cx R = T[idx], L = T[ts + idx], Z = T[2 * ts + idx];
    const double r1 = R.re, r2 = R.im, l1 = L.re, l2 = L.im;
    F[0 * fpl + idx] = float(0.5 * (r1 + l1));
    F[1 * fpl + idx] = float(0.5 * (r2 + l2));
    F[3 * fpl + idx] = float(0.5 * (l2 - r2));
    F[4 * fpl + idx] = float(0.5 * (r1 - l1));
    F[2 * fpl + idx] = float(Z.re);
    F[5 * fpl + idx] = float(Z.im);
}

// ---------------------------------------------------------------- streaming current pack (src/sim_plasma.cpp:99)
__global__ void k_sspec_pack(double* J, long nk, int nx, int ny, cx* pk, double* rl, double* rli, int i0, int j0,
                             TileGeom tg) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int ixx = int(i % nx), j = int(i / nx);
    double v[5];
    for (int c = 0; c < 5; ++c) v[c] = J[c * nk + i];
    if (tg.ntile) {
        // the windows of the tiles next to the point's own (a window reaches M + 1 nodes beyond its tile)
        const int tpx = ixx / tg.TX, tpy = j / tg.TY;
        const long wpl = long(tg.WX) * tg.WY;
        for (int dy = -1; dy <= 1; ++dy) {
            const int ty = wrap_index(tpy + dy, tg.nty);
            if ((dy == 1 && tg.nty <= 2) || (dy != 0 && tg.nty == 1)) continue;
            int vv = j - (ty * tg.TY - 1 - tg.M);
            vv = wrap_index(vv, ny);
            if (vv >= tg.WY) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int tx = wrap_index(tpx + dx, tg.ntx);
                if ((dx == 1 && tg.ntx <= 2) || (dx != 0 && tg.ntx == 1)) continue;
                int uu = ixx - (tx * tg.TX - 1 - tg.M);
                uu = wrap_index(uu, nx);
                if (uu >= tg.WX) continue;
                const double* sl = tg.slab + (long(ty) * tg.ntx + tx) * 5 * wpl + long(vv) * tg.WX + uu;
                for (int c = 0; c < 5; ++c) v[c] += sl[c * wpl];
            }
        }
    }
    pk[i] = mk(v[0], v[1]);
    pk[nk + i] = mk(v[2], 0);
    if (rl) {                                   // density lines y = 0 and x = 0 (src/sim_plasma.cpp:363-374)
        if (j == j0) rl[ixx] = v[3];
        if (ixx == i0) rl[nx + j] = v[3];
        if (rli) {
            if (j == j0) rli[ixx] = v[4];
            if (ixx == i0) rli[nx + j] = v[4];
        }
    }
    for (int c = 0; c < 5; ++c) J[c * nk + i] = 0;
}

__global__ void k_sspec_unpack(SpecGeom g, const cx* pk, cx* ox, double a2, int keep0) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= g.nk) return;
    const int ixx = int(i % g.nx), j = int(i / g.nx);
    const cx z = pk[i], zm = conjg(pk[mirror(i, g.nx, g.ny)]);
    const double k2 = g.kxs[ixx] * g.kxs[ixx] + g.kys[j] * g.kys[j];
    const double f = (ixx == 0 && j == 0 && !keep0) ? 0.0 : g.filt[ixx] * g.filty[j] / (1 + a2 * k2);
    cx X, Y;
    unpack_pair(z, zm, X, Y);
    ox[i] = f * X; ox[g.nk + i] = f * Y; ox[2 * g.nk + i] = f * pk[g.nk + i];
}

// ---------------------------------------------------------------- grid laser (src/sim_fields.cpp:321-374)
// complex Gaussian beam E = amp env exp(-(x^2 + y^2)/(w^2 q)) / q car at z = 0, q = 1 - i z_focus/z_R
// the paraxial Gaussian at transverse point (xp, yp) of slice n: 2 E_x, 2 E_y
__device__ void glaser_field(const GridLaserArgs& a, int n, double xp, double yp, double& ex, double& ey) {
    const double d = a.zeta[n] - a.zc;
    const double env = a.env ? a.env[n] : exp(-2 * 0.69314718055994531 * d * d / (a.tau * a.tau));
    // 1/q with q = 1 + i qi
    const double qi = (0.0 - a.zfoc) / a.zR;
    const double qq = 1 + qi * qi;
    const cx iq = mk(1 / qq, -qi / qq);
    const double r2 = (xp * xp + yp * yp) / (a.waist * a.waist);
    // exp(-r2 / q) = exp(-r2 Re(1/q)) (cos(-r2 Im(1/q)) + i sin(...))
    const double mag = exp(-r2 * iq.re), ph0 = -r2 * iq.im;
    const cx e1 = mk(mag * cos(ph0), mag * sin(ph0));
    const double ph = a.kL * d + a.cep;
    const cx car = mk(cos(ph), sin(ph));
    const cx E = (a.amp * env) * (e1 * iq) * car;
    if (a.circular) { ex = E.re; ey = a.helicity * E.im; }
    else { ex = E.re * a.rc; ey = E.re * a.rs; }
}
// R_x + i R_y = 2 E_perp on the grid; free-hand y (yn != null): R_x on the nodes (zero on the wall), R_y on the
// half-nodes
__global__ void k_glaser_fill(GridLaserArgs a, cx* T, int n0, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const int l = int(idx / a.nk);
    const long i = idx % a.nk;
    const int ixx = int(i % a.nx), j = int(i / a.nx), n = n0 + l;
    const double d = a.zeta[n] - a.zc;
    const double env = a.env ? a.env[n] : exp(-2 * 0.69314718055994531 * d * d / (a.tau * a.tau));
    if (fabs(env) < 1e-12) { T[idx] = mk(0); return; }
    const double xp = ixx * a.dx + a.X0;
    double ex, ey;
    if (!a.yn) {
        glaser_field(a, n, xp, j * a.dy + a.Y0, ex, ey);
        T[idx] = mk(2 * ex, 2 * ey);
        return;
    }
    double ex2, ey2;
    glaser_field(a, n, xp, a.yn[j], ex, ey);
    glaser_field(a, n, xp, 0.5 * (a.yn[j] + a.yn[j + 1]), ex2, ey2);
    T[idx] = mk(j == 0 ? 0.0 : 2 * ex, 2 * ey2);
}
__global__ void k_glaser_add(SpecGeom g, const cx* T, cx* rx, cx* ry, long tot, int fh) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % g.nk, base = idx - i;
    const int ixx = int(i % g.nx);
    const long m = fh ? idx - ixx + (g.nx - ixx) % g.nx : base + mirror(i, g.nx, g.ny);
    cx X, Y;
    unpack_pair(T[idx], conjg(T[m]), X, Y);
    rx[idx] = rx[idx] + X; ry[idx] = ry[idx] + Y;
}

// ---------------------------------------------------------------- window shift (src/sim_setup.cpp:249-281)
__global__ void k_rotate(cx* A, long n, long nk, int nx, const double* kx, const double* ky, double sx, double sy, int ys) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const long i = idx % nk;
    const double ph = kx[i % nx] * sx + (ys ? ky[i / nx] * sy : 0.0);
    A[idx] = A[idx] * mk(cos(ph), sin(ph));
}
__global__ void k_roll(const float* f, float* t, long n, long nk, int nx, int ny, int nxs, int nys) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const long i = idx % nk, base = idx - i;
    const int ixx = int(i % nx), j = int(i / nx);
    const int sxx = ((ixx + nxs) % nx + nx) % nx, syy = ((j + nys) % ny + ny) % ny;
    t[idx] = f[base + long(syy) * nx + sxx];
}

// ---------------------------------------------------------------- co-moving field (src/sim_fields.cpp:76-150)
// evt: eigenvalue per slot s of the free-hand y grid (slot 0 unused); null: Fourier y, k_y^2
__global__ void k_comoving(SpecGeom g, const cx* e1, const double* evt, const double* hc, const double* wl,
                           const double* wu, double beta, double gm2, cx* cp, cx* dp) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= g.nk) return;
    const int s = int(i / g.nx), ixx = int(i % g.nx);
    const double ev = evt ? evt[s] : g.ky[s] * g.ky[s];
    cx cprev = mk(0), dprev = mk(0);
    for (int l = 0; l < g.ns; ++l) {
        const long o = long(l) * g.nk + i;
        const int n = g.a + l;
        const double hm = hc[n + 1], hp = hc[n + 2];
        const double lo = -gm2 * 2 / ((hm + hp) * hm), up = -gm2 * 2 / ((hm + hp) * hp);
        const double dg = g.kx[ixx] * g.kx[ixx] + ev + gm2 * 2 / (hm * hp);
        if ((evt && s == 0) || dg == 0) { cp[o] = mk(0); dp[o] = mk(0); continue; }
        const cx rhs = (1 / beta) * (wl[n] * e1[o] + wu[n] * e1[o + g.nk]);
        if (n == 0) { cprev = mk(0); dprev = mk(0); }
        const cx den = mk(dg) - lo * cprev;
        // complex division in the reference's form (src/portable.hpp:34-37)
        const double dd = den.re * den.re + den.im * den.im;
        const cx u = mk(up);
        cp[o] = mk((u.re * den.re + u.im * den.im) / dd, (u.im * den.re - u.re * den.im) / dd);
        const cx num = rhs - lo * dprev;
        dp[o] = mk((num.re * den.re + num.im * den.im) / dd, (num.im * den.re - num.re * den.im) / dd);
        cprev = cp[o]; dprev = dp[o];
    }
    cx next = mk(0);
    for (int l = g.ns - 1; l >= 0; --l) {
        const long o = long(l) * g.nk + i;
        next = dp[o] - cp[o] * next;
        dp[o] = next;
    }
}
__global__ void k_comoving_R(SpecGeom g, const cx* psi, double beta, cx* rx, cx* ry, cx* ez, cx* bz, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % g.nk;
    const int j = int(i / g.nx), ixx = int(i % g.nx);
    rx[idx] = -(1 + beta) * (ik(g.kx[ixx]) * psi[idx]);
    ry[idx] = -(1 + beta) * (ik(g.ky[j]) * psi[idx]);
    ez[idx] = mk(0); bz[idx] = mk(0);
}

__global__ void k_sub(cx* a, const cx* b, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = a[i] - b[i];
}
// dst = src transposed, plane by plane: dst[p][c][r] = src[p][r][c] (rows x cols complex per plane), through LDS
// tiles of 32 x 32 so that both the reads and the writes run along rows of 32 elements
__global__ void __launch_bounds__(256) k_transpose(const cx* src, cx* dst, int rows, int cols, int ntr, int ntc) {
    __shared__ cx tile[32][33];
    const long per = long(ntr) * ntc;
    const long p = blockIdx.x / per;
    const int t = int(blockIdx.x - p * per), tr = t / ntc, tc = t - tr * ntc;
    const long off = p * long(rows) * cols;
    const int x = threadIdx.x & 31, y = threadIdx.x >> 5;       // 32 x 8 lanes
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int r = tr * 32 + y + k, c = tc * 32 + x;
        if (r < rows && c < cols) tile[y + k][x] = src[off + long(r) * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
        const int c = tc * 32 + y + k, r = tr * 32 + x;
        if (r < rows && c < cols) dst[off + long(c) * rows + r] = tile[x][y + k];
    }
}
__global__ void k_copyc(cx* a, const cx* b, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = b[i];
}
__global__ void k_zeroc(cx* a, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a[i] = mk(0);
}
__global__ void k_axpy(cx* y, const cx* x, double s, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = s > 0 ? y[i] + x[i] : y[i] - x[i];
}
}  // namespace

#define G3D_LAUNCH(kern, n, ...)                                                                             \
    do {                                                                                                     \
        const long n_ = (n);                                                                                 \
        if (n_ > 0) {                                                                                        \
            hipLaunchKernelGGL(kern, dim3(blocks_for(n_, kT)), dim3(kT), 0, stream(), __VA_ARGS__);          \
            launch_check(#kern);                                                                             \
        }                                                                                                    \
    } while (0)

void currents_zfilter(const SpecGeom& g, const double* jw, const double* hcg, int zf, int zu, cx* S1, cx* jkz) {
    G3D_LAUNCH(k_zfilter, long(g.ns + 1) * g.nk, g, jw, hcg, zf, zu, S1, jkz);
}
void currents_unpack(const SpecGeom& g, const cx* S1, cx* jkx, cx* jky, cx* jkz, int keep0) {
    G3D_LAUNCH(k_cur_unpack, long(g.ns + 1) * g.nk, g, S1, jkx, jky, jkz, keep0);
}
void sponge_pack(const SpecGeom& g, const cx* Rx, const cx* Ry, const cx* jkz, const double* hcg, const double* wlo,
                 const double* whi, double g2, double opb, cx* S1, cx* S2) {
    if (g.cm) {
        if (g.nx % kTT || g.ny % kTT) throw std::runtime_error("sponge: column-major modes need nx, ny multiples of 16");
        const long nb = long(g.ns) * (g.nx / kTT) * (g.ny / kTT);
        hipLaunchKernelGGL(k_sponge_pack_cm, dim3(unsigned(nb)), dim3(256), 0, stream(), g, Rx, Ry, jkz, hcg, wlo, whi,
                           g2, opb, S1, S2);
        launch_check("sponge pack (column-major)");
        return;
    }
    G3D_LAUNCH(k_sponge_pack, long(g.ns) * g.nk, g, Rx, Ry, jkz, hcg, wlo, whi, g2, opb, S1, S2);
}
void sponge_mask(const SpecGeom& g, cx* S1, const double* mx, const double* mn, const double* mh, int count) {
    G3D_LAUNCH(k_sponge_mask, long(count) * g.nk, g, S1, mx, mn, mh, long(count) * g.nk);
}
void fields_pack3(const SpecGeom& g, const cx* S1, const cx* S2, double opb, cx* Rx, cx* Ry, const cx* Lx,
                  const cx* Ly, const cx* Ez, const cx* Bz, cx* T, long ts, int count, cxf* Tf) {
    if (Tf && !g.cm) throw std::runtime_error("pack: FP32 planes need the column-major modes");
    if (g.cm) {
        if (g.nx % kTT || g.ny % kTT) throw std::runtime_error("pack: column-major modes need nx, ny multiples of 16");
        const long nb = long(count) * (g.nx / kTT) * (g.ny / kTT);
        if (nb <= 0) return;
        if (Tf)
            hipLaunchKernelGGL(k_pack3_cm<cxf>, dim3(unsigned(nb)), dim3(256), 0, stream(), g, S1, S2, opb, Rx, Ry, Lx,
                               Ly, Ez, Bz, Tf, ts);
        else
            hipLaunchKernelGGL(k_pack3_cm<cx>, dim3(unsigned(nb)), dim3(256), 0, stream(), g, S1, S2, opb, Rx, Ry, Lx,
                               Ly, Ez, Bz, T, ts);
        launch_check("fields pack (column-major)");
        return;
    }
    G3D_LAUNCH(k_pack3, long(count) * g.nk, g, S1, S2, opb, Rx, Ry, Lx, Ly, Ez, Bz, T, ts, long(count) * g.nk);
}
void fields_to_F(const cx* T, long ts, float* F, long fpl, long nk, int count) {
    G3D_LAUNCH(k_to_F, long(count) * nk, T, ts, F, fpl, long(count) * nk);
}
void fields_to_F_f(const cxf* T, long ts, float* F, long fpl, long nk, int count) {
    G3D_LAUNCH(k_to_F_f, long(count) * nk, T, ts, F, fpl, long(count) * nk);
}
void sspec_pack(double* J, long nk, int nx, int ny, cx* pk, double* rl, double* rli, int i0, int j0,
                const TileGeom& tg) {
    G3D_LAUNCH(k_sspec_pack, nk, J, nk, nx, ny, pk, rl, rli, i0, j0, tg);
}
void sspec_unpack(const SpecGeom& g, const cx* pk, cx* jsn, double a2, int keep0) {
    G3D_LAUNCH(k_sspec_unpack, g.nk, g, pk, jsn, a2, keep0);
}
void grid_laser_fill(const GridLaserArgs& a, cx* T, int n0, int count) {
    G3D_LAUNCH(k_glaser_fill, long(count) * a.nk, a, T, n0, long(count) * a.nk);
}
void grid_laser_add(const SpecGeom& g, const cx* T, cx* Rx, cx* Ry, int count, int fh) {
    G3D_LAUNCH(k_glaser_add, long(count) * g.nk, g, T, Rx, Ry, long(count) * g.nk, fh);
}
void rotate_planes(cx* A, long n, long nk, int nx, const double* kxs, const double* kys, double sx, double sy, int ys) {
    G3D_LAUNCH(k_rotate, n, A, n, nk, nx, kxs, kys, sx, sy, ys);
}
void roll_F(const float* F, float* out, long n, long nk, int nx, int ny, int nxs, int nys) {
    G3D_LAUNCH(k_roll, n, F, out, n, nk, nx, ny, nxs, nys);
}
void comoving_solve(const SpecGeom& g, const cx* jkz, const double* hcg, const double* wlo, const double* whi,
                    double beta, double gm2, cx* cp, cx* dp) {
    G3D_LAUNCH(k_comoving, g.nk, g, jkz, (const double*)nullptr, hcg, wlo, whi, beta, gm2, cp, dp);
}
void comoving_solve_ev(const SpecGeom& g, const cx* e1, const double* ev, const double* hcg, const double* wlo,
                       const double* whi, double beta, double gm2, cx* cp, cx* dp) {
    G3D_LAUNCH(k_comoving, g.nk, g, e1, ev, hcg, wlo, whi, beta, gm2, cp, dp);
}
void comoving_R(const SpecGeom& g, const cx* psi, double beta, cx* Rx, cx* Ry, cx* Ez, cx* Bz) {
    G3D_LAUNCH(k_comoving_R, long(g.ns) * g.nk, g, psi, beta, Rx, Ry, Ez, Bz, long(g.ns) * g.nk);
}
void sub_planes(cx* a, const cx* b, long n) { G3D_LAUNCH(k_sub, n, a, b, n); }
void copy_planes(cx* dst, const cx* src, long n) { G3D_LAUNCH(k_copyc, n, dst, src, n); }
void transpose_planes(const cx* src, cx* dst, long planes, int rows, int cols) {
    if (planes <= 0) return;
    const int ntr = (rows + 31) / 32, ntc = (cols + 31) / 32;
    const long nb = planes * ntr * ntc;
    if (nb > 0x7fffffffL) throw std::runtime_error("transpose: too many tiles");
    hipLaunchKernelGGL(k_transpose, dim3(unsigned(nb)), dim3(256), 0, stream(), src, dst, rows, cols, ntr, ntc);
    launch_check("transpose");
}
void zero_planes(cx* a, long n) { G3D_LAUNCH(k_zeroc, n, a, n); }
void axpy_planes(cx* y, const cx* x, double s, long n) { G3D_LAUNCH(k_axpy, n, y, x, s, n); }

}  // namespace g3d
