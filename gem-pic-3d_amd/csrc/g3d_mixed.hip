// The FP32-mixed form of the plasma pass (Fourier y, electron streams): the deposit, the transforms and the
// per-mode implicit update run in FP32 on the vector ALUs.  FP64 is kept for
//   - the particle coordinates x, y, zeta, z and their accumulation along s (the increments are FP32),
//   - the trapezoidal s-integration of R (R and its explicit part R + D (F - J)),
//   - the running currents jk, j0, js0 and the prediction bookkeeping (the charge totals),
// and the spectra stay FP64 in HBM, so the FP64 and mixed passes share every array outside the pass.
//
//   k_stream_mixed      one half-cell pass of the streaming electrons (src/sim_stream.hpp:104-171): gather and
//                       deposit through the tile's LDS windows in FP32, FP32 Boris rotations
//   k_rows_fwd_mixed    jsr + slabs (float) -> x transforms in FP32 -> pair unpacked per row
//   k_cols_mixed        y transforms of three planes in FP32 (forward or inverse)
//   k_modes_mixed       filters, bookkeeping (FP64), march with the FP32 implicit solve, packed planes
//   k_rows_inv_mixed    x inverse transforms in FP32 and E, B of the slice into F
#include "g3d_fft_core.hpp"
#include "g3d_pass.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
using namespace fftcore;
constexpr int kT = 256;

// TSC weights in FP32 on a periodic uniform axis: the cell from the FP64 index coordinate, the fraction in FP32
__device__ __forceinline__ void tsc_f(double x, double d, int n, int idx[3], float w[3]) {
    const double xi = x / d;
    const int i0 = int(floor(xi + 0.5));
    const float f = float(xi - i0);
    const float a = 0.5f - f, b = 0.5f + f;
    w[0] = 0.5f * a * a;
    w[1] = 0.75f - f * f;
    w[2] = 0.5f * b * b;
    idx[0] = i0 - 1; idx[1] = i0; idx[2] = i0 + 1;
#pragma unroll
    for (int k = 0; k < 3; ++k) { const int m = idx[k] % n; idx[k] = m < 0 ? m + n : m; }
}

__device__ __forceinline__ void boris_f(float& ux, float& uy, float& uz, const float E[3], const float B[3], float qm,
                                        float dt) {
    const float qh = 0.5f * qm * dt;
    const float ax = ux + qh * E[0], ay = uy + qh * E[1], az = uz + qh * E[2];
    const float gm = sqrtf(1.0f + ax * ax + ay * ay + az * az);
    const float tx = qh * B[0] / gm, ty = qh * B[1] / gm, tz = qh * B[2] / gm;
    const float t2 = tx * tx + ty * ty + tz * tz;
    const float px = ax + (ay * tz - az * ty), py = ay + (az * tx - ax * tz), pz = az + (ax * ty - ay * tx);
    const float sf = 2.0f / (1.0f + t2);
    const float bx = ax + sf * (py * tz - pz * ty), by = ay + sf * (pz * tx - px * tz), bz = az + sf * (px * ty - py * tx);
    ux = bx + qh * E[0]; uy = by + qh * E[1]; uz = bz + qh * E[2];
}

// LDS windows of a tile: the 5 current planes (FP64 sums of the FP32 contributions: the FP32 LDS atomic add
// serialises on the same-address conflicts of a tile's particles, the FP64 one does not) and the 12 field planes
// (slices fl, fl+1, float)
struct WinF {
    double* lds;
    const float* fld;
    int x0, y0, WX, WY, fl;
};

// E, B at a point: the tile's field window when the stencil and the slice pair are in it, else F
__device__ __forceinline__ void gather_f(const FieldView& c, const WinF* w, double x, double y, double zeta, float E[3],
                                         float B[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) { E[k] = 0; B[k] = 0; }
    if (!c.self) return;
    int n; double fz;
    c.zax.locate(zeta, n, fz);
    if (n == c.a - 1 && fz > 1 - 1e-9) { n = c.a; fz = 0; }
    const int fl = n - c.a;
    if (fl < 0 || fl > c.ns) return;
    int ix[3], jn[3]; float wx[3], wn[3];
    tsc_f(x, c.dx, c.nx, ix, wx);
    tsc_f(y, c.dy, c.ny, jn, wn);
    const float f1 = float(fz), f0 = 1.0f - f1;
    float wt[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) wt[3 * j + i] = wx[i] * wn[j];
    float v[6];
    if (w && w->fld && fl == w->fl) {
        int u0 = ix[0] - w->x0; u0 = ((u0 % c.nx) + c.nx) % c.nx;
        int v0 = jn[0] - w->y0; v0 = ((v0 % c.ny) + c.ny) % c.ny;
        if (u0 + 2 < w->WX && v0 + 2 < w->WY) {
            const int pl = w->WX * w->WY;
#pragma unroll
            for (int comp = 0; comp < 6; ++comp) {
                const float* q0 = w->fld + comp * pl;
                const float* q1 = w->fld + (6 + comp) * pl;
                float s = 0;
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        const int k = (v0 + j) * w->WX + u0 + i;
                        s += wt[3 * j + i] * (f0 * q0[k] + f1 * q1[k]);
                    }
                v[comp] = s;
            }
            E[0] = v[0]; E[1] = v[1]; E[2] = v[2]; B[0] = v[3]; B[1] = v[4]; B[2] = v[5];
            return;
        }
    }
    const long cs = long(c.ns + 2) * c.nk;
    const float* p0 = c.F + long(fl) * c.nk;
#pragma unroll
    for (int comp = 0; comp < 6; ++comp) {
        const float* q0 = p0 + comp * cs;
        const float* q1 = q0 + c.nk;
        float s = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const long k = long(jn[j]) * c.nx + ix[i];
                s += wt[3 * j + i] * (f0 * q0[k] + f1 * q1[k]);
            }
        v[comp] = s;
    }
    E[0] = v[0]; E[1] = v[1]; E[2] = v[2]; B[0] = v[3]; B[1] = v[4]; B[2] = v[5];
}

// TSC deposit of f v into the FP32 current planes (window or jsr), as the FP64 deposit
__device__ __forceinline__ void sdeposit_f(const StreamArgs& c, float* J, const WinF* win, double x, double y, float f,
                                           float vx, float vy, float vz) {
    const FieldView& P = c.fv;
    int ix[3], jn[3]; float wx[3], wn[3];
    tsc_f(x, P.dx, P.nx, ix, wx);
    tsc_f(y, P.dy, P.ny, jn, wn);
    const long nk = P.nk;
    const int cq = f < 0 ? 3 : 4;
    const bool mv = vx != 0 || vy != 0 || vz != 0;
    const float fi = f * float(c.ia[0]);
    if (win) {
        int u0 = ix[0] - win->x0; u0 = ((u0 % P.nx) + P.nx) % P.nx;
        int v0 = jn[0] - win->y0; v0 = ((v0 % P.ny) + P.ny) % P.ny;
        if (u0 + 2 < win->WX && v0 + 2 < win->WY) {
            const int wpl = win->WX * win->WY;
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    double* L = win->lds + (v0 + v) * win->WX + (u0 + u);
                    const float ww = fi * wx[u] * wn[v];
                    if (mv) {
                        atomicAdd(L, double(ww * vx));
                        atomicAdd(L + wpl, double(ww * vy));
                        atomicAdd(L + 2 * wpl, double(ww * vz));
                    }
                    atomicAdd(L + cq * wpl, double(ww));
                }
            return;
        }
    }
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            const long k = long(jn[v]) * P.nx + ix[u];
            const float ww = fi * wx[u] * wn[v];
            if (mv) {
                atomicAdd(&J[k], ww * vx);
                atomicAdd(&J[nk + k], ww * vy);
                atomicAdd(&J[2 * nk + k], ww * vz);
            }
            atomicAdd(&J[cq * nk + k], ww);
        }
}

// one electron through half-cell n (src/sim_stream.hpp:104-171), the physics of the FP64 pass in FP32
__device__ __forceinline__ void stream_one_mixed(const StreamArgs& c, float* J, const WinF* win, Parts S, Parts D,
                                                 long i) {
    const int n = c.n;
    const double zn = c.zs[n], zn1 = c.zs[n + 1], h = zn1 - zn;
    // the whole state up front, so that its loads overlap
    double x = S.x[i], y = S.y[i], zeta = S.zeta[i], z = S.z[i];
    const double pxd = S.px[i], pyd = S.py[i], pzd = S.pz[i];
    const double q = S.q[i], qmd = S.qm[i], w = S.w[i];
    int alive = S.alive[i];
    auto store = [&](double ux, double uy, double uz) {
        D.x[i] = x; D.y[i] = y; D.zeta[i] = zeta; D.px[i] = ux; D.py[i] = uy; D.pz[i] = uz;
        D.z[i] = z; D.q[i] = q; D.qm[i] = qmd; D.alive[i] = alive;
    };
    if (alive == 6) {
        if (zeta > zn) alive = 1;
        else { store(pxd, pyd, pzd); return; }
    }
    if (alive != 1) { store(pxd, pyd, pzd); return; }
    const FieldView& P = c.fv;
    const double hl = zeta >= zn1 - 1e-9 * h ? h : (zeta - zn > 0 ? zeta - zn : 0.0);
    const float hf = float(hl), qm = float(qmd);
    const float px = float(pxd), py = float(pyd), pz = float(pzd);
    const float g0 = sqrtf(1 + px * px + py * py + pz * pz);
    const float omv0 = (1 + px * px + py * py) / (g0 * (g0 + pz));
    const float dt0 = hf / omv0;
    const double xm = x + double(0.5f * dt0 * px / g0), ym = y + double(0.5f * dt0 * py / g0);
    const double zm = z + double(0.5f * dt0 * pz / g0), zetam = zeta - 0.5 * hl;
    float E[3], B[3];
    gather_f(P, win, wrap_mod(xm, P.Lx), wrap_mod(ym, P.Ly), zetam, E, B);
    if (c.laser.on) {
        double Ee[3], Be[3];
        c.laser.field(xm + c.X0, zm, zm - (zetam - c.zc), Ee, Be);
        for (int k = 0; k < 3; ++k) { E[k] += float(Ee[k]); B[k] += float(Be[k]); }
    }
    float ux = px, uy = py, uz = pz;
    boris_f(ux, uy, uz, E, B, qm, dt0);
    float g1 = sqrtf(1 + ux * ux + uy * uy + uz * uz);
    float dt = 2 * hf / (omv0 + (1 + ux * ux + uy * uy) / (g1 * (g1 + uz)));
    ux = px; uy = py; uz = pz;
    boris_f(ux, uy, uz, E, B, qm, dt);
    g1 = sqrtf(1 + ux * ux + uy * uy + uz * uz);
    float vx = 0.5f * (px / g0 + ux / g1), vy = 0.5f * (py / g0 + uy / g1), vz = 0.5f * (pz / g0 + uz / g1);
    double zeta_end = zn;
    bool trapped = false;
    if (z + double(vz * dt) >= c.sTrap) {
        const float vzc = vz > 1e-3f ? vz : 1e-3f;
        const float dtp = float((c.sTrap - z) / vzc) > 0 ? float((c.sTrap - z) / vzc) : 0.0f;
        ux = px; uy = py; uz = pz;
        boris_f(ux, uy, uz, E, B, qm, dtp);
        g1 = sqrtf(1 + ux * ux + uy * uy + uz * uz);
        vx = 0.5f * (px / g0 + ux / g1); vy = 0.5f * (py / g0 + uy / g1); vz = 0.5f * (pz / g0 + uz / g1);
        dt = dtp;
        zeta_end = zeta - double((1 - vz) * dtp);
        trapped = true;
    }
    const double xo = x, yo = y;
    x += double(vx * dt); y += double(vy * dt);
    z = trapped ? c.s1 : z + double(vz * dt);
    zeta = zeta_end;
    const float f = float(q * w * double(dt) / (c.Delta * h * P.dx));
    if (f != 0) sdeposit_f(c, J, win, 0.5 * (xo + x), 0.5 * (yo + y), f, vx, vy, vz);
    if (trapped) {
        alive = 5;
        const float vze = uz / g1;
        if (vze >= float(c.vz_beam)) sdeposit_f(c, J, win, x, y, float(q * w / (h * P.dx * vze)), ux / g1, uy / g1, vze);
    }
    store(ux, uy, uz);
}

__global__ void __launch_bounds__(kT) k_stream_mixed(StreamArgs c, float* J, float* slab, Parts S, Parts D, long n,
                                                     TileGeom tg) {
    extern __shared__ double winf[];
    const long b = blockIdx.x;
    if (b < tg.ntile) {
        const int pl = tg.WX * tg.WY;
        const int wn = 5 * pl;
        const int tx = int(b % tg.ntx), ty = int(b / tg.ntx);
        const int x0 = tx * tg.TX - 1 - tg.M, y0 = ty * tg.TY - 1 - tg.M;
        const FieldView& F = c.fv;
        const int fl = c.n - F.a;
        float* fw = reinterpret_cast<float*>(winf + wn);
        for (int k = threadIdx.x; k < wn; k += blockDim.x) winf[k] = 0;
        if (F.self)
            for (int r = threadIdx.x; r < pl; r += blockDim.x) {
                const int v = r / tg.WX, u = r - v * tg.WX;
                const int gx = ((x0 + u) % F.nx + F.nx) % F.nx, gy = ((y0 + v) % F.ny + F.ny) % F.ny;
                const float* src = F.F + long(fl) * F.nk + long(gy) * F.nx + gx;
                const long cs = long(F.ns + 2) * F.nk;
#pragma unroll
                for (int comp = 0; comp < 6; ++comp) {
                    fw[comp * pl + r] = src[comp * cs];
                    fw[(6 + comp) * pl + r] = src[comp * cs + F.nk];
                }
            }
        __syncthreads();
        WinF w{winf, F.self ? fw : nullptr, x0, y0, tg.WX, tg.WY, fl};
        const long hi = tg.off[b + 1];
        for (long k = tg.off[b] + threadIdx.x; k < hi; k += blockDim.x)
            stream_one_mixed(c, J, &w, S, D, tg.idx ? long(tg.idx[k]) : k);
        __syncthreads();
        float* sl = slab + b * long(wn);
        for (int k = threadIdx.x; k < wn; k += blockDim.x) sl[k] = float(winf[k]);
        return;
    }
    const long i = tg.n_tiled + (b - tg.ntile) * blockDim.x + threadIdx.x;
    if (i < n) stream_one_mixed(c, J, nullptr, S, D, i);
}

// ---------------------------------------------------------------- rows forward (FP32)
__global__ void __launch_bounds__(kT) k_rows_fwd_mixed(float* J, const float* slab, TileGeom tg, int nx, int ny, int lx,
                                                       int rpw, const cx* twx, cxf* X, double* rl, double* rli, int i0,
                                                       int j0) {
    extern __shared__ cxf smf[];
    const int sp = padded(nx);
    cxf* buf = smf;
    cxf* tws = smf + size_t(2 * rpw) * sp;
    load_twiddles(tws, twx, nx >> 1, false);
    const long nk = long(nx) * ny;
    const int row0 = blockIdx.x * rpw;
    const int tot = rpw * nx;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, ixx = e - r * nx, j = row0 + r;
        const long i = long(j) * nx + ixx;
        float v[5];
        for (int cc = 0; cc < 5; ++cc) v[cc] = J[cc * nk + i];
        if (tg.ntile) {
            const int tpx = ixx / tg.TX, tpy = j / tg.TY;
            const long wpl = long(tg.WX) * tg.WY;
            for (int dy = -1; dy <= 1; ++dy) {
                if ((dy == 1 && tg.nty <= 2) || (dy != 0 && tg.nty == 1)) continue;
                const int ty = ((tpy + dy) % tg.nty + tg.nty) % tg.nty;
                int vv = j - (ty * tg.TY - 1 - tg.M);
                vv = ((vv % ny) + ny) % ny;
                if (vv >= tg.WY) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dx == 1 && tg.ntx <= 2) || (dx != 0 && tg.ntx == 1)) continue;
                    const int tx = ((tpx + dx) % tg.ntx + tg.ntx) % tg.ntx;
                    int uu = ixx - (tx * tg.TX - 1 - tg.M);
                    uu = ((uu % nx) + nx) % nx;
                    if (uu >= tg.WX) continue;
                    const float* sl = slab + (long(ty) * tg.ntx + tx) * 5 * wpl + long(vv) * tg.WX + uu;
                    for (int cc = 0; cc < 5; ++cc) v[cc] += sl[cc * wpl];
                }
            }
        }
        if (rl) {
            if (j == j0) rl[ixx] = v[3];
            if (ixx == i0) rl[nx + j] = v[3];
            if (rli) {
                if (j == j0) rli[ixx] = v[4];
                if (ixx == i0) rli[nx + j] = v[4];
            }
        }
        for (int cc = 0; cc < 5; ++cc) J[cc * nk + i] = 0;
        const int b = lpos(bitrev(ixx, lx));
        buf[(2 * r) * sp + b] = mkf(v[0], v[1]);
        buf[(2 * r + 1) * sp + b] = mkf(v[2], 0);
    }
    __syncthreads();
    lds_fft(buf, nx, lx, 2 * rpw, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, kx = e - r * nx, j = row0 + r;
        const long i = long(j) * nx + kx;
        const cxf z = buf[(2 * r) * sp + lpos(kx)], zm = conjg(buf[(2 * r) * sp + lpos((nx - kx) & (nx - 1))]);
        const cxf d = z - zm;
        X[i] = 0.5f * (z + zm);
        X[nk + i] = 0.5f * mkf(d.im, -d.re);
        X[2 * nk + i] = buf[(2 * r + 1) * sp + lpos(kx)];
    }
}

// ---------------------------------------------------------------- columns (FP32), nc columns of 3 planes
__global__ void __launch_bounds__(kT) k_cols_mixed(cxf* data, int nx, int ny, int ly, int nc, const cx* tw, int inv) {
    extern __shared__ cxf smf[];
    const int sp = padded(ny);
    cxf* buf = smf;
    cxf* tws = smf + size_t(nc) * sp;
    const int tiles = nx / nc;
    const long p = blockIdx.x / tiles;
    const int c0 = (blockIdx.x % tiles) * nc;
    cxf* g = data + p * long(nx) * ny + c0;
    load_twiddles(tws, tw, ny >> 1, inv);
    const int tot = nc * ny;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int j = e / nc, c = e - j * nc;
        buf[c * sp + lpos(bitrev(j, ly))] = g[long(j) * nx + c];
    }
    __syncthreads();
    lds_fft(buf, ny, ly, nc, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int j = e / nc, c = e - j * nc;
        g[long(j) * nx + c] = buf[c * sp + lpos(j)];
    }
}

// ---------------------------------------------------------------- the per-mode work (bookkeeping FP64)
__global__ void __launch_bounds__(kT) k_modes_mixed(ModeArgs m, const cxf* X, cxf* T, int kind) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nk = m.mv.nk;
    if (i >= nk) return;
    const int kx = int(i % m.mv.nx), j = int(i / m.mv.nx);
    cx u[3] = {mk(0), mk(0), mk(0)};
    if (kind != 1) {
        const double k2 = m.kxs[kx] * m.kxs[kx] + m.kys[j] * m.kys[j];
        const float f = float((kx == 0 && j == 0 && !m.keep0) ? 0.0 : m.filt[kx] * m.filty[j] / (1 + m.a2 * k2));
        for (int k = 0; k < 3; ++k) u[k] = to_d(f * X[k * nk + i]);
    }
    cx t[3];
    mode_work<true>(m, i, u, t);
    for (int k = 0; k < 3; ++k) T[k * nk + i] = to_f(t[k]);
}

// ---------------------------------------------------------------- rows inverse (FP32) and F of the slice
__global__ void __launch_bounds__(kT) k_rows_inv_mixed(const cxf* T, float* Fn, long fpl, int nx, int ny, int lx,
                                                       int rpw, const cx* twx, float scale) {
    extern __shared__ cxf smf[];
    const int sp = padded(nx);
    cxf* buf = smf;
    cxf* tws = smf + size_t(3 * rpw) * sp;
    load_twiddles(tws, twx, nx >> 1, true);
    const long nk = long(nx) * ny;
    const int row0 = blockIdx.x * rpw;
    const int tot = rpw * nx;
    for (int e = threadIdx.x; e < 3 * tot; e += blockDim.x) {
        const int k = e / tot, rem = e - k * tot, r = rem / nx, ixx = rem - r * nx;
        buf[(3 * r + k) * sp + lpos(bitrev(ixx, lx))] = T[k * nk + long(row0 + r) * nx + ixx];
    }
    __syncthreads();
    lds_fft(buf, nx, lx, 3 * rpw, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, ixx = e - r * nx;
        const long i = long(row0 + r) * nx + ixx;
        const cxf R = scale * buf[(3 * r) * sp + lpos(ixx)], L = scale * buf[(3 * r + 1) * sp + lpos(ixx)];
        const cxf Z = scale * buf[(3 * r + 2) * sp + lpos(ixx)];
        Fn[0 * fpl + i] = 0.5f * (R.re + L.re);
        Fn[1 * fpl + i] = 0.5f * (R.im + L.im);
        Fn[3 * fpl + i] = 0.5f * (L.im - R.im);
        Fn[4 * fpl + i] = 0.5f * (R.re - L.re);
        Fn[2 * fpl + i] = Z.re;
        Fn[5 * fpl + i] = Z.im;
    }
}

int lanes(int work) { int t = 64; while (t < work && t < kT) t <<= 1; return t; }
}  // namespace

void stream_mixed(const StreamArgs& a, float* J, float* slab, Parts src, Parts dst, long nstrm, const TileGeom& tg) {
    if (nstrm <= 0) return;
    const long tail_blocks = nstrm > tg.n_tiled ? long(blocks_for(nstrm - tg.n_tiled, kT)) : 0;
    const size_t lds = tg.ntile ? size_t(tg.WX) * tg.WY * (5 * sizeof(double) + 12 * sizeof(float)) : 0;
    hipLaunchKernelGGL(k_stream_mixed, dim3(unsigned(tg.ntile + tail_blocks)), dim3(kT), lds, stream(), a, J, slab, src,
                       dst, nstrm, tg);
    launch_check("stream half-cell (mixed)");
}
void pass_mixed(const ModeArgs& m, const FftPlan& p, float* J, const float* slab, const TileGeom& tg, cxf* X, cxf* T,
                float* Fn, long fpl, double* rl, double* rli, int i0, int j0, int kind, int rpw) {
    const long nk = long(p.nx) * p.ny;
    if (kind != 1) {
        const size_t lds = (size_t(2 * rpw) * (p.nx + p.nx / 8) + p.nx / 2) * sizeof(cxf);
        hipLaunchKernelGGL(k_rows_fwd_mixed, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), lds, stream(), J, slab, tg, p.nx,
                           p.ny, p.lx, rpw, p.twx, X, rl, rli, i0, j0);
    }
    int nc = 8;
    while (nc > 4 && 3L * (p.nx / nc) < 512) nc >>= 1;
    if (nc > p.nx) nc = p.nx;
    const size_t lc = (size_t(nc) * (p.ny + p.ny / 8) + p.ny / 2) * sizeof(cxf);
    const int tc = lanes(nc * (p.ny / 8 > 0 ? p.ny / 8 : 1));
    if (kind != 1)
        hipLaunchKernelGGL(k_cols_mixed, dim3(unsigned(3 * (p.nx / nc))), dim3(tc), lc, stream(), X, p.nx, p.ny, p.ly, nc,
                           p.twy, 0);
    hipLaunchKernelGGL(k_modes_mixed, dim3(blocks_for(nk, kT)), dim3(kT), 0, stream(), m, X, T, kind);
    hipLaunchKernelGGL(k_cols_mixed, dim3(unsigned(3 * (p.nx / nc))), dim3(tc), lc, stream(), T, p.nx, p.ny, p.ly, nc,
                       p.twy, 1);
    const size_t li = (size_t(3 * rpw) * (p.nx + p.nx / 8) + p.nx / 2) * sizeof(cxf);
    hipLaunchKernelGGL(k_rows_inv_mixed, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), li, stream(), T, Fn, fpl, p.nx, p.ny,
                       p.lx, rpw, p.twx, float(1.0 / (double(p.nx) * p.ny)));
    launch_check("plasma pass (mixed)");
}

}  // namespace g3d
