// One pass of a plasma half-cell after the particles, fused into three kernels around the transverse FFT:
//
//  k_pass_rows_fwd   the streaming current of the half-cell (jsr plus the tile slabs), the density lines,
//                    (j_x + i j_y, j_z) of a row -> x transform -> the pair unpacked per row (a real row has a
//                    Hermitian x spectrum, so the unpack needs only the row itself); jsr is zeroed
//  k_pass_cols       y transforms of the three planes of a block of columns in LDS, then per mode: the filters
//                    and smoothing of the streaming current (src/sim_plasma.cpp:105-115), the replacement of the
//                    streaming part of jk(n) and the extrapolated prediction of jk(n-1) (src/sim_plasma.cpp:
//                    255-351), the march of slice n with the pass backup (src/sim_kernels.hpp:173-210), on the last
//                    pass the correction of R at slice n+1 (src/sim_plasma.cpp:381-396), and the packed planes
//                    R_x + i R_y, L_x + i L_y, E_z + i B_z, which are y-transformed back in LDS
//  k_pass_rows_inv   x inverse transforms, scaled, and E, B of slice n into the float F (src/sim_plasma.cpp:205-218)
//
// In the reference these are about a dozen kernels, two batched FFTs and a host synchronisation after most of
// them; here they are three launches with no host synchronisation.
#include <cstdlib>

#include "g3d_fft_core.hpp"
#include "g3d_pass.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
using namespace fftcore;
constexpr int kT = 256;
constexpr int kMaxModes = 4;          // modes per lane in k_pass_cols (nc * ny <= kMaxModes * kT)

// ---------------------------------------------------------------- rows forward
// tr: the spectra written column-major (mode ix ny + j) for k_colmodes; xcd: the row blocks of one XCD consecutive
__global__ void __launch_bounds__(kT) k_pass_rows_fwd(double* J, TileGeom tg, int nx, int ny, int lx, int rpw,
                                                     const cx* twx, cx* X, double* rl, double* rli, int i0, int j0,
                                                     int tr, int xcd) {
    extern __shared__ cx sm[];
    const int sp = padded(nx);
    cx* buf = sm;                                    // 2 sequences per row: j_x + i j_y, j_z
    cx* tws = sm + size_t(2 * rpw) * sp;
    load_twiddles(tws, twx, nx >> 1, false);
    const long nk = long(nx) * ny;
    const int row0 = row_block(blockIdx.x, gridDim.x, xcd) * rpw;
    const int tot = rpw * nx;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, ixx = e - r * nx, j = row0 + r;
        const long i = long(j) * nx + ixx;
        double v[5];
        for (int cc = 0; cc < 5; ++cc) v[cc] = J[cc * nk + i];
        if (tg.ntile) {
            const int tpx = ixx / tg.TX, tpy = j / tg.TY;
            const long wpl = long(tg.WX) * tg.WY;
            for (int dy = -1; dy <= 1; ++dy) {
                if ((dy == 1 && tg.nty <= 2) || (dy != 0 && tg.nty == 1)) continue;
                const int ty = wrap_index(tpy + dy, tg.nty);
                int vv = j - (ty * tg.TY - 1 - tg.M);
                vv = wrap_index(vv, ny);
                if (vv >= tg.WY) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dx == 1 && tg.ntx <= 2) || (dx != 0 && tg.ntx == 1)) continue;
                    const int tx = wrap_index(tpx + dx, tg.ntx);
                    int uu = ixx - (tx * tg.TX - 1 - tg.M);
                    uu = wrap_index(uu, nx);
                    if (uu >= tg.WX) continue;
                    const double* sl = tg.slab + (long(ty) * tg.ntx + tx) * 5 * wpl + long(vv) * tg.WX + uu;
                    for (int cc = 0; cc < 5; ++cc) v[cc] += sl[cc * wpl];
                }
            }
        }
        if (rl) {                                    // density lines y = 0 and x = 0 (src/sim_plasma.cpp:363-374)
            if (j == j0) rl[ixx] = v[3];
            if (ixx == i0) rl[nx + j] = v[3];
            if (rli) {
                if (j == j0) rli[ixx] = v[4];
                if (ixx == i0) rli[nx + j] = v[4];
            }
        }
        for (int cc = 0; cc < 5; ++cc) J[cc * nk + i] = 0;
        const int b = lpos(bitrev(ixx, lx));
        buf[(2 * r) * sp + b] = mk(v[0], v[1]);
        buf[(2 * r + 1) * sp + b] = mk(v[2], 0);
    }
    __syncthreads();
    lds_fft(buf, nx, lx, 2 * rpw, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, kx = e - r * nx, j = row0 + r;
        const long i = tr ? long(kx) * ny + j : long(j) * nx + kx;
        const cx z = buf[(2 * r) * sp + lpos(kx)], zm = conjg(buf[(2 * r) * sp + lpos((nx - kx) & (nx - 1))]);
        const cx d = z - zm;
        X[i] = 0.5 * (z + zm);
        X[nk + i] = 0.5 * mk(d.im, -d.re);
        X[2 * nk + i] = buf[(2 * r + 1) * sp + lpos(kx)];
    }
}

// ---------------------------------------------------------------- columns: transforms and the per-mode work
// kind 0: a pass of half-cell n; 1: the front slice (no current); 2: spectra only (jsn, for the stand-alone
// stream_spectra)
__global__ void __launch_bounds__(kT) k_pass_cols(ModeArgs m, const cx* X, cx* T, cx* jsn, const cx* twy, int nc,
                                                 int ly, int kind) {
    extern __shared__ cx sm[];
    const MarchView& c = m.mv;
    const int nx = c.nx, ny = m.ny, sp = padded(ny);
    const long nk = c.nk;
    cx* buf = sm;                                    // 3 nc sequences: component k of column c at (k nc + c) sp
    cx* twf = sm + size_t(3 * nc) * sp;
    cx* twi = twf + (ny >> 1);
    load_twiddles(twf, twy, ny >> 1, false);
    load_twiddles(twi, twy, ny >> 1, true);
    const int c0 = blockIdx.x * nc;
    const int nm = nc * ny;
    if (kind != 1) {
        for (int e = threadIdx.x; e < 3 * nm; e += blockDim.x) {
            const int k = e / nm, rem = e - k * nm, j = rem / nc, cc = rem - j * nc;
            buf[(k * nc + cc) * sp + lpos(bitrev(j, ly))] = X[k * nk + long(j) * nx + c0 + cc];
        }
        __syncthreads();
        lds_fft(buf, ny, ly, 3 * nc, sp, twf);
    }
    cx u[kMaxModes][3];
#pragma unroll
    for (int q = 0; q < kMaxModes; ++q) {
        const int e = threadIdx.x + q * blockDim.x;
        if (e >= nm || kind == 1) continue;
        const int j = e / nc, cc = e - j * nc, kx = c0 + cc;
        const double k2 = m.kxs[kx] * m.kxs[kx] + m.kys[j] * m.kys[j];
        const double f = (kx == 0 && j == 0 && !m.keep0) ? 0.0 : m.filt[kx] * m.filty[j] / (1 + m.a2 * k2);
        for (int k = 0; k < 3; ++k) u[q][k] = f * buf[(k * nc + cc) * sp + lpos(j)];
    }
    __syncthreads();
    if (kind == 2) {
#pragma unroll
        for (int q = 0; q < kMaxModes; ++q) {
            const int e = threadIdx.x + q * blockDim.x;
            if (e >= nm) continue;
            const int j = e / nc, cc = e - j * nc;
            const long i = long(j) * nx + c0 + cc;
            for (int k = 0; k < 3; ++k) jsn[k * nk + i] = u[q][k];
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < kMaxModes; ++q) {
        const int e = threadIdx.x + q * blockDim.x;
        if (e >= nm) continue;
        const int j = e / nc, cc = e - j * nc;
        const long i = long(j) * nx + c0 + cc;
        cx t[3];
        mode_work<false>(m, i, u[q], t);
        const int b = lpos(bitrev(j, ly));
        for (int k = 0; k < 3; ++k) buf[(k * nc + cc) * sp + b] = t[k];
    }
    __syncthreads();
    lds_fft(buf, ny, ly, 3 * nc, sp, twi);
    for (int e = threadIdx.x; e < 3 * nm; e += blockDim.x) {
        const int k = e / nm, rem = e - k * nm, j = rem / nc, cc = rem - j * nc;
        T[k * nk + long(j) * nx + c0 + cc] = buf[(k * nc + cc) * sp + lpos(j)];
    }
}

// ---------------------------------------------------------------- the per-mode work on its own
// X: the three planes after both forward transforms; T: the packed planes before the inverse y transforms.  One
// lane per mode (kind as k_pass_cols).
__global__ void __launch_bounds__(kT) k_pass_modes(ModeArgs m, const cx* X, cx* T, cx* jsn, int kind) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nk = m.mv.nk;
    if (i >= nk) return;
    int kx, j;
    mode_xy(m.mv, i, kx, j);
    cx u[3] = {mk(0), mk(0), mk(0)};
    if (kind != 1) {
        const double k2 = m.kxs[kx] * m.kxs[kx] + m.kys[j] * m.kys[j];
        const double f = (kx == 0 && j == 0 && !m.keep0) ? 0.0 : m.filt[kx] * m.filty[j] / (1 + m.a2 * k2);
        for (int k = 0; k < 3; ++k) u[k] = f * X[k * nk + i];
        if (kind == 2) {
            for (int k = 0; k < 3; ++k) jsn[k * nk + i] = u[k];
            return;
        }
    }
    cx t[3];
    mode_work<false>(m, i, u, t);
    for (int k = 0; k < 3; ++k) T[k * nk + i] = t[k];
}

// Columns + modes + inverse columns with the modes column-major (mode ix ny + j, the layout k_pass_rows_fwd writes
// with tr): one workgroup per k_x column, one lane per mode.  The lane issues every read of its mode's state first
// (mode_load, in flight during the forward y transforms), then the same operations as k_fft_cols + k_pass_modes +
// k_fft_cols in the same order (the same transform core and twiddles), so the results are those of the split pass.
constexpr int kColMax = 1024;
template <int BS>
__global__ void __launch_bounds__(BS) k_colmodes(ModeArgs m, const cx* X, cx* T, int ny, int ly, const cx* twy, int kind) {
    extern __shared__ cx sm[];
    const int sp = padded(ny);
    cx* buf = sm;                                    // the three planes of the column
    cx* twf = sm + 3 * sp;
    cx* twi = twf + (ny >> 1);
    load_twiddles(twf, twy, ny >> 1, false);
    load_twiddles(twi, twy, ny >> 1, true);
    const long nk = m.mv.nk;
    const int kx = blockIdx.x;
    const long c0 = long(kx) * ny;
    const int jj = threadIdx.x;                      // this lane's mode (kx, jj)
    const long i = c0 + jj;
    ModeLoad L;
    if constexpr (BS <= 256) L = mode_load<false>(m, i);
    if (kind != 1)
        for (int e = jj; e < 3 * ny; e += blockDim.x) {
            const int p = e / ny, j = e - p * ny;
            buf[p * sp + lpos(bitrev(j, ly))] = X[p * nk + c0 + j];
        }
    __syncthreads();
    if (kind != 1) lds_fft(buf, ny, ly, 3, sp, twf);
    cx u[3] = {mk(0), mk(0), mk(0)};
    if (kind != 1) {
        const double k2 = m.kxs[kx] * m.kxs[kx] + m.kys[jj] * m.kys[jj];
        const double f = (kx == 0 && jj == 0 && !m.keep0) ? 0.0 : m.filt[kx] * m.filty[jj] / (1 + m.a2 * k2);
        for (int k = 0; k < 3; ++k) u[k] = f * buf[k * sp + lpos(jj)];
    }
    cx t[3];
    if constexpr (BS <= 256) mode_compute<false>(m, i, L, u, t);
    else mode_work<false>(m, i, u, t);
    __syncthreads();                                 // every lane has read its X before T overwrites the buffer
    const int b = lpos(bitrev(jj, ly));
    for (int k = 0; k < 3; ++k) buf[k * sp + b] = t[k];
    __syncthreads();
    lds_fft(buf, ny, ly, 3, sp, twi);
    for (int k = 0; k < 3; ++k) T[k * nk + i] = buf[k * sp + lpos(jj)];
}

// ---------------------------------------------------------------- rows inverse and the real fields of slice n
// tr: T column-major (mode ix ny + j); xcd as k_pass_rows_fwd
__global__ void __launch_bounds__(kT) k_pass_rows_inv(const cx* T, float* Fn, long fpl, int nx, int ny, int lx, int rpw,
                                                     const cx* twx, double scale, int tr, int xcd) {
    extern __shared__ cx sm[];
    const int sp = padded(nx);
    cx* buf = sm;                                    // 3 sequences per row: R, L, E_z/B_z packed
    cx* tws = sm + size_t(3 * rpw) * sp;
    load_twiddles(tws, twx, nx >> 1, true);
    const long nk = long(nx) * ny;
    const int row0 = row_block(blockIdx.x, gridDim.x, xcd) * rpw;
    const int tot = rpw * nx;
    for (int e = threadIdx.x; e < 3 * tot; e += blockDim.x) {
        const int k = e / tot, rem = e - k * tot, r = rem / nx, ixx = rem - r * nx;
        const long i = tr ? long(ixx) * ny + row0 + r : long(row0 + r) * nx + ixx;
        buf[(3 * r + k) * sp + lpos(bitrev(ixx, lx))] = T[k * nk + i];
    }
    __syncthreads();
    lds_fft(buf, nx, lx, 3 * rpw, sp, tws);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int r = e / nx, ixx = e - r * nx;
        const long i = long(row0 + r) * nx + ixx;
        const cx R = scale * buf[(3 * r) * sp + lpos(ixx)], L = scale * buf[(3 * r + 1) * sp + lpos(ixx)];
        const cx Z = scale * buf[(3 * r + 2) * sp + lpos(ixx)];
        const double r1 = R.re, r2 = R.im, l1 = L.re, l2 = L.im;
        Fn[0 * fpl + i] = float(0.5 * (r1 + l1));
        Fn[1 * fpl + i] = float(0.5 * (r2 + l2));
        Fn[3 * fpl + i] = float(0.5 * (l2 - r2));
        Fn[4 * fpl + i] = float(0.5 * (r1 - l1));
        Fn[2 * fpl + i] = float(Z.re);
        Fn[5 * fpl + i] = float(Z.im);
    }
}

int cols_per_wg(int nx, int ny) {
    static const int forced = [] { const char* e = getenv("G3D_COLS_PER_WG"); return e ? atoi(e) : 0; }();
    if (forced > 0 && nx % forced == 0 && forced * ny <= kMaxModes * kT) return forced;
    int nc = 1;
    while (nc < 8 && 2 * nc * ny <= kMaxModes * kT && nx / (2 * nc) >= 128) nc <<= 1;
    while (nc * ny > kMaxModes * kT && nc > 1) nc >>= 1;
    return nc;
}
int lanes(int work) { int t = 64; while (t < work && t < kT) t <<= 1; return t; }
}  // namespace

// rows per workgroup of the row kernels: at least 256 workgroups (one per CU) before rows are grouped
int rows_per_wg(int ny) {
    static const int forced = [] { const char* e = getenv("G3D_ROWS_PER_WG"); return e ? atoi(e) : 0; }();
    if (forced > 0 && ny % forced == 0) return forced;
    int r = 1;
    while (r < 8 && ny / (2 * r) >= 256) r <<= 1;
    return r;
}
// the row blocks of one XCD consecutive in the row kernels whose spectra are column-major, and the k_x lines of one
// XCD consecutive in the free-hand line kernels (G3D_XCD_ROWS=0: off)
int xcd_rows() {
    static const int on = [] { const char* e = getenv("G3D_XCD_ROWS"); return e ? atoi(e) : 1; }();
    return on;
}
void pass_rows_fwd(double* J, const TileGeom& tg, const FftPlan& p, cx* X, double* rl, double* rli, int i0, int j0,
                   int tr) {
    const int rpw = rows_per_wg(p.ny);
    if (p.ny % rpw) throw std::runtime_error("pass rows: ny not a multiple of the row tile");
    const size_t lds = (size_t(2 * rpw) * (p.nx + p.nx / 8) + p.nx / 2) * sizeof(cx);
    hipLaunchKernelGGL(k_pass_rows_fwd, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), lds, stream(), J, tg, p.nx, p.ny, p.lx,
                       rpw, p.twx, X, rl, rli, i0, j0, tr, tr && xcd_rows());
    launch_check("pass rows forward");
}
void pass_colmodes(const ModeArgs& m, const FftPlan& p, const cx* X, cx* T, int kind) {
    if (p.ny > kColMax || p.ny < 2 || !m.mv.cm) throw std::runtime_error("plasma pass: column kernel needs 2 <= ny <= 1024, modes column-major");
    const size_t lds = (size_t(3) * (p.ny + p.ny / 8) + p.ny) * sizeof(cx);
    auto* kc = p.ny <= 256 ? k_colmodes<256> : k_colmodes<kColMax>;
    hipLaunchKernelGGL(kc, dim3(unsigned(p.nx)), dim3(unsigned(p.ny)), lds, stream(), m, X, T, p.ny, p.ly, p.twy, kind);
    launch_check("plasma pass (column-major modes)");
}
void pass_cols(const ModeArgs& m, const FftPlan& p, const cx* X, cx* T, cx* jsn, int kind) {
    const int nc = cols_per_wg(p.nx, p.ny);
    if (nc * p.ny > kMaxModes * kT) throw std::runtime_error("pass columns: ny above " + std::to_string(kMaxModes * kT));
    const size_t lds = (size_t(3 * nc) * (p.ny + p.ny / 8) + p.ny) * sizeof(cx);
    hipLaunchKernelGGL(k_pass_cols, dim3(p.nx / nc), dim3(lanes(nc * p.ny)), lds, stream(), m, X, T, jsn, p.twy, nc, p.ly,
                       kind);
    launch_check("pass columns");
}
void pass_modes(const ModeArgs& m, const FftPlan& p, cx* X, cx* T, cx* jsn, int kind) {
    if (kind != 1) fft_cols(p, X, 3, false, 1.0);
    hipLaunchKernelGGL(k_pass_modes, dim3(blocks_for(m.mv.nk, kT)), dim3(kT), 0, stream(), m, X, T, jsn, kind);
    launch_check("pass modes");
    if (kind != 2) fft_cols(p, T, 3, true, 1.0);
}
void pass_rows_inv(const cx* T, float* Fn, long fpl, const FftPlan& p, int tr) {
    const int rpw = rows_per_wg(p.ny);
    const size_t lds = (size_t(3 * rpw) * (p.nx + p.nx / 8) + p.nx / 2) * sizeof(cx);
    hipLaunchKernelGGL(k_pass_rows_inv, dim3(p.ny / rpw), dim3(lanes(rpw * p.nx)), lds, stream(), T, Fn, fpl, p.nx, p.ny,
                       p.lx, rpw, p.twx, p.two_d ? 1.0 / (double(p.nx) * p.ny) : 1.0 / p.nx, tr, tr && xcd_rows());
    launch_check("pass rows inverse");
}

}  // namespace g3d
