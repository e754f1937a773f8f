// The per-mode work of one pass of a plasma half-cell, fused into one kernel (one lane per mode of the plane):
//   unpack + filter the spectra of the streaming current (src/sim_plasma.cpp:105-115)
//   replace the streaming part of jk(n) and the extrapolated prediction of jk(n-1) (src/sim_plasma.cpp:255-351)
//   march slice n with the pass backup of R, E_z, B_z (src/sim_plasma.cpp:306-326, 352)
//   on the last pass the correction of R at slice n+1 and the level-s0 current (src/sim_plasma.cpp:381-396)
//   pack R_x + i R_y, L_x + i L_y, E_z + i B_z of slice n for the inverse transforms (src/sim_plasma.cpp:205-209)
// In the reference these are about a dozen kernels with a host synchronisation after most of them.
#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {

__global__ void __launch_bounds__(256) k_plasma_modes(ModeArgs m) {
    const MarchView& c = m.mv;
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= c.nk) return;
    const long nk = c.nk;
    const int n = m.n;
    const long hp = long(n + 1) * nk + i;           // half-cell n   (plane n + 1 of jk, js0)
    const long hq = long(n) * nk + i;               // half-cell n-1 (plane n)
    cx* jk[3] = {m.jkx, m.jky, m.jkz};
    cx* js0[3] = {m.js0x, m.js0y, m.js0z};
    // ------------------------------------------------ start of the half-cell (first pass)
    if (m.first) {
        if (m.top) {
            // front: half-cell n-1 = N-2 predicted with its level-s0 value (the prediction buffer starts at zero)
            for (int k = 0; k < 3; ++k) {
                const cx t = js0[k][hq];
                jk[k][hq] = jk[k][hq] + t - m.pin[k * nk + i];
                m.pin[k * nk + i] = t;
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                const cx p = m.pin[k * nk + i];
                m.sin_[k * nk + i] = p;
                m.si0[k * nk + i] = p;
            }
            if (n >= 1)
                for (int k = 0; k < 3; ++k) {
                    const cx t = js0[k][hq];
                    jk[k][hq] = jk[k][hq] + t - mk(0);
                    m.pin[k * nk + i] = t;
                }
        }
    }
    // ------------------------------------------------ the streaming current of this pass
    cx u[3];
    if (!m.top) {
        const int ixx = int(i % c.nx), j = int(i / c.nx);
        const long mi = long((m.ny - j) % m.ny) * c.nx + (c.nx - ixx) % c.nx;
        const cx z = m.pk[i], zm = conjg(m.pk[mi]);
        const double k2 = m.kxs[ixx] * m.kxs[ixx] + m.kys[j] * m.kys[j];
        const double f = (ixx == 0 && j == 0 && !m.keep0) ? 0.0 : m.filt[ixx] * m.filty[j] / (1 + m.a2 * k2);
        const cx d = z - zm;
        u[0] = f * (0.5 * (z + zm));
        u[1] = f * (0.5 * mk(d.im, -d.re));
        u[2] = f * m.pk[nk + i];
        for (int k = 0; k < 3; ++k) {
            m.jsn[k * nk + i] = u[k];
            const cx s = m.sin_[k * nk + i];
            jk[k][hp] = jk[k][hp] + u[k] - s;
            m.sin_[k * nk + i] = u[k];
        }
        if (n >= 1 && m.extrap)
            for (int k = 0; k < 3; ++k) {
                const cx d0 = u[k] - js0[k][hp];
                const cx t = js0[k][hq] + d0 + m.wex * (d0 - m.dl[k * nk + i]);
                jk[k][hq] = jk[k][hq] + t - m.pin[k * nk + i];
                m.pin[k * nk + i] = t;
            }
    }
    // ------------------------------------------------ march slice n
    const long o = long(n - c.a) * nk + i;
    Mode6 old;
    if (!m.first && m.bRx) { old.Rx = m.bRx[i]; old.Ry = m.bRy[i]; old.Ez = m.bEz[i]; old.Bz = m.bBz[i]; }
    else {
        old.Rx = c.Rx[o]; old.Ry = c.Ry[o]; old.Ez = c.Ez[o]; old.Bz = c.Bz[o];
        if (m.bRx) { m.bRx[i] = old.Rx; m.bRy[i] = old.Ry; m.bEz[i] = old.Ez; m.bBz[i] = old.Bz; }
    }
    Mode6 up{};
    const long on = o + nk;
    if (n < c.N - 1) up = Mode6{c.Rx[on], c.Ry[on], c.Ez[on], c.Bz[on], c.Lx[on], c.Ly[on]};
    const Mode6 r = march_mode(c, i, n, old, up);
    c.Rx[o] = r.Rx; c.Ry[o] = r.Ry; c.Ez[o] = r.Ez; c.Bz[o] = r.Bz; c.Lx[o] = r.Lx; c.Ly[o] = r.Ly;
    // ------------------------------------------------ half-cell n final: correction of R at slice n+1
    if (m.last && !m.top) {
        const cx s0x = m.si0[i], s0y = m.si0[nk + i];
        c.Rx[on] = c.Rx[on] - m.qcorr * (u[0] - s0x);
        c.Ry[on] = c.Ry[on] - m.qcorr * (u[1] - s0y);
        for (int k = 0; k < 3; ++k) {
            m.dl[k * nk + i] = u[k] - js0[k][hp];
            js0[k][hp] = u[k];
        }
    }
    // ------------------------------------------------ packed planes for the real fields of slice n
    m.T[i] = mk(r.Rx.re - r.Ry.im, r.Rx.im + r.Ry.re);
    m.T[nk + i] = mk(r.Lx.re - r.Ly.im, r.Lx.im + r.Ly.re);
    m.T[2 * nk + i] = mk(r.Ez.re - r.Bz.im, r.Ez.im + r.Bz.re);
}
}  // namespace

void plasma_modes(const ModeArgs& m) {
    hipLaunchKernelGGL(k_plasma_modes, dim3(blocks_for(m.mv.nk, 256)), dim3(256), 0, stream(), m);
    launch_check("plasma modes");
}

}  // namespace g3d
