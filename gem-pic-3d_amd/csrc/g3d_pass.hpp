// The per-point bookkeeping of one pass of a plasma half-cell, shared by the Fourier-y mode kernel
// (g3d_plasma.hip) and the free-hand-y line kernel (g3d_freehand.hip).  u: the filtered streaming current of this
// pass at point i (3 components).
#pragma once
#include "g3d_kernels.hpp"

namespace g3d {

// Before the march of slice n (src/sim_plasma.cpp:300-351): on the first pass the prediction of half-cell n-1
// starts from its level-s0 value; the streaming part of jk(n) is replaced by u, and that of jk(n-1) by the
// extrapolated prediction s0 + change.  At the front (top) only half-cell N-2 is predicted.
__device__ __forceinline__ void pass_currents(const ModeArgs& m, long i, const cx u[3]) {
    const long nk = m.mv.nk;
    const int n = m.n;
    const long hp = long(n + 1) * nk + i;           // half-cell n   (plane n + 1 of jk, js0)
    const long hq = long(n) * nk + i;               // half-cell n-1 (plane n)
    cx* jk[3] = {m.jkx, m.jky, m.jkz};
    cx* js0[3] = {m.js0x, m.js0y, m.js0z};
    if (m.first) {
        if (m.top) {
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
    if (m.top) return;
    for (int k = 0; k < 3; ++k) {
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

// After the march of slice n on the last pass (src/sim_plasma.cpp:381-396): half-cell n is final; R at slice n+1
// is corrected for the prediction it was solved with, the change s0 -> s1 is kept for the extrapolation and u
// becomes the level-s0 current of the next step.
__device__ __forceinline__ void pass_correct(const ModeArgs& m, long i, const cx u[3]) {
    const MarchView& c = m.mv;
    const long nk = c.nk;
    const long on = long(m.n + 1 - c.a) * nk + i, hp = long(m.n + 1) * nk + i;
    cx* js0[3] = {m.js0x, m.js0y, m.js0z};
    c.Rx[on] = c.Rx[on] - m.qcorr * (u[0] - m.si0[i]);
    c.Ry[on] = c.Ry[on] - m.qcorr * (u[1] - m.si0[nk + i]);
    for (int k = 0; k < 3; ++k) {
        m.dl[k * nk + i] = u[k] - js0[k][hp];
        js0[k][hp] = u[k];
    }
}

// R_x + i R_y, L_x + i L_y, E_z + i B_z: the three planes of the inverse transforms (src/sim_plasma.cpp:205-209)
__device__ __forceinline__ void pack3(cx T[3], cx Rx, cx Ry, cx Lx, cx Ly, cx Ez, cx Bz) {
    T[0] = mk(Rx.re - Ry.im, Rx.im + Ry.re);
    T[1] = mk(Lx.re - Ly.im, Lx.im + Ly.re);
    T[2] = mk(Ez.re - Bz.im, Ez.im + Bz.re);
}

}  // namespace g3d
