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

// The per-mode work of a pass: u is the filtered streaming current of this pass.  The same
// operations as pass_currents + march_mode + pass_correct in the same order, with every value read once and kept
// in registers (the bookkeeping planes alias as far as the compiler knows, so it would reload them).
// MIXED: the implicit part of the march in FP32 (march_mode_mixed)
template <bool MIXED>
__device__ __forceinline__ void mode_work(const ModeArgs& m, long i, const cx u[3], cx T[3]) {
    const MarchView& c = m.mv;
    const long nk = c.nk;
    const int n = m.n;
    const long hp = long(n + 1) * nk + i;           // half-cell n   (plane n + 1 of jk, js0)
    const long hq = long(n) * nk + i;               // half-cell n-1 (plane n)
    cx* jk[3] = {m.jkx, m.jky, m.jkz};
    cx* js0[3] = {m.js0x, m.js0y, m.js0z};
    cx q[3], pp[3], si0[3], js0p[3];
    bool have_si0 = false, have_js0p = false;
    if (m.top) {                                    // front: half-cell N-2 predicted with its level-s0 value
        for (int k = 0; k < 3; ++k) {
            const cx t = js0[k][hq];
            q[k] = jk[k][hq] + t - m.pin[k * nk + i];
            m.pin[k * nk + i] = t;
            jk[k][hq] = q[k];
            pp[k] = jk[k][hp];
        }
    } else {
        cx pin[3], sp[3];
        for (int k = 0; k < 3; ++k) pin[k] = m.pin[k * nk + i];
        if (m.first) {
            for (int k = 0; k < 3; ++k) { sp[k] = pin[k]; si0[k] = pin[k]; m.si0[k * nk + i] = pin[k]; }
            have_si0 = true;
        } else
            for (int k = 0; k < 3; ++k) sp[k] = m.sin_[k * nk + i];
        cx t0[3];
        bool qset = false, pin_set = false;
        if (m.first && n >= 1) {
            for (int k = 0; k < 3; ++k) {
                t0[k] = js0[k][hq];
                q[k] = jk[k][hq] + t0[k] - mk(0);
                pin[k] = t0[k];
            }
            qset = pin_set = true;
        }
        for (int k = 0; k < 3; ++k) {
            pp[k] = jk[k][hp] + u[k] - sp[k];
            jk[k][hp] = pp[k];
            m.sin_[k * nk + i] = u[k];
        }
        if (n >= 1 && m.extrap) {
            for (int k = 0; k < 3; ++k) {
                js0p[k] = js0[k][hp];
                const cx d0 = u[k] - js0p[k];
                const cx jq = (m.first && n >= 1) ? t0[k] : js0[k][hq];
                const cx t = jq + d0 + m.wex * (d0 - m.dl[k * nk + i]);
                if (!qset) q[k] = jk[k][hq];
                q[k] = q[k] + t - pin[k];
                pin[k] = t;
            }
            have_js0p = true;
            qset = pin_set = true;
        }
        if (qset) for (int k = 0; k < 3; ++k) jk[k][hq] = q[k];
        else for (int k = 0; k < 3; ++k) q[k] = jk[k][hq];
        if (pin_set) for (int k = 0; k < 3; ++k) m.pin[k * nk + i] = pin[k];
    }
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
    const Cur9 J{c.j0x[o], c.j0x[on], c.j0y[o], c.j0y[on], q[0], pp[0], q[1], pp[1], pp[2]};
    const Mode6 r = MIXED ? march_mode_mixed(c, i, n, old, up, J) : march_mode(c, i, n, old, up, J);
    c.Rx[o] = r.Rx; c.Ry[o] = r.Ry; c.Ez[o] = r.Ez; c.Bz[o] = r.Bz; c.Lx[o] = r.Lx; c.Ly[o] = r.Ly;
    if (m.last && !m.top) {                         // src/sim_plasma.cpp:381-396
        if (!have_si0) { si0[0] = m.si0[i]; si0[1] = m.si0[nk + i]; }
        c.Rx[on] = up.Rx - m.qcorr * (u[0] - si0[0]);
        c.Ry[on] = up.Ry - m.qcorr * (u[1] - si0[1]);
        for (int k = 0; k < 3; ++k) {
            if (!have_js0p) js0p[k] = js0[k][hp];
            m.dl[k * nk + i] = u[k] - js0p[k];
            js0[k][hp] = u[k];
        }
    }
    pack3(T, r.Rx, r.Ry, r.Lx, r.Ly, r.Ez, r.Bz);
}

}  // namespace g3d
