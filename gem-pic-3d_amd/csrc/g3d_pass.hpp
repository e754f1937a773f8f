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
// MIXED: the implicit part of the march in FP32 (march_mode_mixed), the bookkeeping stored as complex float.
//
// mode_load issues every read of a mode up front (under the launch-uniform flags of the pass), so that a kernel
// can overlap them with other work before mode_compute needs them; no read follows a write of the same mode.
template <bool MIXED>
__device__ __forceinline__ cx bk_load(const cx* p, long k) {
    if constexpr (MIXED) return to_d(reinterpret_cast<const cxf*>(p)[k]);
    else return p[k];
}
template <bool MIXED>
__device__ __forceinline__ void bk_store(cx* p, long k, cx v) {
    if constexpr (MIXED) reinterpret_cast<cxf*>(p)[k] = to_f(v);
    else p[k] = v;
}

struct ModeLoad {
    cx pin[3], sp[3], jkq[3], jkp[3], js0q[3], js0p[3], dl[3], si0[2];
    cx j0xo, j0xn, j0yo, j0yn;
    Mode6 old, up;
};

template <bool MIXED>
__device__ __forceinline__ ModeLoad mode_load(const ModeArgs& m, long i) {
    const MarchView& c = m.mv;
    const long nk = c.nk;
    const int n = m.n;
    const long hp = long(n + 1) * nk + i, hq = long(n) * nk + i;
    const cx* jk[3] = {m.jkx, m.jky, m.jkz};
    const cx* js0[3] = {m.js0x, m.js0y, m.js0z};
    ModeLoad L;
    const bool ext = !m.top && n >= 1 && m.extrap;
    const bool need_q0 = m.top || (m.first && n >= 1) || ext;          // js0 of half-cell n-1
    const bool need_p0 = ext || (m.last && !m.top);                     // js0 of half-cell n
    for (int k = 0; k < 3; ++k) {
        L.pin[k] = bk_load<MIXED>(m.pin, k * nk + i);
        L.jkq[k] = jk[k][hq];
        L.jkp[k] = jk[k][hp];
        if (need_q0) L.js0q[k] = js0[k][hq];
        if (need_p0) L.js0p[k] = js0[k][hp];
        if (!m.top && !m.first) L.sp[k] = bk_load<MIXED>(m.sin_, k * nk + i);
        if (ext) L.dl[k] = bk_load<MIXED>(m.dl, k * nk + i);
    }
    if (m.last && !m.top && !m.first) { L.si0[0] = bk_load<MIXED>(m.si0, i); L.si0[1] = bk_load<MIXED>(m.si0, nk + i); }
    const long o = long(n - c.a) * nk + i, on = o + nk;
    // slice n at s0: the passes before the last leave it in place (below), so every pass reads it there
    L.old.Rx = c.Rx[o]; L.old.Ry = c.Ry[o]; L.old.Ez = c.Ez[o]; L.old.Bz = c.Bz[o];
    L.up = Mode6{};
    if (n < c.N - 1) L.up = Mode6{c.Rx[on], c.Ry[on], c.Ez[on], c.Bz[on], c.Lx[on], c.Ly[on]};
    L.j0xo = c.j0x[o]; L.j0xn = c.j0x[on]; L.j0yo = c.j0y[o]; L.j0yn = c.j0y[on];
    return L;
}

template <bool MIXED>
__device__ __forceinline__ void mode_compute(const ModeArgs& m, long i, const ModeLoad& L, const cx u[3], cx T[3]) {
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
            const cx t = L.js0q[k];
            q[k] = L.jkq[k] + t - L.pin[k];
            bk_store<MIXED>(m.pin, k * nk + i, t);
            jk[k][hq] = q[k];
            pp[k] = L.jkp[k];
        }
    } else {
        cx pin[3], sp[3];
        for (int k = 0; k < 3; ++k) pin[k] = L.pin[k];
        if (m.first) {
            for (int k = 0; k < 3; ++k) { sp[k] = pin[k]; si0[k] = pin[k]; bk_store<MIXED>(m.si0, k * nk + i, pin[k]); }
            have_si0 = true;
        } else
            for (int k = 0; k < 3; ++k) sp[k] = L.sp[k];
        cx t0[3];
        bool qset = false, pin_set = false;
        if (m.first && n >= 1) {
            for (int k = 0; k < 3; ++k) {
                t0[k] = L.js0q[k];
                q[k] = L.jkq[k] + t0[k] - mk(0);
                pin[k] = t0[k];
            }
            qset = pin_set = true;
        }
        for (int k = 0; k < 3; ++k) {
            pp[k] = L.jkp[k] + u[k] - sp[k];
            jk[k][hp] = pp[k];
            bk_store<MIXED>(m.sin_, k * nk + i, u[k]);
        }
        if (n >= 1 && m.extrap) {
            for (int k = 0; k < 3; ++k) {
                js0p[k] = L.js0p[k];
                const cx d0 = u[k] - js0p[k];
                const cx jq = (m.first && n >= 1) ? t0[k] : L.js0q[k];
                const cx t = jq + d0 + m.wex * (d0 - L.dl[k]);
                if (!qset) q[k] = L.jkq[k];
                q[k] = q[k] + t - pin[k];
                pin[k] = t;
            }
            have_js0p = true;
            qset = pin_set = true;
        }
        if (qset) for (int k = 0; k < 3; ++k) jk[k][hq] = q[k];
        else for (int k = 0; k < 3; ++k) q[k] = L.jkq[k];
        if (pin_set) for (int k = 0; k < 3; ++k) bk_store<MIXED>(m.pin, k * nk + i, pin[k]);
    }
    const long o = long(n - c.a) * nk + i;
    Mode6 old = L.old;
    // FP32-mixed with several passes: E_z and B_z of slice n enter every pass rounded to float (the precision of the
    // pass bookkeeping)
    if constexpr (MIXED) if (m.bRx) { old.Ez = to_d(to_f(old.Ez)); old.Bz = to_d(to_f(old.Bz)); }
    const Mode6& up = L.up;
    const long on = o + nk;
    const Cur9 J{L.j0xo, L.j0xn, L.j0yo, L.j0yn, q[0], pp[0], q[1], pp[1], pp[2]};
    const Mode6 r = MIXED ? march_mode_mixed(c, i, n, old, up, J) : march_mode(c, i, n, old, up, J);
    // a pass before the last (m.bRx set: several passes) leaves slice n at s0 for the next pass to start from; only
    // its packed fields (T) go on to the particles
    if (m.last || !m.bRx) { c.Rx[o] = r.Rx; c.Ry[o] = r.Ry; c.Ez[o] = r.Ez; c.Bz[o] = r.Bz; c.Lx[o] = r.Lx; c.Ly[o] = r.Ly; }
    if (m.last && !m.top) {                         // src/sim_plasma.cpp:381-396
        if (!have_si0) { si0[0] = L.si0[0]; si0[1] = L.si0[1]; }
        c.Rx[on] = up.Rx - m.qcorr * (u[0] - si0[0]);
        c.Ry[on] = up.Ry - m.qcorr * (u[1] - si0[1]);
        for (int k = 0; k < 3; ++k) {
            if (!have_js0p) js0p[k] = L.js0p[k];
            bk_store<MIXED>(m.dl, k * nk + i, u[k] - js0p[k]);
            js0[k][hp] = u[k];
        }
    }
    pack3(T, r.Rx, r.Ry, r.Lx, r.Ly, r.Ez, r.Bz);
}

template <bool MIXED>
__device__ __forceinline__ void mode_work(const ModeArgs& m, long i, const cx u[3], cx T[3]) {
    mode_compute<MIXED>(m, i, mode_load<MIXED>(m, i), u, T);
}

}  // namespace g3d
