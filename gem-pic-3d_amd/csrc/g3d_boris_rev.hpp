// The reverse mode of the Boris rotation boris() (g3d_common.hpp, src/sim_stream.hpp:8-18), statement by statement:
// the adjoints of its inputs from the adjoint of the rotated momentum.  Shared by the Boris reverse of the C ABI
// (g3d_adjoint.hip) and the reverse of the streaming push (g3d_spush.hip).  No counterpart in the reference.
#pragma once
#include "g3d_common.hpp"

namespace g3d {

G3D_HD void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1]; c[1] = a[2] * b[0] - a[0] * b[2]; c[2] = a[0] * b[1] - a[1] * b[0];
}
// u: the momentum before the rotation, ub: the adjoint of the one after it; ab receives the adjoint of u (set),
// Eb, Bb, qmb and dtb are added to
G3D_HD void boris_rev(const double u[3], const double E[3], const double B[3], double qm, double dt,
                      const double ub[3], double ab[3], double Eb[3], double Bb[3], double& qmb, double& dtb) {
    // forward values
    const double qh = 0.5 * qm * dt;
    double av[3], t[3], pv[3], c[3];
    for (int k = 0; k < 3; ++k) av[k] = u[k] + qh * E[k];
    const double gm = sqrt(1.0 + av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
    for (int k = 0; k < 3; ++k) t[k] = qh * B[k] / gm;
    const double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    cross3(av, t, c);
    for (int k = 0; k < 3; ++k) pv[k] = av[k] + c[k];
    const double sf = 2.0 / (1.0 + t2);
    // reverse: u' = b + qh E, b = a + sf (p x t)
    double tb[3], pb[3];
    double qhb = ub[0] * E[0] + ub[1] * E[1] + ub[2] * E[2];
    for (int k = 0; k < 3; ++k) { Eb[k] += qh * ub[k]; ab[k] = ub[k]; }
    double pxt[3];
    cross3(pv, t, pxt);
    const double sfb = ub[0] * pxt[0] + ub[1] * pxt[1] + ub[2] * pxt[2];
    double txb[3], bxp[3];
    cross3(t, ub, txb);                  // d(p x t) . ub / dp = t x ub
    cross3(ub, pv, bxp);                 // d(p x t) . ub / dt = ub x p
    for (int k = 0; k < 3; ++k) { pb[k] = sf * txb[k]; tb[k] = sf * bxp[k]; }
    // sf = 2 / (1 + t2)
    const double t2b = -sfb * sf * sf * 0.5;
    // p = a + a x t
    double txp[3], pxa[3];
    cross3(t, pb, txp);                  // d(a x t) . pb / da = t x pb
    cross3(pb, av, pxa);                 // d(a x t) . pb / dt = pb x a
    for (int k = 0; k < 3; ++k) { ab[k] += pb[k] + txp[k]; tb[k] += pxa[k]; }
    // t2 = t . t
    for (int k = 0; k < 3; ++k) tb[k] += 2 * t2b * t[k];
    // t = qh B / gm
    double gmb = 0;
    for (int k = 0; k < 3; ++k) {
        Bb[k] += qh * tb[k] / gm;
        qhb += tb[k] * B[k] / gm;
        gmb -= tb[k] * t[k] / gm;
    }
    // gm = sqrt(1 + a . a)
    for (int k = 0; k < 3; ++k) ab[k] += gmb * av[k] / gm;
    // a = u + qh E
    for (int k = 0; k < 3; ++k) { Eb[k] += qh * ab[k]; qhb += ab[k] * E[k]; }
    qmb += 0.5 * dt * qhb;
    dtb += 0.5 * qm * qhb;
}

}  // namespace g3d
