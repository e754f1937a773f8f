/* g3d_oracle.c - CPU restatement of the per-slice march (test infrastructure; see g3d_oracle.h).
 * Built with -ffp-contract=off: no fused multiply-add, as in the reference's x86-64 g++ build. */
#include "g3d_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef g3d_cpx cx;
static cx mk(double r, double i) { cx z; z.re = r; z.im = i; return z; }
static cx add(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static cx sub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
static cx neg(cx a) { return mk(-a.re, -a.im); }
static cx mul(cx a, cx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cx scl(double s, cx a) { return mk(s * a.re, s * a.im); }
static cx dvs(cx a, double s) { return mk(a.re / s, a.im / s); }
static cx cjg(cx a) { return mk(a.re, -a.im); }
static cx ik(double k) { return mk(0, k); }

/* ---------------------------------------------------------------- FFT (src/fft.hpp:21-85, fftmany.hpp:95-118) */
void or_fft1(cx* a, int n, int inverse) {
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    if (n <= 1) return;
    for (int i = 0; i < n; ++i) {                 /* bit-reversal permutation */
        int r = 0;
        for (int b = 0; b < lg; ++b) if (i >> b & 1) r |= 1 << (lg - 1 - b);
        if (i < r) { cx t = a[i]; a[i] = a[r]; a[r] = t; }
    }
    cx* w = (cx*)malloc(sizeof(cx) * (size_t)(n / 2 > 0 ? n / 2 : 1));
    for (int k = 0; k < n / 2; ++k) {             /* std::polar(1, -2 pi k / n) */
        const double ph = -2.0 * M_PI * k / n;
        w[k] = mk(cos(ph), sin(ph));
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len / 2, step = n / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < half; ++j) {
                const cx ww = inverse ? cjg(w[j * step]) : w[j * step];
                const cx u = a[i + j], v = mul(a[i + j + half], ww);
                a[i + j] = add(u, v);
                a[i + j + half] = sub(u, v);
            }
    }
    free(w);
}

void or_fft_many(int nx, int ny, int two_d, cx* z, int count, int inverse) {
    const long rows = (long)count * ny, nk = (long)nx * ny;
    const double s = 1.0 / (double)(two_d ? nk : nx);
    for (long r = 0; r < rows; ++r) or_fft1(z + r * nx, nx, inverse);
    if (two_d) {
        cx* col = (cx*)malloc(sizeof(cx) * (size_t)ny);
        for (long c = 0; c < (long)count * nx; ++c) {
            cx* p = z + (c / nx) * nk + c % nx;
            for (int j = 0; j < ny; ++j) col[j] = p[(long)j * nx];
            or_fft1(col, ny, inverse);
            for (int j = 0; j < ny; ++j) p[(long)j * nx] = col[j];
        }
        free(col);
    }
    if (inverse)
        for (long i = 0; i < rows * nx; ++i) z[i] = scl(s, z[i]);
}

/* ---------------------------------------------------------------- grids */
void or_tsc_index(double xi, int idx[3], double w[3]) {      /* src/sim_kernels.hpp:21-26 */
    const int i0 = (int)floor(xi + 0.5);
    const double f = xi - i0;
    w[0] = 0.5 * (0.5 - f) * (0.5 - f);
    w[1] = 0.75 - f * f;
    w[2] = 0.5 * (0.5 + f) * (0.5 + f);
    for (int k = 0; k < 3; ++k) idx[k] = i0 - 1 + k;
}
static void tsc_per(double x, double d, int n, int idx[3], double w[3]) {   /* src/sim_kernels.hpp:28-31 */
    or_tsc_index(x / d, idx, w);
    for (int k = 0; k < 3; ++k) idx[k] = (idx[k] % n + n) % n;
}
void or_axis_locate(const or_axis* ax, double x, int* i, double* f) {      /* src/grid.hpp:35-46 */
    if (ax->uniform) { const double t = (x - ax->x0) * ax->ih; *i = (int)floor(t); *f = t - *i; return; }
    const double* X = ax->X;
    const int n = ax->n;
    if (x < X[0]) { const double t = (x - X[0]) / (X[1] - X[0]); *i = (int)floor(t); *f = t - *i; return; }
    if (x >= X[n - 1]) {
        const double t = (x - X[n - 1]) / (X[n - 1] - X[n - 2]);
        const int k = (int)floor(t);
        *i = n - 1 + k; *f = t - k;
        return;
    }
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (X[m] <= x) lo = m; else hi = m; }
    *i = lo; *f = (x - X[lo]) / (X[lo + 1] - X[lo]);
}
static int is_half(int c) { return c == 1 || c == 3 || c == 5; }           /* src/sim_internal.hpp:21 */
static double wrapm(double x, double L) { return fmod(fmod(x, L) + L, L); }

void or_gather(const or_fields* c, double x, double y, double zeta, double E[3], double B[3]) {  /* sim_kernels.hpp:44-72 */
    E[0] = E[1] = E[2] = B[0] = B[1] = B[2] = 0;
    if (!c->self) return;
    int n; double fz;
    or_axis_locate(&c->zax, zeta, &n, &fz);
    if (n == c->a - 1 && fz > 1 - 1e-9) { n = c->a; fz = 0; }
    const int fl = n - c->a;
    if (fl < 0 || fl > c->ns) return;
    int ix[3], jn[3], jh[3]; double wx[3], wn[3], wh[3];
    tsc_per(x, c->dx, c->nx, ix, wx);
    if (c->yf) tsc_per(y, c->dy, c->ny, jn, wn);
    else {
        int i; double f;
        or_axis_locate(&c->yax, y, &i, &f);
        or_tsc_index(i + f, jn, wn);
        or_tsc_index(i + f - 0.5, jh, wh);
    }
    for (int comp = 0; comp < 6; ++comp) {
        const float* p0 = c->F + ((size_t)comp * (c->ns + 2) + fl) * c->nk;
        const float* p1 = p0 + c->nk;
        const int hf = !c->yf && is_half(comp);
        double v = 0;
        for (int j = 0; j < 3; ++j) {
            const int jj = hf ? jh[j] : jn[j];
            const double wy = hf ? wh[j] : wn[j];
            if (jj < 0 || jj >= c->ny) continue;
            for (int i = 0; i < 3; ++i) {
                const long k = (long)jj * c->nx + ix[i];
                v += wx[i] * wy * ((1 - fz) * p0[k] + fz * p1[k]);
            }
        }
        if (comp < 3) E[comp] = v; else B[comp - 3] = v;
    }
}

void or_laser_field(const or_laser* L, double x, double z, double t, double E[3], double B[3]) {   /* sim.hpp:196-209 */
    E[0] = E[1] = E[2] = B[0] = B[1] = B[2] = 0;
    if (!L->on) return;
    const double xi = t - L->kx * x - L->kz * z - L->eta_c;
    if (fabs(xi) > 5 * L->tau) return;
    const double g = exp(-1.3862943611198906 * xi * xi / (L->tau * L->tau));
    const double gp = -2.7725887222397811 * xi / (L->tau * L->tau) * g;
    const double ph = L->kL * xi + L->cep, cc = cos(ph), ss = sin(ph);
    const double d1 = L->a0 * (gp * cc - L->kL * g * ss), d2 = L->a0 * (gp * ss + L->kL * g * cc);
    if (L->pol == 0) E[1] = -d1;
    else if (L->pol == 1) { E[0] = -d1 * L->kz; E[2] = d1 * L->kx; }
    else { E[0] = -d1 * L->kz; E[2] = d1 * L->kx; E[1] = -d2; }
    B[0] = -L->kz * E[1]; B[1] = L->kz * E[0] - L->kx * E[2]; B[2] = L->kx * E[1];
}

void or_boris(double u[3], const double E[3], const double B[3], double qm, double dt) {   /* src/sim_stream.hpp:8-18 */
    const double qh = 0.5 * qm * dt;
    const double mx = u[0] + qh * E[0], my = u[1] + qh * E[1], mz = u[2] + qh * E[2];
    const double gm = sqrt(1.0 + mx * mx + my * my + mz * mz);
    const double tx = qh * B[0] / gm, ty = qh * B[1] / gm, tz = qh * B[2] / gm;
    const double t2 = tx * tx + ty * ty + tz * tz;
    const double px = mx + (my * tz - mz * ty), py = my + (mz * tx - mx * tz), pz = mz + (mx * ty - my * tx);
    const double sf = 2.0 / (1.0 + t2);
    const double vx = mx + sf * (py * tz - pz * ty), vy = my + sf * (pz * tx - px * tz), vz = mz + sf * (px * ty - py * tx);
    u[0] = vx + qh * E[0]; u[1] = vy + qh * E[1]; u[2] = vz + qh * E[2];
}

/* ---------------------------------------------------------------- RK4 push in z (src/sim_kernels.hpp:75-149) */
typedef struct { double x, y, zeta, px, py, pz; } st6;
static int rhs(const or_push_ctx* c, const st6* Y, double z, double qm, st6* d) {
    const double g = sqrt(1 + Y->px * Y->px + Y->py * Y->py + Y->pz * Y->pz);
    if (Y->pz < 0.3 * g) return 0;
    const double gm = (1 + Y->px * Y->px + Y->py * Y->py) / (g + Y->pz);
    double E[3], B[3], Ee[3], Be[3];
    or_gather(&c->f, wrapm(Y->x, c->f.Lx), c->f.yf ? wrapm(Y->y, c->f.Ly) : Y->y, Y->zeta, E, B);
    or_laser_field(&c->laser, Y->x + c->X0, z, z - (Y->zeta - c->zc), Ee, Be);
    for (int k = 0; k < 3; ++k) { E[k] += Ee[k]; B[k] += Be[k]; }
    const double ip = 1 / Y->pz;
    d->x = Y->px * ip; d->y = Y->py * ip; d->zeta = -gm * ip;
    d->px = qm * (g * E[0] + Y->py * B[2] - Y->pz * B[1]) * ip;
    d->py = qm * (g * E[1] + Y->pz * B[0] - Y->px * B[2]) * ip;
    d->pz = qm * (g * E[2] + Y->px * B[1] - Y->py * B[0]) * ip;
    return 1;
}
static st6 axpy6(const st6* a, double h, const st6* b) {
    st6 r = {a->x + h * b->x, a->y + h * b->y, a->zeta + h * b->zeta, a->px + h * b->px, a->py + h * b->py,
             a->pz + h * b->pz};
    return r;
}
static void push_one(const or_push_ctx* c, g3d_particle* p, double z0) {
    if (c->freeze && p->sp <= 0) { p->z = c->s1; return; }
    st6 Y = {p->x, p->y, p->zeta, p->px, p->py, p->pz};
    double z = z0;
    while (z < c->s1 - 1e-12 * fabs(c->s1)) {
        if (Y.zeta < c->zhand) {
            p->x = Y.x; p->y = Y.y; p->zeta = Y.zeta; p->px = Y.px; p->py = Y.py; p->pz = Y.pz;
            p->alive = 2; p->rem = c->s1 - z;
            return;
        }
        if (c->vz_min > 0 && Y.pz < c->vz_min * sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz)) {
            p->x = Y.x; p->y = Y.y; p->zeta = Y.zeta; p->px = Y.px; p->py = Y.py; p->pz = Y.pz; p->z = z;
            p->alive = 4;
            return;
        }
        double ds = c->dz_push < c->s1 - z ? c->dz_push : c->s1 - z;
        {
            const double g = sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
            const double lz = 0.5 * c->hmin * Y.pz * (g + Y.pz) / (1 + Y.px * Y.px + Y.py * Y.py);
            if (lz < ds) ds = (c->s1 - z) / ceil((c->s1 - z) / lz);
        }
        if (c->laser.on) {
            const double g = sqrt(1 + Y.px * Y.px + Y.py * Y.py + Y.pz * Y.pz);
            const double hl = g - Y.px * c->laser.kx - Y.pz * c->laser.kz;
            const double lim = c->push_phase * Y.pz / (c->laser.kL * (hl > 1e-12 ? hl : 1e-12));
            if (lim < ds) ds = (c->s1 - z) / ceil((c->s1 - z) / lim);
        }
        st6 k1, k2, k3, k4, t;
        const double qm = p->qm;
        int ok = rhs(c, &Y, z, qm, &k1);
        if (ok) { t = axpy6(&Y, 0.5 * ds, &k1); ok = rhs(c, &t, z + 0.5 * ds, qm, &k2); }
        if (ok) { t = axpy6(&Y, 0.5 * ds, &k2); ok = rhs(c, &t, z + 0.5 * ds, qm, &k3); }
        if (ok) { t = axpy6(&Y, ds, &k3); ok = rhs(c, &t, z + ds, qm, &k4); }
        if (!ok) { p->alive = 0; return; }
        const double w = ds / 6;
        Y.x += w * (k1.x + 2 * k2.x + 2 * k3.x + k4.x); Y.y += w * (k1.y + 2 * k2.y + 2 * k3.y + k4.y);
        Y.zeta += w * (k1.zeta + 2 * k2.zeta + 2 * k3.zeta + k4.zeta);
        Y.px += w * (k1.px + 2 * k2.px + 2 * k3.px + k4.px); Y.py += w * (k1.py + 2 * k2.py + 2 * k3.py + k4.py);
        Y.pz += w * (k1.pz + 2 * k2.pz + 2 * k3.pz + k4.pz);
        z += ds;
    }
    p->x = Y.x; p->y = Y.y; p->zeta = Y.zeta; p->px = Y.px; p->py = Y.py; p->pz = Y.pz; p->z = c->s1;
    p->alive = (p->zeta < 0 || p->zeta >= c->zmax || p->y < c->ylo || p->y > c->yhi) ? 0 : 1;
    if (p->alive == 1 && p->zeta < c->zhand) { p->alive = 2; p->rem = 0; }
}
void or_bunch_push(const or_push_ctx* c, g3d_particle* p, long n, double s0) {   /* src/sim_particles.cpp:17-22 */
    for (long i = 0; i < n; ++i) {
        if (p[i].alive == 1) push_one(c, &p[i], s0);
        else if (p[i].alive == 3) push_one(c, &p[i], c->s1 - p[i].rem);
    }
}

void or_bunch_deposit(const g3d_particle* P, long np, const or_axis* za, const double* hc, const double* ia, double dx,
                      double dy, int nx, int ny, int a, int ns, int N, double* J) {   /* src/sim_particles.cpp:66-89 */
    const int nh = ns + 3;
    const long nk = (long)nx * ny;
    for (long ip = 0; ip < np; ++ip) {
        const g3d_particle* p = &P[ip];
        if (p->alive != 1) continue;
        int m; double fz;
        or_axis_locate(za, p->zeta, &m, &fz);
        const int hw = m - (a - 2);
        if (hw < 0 || hw >= nh || m < 0 || m > N - 2) continue;
        const double g = sqrt(1 + p->px * p->px + p->py * p->py + p->pz * p->pz);
        const double vx = p->px / g, vy = p->py / g, vz = p->pz / g;
        const double f = p->q * p->w / (hc[m + 2] * dx);
        int ix[3], jn[3]; double wx[3], wn[3];
        tsc_per(p->x, dx, nx, ix, wx);
        tsc_per(p->y, dy, ny, jn, wn);
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) {
                const long k = (long)jn[v] * nx + ix[u];
                const double w = f * wx[u] * wn[v] * ia[0];
                J[(0L * nh + hw) * nk + k] += w * vx / vz;
                J[(1L * nh + hw) * nk + k] += w * vy / vz;
                J[(2L * nh + hw) * nk + k] += w;
            }
    }
}

/* ---------------------------------------------------------------- ADK and draws */
double or_adk_rate(const or_adk* L, double E) {                        /* src/ionization.hpp:26-30 */
    if (!(E >= L->Emin) || !(E > 0)) return 0.0;
    E = E < L->Epeak ? E : L->Epeak;
    return exp(L->logpre - L->power * log(E) + L->expo / E);
}
static double adk_raw(const or_adk* L, double E) {                     /* src/ionization.cpp:17-20 */
    E = E < L->Epeak ? E : L->Epeak;
    return E > 0 ? exp(L->logpre - L->power * log(E) + L->expo / E) : 0.0;
}
void or_adk_levels(const double* ip, int n_ip, int Z0, int Zmax, double E0, double omega0, double T, or_adk* lev) {
    /* src/ionization.cpp:9-16 (coefficients) and 30-53 (code units, peak, skip threshold by bisection) */
    const double UH = 13.605693122994, Ea = 5.14220674763e11, wa = 4.1341373335e16;
    for (int k = 0; k < Zmax; ++k) { lev[k].logpre = lev[k].power = lev[k].expo = 0; lev[k].Epeak = lev[k].Emin = 1e300; }
    for (int k = Z0; k < Zmax && k < n_ip; ++k) {
        const double x = ip[k] / UH;
        const double nstar = (k + 1) / sqrt(x);
        const double logC2 = 2.0 * nstar * log(2.0) - log(nstar) - lgamma(nstar + 1.0) - lgamma(nstar);
        const double power = 2.0 * nstar - 1.0;
        const double logpre = log(wa) + logC2 + log(0.5 * x) + power * log(2.0 * pow(x, 1.5) * Ea);
        const double expo = -(2.0 / 3.0) * pow(x, 1.5) * Ea;
        or_adk* L = &lev[k];
        L->power = power;
        L->logpre = logpre - log(omega0) - power * log(E0);
        L->expo = expo / E0;
        L->Epeak = power > 0 ? -L->expo / power : 1e300;
        double lo = 1e-12, hi = L->Epeak < 1e6 ? L->Epeak : 1e6;
        if (adk_raw(L, hi) * T < 1e-10) L->Emin = hi;
        else {
            for (int it = 0; it < 200; ++it) {
                const double mid = sqrt(lo * hi);
                if (adk_raw(L, mid) * T < 1e-10) lo = mid; else hi = mid;
            }
            L->Emin = lo;
        }
    }
}
uint64_t or_mix64(uint64_t z) {                                       /* src/ionization.hpp:37-42 (SplitMix64) */
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
double or_uniform01(uint64_t key) { return (double)(or_mix64(key) >> 11) * (1.0 / 9007199254740992.0); }
uint64_t or_key_step(uint64_t seed, long step) { return or_mix64(seed * 0x9E3779B97F4A7C15ULL + (uint64_t)step); }
uint64_t or_key_slice(uint64_t ks, int n, int pass) { return or_mix64(ks ^ (((uint64_t)n << 8) | (uint64_t)pass)); }
uint64_t or_key_particle(uint64_t kl, uint64_t id) { return or_mix64(kl ^ or_mix64(id)); }

/* ---------------------------------------------------------------- streaming half-cell (src/sim_stream.hpp:33-171) */
static void sdep(const or_stream_ctx* c, double x, double y, double f, double vx, double vy, double vz) {
    const or_fields* P = &c->f;
    int ix[3], jn[3], jh[3]; double wx[3], wn[3], wh[3];
    tsc_per(x, P->dx, P->nx, ix, wx);
    const long nk = P->nk, cq = (f < 0 ? 3 : 4) * nk;
    if (P->yf) {
        tsc_per(y, P->dy, P->ny, jn, wn);
        for (int v = 0; v < 3; ++v)
            for (int u = 0; u < 3; ++u) {
                const long k = (long)jn[v] * P->nx + ix[u];
                const double w = f * wx[u] * wn[v] * c->ia[0];
                c->J[k] += w * vx; c->J[nk + k] += w * vy; c->J[2 * nk + k] += w * vz;
                c->J[cq + k] += w;
            }
        return;
    }
    int i; double fr;
    or_axis_locate(&P->yax, y, &i, &fr);
    or_tsc_index(i + fr, jn, wn);
    or_tsc_index(i + fr - 0.5, jh, wh);
    for (int v = 0; v < 3; ++v) {
        const int nod = jn[v] >= 1 && jn[v] <= P->ny - 1, hal = jh[v] >= 0 && jh[v] <= P->ny - 1;
        for (int u = 0; u < 3; ++u) {
            if (nod) {
                const long k = (long)jn[v] * P->nx + ix[u];
                const double w = f * wx[u] * wn[v] * c->ia[jn[v]];
                c->J[k] += w * vx; c->J[2 * nk + k] += w * vz; c->J[cq + k] += w;
            }
            if (hal) c->J[nk + (long)jh[v] * P->nx + ix[u]] += f * wx[u] * wh[v] * c->id[jh[v]] * vy;
        }
    }
}
static double field_abs(const or_stream_ctx* c, double x, double y, double z, double zeta) {
    const or_fields* P = &c->f;
    double E[3], B[3], Ee[3], Be[3];
    or_gather(P, wrapm(x, P->Lx), P->yf ? wrapm(y, P->Ly) : y, zeta, E, B);
    or_laser_field(&c->laser, x + c->X0, z, z - (zeta - c->zc), Ee, Be);
    for (int k = 0; k < 3; ++k) E[k] += Ee[k];
    return sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
}
static void ionize(const or_stream_ctx* c, g3d_particle* p, double E1, double E0, double dt) {
    const or_ion* I = &c->ions[p->sp - 3];
    if (!I->ionizing || (E1 < I->Emin && E0 < I->Emin)) return;
    const uint64_t base = or_key_particle(c->key, p->id);
    int Z = (int)floor(p->q + 0.5);
    const int Z0 = Z;
    unsigned draw = 0;
    double tleft = dt;
    while (Z < I->Zmax) {
        const or_adk* L = &c->adk[I->off + Z];
        const double w = 0.5 * (or_adk_rate(L, E1) + or_adk_rate(L, E0));
        if (!(w > 0)) break;
        const double tw = -log(1.0 - or_uniform01(base + draw++)) / w;
        if (tw > tleft) break;
        tleft -= tw; ++Z;
        g3d_particle e = *p;
        e.q = -1; e.qm = -1; e.sp = 50 + (p->sp - 3); e.alive = 1;
        const int j = (*c->bcnt)++;
        if (j < c->bcap) c->born[j] = e;
    }
    if (Z != Z0) { p->q = Z; p->qm = Z / I->mass; }
}
static void stream_one(const or_stream_ctx* c, g3d_particle* p) {
    const int n = c->n;
    const double zn = c->zs[n], zn1 = c->zs[n + 1], h = zn1 - zn;
    if (p->alive == 6) { if (p->zeta > zn) p->alive = 1; else return; }
    if (p->alive != 1) return;
    const double hl = p->zeta >= zn1 - 1e-9 * h ? h : (p->zeta - zn > 0 ? p->zeta - zn : 0.0);
    const or_fields* P = &c->f;
    const int ion = p->sp >= 3 && p->sp < 50;
    if (ion && !c->ions[p->sp - 3].mobile) {
        const double f = p->q * p->w * hl / (c->Delta * h * P->dx);
        if (f != 0) sdep(c, p->x, p->y, f, 0, 0, 0);
        const double E1 = field_abs(c, p->x, p->y, p->z, p->zeta), E0 = field_abs(c, p->x, p->y, p->z, zn);
        p->zeta = zn;
        ionize(c, p, E1, E0, hl);
        return;
    }
    const double zeta0 = p->zeta;
    const double g0 = sqrt(1 + p->px * p->px + p->py * p->py + p->pz * p->pz);
    const double omv0 = (1 + p->px * p->px + p->py * p->py) / (g0 * (g0 + p->pz));
    const double dt0 = hl / omv0;
    const double xm = p->x + 0.5 * dt0 * p->px / g0, ym = p->y + 0.5 * dt0 * p->py / g0;
    const double zm = p->z + 0.5 * dt0 * p->pz / g0, zetam = p->zeta - 0.5 * hl;
    double E[3], B[3], Ee[3], Be[3];
    or_gather(P, wrapm(xm, P->Lx), P->yf ? wrapm(ym, P->Ly) : ym, zetam, E, B);
    or_laser_field(&c->laser, xm + c->X0, zm, zm - (zetam - c->zc), Ee, Be);
    for (int k = 0; k < 3; ++k) { E[k] += Ee[k]; B[k] += Be[k]; }
    double u[3] = {p->px, p->py, p->pz};
    or_boris(u, E, B, p->qm, dt0);
    double g1 = sqrt(1 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    double dt = 2 * hl / (omv0 + (1 + u[0] * u[0] + u[1] * u[1]) / (g1 * (g1 + u[2])));
    u[0] = p->px; u[1] = p->py; u[2] = p->pz;
    or_boris(u, E, B, p->qm, dt);
    g1 = sqrt(1 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    double vx = 0.5 * (p->px / g0 + u[0] / g1), vy = 0.5 * (p->py / g0 + u[1] / g1), vz = 0.5 * (p->pz / g0 + u[2] / g1);
    double zeta_end = zn;
    int trapped = 0;
    if (p->z + vz * dt >= c->sTrap) {
        const double dtp = (c->sTrap - p->z) / (vz > 1e-3 ? vz : 1e-3) > 0 ? (c->sTrap - p->z) / (vz > 1e-3 ? vz : 1e-3) : 0.0;
        u[0] = p->px; u[1] = p->py; u[2] = p->pz;
        or_boris(u, E, B, p->qm, dtp);
        g1 = sqrt(1 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        vx = 0.5 * (p->px / g0 + u[0] / g1); vy = 0.5 * (p->py / g0 + u[1] / g1); vz = 0.5 * (p->pz / g0 + u[2] / g1);
        dt = dtp;
        zeta_end = p->zeta - (1 - vz) * dtp;
        trapped = 1;
    }
    const double xo = p->x, yo = p->y;
    p->x += vx * dt; p->y += vy * dt;
    p->z = trapped ? c->s1 : p->z + vz * dt;
    p->px = u[0]; p->py = u[1]; p->pz = u[2];
    p->zeta = zeta_end;
    const double f = p->q * p->w * dt / (c->Delta * h * P->dx);
    if (f != 0) sdep(c, 0.5 * (xo + p->x), 0.5 * (yo + p->y), f, vx, vy, vz);
    if (!P->yf && (p->y < c->ylo || p->y > c->yhi)) { p->alive = 0; return; }
    if (ion && !trapped && c->ions[p->sp - 3].ionizing)
        ionize(c, p, field_abs(c, xm, ym, zm, zeta0), field_abs(c, xm, ym, zm, zn), dt);
    if (trapped) {
        p->alive = 5;
        const double vze = u[2] / g1;
        if (vze >= c->vz_beam) sdep(c, p->x, p->y, p->q * p->w / (h * P->dx * vze), u[0] / g1, u[1] / g1, vze);
    }
}
void or_stream_halfcell(const or_stream_ctx* c, g3d_particle* p, long n) {
    for (long i = 0; i < n; ++i) stream_one(c, &p[i]);
}

/* ---------------------------------------------------------------- march (src/sim_kernels.hpp:173-210) */
static void march_point(const or_march_ctx* c, long i, int n) {
    const int l = n - c->a;
    const long o = (long)l * c->nk + i;
    const int ixx = (int)(i % c->nx), j = (int)(i / c->nx);
    const cx a = ik(c->kx[ixx]), b = ik(c->ky[j]);
    const double k2 = c->kx[ixx] * c->kx[ixx] + c->ky[j] * c->ky[j], D = c->D, wl = c->wlo[n], wu = c->whi[n];
    const long od = o, ou = o + c->nk;
    const cx Jx = add(scl(wl, add(c->j0x[od], c->jkx[od])), scl(wu, add(c->j0x[ou], c->jkx[ou])));
    const cx Jy = add(scl(wl, add(c->j0y[od], c->jky[od])), scl(wu, add(c->j0y[ou], c->jky[ou])));
    const cx Rsx = sub(add(c->Rx[o], scl(D, add(mul(a, c->Ez[o]), mul(b, c->Bz[o])))), scl(D, Jx));
    const cx Rsy = sub(add(c->Ry[o], scl(D, sub(mul(b, c->Ez[o]), mul(a, c->Bz[o])))), scl(D, Jy));
    if (n == c->N - 1) {
        c->Rx[o] = Rsx; c->Ry[o] = Rsy;
        c->Ez[o] = c->Bz[o] = c->Lx[o] = c->Ly[o] = mk(0, 0);
        return;
    }
    const long p = o + c->nk;
    const cx pRx = c->Rx[p], pRy = c->Ry[p], pEz = c->Ez[p], pBz = c->Bz[p], pLx = c->Lx[p], pLy = c->Ly[p];
    const double h = c->hcg[n + 2];
    const cx upx = c->jkx[ou], upy = c->jky[ou], upz = c->jkz[ou];
    const cx Lsx = sub(sub(pLx, scl(0.25 * h, sub(mul(a, pEz), mul(b, pBz)))), scl(0.5 * h, upx));
    const cx Lsy = sub(sub(pLy, scl(0.25 * h, add(mul(a, pBz), mul(b, pEz)))), scl(0.5 * h, upy));
    const cx Cp = scl(0.5, sub(mul(a, add(pRy, pLy)), mul(b, add(pRx, pLx))));
    const cx Qp = add(scl(-0.5, mul(a, sub(pRx, pLx))), scl(0.5, mul(b, sub(pLy, pRy))));
    const cx SB = sub(sub(add(mul(a, Rsy), mul(a, Lsy)), mul(b, Rsx)), mul(b, Lsx));
    const cx SE = sub(add(add(neg(mul(a, Rsx)), mul(a, Lsx)), mul(b, Lsy)), mul(b, Rsy));
    const double den = 1 + 0.25 * h * (D + 0.25 * h) * k2;
    const cx Bzn = dvs(sub(sub(pBz, scl(0.5 * h, Cp)), scl(0.25 * h, SB)), den);
    const cx Ezn = dvs(sub(sub(sub(pEz, scl(0.5 * h, Qp)), scl(0.25 * h, SE)), scl(h, upz)), den);
    c->Ez[o] = Ezn; c->Bz[o] = Bzn;
    c->Rx[o] = add(Rsx, scl(D, add(mul(a, Ezn), mul(b, Bzn))));
    c->Ry[o] = add(Rsy, scl(D, sub(mul(b, Ezn), mul(a, Bzn))));
    c->Lx[o] = sub(Lsx, scl(0.25 * h, sub(mul(a, Ezn), mul(b, Bzn))));
    c->Ly[o] = sub(Lsy, scl(0.25 * h, add(mul(a, Bzn), mul(b, Ezn))));
}
/* ---------------------------------------------------------------- march of a k_x line, free-hand y
 * (src/sim_kernels.hpp:213-313): explicit parts of R and L in scratch lines, the E_z system on the nodes 1..ny-1
 * and the B_z system on the half-nodes 0..ny-1 by Thomas elimination, then R and L.  Scratch: 5 lines of cx and
 * one of double per k_x, element j of line ik at j nx + ik. */
typedef struct { cx *sRx, *sRy, *sLx, *sLy, *rhs; double* cp; } line_scratch;
static void march_line(const or_march_ctx* c, const line_scratch* w, int kx, int n) {
    const int nx = c->nx, ny = c->ny;
    const long nk = c->nk;
    const long P0 = (long)(n - c->a) * nk + kx, P1 = P0 + nk;
    const cx a = ik(c->kx[kx]);
    const double kx2 = c->kx[kx] * c->kx[kx], D = c->D, wl = c->wlo[n], wu = c->whi[n];
    const double *iD = c->iD, *iA = c->iA;
#define SL(A, j) (A)[(long)(j) * nx + kx]
#define NX(A, j) (((j) < 0 || (j) >= ny) ? mk(0, 0) : (A)[P1 + (long)(j) * nx])
    for (int j = 0; j < ny; ++j) {
        const long o = P0 + (long)j * nx;
        if (j == 0) SL(w->sRx, 0) = mk(0, 0);
        else {
            const cx Jx = add(scl(wl, add(c->j0x[o], c->jkx[o])), scl(wu, add(c->j0x[o + nk], c->jkx[o + nk])));
            const cx bm = scl(iA[j], sub(c->Bz[o], c->Bz[o - nx]));
            SL(w->sRx, j) = sub(add(c->Rx[o], scl(D, add(mul(a, c->Ez[o]), bm))), scl(D, Jx));
        }
        const cx Jy = add(scl(wl, add(c->j0y[o], c->jky[o])), scl(wu, add(c->j0y[o + nk], c->jky[o + nk])));
        const cx ezu = j + 1 < ny ? c->Ez[o + nx] : mk(0, 0);
        SL(w->sRy, j) = sub(add(c->Ry[o], scl(D, sub(scl(iD[j], sub(ezu, c->Ez[o])), mul(a, c->Bz[o])))), scl(D, Jy));
    }
    if (n == c->N - 1) {
        for (int j = 0; j < ny; ++j) {
            const long o = P0 + (long)j * nx;
            c->Rx[o] = SL(w->sRx, j); c->Ry[o] = SL(w->sRy, j);
            c->Ez[o] = c->Bz[o] = c->Lx[o] = c->Ly[o] = mk(0, 0);
        }
        return;
    }
    const double h = c->hcg[n + 2], cc = 0.25 * h * (D + 0.25 * h);
#define PEZ(j) ((j) == 0 ? mk(0, 0) : NX(c->Ez, j))
    for (int j = 0; j < ny; ++j) {
        const long u = P1 + (long)j * nx;
        if (j == 0) SL(w->sLx, 0) = mk(0, 0);
        else {
            const cx bm = scl(iA[j], sub(NX(c->Bz, j), NX(c->Bz, j - 1)));
            SL(w->sLx, j) = sub(sub(NX(c->Lx, j), scl(0.25 * h, sub(mul(a, PEZ(j)), bm))), scl(0.5 * h, c->jkx[u]));
        }
        const cx bp = scl(iD[j], sub(PEZ(j + 1), PEZ(j)));
        SL(w->sLy, j) = sub(sub(NX(c->Ly, j), scl(0.25 * h, add(mul(a, NX(c->Bz, j)), bp))), scl(0.5 * h, c->jky[u]));
    }
    double cp = 0;
    cx dp = mk(0, 0);
    for (int j = 1; j < ny; ++j) {                       /* E_z, nodes */
        const cx Qp = scl(0.5, add(neg(mul(a, sub(NX(c->Rx, j), NX(c->Lx, j)))),
                                   scl(iA[j], sub(sub(NX(c->Ly, j), NX(c->Ry, j)), sub(NX(c->Ly, j - 1), NX(c->Ry, j - 1))))));
        const cx SE = add(neg(mul(a, sub(SL(w->sRx, j), SL(w->sLx, j)))),
                          scl(iA[j], sub(sub(SL(w->sLy, j), SL(w->sRy, j)), sub(SL(w->sLy, j - 1), SL(w->sRy, j - 1)))));
        const cx r = sub(sub(sub(PEZ(j), scl(0.5 * h, Qp)), scl(0.25 * h, SE)), scl(h, c->jkz[P1 + (long)j * nx]));
        const double lo = j > 1 ? -cc * iA[j] * iD[j - 1] : 0, up = j < ny - 1 ? -cc * iA[j] * iD[j] : 0;
        const double den = 1 + cc * kx2 + cc * iA[j] * (iD[j - 1] + iD[j]) - lo * cp;
        dp = dvs(sub(r, scl(lo, dp)), den);
        cp = up / den;
        SL(w->rhs, j) = dp; SL(w->cp, j) = cp;
    }
    {
        cx next = mk(0, 0);
        for (int j = ny - 1; j >= 1; --j) { next = sub(SL(w->rhs, j), scl(SL(w->cp, j), next)); c->Ez[P0 + (long)j * nx] = next; }
        c->Ez[P0] = mk(0, 0);
    }
    cp = 0; dp = mk(0, 0);
    for (int j = 0; j < ny; ++j) {                       /* B_z, half-nodes */
        const cx fu = j + 1 < ny ? add(NX(c->Rx, j + 1), NX(c->Lx, j + 1)) : mk(0, 0), f0 = add(NX(c->Rx, j), NX(c->Lx, j));
        const cx Cp = scl(0.5, sub(mul(a, add(NX(c->Ry, j), NX(c->Ly, j))), scl(iD[j], sub(fu, f0))));
        const cx su = j + 1 < ny ? add(SL(w->sRx, j + 1), SL(w->sLx, j + 1)) : mk(0, 0), s0 = add(SL(w->sRx, j), SL(w->sLx, j));
        const cx SB = sub(mul(a, add(SL(w->sRy, j), SL(w->sLy, j))), scl(iD[j], sub(su, s0)));
        const cx r = sub(sub(NX(c->Bz, j), scl(0.5 * h, Cp)), scl(0.25 * h, SB));
        const double al = j > 0 ? iA[j] : 0, au = j < ny - 1 ? iA[j + 1] : 0;
        const double lo = -cc * iD[j] * al, up = -cc * iD[j] * au;
        const double den = 1 + cc * kx2 + cc * iD[j] * (al + au) - lo * cp;
        dp = dvs(sub(r, scl(lo, dp)), den);
        cp = up / den;
        SL(w->rhs, j) = dp; SL(w->cp, j) = cp;
    }
    {
        cx next = mk(0, 0);
        for (int j = ny - 1; j >= 0; --j) { next = sub(SL(w->rhs, j), scl(SL(w->cp, j), next)); c->Bz[P0 + (long)j * nx] = next; }
    }
    for (int j = 0; j < ny; ++j) {                       /* R and L */
        const long o = P0 + (long)j * nx;
        if (j == 0) { c->Rx[o] = mk(0, 0); c->Lx[o] = mk(0, 0); }
        else {
            const cx bm = scl(iA[j], sub(c->Bz[o], c->Bz[o - nx]));
            c->Rx[o] = add(SL(w->sRx, j), scl(D, add(mul(a, c->Ez[o]), bm)));
            c->Lx[o] = sub(SL(w->sLx, j), scl(0.25 * h, sub(mul(a, c->Ez[o]), bm)));
        }
        const cx ezu = j + 1 < ny ? c->Ez[o + nx] : mk(0, 0);
        const cx bp = scl(iD[j], sub(ezu, c->Ez[o]));
        c->Ry[o] = add(SL(w->sRy, j), scl(D, sub(bp, mul(a, c->Bz[o]))));
        c->Ly[o] = sub(SL(w->sLy, j), scl(0.25 * h, add(mul(a, c->Bz[o]), bp)));
    }
#undef PEZ
#undef NX
#undef SL
}

void or_march(const or_march_ctx* c, int lo, int hi) {               /* src/sim_fields.cpp:19-29 (CPU order) */
    if (c->ky) {
        for (int n = hi - 1; n >= lo; --n)
            for (long i = 0; i < c->nk; ++i) march_point(c, i, n);
        return;
    }
    line_scratch w;
    w.sRx = (cx*)malloc(sizeof(cx) * 5 * (size_t)c->nk);
    w.sRy = w.sRx + c->nk; w.sLx = w.sRy + c->nk; w.sLy = w.sLx + c->nk; w.rhs = w.sLy + c->nk;
    w.cp = (double*)malloc(sizeof(double) * (size_t)c->nk);
    for (int n = hi - 1; n >= lo; --n)
        for (int i = 0; i < c->nx; ++i) march_line(c, &w, i, n);
    free(w.cp);
    free(w.sRx);
}

/* ---------------------------------------------------------------- spectra and real fields */
void or_stream_spectra(int nx, int ny, const double* J, const double* kxs, const double* kys, const double* fl,
                       const double* fly, double a2, int keep0, cx* out) {   /* src/sim_plasma.cpp:94-117 */
    const long nk = (long)nx * ny;
    cx* pk = (cx*)malloc(sizeof(cx) * 2 * (size_t)nk);
    for (long i = 0; i < nk; ++i) { pk[i] = mk(J[i], J[nk + i]); pk[nk + i] = mk(J[2 * nk + i], 0); }
    or_fft_many(nx, ny, 1, pk, 2, 0);
    cx* ox = out; cx* oy = out + nk; cx* oz = out + 2 * nk;
    for (long i = 0; i < nk; ++i) {
        const int ixx = (int)(i % nx), j = (int)(i / nx);
        const long m = (long)((ny - j) % ny) * nx + (nx - ixx) % nx;
        const cx z = pk[i], zm = cjg(pk[m]);
        const double k2 = kxs[ixx] * kxs[ixx] + kys[j] * kys[j];
        const double f = (ixx == 0 && j == 0 && !keep0) ? 0.0 : fl[ixx] * fly[j] / (1 + a2 * k2);
        const cx d = sub(z, zm);
        ox[i] = scl(f, scl(0.5, add(z, zm)));
        oy[i] = scl(f, scl(0.5, mk(d.im, -d.re)));
        oz[i] = scl(f, pk[nk + i]);
    }
    free(pk);
}

void or_real_fields(int nx, int ny, int count, const cx* rx, const cx* ry, const cx* lx, const cx* ly, const cx* ez,
                    const cx* bz, float* F, long fpl) {   /* src/sim_fields.cpp:270-293 */
    const long nk = (long)nx * ny, n = (long)count * nk;
    cx* s1 = (cx*)malloc(sizeof(cx) * (size_t)n);
    cx* s2 = (cx*)malloc(sizeof(cx) * (size_t)n);
    for (long i = 0; i < n; ++i) {
        s1[i] = mk(rx[i].re - ry[i].im, rx[i].im + ry[i].re);
        s2[i] = mk(lx[i].re - ly[i].im, lx[i].im + ly[i].re);
    }
    or_fft_many(nx, ny, 1, s1, count, 1);
    or_fft_many(nx, ny, 1, s2, count, 1);
    for (long i = 0; i < n; ++i) {
        const double r1 = s1[i].re, r2 = s1[i].im, l1 = s2[i].re, l2 = s2[i].im;
        F[0 * fpl + i] = (float)(0.5 * (r1 + l1));
        F[1 * fpl + i] = (float)(0.5 * (r2 + l2));
        F[3 * fpl + i] = (float)(0.5 * (l2 - r2));
        F[4 * fpl + i] = (float)(0.5 * (r1 - l1));
    }
    for (long i = 0; i < n; ++i) s1[i] = mk(ez[i].re - bz[i].im, ez[i].im + bz[i].re);
    or_fft_many(nx, ny, 1, s1, count, 1);
    for (long i = 0; i < n; ++i) { F[2 * fpl + i] = (float)s1[i].re; F[5 * fpl + i] = (float)s1[i].im; }
    free(s1); free(s2);
}
