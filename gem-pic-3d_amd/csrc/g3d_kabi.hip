// Kernel-level C ABI (include/g3d.h): the reference's device contexts (PushCtx, StreamCtx) and the member functions
// of the drop-in files (push, the streaming loop, stream_spectra, real_fields_slice, currents, the window centroid
// and shift) over caller-owned device buffers in the reference's layouts.  Each entry point converts its POD
// context into the engine's launch arguments and runs the same kernels as the engine's member function of the
// same name; nothing here allocates except the beam deposit's sort buffers and the centroid's partial sums, which
// the library holds between calls.
#include <stdexcept>
#include "g3d_kernels.hpp"

namespace g3d {
namespace {

static_assert(sizeof(g3d_adk_level) == sizeof(AdkLevel) && offsetof(g3d_adk_level, Emin) == offsetof(AdkLevel, Emin),
              "g3d_adk_level mirrors AdkLevel");
static_assert(sizeof(g3d_ion_dev) == sizeof(IonDev) && offsetof(g3d_ion_dev, off) == offsetof(IonDev, off) &&
                  offsetof(g3d_ion_dev, Emin) == offsetof(IonDev, Emin),
              "g3d_ion_dev mirrors IonDev");

Axis axis_of(const g3d_axis& a) {
    Axis x{};
    x.X = a.X; x.n = a.n; x.uniform = a.uniform; x.x0 = a.x0; x.ih = a.ih;
    if (!a.uniform && (!a.X || a.n < 2)) throw std::invalid_argument("g3d axis: a table axis needs X and n >= 2");
    return x;
}
PlaneWave wave_of(const g3d_plane_wave& w) {
    PlaneWave p{};
    p.on = w.on; p.pol = w.pol; p.a0 = w.a0; p.kL = w.kL; p.tau = w.tau; p.cep = w.cep; p.eta_c = w.eta_c;
    p.kx = w.kx; p.kz = w.kz;
    return p;
}
FieldView fields_of(const g3d_push_ctx& c) {
    if (c.nx < 1 || c.ny < 1 || c.nk != long(c.nx) * c.ny || c.ns < 1 || (c.self && !c.F))
        throw std::invalid_argument("g3d push context: bad grid or missing F");
    FieldView f{};
    f.F = c.F; f.a = c.a; f.ns = c.ns; f.nx = c.nx; f.ny = c.ny; f.yf = c.yf; f.self = c.self; f.nk = c.nk;
    f.dx = c.dx; f.dy = c.dy; f.Lx = c.Lx; f.Ly = c.Ly; f.zax = axis_of(c.zax); f.yax = axis_of(c.yax);
    return f;
}
PushArgs push_of(const g3d_push_ctx& c) {
    PushArgs a{};
    a.fv = fields_of(c); a.laser = wave_of(c.laser);
    a.X0 = c.X0; a.Y0 = c.Y0; a.zc = c.zc; a.dz_push = c.dz_push; a.push_phase = c.push_phase; a.zhand = c.zhand;
    a.zmax = c.zmax; a.ylo = c.ylo; a.yhi = c.yhi; a.s1 = c.s1; a.vz_min = c.vz_min; a.hmin = c.hmin;
    a.freeze = c.freeze;
    return a;
}
SpecGeom geom_of(const g3d_spec_ctx& c) {
    if (!c.plan || c.plan->p.nx != c.nx || c.plan->p.ny != c.ny || !c.plan->p.two_d)
        throw std::invalid_argument("g3d spectral context: the plan must be FftMany(nx, ny, two_d = 1)");
    if (c.nk != long(c.nx) * c.ny || !c.kxs || !c.kys || !c.filt || !c.filty)
        throw std::invalid_argument("g3d spectral context: bad grid or missing k tables");
    SpecGeom g{};
    g.nx = c.nx; g.ny = c.ny; g.ns = c.ns; g.N = c.N; g.a = c.a; g.nk = c.nk;
    g.kx = c.kx; g.ky = c.ky; g.kxs = c.kxs; g.kys = c.kys; g.filt = c.filt; g.filty = c.filty; g.cm = 0;
    return g;
}
void need(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string("g3d: missing ") + what);
}

DepositScratch& deposit_scratch() {
    static DepositScratch s;          // the sort buffers of the ordered beam deposit, grown on demand
    return s;
}
}  // namespace
}  // namespace g3d

extern "C" {
int g3d_bunch_push(const g3d_push_ctx* c, const g3d_parts* p, long i0, long i1, double s0) {
    return g3d::guard([&] {
        g3d::need(c, "push context");
        g3d::need(p, "particles");
        if (i0 < 0 || i1 < i0) throw std::invalid_argument("g3d_bunch_push: bad range");
        if (i1 == i0) return;
        g3d_parts q = *p;               // the range [i0, i1) as a list of its own
        q.x += i0; q.y += i0; q.zeta += i0; q.px += i0; q.py += i0; q.pz += i0; q.w += i0; q.z += i0; q.q += i0;
        q.qm += i0; q.rem += i0; q.alive += i0; q.sp += i0; q.id += i0;
        g3d::bunch_push(g3d::push_of(*c), q, i1 - i0, s0);
    });
}

int g3d_stream_halfcell(const g3d_stream_ctx* c, const g3d_parts* src, const g3d_parts* dst, long nstrm) {
    return g3d::guard([&] {
        g3d::need(c, "stream context");
        g3d::need(src, "source list");
        g3d::need(dst, "destination list");
        g3d::need(c->zs, "zs");
        g3d::need(c->J, "J");
        g3d::need(c->ia, "ia");
        if (c->n < 0 || c->n + 1 >= c->pc.N) throw std::invalid_argument("g3d_stream_halfcell: half-cell out of range");
        if (c->ions && (!c->adk || !c->bcnt)) throw std::invalid_argument("g3d_stream_halfcell: ions need adk and bcnt");
        g3d::StreamArgs a{};
        a.fv = g3d::fields_of(c->pc);
        a.laser = g3d::wave_of(c->pc.laser);
        a.X0 = c->pc.X0; a.Y0 = c->pc.Y0; a.zc = c->pc.zc; a.ylo = c->pc.ylo; a.yhi = c->pc.yhi;
        a.zs = c->zs; a.J = c->J; a.ia = c->ia; a.id = c->id ? c->id : c->ia;
        a.Delta = c->Delta; a.s1 = c->s1; a.sTrap = c->sTrap; a.vz_beam = c->vz_beam; a.n = c->n;
        a.ions = reinterpret_cast<const g3d::IonDev*>(c->ions);
        a.adk = reinterpret_cast<const g3d::AdkLevel*>(c->adk);
        a.born = c->born; a.bcnt = c->bcnt; a.bcap = c->bcap; a.key = c->key;
        a.npush = nullptr; a.npush_cap = 0;
        g3d::stream_halfcell(a, *src, *dst, nstrm, g3d::TileGeom{});
    });
}

int g3d_stream_spectra(const g3d_spec_ctx* c, double* jsr, g3d_cpx* jsn, g3d_cpx* scratch) {
    return g3d::guard([&] {
        g3d::need(c, "spectral context");
        g3d::need(jsr, "jsr");
        g3d::need(jsn, "jsn");
        g3d::need(scratch, "scratch");
        const g3d::SpecGeom g = g3d::geom_of(*c);
        g3d::cx* X = scratch;
        g3d::cx* T = scratch + 3 * g.nk;
        g3d::pass_rows_fwd(jsr, g3d::TileGeom{}, c->plan->p, X, nullptr, nullptr, 0, 0);
        g3d::ModeArgs m{};
        m.mv.nx = g.nx; m.mv.ny = g.ny; m.mv.nk = g.nk; m.mv.a = g.a; m.mv.N = g.N; m.mv.kx = c->kx; m.mv.ky = c->ky;
        m.kxs = c->kxs; m.kys = c->kys; m.filt = c->filt; m.filty = c->filty;
        m.a2 = c->a2; m.keep0 = c->keep0; m.ny = g.ny; m.kf = 1;
        g3d::pass_cols(m, c->plan->p, X, T, jsn, 2);
    });
}

int g3d_real_fields_slice(const g3d_spec_ctx* c, const g3d_cpx* Rx, const g3d_cpx* Ry, const g3d_cpx* Lx,
                          const g3d_cpx* Ly, const g3d_cpx* Ez, const g3d_cpx* Bz, float* F_slice, long fpl,
                          g3d_cpx* scratch) {
    return g3d::guard([&] {
        g3d::need(c, "spectral context");
        g3d::need(Rx, "Rx"); g3d::need(Ry, "Ry"); g3d::need(Lx, "Lx"); g3d::need(Ly, "Ly");
        g3d::need(Ez, "Ez"); g3d::need(Bz, "Bz"); g3d::need(F_slice, "F"); g3d::need(scratch, "scratch");
        const g3d::SpecGeom g = g3d::geom_of(*c);
        if (fpl < g.nk) throw std::invalid_argument("g3d_real_fields_slice: fpl below nk");
        g3d::fields_pack3(g, nullptr, nullptr, 0, const_cast<g3d::cx*>(Rx), const_cast<g3d::cx*>(Ry), Lx, Ly, Ez, Bz,
                          scratch, g.nk, 1);
        g3d::fft_run(c->plan->p, scratch, 3, true);
        g3d::fields_to_F(scratch, g.nk, F_slice, fpl, g.nk, 1);
    });
}

int g3d_bunch_currents(const g3d_spec_ctx* c, const g3d_axis* zax, const double* ia, const double* id, double dx,
                       double dy, int zfilter, int zuniform, const g3d_parts* p, long np, double* jw, g3d_cpx* jkx,
                       g3d_cpx* jky, g3d_cpx* jkz, g3d_cpx* scratch) {
    return g3d::guard([&] {
        g3d::need(c, "spectral context"); g3d::need(zax, "zeta axis"); g3d::need(ia, "ia"); g3d::need(p, "particles");
        g3d::need(jw, "jw"); g3d::need(jkx, "jkx"); g3d::need(jky, "jky"); g3d::need(jkz, "jkz");
        g3d::need(scratch, "scratch"); g3d::need(c->hcg, "hcg");
        const g3d::SpecGeom g = g3d::geom_of(*c);
        const long nw = 3L * (g.ns + 3) * g.nk;
        g3d::hip_check(hipMemsetAsync(jw, 0, size_t(nw) * sizeof(double), g3d::stream()), "hipMemsetAsync");
        g3d::DepositArgs d{};
        d.zax = g3d::axis_of(*zax); d.yax = g3d::Axis{}; d.yax.uniform = 1;
        d.hcg = c->hcg; d.ia = ia; d.id = id ? id : ia; d.J = jw; d.dx = dx; d.dy = dy;
        d.nh = g.ns + 3; d.a = g.a; d.N = g.N; d.nx = g.nx; d.ny = g.ny; d.yf = 1; d.nk = g.nk;
        if (np > 0) g3d::bunch_deposit(d, *p, np, g3d::deposit_scratch());
        g3d::currents_zfilter(g, jw, c->hcg, zfilter, zuniform, scratch, jkz);
        g3d::fft_run(c->plan->p, scratch, g.ns + 1, false);
        g3d::fft_run(c->plan->p, jkz, g.ns + 1, false);
        g3d::currents_unpack(g, scratch, jkx, jky, jkz, c->keep0);
    });
}

int g3d_sponge_real_fields(const g3d_spec_ctx* c, int sponge, const double* maskx, const double* masky,
                           double beta_b, double gm2_b, g3d_cpx* Rx, g3d_cpx* Ry, const g3d_cpx* Lx, const g3d_cpx* Ly,
                           const g3d_cpx* Ez, const g3d_cpx* Bz, const g3d_cpx* jkz, float* F, g3d_cpx* scratch) {
    return g3d::guard([&] {
        g3d::need(c, "spectral context"); g3d::need(Rx, "Rx"); g3d::need(Ry, "Ry"); g3d::need(Lx, "Lx");
        g3d::need(Ly, "Ly"); g3d::need(Ez, "Ez"); g3d::need(Bz, "Bz"); g3d::need(F, "F"); g3d::need(scratch, "scratch");
        if (sponge) { g3d::need(jkz, "jkz"); g3d::need(maskx, "maskx"); g3d::need(masky, "masky"); g3d::need(c->hcg, "hcg"); }
        const g3d::SpecGeom g = g3d::geom_of(*c);
        const long ts = long(g.ns) * g.nk;
        g3d::cx* S1 = scratch;
        g3d::cx* S2 = scratch + ts;
        g3d::cx* T3 = scratch + 2 * ts;
        const double opb = 1 + beta_b;
        if (sponge) {
            g3d::sponge_pack(g, Rx, Ry, jkz, c->hcg, c->wlo, c->whi, gm2_b, opb, S1, S2);
            g3d::fft_run(c->plan->p, S1, g.ns, true);
            g3d::sponge_mask(g, S1, maskx, masky, masky, g.ns);
            g3d::fft_run(c->plan->p, S1, g.ns, false);
        }
        g3d::fields_pack3(g, sponge ? S1 : nullptr, S2, opb, Rx, Ry, Lx, Ly, Ez, Bz, T3, ts, g.ns);
        g3d::fft_run(c->plan->p, T3, 3 * g.ns, true);
        g3d::fields_to_F(T3, ts, F, long(g.ns + 2) * g.nk, g.nk, g.ns);
    });
}

int g3d_comoving(const g3d_spec_ctx* c, const g3d_cpx* jkz, double beta, double gm2, g3d_cpx* Rx, g3d_cpx* Ry,
                 g3d_cpx* Ez, g3d_cpx* Bz, g3d_cpx* scratch) {
    return g3d::guard([&] {
        g3d::need(c, "spectral context"); g3d::need(jkz, "jkz"); g3d::need(Rx, "Rx"); g3d::need(Ry, "Ry");
        g3d::need(Ez, "Ez"); g3d::need(Bz, "Bz"); g3d::need(scratch, "scratch"); g3d::need(c->hcg, "hcg");
        const g3d::SpecGeom g = g3d::geom_of(*c);
        g3d::cx* S1 = scratch;
        g3d::cx* S2 = scratch + long(g.ns) * g.nk;
        g3d::comoving_solve(g, jkz, c->hcg, c->wlo, c->whi, beta, gm2, S1, S2);
        g3d::comoving_R(g, S2, beta, Rx, Ry, Ez, Bz);
    });
}

int g3d_bunch_centroid(const g3d_parts* p, long np, double X0, double Y0, double* out) {
    return g3d::guard([&] {
        g3d::need(p, "particles");
        g3d::need(out, "out");
        static g3d::DevBuf<double> dev;          // the three sums on the device
        if (dev.size() < 3) dev.resize(3);
        g3d::bunch_centroid(*p, np, X0, Y0, dev.data());
        dev.to_host(out, 0, 3);
    });
}

int g3d_window_shift(const g3d_spec_ctx* c, int nxs, int nys, double dx, double dy, g3d_cpx* const* planes,
                     const long* counts, int nplanes, float* F, float* F_scratch, const g3d_parts* p, long np) {
    return g3d::guard([&] {
        g3d::need(c, "spectral context");
        if (nplanes > 0) { g3d::need(planes, "planes"); g3d::need(counts, "counts"); }
        if (c->nk != long(c->nx) * c->ny || !c->kxs || !c->kys) throw std::invalid_argument("g3d_window_shift: bad grid");
        const double sx = nxs * dx, sy = nys * dy;
        if (p && np > 0) g3d::parts_shift(*p, np, sx, sy);
        for (int k = 0; k < nplanes; ++k) {
            g3d::need(planes[k], "plane array");
            g3d::rotate_planes(planes[k], counts[k] * c->nk, c->nk, c->nx, c->kxs, c->kys, sx, sy, nys != 0);
        }
        if (F) {
            g3d::need(F_scratch, "F scratch");
            const long n = 6L * (c->ns + 2) * c->nk;
            g3d::roll_F(F, F_scratch, n, c->nk, c->nx, c->ny, nxs, nys);
            g3d::hip_check(hipMemcpyAsync(F, F_scratch, size_t(n) * sizeof(float), hipMemcpyDeviceToDevice,
                                          g3d::stream()), "hipMemcpy D2D");
        }
        g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize");
    });
}
}
