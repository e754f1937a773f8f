// Test-only harness around the reference's own CPU build (oracle/_ref): builds a reference Config from the
// g3d_config POD, owns a reference Simulation and exposes its member functions and arrays over a C ABI, so that
// the tests can run the reference and the MI355X library on the same inputs.  Compiled against the sources under
// /root/reference (never copied); only tests/, smoke() and bench.py's CPU baseline load the result.
#include "deck.hpp"
#include "sim.hpp"
#include "sim_stream.hpp"
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#include <omp.h>
#include "../include/g3d.h"

static_assert(sizeof(Particle) == sizeof(g3d_particle), "particle record");
static_assert(sizeof(cpx) == sizeof(g3d_cpx), "cpx");

namespace {
thread_local std::string err_;
Comm comm_;

struct Ref {
    Config cfg;
    std::unique_ptr<Simulation> sim;
};

int fail(const std::exception& e) { err_ = e.what(); return 1; }

// one pass of the streaming particles through half-cell n, with the context march_plasma builds
// (src/sim_plasma.cpp:280-291, 327-337); jsr is zeroed first
void stream_pass(Simulation& s, int n, int pass, double s0, double s1) {
    const Config& cfg = s.cfg;
    const double Delta = s1 - s0;
    StreamCtx sc;
    {
        PushCtx& c = sc.pc;
        c.F = s.F.data(); c.a = s.a; c.ns = s.ns; c.N = s.N; c.nx = cfg.nx; c.ny = cfg.ny; c.nk = s.nk; c.self = 1;
        c.yf = s.yf; c.dx = s.dx; c.dy = s.dy; c.Lx = cfg.Lx; c.Ly = cfg.Ly; c.zax = s.zax; c.yax = s.yax;
        c.laser = s.laser; c.X0 = s.X0; c.zc = cfg.zeta_center; c.dz_push = cfg.dz_push; c.push_phase = cfg.push_phase;
        c.zhand = -1e300; c.zmax = cfg.zeta[s.N - 1]; c.s1 = s1; c.vz_min = 0; c.freeze = 0; c.hmin = 1e300; c.Y0 = s.Y0;
        c.ext = s.external();
        c.ylo = s.yf ? -1e300 : cfg.ynode[1]; c.yhi = s.yf ? 1e300 : cfg.ynode[cfg.ny - 1];
    }
    sc.zs = s.zs.data(); sc.J = s.jsr.data(); sc.ia = s.iA.data(); sc.id = s.iD.data();
    sc.Delta = Delta; sc.s1 = s1; sc.sTrap = s1 + (cfg.trap_advance - 1) * Delta; sc.vz_beam = cfg.vz_beam;
    sc.ions = s.ion_dev.data(); sc.adk = s.adk_dev.data(); sc.born = s.born.data(); sc.bcnt = s.bcnt.data();
    sc.bcap = 0; sc.key = 0; sc.n = n;
    s.jsr.zero();
    if (s.born_cap > 0) {
        s.bcnt.zero();
        sc.key = ion_key_slice(ion_key_step(cfg.seed, s.istep), n, pass);
        sc.bcap = int(std::min<long>(long(s.born.size()), long(s.strm.size()) - s.nstrm));
    }
    Particle* P = s.strm.data();
    const StreamCtx scc = sc;
    par_for(s.nstrm, [=](long i) { stream_one(scc, P[i]); });
}

// the text of a piecewise-linear function as the sum of its hat functions, the numbers printed to round-trip:
// v[0]*max(0,1-abs(zeta-x_0)/h)+v[1]*max(...)+...
std::string hat_text(const double* v, int n, double x0, double h) {
    std::string t;
    char b[256];
    for (int k = 0; k < n; ++k) {
        const double xk = x0 + k * h;
        std::snprintf(b, sizeof b, "%s(%.17g)*max(0,1-abs(zeta-(%.17g))/(%.17g))", k ? "+" : "", v[k], xk, h);
        t += b;
    }
    return t;
}
Expr hat_expr(const double* v, int n, double x0, double h) {
    return Expr(hat_text(v, n, x0, h), {"zeta"}, {}, [](const std::string&, double&) { return false; });
}

Config make_config(const g3d_config& g) {
    Config c = Config::base_units(g.lambda0);
    c.nx = g.nx; c.Lx = g.Lx;
    c.ny = g.ny; c.Ly = g.Ly; c.y_fourier = g.y_fourier != 0;
    c.nzeta = g.nzeta;
    c.zeta.assign(g.zeta, g.zeta + g.nzeta);
    c.ynode.assign(g.ynode, g.ynode + g.ny + 1);
    c.Lwin = c.zeta.back();
    c.dz = g.dz; c.z_end = g.z_end; c.dz_push = g.dz_push;
    c.sponge = g.sponge; c.sponge_y = g.sponge_y; c.sponge_rate = g.sponge_rate;
    c.move_x = g.move_x; c.move_y = g.move_y; c.window_lag = g.window_lag;
    c.self_fields = g.self_fields; c.zeta_filter = g.zeta_filter; c.k_filter = g.k_filter;
    c.push_phase = g.push_phase;
    c.laser_on = g.laser_on; c.a0 = g.a0; c.kL = g.kL; c.theta = g.theta; c.tau = g.tau; c.cep = g.cep;
    c.pol = g.laser_pol == 0 ? "s" : g.laser_pol == 1 ? "p" : "circular";
    c.start_fwhm = g.start_fwhm; c.stop_fwhm = g.stop_fwhm;
    c.charge = g.charge; c.gamma = g.gamma; c.sigx = g.sigx; c.sigy = g.sigy; c.sigz = g.sigz;
    c.emitx = g.emitx; c.emity = g.emity; c.espread = g.espread; c.zeta_center = g.zeta_center;
    c.nmacro = g.nmacro; c.seed = g.seed;
    for (int k = 0; k < g.n_glasers; ++k) {
        Config::GridLaser L;
        const g3d_glaser& s = g.glasers[k];
        L.a0 = s.a0; L.kL = s.kL; L.waist = s.waist; L.tau = s.tau; L.zeta_c = s.zeta_c; L.z_focus = s.z_focus;
        L.cep = s.cep; L.pol_angle = s.pol_angle; L.circular = s.circular; L.helicity = s.helicity;
        if (s.n_env > 0) L.envelope = hat_expr(s.env, s.n_env, s.env_z0, s.env_dz);
        c.glasers.push_back(L);
    }
    c.plasma_on = g.plasma_on; c.n0 = g.n0; c.r_max = g.r_max; c.ppc_x = g.ppc_x; c.ppc_y = g.ppc_y;
    c.smooth_length = g.smooth_length;
    c.pz_points.assign(g.pz_points, g.pz_points + g.n_pz);
    c.pz_values.assign(g.pz_values, g.pz_values + g.n_pz);
    c.parabolic = g.parabolic; c.channel_depth = g.channel_depth; c.r_channel = g.r_channel;
    c.iterations = g.iterations; c.vz_min = g.vz_min; c.vz_beam = g.vz_beam; c.trap_advance = g.trap_advance;
    c.predictor_extrapolate = g.predictor_extrapolate; c.free_electrons = g.free_electrons;
    for (int k = 0; k < g.n_ions; ++k) {
        Config::IonSpecies s;
        const g3d_ion& i = g.ions[k];
        s.name = i.name[0] ? std::string(i.name, strnlen(i.name, sizeof i.name)) : "ion" + std::to_string(k + 1);
        s.mass = i.mass; s.atoms = i.atoms; s.Z0 = i.Z0; s.Zmax = i.Zmax; s.mobile = i.mobile; s.ionization = i.ionization;
        s.ip_eV.assign(i.ip_eV, i.ip_eV + i.n_ip);
        c.ions.push_back(s);
    }
    c.freeze_bunch = g.freeze_bunch; c.start_relax = g.start_relax; c.box_mean = g.box_mean;
    if (g.n_prof > 0 || g.n_bunch_more > 0) {   // the bunch list: the first bunch (distribution_z = function
        Config::Bunch u;                            // with profile_z), then the further bunch blocks
        u.charge = g.charge; u.gamma = g.gamma; u.sigx = g.sigx; u.sigy = g.sigy; u.sigz = g.sigz;
        u.emitx = g.emitx; u.emity = g.emity; u.espread = g.espread; u.nmacro = g.nmacro; u.seed = g.seed;
        u.zeta_center = g.zeta_center;
        if (g.n_prof > 0) {
            u.zeta_min = g.prof_zmin; u.zeta_max = g.prof_zmax;
            u.profile_z = hat_expr(g.prof, g.n_prof, g.prof_zmin, (g.prof_zmax - g.prof_zmin) / (g.n_prof - 1));
            u.zeta_center = 0.5 * (u.zeta_min + u.zeta_max);
            c.zeta_center = u.zeta_center;
        }
        c.bunches.push_back(u);
        for (int k = 0; k < g.n_bunch_more; ++k) {
            const g3d_bunch& b = g.bunch_more[k];
            Config::Bunch v;
            v.name = b.name[0] ? std::string(b.name, strnlen(b.name, sizeof b.name)) : "bunch" + std::to_string(k + 2);
            v.charge = b.charge; v.gamma = b.gamma; v.sigx = b.sigx; v.sigy = b.sigy; v.sigz = b.sigz;
            v.emitx = b.emitx; v.emity = b.emity; v.espread = b.espread; v.zeta_center = b.zeta_center;
            v.nmacro = b.nmacro; v.seed = b.seed;
            c.bunches.push_back(v);
        }
    }
    return c;
}

// the freed electrons stored by the last stream pass (the count bcnt, capped at the born buffer)
size_t born_count(Simulation& s) {
    if (s.bcnt.size() == 0) return 0;
    int c = 0;
    s.bcnt.to_host(&c, 0, 1);
    return std::min(size_t(c < 0 ? 0 : c), s.born.size());
}

template <class T> int copy_out(const Buf<T>& b, void* out, size_t bytes) {
    if (bytes != b.size() * sizeof(T)) { err_ = "size mismatch"; return 2; }
    b.to_host(static_cast<T*>(out), 0, b.size());
    return 0;
}
template <class T> int copy_in(Buf<T>& b, const void* in, size_t bytes) {
    if (bytes != b.size() * sizeof(T)) { err_ = "size mismatch"; return 2; }
    b.from_host(static_cast<const T*>(in), 0, b.size());
    return 0;
}
}  // namespace

extern "C" {
const char* ref_last_error() { return err_.c_str(); }
void ref_set_threads(int n) { omp_set_num_threads(n); }

void* ref_sim_create(const g3d_config* g) {
    try {
        auto* r = new Ref;
        r->cfg = make_config(*g);
        r->sim.reset(new Simulation(r->cfg, comm_));
        return r;
    } catch (const std::exception& e) { err_ = e.what(); return nullptr; }
}
void ref_sim_destroy(void* h) { delete static_cast<Ref*>(h); }

int ref_sim_info(void* h, g3d_sim_info* o) {
    Simulation& s = *static_cast<Ref*>(h)->sim;
    o->N = s.N; o->ns = s.ns; o->nk = s.nk; o->np = s.np; o->nstrm = s.nstrm; o->istep = s.istep;
    o->born_cap = s.born_cap; o->s = s.s; o->dx = s.dx; o->dy = s.dy; o->X0 = s.X0; o->Y0 = s.Y0;
    o->beta_b = s.beta_b; o->gm2_b = s.gm2_b; o->trapped_total = s.trapped_total;
    o->restreamed_total = s.restreamed_total; o->lost_total = s.lost_total;
    o->pushes_stream = 0; o->pushes_beam = 0;      // not counted by the reference
    return 0;
}

int ref_sim_call(void* h, const char* op_, const double* a, int na) {
    try {
        Simulation& s = *static_cast<Ref*>(h)->sim;
        const std::string op(op_);
        auto need = [&](int k) { if (na < k) throw std::runtime_error(op + ": needs " + std::to_string(k) + " arguments"); };
        if (op == "init") s.init();
        else if (op == "step") s.step();
        else if (op == "march") { need(1); s.march(a[0]); }
        else if (op == "currents") s.currents();
        else if (op == "sponge_and_real_fields") s.sponge_and_real_fields();
        else if (op == "solve_fields") { need(1); s.solve_fields(a[0]); }
        else if (op == "stream_spectra") s.stream_spectra();
        else if (op == "march_slice") { need(2); s.march_slice(int(a[0]), a[1]); }
        else if (op == "real_fields_slice") { need(1); s.real_fields_slice(int(a[0])); }
        else if (op == "push") { need(2); s.push(a[0], a[1], 0, s.np); }
        else if (op == "march_plasma") { need(2); s.march_plasma(a[0], a[1]); }
        else if (op == "comoving_field") s.comoving_field();
        else if (op == "relax_start_field") s.relax_start_field();
        else if (op == "add_grid_lasers") s.add_grid_lasers();
        else if (op == "load_plasma") { need(2); s.load_plasma(a[0], a[1], s.joining_, nullptr); }
        else if (op == "set_s") { need(2); s.s = a[0]; s.istep = long(a[1]); }
        else if (op == "set_starting") { need(1); s.starting_ = a[0] != 0; }
        else if (op == "stream_pass") { need(4); stream_pass(s, int(a[0]), int(a[1]), a[2], a[3]); }
        else if (op == "shift_window") { need(2); s.shift_window(int(a[0]), int(a[1])); }
        else if (op == "reserve_born") {
            // room for `cap` freed electrons behind the streaming list, as load_plasma leaves it
            // (src/sim_plasma.cpp:82-88): born_cap, the born buffer, and the list's spare capacity
            need(1);
            const long cap = long(a[0]);
            std::vector<Particle> v(size_t(s.nstrm));
            s.strm.to_host(v.data(), 0, v.size());
            s.strm.resize(size_t(s.nstrm + cap)); s.strm_bk.resize(s.strm.size());
            s.strm.from_host(v.data(), 0, v.size());
            if (s.born.size() < size_t(cap)) s.born.resize(size_t(cap));
            s.born_cap = cap;
        }
        else if (op == "j0_from_jk") {
            const long n = long(s.ns + 1) * s.nk;
            std::vector<cpx> t(n);
            s.jkx.to_host(t.data(), 0, n); s.j0x.from_host(t.data(), 0, n);
            s.jky.to_host(t.data(), 0, n); s.j0y.from_host(t.data(), 0, n);
        } else { err_ = "unknown op " + op; return 3; }
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

// the reference's output and checkpoint member functions: what = bunches | diag (appended to path, arg: header),
// slices (darg: slice length), energy (*out), particles, frame (arg: comps), axis, write_ckpt / read_ckpt (path: the
// directory, arg: the step)
int ref_sim_output(void* h, const char* what_, const char* path, long arg, double darg, double* out) {
    try {
        Simulation& s = *static_cast<Ref*>(h)->sim;
        const std::string what(what_);
        if (what == "bunches" || what == "diag") {
            FILE* f = std::fopen(path, "a");
            if (!f) { err_ = std::string("cannot write ") + path; return 1; }
            if (what == "bunches") s.write_bunches(f, arg != 0); else s.diagnostics(f, arg != 0);
            std::fclose(f);
        } else if (what == "slices") { s.cfg.slice_length = darg; s.write_slices(path); }
        else if (what == "energy") *out = s.total_field_energy();
        else if (what == "particles") s.dump_particles(path);
        else if (what == "frame") s.write_frame(path, int(arg));
        else if (what == "axis") s.write_axis(path);
        else if (what == "write_ckpt") { s.cfg.dir = path; s.write_checkpoint(); }
        else if (what == "read_ckpt") { s.cfg.dir = path; s.read_checkpoint(arg); }
        else { err_ = "unknown output " + what; return 3; }
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

int ref_sim_array_size(void* h, const char* name_, size_t* bytes) {
    Simulation& s = *static_cast<Ref*>(h)->sim;
    const std::string n(name_);
    const std::map<std::string, size_t> m = {
        {"Rx", s.Rx.size() * 16}, {"Ry", s.Ry.size() * 16}, {"Ez", s.Ez.size() * 16}, {"Bz", s.Bz.size() * 16},
        {"Lx", s.Lx.size() * 16}, {"Ly", s.Ly.size() * 16}, {"j0x", s.j0x.size() * 16}, {"j0y", s.j0y.size() * 16},
        {"jkx", s.jkx.size() * 16}, {"jky", s.jky.size() * 16}, {"jkz", s.jkz.size() * 16},
        {"js0x", s.js0x.size() * 16}, {"js0y", s.js0y.size() * 16}, {"js0z", s.js0z.size() * 16},
        {"jw", s.jw.size() * 8}, {"F", s.F.size() * 4}, {"jsr", s.jsr.size() * 8}, {"jsn", s.jsn.size() * 16},
        {"rho_lines", s.rho_lines.size() * 8},
        {"parts", size_t(s.np) * sizeof(Particle)}, {"strm", size_t(s.nstrm) * sizeof(Particle)},
        {"born", born_count(s) * sizeof(Particle)}};
    auto it = m.find(n);
    if (it == m.end()) { err_ = "unknown array " + n; return 3; }
    *bytes = it->second;
    return 0;
}

int ref_sim_get(void* h, const char* name_, void* out, size_t bytes) {
    try {
        Simulation& s = *static_cast<Ref*>(h)->sim;
        const std::string n(name_);
        if (n == "Rx") return copy_out(s.Rx, out, bytes);
        if (n == "Ry") return copy_out(s.Ry, out, bytes);
        if (n == "Ez") return copy_out(s.Ez, out, bytes);
        if (n == "Bz") return copy_out(s.Bz, out, bytes);
        if (n == "Lx") return copy_out(s.Lx, out, bytes);
        if (n == "Ly") return copy_out(s.Ly, out, bytes);
        if (n == "j0x") return copy_out(s.j0x, out, bytes);
        if (n == "j0y") return copy_out(s.j0y, out, bytes);
        if (n == "jkx") return copy_out(s.jkx, out, bytes);
        if (n == "jky") return copy_out(s.jky, out, bytes);
        if (n == "jkz") return copy_out(s.jkz, out, bytes);
        if (n == "js0x") return copy_out(s.js0x, out, bytes);
        if (n == "js0y") return copy_out(s.js0y, out, bytes);
        if (n == "js0z") return copy_out(s.js0z, out, bytes);
        if (n == "jw") return copy_out(s.jw, out, bytes);
        if (n == "F") return copy_out(s.F, out, bytes);
        if (n == "jsr") return copy_out(s.jsr, out, bytes);
        if (n == "jsn") return copy_out(s.jsn, out, bytes);
        if (n == "rho_lines") return copy_out(s.rho_lines, out, bytes);
        if (n == "parts") {
            if (bytes != size_t(s.np) * sizeof(Particle)) { err_ = "size mismatch"; return 2; }
            s.parts.to_host(static_cast<Particle*>(out), 0, size_t(s.np));
            return 0;
        }
        if (n == "strm") {
            if (bytes != size_t(s.nstrm) * sizeof(Particle)) { err_ = "size mismatch"; return 2; }
            s.strm.to_host(static_cast<Particle*>(out), 0, size_t(s.nstrm));
            return 0;
        }
        if (n == "born") {
            const size_t c = born_count(s);
            if (bytes != c * sizeof(Particle)) { err_ = "size mismatch"; return 2; }
            s.born.to_host(static_cast<Particle*>(out), 0, c);
            return 0;
        }
        err_ = "unknown array " + n;
        return 3;
    } catch (const std::exception& e) { return fail(e); }
}

int ref_sim_set(void* h, const char* name_, const void* in, size_t bytes) {
    try {
        Simulation& s = *static_cast<Ref*>(h)->sim;
        const std::string n(name_);
        if (n == "Rx") return copy_in(s.Rx, in, bytes);
        if (n == "Ry") return copy_in(s.Ry, in, bytes);
        if (n == "Ez") return copy_in(s.Ez, in, bytes);
        if (n == "Bz") return copy_in(s.Bz, in, bytes);
        if (n == "Lx") return copy_in(s.Lx, in, bytes);
        if (n == "Ly") return copy_in(s.Ly, in, bytes);
        if (n == "j0x") return copy_in(s.j0x, in, bytes);
        if (n == "j0y") return copy_in(s.j0y, in, bytes);
        if (n == "jkx") return copy_in(s.jkx, in, bytes);
        if (n == "jky") return copy_in(s.jky, in, bytes);
        if (n == "jkz") return copy_in(s.jkz, in, bytes);
        if (n == "js0x") return copy_in(s.js0x, in, bytes);
        if (n == "js0y") return copy_in(s.js0y, in, bytes);
        if (n == "js0z") return copy_in(s.js0z, in, bytes);
        if (n == "jw") return copy_in(s.jw, in, bytes);
        if (n == "F") return copy_in(s.F, in, bytes);
        if (n == "jsr") return copy_in(s.jsr, in, bytes);
        if (n == "strm") {
            const size_t cnt = bytes / sizeof(Particle);
            std::vector<Particle> v(static_cast<const Particle*>(in), static_cast<const Particle*>(in) + cnt);
            if (s.strm.size() < cnt) { s.strm.resize(cnt); s.strm_bk.resize(cnt); }
            s.strm.from_host(v.data(), 0, cnt);
            s.nstrm = long(cnt);
            return 0;
        }
        err_ = "unknown array " + n;
        return 3;
    } catch (const std::exception& e) { return fail(e); }
}

int ref_sim_set_particles(void* h, const g3d_particle* p, long n) {
    try {
        Simulation& s = *static_cast<Ref*>(h)->sim;
        std::vector<Particle> v(reinterpret_cast<const Particle*>(p), reinterpret_cast<const Particle*>(p) + n);
        s.set_particles(v);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}
int ref_sim_set_joining(void* h, const g3d_particle* p, long n) {
    Simulation& s = *static_cast<Ref*>(h)->sim;
    s.joining_.assign(reinterpret_cast<const Particle*>(p), reinterpret_cast<const Particle*>(p) + n);
    return 0;
}
int ref_sim_get_joining(void* h, g3d_particle* p, long cap, long* n) {
    Simulation& s = *static_cast<Ref*>(h)->sim;
    *n = long(s.joining_.size());
    if (*n > cap) { err_ = "joining list larger than the buffer"; return 2; }
    std::memcpy(p, s.joining_.data(), s.joining_.size() * sizeof(Particle));
    return 0;
}

// the reference's own batched FFT (src/fftmany.hpp) on host data
int ref_fft_many(int nx, int ny, int two_d, g3d_cpx* data, int count, int inverse) {
    try {
        FftMany f(nx, ny, two_d != 0);
        f.run(reinterpret_cast<cpx*>(data), count, inverse != 0);
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

// ADK table of one species (src/ionization.cpp:30-53): 5 doubles per level Z0 ... Zmax-1
int ref_adk_levels(const double* ip_eV, int n_ip, int Z0, int Zmax, double E0, double omega0, double T, double* out) {
    try {
        std::vector<double> ip(ip_eV, ip_eV + n_ip);
        const std::vector<AdkLevel> L = adk_levels(ip, Z0, Zmax, E0, omega0, T);
        for (int k = 0; k < Zmax; ++k) {
            out[5 * k + 0] = L[k].logpre; out[5 * k + 1] = L[k].power; out[5 * k + 2] = L[k].expo;
            out[5 * k + 3] = L[k].Epeak; out[5 * k + 4] = L[k].Emin;
        }
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}
double ref_adk_rate_SI(double U_eV, int Zafter, double E) { return adk_rate_SI(U_eV, Zafter, E); }
// the reference's own deck reader (src/deck.cpp Deck::read, Deck::override_from; src/sim_config.cpp:69-343
// Config::from_deck) into the g3d_config POD; zeta (zcap) and ynode (ycap) receive the grid tables, out[] the
// main output group: every, beam_every, fields, beams, axis, particles, frames, slice_length, checkpoint_every,
// walltime_hours; dir (dcap bytes) the output folder
int ref_deck_config(const char* path, const char* const* ov, int nov, g3d_config* g, double* zeta, int zcap,
                    double* ynode, int ycap, double* out, char* dir, int dcap) {
    try {
        Deck d = Deck::read(path);
        for (int i = 0; i < nov; ++i) d.override_from(ov[i]);
        const Config c = Config::from_deck(d);
        if (c.nzeta > zcap || c.ny + 1 > ycap) throw std::runtime_error("grid tables larger than the buffers");
        if (c.glasers.size() > 4 || c.ions.size() > 4 || c.pz_points.size() > 16)
            throw std::runtime_error("more lasers, ions or ramp points than the POD holds");
        std::memset(g, 0, sizeof *g);
        g->lambda0 = c.lambda0;
        g->nx = c.nx; g->ny = c.ny; g->nzeta = c.nzeta; g->Lx = c.Lx; g->Ly = c.Ly; g->y_fourier = c.y_fourier;
        std::copy(c.zeta.begin(), c.zeta.end(), zeta); std::copy(c.ynode.begin(), c.ynode.end(), ynode);
        g->zeta = zeta; g->ynode = ynode;
        g->dz = c.dz; g->z_end = c.z_end; g->dz_push = c.dz_push;
        g->sponge = c.sponge; g->sponge_y = c.sponge_y; g->sponge_rate = c.sponge_rate;
        g->move_x = c.move_x; g->move_y = c.move_y; g->window_lag = c.window_lag;
        g->self_fields = c.self_fields; g->zeta_filter = c.zeta_filter; g->k_filter = c.k_filter;
        g->push_phase = c.push_phase;
        g->laser_on = c.laser_on; g->laser_pol = c.pol == "s" ? 0 : c.pol == "p" ? 1 : 2;
        g->a0 = c.a0; g->kL = c.kL; g->theta = c.theta; g->tau = c.tau; g->cep = c.cep;
        g->start_fwhm = c.start_fwhm; g->stop_fwhm = c.stop_fwhm;
        g->charge = c.charge; g->gamma = c.gamma; g->sigx = c.sigx; g->sigy = c.sigy; g->sigz = c.sigz;
        g->emitx = c.emitx; g->emity = c.emity; g->espread = c.espread; g->zeta_center = c.zeta_center;
        g->nmacro = c.nmacro; g->seed = c.seed;
        g->n_glasers = int(c.glasers.size());
        for (int k = 0; k < g->n_glasers; ++k) {
            const Config::GridLaser& L = c.glasers[k];
            g3d_glaser& s = g->glasers[k];
            s.a0 = L.a0; s.kL = L.kL; s.waist = L.waist; s.tau = L.tau; s.zeta_c = L.zeta_c; s.z_focus = L.z_focus;
            s.cep = L.cep; s.pol_angle = L.pol_angle; s.circular = L.circular; s.helicity = L.helicity;
        }
        g->plasma_on = c.plasma_on; g->n0 = c.n0; g->r_max = c.r_max; g->ppc_x = c.ppc_x; g->ppc_y = c.ppc_y;
        g->smooth_length = c.smooth_length;
        g->n_pz = int(c.pz_points.size());
        std::copy(c.pz_points.begin(), c.pz_points.end(), g->pz_points);
        std::copy(c.pz_values.begin(), c.pz_values.end(), g->pz_values);
        g->parabolic = c.parabolic; g->channel_depth = c.channel_depth; g->r_channel = c.r_channel;
        g->iterations = c.iterations; g->vz_min = c.vz_min; g->vz_beam = c.vz_beam; g->trap_advance = c.trap_advance;
        g->predictor_extrapolate = c.predictor_extrapolate; g->free_electrons = c.free_electrons;
        g->n_ions = int(c.ions.size());
        for (int k = 0; k < g->n_ions; ++k) {
            const Config::IonSpecies& s = c.ions[k];
            g3d_ion& i = g->ions[k];
            if (s.ip_eV.size() > 16) throw std::runtime_error("more ionization levels than the POD holds");
            i.mass = s.mass; i.atoms = s.atoms; i.Z0 = s.Z0; i.Zmax = s.Zmax; i.mobile = s.mobile;
            i.ionization = s.ionization; i.n_ip = int(s.ip_eV.size());
            std::copy(s.ip_eV.begin(), s.ip_eV.end(), i.ip_eV);
            std::strncpy(i.name, s.name.c_str(), sizeof i.name - 1);
        }
        g->freeze_bunch = c.freeze_bunch; g->start_relax = c.start_relax; g->box_mean = c.box_mean;
        if (c.bunches.size() > 4) throw std::runtime_error("more bunches than the POD holds");
        g->n_bunch_more = c.bunches.empty() ? 0 : int(c.bunches.size()) - 1;
        for (int k = 0; k < g->n_bunch_more; ++k) {
            const Config::Bunch& u = c.bunches[size_t(k) + 1];
            g3d_bunch& b = g->bunch_more[k];
            b.charge = u.charge; b.gamma = u.gamma; b.sigx = u.sigx; b.sigy = u.sigy; b.sigz = u.sigz;
            b.emitx = u.emitx; b.emity = u.emity; b.espread = u.espread; b.zeta_center = u.zeta_center;
            b.nmacro = u.nmacro; b.seed = u.seed;
            std::strncpy(b.name, u.name.c_str(), sizeof b.name - 1);
        }
        const auto& o = c.groups.at(0);
        const double v[10] = {double(c.every), double(o.beam_every), double(o.fields), double(o.beams),
                              double(o.axis), double(c.dump_particles), double(c.frames), c.slice_length,
                              double(c.checkpoint_every), c.walltime_hours};
        std::copy(v, v + 10, out);
        std::snprintf(dir, dcap, "%s", c.dir.c_str());
        return 0;
    } catch (const std::exception& e) { return fail(e); }
}

unsigned long long ref_mix64(unsigned long long z) { return mix64(z); }
// the key chain of the ionization draws (src/ionization.hpp:44-46)
unsigned long long ref_key_step(unsigned long long seed, long step) { return ion_key_step(seed, step); }
unsigned long long ref_key_slice(unsigned long long ks, int n, int pass) { return ion_key_slice(ks, n, pass); }
unsigned long long ref_key_particle(unsigned long long kl, unsigned long long id) { return ion_key_particle(kl, id); }
double ref_uniform01(unsigned long long k) { return uniform01(k); }
}
