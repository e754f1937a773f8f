// C ABI of the engine (include/g3d.h): create a Simulation of the march, run init/step or single member
// functions, and move its arrays between host and device in the reference's layouts.
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include "g3d_engine.hpp"

struct g3d_sim { g3d::Sim* s; };

namespace {
using g3d::DevBuf;
using g3d::cx;

template <class T> size_t bytes_of(const DevBuf<T>& b) { return b.size() * sizeof(T); }

// name -> (pointer to the device buffer as bytes, size in bytes); particles are handled separately
bool lookup(g3d::Sim& s, const std::string& n, void*& dev, size_t& bytes) {
    std::map<std::string, std::pair<void*, size_t>> m = {
        {"Rx", {s.Rx.data(), bytes_of(s.Rx)}}, {"Ry", {s.Ry.data(), bytes_of(s.Ry)}},
        {"Ez", {s.Ez.data(), bytes_of(s.Ez)}}, {"Bz", {s.Bz.data(), bytes_of(s.Bz)}},
        {"Lx", {s.Lx.data(), bytes_of(s.Lx)}}, {"Ly", {s.Ly.data(), bytes_of(s.Ly)}},
        {"j0x", {s.j0x.data(), bytes_of(s.j0x)}}, {"j0y", {s.j0y.data(), bytes_of(s.j0y)}},
        {"jkx", {s.jkx.data(), bytes_of(s.jkx)}}, {"jky", {s.jky.data(), bytes_of(s.jky)}},
        {"jkz", {s.jkz.data(), bytes_of(s.jkz)}}, {"js0x", {s.js0x.data(), bytes_of(s.js0x)}},
        {"js0y", {s.js0y.data(), bytes_of(s.js0y)}}, {"js0z", {s.js0z.data(), bytes_of(s.js0z)}},
        {"jw", {s.jw.data(), bytes_of(s.jw)}}, {"F", {s.F.data(), bytes_of(s.F)}},
        {"jsr", {s.jsr.data(), bytes_of(s.jsr)}}, {"jsn", {s.jsn.data(), bytes_of(s.jsn)}},
        {"rho_lines", {s.rho_lines.data(), bytes_of(s.rho_lines)}},
        {"rho_lines_i", {s.rho_lines_i.data(), bytes_of(s.rho_lines_i)}},
        // the march geometry (src/sim_setup.cpp:5-60, 148-164): k vectors with the Nyquist mode zeroed, the
        // half-cell lengths and the two-point current weights of each slice
        {"kx", {s.kxd.data(), bytes_of(s.kxd)}}, {"ky", {s.kyd.data(), bytes_of(s.kyd)}},
        {"hcg", {s.hcg.data(), bytes_of(s.hcg)}}, {"wlo", {s.wlo.data(), bytes_of(s.wlo)}},
        {"whi", {s.whi.data(), bytes_of(s.whi)}},
        // the filters, phase-shift k vectors, slice positions and area factors of the same constructor
        {"kxs", {s.kxs.data(), bytes_of(s.kxs)}}, {"kys", {s.kys.data(), bytes_of(s.kys)}},
        {"filt", {s.filt.data(), bytes_of(s.filt)}}, {"filty", {s.filty.data(), bytes_of(s.filty)}},
        {"zs", {s.zs.data(), bytes_of(s.zs)}}, {"iA", {s.iA.data(), bytes_of(s.iA)}},
        {"iD", {s.iD.data(), bytes_of(s.iD)}}, {"wsl", {s.wsl.data(), bytes_of(s.wsl)}},
        {"maskx", {s.maskx.data(), bytes_of(s.maskx)}}, {"masky", {s.masky_n.data(), bytes_of(s.masky_n)}}};
    auto it = m.find(n);
    if (it == m.end()) return false;
    dev = it->second.first; bytes = it->second.second;
    return true;
}
// "ion_statsK": the ion statistics of ion species K after the last step (src/sim_plasma.cpp:438-446: the weight
// of each charge state 0 .. Zmax of the ions still streaming, then the freed charge), a host array
const std::vector<double>* ion_stats_of(g3d::Sim& s, const std::string& n) {
    if (n.rfind("ion_stats", 0) != 0 || n.size() != 10) return nullptr;
    const int k = n[9] - '0';
    if (k < 0 || k >= int(s.ion_stats.size())) return nullptr;
    return &s.ion_stats[size_t(k)];
}
// the Simulation of a handle; a null handle is an error, not a fault
g3d::Sim& S(g3d_sim* h) {
    if (!h || !h->s) throw std::invalid_argument("null simulation handle");
    return *h->s;
}
int check_size(size_t want, size_t got) {
    if (want != got) { g3d::set_error("size mismatch: " + std::to_string(got) + " bytes, array has " + std::to_string(want)); return 2; }
    return 0;
}
}  // namespace

extern "C" {
int g3d_sim_create(const g3d_config* cfg, g3d_sim** out) {
    return g3d::guard([&] {
        *out = nullptr;
        auto* h = new g3d_sim{nullptr};
        try { h->s = new g3d::Sim(g3d::Config::from_abi(*cfg)); }
        catch (...) { delete h; throw; }
        *out = h;
    });
}
int g3d_sim_destroy(g3d_sim* sim) {
    return g3d::guard([&] { if (sim) { delete sim->s; delete sim; } });
}
int g3d_sim_info_get(g3d_sim* sim, g3d_sim_info* o) {
    return g3d::guard([&] {
        g3d::Sim& s = S(sim);
        o->N = s.N; o->ns = s.ns; o->nk = s.nk; o->np = s.np; o->nstrm = s.nstrm; o->istep = s.istep;
        o->born_cap = s.born_cap; o->s = s.s; o->dx = s.dx; o->dy = s.dy; o->X0 = s.X0; o->Y0 = s.Y0;
        o->beta_b = s.beta_b; o->gm2_b = s.gm2_b; o->trapped_total = s.trapped_total;
        o->restreamed_total = s.restreamed_total; o->lost_total = s.lost_total;
        o->pushes_stream = s.pushes_stream; o->pushes_beam = s.pushes_beam;
    });
}
int g3d_sim_init(g3d_sim* sim) { return g3d::guard([&] { S(sim).init(); g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize"); }); }
int g3d_sim_step(g3d_sim* sim) { return g3d::guard([&] { S(sim).step(); }); }

int g3d_sim_call(g3d_sim* sim, const char* op_, const double* a, int na) {
    int rc = 0;
    const int g = g3d::guard([&] {
        g3d::Sim& s = S(sim);
        const std::string op(op_);
        auto need = [&](int k) {
            if (na < k) throw std::runtime_error(op + ": needs " + std::to_string(k) + " arguments");
        };
        if (op == "init") s.init();
        else if (op == "step") s.step();
        else if (op == "march") { need(1); s.march(a[0]); }
        else if (op == "currents") s.currents();
        else if (op == "sponge_and_real_fields") s.sponge_and_real_fields();
        else if (op == "solve_fields") { need(1); s.solve_fields(a[0]); }
        else if (op == "stream_spectra") s.stream_spectra();
        else if (op == "march_slice") { need(2); s.march_slice(int(a[0]), a[1]); }
        else if (op == "real_fields_slice") { need(1); s.real_fields_slice(int(a[0])); }
        else if (op == "push") { need(2); s.push(a[0], a[1]); }
        else if (op == "march_plasma") { need(2); s.march_plasma(a[0], a[1]); }
        else if (op == "comoving_field") s.comoving_field();
        else if (op == "relax_start_field") s.relax_start_field();
        else if (op == "add_grid_lasers") s.add_grid_lasers();
        else if (op == "load_plasma") { need(2); s.load_plasma(a[0], a[1]); }
        else if (op == "stream_pass") { need(4); s.stream_pass(int(a[0]), int(a[1]), a[2], a[3]); }
        else if (op == "set_s") { need(2); s.s = a[0]; s.istep = long(a[1]); }
        else if (op == "set_starting") { need(1); s.starting_ = a[0] != 0; }
        else if (op == "j0_from_jk") s.j0_from_jk();
        else if (op == "line_pivots") { need(1); s.line_pivots_on = a[0] != 0; s.piv_ok = false; }
        else if (op == "bpush_windows") { need(1); s.bpush_windows = a[0] != 0; }
        else { g3d::set_error("unknown op " + op); rc = 3; return; }
        g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize");
    });
    return g ? g : rc;
}

int g3d_sim_array_size(g3d_sim* sim, const char* name, size_t* bytes) {
    int rc = 0;
    const int g = g3d::guard([&] {
        g3d::Sim& s = S(sim);
        const std::string n(name);
        void* d;
        if (n == "parts") *bytes = size_t(s.np) * sizeof(g3d_particle);
        else if (n == "strm") *bytes = size_t(s.nstrm) * sizeof(g3d_particle);
        else if (const std::vector<double>* h = ion_stats_of(s, n)) *bytes = h->size() * sizeof(double);
        else if (n == "laser") *bytes = 9 * sizeof(double);
        else if (!lookup(s, n, d, *bytes)) { g3d::set_error("unknown array " + n); rc = 3; }
    });
    return g ? g : rc;
}
int g3d_sim_get(g3d_sim* sim, const char* name, void* host, size_t bytes) {
    int rc = 0;
    const int g = g3d::guard([&] {
        g3d::Sim& s = S(sim);
        s.ensure_layout(false, false);      // arrays leave and enter in the reference's layout
        const std::string n(name);
        if (n == "parts" || n == "strm") {
            const std::vector<g3d_particle> v = n == "parts" ? s.host_particles() : s.host_strm();
            if ((rc = check_size(v.size() * sizeof(g3d_particle), bytes))) return;
            std::memcpy(host, v.data(), bytes);
            return;
        }
        if (n == "laser") {          // the analytic plane wave: on, pol, a0, kL, tau, cep, eta_c, kx, kz
            const g3d::PlaneWave& L = s.laser;
            const double v[9] = {double(L.on), double(L.pol), L.a0, L.kL, L.tau, L.cep, L.eta_c, L.kx, L.kz};
            if ((rc = check_size(sizeof v, bytes))) return;
            std::memcpy(host, v, sizeof v);
            return;
        }
        if (const std::vector<double>* h = ion_stats_of(s, n)) {
            if ((rc = check_size(h->size() * sizeof(double), bytes))) return;
            std::memcpy(host, h->data(), bytes);
            return;
        }
        void* d; size_t b;
        if (!lookup(s, n, d, b)) { g3d::set_error("unknown array " + n); rc = 3; return; }
        if ((rc = check_size(b, bytes))) return;
        if (!bytes) return;
        g3d::hip_check(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, g3d::stream()), "hipMemcpy D2H");
        g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize");
    });
    return g ? g : rc;
}
int g3d_sim_set(g3d_sim* sim, const char* name, const void* host, size_t bytes) {
    int rc = 0;
    const int g = g3d::guard([&] {
        g3d::Sim& s = S(sim);
        s.ensure_layout(false, false);
        s.jw_zero = false;
        s.piv_ok = false;          // the march geometry (kx, hcg) may change: the pivot table is formed again
        if (std::string(name) == "strm") {
            const g3d_particle* p = static_cast<const g3d_particle*>(host);
            s.set_strm(std::vector<g3d_particle>(p, p + bytes / sizeof(g3d_particle)));
            return;
        }
        void* d; size_t b;
        if (!lookup(s, name, d, b)) { g3d::set_error(std::string("unknown array ") + name); rc = 3; return; }
        if ((rc = check_size(b, bytes))) return;
        if (!bytes) return;
        g3d::hip_check(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, g3d::stream()), "hipMemcpy H2D");
        g3d::hip_check(hipStreamSynchronize(g3d::stream()), "synchronize");
    });
    return g ? g : rc;
}
int g3d_sim_set_particles(g3d_sim* sim, const g3d_particle* p, long n) {
    return g3d::guard([&] { S(sim).set_particles(std::vector<g3d_particle>(p, p + n)); });
}
int g3d_sim_set_joining(g3d_sim* sim, const g3d_particle* p, long n) {
    return g3d::guard([&] { S(sim).joining_.assign(p, p + n); });
}
int g3d_sim_get_joining(g3d_sim* sim, g3d_particle* p, long cap, long* n) {
    int rc = 0;
    const int g = g3d::guard([&] {
        const auto& j = S(sim).joining_;
        *n = long(j.size());
        if (*n > cap) { g3d::set_error("joining list larger than the buffer"); rc = 2; return; }
        std::memcpy(p, j.data(), j.size() * sizeof(g3d_particle));
    });
    return g ? g : rc;
}
static FILE* open_or_throw(const char* path, const char* mode) {
    if (!path) throw std::invalid_argument("missing path");
    FILE* f = std::fopen(path, mode);
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    return f;
}
int g3d_sim_write_bunches(g3d_sim* sim, const char* path, int header) {
    return g3d::guard([&] {
        g3d::Sim& s = S(sim);
        FILE* f = open_or_throw(path, "a");
        try { s.write_bunches(f, header != 0); } catch (...) { std::fclose(f); throw; }
        std::fclose(f);
    });
}
int g3d_sim_write_slices(g3d_sim* sim, const char* path, double len) {
    return g3d::guard([&] { if (!path) throw std::invalid_argument("missing path"); S(sim).write_slices(path, len); });
}
int g3d_sim_diagnostics(g3d_sim* sim, const char* path, int header) {
    return g3d::guard([&] {
        g3d::Sim& s = S(sim);
        FILE* f = open_or_throw(path, "a");
        try { s.diagnostics(f, header != 0); } catch (...) { std::fclose(f); throw; }
        std::fclose(f);
    });
}
int g3d_sim_field_energy(g3d_sim* sim, double* out) {
    return g3d::guard([&] { if (!out) throw std::invalid_argument("missing out"); *out = S(sim).total_field_energy(); });
}
int g3d_sim_dump_particles(g3d_sim* sim, const char* path) {
    return g3d::guard([&] { if (!path) throw std::invalid_argument("missing path"); S(sim).dump_particles(path); });
}
int g3d_sim_write_frame(g3d_sim* sim, const char* path, int comps) {
    return g3d::guard([&] { if (!path) throw std::invalid_argument("missing path"); S(sim).write_frame(path, comps); });
}
int g3d_sim_write_axis(g3d_sim* sim, const char* path) {
    return g3d::guard([&] { if (!path) throw std::invalid_argument("missing path"); S(sim).write_axis(path); });
}
int g3d_sim_write_checkpoint(g3d_sim* sim, const char* dir) {
    return g3d::guard([&] { if (!dir) throw std::invalid_argument("missing dir"); S(sim).write_checkpoint(dir); });
}
int g3d_sim_read_checkpoint(g3d_sim* sim, const char* dir, long step) {
    return g3d::guard([&] { if (!dir) throw std::invalid_argument("missing dir"); S(sim).read_checkpoint(dir, step); });
}
int g3d_sim_gradient(g3d_sim* sim, int steps, int every, double* loss, double* genv, int nenv, double* gprof,
                     int nprof) {
    return g3d::guard([&] {
        if (!loss || (nenv > 0 && !genv) || (nprof > 0 && !gprof)) throw std::invalid_argument("gradient: missing output");
        S(sim).gradient(steps, every, loss, genv, nenv, gprof, nprof);
    });
}
int g3d_sim_tangent(g3d_sim* sim, int steps, const double* denv, int nenv, const double* dprof, int nprof,
                    double* loss, double* dloss) {
    return g3d::guard([&] {
        if (!loss || !dloss || (nenv > 0 && !denv) || (nprof > 0 && !dprof))
            throw std::invalid_argument("tangent: missing direction or output");
        S(sim).tangent(steps, denv, nenv, dprof, nprof, loss, dloss);
    });
}
int g3d_sim_timing(g3d_sim* sim, int enable) {
    return g3d::guard([&] {
        g3d::Sim& s = S(sim);
        s.flush_timers();
        s.timing = unsigned(enable);
        for (int k = 0; k < g3d::Sim::kFam; ++k) { s.fam_ms[k] = 0; s.fam_n[k] = 0; s.fam_seen[k] = 0; }
    });
}
int g3d_sim_timing_every(g3d_sim* sim, int every) {
    return g3d::guard([&] {
        if (every < 1) throw std::invalid_argument("timing: every must be at least 1");
        S(sim).timing_every = every;
    });
}
int g3d_sim_timers(g3d_sim* sim, double* ms, long* counts, int nfam, char* names, int name_len) {
    return g3d::guard([&] {
        g3d::Sim& s = S(sim);
        s.flush_timers();
        for (int k = 0; k < nfam && k < g3d::Sim::kFam; ++k) {
            ms[k] = s.fam_ms[k]; counts[k] = s.fam_n[k];
            if (names && name_len > 0) {
                std::strncpy(names + size_t(k) * name_len, g3d::kFamilyNames[k], size_t(name_len) - 1);
                names[size_t(k) * name_len + name_len - 1] = 0;
            }
        }
    });
}
}
