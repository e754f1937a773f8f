// Engine of the per-slice march: setup, init, step and the member functions of the reference's drop-in files
// (src/sim_setup.cpp, sim_particles.cpp, sim_fields.cpp, sim_plasma.cpp, sim_ions.cpp) over the gfx950 kernels.
#include "g3d_engine.hpp"
#include "g3d_grad.hpp"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>

namespace g3d {


const char* const kFamilyNames[Sim::kFam] = {"stream_halfcell", "pass_rows_fwd", "pass_cols", "pass_rows_inv",
                                             "unused", "beam_push", "currents", "sponge_real_fields"};
enum { F_STREAM = 0, F_SFFT, F_MODES, F_IFFT, F_TOF, F_PUSH, F_CUR, F_SPONGE };

// ---------------------------------------------------------------- configuration
// a piecewise-linear function as the sum of its hat functions v[k] max(0, 1 - |x - x_k| / h), x_k = x0 + k h, summed
// in order: the arithmetic of the reference's expression of this form (an envelope_function or profile_z)
double hat_sum(const double* v, int n, double x0, double h, double x) {
    double s = 0;
    for (int k = 0; k < n; ++k) {
        const double xk = x0 + k * h;
        const double t = v[k] * std::max(0.0, 1 - std::fabs(x - xk) / h);
        s = k ? s + t : t;
    }
    return s;
}

Config Config::from_abi(const g3d_config& g) {
    Config c;
    c.lambda0 = g.lambda0;
    c.Lu = g.lambda0 / (2 * M_PI); c.Tu = c.Lu / kC;
    const double w0 = kC / c.Lu;
    c.nc = kEps0 * kMe * w0 * w0 / (kE * kE);
    c.nx = g.nx; c.ny = g.ny; c.nzeta = g.nzeta; c.Lx = g.Lx; c.Ly = g.Ly; c.y_fourier = g.y_fourier != 0;
    if (!g.zeta || !g.ynode || c.nzeta < 4 || c.nx < 2 || c.ny < 2) throw std::runtime_error("grids not set");
    c.zeta.assign(g.zeta, g.zeta + g.nzeta);
    c.ynode.assign(g.ynode, g.ynode + g.ny + 1);
    c.Lwin = c.zeta.back();
    c.dz = g.dz; c.z_end = g.z_end; c.dz_push = g.dz_push;
    c.sponge = g.sponge; c.sponge_y = g.sponge_y; c.sponge_rate = g.sponge_rate;
    c.move_x = g.move_x; c.move_y = g.move_y; c.window_lag = g.window_lag;
    c.self_fields = g.self_fields; c.zeta_filter = g.zeta_filter; c.k_filter = g.k_filter;
    c.push_phase = g.push_phase;
    c.laser_on = g.laser_on; c.pol = g.laser_pol; c.a0 = g.a0; c.kL = g.kL; c.theta = g.theta; c.tau = g.tau;
    c.cep = g.cep; c.start_fwhm = g.start_fwhm; c.stop_fwhm = g.stop_fwhm;
    c.charge = g.charge; c.gamma = g.gamma; c.sigx = g.sigx; c.sigy = g.sigy; c.sigz = g.sigz; c.emitx = g.emitx;
    c.emity = g.emity; c.espread = g.espread; c.zeta_center = g.zeta_center; c.nmacro = g.nmacro; c.seed = g.seed;
    if (g.n_prof < 0 || g.n_prof > 16 || g.n_prof == 1 || (g.n_prof && !(g.prof_zmax > g.prof_zmin)))
        throw std::invalid_argument("bunch profile: 2 ... 16 points on zeta_min < zeta_max");
    c.prof.assign(g.prof, g.prof + g.n_prof); c.prof_zmin = g.prof_zmin; c.prof_zmax = g.prof_zmax;
    if (g.n_bunch_more < 0 || g.n_bunch_more > 3) throw std::invalid_argument("bunches: at most 4");
    c.bunches.push_back(Config::Bunch{g.charge, g.gamma, g.sigx, g.sigy, g.sigz, g.emitx, g.emity, g.espread,
                                      g.zeta_center, g.nmacro, g.seed, "bunch"});
    for (int k = 0; k < g.n_bunch_more; ++k) {
        const g3d_bunch& b = g.bunch_more[k];
        std::string nm(b.name, strnlen(b.name, sizeof b.name));
        if (nm.empty()) nm = "bunch" + std::to_string(k + 2);
        for (const auto& o : c.bunches)
            if (o.name == nm) throw std::invalid_argument("bunches: the name '" + nm + "' given twice");
        c.bunches.push_back(Config::Bunch{b.charge, b.gamma, b.sigx, b.sigy, b.sigz, b.emitx, b.emity, b.espread,
                                          b.zeta_center, b.nmacro, b.seed, nm});
    }
    for (int k = 0; k < g.n_glasers; ++k)
        if (g.glasers[k].n_env < 0 || g.glasers[k].n_env > 32 || (g.glasers[k].n_env && !(g.glasers[k].env_dz > 0)))
            throw std::invalid_argument("grid laser envelope: 1 ... 32 points with a positive spacing");
    c.glasers.assign(g.glasers, g.glasers + g.n_glasers);
    c.plasma_on = g.plasma_on; c.n0 = g.n0; c.r_max = g.r_max; c.ppc_x = g.ppc_x; c.ppc_y = g.ppc_y;
    c.smooth_length = g.smooth_length;
    c.pz_points.assign(g.pz_points, g.pz_points + g.n_pz);
    c.pz_values.assign(g.pz_values, g.pz_values + g.n_pz);
    c.parabolic = g.parabolic; c.channel_depth = g.channel_depth; c.r_channel = g.r_channel;
    c.iterations = g.iterations; c.vz_min = g.vz_min; c.vz_beam = g.vz_beam; c.trap_advance = g.trap_advance;
    c.predictor_extrapolate = g.predictor_extrapolate; c.free_electrons = g.free_electrons;
    c.ions.assign(g.ions, g.ions + g.n_ions);
    c.freeze_bunch = g.freeze_bunch; c.start_relax = g.start_relax; c.box_mean = g.box_mean;
    c.mixed = g.mixed != 0;
    if (c.iterations < 1) throw std::runtime_error("numerics.iterations must be >= 1");
    return c;
}

// n = n0 f(z) g(r) (src/sim_config.cpp:49-67; the deck's density_function is not part of the march)
double Config::plasma_zfactor(double z) const {
    double f = 1;
    if (!pz_points.empty()) {
        if (z <= pz_points.front()) f = pz_values.front();
        else if (z >= pz_points.back()) f = pz_values.back();
        else {
            size_t k = 1;
            while (pz_points[k] < z) ++k;
            const double t = (z - pz_points[k - 1]) / (pz_points[k] - pz_points[k - 1]);
            f = (1 - t) * pz_values[k - 1] + t * pz_values[k];
        }
    }
    return f;
}
double Config::plasma_density(double x, double y, double z) const {
    if (!plasma_on) return 0;
    const double r2 = x * x + y * y;
    if (r2 >= r_max * r_max) return 0;
    const double f = plasma_zfactor(z);
    const double g = parabolic ? 1 + channel_depth * r2 / (r_channel * r_channel) : 1.0;
    return f * g > 0 ? n0 * f * g : 0.0;
}

// ---------------------------------------------------------------- particle storage
void PartStore::resize(size_t n, bool own) {
    cap = n;
    for (DevBuf<double>* b : {&x, &y, &zeta, &px, &py, &pz, &z, &q, &qm}) b->resize(n);
    alive.resize(n);
    if (own) { w.resize(n); rem.resize(n); sp.resize(n); id.resize(n); }
}
Parts PartStore::view(const PartStore* sh) {
    const PartStore& c = sh ? *sh : *this;
    Parts p;
    p.x = x.data(); p.y = y.data(); p.zeta = zeta.data(); p.px = px.data(); p.py = py.data(); p.pz = pz.data();
    p.z = z.data(); p.q = q.data(); p.qm = qm.data(); p.alive = alive.data();
    p.w = const_cast<double*>(c.w.data()); p.rem = const_cast<double*>(c.rem.data());
    p.sp = const_cast<int*>(c.sp.data()); p.id = const_cast<unsigned long long*>(c.id.data());
    return p;
}

// ---------------------------------------------------------------- ADK tables (src/ionization.cpp:9-53)
namespace {
constexpr double kUH = 13.605693122994, kEa = 5.14220674763e11, kWa = 4.1341373335e16;
void adk_coefficients(double U, int Zafter, double& logpre, double& power, double& expo) {
    const double x = U / kUH;
    const double nstar = Zafter / std::sqrt(x);
    const double logC2 = 2.0 * nstar * std::log(2.0) - std::log(nstar) - std::lgamma(nstar + 1.0) - std::lgamma(nstar);
    power = 2.0 * nstar - 1.0;
    logpre = std::log(kWa) + logC2 + std::log(0.5 * x) + power * std::log(2.0 * std::pow(x, 1.5) * kEa);
    expo = -(2.0 / 3.0) * std::pow(x, 1.5) * kEa;
}
double adk_raw(const AdkLevel& L, double E) {
    E = std::min(E, L.Epeak);
    return E > 0 ? std::exp(L.logpre - L.power * std::log(E) + L.expo / E) : 0.0;
}
std::vector<AdkLevel> adk_levels(const g3d_ion& s, double E0, double omega0, double T) {
    std::vector<AdkLevel> lev(std::max(s.Zmax, 0), AdkLevel{0, 0, 0, 1e300, 1e300});
    for (int k = s.Z0; k < s.Zmax; ++k) {
        if (k >= s.n_ip) throw std::runtime_error("ion: ionization potentials missing");
        AdkLevel& L = lev[k];
        double logpre, power, expo;
        adk_coefficients(s.ip_eV[k], k + 1, logpre, power, expo);
        L.power = power;
        L.logpre = logpre - std::log(omega0) - power * std::log(E0);
        L.expo = expo / E0;
        L.Epeak = power > 0 ? -L.expo / power : 1e300;
        double lo = 1e-12, hi = std::min(L.Epeak, 1e6);
        if (adk_raw(L, hi) * T < 1e-10) L.Emin = hi;
        else {
            for (int it = 0; it < 200; ++it) {
                const double mid = std::sqrt(lo * hi);
                (adk_raw(L, mid) * T < 1e-10 ? lo : hi) = mid;
            }
            L.Emin = lo;
        }
    }
    return lev;
}
}  // namespace

// ---------------------------------------------------------------- setup (src/sim_setup.cpp:5-164)
Sim::Sim(const Config& c) : cfg(c) {
    N = cfg.nzeta; nk = long(cfg.nx) * cfg.ny;
    dx = cfg.Lx / cfg.nx;
    if (int(cfg.zeta.size()) != N || int(cfg.ynode.size()) != cfg.ny + 1) throw std::runtime_error("grids not set");
    a = 0; ns = N; b = N;
    const int nx = cfg.nx, ny = cfg.ny;
    std::vector<double> hkx(nx), hsx(nx), hf(nx), hm(nx);
    auto prof = [&](double dd, double w) {
        if (w <= 0 || dd >= w) return 1.0;
        const double r = 1 - dd / w;
        return 1 - cfg.sponge_rate * r * r;
    };
    for (int ixx = 0; ixx < nx; ++ixx) {
        const int mx = ixx <= nx / 2 ? ixx : ixx - nx;
        hsx[ixx] = 2 * M_PI * mx / cfg.Lx;
        hkx[ixx] = (2 * ixx == nx) ? 0 : hsx[ixx];
        hf[ixx] = cfg.k_filter ? std::pow(std::cos(0.5 * hsx[ixx] * dx), 2) : 1.0;
        if (2 * ixx == nx) hf[ixx] = 0;
        hm[ixx] = prof(std::min(ixx * dx, cfg.Lx - ixx * dx), cfg.sponge);
    }
    auto up = [](DevBuf<double>& d, const std::vector<double>& h) { d.resize(h.size()); d.from_host(h.data(), 0, h.size()); };
    up(kxd, hkx); up(kxs, hsx); up(filt, hf); up(maskx, hm);
    yf = cfg.y_fourier;
    if (!yf) setup_y();
    else {   // Fourier y (src/sim_setup.cpp:63-82)
        dy = cfg.Ly / ny;
        std::vector<double> hky(ny), hsy(ny), hfy(ny), hmy(ny), hi(ny, 1 / dy);
        for (int j = 0; j < ny; ++j) {
            const int my = j <= ny / 2 ? j : j - ny;
            hsy[j] = 2 * M_PI * my / cfg.Ly;
            hky[j] = (2 * j == ny) ? 0 : hsy[j];
            hfy[j] = cfg.k_filter ? std::pow(std::cos(0.5 * hsy[j] * dy), 2) : 1.0;
            if (2 * j == ny) hfy[j] = 0;
            const double dd = std::min(j * dy, cfg.Ly - j * dy), w = cfg.sponge_y;
            double m = 1;
            if (w > 0 && dd < w) { const double r = 1 - dd / w; m = 1 - cfg.sponge_rate * r * r; }
            hmy[j] = m;
        }
        up(kyd, hky); up(kys, hsy); up(filty, hfy); up(masky_n, hmy); up(masky_h, hmy); up(iD, hi); up(iA, hi);
        yax.X = nullptr; yax.n = ny + 1; yax.uniform = 1; yax.x0 = 0; yax.ih = 1 / dy;
    }
    {   // zeta (src/sim_setup.cpp:148-164)
        const std::vector<double>& Z = cfg.zeta;
        std::vector<double> hh(N - 1);
        for (int m = 0; m < N - 1; ++m) hh[m] = Z[m + 1] - Z[m];
        bool uni = true;
        for (int m = 0; m < N - 1; ++m) uni = uni && hh[m] == hh[0];
        if (uni) for (double& v : hh) v = cfg.Lwin / (N - 1);
        std::vector<double> g(N + 3), lo(N), hi(N), ws(N);
        for (int m = -2; m <= N; ++m) g[m + 2] = hh[std::min(std::max(m, 0), N - 2)];
        for (int n = 0; n < N; ++n) {
            const double hm_ = g[n - 1 + 2], hp = g[n + 2];
            lo[n] = hp / (hm_ + hp); hi[n] = hm_ / (hm_ + hp);
            ws[n] = 0.5 * ((n > 0 ? hh[n - 1] : 0) + (n < N - 1 ? hh[n] : 0));
        }
        up(zs, Z); up(hcg, g); up(wlo, lo); up(whi, hi); up(wsl, ws);
        wlo_h = lo; hcg_h = g;
        zax.X = zs.data(); zax.n = N; zax.uniform = uni; zax.x0 = 0; zax.ih = 1 / hh[0];
    }
    const size_t pl = size_t(nk);
    for (DevBuf<cx>* p : {&Rx, &Ry, &Ez, &Bz, &Lx, &Ly}) p->resize(ns * pl);
    for (DevBuf<cx>* p : {&j0x, &j0y, &jkx, &jky, &jkz, &S1, &S2}) p->resize((ns + 1) * pl);
    T3.resize(3 * size_t(ns) * pl);
    jw.resize(3 * (ns + 3) * pl);
    F.resize(6 * (ns + 2) * pl);
    fft_plan_init(fft, cfg.nx, cfg.ny, yf);

    laser.on = cfg.laser_on; laser.a0 = cfg.a0; laser.kL = cfg.kL; laser.tau = cfg.tau; laser.cep = cfg.cep;
    laser.kx = std::sin(cfg.theta); laser.kz = std::cos(cfg.theta); laser.pol = cfg.pol;
    laser.eta_c = cfg.start_fwhm * cfg.tau;
    if (cfg.plasma_on) {   // src/sim_plasma.cpp:21-38
        for (DevBuf<cx>* p : {&js0x, &js0y, &js0z}) p->resize((ns + 1) * pl);
        jsr.resize(5 * pl);
        rho_lines.resize(size_t(N) * (cfg.nx + cfg.ny));
        bool mob = false;
        for (const auto& s_ : cfg.ions) mob = mob || cfg.loaded(s_);
        if (mob) rho_lines_i.resize(size_t(N) * (cfg.nx + cfg.ny));
        jsn.resize(3 * pl); jsin.resize(3 * pl); bk.resize(4 * pl); sp3.resize(3 * pl); tp3.resize(3 * pl);
        jpred.resize(3 * pl); jsi0.resize(3 * pl); dlt.resize(3 * pl);
    }
    counters.resize(4);
    cent_dev.resize(3);
    { const char* e = getenv("G3D_LINE_PIVOTS"); line_pivots_on = !(e && atoi(e) == 0); }
    { const char* e = getenv("G3D_BPUSH_WIN"); bpush_windows = !(e && atoi(e) == 0); }
    setup_ions();
    if (cfg.plasma_on) setup_tiles();
}

// Tiles of TX x TY cells (16 x 8 on large grids) with a margin of M nodes around the TSC reach of the tile; a
// window of WX = TX + 3 + 2M nodes must fit the period, and a window may reach only the neighbouring tiles.
void Sim::setup_tiles() {
    auto pick = [](int n, int pref) { for (int t = pref; t >= 8; t >>= 1) if (n % t == 0 && n >= 2 * t) return t; return 0; };
    int px = 16, py = 8;                      // G3D_TILE=<TX>x<TY> (powers of two from 8) for launch-shape sweeps
    if (const char* t = getenv("G3D_TILE")) {
        if (sscanf(t, "%dx%d", &px, &py) != 2 || px < 8 || py < 8)
            throw std::runtime_error("G3D_TILE: expected <TX>x<TY>, both at least 8");
    }
    // free-hand y: rows of py cells from the first node (the last row may be short), the windows end at the walls
    const int TX = pick(cfg.nx, px), TY = yf ? pick(cfg.ny, py) : (cfg.ny >= 2 * py ? py : 0);
    if (!TX || !TY) return;
    // G3D_TILE_MARGIN: nodes of drift a window allows (sweeps); 2 for the FP32-mixed electron kernel (its smaller
    // windows load faster than the few more fallbacks to F cost: 24.1 -> 23.7 us at 256 x 256), else 3
    bool ions = false;
    for (const auto& is : cfg.ions) ions = ions || cfg.loaded(is);
    int M = cfg.mixed && yf && !ions ? 2 : 3;
    if (const char* e = getenv("G3D_TILE_MARGIN")) M = std::max(0, atoi(e));
    M = std::min({M, TX - 1, TY - 1, (cfg.nx - TX - 3) / 2});
    if (yf) M = std::min(M, (cfg.ny - TY - 3) / 2);
    if (M < 0) return;
    tiles.TX = TX; tiles.TY = TY; tiles.M = M; tiles.WX = TX + 3 + 2 * M; tiles.WY = TY + 3 + 2 * M;
    tiles.ntx = cfg.nx / TX; tiles.nty = (cfg.ny + TY - 1) / TY; tiles.ntile = long(tiles.ntx) * tiles.nty;
    tiles.fh = !yf;
    tiles.fy = yax;
    tile_off.resize(tiles.ntile + 1);
    slab.resize(size_t(tiles.ntile) * 5 * tiles.WX * tiles.WY);
    tiles.off = tile_off.data(); tiles.slab = slab.data(); tiles.n_tiled = 0;
    const char* e = getenv("G3D_STREAM_FWIN");
    tiles.fwin = e ? atoi(e) : 1;
    // lanes per tile workgroup (about the particles of a tile; free-hand y with its ions: 256, 2 % faster than 512
    // on config 3)
    const char* b = getenv("G3D_STREAM_BLOCK");
    tiles.block = b ? atoi(b) : (yf ? 512 : 256);
    const char* w = getenv("G3D_STREAM_WPE");     // register cap of the FP32-mixed stream kernel (0 or 4)
    tiles.wpe = w ? atoi(w) : 4;
    if (tiles.block < 64 || tiles.block > kStreamBlockMax || tiles.block % 64)
        throw std::runtime_error("G3D_STREAM_BLOCK: a multiple of 64 up to 512");
}

Sim::~Sim() {
    (void)hipStreamSynchronize(stream());
    fft_plan_free(fft);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (auto& p : ev_pending) { (void)hipEventDestroy(p.second.first); (void)hipEventDestroy(p.second.second); }
}

// Symmetric tridiagonal eigenproblem by the implicit QL method with Wilkinson shifts (EISPACK tql2: Bowdler,
// Martin, Reinsch and Wilkinson, Numer. Math. 11 (1968) 293).  d: the diagonal, e[k]: the coupling of k and k+1;
// on return d holds the eigenvalues (unordered) and column m of z (z[k n + m]) the unit eigenvector of d[m].
// The floating-point operations run in the order of src/grid.cpp:56-100: the smallest eigenvalues of -Lap_y on a
// graded grid are fixed by any method only to about 1e-12 relative, so the co-moving field of those modes stays
// within the FP64 parity bound only with the reference's rounding.
static bool ql_implicit(std::vector<double>& d, std::vector<double> e, std::vector<double>& z) {
    const int n = int(d.size());
    z.assign(size_t(n) * n, 0.0);
    for (int k = 0; k < n; ++k) z[size_t(k) * n + k] = 1;
    if (n == 0) return true;
    e.resize(n);
    e[n - 1] = 0;
    constexpr double kEps = 2.2e-16;
    // end of the unreduced block that starts at l: the first m >= l whose coupling is negligible
    auto block_end = [&](int l) {
        int m = l;
        while (m < n - 1 && !(std::fabs(e[m]) <= kEps * (std::fabs(d[m]) + std::fabs(d[m + 1])))) ++m;
        return m;
    };
    // plane rotation of the columns i, i+1 of z
    auto rotate = [&](int i, double sn, double cs) {
        for (int k = 0; k < n; ++k) {
            double* row = &z[size_t(k) * n];
            const double t = row[i + 1];
            row[i + 1] = sn * row[i] + cs * t;
            row[i] = cs * row[i] - sn * t;
        }
    };
    for (int l = 0; l < n; ++l) {
        for (int sweep = 0;; ++sweep) {
            const int m = block_end(l);
            if (m == l) break;
            if (sweep == 60) return false;
            // shifted start of the chase: the eigenvalue of the leading 2 x 2 block nearer to d[l]
            double g = (d[l + 1] - d[l]) / (2 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double sn = 1, cs = 1, p = 0;
            bool split = false;
            for (int i = m - 1; i >= l; --i) {       // chase the bulge from m up to l
                const double f = sn * e[i], bb = cs * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0) { d[i + 1] -= p; e[m] = 0; split = true; break; }
                sn = f / r; cs = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * sn + 2 * cs * bb;
                p = sn * r;
                d[i + 1] = g + p;
                g = cs * r - bb;
                rotate(i, sn, cs);
            }
            if (split) continue;
            d[l] -= p; e[l] = g; e[m] = 0;
        }
    }
    return true;
}

// Free-hand y (src/sim_setup.cpp:86-143): nodes 0..ny (walls at 0 and ny), half-nodes between; 1/Delta_j,
// 1/A_j, the sponge on nodes and half-nodes, the LU factors of k_x^2 - Lap_y (bound field of the sponge) and the
// eigenvectors of -Lap_y (co-moving field).
void Sim::setup_y() {
    const int ny = cfg.ny, nx = cfg.nx;
    if (ny < 4) throw std::runtime_error("free-hand y grid: at least 4 cells between the walls");
    const std::vector<double>& Y = cfg.ynode;
    dy = 0;
    hDel.assign(ny, 0); hA.assign(ny, 0);
    for (int j = 0; j < ny; ++j) hDel[j] = Y[j + 1] - Y[j];
    hA[0] = 0.5 * hDel[0];
    for (int j = 1; j < ny; ++j) hA[j] = 0.5 * (hDel[j - 1] + hDel[j]);
    std::vector<double> hiD(ny), hiA(ny), hmn(ny), hmh(ny);
    auto prof = [&](double y) {
        const double dd = 0.5 * cfg.Ly - std::fabs(y), w = cfg.sponge_y;
        if (w <= 0 || dd >= w) return 1.0;
        const double r = 1 - dd / w;
        return 1 - cfg.sponge_rate * r * r;
    };
    for (int j = 0; j < ny; ++j) {
        hiD[j] = 1 / hDel[j]; hiA[j] = 1 / hA[j];
        hmn[j] = prof(Y[j]); hmh[j] = prof(0.5 * (Y[j] + Y[j + 1]));
    }
    auto up = [](DevBuf<double>& d, const std::vector<double>& h) { d.resize(h.size()); d.from_host(h.data(), 0, h.size()); };
    up(yn, Y); up(iD, hiD); up(iA, hiA); up(masky_n, hmn); up(masky_h, hmh);
    yax.X = yn.data(); yax.n = ny + 1;
    bool uni = true;
    for (int j = 0; j < ny; ++j) uni = uni && hDel[j] == hDel[0];
    yax.uniform = uni; yax.x0 = Y[0]; yax.ih = 1 / hDel[0];
    std::vector<double> kx(nx);
    kxd.to_host(kx.data(), 0, nx);
    std::vector<double> hcp(size_t(nk), 0.0), hid(size_t(nk), 0.0);
    for (int i = 0; i < nx; ++i) {
        double cprev = 0;
        for (int j = 1; j < ny; ++j) {
            const double lo = j > 1 ? -hiA[j] * hiD[j - 1] : 0, upv = j < ny - 1 ? -hiA[j] * hiD[j] : 0;
            const double den = kx[i] * kx[i] + hiA[j] * (hiD[j - 1] + hiD[j]) - lo * cprev;
            hid[size_t(j) * nx + i] = 1 / den;
            hcp[size_t(j) * nx + i] = cprev = upv / den;
        }
    }
    up(pcp, hcp); up(pid, hid);
    if (cfg.self_fields) {
        const int M = ny - 1;
        std::vector<double> d(M), e(M, 0.0), z;
        for (int j = 1; j < ny; ++j) d[j - 1] = hiA[j] * (hiD[j - 1] + hiD[j]);
        for (int j = 1; j < ny - 1; ++j) e[j - 1] = -hiD[j] / std::sqrt(hA[j] * hA[j + 1]);
        if (!ql_implicit(d, e, z)) throw std::runtime_error("eigenvectors of the y Laplacian: no convergence");
        std::vector<double> hf(size_t(ny) * ny, 0.0), hb(size_t(ny) * ny, 0.0), hl(ny, 0.0);
        for (int m = 0; m < M; ++m) {
            hl[m + 1] = d[m];
            for (int j = 1; j < ny; ++j) {
                const double u = z[size_t(j - 1) * M + m], sq = std::sqrt(hA[j]);
                hf[size_t(m + 1) * ny + j] = u * sq;
                hb[size_t(j) * ny + m + 1] = u / sq;
            }
        }
        up(evf, hf); up(evb, hb); up(evl, hl);
    }
    scp.resize(2 * size_t(nk));
    // the y wave numbers of the Fourier form are not used; filty = 1 keeps the shared kernels valid
    std::vector<double> zero(ny, 0.0), one(ny, 1.0);
    up(kyd, zero); up(kys, zero); up(filty, one);
}

void Sim::setup_ions() {   // src/sim_ions.cpp:68-87
    if (cfg.ions.empty()) return;
    std::vector<IonDev> iv;
    std::vector<AdkLevel> lev;
    const double omega0 = kC / cfg.Lu, E0 = kMe * kC * omega0 / kE;
    for (const auto& s_ : cfg.ions) {
        IonDev d{};
        d.mass = s_.mass; d.mobile = s_.mobile; d.ionizing = s_.ionization; d.Zmax = s_.Zmax; d.off = int(lev.size());
        d.Emin = 1e300;
        if (s_.ionization) {
            const std::vector<AdkLevel> L = adk_levels(s_, E0, omega0, 2 * cfg.Lwin);
            for (int k = s_.Z0; k < s_.Zmax; ++k) d.Emin = std::min(d.Emin, L[k].Emin);
            lev.insert(lev.end(), L.begin(), L.end());
        }
        iv.push_back(d);
    }
    ion_dev.resize(iv.size()); ion_dev.from_host(iv.data(), 0, iv.size());
    if (!lev.empty()) { adk_dev.resize(lev.size()); adk_dev.from_host(lev.data(), 0, lev.size()); }
    ion_stats.assign(cfg.ions.size(), std::vector<double>());
}

SpecGeom Sim::geom() const {
    SpecGeom g;
    g.nx = cfg.nx; g.ny = cfg.ny; g.ns = ns; g.N = N; g.a = a; g.nk = nk;
    g.kx = kxd.data(); g.ky = kyd.data(); g.kxs = kxs.data(); g.kys = kys.data(); g.filt = filt.data();
    g.filty = filty.data();
    return g;
}
MarchView Sim::mview(double D) {
    MarchView v;
    v.Rx = Rx.data(); v.Ry = Ry.data(); v.Ez = Ez.data(); v.Bz = Bz.data(); v.Lx = Lx.data(); v.Ly = Ly.data();
    v.j0x = j0x.data(); v.j0y = j0y.data(); v.jkx = jkx.data(); v.jky = jky.data(); v.jkz = jkz.data();
    v.kx = kxd.data(); v.ky = kyd.data(); v.hcg = hcg.data(); v.wlo = wlo.data(); v.whi = whi.data();
    v.D = D; v.a = a; v.N = N; v.nx = cfg.nx; v.nk = nk; v.ny = cfg.ny;
    return v;
}
LineView Sim::lview(double D) {
    LineView L;
    L.mv = mview(D); L.iD = iD.data(); L.iA = iA.data(); L.scp = scp.data(); L.ny = cfg.ny; L.piv = nullptr;
    // free-hand y march (D != 0): the pivots and elimination factors of every line and slice, formed once per D
    // and geometry (4 doubles per node of every line and slice, at most 24 GiB and only while the device has that
    // much free memory beside 1 GiB; otherwise, or with line_pivots off, march_line forms them as it goes)
    const size_t need = size_t(N) * 4 * size_t(cfg.ny) * size_t(cfg.nx);
    if (!yf && D != 0 && line_pivots_on && N >= 2 && need * sizeof(double) <= (size_t(24) << 30)) {
        if (!(piv_ok && piv_D == D)) {
            if (piv.size() < need) {
                piv.resize(0);
                size_t fr = 0, tot = 0;
                if (hipMemGetInfo(&fr, &tot) != hipSuccess || need * sizeof(double) + (size_t(1) << 30) > fr) {
                    (void)hipGetLastError();
                    return L;
                }
                piv.resize(need);
            }
            line_pivots(L, piv.data());
            piv_D = D; piv_ok = true;
        }
        L.piv = piv.data();
    }
    return L;
}
FieldView Sim::fview() {
    FieldView f;
    f.F = F.data(); f.a = a; f.ns = ns; f.nx = cfg.nx; f.ny = cfg.ny; f.yf = yf; f.self = cfg.self_fields; f.nk = nk;
    f.dx = dx; f.dy = dy; f.Lx = cfg.Lx; f.Ly = cfg.Ly; f.zax = zax; f.yax = yax;
    return f;
}

// ---------------------------------------------------------------- timing
void Sim::tic(int fam) {
    if (!(timing >> fam & 1u)) return;
    hipEvent_t e;
    if (!ev_pool.empty()) { e = ev_pool.back(); ev_pool.pop_back(); }
    else hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipEventRecord(e, stream()), "hipEventRecord");
    ev_open = e; ev_open_fam = fam;
}
void Sim::toc(int fam) {
    if (!(timing >> fam & 1u) || ev_open_fam != fam) return;
    hipEvent_t e;
    if (!ev_pool.empty()) { e = ev_pool.back(); ev_pool.pop_back(); }
    else hip_check(hipEventCreate(&e), "hipEventCreate");
    hip_check(hipEventRecord(e, stream()), "hipEventRecord");
    ev_pending.push_back({fam, {ev_open, e}});
    ev_open = nullptr; ev_open_fam = -1;
    if (ev_pending.size() > 8192) flush_timers();
}
// the events of one launch (hipExtLaunchKernel records them with the dispatch itself: no marker packets, so the
// timed run keeps the launch spacing of the untimed one)
LaunchTimer Sim::launch_timer(int fam) {
    if (!(timing >> fam & 1u)) return {};
    if (fam_seen[fam]++ % timing_every) return {};
    LaunchTimer lt;
    for (hipEvent_t* e : {&lt.start, &lt.stop}) {
        if (!ev_pool.empty()) { *e = ev_pool.back(); ev_pool.pop_back(); }
        else hip_check(hipEventCreate(e), "hipEventCreate");
    }
    return lt;
}
void Sim::launch_timed(int fam, const LaunchTimer& lt) {
    if (!lt.start) return;
    ev_pending.push_back({fam, {lt.start, lt.stop}});
    if (ev_pending.size() > 8192) flush_timers();
}
void Sim::flush_timers() {
    if (ev_pending.empty()) return;
    hip_check(hipStreamSynchronize(stream()), "synchronize");
    for (auto& p : ev_pending) {
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, p.second.first, p.second.second), "hipEventElapsedTime");
        fam_ms[p.first] += ms; fam_n[p.first] += 1;
        ev_pool.push_back(p.second.first); ev_pool.push_back(p.second.second);
    }
    ev_pending.clear();
}

// ---------------------------------------------------------------- particles of the beam list
std::vector<g3d_particle> Sim::host_particles() {
    std::vector<g3d_particle> v(static_cast<size_t>(np));
    if (np) {
        if (aos.size() < size_t(np)) aos.resize(size_t(np));
        parts_to_aos(beam.view(), aos.data(), np);
        aos.to_host(v.data(), 0, size_t(np));
    }
    return v;
}
void Sim::set_particles(const std::vector<g3d_particle>& v) {
    np = long(v.size());
    if (beam.cap < v.size() || beam.cap > 2 * v.size() + 1024) beam.resize(v.size() + v.size() / 4 + 64);
    if (!np) return;
    if (aos.size() < v.size()) aos.resize(v.size());
    aos.from_host(v.data(), 0, v.size());
    parts_from_aos(aos.data(), beam.view(), np);
}
// the streaming list in the reference's order: the fresh plasma by id (load order), then the rest as listed
std::vector<g3d_particle> Sim::host_strm() {
    std::vector<g3d_particle> v(static_cast<size_t>(nstrm));
    if (nstrm) {
        if (aos.size() < size_t(nstrm)) aos.resize(size_t(nstrm));
        parts_to_aos(strmA.view(), aos.data(), nstrm);
        aos.to_host(v.data(), 0, size_t(nstrm));
        std::stable_sort(v.begin(), v.begin() + std::min(n_tiled, nstrm),
                         [](const g3d_particle& a, const g3d_particle& b) { return a.id < b.id; });
    }
    return v;
}

// bunch loading with the reference's generator and draw order (src/sim_bunches.cpp:92-217, built-in Gaussian
// profiles); libstdc++'s mt19937_64 and normal_distribution give the reference's particles bit for bit
void Sim::load_bunch() {
    const double ylo = yf ? -1e300 : cfg.ynode[1], yhi = yf ? 1e300 : cfg.ynode[cfg.ny - 1];
    const double yoff = yf ? 0.5 * cfg.Ly : 0;   // Fourier y: box coordinate, physical y = p.y + Y0
    std::vector<g3d_particle> mine;
    // the bunches one after the other, each from its own seed (species 0, -1, -2, ...; src/sim_bunches.cpp:92-217)
    for (size_t kb = 0; kb < cfg.bunches.size(); ++kb) {
        const Config::Bunch& B = cfg.bunches[kb];
        if (B.nmacro <= 0) continue;
        if (kb == 0 && !cfg.prof.empty()) { load_bunch_profile(mine, ylo, yhi, yoff); continue; }
        std::mt19937_64 rng(B.seed);
        std::normal_distribution<double> G(0, 1);
        const double Ne = B.charge / kE;
        const double W = (Ne / B.nmacro) / (cfg.nc * std::pow(cfg.Lu, 3));
        const double p0 = std::sqrt(B.gamma * B.gamma - 1);
        Axis zh = zax; zh.X = cfg.zeta.data();
        for (long ib = 0; ib < B.nmacro; ++ib) {
            double x, y, z, px, py, pz;
            x = 0 + B.sigx * G(rng); y = 0 + B.sigy * G(rng);
            z = B.sigz * G(rng);
            px = B.sigx > 0 ? B.emitx / B.sigx * G(rng) : 0;
            py = B.sigy > 0 ? B.emity / B.sigy * G(rng) : 0;
            pz = p0 * (1 + B.espread * G(rng));
            const double g = std::sqrt(1 + px * px + py * py + pz * pz);
            const double vx = px / g, vy = py / g, vz = pz / g;
            const double zz = z + 0.0;
            g3d_particle p{};
            p.x = x - vx * zz / vz + 0.5 * cfg.Lx; p.y = y - vy * zz / vz + yoff;
            p.zeta = B.zeta_center + zz / vz;
            p.px = px; p.py = py; p.pz = pz; p.w = W; p.alive = 1;
            p.q = -1; p.qm = -1; p.sp = -int(kb); p.z = 0;
            int hc; double f;
            zh.locate(p.zeta, hc, f);
            if (p.zeta < 0 || hc > N - 2 || p.y < ylo || p.y > yhi) { ++lost_total; continue; }
            mine.push_back(p);
        }
    }
    set_particles(mine);
    X0 = -0.5 * cfg.Lx; Y0 = -yoff;
}

// the bunch with a current profile along zeta (distribution_z = function, src/sim_bunches.cpp:115-132, 160-177):
// zeta from the inverse of the profile's cumulative distribution (4096 trapezoids), the draws in the reference's
// order (uniform for zeta, then x, y, the transverse momenta if there is an emittance, the energy spread)
void Sim::load_bunch_profile(std::vector<g3d_particle>& mine, double ylo, double yhi, double yoff) {
    std::mt19937_64 rng(cfg.seed);
    std::normal_distribution<double> G(0, 1);
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    const double Ne = cfg.charge / kE;
    const double W = (Ne / cfg.nmacro) / (cfg.nc * std::pow(cfg.Lu, 3));
    const double zmin = cfg.prof_zmin, zmax = cfg.prof_zmax, zc = cfg.zeta_center;
    const int np = int(cfg.prof.size());
    const double hk = (zmax - zmin) / (np - 1);
    auto prof = [&](double z) { return hat_sum(cfg.prof.data(), np, zmin, hk, z); };
    const int K = 4096;
    std::vector<double> cz(K + 1, 0.0);
    double prev = std::max(0.0, prof(zmin));
    for (int k = 1; k <= K; ++k) {
        const double f = std::max(0.0, prof(zmin + (zmax - zmin) * k / K));
        cz[k] = cz[k - 1] + 0.5 * (prev + f); prev = f;
    }
    if (!(cz[K] > 0)) throw std::runtime_error("bunch: profile_z is zero on [zeta_min, zeta_max]");
    auto sample_zeta = [&](double u01) {
        const double t = u01 * cz[K];
        const int k = int(std::upper_bound(cz.begin(), cz.end(), t) - cz.begin()) - 1;
        const int kk = std::min(std::max(k, 0), K - 1);
        const double f = cz[kk + 1] > cz[kk] ? (t - cz[kk]) / (cz[kk + 1] - cz[kk]) : 0.5;
        return zmin + (zmax - zmin) * (kk + f) / K;
    };
    // for the gradient in the profile values (record_dprof_): d cz[k] / d prof[m] of the trapezoid sums, and per
    // loaded particle d zeta_f / d prof[m] at its draw (the trapezoid k it falls in held fixed)
    std::vector<double> dcz;
    if (record_dprof_) {
        dcz.assign(size_t(K + 1) * np, 0.0);
        auto hat = [&](int m, double z) { return std::max(0.0, 1 - std::fabs(z - (zmin + m * hk)) / hk); };
        for (int k = 1; k <= K; ++k) {
            const double z0 = zmin + (zmax - zmin) * (k - 1) / K, z1 = zmin + (zmax - zmin) * k / K;
            const bool p0 = prof(z0) > 0, p1 = prof(z1) > 0;
            for (int m = 0; m < np; ++m)
                dcz[size_t(k) * np + m] = dcz[size_t(k - 1) * np + m] + 0.5 * ((p0 ? hat(m, z0) : 0) + (p1 ? hat(m, z1) : 0));
        }
    }
    auto dzeta_dprof = [&](double u01, double* out) {
        const double t = u01 * cz[K];
        const int k = int(std::upper_bound(cz.begin(), cz.end(), t) - cz.begin()) - 1;
        const int kk = std::min(std::max(k, 0), K - 1);
        const double dl = cz[kk + 1] - cz[kk];
        for (int m = 0; m < np; ++m) {
            double df = 0;
            if (dl > 0) {
                const double dt = u01 * dcz[size_t(K) * np + m], d0 = dcz[size_t(kk) * np + m], d1 = dcz[size_t(kk + 1) * np + m];
                df = (dt - d0) / dl - (t - cz[kk]) * (d1 - d0) / (dl * dl);
            }
            out[m] = (zmax - zmin) * df / K;
        }
    };
    Axis zh = zax; zh.X = cfg.zeta.data();
    std::vector<double> dz(static_cast<size_t>(np));
    for (long ib = 0; ib < cfg.nmacro; ++ib) {
        const double u01 = U(rng) + 0.5;
        const double zf = sample_zeta(u01);
        if (record_dprof_) dzeta_dprof(u01, dz.data());
        const double xr = cfg.sigx * G(rng), yr = cfg.sigy * G(rng);
        const double x = 0 + xr, y = 0 + yr;
        const double px = cfg.emitx > 0 && cfg.sigx > 0 ? cfg.emitx / cfg.sigx * G(rng) : 0;
        const double py = cfg.emity > 0 && cfg.sigy > 0 ? cfg.emity / cfg.sigy * G(rng) : 0;
        const double ptot = std::sqrt(cfg.gamma * cfg.gamma - 1) * (1 + cfg.espread * G(rng));
        if (px * px + py * py >= ptot * ptot) throw std::runtime_error("bunch: transverse momentum above the total");
        const double pz = std::sqrt(ptot * ptot - px * px - py * py);
        const double z = (zf - zc) * pz / std::sqrt(1 + ptot * ptot);
        const double g = std::sqrt(1 + px * px + py * py + pz * pz);
        const double vx = px / g, vy = py / g, vz = pz / g;
        const double zz = z + 0.0;
        g3d_particle p{};
        p.x = x - vx * zz / vz + 0.5 * cfg.Lx; p.y = y - vy * zz / vz + yoff;
        p.zeta = zc + zz / vz;
        p.px = px; p.py = py; p.pz = pz; p.w = W; p.alive = 1;
        p.q = -1; p.qm = -1; p.sp = 0; p.z = 0;
        int hc; double f;
        zh.locate(p.zeta, hc, f);
        if (p.zeta < 0 || hc > N - 2 || p.y < ylo || p.y > yhi) { ++lost_total; continue; }
        mine.push_back(p);
        // zeta = zc + (zeta_f - zc) (pz / sqrt(1 + ptot^2)) / vz
        if (record_dprof_)
            for (int m = 0; m < np; ++m) dprof_.push_back(dz[size_t(m)] * (pz / std::sqrt(1 + ptot * ptot)) / vz);
    }
}

// ---------------------------------------------------------------- time step (src/sim_setup.cpp:178-217)
void Sim::init() {
    ensure_layout(false, false);
    load_bunch();
    s = 0; istep = 0;
    if (cfg.self_fields) {
        starting_ = true;
        if (cfg.bunches.size() > 1) comoving_fields_per_bunch();
        else { currents(); comoving_field(); }
        if (cfg.start_relax) relax_start_field();
        add_grid_lasers();
        march(0.0);
        starting_ = false;
        sponge_and_real_fields();
        j0_from_jk();
    }
    window_shift_post();
}

void Sim::step() {
    window_shift_start();
    push(s, s + cfg.dz);
    if (cfg.plasma_on) collect_restreaming();
    if (tape_) tape_beam(*tape_, false);
    s += cfg.dz; ++istep;
    if (cfg.self_fields && cfg.plasma_on) {
        currents();
        march_plasma(s - cfg.dz, s);
        sponge_and_real_fields();
        j0_from_jk();
    } else if (cfg.self_fields) solve_fields(0.5 * cfg.dz);
    window_shift_post();
    hip_check(hipStreamSynchronize(stream()), "synchronize");
}

void Sim::j0_from_jk() {
    if (cmR != cmJ) ensure_layout(cmR, cmR);          // j0 takes jk's values in its own layout
    const long n = long(ns + 1) * nk;
    copy_planes(j0x.data(), jkx.data(), n);
    copy_planes(j0y.data(), jky.data(), n);
}

// ---------------------------------------------------------------- beam push (src/sim_particles.cpp:6-24)
void Sim::push(double s0, double s1) {
    if (np <= 0) return;
    pushes_beam += np;
    PushArgs c;
    c.fv = fview(); c.laser = laser; c.X0 = X0; c.Y0 = Y0; c.zc = cfg.zeta_center; c.dz_push = cfg.dz_push;
    c.push_phase = cfg.push_phase; c.zhand = -1e300; c.zmax = cfg.zeta[N - 1]; c.s1 = s1;
    c.vz_min = cfg.plasma_on ? cfg.vz_min : 0; c.freeze = cfg.freeze_bunch;
    c.hmin = 1e300;
    for (int m = 0; m + 1 < N; ++m) c.hmin = std::min(c.hmin, cfg.zeta[m + 1] - cfg.zeta[m]);
    c.ylo = yf ? -1e300 : cfg.ynode[1]; c.yhi = yf ? 1e300 : cfg.ynode[cfg.ny - 1];
    // the binned push with field windows (off: one lane per particle gathering from F; the same results)
    tic(F_PUSH);
    if (bpush_windows && np >= 4096) bunch_push_windowed(c, beam.view(), np, s0, bpush_scratch);
    else bunch_push(c, beam.view(), np, s0);
    toc(F_PUSH);
}

// ---------------------------------------------------------------- currents (src/sim_particles.cpp:56-207)
void Sim::currents() {
    tic(F_CUR);
    cmJ = false;                 // jk is written whole below, in the reference's layout
    if (np == 0) {               // no beam particles: the beam current and its spectra are zero
        if (!jw_zero) { jw.zero(); jw_zero = true; }
        const long n = long(ns + 1) * nk;
        zero_planes(jkx.data(), n); zero_planes(jky.data(), n); zero_planes(jkz.data(), n);
        toc(F_CUR);
        return;
    }
    jw.zero();
    jw_zero = false;
    DepositArgs d;
    d.zax = zax; d.yax = yax; d.hcg = hcg.data(); d.ia = iA.data(); d.id = iD.data(); d.J = jw.data(); d.dx = dx;
    d.dy = dy; d.nh = ns + 3; d.a = a; d.N = N; d.nx = cfg.nx; d.ny = cfg.ny; d.yf = yf; d.nk = nk;
    bunch_deposit(d, beam.view(), np, dep_scratch);
    const SpecGeom g = geom();
    currents_zfilter(g, jw.data(), hcg.data(), cfg.zeta_filter, zax.uniform, S1.data(), jkz.data());
    fft_run(fft, S1.data(), ns + 1, false);
    fft_run(fft, jkz.data(), ns + 1, false);
    const bool keep0 = cfg.plasma_on && cfg.box_mean && !starting_;
    if (yf) currents_unpack(g, S1.data(), jkx.data(), jky.data(), jkz.data(), keep0);
    else fh_currents(S1.data(), jkx.data(), jky.data(), jkz.data(), ns + 1, lview(0.0), filt.data(), cfg.k_filter,
                     keep0, cfg.Ly);
    toc(F_CUR);
}

void Sim::march(double D) {
    ensure_layout(false, false);
    if (yf) march_all(mview(D), a, b);
    else march_lines(lview(D), a, b);
}

// ---------------------------------------------------------------- sponge and real fields (src/sim_fields.cpp:156-307)
void Sim::sponge_and_real_fields() {
    tic(F_SPONGE);
    // Fourier y reads R, L, E_z, B_z and jk in whichever layout they are (both the same); free-hand y in the
    // reference's
    if (!yf) ensure_layout(false, false);
    else if (cmR != cmJ) ensure_layout(cmR, cmR);
    SpecGeom g = geom();
    g.cm = cmR;
    const long ts = long(ns) * nk;
    const bool sp = cfg.sponge > 0 || cfg.sponge_y > 0;
    const double opb = 1 + beta_b;
    FhSpongeArgs fs{};
    const LineView L = lview(0.0);
    if (!yf) {
        fs.jz = jkz.data(); fs.kx = kxd.data(); fs.pc = pcp.data(); fs.pd = pid.data(); fs.hc = hcg.data();
        fs.wl = wlo.data(); fs.wu = whi.data(); fs.g2 = gm2_b; fs.opb = opb; fs.ns = ns; fs.N = N; fs.a = a;
    }
    if (sp) {
        if (yf)
            sponge_pack(g, Rx.data(), Ry.data(), jkz.data(), hcg.data(), wlo.data(), whi.data(), gm2_b, opb, S1.data(),
                        S2.data());
        else fh_sponge_psi(fs, L, Rx.data(), Ry.data(), S1.data(), S2.data());
        fft_run(fft, S1.data(), ns, true);
        sponge_mask(g, S1.data(), maskx.data(), masky_n.data(), masky_h.data(), ns);
        fft_run(fft, S1.data(), ns, false);
        if (!yf) fh_sponge_unpack(fs, L, S1.data(), S2.data(), Rx.data(), Ry.data());
    }
    // the FP32-mixed form: the three planes of the real fields transformed in FP32 (F is float in both forms)
    bool ions = false;
    for (const auto& is : cfg.ions) ions = ions || cfg.loaded(is);
    const bool f32 = cfg.mixed && yf && !ions && g.cm;
    cxf* T3f = f32 ? reinterpret_cast<cxf*>(T3.data()) : nullptr;
    fields_pack3(g, sp && yf ? S1.data() : nullptr, S2.data(), opb, Rx.data(), Ry.data(), Lx.data(), Ly.data(),
                 Ez.data(), Bz.data(), T3.data(), ts, ns, T3f);
    if (f32) {
        fft_run_f(fft, T3f, 3 * ns, true);
        fields_to_F_f(T3f, ts, F.data(), long(ns + 2) * nk, nk, ns);
    } else {
        fft_run(fft, T3.data(), 3 * ns, true);
        fields_to_F(T3.data(), ts, F.data(), long(ns + 2) * nk, nk, ns);
    }
    toc(F_SPONGE);
}

void Sim::solve_fields(double D) {
    currents();
    march(D);
    sponge_and_real_fields();
    j0_from_jk();
}

// ---------------------------------------------------------------- start field (src/sim_fields.cpp:44-151, 384-410)
void Sim::comoving_field() {
    ensure_layout(false, false);
    double sums[2] = {0, 0};
    for (auto& p : host_particles())
        if (p.alive == 1) { sums[0] += p.w * p.pz / std::sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz); sums[1] += p.w; }
    const double beta = comoving_c_ ? 1 : sums[1] > 0 ? sums[0] / sums[1] : 1;
    const double gm2 = 1 - beta * beta;
    if (!comoving_c_) { beta_b = beta; gm2_b = gm2; }
    const SpecGeom g = geom();
    if (yf) {
        comoving_solve(g, jkz.data(), hcg.data(), wlo.data(), whi.data(), beta, gm2, S1.data(), S2.data());
        comoving_R(g, S2.data(), beta, Rx.data(), Ry.data(), Ez.data(), Bz.data());
        return;
    }
    // free-hand y: the half-cell charges in the eigenvectors of -Lap_y (T3 as scratch), one solve along zeta per
    // (k_x, eigenvalue), back to the nodes
    const int nx = cfg.nx, ny = cfg.ny;
    fh_eigen_forward(jkz.data(), evf.data(), T3.data(), ns + 1, nx, ny, nk);
    comoving_solve_ev(g, T3.data(), evl.data(), hcg.data(), wlo.data(), whi.data(), beta, gm2, S1.data(), S2.data());
    fh_eigen_back(S2.data(), evb.data(), S1.data(), ns, nx, ny, nk);
    fh_comoving_R(S1.data(), kxd.data(), iD.data(), beta, Rx.data(), Ry.data(), Ez.data(), Bz.data(), ns, nx, ny, nk);
}

// several bunches: the co-moving field of each with its own beta, the forward fields R summed (the equation is
// linear), then the currents of all; the sponge's bound field keeps the mean beta of all bunch particles
// (src/sim_bunches.cpp:218-248)
void Sim::comoving_fields_per_bunch() {
    ensure_layout(false, false);
    const std::vector<g3d_particle> all = host_particles();
    const long n = long(ns) * nk;
    DevBuf<cx> sx(static_cast<size_t>(n)), sy(static_cast<size_t>(n));
    for (int k = 0; k < int(cfg.bunches.size()); ++k) {
        std::vector<g3d_particle> one;
        for (const auto& p : all) if (p.sp == -k) one.push_back(p);
        set_particles(one);
        currents();
        comoving_field();
        axpy_planes(sx.data(), Rx.data(), 1, n);
        axpy_planes(sy.data(), Ry.data(), 1, n);
    }
    set_particles(all);
    currents();
    copy_planes(Rx.data(), sx.data(), n);
    copy_planes(Ry.data(), sy.data(), n);
    double m[2] = {0, 0};
    for (const auto& p : all)
        if (p.alive == 1) { m[0] += p.w * p.pz / std::sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz); m[1] += p.w; }
    beta_b = m[1] > 0 ? m[0] / m[1] : 1;
    gm2_b = 1 - beta_b * beta_b;
}

void Sim::relax_start_field() {
    ensure_layout(false, false);
    const long n = long(ns) * nk, nj = long(ns + 1) * nk;
    DevBuf<cx> dx_(n), dy_(n);
    copy_planes(dx_.data(), Rx.data(), n);
    copy_planes(dy_.data(), Ry.data(), n);
    comoving_c_ = true; comoving_field(); comoving_c_ = false;
    axpy_planes(dx_.data(), Rx.data(), -1, n);
    axpy_planes(dy_.data(), Ry.data(), -1, n);
    const double D = 100 * cfg.dz;
    for (int it = 0; it < 3; ++it) {
        zero_planes(Ez.data(), n); zero_planes(Bz.data(), n);
        zero_planes(j0x.data(), nj); zero_planes(j0y.data(), nj);
        march(D);
    }
    axpy_planes(Rx.data(), dx_.data(), 1, n);
    axpy_planes(Ry.data(), dy_.data(), 1, n);
    hip_check(hipStreamSynchronize(stream()), "synchronize");
}

GridLaserArgs Sim::grid_laser_args(const g3d_glaser& Lz) {
    GridLaserArgs la;
    la.amp = Lz.a0 * Lz.kL; la.waist = Lz.waist; la.zR = 0.5 * Lz.kL * Lz.waist * Lz.waist; la.zfoc = Lz.z_focus;
    la.zc = Lz.zeta_c; la.tau = Lz.tau; la.kL = Lz.kL; la.cep = Lz.cep;
    la.rc = std::cos(Lz.pol_angle); la.rs = std::sin(Lz.pol_angle); la.dx = dx; la.dy = dy; la.X0 = X0; la.Y0 = Y0;
    la.circular = Lz.circular; la.helicity = Lz.helicity; la.nx = cfg.nx; la.ny = cfg.ny; la.nk = nk;
    la.zeta = zs.data(); la.yn = yf ? nullptr : yn.data();
    la.env = nullptr;
    return la;
}

void Sim::add_grid_lasers() {
    ensure_layout(false, false);
    if (cfg.glasers.empty()) return;
    const SpecGeom g = geom();
    for (const auto& Lz : cfg.glasers) {
        GridLaserArgs la = grid_laser_args(Lz);
        DevBuf<double> envd;
        if (Lz.n_env > 0) {          // envelope_function: a(zeta) of each slice (src/sim_fields.cpp:335)
            std::vector<double> e(static_cast<size_t>(N));
            for (int n = 0; n < N; ++n) e[size_t(n)] = hat_sum(Lz.env, Lz.n_env, Lz.env_z0, Lz.env_dz, cfg.zeta[n]);
            envd.resize(size_t(N));
            envd.from_host(e.data(), 0, size_t(N));
            la.env = envd.data();
        }
        grid_laser_fill(la, T3.data(), a, ns);
        fft_run(fft, T3.data(), ns, false);
        grid_laser_add(g, T3.data(), Rx.data(), Ry.data(), ns, !yf);
    }
}

// ---------------------------------------------------------------- plasma (src/sim_plasma.cpp)
void Sim::ensure_strm(size_t need) {
    if (strmA.cap >= need) return;
    const size_t cap = need + need / 8 + 64;
    strmA.resize(cap); strmB.resize(cap, false);
}

void Sim::load_plasma(double s0, double s1) {
    if (load_plasma_device(s0, s1)) return;
    // the list in the reference's order (src/sim_plasma.cpp:43-90): lattice j, i, b2, a2 with the ions of every
    // loaded species after each electron, ids in that order; then regrouped tile by tile (stable) for the deposit
    std::vector<g3d_particle> v;
    std::vector<int> tile_of;
    long cap = 0;
    const unsigned long long id0 = (unsigned long long)(istep + 1) << 32;
    id0_ = id0;
    const double zmid = 0.5 * (s0 + s1), Delta = s1 - s0;
    const int nx = cfg.nx, ny = cfg.ny, px = cfg.ppc_x, py = cfg.ppc_y;
    const double zf = cfg.zeta[N - 1];
    const bool tiled = tiles.ntile > 0;
    v.reserve(size_t(nx) * ny * px * py * (1 + cfg.ions.size()) + joining_.size());
    for (int j = 0; j < ny; ++j) {
        double y0 = j * dy, hy = dy;
        if (!yf) {
            if (j == 0 || j == ny - 1) continue;      // no plasma in the cells at the walls
            y0 = cfg.ynode[j]; hy = cfg.ynode[j + 1] - cfg.ynode[j];
        }
        for (int i = 0; i < nx; ++i) {
            const int t = tiled ? (j / tiles.TY) * tiles.ntx + i / tiles.TX : 0;
            for (int b2 = 0; b2 < py; ++b2)
                for (int a2 = 0; a2 < px; ++a2) {
                    g3d_particle p{};
                    p.x = (i + (a2 + 0.5) / px) * dx;
                    p.y = y0 + (b2 + 0.5) / py * hy;
                    const double n = cfg.plasma_density(p.x + X0, p.y + Y0, zmid);
                    if (n <= 0) continue;
                    p.w = n * dx * hy * Delta / (px * py);
                    p.q = -1; p.qm = -1; p.sp = SP_PLASMA; p.alive = 1;
                    p.zeta = zf; p.z = s0;
                    if (cfg.free_electrons) { p.id = id0 + v.size(); v.push_back(p); tile_of.push_back(t); }
                    for (size_t k = 0; k < cfg.ions.size(); ++k) {
                        const g3d_ion& is = cfg.ions[k];
                        if (!cfg.loaded(is)) continue;
                        g3d_particle q = p;
                        q.q = is.Z0; q.qm = is.Z0 / is.mass; q.w = is.atoms * p.w; q.sp = SP_ION + int(k);
                        q.id = id0 + v.size();
                        v.push_back(q); tile_of.push_back(t);
                    }
                }
        }
    }
    fresh_last = long(v.size());
    if (tiled) {
        std::vector<long> off(tiles.ntile + 1, 0);
        for (int t : tile_of) ++off[t + 1];
        for (long t = 0; t < tiles.ntile; ++t) off[t + 1] += off[t];
        std::vector<long> pos(off.begin(), off.end() - 1);
        std::vector<g3d_particle> w(v.size());
        for (size_t k = 0; k < v.size(); ++k) w[pos[tile_of[k]]++] = v[k];
        v.swap(w);
        tile_off.from_host(off.data(), 0, off.size());
        n_tiled = fresh_last;
        tiles.n_tiled = n_tiled;
    } else n_tiled = 0;
    for (g3d_particle p : joining_) { p.alive = 6; v.push_back(p); }
    for (const auto& p : v)
        if (is_ion(p.sp) && cfg.ions[p.sp - SP_ION].ionization) cap += cfg.ions[p.sp - SP_ION].Zmax - std::lround(p.q);
    nstrm = long(v.size());
    born_cap = cap;
    ensure_strm(v.size() + size_t(cap));
    if (cap > 0 && born.cap < size_t(cap)) born.resize(size_t(cap));
    if (!nstrm) return;
    if (aos.size() < v.size()) aos.resize(v.size());
    aos.from_host(v.data(), 0, v.size());
    parts_from_aos(aos.data(), strmA.view(), nstrm);
    parts_copy(strmA.view(), strmB.view(&strmA), nstrm);
}

// The common case on the device: Fourier y, tiled, electrons only and every lattice point loaded (the density is
// positive on the whole box: no radial cut inside it, no negative channel, a positive longitudinal factor).  The
// list is then the host loader's, regrouped by tile, with the same values and ids.
bool Sim::load_plasma_device(double s0, double s1) {
    const char* e = getenv("G3D_HOST_LATTICE");                 // 1: always the host loader (for comparisons)
    if ((e && atoi(e) > 0) || !yf || tiles.ntile <= 0 || !cfg.free_electrons || !cfg.plasma_on || cfg.n0 <= 0) return false;
    for (const auto& is : cfg.ions) if (cfg.loaded(is)) return false;
    const double zmid = 0.5 * (s0 + s1);
    const double f = cfg.plasma_zfactor(zmid);
    if (!(f > 0) || (cfg.parabolic && !(cfg.channel_depth >= 0))) return false;
    const double xm = std::max(std::fabs(X0), std::fabs(X0 + cfg.Lx)), ym = std::max(std::fabs(Y0), std::fabs(Y0 + cfg.Ly));
    if (!(xm * xm + ym * ym < 0.99 * cfg.r_max * cfg.r_max)) return false;      // no point beyond the radial cut
    const int nx = cfg.nx, ny = cfg.ny, px = cfg.ppc_x, py = cfg.ppc_y;
    if (px < 1 || py < 1) return false;
    LatticeArgs L;
    L.nx = nx; L.ny = ny; L.px = px; L.py = py; L.TX = tiles.TX; L.TY = tiles.TY; L.ntx = tiles.ntx;
    L.parabolic = cfg.parabolic;
    L.per_tile = long(tiles.TX) * tiles.TY * px * py;
    L.dx = dx; L.dy = dy; L.X0 = X0; L.Y0 = Y0; L.zf = cfg.zeta[N - 1]; L.s0 = s0; L.Delta = s1 - s0;
    L.n0 = cfg.n0; L.f = f; L.channel_depth = cfg.channel_depth; L.r_channel = cfg.r_channel;
    const unsigned long long id0 = (unsigned long long)(istep + 1) << 32;
    id0_ = id0;
    L.id0 = id0;
    const long fresh = L.per_tile * tiles.ntile;
    const long total = fresh + long(joining_.size());
    fresh_last = fresh;
    std::vector<long> off(tiles.ntile + 1);
    for (long t = 0; t <= tiles.ntile; ++t) off[t] = t * L.per_tile;
    tile_off.from_host(off.data(), 0, off.size());
    n_tiled = fresh;
    tiles.n_tiled = n_tiled;
    nstrm = total;
    born_cap = 0;
    ensure_strm(size_t(total));
    plasma_lattice(L, strmA.view(), strmB.view(&strmA), fresh);
    if (!joining_.empty()) {      // trapped particles too slow to join the beam stream again, waiting (alive 6)
        std::vector<g3d_particle> v(joining_);
        for (auto& p : v) p.alive = 6;
        if (aos.size() < v.size()) aos.resize(v.size());
        aos.from_host(v.data(), 0, v.size());
        auto at = [&](Parts p) {
            p.x += fresh; p.y += fresh; p.zeta += fresh; p.px += fresh; p.py += fresh; p.pz += fresh; p.w += fresh;
            p.z += fresh; p.q += fresh; p.qm += fresh; p.rem += fresh; p.alive += fresh; p.sp += fresh; p.id += fresh;
            return p;
        };
        parts_from_aos(aos.data(), at(strmA.view()), long(v.size()));
        parts_copy(at(strmA.view()), at(strmB.view(&strmA)), long(v.size()));
    }
    return true;
}

void Sim::ensure_layout(bool r_cm, bool j_cm) {
    auto flip = [&](DevBuf<cx>& A, DevBuf<cx>& scratch, bool to_cm) {
        if (!A.size()) return;
        if (scratch.size() != A.size()) scratch.resize(A.size());
        transpose_planes(A.data(), scratch.data(), long(A.size()) / nk, to_cm ? cfg.ny : cfg.nx,
                         to_cm ? cfg.nx : cfg.ny);
        A.swap(scratch);
    };
    if (cmR != r_cm) {
        for (DevBuf<cx>* A : {&Rx, &Ry, &Ez, &Bz, &Lx, &Ly}) flip(*A, trF, r_cm);
        for (DevBuf<cx>* A : {&js0x, &js0y, &js0z, &j0x, &j0y}) flip(*A, trC, r_cm);
        cmR = r_cm;
    }
    if (cmJ != j_cm) {
        for (DevBuf<cx>* A : {&jkx, &jky, &jkz}) flip(*A, trC, j_cm);
        cmJ = j_cm;
    }
}

void Sim::stream_spectra() {
    ensure_layout(false, false);
    // stand-alone form (jsr is consumed): the same row and column kernels as the march, spectra only
    pass_rows_fwd(jsr.data(), TileGeom{}, fft, sp3.data(), nullptr, nullptr, 0, 0);
    ModeArgs m{};
    m.mv = mview(0.0);
    m.kxs = kxs.data(); m.kys = kys.data(); m.filt = filt.data(); m.filty = filty.data();
    m.a2 = cfg.smooth_length * cfg.smooth_length; m.keep0 = cfg.box_mean; m.ny = cfg.ny; m.kf = cfg.k_filter;
    if (yf) pass_cols(m, fft, sp3.data(), tp3.data(), jsn.data(), 2);
    else pass_lines(m, lview(0.0), sp3.data(), tp3.data(), jsn.data(), 2);
}

void Sim::march_slice(int n, double D) {
    ensure_layout(false, false);
    if (yf) march_one(mview(D), n, nullptr, nullptr, nullptr, nullptr, 0);
    else march_lines(lview(D), n, n + 1);
}

void Sim::real_fields_slice(int n) {
    ensure_layout(false, false);
    const SpecGeom g = geom();
    const long l = n - a;
    fields_pack3(g, nullptr, nullptr, 0, Rx.data() + l * nk, Ry.data() + l * nk, Lx.data() + l * nk, Ly.data() + l * nk,
                 Ez.data() + l * nk, Bz.data() + l * nk, tp3.data(), nk, 1);
    fft_run(fft, tp3.data(), 3, true);
    fields_to_F(tp3.data(), nk, F.data() + l * nk, long(ns + 2) * nk, nk, 1);
}

// the streaming context of the step s0 -> s1 (src/sim_plasma.cpp:280-291)
StreamArgs Sim::stream_args(double s0, double s1) {
    const double Delta = s1 - s0;
    StreamArgs sc;
    sc.fv = fview(); sc.laser = laser; sc.X0 = X0; sc.Y0 = Y0; sc.zc = cfg.zeta_center;
    sc.ylo = yf ? -1e300 : cfg.ynode[1]; sc.yhi = yf ? 1e300 : cfg.ynode[cfg.ny - 1];
    sc.zs = zs.data(); sc.J = jsr.data(); sc.ia = iA.data(); sc.id = iD.data();
    sc.Delta = Delta; sc.s1 = s1; sc.sTrap = s1 + (cfg.trap_advance - 1) * Delta; sc.vz_beam = cfg.vz_beam;
    sc.ions = ion_dev.data(); sc.adk = adk_dev.data(); sc.born = born.cap ? born.view() : Parts{};
    sc.bcnt = counters.data(); sc.bcap = 0; sc.key = 0; sc.n = 0;
    sc.npush = nullptr; sc.npush_cap = 0; sc.Jf = nullptr; sc.skip_state = 0;
    return sc;
}

// one pass of the (untiled) streaming list through half-cell n into a zeroed jsr; the list keeps the result
void Sim::stream_pass(int n, int pass, double s0, double s1) {
    ensure_layout(false, false);
    StreamArgs sc = stream_args(s0, s1);
    sc.n = n;
    jsr.zero();
    if (born_cap > 0) {
        counters.zero();
        sc.key = key_slice(key_step(cfg.seed, istep), n, pass);
        sc.bcap = int(std::min<long>(long(born.cap), long(strmA.cap) - nstrm));
    }
    stream_halfcell(sc, strmA.view(), strmB.view(&strmA), nstrm, TileGeom{});
    parts_copy(strmB.view(&strmA), strmA.view(), nstrm);
}

// a streaming list from the host (untiled)
void Sim::set_strm(const std::vector<g3d_particle>& v) {
    nstrm = long(v.size());
    n_tiled = 0;
    ensure_strm(v.size() + 1);
    if (!nstrm) return;
    if (aos.size() < v.size()) aos.resize(v.size());
    aos.from_host(v.data(), 0, v.size());
    parts_from_aos(aos.data(), strmA.view(), nstrm);
    parts_copy(strmA.view(), strmB.view(&strmA), nstrm);
}

void Sim::march_plasma(double s0, double s1) {
    const double Delta = s1 - s0, D = 0.5 * Delta;
    const size_t pl = size_t(nk);
    load_plasma(s0, s1);
    joining_.clear();
    const bool extrap = cfg.predictor_extrapolate;
    std::vector<double> cen(N - 1);
    for (int m = 0; m < N - 1; ++m) cen[m] = 0.5 * (cfg.zeta[m] + cfg.zeta[m + 1]);
    const SpecGeom g = geom();
    jpred.zero();
    jsr.zero();
    // the FP32-mixed pass: with Fourier y and electrons only the mixed stream kernel (mx); with ion species or the
    // free-hand y grid (mxi) the FP64 stream kernel (the ion push and ionization as in the FP64 form) depositing into
    // float current planes, the FP32 forward row transforms, and the mixed column pass with its FP32 inverse rows
    // (Fourier y) or the FP64 line pass and FP64 inverse rows (free-hand y: the elimination stays FP64).  With Fourier y the modes of either form run column-major (one
    // fused column launch per pass) unless G3D_PASS_CM=0
    bool ions = false;
    for (const auto& is : cfg.ions) ions = ions || cfg.loaded(is);
    const bool mx = cfg.mixed && yf && !ions;
    const bool mxi = cfg.mixed && !mx;
    const char* ecm = getenv("G3D_PASS_CM");
    const bool cm = yf && !(ecm && atoi(ecm) == 0) && cfg.ny <= 1024 && cfg.nx % 16 == 0 && cfg.ny % 16 == 0;
    ensure_layout(cm, cm);
    ModeArgs m{};
    m.mv = mview(D);
    m.mv.cm = cm;
    m.jsn = jsn.data(); m.sin_ = jsin.data(); m.si0 = jsi0.data(); m.pin = jpred.data(); m.dl = dlt.data();
    m.js0x = js0x.data(); m.js0y = js0y.data(); m.js0z = js0z.data();
    m.jkx = jkx.data(); m.jky = jky.data(); m.jkz = jkz.data();
    m.T = tp3.data(); m.kxs = kxs.data(); m.kys = kys.data(); m.filt = filt.data(); m.filty = filty.data();
    m.a2 = cfg.smooth_length * cfg.smooth_length; m.extrap = extrap; m.keep0 = cfg.box_mean; m.ny = cfg.ny;
    m.kf = cfg.k_filter;
    const LineView L = lview(D);
    // the columns (Fourier y) or lines (free-hand y) of a pass
    static const bool fused = [] { const char* e = getenv("G3D_PASS_FUSED"); return e && atoi(e) > 0; }();
    auto pass_work = [&](cx* X, int kind) {
        if (!yf) pass_lines(m, L, X, tp3.data(), nullptr, kind);
        else if (cm) pass_colmodes(m, fft, X, tp3.data(), kind);
        else if (fused) pass_cols(m, fft, X, tp3.data(), nullptr, kind);
        else pass_modes(m, fft, X, tp3.data(), nullptr, kind);
    };
    const long fpl = long(ns + 2) * nk;
    auto real_fields = [&](int n) {
        tic(F_IFFT);
        pass_rows_inv(tp3.data(), F.data() + long(n - a) * nk, fpl, fft, cm);
        toc(F_IFFT);
    };
    StreamArgs sc = stream_args(s0, s1);
    const bool ionz = born_cap > 0;
    // tiles only when this step loaded tiled plasma (otherwise the slabs would hold an earlier step's current)
    TileGeom tg = n_tiled > 0 ? tiles : TileGeom{};
    tg.idx = nullptr;
    tg.chunk = 0; tg.max_extra = 0; tg.xtile = nullptr; tg.xstart = nullptr; tg.xcount = nullptr;
    // FP32-mixed: the tile workgroups XCD-contiguous (G3D_STREAM_XCD=0: in tile order; the FP64 kernel ran 4 %
    // slower so)
    { const char* ex = getenv("G3D_STREAM_XCD"); tg.xcd = mx && (ex ? atoi(ex) : 1) && tg.ntile % 8 == 0 && tg.ntile < (1L << 30); }
    // FP32-mixed: the tile windows go to the current planes with float atomics instead of slabs (3 % faster at
    // 256 x 256: the row kernel then reads one plane set); G3D_STREAM_AFLUSH=0 keeps the slabs
    { const char* e = getenv("G3D_STREAM_AFLUSH"); tg.aflush = (mx && !(e && atoi(e) == 0)) || mxi; }
    // FP32-mixed: tiles denser than G3D_STREAM_CHUNK (default 1) times one workgroup's lanes are streamed in
    // chunks of that many particles (0: off)
    const char* ech = getenv("G3D_STREAM_CHUNK");
    const int chunk_mult = ech ? atoi(ech) : 1;
    if (mx && tg.ntile && chunk_mult > 0) {
        tg.chunk = chunk_mult * tg.block;
        tg.max_extra = int((nstrm + tg.chunk - 1) / tg.chunk);
        // XCD-contiguous tiles: the further-chunk workgroups padded to a multiple of 8 so that tile workgroup
        // b - max_extra runs on XCD (b - max_extra) mod 8
        if (tg.xcd) tg.max_extra = (tg.max_extra + 7) & ~7;
        if (chunk_tile.size() < size_t(tg.max_extra)) { chunk_tile.resize(size_t(tg.max_extra)); chunk_start.resize(size_t(tg.max_extra)); }
        if (chunk_cnt.size() < 1) chunk_cnt.resize(1);
        tg.xtile = chunk_tile.data(); tg.xstart = chunk_start.data(); tg.xcount = chunk_cnt.data();
        tile_chunks(tg, chunk_tile.data(), chunk_start.data(), chunk_cnt.data());
    }
    if (mxi) {
        if (jsr_f.size() < 5 * pl) jsr_f.resize(5 * pl);
        sc.Jf = jsr_f.data();
    }
    if (mx) {
        if (jsr_f.size() < 5 * pl) jsr_f.resize(5 * pl);
        if (tiles.ntile && slab_f.size() < size_t(tiles.ntile) * 5 * tiles.WX * tiles.WY)
            slab_f.resize(size_t(tiles.ntile) * 5 * tiles.WX * tiles.WY);
    }
    // the stream launches count their pushes per workgroup: at most every tile, every further chunk and the
    // untiled rest of the list in workgroups of at least 64 lanes
    const long npc = tg.ntile + tg.max_extra + (long(strmA.cap) + 63) / 64 + 1;
    if (npush.size() < size_t(npc)) npush.resize(size_t(npc));
    else npush.zero();
    sc.npush = npush.data(); sc.npush_cap = long(npush.size());
    cxf* Xf = reinterpret_cast<cxf*>(sp3.data());
    cxf* Tf = reinterpret_cast<cxf*>(tp3.data());
    const int rpw = rows_per_wg(cfg.ny);
    // front: half-cell N-2 predicted with its level-s0 value; slice N-1
    m.pk = nullptr; m.n = N - 1; m.first = 1; m.last = 1; m.top = 1; m.wex = 0; m.qcorr = 0;
    m.bRx = m.bRy = m.bEz = m.bBz = nullptr;
    tic(F_MODES);
    if (mx || (mxi && yf))
        pass_mixed(m, fft, jsr_f.data(), slab_f.data(), tg, Xf, Tf, F.data() + long(N - 1 - a) * nk, fpl, nullptr,
                   nullptr, 0, 0, 1, rpw);
    else pass_work(sp3.data(), 1);
    toc(F_MODES);
    if (!mx && !(mxi && yf)) real_fields(N - 1);
    // every rebin half-cells the live particles are binned again by their current tile (0: never)
    static const int rebin = [] { const char* e = getenv("G3D_REBIN"); return e ? atoi(e) : 8; }();
    const uint64_t kst = key_step(cfg.seed, istep);
    const int passes = cfg.iterations;
    const int nx = cfg.nx, ny = cfg.ny;
    const int i0 = ((int(std::lround(-X0 / dx)) % nx) + nx) % nx;
    const int j0 = yf ? ((int(std::lround(-Y0 / dy)) % ny) + ny) % ny : ny / 2;
    bool have_dl = false;
    long born_base = nstrm;
    long nlist = nstrm;          // particles in the list (freed electrons are appended on the device)
    // the record for the reverse of the step (Sim::gradient): the streaming particles before every half-cell, the
    // real fields of every slice after each pass but the last, and after the march
    auto tape_copy = [&](void* dst, const void* src, size_t bytes) {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()), "hipMemcpy D2D");
    };
    // the six E, B planes of one slice of F (planes fpl apart) in one strided copy
    auto tape_slice = [&](float* dst, const float* src) {
        hip_check(hipMemcpy2DAsync(dst, size_t(fpl) * sizeof(float), src, size_t(fpl) * sizeof(float),
                                   size_t(nk) * sizeof(float), 6, hipMemcpyDeviceToDevice, stream()),
                  "hipMemcpy2D D2D");
    };
    if (tape_) {
        StepTape& t = *tape_;
        const size_t np7 = size_t(N - 1) * 7 * size_t(nlist);
        t.nstrm = nlist;
        if (t.ps.size() != np7) {
            t.ps.resize(np7); t.pa.resize(size_t(N - 1) * size_t(nlist)); t.pq.resize(3 * size_t(nlist));
            t.ids.resize(size_t(nlist));
        }
        tape_copy(t.ids.data(), strmA.id.data(), size_t(nlist) * sizeof(unsigned long long));
        t.added.clear();
        if (t.Fmid.size() != size_t(std::max(passes - 1, 0)) * F.size()) t.Fmid.resize(size_t(std::max(passes - 1, 0)) * F.size());
        if (t.Fmarch.size() != F.size()) t.Fmarch.resize(F.size());
        const size_t b = size_t(nlist) * sizeof(double);
        tape_copy(t.pq.data(), strmA.q.data(), b);
        tape_copy(t.pq.data() + nlist, strmA.qm.data(), b);
        tape_copy(t.pq.data() + 2 * nlist, strmA.w.data(), b);
    }
    for (int n = N - 2; n >= a; --n) {
        const double wex = (n >= 1 && n + 1 <= N - 2 && have_dl) ? (cen[n] - cen[n - 1]) / (cen[n + 1] - cen[n]) : 0.0;
        if (tape_) {
            StepTape& t = *tape_;
            const size_t b = size_t(nlist) * sizeof(double);
            double* dst = t.ps.data() + size_t(n) * 7 * size_t(nlist);
            const DevBuf<double>* src[7] = {&strmA.x, &strmA.y, &strmA.zeta, &strmA.z, &strmA.px, &strmA.py, &strmA.pz};
            CopySpan sp[8];                  // the seven state arrays and alive in one launch
            for (int c = 0; c < 7; ++c) sp[c] = CopySpan{dst + size_t(c) * nlist, src[c]->data(), long(b / 4)};
            sp[7] = CopySpan{t.pa.data() + size_t(n) * nlist, strmA.alive.data(), nlist};
            copy_spans(sp, 8);
        }
        if (tg.ntile && rebin > 0 && n < N - 2 && (N - 2 - n) % rebin == 0) {
            if (bin_idx.size() < strmA.cap) bin_idx.resize(strmA.cap);
            if (bin_cnt.size() < size_t(tg.ntile)) { bin_cnt.resize(size_t(tg.ntile)); bin_cur.resize(size_t(tg.ntile)); }
            bin_tiles(strmA.view(), strmB.view(&strmA), nlist, tg, dx, dy, nx, ny, bin_cnt.data(), bin_cur.data(),
                      tile_off.data(), bin_idx.data());
            tg.idx = bin_idx.data();
            tg.n_tiled = nlist;
            if (tg.chunk) tile_chunks(tg, chunk_tile.data(), chunk_start.data(), chunk_cnt.data());
        }
        for (int pass = 0; pass < passes; ++pass) {
            sc.n = n;
            if (ionz) {
                hip_check(hipMemsetAsync(counters.data(), 0, sizeof(int), stream()), "hipMemset");
                sc.key = key_slice(kst, n, pass);
                sc.bcap = int(std::min<long>(long(born.cap), long(strmA.cap) - nlist));
            }
            const bool last = pass == passes - 1;
            sc.skip_state = !last;
            m.n = n; m.first = pass == 0; m.last = last; m.top = 0; m.wex = wex;
            m.qcorr = D * wlo_h[n + 1];
            if (passes > 1) { m.bRx = bk.data(); m.bRy = bk.data() + pl; m.bEz = bk.data() + 2 * pl; m.bBz = bk.data() + 3 * pl; }
            if (mx) {
                const LaunchTimer lt = launch_timer(F_STREAM);
                stream_mixed(sc, jsr_f.data(), slab_f.data(), strmA.view(), strmB.view(&strmA), nlist, tg, lt, last);
                launch_timed(F_STREAM, lt);
                tic(F_MODES);
                pass_mixed(m, fft, jsr_f.data(), slab_f.data(), tg, Xf, Tf, F.data() + long(n - a) * nk, fpl,
                           last ? rho_lines.data() + long(n) * (nx + ny) : nullptr, nullptr, i0, j0, 0, rpw);
                toc(F_MODES);
                if (tape_ && !last)
                    tape_slice(tape_->Fmid.data() + size_t(pass) * F.size() + size_t(n - a) * nk,
                               F.data() + size_t(n - a) * nk);
                continue;
            }
            const LaunchTimer lt = launch_timer(F_STREAM);
            stream_halfcell(sc, strmA.view(), strmB.view(&strmA), nlist, tg, lt);
            launch_timed(F_STREAM, lt);
            if (mxi) {
                double* rl = last ? rho_lines.data() + long(n) * (nx + ny) : nullptr;
                double* rli = last && rho_lines_i.size() ? rho_lines_i.data() + long(n) * (nx + ny) : nullptr;
                tic(F_MODES);
                if (yf) {
                    TileGeom tr = tg;
                    tr.ntile = 0;                   // the windows are already in the float planes
                    pass_mixed(m, fft, jsr_f.data(), slab_f.data(), tr, Xf, Tf, F.data() + long(n - a) * nk, fpl, rl,
                               rli, i0, j0, 0, rpw);
                } else {
                    // free-hand y: the inverse x transforms stay FP64 (F from FP32 transforms moved the rear slices
                    // of config 3 to 3e-4 in 5 steps, FP64 ones keep them at 2e-5)
                    rows_fwd_mixed_fh(fft, jsr_f.data(), sp3.data(), rl, rli, i0, j0, rpw);
                    pass_work(sp3.data(), 0);
                    real_fields(n);
                }
                toc(F_MODES);
                continue;
            }
            tic(F_SFFT);
            TileGeom tr = tg;
            if (tg.fh) tr.ntile = 0;             // free-hand: the windows are already in jsr
            pass_rows_fwd(jsr.data(), tr, fft, sp3.data(), last ? rho_lines.data() + long(n) * (nx + ny) : nullptr,
                          last && rho_lines_i.size() ? rho_lines_i.data() + long(n) * (nx + ny) : nullptr, i0, j0, cm);
            toc(F_SFFT);
            tic(F_MODES);
            pass_work(sp3.data(), 0);
            toc(F_MODES);
            real_fields(n);
            if (tape_ && !last)
                tape_slice(tape_->Fmid.data() + size_t(pass) * F.size() + size_t(n - a) * nk,
                           F.data() + size_t(n - a) * nk);
        }
        // the passes wrote the new state into strmB: it becomes the list
        std::swap(strmA, strmB);
        // strmB borrowed the constants of strmA; after the swap the owner is strmA again only if the constant arrays
        // move with it, so swap the constants back
        strmA.w.swap(strmB.w); strmA.rem.swap(strmB.rem); strmA.sp.swap(strmB.sp); strmA.id.swap(strmB.id);
        if (ionz) {
            append_born(born.view(), counters.data(), strmA.view(), strmB.view(&strmA), nlist, nlist + born_cap);
            int cnt = 0;
            hip_check(hipMemcpyAsync(&cnt, counters.data(), sizeof(int), hipMemcpyDeviceToHost, stream()), "hipMemcpy D2H");
            hip_check(hipStreamSynchronize(stream()), "synchronize");
            if (cnt > sc.bcap) throw std::runtime_error("ionization: list of freed electrons overflowed");
            nlist += cnt;
        }
        have_dl = true;
    }
    (void)born_base;
    if (tape_) tape_copy(tape_->Fmarch.data(), F.data(), F.size() * sizeof(float));
    nstrm = nlist;
    {
        std::vector<unsigned long long> cnt(npush.size());
        npush.to_host(cnt.data(), 0, cnt.size());
        for (unsigned long long c : cnt) pushes_stream += long(c);
    }
    // trapped particles (alive 5): fast ones join the beam list, slow ones stream again in the next step
    std::vector<g3d_particle> sv;
    bool need_all = false;
    for (const auto& is : cfg.ions) need_all = need_all || is.ionization;
    if (need_all) sv = host_strm();
    else {
        counters.zero();
        if (aos.size() < size_t(std::max<long>(nstrm, 1))) aos.resize(size_t(std::max<long>(nstrm, 1)));
        if (aos_idx.size() < aos.size()) aos_idx.resize(aos.size());
        collect_alive(strmA.view(), nstrm, 5, aos.data(), aos_idx.data(), counters.data() + 1, long(aos.size()));
        int cnt = 0;
        hip_check(hipMemcpyAsync(&cnt, counters.data() + 1, sizeof(int), hipMemcpyDeviceToHost, stream()), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(stream()), "synchronize");
        if (cnt > 0) {
            std::vector<g3d_particle> t(cnt);
            std::vector<long> idx(cnt);
            aos.to_host(t.data(), 0, cnt);
            aos_idx.to_host(idx.data(), 0, cnt);
            std::vector<int> ord(cnt);
            std::iota(ord.begin(), ord.end(), 0);
            // the reference's list order: the fresh plasma by id, then the untiled tail in place
            auto key = [&](int k) { return idx[k] < n_tiled ? long(t[k].id - id0_) : idx[k]; };
            std::sort(ord.begin(), ord.end(), [&](int x, int y) { return key(x) < key(y); });
            for (int k : ord) sv.push_back(t[k]);
        }
    }
    std::vector<g3d_particle> add;
    for (auto& h : ion_stats) h.clear();
    for (auto& p : sv) {
        if (is_ion(p.sp) && cfg.ions[p.sp - SP_ION].ionization) {
            const auto& is = cfg.ions[p.sp - SP_ION];
            std::vector<double>& h = ion_stats[p.sp - SP_ION];
            if (h.empty()) h.assign(is.Zmax + 2, 0.0);
            const int Z = int(std::lround(p.q));
            if (p.alive == 1) h[Z] += p.w;
            h[is.Zmax + 1] += (Z - is.Z0) * p.w;
        }
        if (p.alive != 5) continue;
        const double gg = std::sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz);
        p.z = s1;
        if (p.pz / gg >= cfg.vz_beam) {
            p.alive = 1; p.sp = trapped_species(p.sp); ++trapped_total;
            add.push_back(p);
        } else { p.alive = 6; joining_.push_back(p); }
    }
    if (!add.empty()) {
        if (tape_) for (const auto& p : add) tape_->added.push_back(p.id);
        std::vector<g3d_particle> all = host_particles(), keep;
        keep.reserve(all.size() + add.size());
        for (auto& p : all) if (p.alive == 1) keep.push_back(p);
        keep.insert(keep.end(), add.begin(), add.end());
        set_particles(keep);
    }
    sub_planes(jkz.data(), js0z.data(), long(ns + 1) * nk);
    (void)g;
}

void Sim::collect_restreaming() {
    if (np <= 0) return;
    counters.zero();
    count_alive(beam.view(), np, 4, counters.data() + 2);
    int cnt = 0;
    hip_check(hipMemcpyAsync(&cnt, counters.data() + 2, sizeof(int), hipMemcpyDeviceToHost, stream()), "hipMemcpy D2H");
    hip_check(hipStreamSynchronize(stream()), "synchronize");
    if (!cnt) return;
    std::vector<g3d_particle> all = host_particles(), keep;
    for (auto& p : all) {
        if (p.alive == 4) { g3d_particle q = p; q.alive = 6; ++restreamed_total; joining_.push_back(q); }
        else if (p.alive == 1) keep.push_back(p);
    }
    set_particles(keep);
}

// ---------------------------------------------------------------- window (src/sim_setup.cpp:222-281)
void Sim::window_shift_post() {
    if (!cfg.move_x && !(yf && cfg.move_y)) return;
    std::array<double, 3> c{0, 0, 0};
    if (np > 0) {
        bunch_centroid(beam.view(), np, X0, Y0, cent_dev.data());
        cent_dev.to_host(c.data(), 0, 3);
    }
    cent_.push_back(c);
}
void Sim::window_shift_start() {
    if (!cfg.move_x && !(yf && cfg.move_y)) return;
    if (long(cent_.size()) < cfg.window_lag) return;
    const std::array<double, 3> c = cent_.front();
    cent_.pop_front();
    if (c[2] <= 0) return;
    const int nxs = cfg.move_x ? int(std::lround((c[0] / c[2] - (X0 + 0.5 * cfg.Lx)) / dx)) : 0;
    const int nys = yf && cfg.move_y ? int(std::lround((c[1] / c[2] - (Y0 + 0.5 * cfg.Ly)) / dy)) : 0;
    if (std::abs(nxs) >= 2 || std::abs(nys) >= 2) shift_window(nxs, nys);
}
void Sim::shift_window(int nxs, int nys) {
    ensure_layout(false, false);
    X0 += nxs * dx; Y0 += nys * dy;
    const double sx = nxs * dx, sy = nys * dy;
    parts_shift(beam.view(), np, sx, sy);
    for (auto& p : joining_) { p.x -= sx; p.y -= sy; }
    if (!cfg.self_fields) return;
    auto rot = [&](DevBuf<cx>& A) {
        rotate_planes(A.data(), long(A.size()), nk, cfg.nx, kxs.data(), kys.data(), sx, sy, nys != 0);
    };
    rot(Rx); rot(Ry); rot(Ez); rot(Bz); rot(j0x); rot(j0y);
    if (cfg.plasma_on) { rot(js0x); rot(js0y); rot(js0z); }
    DevBuf<float> tmp(F.size());
    roll_F(F.data(), tmp.data(), long(F.size()), nk, cfg.nx, cfg.ny, nxs, nys);
    F.swap(tmp);
    hip_check(hipStreamSynchronize(stream()), "synchronize");
}

}  // namespace g3d
