// Output and checkpoints of the engine, in the reference's formats: bunches.txt and slices_NNNNNN.txt (centred beam
// moments, src/sim_bunch_diag.cpp:9-97), diag.txt (src/sim_output.cpp:5-52), the field energy (54-79), particle
// dumps (81-97), the frame / fields_NNNNNN.bin files (103-198), the axis lineouts (204-245) and the G3DCKPT1
// checkpoint (src/sim_checkpoint.cpp:9-105).  The records, the column formats and the summation orders are the
// reference's, so that a run of the reference on one thread writes the same text and the same bytes from the same
// state.  Everything here is host code over host copies of the device arrays (an output step, not the march).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include "g3d_engine.hpp"

namespace g3d {
namespace {

// physical constants of the reference's unit system (src/sim_internal.hpp:10)
constexpr double kMeV = 0.51099895;

// moments of groups of beam particles (src/sim_bunch_diag.cpp:9-28): per group w, n, w x, w y, w zeta, w ux, w uy,
// w gamma; second moments centred on the first (two passes)
struct Moments {
    static constexpr int N1 = 8, N2 = 8;
    std::vector<double> s1, s2;
    explicit Moments(size_t groups) : s1(groups * N1, 0.0), s2(groups * N2, 0.0) {}
    void add1(size_t g, double w, const double v[6]) {
        double* a = &s1[g * N1];
        a[0] += w; a[1] += 1;
        for (int k = 0; k < 6; ++k) a[2 + k] += w * v[k];
    }
    double mean(size_t g, int k) const { const double* a = &s1[g * N1]; return a[0] > 0 ? a[2 + k] / a[0] : 0; }
    void add2(size_t g, double w, const double v[6]) {
        double d[6];
        for (int k = 0; k < 6; ++k) d[k] = v[k] - mean(g, k);
        double* a = &s2[g * N2];
        a[0] += w * d[0] * d[0]; a[1] += w * d[1] * d[1]; a[2] += w * d[2] * d[2]; a[3] += w * d[3] * d[3];
        a[4] += w * d[4] * d[4]; a[5] += w * d[0] * d[3]; a[6] += w * d[1] * d[4]; a[7] += w * d[5] * d[5];
    }
    double var(size_t g, int k) const { const double w = s1[g * N1]; return w > 0 ? s2[g * N2 + k] / w : 0; }
};

// one group's line: N, Q [pC], E [MeV], dE/E, sig_x, sig_y, eps_nx, eps_ny [um], <zeta>, sig_z [um], <x>, <y> [um]
void print_group(FILE* f, const Moments& M, size_t g, double pC, double um) {
    const double w = M.s1[g * Moments::N1], gm = M.mean(g, 5);
    const double ex = std::sqrt(std::max(0.0, M.var(g, 0) * M.var(g, 3) - M.var(g, 5) * M.var(g, 5)));
    const double ey = std::sqrt(std::max(0.0, M.var(g, 1) * M.var(g, 4) - M.var(g, 6) * M.var(g, 6)));
    std::fprintf(f, " %10ld %12.6f %13.6f %10.3e %10.5f %10.5f %10.5f %10.5f %11.4f %10.4f %11.5f %11.5f\n",
                 long(M.s1[g * Moments::N1 + 1]), w * pC, w > 0 ? (gm - 1) * kMeV : 0.0,
                 gm > 0 ? std::sqrt(M.var(g, 7)) / gm : 0.0, std::sqrt(M.var(g, 0)) * um, std::sqrt(M.var(g, 1)) * um,
                 ex * um, ey * um, M.mean(g, 2) * um, std::sqrt(M.var(g, 2)) * um, M.mean(g, 0) * um,
                 M.mean(g, 1) * um);
}

bool is_half_comp(int comp) { return comp == 1 || comp == 3 || comp == 5; }     // E_y, B_x, B_z (free-hand y)

const char kMagic[8] = {'G', '3', 'D', 'C', 'K', 'P', 'T', '1'};
// the checkpoint header, field for field and padded as the reference's (src/sim_checkpoint.cpp:11-17)
struct CkptHeader {
    char magic[8];
    int ranks, rank, N, a, b, ns, nk, nx, ny, plasma;
    long istep, lost, trapped, restreamed;
    double s, X0, Y0, beta_b, gm2_b;
    int stop, ncent;
};
static_assert(sizeof(CkptHeader) == 128, "the G3DCKPT1 header is 128 bytes");
template <class T> void put(FILE* f, const std::vector<T>& v) {
    const unsigned long long n = v.size();
    if (std::fwrite(&n, sizeof n, 1, f) != 1 || (n && std::fwrite(v.data(), sizeof(T), n, f) != n))
        throw std::runtime_error("checkpoint: write error");
}
template <class T> std::vector<T> get(FILE* f) {
    unsigned long long n = 0;
    if (std::fread(&n, sizeof n, 1, f) != 1) throw std::runtime_error("checkpoint: file too short");
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw std::runtime_error("checkpoint: file too short");
    return v;
}
template <class T> std::vector<T> host_copy(const DevBuf<T>& b) {
    std::vector<T> v(b.size());
    b.to_host(v.data(), 0, v.size());
    return v;
}
}  // namespace

// beam-list species of the output (src/sim.hpp:132-143): the bunch, the trapped plasma electrons, then the electrons
// injected by the ionization of each ion species
std::vector<std::string> Sim::beam_names() const {
    std::vector<std::string> v;
    for (const auto& b : cfg.bunches) v.push_back(b.name);
    v.push_back("trapped");
    for (size_t k = 0; k < cfg.ions.size(); ++k) {
        const char* nm = cfg.ions[k].name;
        v.push_back("inj_" + (nm[0] ? std::string(nm, strnlen(nm, sizeof(cfg.ions[k].name))) : "ion" + std::to_string(k + 1)));
    }
    return v;
}
int Sim::beam_index(int sp) const {
    const int nb = int(cfg.bunches.size());
    return is_bunch(sp) ? -sp : sp >= SP_INJECTED ? nb + 1 + sp - SP_INJECTED : nb;
}

static Moments beam_moments(Sim& S, int nsl, double len) {
    const std::vector<g3d_particle> P = S.host_particles();
    const size_t ng = S.beam_names().size() * size_t(nsl);
    Moments M(ng);
    auto group = [&](const g3d_particle& p, size_t& g, double v[6]) {
        if (p.alive != 1) return false;
        const int sl = nsl > 1 ? int(std::floor(p.zeta / len)) : 0;
        if (sl < 0 || sl >= nsl) return false;
        g = size_t(S.beam_index(p.sp)) * nsl + sl;
        v[0] = p.x + S.X0; v[1] = p.y + S.Y0; v[2] = p.zeta; v[3] = p.px; v[4] = p.py;
        v[5] = std::sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz);
        return true;
    };
    size_t g; double v[6];
    for (const auto& p : P) if (group(p, g, v)) M.add1(g, p.w, v);
    for (const auto& p : P) if (group(p, g, v)) M.add2(g, p.w, v);
    return M;
}

void Sim::write_bunches(FILE* f, bool header) {
    if (!f) return;
    const Moments M = beam_moments(*this, 1, 0);
    if (header)
        std::fprintf(f, "# step  z[mm]  bunch  N  Q[pC]  E[MeV]  dE/E(rms)  sig_x[um]  sig_y[um]  eps_nx[um]  eps_ny[um]  "
                        "<zeta>[um]  sig_z[um]  <x>[um]  <y>[um]    (trapped = self-injected plasma electrons, "
                        "inj_<ion> = from field ionization)\n");
    const double pC = cfg.nc * std::pow(cfg.Lu, 3) * kE * 1e12, um = cfg.Lu * 1e6;
    const std::vector<std::string> names = beam_names();
    for (size_t k = 0; k < names.size(); ++k) {
        std::fprintf(f, "%6ld %10.5f %-12s", istep, s * cfg.Lu * 1e3, names[k].c_str());
        print_group(f, M, k, pC, um);
    }
    std::fflush(f);
}

void Sim::write_slices(const std::string& file, double len) {
    if (!(len > 0)) throw std::invalid_argument("write_slices: slice_length must be positive");
    const int nsl = std::max(1, int(std::ceil(cfg.Lwin / len - 1e-9)));
    const Moments M = beam_moments(*this, nsl, len);
    FILE* f = std::fopen(file.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + file);
    std::fprintf(f, "# GEM-PIC-3D beam slices of %.9g um, step %ld, z = %.9g mm\n", len * cfg.Lu * 1e6, istep,
                 s * cfg.Lu * 1e3);
    std::fprintf(f, "# bunch  zeta_centre[um]  N  Q[pC]  E[MeV]  dE/E(rms)  sig_x[um]  sig_y[um]  eps_nx[um]  eps_ny[um]  "
                    "<zeta>[um]  sig_z[um]  <x>[um]  <y>[um]\n");
    const double pC = cfg.nc * std::pow(cfg.Lu, 3) * kE * 1e12, um = cfg.Lu * 1e6;
    const std::vector<std::string> names = beam_names();
    for (size_t k = 0; k < names.size(); ++k)
        for (int sl = 0; sl < nsl; ++sl) {
            const size_t g = k * nsl + sl;
            if (!(M.s1[g * Moments::N1] > 0)) continue;
            std::fprintf(f, "%-12s %11.4f", names[k].c_str(), (sl + 0.5) * len * um);
            print_group(f, M, g, pC, um);
        }
    std::fclose(f);
}

// the field energy of the slices in code units, dual-cell weights along zeta and y (src/sim_output.cpp:54-79), the
// points summed in the reference's order
double Sim::total_field_energy() {
    if (!cfg.self_fields) return 0;
    const std::vector<float> Fp = host_copy(F);
    std::vector<double> ws(wsl.size()), ia(iA.size()), id(iD.size());
    wsl.to_host(ws.data(), 0, ws.size()); iA.to_host(ia.data(), 0, ia.size()); iD.to_host(id.data(), 0, id.size());
    const long fpl = long(ns + 2) * nk;
    const int nx = cfg.nx;
    double A = 0;
    for (long i = 0; i < nk; ++i) {
        const int j = int(i / nx);
        const double wn = 1 / ia[j], wh = 1 / id[j];
        double e = 0;
        for (int l = 0; l < ns; ++l) {
            const double w = ws[a + l];
            for (int comp = 0; comp < 6; ++comp) {
                const double v = Fp[comp * fpl + l * nk + i];
                e += 0.5 * w * (is_half_comp(comp) ? wh : wn) * v * v;
            }
        }
        A += e;
    }
    return A * dx;
}

void Sim::diagnostics(FILE* f, bool header) {
    if (header && f)
        std::fprintf(f, "# step  z[mm]  N_alive  <x>[um] <y>[um] <zeta>[um]  rms_x rms_y rms_zeta[um]  <px> <py> <pz> "
                        "<gamma> <p.khat>  rms_pz  eta_rel[FWHM] a_env  W_field[J] W_kin[J]  N_trapped Q_trapped[pC]%s\n",
                     rho_lines_i.size() ? "  max_Zni/n0(axis)" : "");
    double S[19] = {0};
    for (auto& p : host_particles()) {
        if (p.alive != 1) continue;
        if (!is_bunch(p.sp)) { S[17] += 1; S[18] += p.q * p.w; continue; }
        const double g = std::sqrt(1 + p.px * p.px + p.py * p.py + p.pz * p.pz);
        const double X = p.x + X0, Y = p.y + Y0;
        const double et = s - (p.zeta - cfg.zeta_center) - laser.kx * X - laser.kz * s;
        const double v[] = {1, X, Y, p.zeta, X * X, Y * Y, p.zeta * p.zeta, p.px, p.py, p.pz, g,
                            p.px * laser.kx + p.pz * laser.kz, p.pz * p.pz, et};
        for (int k = 0; k < 14; ++k) S[k] += p.w * v[k];
        S[14] += 1;
    }
    S[15] = cfg.self_fields ? total_field_energy() : 0;
    S[16] = double(lost_total);
    const double w = S[0] > 0 ? S[0] : 1;
    auto m = [&](int k) { return S[k] / w; };
    auto rms = [&](int k, int k2) { return std::sqrt(std::max(0.0, m(k2) - m(k) * m(k))); };
    const double eta_rel = laser.on ? (m(13) - laser.eta_c) / laser.tau : 0;
    if (laser.on && cfg.stop_fwhm > 0 && eta_rel > cfg.stop_fwhm) stop_ = true;
    if (!f) return;
    const double um = cfg.Lu * 1e6;
    const double ex = m(13) - laser.eta_c;
    const double aenv = laser.on ? laser.a0 * std::exp(-1.3862943611198906 * ex * ex / (laser.tau * laser.tau)) : 0;
    const double Efield = kMe * kC * kC / (cfg.Lu * kE);
    const double Wf = S[15] * kEps0 * Efield * Efield * std::pow(cfg.Lu, 3);
    const double Wk = (S[10] - S[0]) * cfg.nc * std::pow(cfg.Lu, 3) * kMe * kC * kC;
    const double Qtr = S[18] * cfg.nc * std::pow(cfg.Lu, 3) * kE * 1e12;
    char line[600];
    std::snprintf(line, sizeof line,
                  "%6ld %10.5f %8ld %10.4f %10.4f %10.4f %9.4f %9.4f %9.4f %12.6e %12.6e %14.8e %14.8e %14.8e %11.4e "
                  "%8.4f %8.4f %11.4e %11.4e %8ld %11.4e\n",
                  istep, s * cfg.Lu * 1e3, long(S[14]), m(1) * um, m(2) * um, m(3) * um, rms(1, 4) * um, rms(2, 5) * um,
                  rms(3, 6) * um, m(7), m(8), m(9), m(10), m(11), rms(9, 12), eta_rel, aenv, Wf, Wk, long(S[17]), Qtr);
    if (rho_lines_i.size()) {                 // largest mobile-ion charge density on the axis
        const std::vector<double> rl = host_copy(rho_lines_i);
        const int nx = cfg.nx, i0 = ((int(std::lround(-X0 / dx)) % nx) + nx) % nx;
        double mx = 0;
        for (int n = 0; n + 1 < N; ++n) mx = std::max(mx, rl[size_t(n) * (nx + cfg.ny) + i0] / cfg.n0);
        const size_t l = std::strlen(line);
        std::snprintf(line + l - 1, sizeof line - l + 1, " %11.4e\n", mx);
    }
    std::fputs(line, f);
    std::fflush(f);
}

void Sim::dump_particles(const std::string& file) {
    std::ofstream o(file);
    if (!o) throw std::runtime_error("cannot write " + file);
    o.precision(10);
    o << "# x[um] y[um] zeta[um] px py pz (m c), z = " << s * cfg.Lu * 1e3 << " mm\n";
    const double um = cfg.Lu * 1e6;
    for (auto& p : host_particles())
        if (p.alive == 1)
            o << (p.x + X0) * um << ' ' << (p.y + Y0) * um << ' ' << p.zeta * um << ' ' << p.px << ' ' << p.py << ' '
              << p.pz << '\n';
}

// E and B (and with plasma the densities) in the planes y = 0 and x = 0 through the axis: a JSON header line, then
// float32 [nzeta][nx] and [nzeta][ny] per component (src/sim_output.cpp:103-198)
void Sim::write_frame(const std::string& file, int comps) {
    const int nx = cfg.nx, ny = cfg.ny, ncomp = 6;
    const int i0 = ((int(std::lround(-X0 / dx)) % nx) + nx) % nx;
    const int j0 = yf ? ((int(std::lround(-Y0 / dy)) % ny) + ny) % ny : ny / 2;
    const std::vector<float> Fh = cfg.self_fields ? host_copy(F) : std::vector<float>(F.size(), 0.0f);
    std::vector<float> A(size_t(ns) * ncomp * (nx + ny));
    const long fpl = long(ns + 2) * nk;
    for (int l = 0; l < ns; ++l)
        for (int c = 0; c < ncomp; ++c) {
            const float* P = Fh.data() + c * fpl + long(l) * nk;
            float* o = &A[(size_t(l) * ncomp + c) * (nx + ny)];
            const bool half = !yf && is_half_comp(c);
            for (int i = 0; i < nx; ++i)
                o[i] = half ? 0.5f * (P[long(j0 - 1) * nx + i] + P[long(j0) * nx + i]) : P[long(j0) * nx + i];
            for (int j = 0; j < ny; ++j) o[nx + j] = P[long(j) * nx + i0];
        }
    const bool dens = cfg.plasma_on && (comps >> 6 & 1);
    const bool idens = rho_lines_i.size() && (comps >> 7 & 1);
    const std::vector<double> rle = dens ? host_copy(rho_lines) : std::vector<double>();
    const std::vector<double> rli = idens ? host_copy(rho_lines_i) : std::vector<double>();
    const double um = cfg.Lu * 1e6;
    std::string h = "{\"step\": " + std::to_string(istep) + ", \"z_mm\": " + std::to_string(s * cfg.Lu * 1e3) +
                    ", \"nzeta\": " + std::to_string(N) + ", \"nx\": " + std::to_string(nx) + ", \"ny\": " +
                    std::to_string(ny);
    auto arr = [&](const std::vector<double>& v) {
        std::string t = "[";
        char b[32];
        for (size_t k = 0; k < v.size(); ++k) { std::snprintf(b, sizeof b, "%s%.6g", k ? ", " : "", v[k]); t += b; }
        return t + "]";
    };
    std::vector<double> zeta(N), xs(nx), yn(ny), yh(ny);
    for (int n = 0; n < N; ++n) zeta[n] = cfg.zeta[n] * um;
    for (int i = 0; i < nx; ++i) xs[i] = (((i - i0) % nx + nx + nx / 2) % nx - nx / 2) * dx * um;
    for (int j = 0; j < ny; ++j) {
        yn[j] = yf ? (((j - j0) % ny + ny + ny / 2) % ny - ny / 2) * dy * um : cfg.ynode[j] * um;
        yh[j] = yf ? yn[j] : 0.5 * (cfg.ynode[j] + cfg.ynode[j + 1]) * um;
    }
    std::string cl, hl;
    const char* cn[6] = {"ex", "ey", "ez", "bx", "by", "bz"};
    for (int c = 0; c < ncomp; ++c)
        if (comps >> c & 1) {
            cl += std::string(cl.empty() ? "" : ", ") + "\"" + cn[c] + "\"";
            hl += std::string(hl.empty() ? "" : ", ") + (!yf && is_half_comp(c) ? "true" : "false");
        }
    if (dens) { cl += std::string(cl.empty() ? "" : ", ") + "\"ne\""; hl += std::string(hl.empty() ? "" : ", ") + "false"; }
    if (idens) { cl += std::string(cl.empty() ? "" : ", ") + "\"ni\""; hl += std::string(hl.empty() ? "" : ", ") + "false"; }
    h += ", \"zeta_um\": " + arr(zeta) + ", \"x_um\": " + arr(xs) + ", \"y_um\": " + arr(yn) + ", \"y_half_um\": " +
         arr(yh) + ", \"comps\": [" + cl + "], \"half\": [" + hl + "], \"n0\": " + std::to_string(cfg.n0) +
         ", \"lambda0_um\": " + std::to_string(cfg.lambda0 * 1e6) + "}\n";
    const std::string tmp = file + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + tmp);
    std::fwrite(h.data(), 1, h.size(), f);
    std::vector<float> buf(size_t(N) * std::max(nx, ny));
    for (int c = 0; c < ncomp; ++c) {
        if (!(comps >> c & 1)) continue;
        for (int n = 0; n < N; ++n)
            for (int i = 0; i < nx; ++i) buf[size_t(n) * nx + i] = A[(size_t(n) * ncomp + c) * (nx + ny) + i];
        std::fwrite(buf.data(), sizeof(float), size_t(N) * nx, f);
        for (int n = 0; n < N; ++n)
            for (int j = 0; j < ny; ++j) buf[size_t(n) * ny + j] = A[(size_t(n) * ncomp + c) * (nx + ny) + nx + j];
        std::fwrite(buf.data(), sizeof(float), size_t(N) * ny, f);
    }
    for (int sgn = 0; sgn < 2; ++sgn) {          // n_e / n0, Z n_i / n0 at the slices: mean of the half-cells
        if (!(sgn ? idens : dens)) continue;
        const std::vector<double>& rl = sgn ? rli : rle;
        const int L = nx + ny;
        auto at = [&](int n, int k) {
            const double lo = n > 0 ? rl[size_t(n - 1) * L + k] : 0, hi = n < N - 1 ? rl[size_t(n) * L + k] : 0;
            return float((sgn ? 0.5 : -0.5) * (lo + hi) / cfg.n0);
        };
        for (int n = 0; n < N; ++n) for (int i = 0; i < nx; ++i) buf[size_t(n) * nx + i] = at(n, i);
        std::fwrite(buf.data(), sizeof(float), size_t(N) * nx, f);
        for (int n = 0; n < N; ++n) for (int j = 0; j < ny; ++j) buf[size_t(n) * ny + j] = at(n, nx + j);
        std::fwrite(buf.data(), sizeof(float), size_t(N) * ny, f);
    }
    std::fclose(f);
    std::rename(tmp.c_str(), file.c_str());
}

// one line per slice on the axis x = 0, y = 0: zeta [um], E_z, E_x - B_y, E_y + B_x and the densities
// (src/sim_output.cpp:200-245)
void Sim::write_axis(const std::string& file) {
    const int nx = cfg.nx, ny = cfg.ny;
    const int i0 = ((int(std::lround(-X0 / dx)) % nx) + nx) % nx;
    const int j0 = yf ? ((int(std::lround(-Y0 / dy)) % ny) + ny) % ny : ny / 2;
    const std::vector<float> Fh = cfg.self_fields ? host_copy(F) : std::vector<float>(F.size(), 0.0f);
    const long fpl = long(ns + 2) * nk;
    std::vector<double> A(size_t(ns) * 3);
    for (int l = 0; l < ns; ++l) {
        const long k = long(l) * nk + long(j0) * nx + i0, kh = yf ? k : k - nx;
        auto hv = [&](int comp) { return yf ? double(Fh[comp * fpl + k]) : 0.5 * (Fh[comp * fpl + k] + Fh[comp * fpl + kh]); };
        A[3 * l] = Fh[2 * fpl + k];
        A[3 * l + 1] = Fh[0 * fpl + k] - Fh[4 * fpl + k];
        A[3 * l + 2] = hv(1) + hv(3);
    }
    const std::vector<double> rl = cfg.plasma_on ? host_copy(rho_lines) : std::vector<double>();
    const std::vector<double> rli = rho_lines_i.size() ? host_copy(rho_lines_i) : std::vector<double>();
    const int L = nx + ny;
    FILE* f = std::fopen(file.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + file);
    std::fprintf(f, "# GEM-PIC-3D axis x = 0, y = 0, step %ld, z = %.9g mm; fields in m c omega0 / e\n", istep,
                 s * cfg.Lu * 1e3);
    std::fprintf(f, "# zeta_um Ez Ex-By Ey+Bx%s%s\n", cfg.plasma_on ? " ne/n0" : "", rli.empty() ? "" : " Zni/n0");
    for (int n = 0; n < N; ++n) {
        std::fprintf(f, "%.10g %.10e %.10e %.10e", cfg.zeta[n] * cfg.Lu * 1e6, A[3 * n], A[3 * n + 1], A[3 * n + 2]);
        if (cfg.plasma_on) {
            const double lo = n > 0 ? rl[size_t(n - 1) * L + i0] : 0, hi = n < N - 1 ? rl[size_t(n) * L + i0] : 0;
            std::fprintf(f, " %.10e", -0.5 * (lo + hi) / cfg.n0);
        }
        if (!rli.empty()) {
            const double lo = n > 0 ? rli[size_t(n - 1) * L + i0] : 0, hi = n < N - 1 ? rli[size_t(n) * L + i0] : 0;
            std::fprintf(f, " %.10e", 0.5 * (lo + hi) / cfg.n0);
        }
        std::fprintf(f, "\n");
    }
    std::fclose(f);
}

// ---------------------------------------------------------------- checkpoints (src/sim_checkpoint.cpp)
std::string Sim::checkpoint_file(const std::string& dir, long step) const {
    char fn[96];
    std::snprintf(fn, sizeof fn, "/checkpoint_%06ld_r%04d.bin", step, 0);
    return dir + fn;
}

void Sim::write_checkpoint(const std::string& dir) {
    ensure_layout(false, false);
    CkptHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.ranks = 1; h.rank = 0; h.N = N; h.a = a; h.b = b; h.ns = ns; h.nk = int(nk); h.nx = cfg.nx; h.ny = cfg.ny;
    h.plasma = cfg.plasma_on; h.istep = istep; h.lost = lost_total; h.trapped = trapped_total;
    h.restreamed = restreamed_total; h.s = s; h.X0 = X0; h.Y0 = Y0; h.beta_b = beta_b; h.gm2_b = gm2_b;
    h.stop = stop_; h.ncent = int(cent_.size());
    const std::string fn = checkpoint_file(dir, istep), tmp = fn + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + tmp);
    std::fwrite(&h, sizeof h, 1, f);
    for (DevBuf<cx>* p : {&Rx, &Ry, &Ez, &Bz, &Lx, &Ly, &j0x, &j0y, &js0x, &js0y, &js0z}) put(f, host_copy(*p));
    put(f, host_copy(F));
    put(f, host_particles());
    put(f, joining_);
    for (const auto& c : cent_) put(f, std::vector<double>(c.begin(), c.end()));
    std::fclose(f);
    std::rename(tmp.c_str(), fn.c_str());
}

void Sim::read_checkpoint(const std::string& dir, long step) {
    const std::string fn = checkpoint_file(dir, step);
    FILE* f = std::fopen(fn.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot read " + fn);
    try {
        CkptHeader h{};
        if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kMagic, 8))
            throw std::runtime_error(fn + ": not a GEM-PIC-3D checkpoint");
        if (h.ranks != 1) throw std::runtime_error(fn + ": written by " + std::to_string(h.ranks) + " ranks; restart with as many");
        if (h.N != N || h.a != a || h.b != b || h.nx != cfg.nx || h.ny != cfg.ny || h.plasma != int(cfg.plasma_on))
            throw std::runtime_error(fn + ": the grid (nx, ny, zeta slices) or the plasma differ from the deck");
        ensure_layout(false, false);
        istep = h.istep; lost_total = h.lost; trapped_total = h.trapped; restreamed_total = h.restreamed;
        s = h.s; X0 = h.X0; Y0 = h.Y0; beta_b = h.beta_b; gm2_b = h.gm2_b; stop_ = h.stop != 0;
        for (DevBuf<cx>* p : {&Rx, &Ry, &Ez, &Bz, &Lx, &Ly, &j0x, &j0y, &js0x, &js0y, &js0z}) {
            const std::vector<cx> v = get<cx>(f);
            if (v.size() != p->size()) throw std::runtime_error(fn + ": array sizes differ from the deck");
            if (!v.empty()) p->from_host(v.data(), 0, v.size());
        }
        const std::vector<float> fv = get<float>(f);
        if (fv.size() != F.size()) throw std::runtime_error(fn + ": array sizes differ from the deck");
        F.from_host(fv.data(), 0, fv.size());
        set_particles(get<g3d_particle>(f));
        joining_ = get<g3d_particle>(f);
        cent_.clear();
        for (int k = 0; k < h.ncent; ++k) {
            const std::vector<double> c = get<double>(f);
            if (c.size() != 3) throw std::runtime_error(fn + ": bad centroid record");
            cent_.push_back({c[0], c[1], c[2]});
        }
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    jw_zero = false;
    piv_ok = false;
}

}  // namespace g3d
