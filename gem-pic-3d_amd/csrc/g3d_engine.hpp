// The engine: a Simulation of the per-slice march on one MI355X.  The member functions keep the names and the
// contracts of the reference's Simulation (src/sim.hpp:222-360) so that each one is the drop-in of the
// reference's function of the same name; their bodies are sequences of the gfx950 kernels of this library on one
// stream, with no host synchronisation inside the half-cell loop.
#pragma once
#include <array>
#include <cstdio>
#include <deque>
#include <string>
#include <vector>
#include "g3d_kernels.hpp"
#include "g3d_runtime.hpp"

namespace g3d {

// physical constants (SI) of the reference's unit system
constexpr double kC = 299792458.0, kE = 1.602176634e-19, kMe = 9.1093837015e-31, kEps0 = 8.8541878128e-12;

struct Config {
    double lambda0 = 0, Lu = 0, Tu = 0, nc = 0;
    int nx = 0, ny = 0, nzeta = 0;
    double Lx = 0, Ly = 0, Lwin = 0;
    bool y_fourier = true;
    std::vector<double> zeta, ynode;
    double dz = 0, z_end = 0, dz_push = 0, sponge = 0, sponge_y = 0, sponge_rate = 0.1;
    bool move_x = true, move_y = true;
    int window_lag = 4;
    bool self_fields = true, zeta_filter = true, k_filter = true;
    double push_phase = 0.05;
    bool laser_on = false;
    int pol = 0;
    double a0 = 0, kL = 1, theta = 0, tau = 1, cep = 0, start_fwhm = 3.5, stop_fwhm = 3.5;
    double charge = 0, gamma = 1, sigx = 0, sigy = 0, sigz = 0, emitx = 0, emity = 0, espread = 0, zeta_center = 0;
    long nmacro = 0;
    unsigned long seed = 1;
    struct Bunch {                    // the bunch blocks (src/sim.hpp:61-80): [0] the single-bunch fields above
        double charge, gamma, sigx, sigy, sigz, emitx, emity, espread, zeta_center;
        long nmacro;
        unsigned long seed;
        std::string name;
    };
    std::vector<Bunch> bunches;
    std::vector<double> prof;         // profile_z of the first bunch (empty: Gaussian in zeta), prof_zmin ... prof_zmax
    double prof_zmin = 0, prof_zmax = 0;
    std::vector<g3d_glaser> glasers;
    bool plasma_on = false;
    double n0 = 0, r_max = 1e300, channel_depth = 0, r_channel = 1, smooth_length = 0;
    std::vector<double> pz_points, pz_values;
    bool parabolic = false;
    int ppc_x = 2, ppc_y = 2, iterations = 2;
    double vz_min = 0.8, vz_beam = 0.9, trap_advance = 1.0;
    bool predictor_extrapolate = true, free_electrons = true;
    std::vector<g3d_ion> ions;
    bool freeze_bunch = false, start_relax = true, box_mean = true;
    bool mixed = false;               // FP32-mixed plasma pass

    static Config from_abi(const g3d_config& g);
    double plasma_density(double x, double y, double z) const;
    double plasma_zfactor(double z) const;    // the longitudinal profile factor of plasma_density
    bool loaded(const g3d_ion& s) const { return s.mobile || s.ionization; }
};

// a piecewise-linear function as the sum of its hat functions v[k] max(0, 1 - |x - x_k| / h), x_k = x0 + k h
double hat_sum(const double* v, int n, double x0, double h, double x);

// structure-of-arrays particle storage; `shared` borrows w, sp, id, rem of another store (the second buffer of
// the streaming list)
struct PartStore {
    DevBuf<double> x, y, zeta, px, py, pz, w, z, q, qm, rem;
    DevBuf<int> alive, sp;
    DevBuf<unsigned long long> id;
    size_t cap = 0;
    void resize(size_t n, bool own_constants = true);
    Parts view(const PartStore* shared = nullptr);
};

// The record of one step kept for the reverse of the march (Sim::gradient, g3d_grad.cpp): the beam before and
// after its push, the real fields before the step, after the march and after every pass but the last of each
// half-cell, and the streaming particles before every half-cell.  step() and march_plasma fill it while
// Sim::tape_ points to it.
struct StepTape {
    double s0 = 0;
    long np = 0, nstrm = 0;
    DevBuf<double> beam0, beam1;      // x y zeta px py pz z rem, [8][np]: before and after the push
    DevBuf<int> alive0, alive1;
    DevBuf<double> qm;
    DevBuf<float> F0, Fmarch, Fmid;   // Fmid: [passes - 1][6][ns + 2][nk]
    DevBuf<double> ps;                // [N - 1][7][nstrm]: x y zeta z px py pz before half-cell n
    DevBuf<int> pa;                   // [N - 1][nstrm]
    DevBuf<double> pq;                // q qm w, [3][nstrm]
    DevBuf<unsigned long long> ids;   // the ids of the streaming list (its order is that of the whole step)
    std::vector<unsigned long long> added;   // ids of the electrons trapped into the beam at the end of the march
};

class Sim {
public:
    explicit Sim(const Config& c);
    ~Sim();
    void init();
    void step();
    bool finished() const { return s >= cfg.z_end - 1e-9 || stop_; }

    // the reference's member functions of the five drop-in files
    void push(double s0, double s1);
    void currents();
    void march(double D);
    void sponge_and_real_fields();
    void solve_fields(double D);
    void comoving_field();
    void relax_start_field();
    void add_grid_lasers();
    GridLaserArgs grid_laser_args(const g3d_glaser& Lz);     // the fill of one grid laser (env: none)
    void load_plasma(double s0, double s1);
    void stream_spectra();
    void march_slice(int n, double D);
    void real_fields_slice(int n);
    void march_plasma(double s0, double s1);
    void stream_pass(int n, int pass, double s0, double s1);
    void set_strm(const std::vector<g3d_particle>& v);
    StreamArgs stream_args(double s0, double s1);
    void collect_restreaming();
    void load_bunch();
    void load_bunch_profile(std::vector<g3d_particle>& mine, double ylo, double yhi, double yoff);
    void comoving_fields_per_bunch();
    void window_shift_start();
    void window_shift_post();
    void shift_window(int nxs, int nys);
    void j0_from_jk();

    // output and checkpoints in the reference's formats (g3d_output.cpp)
    std::vector<std::string> beam_names() const;
    int beam_index(int sp) const;
    void write_bunches(FILE* f, bool header);
    void write_slices(const std::string& file, double slice_length);
    void diagnostics(FILE* f, bool header);
    double total_field_energy();
    void dump_particles(const std::string& file);
    void write_frame(const std::string& file, int comps = 255);
    void write_axis(const std::string& file);
    std::string checkpoint_file(const std::string& dir, long step) const;
    void write_checkpoint(const std::string& dir);
    void read_checkpoint(const std::string& dir, long step);
    bool stop_ = false;               // the laser has left the bunch by stop_fwhm (diagnostics)

    // The gradient of the mean p_z of the live species-0 beam after `steps` steps from init() in the envelope
    // values of the first grid laser and in the values of the bunch's current profile (the reverse of the march,
    // g3d_grad.cpp; no counterpart in the reference).  State checkpoints every `every` steps; each segment is
    // run again with its tape and reversed.
    void gradient(int steps, int every, double* loss, double* genv, int nenv, double* gprof, int nprof);
    // the derivative of the same loss along the direction (denv, dprof) by the tangent-linear run (g3d_grad.cpp)
    void tangent(int steps, const double* denv, int nenv, const double* dprof, int nprof, double* loss, double* dloss);
    StepTape* tape_ = nullptr;
    void tape_beam(StepTape& t, bool before);     // the beam state into t.beam0 (before the push) or t.beam1
    void init_recording(bool dprof);              // init(), recording the profile derivatives when dprof
    bool record_dprof_ = false;       // load_bunch_profile records the profile derivatives below
    std::vector<double> dprof_;       // per loaded particle: d zeta_f / d prof[m] (nprof), then dx, dy, dzeta / d zeta_f

    std::vector<g3d_particle> host_particles();
    void set_particles(const std::vector<g3d_particle>& v);
    std::vector<g3d_particle> host_strm();

    Config cfg;
    int N = 0, a = 0, b = 0, ns = 0;
    long nk = 0;
    double dx = 0, dy = 0, s = 0, X0 = 0, Y0 = 0, beta_b = 1, gm2_b = 0;
    bool yf = true, starting_ = false, comoving_c_ = false;
    long istep = 0, np = 0, nstrm = 0, born_cap = 0, fresh_last = 0;
    long trapped_total = 0, restreamed_total = 0, lost_total = 0;
    // pushes done since the simulation was made: streaming particles through a half-cell in one pass, and beam
    // particles from s0 to s1 (the particles of the list at the start of push)
    long pushes_stream = 0, pushes_beam = 0;
    DevBuf<unsigned long long> npush;     // per-workgroup push counts of the stream launches of a step
    PlaneWave laser{};
    std::vector<g3d_particle> joining_;
    std::vector<std::vector<double>> ion_stats;

    // device arrays in the reference's layouts
    DevBuf<cx> Rx, Ry, Ez, Bz, Lx, Ly, j0x, j0y, jkx, jky, jkz, S1, S2, T3;
    DepositScratch dep_scratch;      // ordered beam deposit
    BpushScratch bpush_scratch;      // binned beam push with field windows
    bool bpush_windows = true;       // the beam push bins its particles and stages field windows (G3D_BPUSH_WIN=0: off)
    DevBuf<cx> js0x, js0y, js0z, jsn, jsin, jsi0, jpred, dlt, bk, sp3, tp3;
    DevBuf<double> jw, jsr, rho_lines, rho_lines_i;
    DevBuf<float> jsr_f, slab_f;      // FP32-mixed pass: the half-cell current and the tile slabs
    DevBuf<float> F;
    DevBuf<double> kxd, kxs, filt, maskx, kyd, kys, filty, masky_n, masky_h, iD, iA, zs, hcg, wlo, whi, wsl;
    std::vector<double> wlo_h, hcg_h, hDel, hA;
    // free-hand y: node table, LU factors of k_x^2 - Lap_y, eigenvectors of -Lap_y, Thomas factors of the lines
    DevBuf<double> yn, pcp, pid, evf, evb, evl, scp;
    DevBuf<double> piv;                    // free-hand y: the march's pivots of every line and slice (lview)
    double piv_D = 0;
    bool piv_ok = false;              // piv holds the pivots of D = piv_D and the current march geometry
    bool line_pivots_on = true;       // form the pivot table (false: march_line forms the pivots as it goes)
    Axis zax{}, yax{};
    FftPlan fft;
    PartStore beam, strmA, strmB, born;
    DevBuf<IonDev> ion_dev;
    DevBuf<AdkLevel> adk_dev;
    DevBuf<int> counters;             // [0] freed electrons of the pass, [1] collected, [2] counted
    DevBuf<double> cent_dev;
    DevBuf<g3d_particle> aos;         // staging of AoS lists
    DevBuf<long> aos_idx;
    std::deque<std::array<double, 3>> cent_;
    // transverse tiles of the streaming list (tile-ordered fresh plasma, LDS deposit windows)
    TileGeom tiles{};
    DevBuf<long> tile_off;
    DevBuf<int> bin_idx, bin_cnt, bin_cur;   // binning of the live list by current tile
    DevBuf<double> slab;
    long n_tiled = 0;
    unsigned long long id0_ = 0;

    // timing of kernel families with HIP events on the library stream
    unsigned timing = 0;              // bit f: time family f
    int timing_every = 1;             // launch-event families: time one launch in timing_every
    static constexpr int kFam = 8;
    long fam_seen[kFam] = {};         // launches seen per family since timing was set
    double fam_ms[kFam] = {};
    long fam_n[kFam] = {};
    void tic(int fam);
    void toc(int fam);
    void flush_timers();
    LaunchTimer launch_timer(int fam);
    void launch_timed(int fam, const LaunchTimer& lt);
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    std::vector<hipEvent_t> ev_pool;
    hipEvent_t ev_open = nullptr;
    int ev_open_fam = -1;

    // The spectra of the plasma pass in the column-major layout of the fused column kernel (mode ix ny + j) or in
    // the reference's (j nx + ix): R, E_z, B_z, L, js0, j0 (cmR) and jk (cmJ).  The FP32-mixed march sets them
    // column-major and they stay so through the sponge and into the next step; every other member function and
    // array access asks for the reference's layout first.
    void ensure_layout(bool r_cm, bool j_cm);
    bool cmR = false, cmJ = false;
    bool jw_zero = false;             // jw known to be zero (no beam deposit since it was cleared)

private:
    SpecGeom geom() const;
    MarchView mview(double D);
    LineView lview(double D);
    FieldView fview();
    void setup_ions();
    void setup_y();
    void setup_tiles();
    void ensure_strm(size_t need);
    bool load_plasma_device(double s0, double s1);
    DevBuf<cx> trF, trC;
    DevBuf<int> chunk_tile, chunk_cnt;        // further chunks of dense tiles (FP32-mixed stream kernel)
    DevBuf<long> chunk_start;
};

extern const char* const kFamilyNames[Sim::kFam];

}  // namespace g3d
