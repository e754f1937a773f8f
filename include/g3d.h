/* g3d.h - C ABI of the MI355X (gfx950) per-slice march of GEM-PIC-3D.
 *
 * Every entry point returns 0 on success or a nonzero g3d status; g3d_last_error() holds the message
 * ("HIP error in <what>: <msg>", or the logic error), worded as the reference's std::runtime_error texts
 * (src/portable.hpp:43-49, src/sim_plasma.cpp:358).  Pointers passed to kernel entry points are device
 * pointers (g3d_malloc) in the reference's layouts:
 *   - spectra: planes of nk = nx*ny complex values, index j*nx + ix, cpx = {re, im}   (src/sim.hpp:255,268-270)
 *   - real fields F: float [6][ns+2][nk], components Ex Ey Ez Bx By Bz              (src/sim.hpp:273)
 *   - beam deposit jw: double [3][ns+3][nk]; streaming deposit jsr: double [5][nk]   (src/sim.hpp:272,297)
 *   - particles on the device: structure of arrays (g3d_parts); the 104-byte AoS record g3d_particle is the
 *     host-side view (src/sim.hpp:217).
 * All work is enqueued on one HIP stream per process; g3d_sync() waits for it.  Calls that read device data
 * back to the host synchronise first, as the reference does before every host read.
 */
#ifndef G3D_H
#define G3D_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */
typedef struct { double re, im; } g3d_cpx;                     /* src/portable.hpp:25 (cufftDoubleComplex layout) */

/* one particle (host view), field for field the reference's Particle (src/sim.hpp:217), 104 bytes */
typedef struct {
    double x, y, zeta, px, py, pz, w, rem, z, q, qm;
    int alive, sp;
    unsigned long long id;
} g3d_particle;

/* device particle list, structure of arrays (one entry per particle; n used, cap allocated) */
typedef struct {
    double *x, *y, *zeta, *px, *py, *pz, *w, *z, *q, *qm, *rem;
    int *alive, *sp;
    unsigned long long* id;
} g3d_parts;

/* grid laser (laser block with type = gaussian), src/sim.hpp:95-101 */
typedef struct {
    double a0, kL, waist, tau, zeta_c, z_focus, cep, pol_angle;
    int circular, helicity;
    /* envelope_function a(zeta) in place of the Gaussian of tau (src/sim_config.cpp:146-158, sim_fields.cpp:335): the
     * piecewise-linear function through env[k] at zeta = env_z0 + k env_dz, k < n_env (0: Gaussian), zero beyond
     * one step outside; evaluated as the sum of the hat functions env[k] max(0, 1 - |zeta - zeta_k| / env_dz) */
    int n_env;
    double env_z0, env_dz, env[32];
} g3d_glaser;

/* a bunch block after the first (bunch2, bunch3, ...; src/sim.hpp:61-80 Bunch): Gaussian electrons with their own
 * seed and name; the first bunch is the single-bunch fields of g3d_config, named "bunch" */
typedef struct {
    double charge, gamma, sigx, sigy, sigz, emitx, emity, espread, zeta_center;
    long nmacro;
    unsigned long seed;
    char name[16];
} g3d_bunch;

/* ion species (ion blocks), src/sim.hpp:118-127 */
typedef struct {
    double mass, atoms;
    int Z0, Zmax, mobile, ionization, n_ip;
    double ip_eV[16];
    char name[16];            /* the ion block's name (the output's inj_<name>); empty: ion<k> */
} g3d_ion;

/* run configuration: the fields of the reference's Config (src/sim.hpp:36-187) that the march uses, in code
 * units (lengths 1/k0, times 1/omega0, densities n_c).  Tables are borrowed for the duration of g3d_sim_create. */
typedef struct {
    double lambda0;                               /* reference wavelength [m] (units only) */
    int nx, ny, nzeta;                            /* ny: Fourier modes, or cells between the walls (free-hand) */
    double Lx, Ly;
    int y_fourier;
    const double* zeta;                           /* nzeta slice positions, zeta[0] = 0 */
    const double* ynode;                          /* ny + 1 y nodes */
    double dz, z_end, dz_push;
    double sponge, sponge_y, sponge_rate;
    int move_x, move_y, window_lag;
    int self_fields, zeta_filter, k_filter;
    double push_phase;
    /* analytic plane-wave laser (src/sim.hpp:191-210) */
    int laser_on, laser_pol;                      /* pol: 0 s, 1 p, 2 circular */
    double a0, kL, theta, tau, cep, start_fwhm, stop_fwhm;
    /* the bunch (first bunch block) */
    double charge, gamma, sigx, sigy, sigz, emitx, emity, espread, zeta_center;
    long nmacro;
    unsigned long seed;
    /* distribution_z = function (src/sim_bunches.cpp:115-132, 160): profile_z the piecewise-linear function through
     * prof[k] at zeta = prof_zmin + k (prof_zmax - prof_zmin) / (n_prof - 1), k < n_prof (0: Gaussian in zeta of
     * sigz), sampled by the inverse of its cumulative distribution; zeta_center is then the middle of the range */
    int n_prof;
    double prof_zmin, prof_zmax, prof[16];
    /* further bunches (species -1, -2, -3): their co-moving start fields are solved each with its own beta and
     * summed (src/sim_bunches.cpp:218-248) */
    int n_bunch_more;
    g3d_bunch bunch_more[3];
    /* grid lasers */
    int n_glasers;
    g3d_glaser glasers[4];
    /* plasma */
    int plasma_on;
    double n0, r_max;
    int ppc_x, ppc_y;
    double smooth_length;
    int n_pz;
    double pz_points[16], pz_values[16];
    int parabolic;
    double channel_depth, r_channel;
    int iterations;
    double vz_min, vz_beam, trap_advance;
    int predictor_extrapolate, free_electrons;
    int n_ions;
    g3d_ion ions[4];
    int freeze_bunch, start_relax, box_mean;
    int mixed;          /* plasma pass in the FP32-mixed form (no counterpart in the reference: 0 = FP64 as the
                           reference) */
} g3d_config;

/* context of the per-mode march (src/sim_kernels.hpp:161-167 MarchCtx), device pointers */
typedef struct {
    g3d_cpx *Rx, *Ry, *Ez, *Bz, *Lx, *Ly;         /* ns planes each */
    const g3d_cpx *bnd;                           /* 6 planes: slice b from above (unused on one GPU: b = N) */
    const g3d_cpx *j0x, *j0y, *jkx, *jky, *jkz;   /* ns + 1 half-cell planes each */
    g3d_cpx *bkRx, *bkRy, *bkEz, *bkBz;           /* backup of slice n (march_slice with passes), may be null */
    const double *kx, *ky, *iD, *iA, *hcg, *wlo, *whi;   /* ky = NULL: free-hand y (k_x lines between walls) */
    double D;
    int a, b, N, nx, ny;
    long nk;
    double *scp;            /* free-hand y: 2 nk doubles of device scratch (the Thomas factors of a line; the
                               reference's sRx ... scp lines); unused for Fourier y */
} g3d_march_ctx;

/* ------------------------------------------------------------------ runtime and memory (src/portable.hpp) */
int g3d_init(int device);
const char* g3d_last_error(void);
const char* g3d_status_string(int code);
int g3d_sync(void);
int g3d_malloc(void** p, size_t bytes);           /* zero-filled, as Buf<T>::resize (portable.hpp:113-122) */
int g3d_free(void* p);
int g3d_memset0(void* p, size_t bytes);
int g3d_memcpy_h2d(void* dst, const void* src, size_t bytes);
int g3d_memcpy_d2h(void* dst, const void* src, size_t bytes);
int g3d_memcpy_d2d(void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------ batched FFT (src/fftmany.hpp:26-40) */
typedef struct g3d_fft g3d_fft;
/* FftMany(nx, ny, two_d); nx, ny powers of two in 2 ... 1024 */
int g3d_fft_plan(int nx, int ny, int two_d, g3d_fft** plan);
/* FftMany::run: in place over count planes; forward sign -1, inverse scaled by 1/(nx ny) (2D) or 1/nx (rows) */
int g3d_fft_exec(g3d_fft* plan, g3d_cpx* data, int count, int inverse);
int g3d_fft_destroy(g3d_fft* plan);

/* ------------------------------------------------------------------ march (src/sim_kernels.hpp:173-313) */
/* Simulation::march (src/sim_fields.cpp:7-39): every mode (Fourier y) or k_x line (free-hand y) from slice
 * hi-1 down to lo */
int g3d_march(const g3d_march_ctx* c, int lo, int hi);
/* Simulation::march_slice (src/sim_plasma.cpp:189-197): one slice n.  With the bk* pointers set, restore=0
 * copies slice n of R_x, R_y, E_z, B_z into them before the update and restore=1 reads slice n from them
 * (the two-pass backup of src/sim_plasma.cpp:306-326 without the extra copy kernels). */
int g3d_march_slice(const g3d_march_ctx* c, int n, int restore);

/* ------------------------------------------------------------------ adjoint of the march (no counterpart in the
 * reference: the reverse mode of g3d_march for gradients through the march, Fourier y).  The march of slices
 * lo ... hi-1 is linear in the level-s0 slices (R_x, R_y, E_z, B_z), slice hi at level s1 and the currents; this
 * applies its transpose (for the real inner product Re sum conj(u) v) to the adjoints of its outputs. */
typedef struct {
    g3d_cpx *Rx, *Ry, *Ez, *Bz, *Lx, *Ly;         /* in: adjoints of the outputs of slices lo..hi-1 (ns planes,
                                                     plane n - a); out: adjoints of the level-s0 inputs R_x, R_y,
                                                     E_z, B_z of those slices (L_x, L_y: zero, not inputs) */
    g3d_cpx *up;                                  /* out: 6 planes R_x R_y E_z B_z L_x L_y, the adjoint of slice hi
                                                     (hi < N); may be null */
    g3d_cpx *j0x, *j0y, *jkx, *jky, *jkz;         /* out, added to: adjoints of the currents (ns + 1 planes) */
    const double *kx, *ky, *hcg, *wlo, *whi;      /* as g3d_march_ctx */
    double D;
    int a, N, nx, ny;
    long nk;
} g3d_march_adj_ctx;
int g3d_march_adjoint(const g3d_march_adj_ctx* c, int lo, int hi);

/* Adjoint pairs of the particle-grid couplings (no counterpart in the reference; the pieces of the reverse march
 * for fixed trajectories).  TSC on the periodic Fourier-y grid, nodes at (i dx, j dy), weights of
 * src/sim_kernels.hpp:21-31: g3d_tsc_deposit adds q_p W_p to the plane, g3d_tsc_gather returns sum_k W_p(k) P[k]
 * (its transpose) and, when gx, gy are set, the derivatives of that sum in x_p and y_p. */
typedef struct {
    const double *x, *y;          /* box coordinates of np particles (device) */
    long np;
    int nx, ny;
    double dx, dy;
} g3d_tsc_ctx;
int g3d_tsc_deposit(const g3d_tsc_ctx* c, const double* q, double* plane);
int g3d_tsc_gather(const g3d_tsc_ctx* c, const double* plane, double* out, double* gx, double* gy);

/* The relativistic Boris rotation of the streaming push (src/sim_stream.hpp:8-18), in place on u, per particle
 * fields E, B, charge-to-mass qm and step dt; and its vector-Jacobian product: given the adjoint of the pushed u
 * (ub*), the adjoints of u, E, B, qm and dt at the state before the push (u* must still hold that state). */
typedef struct {
    double *ux, *uy, *uz;
    const double *Ex, *Ey, *Ez, *Bx, *By, *Bz, *qm, *dt;
    long np;
} g3d_boris_args;
typedef struct {
    const double *ubx, *uby, *ubz;                 /* in: adjoint of the pushed momentum */
    double *gux, *guy, *guz, *gEx, *gEy, *gEz, *gBx, *gBy, *gBz, *gqm, *gdt;   /* out */
} g3d_boris_adj;
int g3d_boris(const g3d_boris_args* a);
int g3d_boris_vjp(const g3d_boris_args* a, const g3d_boris_adj* g);

/* One streaming electron through half-cell n (from slice n at zn towards slice n + 1 at zn1) on a periodic Fourier-y
 * grid: the push of src/sim_stream.hpp:104-171 for an electron with no ionization and no analytic laser (the gather
 * at the midpoint, the two Boris rotations, the trapping branch, the TSC deposit of the current and, for a trapped
 * electron, of its beam current), with the fields of the two slices given in FP64.  g3d_stream_push runs it in
 * place on the state and adds the deposit to J; g3d_stream_push_vjp is its reverse: given the adjoints of the
 * pushed state and of J, the adjoints of the state before the push (the state arrays must still hold it) and of the
 * two field planes (added to).  The branch decisions (alive, trapped, the sign of the charge) are those of the
 * forward push and are held fixed.  The reverse has no counterpart in the reference. */
typedef struct {
    int nx, ny;
    double dx, dy;                 /* periodic grid: Lx = nx dx, Ly = ny dy; node (ix, j) at (ix dx, j dy) */
    double zeta0, ihz;             /* uniform ζ axis: slice k at zeta0 + k / ihz */
    int n;                         /* the half-cell */
    double zn, zn1;                /* ζ of slices n and n + 1 */
    double Delta, s1, sTrap, vz_beam, ia;   /* step length, end of the step, trapping limit, beam threshold,
                                               inverse cell-area factor of the deposit */
    const double* F0;              /* E_x E_y E_z B_x B_y B_z of slice n, [6][ny][nx] */
    const double* F1;              /* the same of slice n + 1 */
    double *x, *y, *zeta, *z, *px, *py, *pz;
    const double *q, *qm, *w;
    int* alive;                    /* 1: streaming (pushed); 5: trapped by this push; other codes are left alone */
    double* J;                     /* j_x j_y j_z rho- rho+, [5][ny][nx], added to (forward) */
    long np;
} g3d_spush_ctx;
typedef struct {
    const double *bx, *by, *bzeta, *bz, *bpx, *bpy, *bpz;   /* in: adjoints of the pushed state */
    const double* bJ;                                       /* in: adjoint of J, [5][ny][nx] */
    double *gx, *gy, *gzeta, *gz, *gpx, *gpy, *gpz;         /* out: adjoints of the state before the push */
    double *gF0, *gF1;                                      /* out: adjoints of F0, F1 (added to) */
} g3d_spush_adj;
int g3d_stream_push(const g3d_spush_ctx* c);
int g3d_stream_push_vjp(const g3d_spush_ctx* c, const g3d_spush_adj* g);

/* The RK4 push in z of beam particles from s0 (alive 1; alive 3: from s1 - rem) to s1 (src/sim_kernels.hpp:75-149,
 * src/sim_particles.cpp:6-24): the adaptive sub-steps, the gather at every stage from the slices of a uniform ζ
 * axis on a periodic Fourier-y grid (fields in FP64), the analytic plane-wave laser, the hand-over, slow and lost
 * particle codes.  g3d_bunch_push_piece runs it in place; g3d_bunch_push_vjp is its reverse: given the adjoints of
 * the pushed (x, y, ζ, p_x, p_y, p_z), the adjoints of the state before the push (the arrays must still hold it)
 * and of the field planes (added to).  The sub-step lengths and every branch are those of the forward push and
 * are held fixed; a particle whose push fails keeps its state (gradient: the identity).  The reverse has no
 * counterpart in the reference. */
typedef struct {
    int nx, ny, nsl;
    double dx, dy, zeta0, ihz;     /* periodic grid; slice k (k = 0 .. nsl-1) at zeta0 + k / ihz */
    const double* F;               /* E_x E_y E_z B_x B_y B_z, [6][nsl][ny][nx] */
    int laser_on, laser_pol;       /* the plane wave of g3d_config (code units) */
    double a0, kL, tau, cep, eta_c, kx, kz;
    double X0, zc;                 /* window offset in x; ζ of the laser centre */
    double s0, s1, dz_push, push_phase, hmin, zhand, zmax, ylo, yhi, vz_min;
    double *x, *y, *zeta, *px, *py, *pz, *z, *rem;
    const double* qm;
    int* alive;
    long np;
} g3d_bpush_ctx;
typedef struct {
    const double *bx, *by, *bzeta, *bpx, *bpy, *bpz;       /* in: adjoints of the pushed state */
    double *gx, *gy, *gzeta, *gpx, *gpy, *gpz;             /* out: adjoints of the state before the push */
    double* gF;                                            /* out: adjoint of F (added to) */
} g3d_bpush_adj;
int g3d_bunch_push_piece(const g3d_bpush_ctx* c);
int g3d_bunch_push_vjp(const g3d_bpush_ctx* c, const g3d_bpush_adj* g);

/* ------------------------------------------------------------------ kernel level: the device contexts of the
 * reference (src/sim_kernels.hpp:10-17 PushCtx, src/sim_stream.hpp:20-29 StreamCtx) and the member functions of the
 * drop-in files over caller-owned device buffers in the reference's layouts.  Fourier y unless stated; free-hand y
 * runs through the engine (g3d_sim_*). */

/* a grid axis (src/grid.hpp:20-46 Axis): uniform nodes x0 + i / ih (uniform = 1), or the node table X[0 .. n-1]
 * on the device; outside the table the end cells are continued */
typedef struct {
    const double* X;
    int n, uniform;
    double x0, ih;
} g3d_axis;

/* the analytic plane-wave laser in code units (src/sim.hpp:191-210 Laser): pol 0 s, 1 p, 2 circular; eta_c the
 * centre of the envelope; (kx, kz) the unit propagation vector */
typedef struct {
    int on, pol;
    double a0, kL, tau, cep, eta_c, kx, kz;
} g3d_plane_wave;

/* PushCtx (src/sim_kernels.hpp:10-17) without the external elements (undulators, magnets) */
typedef struct {
    const float* F;               /* real fields [6][ns+2][nk] */
    int a, ns, N, nx, ny, self, yf;
    long nk;
    double dx, dy, Lx, Ly;
    g3d_axis zax, yax;
    g3d_plane_wave laser;
    double X0, Y0, zc, dz_push, push_phase, zhand, zmax, ylo, yhi, s1, vz_min, hmin;
    int freeze;
} g3d_push_ctx;

/* Simulation::push (src/sim_particles.cpp:6-24): particles [i0, i1) of the beam list to the plane s1, those with
 * alive = 1 from s0, adopted ones (alive = 3) from s1 - rem; RK4 in z with the gathers of src/sim_kernels.hpp:75-149 */
int g3d_bunch_push(const g3d_push_ctx* c, const g3d_parts* p, long i0, long i1, double s0);

/* ADK level (src/ionization.hpp:25-30) and the device constants of an ion species (src/sim_ions.cpp:68-87) */
typedef struct { double logpre, power, expo, Epeak, Emin; } g3d_adk_level;
typedef struct { double mass; int mobile, ionizing, Zmax, off; double Emin; } g3d_ion_dev;

/* StreamCtx (src/sim_stream.hpp:20-29) */
typedef struct {
    g3d_push_ctx pc;              /* fields, laser, grids */
    const double* zs;             /* slice positions, N */
    double* J;                    /* jsr: j_x j_y j_z rho- rho+ of the half-cell, [5][nk] (added to) */
    const double *ia, *id;        /* 1/A_j, 1/Delta_j (Fourier y: 1/dy) */
    double Delta, s1, sTrap, vz_beam;
    int n;                        /* the half-cell */
    const g3d_ion_dev* ions;      /* ion species (null: no ions), their ADK levels */
    const g3d_adk_level* adk;
    g3d_parts born;               /* electrons freed in this pass at slots *bcnt (atomic), at most bcap */
    int* bcnt;
    int bcap;
    uint64_t key;                 /* ion_key_slice of the pass (src/sim_plasma.cpp:293,331) */
} g3d_stream_ctx;

/* the streaming loop of march_plasma (src/sim_plasma.cpp:334-337, stream_one of src/sim_stream.hpp:104-171): every
 * particle [0, nstrm) of src through half-cell c->n, the pushed state written to dst (dst may be src), the deposit
 * added to c->J with FP64 atomics (summation order as free as the reference's OpenMP atomics) */
int g3d_stream_halfcell(const g3d_stream_ctx* c, const g3d_parts* src, const g3d_parts* dst, long nstrm);

/* the spectral geometry of the constructor (src/sim_setup.cpp:5-60, 145-164) as device tables, for the member
 * functions below (Fourier y) */
typedef struct {
    int nx, ny, ns, N, a;
    long nk;
    const double *kx, *ky;        /* k vectors of the march, the Nyquist mode zeroed (nx, ny) */
    const double *kxs, *kys;      /* k vectors of the filters and phase shifts (nx, ny) */
    const double *filt, *filty;   /* cos^2 (k d/2) filters, Nyquist zeroed, all ones without k_filter (nx, ny) */
    const double *hcg, *wlo, *whi;   /* half-cell lengths (hcg[m+2] = h_m) and current weights of the slices */
    double a2;                    /* smooth_length^2 of the streaming current */
    int keep0;                    /* keep the k = 0 mode of the streaming and beam currents (box_mean) */
    g3d_fft* plan;                /* FftMany(nx, ny, two_d = 1) */
} g3d_spec_ctx;

/* Simulation::stream_spectra (src/sim_plasma.cpp:94-186): jsr [5][nk] -> jsn [3][nk] (j_x, j_y, j_z spectra with the
 * transverse filters and the smoothing); jsr is zeroed; scratch: 6 nk complex values */
int g3d_stream_spectra(const g3d_spec_ctx* c, double* jsr, g3d_cpx* jsn, g3d_cpx* scratch);

/* No counterpart in the reference: g3d_stream_spectra with its workspace in a file-scope device array instead of a
 * caller's buffer, for callers that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  It calls no allocating routine.  Limit: nk <= 65536 (256 x 256 modes), refused beyond; the
 * workspace is 6 MiB per device, held until the process ends.  Every call on one GPU shares that GPU's workspace: calls
 * that overlap on one GPU race. */
int g3d_stream_spectra_static(const g3d_spec_ctx* c, double* jsr, g3d_cpx* jsn);

/* Simulation::real_fields_slice (src/sim_plasma.cpp:200-220): the spectra of one slice (plane pointers of R_x R_y
 * L_x L_y E_z B_z) -> its six real field planes E_x E_y E_z B_x B_y B_z at F_slice + comp * fpl (float);
 * scratch: 3 nk complex values */
int g3d_real_fields_slice(const g3d_spec_ctx* c, const g3d_cpx* Rx, const g3d_cpx* Ry, const g3d_cpx* Lx,
                          const g3d_cpx* Ly, const g3d_cpx* Ez, const g3d_cpx* Bz, float* F_slice, long fpl,
                          g3d_cpx* scratch);

/* Simulation::currents (src/sim_particles.cpp:56-207): the beam deposit of np particles into jw [3][ns+3][nk]
 * (summed in the reference's single-thread order: the same doubles), the (1,2,1)/4 zeta filter, the spectra
 * jk{x,y,z} [ns+1][nk] with the transverse filter and the k = 0 rule (keep0: plasma with box_mean, not starting);
 * zfilter, zuniform: cfg.zeta_filter and a uniform zeta axis; ia, id as g3d_stream_ctx; zax: the zeta axis;
 * scratch: (ns+1) nk complex values.  The deposit's sort buffers are held by the library between calls. */
int g3d_bunch_currents(const g3d_spec_ctx* c, const g3d_axis* zax, const double* ia, const double* id, double dx,
                       double dy, int zfilter, int zuniform, const g3d_parts* p, long np, double* jw, g3d_cpx* jkx,
                       g3d_cpx* jky, g3d_cpx* jkz, g3d_cpx* scratch);

/* Simulation::sponge_and_real_fields (src/sim_fields.cpp:156-307, Fourier y): with sponge != 0 the bound field of
 * the beam (jkz, the beam's co-moving beta_b and gamma_b^-2 = gm2_b) is taken out of R, R is damped by the real-space
 * masks maskx [nx] and masky [ny] (sim_setup.cpp: 1 - rate (1 - d/w)^2 in the sponge layers) and the bound field
 * added back; then the real fields of all ns slices into F [6][ns+2][nk].  R_x, R_y: ns planes (rewritten); L_x,
 * L_y, E_z, B_z: ns planes; jkz: ns + 1; scratch: 5 ns nk complex values */
int g3d_sponge_real_fields(const g3d_spec_ctx* c, int sponge, const double* maskx, const double* masky,
                           double beta_b, double gm2_b, g3d_cpx* Rx, g3d_cpx* Ry, const g3d_cpx* Lx, const g3d_cpx* Ly,
                           const g3d_cpx* Ez, const g3d_cpx* Bz, const g3d_cpx* jkz, float* F, g3d_cpx* scratch);

/* Simulation::comoving_field (src/sim_fields.cpp:44-151, Fourier y): the start field of the beam's half-cell charge
 * jkz (ns + 1 planes) moving at beta (gm2 = 1 - beta^2; the reference's beta is the weighted mean v_z of the live
 * beam particles, computed on the host): the Thomas solve along zeta per mode of -(k^2 + gm2 d^2/dzeta^2) psi = rho
 * and R = (1 + beta) (-grad psi), E_z, B_z of the slices; scratch: 2 ns nk complex values */
int g3d_comoving(const g3d_spec_ctx* c, const g3d_cpx* jkz, double beta, double gm2, g3d_cpx* Rx, g3d_cpx* Ry,
                 g3d_cpx* Ez, g3d_cpx* Bz, g3d_cpx* scratch);

/* the window centroid of src/sim_setup.cpp:221-234: sums over the live bunch particles of w (x + X0), w (y + Y0)
 * and w into out[3] (host), in a fixed tree order; the partial sums are library scratch */
int g3d_bunch_centroid(const g3d_parts* p, long np, double X0, double Y0, double* out);

/* the window shift of src/sim_setup.cpp:248-280 by (nxs, nys) cells: the phase factor e^{i (k_x sx + k_y sy)} on
 * count planes of each spectral array in planes[0 .. nplanes) (R, E_z, B_z, j0, js0 ...), the index roll of F
 * [6][ns+2][nk] (scratch: F's size in floats), and the beam particles moved by (-sx, -sy) */
int g3d_window_shift(const g3d_spec_ctx* c, int nxs, int nys, double dx, double dy, g3d_cpx* const* planes,
                     const long* counts, int nplanes, float* F, float* F_scratch, const g3d_parts* p, long np);

/* ------------------------------------------------------------------ engine: the Simulation of the march
 * g3d_sim mirrors the reference's Simulation (src/sim.hpp:222-360) on one GPU: init() and step() and the
 * member functions of the five drop-in files, each over the kernels of this library. */
typedef struct g3d_sim g3d_sim;
typedef struct {
    long N, ns, nk, np, nstrm, istep, born_cap;
    double s, dx, dy, X0, Y0, beta_b, gm2_b;
    long trapped_total, restreamed_total, lost_total;
    /* pushes done since g3d_sim_create (no counterpart in the reference: the measure of the march's throughput):
     * streaming particles through one half-cell in one pass, counted on the device by the stream launches, and
     * beam particles advanced from s0 to s1 (the list at the start of each push) */
    long pushes_stream, pushes_beam;
} g3d_sim_info;

int g3d_sim_create(const g3d_config* cfg, g3d_sim** out);
int g3d_sim_destroy(g3d_sim* sim);
int g3d_sim_info_get(g3d_sim* sim, g3d_sim_info* info);
/* Simulation::init (src/sim_setup.cpp:178-197) and Simulation::step (199-217) */
int g3d_sim_init(g3d_sim* sim);
int g3d_sim_step(g3d_sim* sim);
/* a member function by name with numeric arguments:
 *   "march" D | "currents" | "sponge_and_real_fields" | "solve_fields" D | "stream_spectra" |
 *   "march_slice" n D | "real_fields_slice" n | "push" s0 s1 | "march_plasma" s0 s1 | "comoving_field" |
 *   "relax_start_field" | "add_grid_lasers" | "load_plasma" s0 s1 | "j0_from_jk" */
int g3d_sim_call(g3d_sim* sim, const char* op, const double* args, int nargs);
/* host copies of the device arrays by name (reference layouts): Rx Ry Ez Bz Lx Ly j0x j0y jkx jky jkz js0x
 * js0y js0z jw F jsr jsn rho_lines; "parts" and "strm" as g3d_particle records.  bytes must match the size. */
int g3d_sim_array_size(g3d_sim* sim, const char* name, size_t* bytes);
int g3d_sim_get(g3d_sim* sim, const char* name, void* host, size_t bytes);
int g3d_sim_set(g3d_sim* sim, const char* name, const void* host, size_t bytes);
/* set the beam list (Simulation::set_particles, src/sim_setup.cpp:171-175) / the joining list */
int g3d_sim_set_particles(g3d_sim* sim, const g3d_particle* p, long n);
int g3d_sim_set_joining(g3d_sim* sim, const g3d_particle* p, long n);
int g3d_sim_get_joining(g3d_sim* sim, g3d_particle* p, long cap, long* n);
/* output and checkpoints in the reference's formats, from the simulation's state after a step:
 * bunches.txt lines (Simulation::write_bunches, src/sim_bunch_diag.cpp:63-77) and diag.txt lines
 * (Simulation::diagnostics, src/sim_output.cpp:5-52) appended to the file at path, with the header line when
 * header != 0; slices_NNNNNN.txt (write_slices, src/sim_bunch_diag.cpp:79-97); the field energy in code units
 * (total_field_energy, src/sim_output.cpp:54-79); the particle dump (dump_particles, 81-97); the frame /
 * fields_NNNNNN.bin file of the planes through the axis (write_frame, 103-198; comps: bits ex ey ez bx by bz ne ni)
 * and axis_NNNNNN.txt (write_axis, 200-245) */
int g3d_sim_write_bunches(g3d_sim* sim, const char* path, int header);
int g3d_sim_write_slices(g3d_sim* sim, const char* path, double slice_length);
int g3d_sim_diagnostics(g3d_sim* sim, const char* path, int header);
int g3d_sim_field_energy(g3d_sim* sim, double* out);
int g3d_sim_dump_particles(g3d_sim* sim, const char* path);
int g3d_sim_write_frame(g3d_sim* sim, const char* path, int comps);
int g3d_sim_write_axis(g3d_sim* sim, const char* path);
/* the G3DCKPT1 checkpoint (src/sim_checkpoint.cpp:9-105): written as dir/checkpoint_NNNNNN_r0000.bin after a step
 * (write_checkpoint), read instead of init() (read_checkpoint); the same header and records as the reference's,
 * one rank */
int g3d_sim_write_checkpoint(g3d_sim* sim, const char* dir);
int g3d_sim_read_checkpoint(g3d_sim* sim, const char* dir, long step);
/* The gradient of the march (no counterpart in the reference; the reverse of Sim::step for gradients): on a new
 * simulation (not yet initialised), runs init() and `steps` steps, with a state checkpoint every `every` steps, and
 * returns the mean p_z of the live bunch particles at the end in *loss, its derivatives in the nenv envelope
 * values of the first grid laser in genv and in the nprof current-profile values of the bunch in gprof (either
 * count may be 0).  The reverse re-runs each segment from its checkpoint with a tape and applies the transposed
 * field updates and the dual-number reverses of the streaming and beam pushes, holding the forward's discrete
 * decisions fixed.  The simulation is left at the end of the steps.  Either form (the reverse in FP64), Fourier y,
 * uniform zeta grid, electrons only, one bunch, a fixed window; a run in which beam particles are re-streamed is
 * refused. */
int g3d_sim_gradient(g3d_sim* sim, int steps, int every, double* loss, double* genv, int nenv, double* gprof,
                     int nprof);
/* The tangent-linear run of the same march (no counterpart in the reference): on a new simulation, the loss of
 * g3d_sim_gradient and its derivative along the direction (denv, dprof) of the envelope and profile values, by
 * carrying the tangents of the state forward through init() and the steps beside the primal ones (the forward's
 * field kernels on the tangent arrays, the pushes and the beam deposit on dual numbers).  The same coverage as
 * g3d_sim_gradient; <gradient, direction> equals *dloss. */
int g3d_sim_tangent(g3d_sim* sim, int steps, const double* denv, int nenv, const double* dprof, int nprof,
                    double* loss, double* dloss);
/* HIP event timing of the kernel families: bit f of mask times family f (-1: all, 0: off); clears the sums.
 * g3d_sim_timers reads the accumulated milliseconds and launch counts with the family names. */
int g3d_sim_timing(g3d_sim* sim, int mask);
int g3d_sim_timers(g3d_sim* sim, double* ms, long* counts, int nfam, char* names, int name_len);
/* Time one launch in `every` (>= 1) of the streaming half-cell kernel, whose launches carry their own events, so
 * that timing perturbs the march little; the counts of g3d_sim_timers are the timed launches.  No counterpart in
 * the reference (a measurement hook). */
int g3d_sim_timing_every(g3d_sim* sim, int every);

#ifdef __cplusplus
}
#endif
#endif
