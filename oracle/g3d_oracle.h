/* g3d_oracle.h - the CPU restatement of the per-slice march (test infrastructure only).
 *
 * Plain C, one thread, deterministic: every function restates one function of the reference (cited file:line) in
 * the reference's operand order.  Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it, as the
 * checker of the MI355X library; nothing in the library links it.
 */
#ifndef G3D_ORACLE_H
#define G3D_ORACLE_H
#include <stdint.h>
#include "../include/g3d.h"

typedef struct { const double* X; int n, uniform; double x0, ih; } or_axis;
typedef struct {
    const float* F;
    int a, ns, nx, ny, yf, self;
    long nk;
    double dx, dy, Lx, Ly;
    or_axis zax, yax;
} or_fields;
typedef struct { int on, pol; double a0, kL, tau, cep, eta_c, kx, kz; } or_laser;
typedef struct {
    or_fields f;
    or_laser laser;
    double X0, Y0, zc, dz_push, push_phase, zhand, zmax, ylo, yhi, s1, vz_min, hmin;
    int freeze;
} or_push_ctx;
typedef struct { double mass; int mobile, ionizing, Zmax, off; double Emin; } or_ion;
typedef struct { double logpre, power, expo, Epeak, Emin; } or_adk;
typedef struct {
    or_fields f;
    or_laser laser;
    double X0, Y0, zc, ylo, yhi;
    const double* zs;
    double* J;                  /* jx jy jz rho- rho+ (5 nk) */
    const double *ia, *id;
    double Delta, s1, sTrap, vz_beam;
    int n;
    const or_ion* ions;
    const or_adk* adk;
    g3d_particle* born;
    int* bcnt;
    int bcap;
    uint64_t key;
} or_stream_ctx;
typedef struct {
    g3d_cpx *Rx, *Ry, *Ez, *Bz, *Lx, *Ly;
    const g3d_cpx *j0x, *j0y, *jkx, *jky, *jkz;
    const double *kx, *ky, *iD, *iA, *hcg, *wlo, *whi;
    double D;
    int a, N, nx, ny;
    long nk;
} or_march_ctx;

/* FFT (src/fft.hpp:30-33, 74-85 radix 2; src/fftmany.hpp:95-118 rows then columns, inverse scaled) */
void or_fft1(g3d_cpx* a, int n, int inverse);
void or_fft_many(int nx, int ny, int two_d, g3d_cpx* data, int count, int inverse);
/* grids (src/sim_kernels.hpp:21-40, src/grid.hpp:35-46) */
void or_tsc_index(double xi, int idx[3], double w[3]);
void or_axis_locate(const or_axis* ax, double x, int* i, double* f);
/* fields at a particle (src/sim_kernels.hpp:44-72; src/sim.hpp:196-209) */
void or_gather(const or_fields* c, double x, double y, double zeta, double E[3], double B[3]);
void or_laser_field(const or_laser* L, double x, double z, double t, double E[3], double B[3]);
/* pushers (src/sim_stream.hpp:8-18; src/sim_kernels.hpp:75-149) */
void or_boris(double u[3], const double E[3], const double B[3], double qm, double dt);
void or_bunch_push(const or_push_ctx* c, g3d_particle* p, long n, double s0);
/* beam deposit into jw [3][ns+3][nk] (src/sim_particles.cpp:66-106, Fourier y) */
void or_bunch_deposit(const g3d_particle* p, long n, const or_axis* zax, const double* hcg, const double* ia, double dx,
                      double dy, int nx, int ny, int a, int ns, int N, double* J);
/* one streaming half-cell over a list, in list order (src/sim_stream.hpp:33-171) */
void or_stream_halfcell(const or_stream_ctx* c, g3d_particle* p, long n);
/* ADK and the counter-based draws (src/ionization.hpp:25-46, src/ionization.cpp:9-53) */
double or_adk_rate(const or_adk* L, double E);
void or_adk_levels(const double* ip_eV, int n_ip, int Z0, int Zmax, double E0, double omega0, double T, or_adk* out);
uint64_t or_mix64(uint64_t z);
double or_uniform01(uint64_t key);
uint64_t or_key_step(uint64_t seed, long step);
uint64_t or_key_slice(uint64_t ks, int n, int pass);
uint64_t or_key_particle(uint64_t kl, uint64_t id);
/* march (src/sim_kernels.hpp:173-210), slices hi-1 ... lo, Fourier y */
/* the march of slices hi-1 .. lo: per mode (Fourier y) or, with ky = NULL, per k_x line of the free-hand y grid */
void or_march(const or_march_ctx* c, int lo, int hi);
/* spectra of a streaming half-cell current (src/sim_plasma.cpp:94-117, Fourier y): J (5 nk) -> jsn (3 nk) */
void or_stream_spectra(int nx, int ny, const double* J, const double* kxs, const double* kys, const double* filt,
                       const double* filty, double a2, int keep0, g3d_cpx* jsn);
/* real fields of count slices (src/sim_fields.cpp:270-293, src/sim_plasma.cpp:200-220): F planes with stride fpl */
void or_real_fields(int nx, int ny, int count, const g3d_cpx* Rx, const g3d_cpx* Ry, const g3d_cpx* Lx, const g3d_cpx* Ly,
                    const g3d_cpx* Ez, const g3d_cpx* Bz, float* F, long fpl);
#endif
