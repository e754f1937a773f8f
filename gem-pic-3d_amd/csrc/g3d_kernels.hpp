// Launch wrappers of the gfx950 kernels used by the engine (g3d_engine.cpp).  All launches go to the library
// stream; none synchronises.
#pragma once
#include <hip/hip_ext.h>
#include "g3d_common.hpp"
#include "g3d_fft.hpp"
#include "g3d_march.hpp"
#include "g3d_runtime.hpp"

namespace g3d {

using Parts = g3d_parts;

// ---------------------------------------------------------------- particles
// the push context of the reference (src/sim_kernels.hpp:10-17) without the external elements
struct PushArgs {
    FieldView fv;
    PlaneWave laser;
    double X0, Y0, zc, dz_push, push_phase, zhand, zmax, ylo, yhi, s1, vz_min, hmin;
    int freeze;
};
// the streaming context (src/sim_stream.hpp:20-29)
struct StreamArgs {
    FieldView fv;
    PlaneWave laser;
    double X0, Y0, zc, ylo, yhi;
    const double* zs;
    double* J;                       // jsr: jx jy jz rho- rho+ (5 nk)
    float* Jf;                       // FP32-mixed form: the current planes in float (same layout); null: J
    int skip_state;                  // a pass before the last of a half-cell: the particle state is not written (the
                                     // next pass restarts from the same start state and writes it)
    const double *ia, *id;
    double Delta, s1, sTrap, vz_beam;
    int n;
    const IonDev* ions;
    const AdkLevel* adk;
    Parts born;                      // electrons freed in this pass
    int* bcnt;
    int bcap;
    uint64_t key;
    // pushes of the launch: workgroup b adds the number of particles it pushed through the half-cell to npush[b]
    // (npush_cap slots; null: not counted)
    unsigned long long* npush;
    long npush_cap;
};

// Transverse tiles of the streaming list: the fresh plasma is loaded tile by tile (TX x TY cells), particles
// [off[t], off[t+1]) belong to tile t.  A tile's workgroup deposits into an LDS window of WX x WY nodes starting
// at node (tx TX - 1 - M, ty TY - 1 - M) and writes it to its slab (5 planes of WX x WY); deposits that leave the
// window go to jsr with global atomics.  Particles from off[ntile] on (joining, freed) are not tiled.
struct TileGeom {
    int TX, TY, M, WX, WY, ntx, nty;
    int fwin;                         // stage the tile's field window in LDS for the gather
    int block;                        // lanes of a tile's workgroup (a multiple of 64, at most kStreamBlockMax)
    int wpe;                          // FP32-mixed stream kernel: 4 = registers capped for 4 waves per SIMD, else 0
    // FP32-mixed stream kernel, dense tiles: a workgroup streams at most `chunk` particles (0: a whole tile); the
    // further chunks of a tile, xtile/xstart[0 .. *xcount), run in max_extra more workgroups after the tiles and
    // add their window to the current planes with atomics
    int chunk, max_extra;
    int aflush;                       // FP32-mixed: every window added to the current planes with atomics (no slabs)
    // FP32-mixed: tile workgroups XCD-contiguous (workgroup b runs on XCD b mod 8; with max_extra and ntile
    // multiples of 8, XCD x takes tiles x ntile/8 ... (x+1) ntile/8 - 1, so the overlapping field windows of
    // neighbouring tiles meet in one L2)
    int xcd;
    const int* xtile;
    const long* xstart;
    const int* xcount;
    // free-hand y: the tile rows follow the node table (cell j of a particle from fy; the last row of tiles may be
    // short), the windows stop at the walls, and the tile workgroups add their deposit window to the current
    // planes with FP64 atomics instead of slabs (no field window)
    int fh;
    Axis fy;
    long ntile, n_tiled;              // tiles; list entries binned (= off[ntile]); from n_tiled on: untiled tail
    const long* off;
    const int* idx;                   // list indices grouped by tile (null: the list itself is tile-ordered)
    double* slab;
};
// bin the live particles of the list by their current tile: off[0 .. ntile], idx[0 .. off[ntile]); the state of
// the other particles (final for the rest of the step) is copied from S to D.  cnt, cur: ntile ints of scratch
void bin_tiles(Parts S, Parts D, long n, const TileGeom& tg, double dx, double dy, int nx, int ny, int* cnt,
               int* cur, long* off, int* idx);
// the chunks beyond the first of every tile holding more than tg.chunk particles (see TileGeom); xtile, xstart hold
// tg.max_extra entries
void tile_chunks(const TileGeom& tg, int* xtile, long* xstart, int* xcount);

void bunch_push(const PushArgs& a, Parts p, long np, double s0);
// the same push with the particles binned by slice pair and transverse tile and the fields of each bin staged in LDS
// (Fourier y with self-fields; otherwise the plain push); the results are the plain push's, bit for bit
struct BpushScratch {
    DevBuf<unsigned long long> key0, key1;
    DevBuf<unsigned> idx0, idx1;
    DevBuf<unsigned char> tmp;
    DevBuf<long> chunk;
    DevBuf<int> nchunk;
};
void bunch_push_windowed(const PushArgs& a, Parts p, long np, double s0, BpushScratch& s);
// beam deposit into jw [3][ns+3][nk] (src/sim_particles.cpp:66-106)
struct DepositArgs {
    Axis zax, yax;
    const double *hcg, *ia, *id;
    double* J;
    double dx, dy;
    int nh, a, N, nx, ny, yf;
    long nk;
};
// scratch of the ordered deposit (records, their sorted copies, the sort's temporary storage)
struct DepositScratch {
    DevBuf<unsigned long long> key0, key1;
    DevBuf<unsigned> rec0, rec1;
    DevBuf<double> val, vs;          // the record values, and the same in sorted order
    DevBuf<unsigned char> tmp;
    DevBuf<long> runs;               // the long runs of equal keys: (first record, end) pairs
    DevBuf<int> nruns;
};
// the reference's single-thread summation order, bit for bit (g3d_deposit.hip)
void bunch_deposit(const DepositArgs& a, Parts p, long np, DepositScratch& s);
// one half-cell pass of the streaming particles: src -> dst (the state of every particle is written); tiles
// with tg.ntile > 0, otherwise every deposit is a global atomic
// start and stop events of one launch, recorded by the launch itself (no marker packets around it); null: untimed
struct LaunchTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};
// a kernel launch on the library stream, through hipExtLaunchKernel with the timer's events when it has them
template <typename... KArgs, typename... Args>
void launch_timed_kernel(const LaunchTimer& lt, void (*k)(KArgs...), dim3 grid, dim3 block, unsigned lds,
                         Args... args) {
    if (lt.start) hipExtLaunchKernelGGL(k, grid, block, lds, stream(), lt.start, lt.stop, 0, args...);
    else hipLaunchKernelGGL(k, grid, block, lds, stream(), args...);
}
constexpr int kStreamBlockMax = 512;
// a list entry is pushed through half-cell n (from slice n + 1 towards slice n) when it streams (alive 1) or when it
// waits (alive 6) above slice n (src/sim_stream.hpp:107-111)
__device__ __forceinline__ bool pushed_through(int alive, double zeta, double zn) {
    return alive == 1 || (alive == 6 && zeta > zn);
}
// the workgroup's pushed particles (cnt per lane) added to npush[blockIdx.x]; every lane of the workgroup calls it
// once (it has a barrier), at most kStreamBlockMax lanes
__device__ __forceinline__ void add_pushes(unsigned long long* npush, unsigned cnt) {
    __shared__ unsigned part[kStreamBlockMax / 64];
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (unsigned w = 0; w < (blockDim.x + 63) / 64; ++w) t += part[w];
        npush[blockIdx.x] += t;
    }
}
void stream_halfcell(const StreamArgs& a, Parts src, Parts dst, long nstrm, const TileGeom& tg,
                     const LaunchTimer& lt = {});
// append the freed electrons of the last pass (count on the device) to both buffers; nstrm_dev += count
void append_born(Parts born, const int* bcnt, Parts p0, Parts p1, long base, long cap);
// window centroid sums of the bunch particles (weight, w x, w y), deterministic tree order
void bunch_centroid(Parts p, long np, double X0, double Y0, double* out3);
// copy the particles with alive == code into a compact AoS list (slot order is arbitrary; idx holds the source
// index for a stable host-side ordering); *count on the device
void collect_alive(Parts p, long n, int code, g3d_particle* out, long* idx, int* count, long cap);
void count_alive(Parts p, long n, int code, int* count);
// AoS <-> SoA
void parts_from_aos(const g3d_particle* in, Parts p, long n);
void parts_to_aos(Parts p, g3d_particle* out, long n);
void parts_copy(Parts src, Parts dst, long n);
// The fresh plasma of a step on the device (src/sim_plasma.cpp:43-90 for electrons only, Fourier y, every lattice
// point loaded): lattice points in tile order (tiles of TX x TY cells, per_tile = TX TY px py points each; inside a
// tile the sub-position (b2, a2) outermost, then the cells j-major, so that the 64 lanes of a wave push electrons
// of 64 different cells and their deposits into the tile's LDS window meet no same-address atomics: the stream
// kernel runs 8 % faster than with the points of a cell adjacent), each with the host loader's values and its id in
// the reference's order, id0 + ((j nx + i) py + b2) px + a2.  The density is n0 f g with the longitudinal factor f of the step and the
// channel factor g.
struct LatticeArgs {
    int nx, ny, px, py, TX, TY, ntx, parabolic;
    long per_tile;
    double dx, dy, X0, Y0, zf, s0, Delta, n0, f, channel_depth, r_channel;
    unsigned long long id0;
};
void plasma_lattice(const LatticeArgs& L, Parts A, Parts B, long n);
void parts_shift(Parts p, long n, double sx, double sy);

// ---------------------------------------------------------------- spectra and fields (Fourier y)
struct SpecGeom {
    int nx, ny, ns, N, a;
    long nk;
    const double *kx, *ky, *kxs, *kys, *filt, *filty;
    int cm = 0;      // sponge and real fields: R, L, E_z, B_z and jk column-major (mode ix ny + j); scratch row-major
};
// currents: zeta filter of jw into S1 (jx + i jy) and jkz (jz + 0 i), ns+1 half-cell planes
void currents_zfilter(const SpecGeom& g, const double* jw, const double* hcg, int zfilter, int zuniform, cx* S1, cx* jkz);
// unpack the pair, transverse filter, k = 0 rule
void currents_unpack(const SpecGeom& g, const cx* S1, cx* jkx, cx* jky, cx* jkz, int keep0);
// sponge: bound field psi (S2) and R - R_b packed (S1) over ns slices
void sponge_pack(const SpecGeom& g, const cx* Rx, const cx* Ry, const cx* jkz, const double* hcg, const double* wlo,
                 const double* whi, double g2, double opb, cx* S1, cx* S2);
void sponge_mask(const SpecGeom& g, cx* S1, const double* mx, const double* mn, const double* mh, int count);
// unpack the sponged R and add R_b (when S1 != null), then pack R, L, E_z/B_z for the inverse transforms (into
// Tf as complex float instead when Tf != null: the FP32 transforms of the FP32-mixed form, column-major modes)
void fields_pack3(const SpecGeom& g, const cx* S1, const cx* S2, double opb, cx* Rx, cx* Ry, const cx* Lx,
                  const cx* Ly, const cx* Ez, const cx* Bz, cx* T, long tstride, int count, cxf* Tf = nullptr);
// real fields from the three inverse transforms into F planes fl0 ... fl0+count-1 (FP64 or FP32 planes)
void fields_to_F(const cx* T, long tstride, float* F, long fpl, long nk, int count);
void fields_to_F_f(const cxf* T, long tstride, float* F, long fpl, long nk, int count);
// streaming current of one half-cell: the sum of jsr and of the tile slabs covering each point (tg.ntile > 0),
// packed as (jx + i jy, jz); the density lines (when rl != null); jsr is zeroed for the next pass
void sspec_pack(double* J, long nk, int nx, int ny, cx* pk, double* rl, double* rli, int i0, int j0,
                const TileGeom& tg);
// unpack + transverse filter + smoothing of the streaming current spectra (src/sim_plasma.cpp:105-115)
void sspec_unpack(const SpecGeom& g, const cx* pk, cx* jsn, double a2, int keep0);

// the per-mode work of one plasma half-cell pass (see g3d_plasma.hip)
struct ModeArgs {
    MarchView mv;
    const cx* pk;                     // spectra of the streaming current (2 planes), null at the front slice
    cx *jsn, *sin_, *si0, *pin, *dl;  // 3 planes each
    cx *js0x, *js0y, *js0z;
    cx *jkx, *jky, *jkz;              // writable views of the march's jk
    cx *bRx, *bRy, *bEz, *bBz;        // several passes: set (the free-hand line kernels keep slice n at s0 here)
    cx* T;                            // 3 packed planes for the inverse transforms
    const double *kxs, *kys, *filt, *filty;
    double wex, qcorr, a2;
    int n, first, last, top, extrap, keep0, ny, kf;
};
// the fused kernels of a plasma pass (g3d_plasma.hip): rows forward (jsr + slabs -> X, 3 planes), columns
// (kind 0 pass, 1 front slice, 2 spectra only into jsn; X -> T, 3 planes), rows inverse (T -> F planes of a slice)
void pass_rows_fwd(double* J, const TileGeom& tg, const FftPlan& p, cx* X, double* rl, double* rli, int i0, int j0,
                   int tr = 0);
void pass_cols(const ModeArgs& m, const FftPlan& p, const cx* X, cx* T, cx* jsn, int kind);
void pass_rows_inv(const cx* T, float* Fn, long fpl, const FftPlan& p, int tr = 0);
// tr = 1: X written and T read column-major (mode ix ny + j), for pass_colmodes: the y transforms, the per-mode
// work and the inverse y transforms in one launch, one workgroup per k_x column (needs m.mv.cm, 2 <= ny <= 1024)
void pass_colmodes(const ModeArgs& m, const FftPlan& p, const cx* X, cx* T, int kind);
int xcd_rows();                     // the row blocks of one XCD consecutive in the column-major row kernels
int rows_per_wg(int ny);            // rows per workgroup of the row kernels
// the FP32-mixed plasma pass (g3d_mixed.hip): the streaming half-cell with FP32 deposit (jsr and slabs float), and
// the field pass of slice m.n (rows forward, y transforms, per-mode work, inverse transforms, F); X, T: 3 planes
// of cxf scratch; kind as pass_cols
// rho: deposit the charge density planes too (the density lines read the last pass's)
void stream_mixed(const StreamArgs& a, float* J, float* slab, Parts src, Parts dst, long nstrm, const TileGeom& tg,
                  const LaunchTimer& lt = {}, bool rho = true);
void pass_mixed(const ModeArgs& m, const FftPlan& p, float* J, const float* slab, const TileGeom& tg, cxf* X, cxf* T,
                float* Fn, long fpl, double* rl, double* rli, int i0, int j0, int kind, int rpw);
void rows_fwd_mixed_fh(const FftPlan& p, float* J, cx* X, double* rl, double* rli, int i0, int j0, int rpw);
// the same columns step as three launches: y transforms of X (in place), the per-mode work (one lane per mode),
// inverse y transforms of T
void pass_modes(const ModeArgs& m, const FftPlan& p, cx* X, cx* T, cx* jsn, int kind);
// jz -= js0z over count planes
void sub_planes(cx* a, const cx* b, long n);
void copy_planes(cx* dst, const cx* src, long n);
// dst[p][c][r] = src[p][r][c] for `planes` planes of rows x cols (out of place)
void transpose_planes(const cx* src, cx* dst, long planes, int rows, int cols);
void zero_planes(cx* a, long n);

// grid laser (src/sim_fields.cpp:321-374, Fourier y): R_x + i R_y = 2 E_perp of the paraxial Gaussian of every
// slice into T (count planes), then the engine transforms and adds
struct GridLaserArgs {
    double amp, waist, zR, zfoc, zc, tau, kL, cep, rc, rs, dx, dy, X0, Y0;
    int circular, helicity, nx, ny;
    long nk;
    const double* zeta;
    const double* yn;                 // free-hand y: the node table (null: Fourier y)
    const double* env;                // envelope_function: a(zeta) per slice (null: the Gaussian of tau)
};
void grid_laser_fill(const GridLaserArgs& a, cx* T, int n0, int count);
void grid_laser_add(const SpecGeom& g, const cx* T, cx* Rx, cx* Ry, int count, int fh);
// phase rotation e^{i(k_x sx + k_y sy)} of count planes (window shift, src/sim_setup.cpp:259-267)
void rotate_planes(cx* A, long n, long nk, int nx, const double* kxs, const double* kys, double sx, double sy, int ys);
void roll_F(const float* F, float* out, long n, long nk, int nx, int ny, int nxs, int nys);
// comoving field: Thomas solve along zeta per mode (src/sim_fields.cpp:81-114, Fourier y)
void comoving_solve(const SpecGeom& g, const cx* jkz, const double* hcg, const double* wlo, const double* whi,
                    double beta, double gm2, cx* cp, cx* dp);
void comoving_R(const SpecGeom& g, const cx* psi, double beta, cx* Rx, cx* Ry, cx* Ez, cx* Bz);
void axpy_planes(cx* y, const cx* x, double s, long n);    // y = y + s x (s = +-1)

// ---------------------------------------------------------------- free-hand y (g3d_freehand.hip)
// A k_x line of the free-hand y grid: element j of line ik sits at j nx + ik of a plane.  R_x, L_x, E_z, j_x, j_z
// live on the nodes (j = 0 is the wall), R_y, L_y, B_z, j_y on the half-nodes.
struct LineView {
    MarchView mv;
    const double *iD, *iA;            // 1/Delta_j (half-nodes), 1/A_j (nodes)
    double* scp;                      // 2 nk: Thomas factors of the E_z and B_z systems (march, smoothing)
    int ny;
    // null, or the march's pivots and elimination factors, [N][4][ny][nx]: E_z pivot, E_z factor, B_z pivot, B_z
    // factor (they depend on the line, the y grid, h of the slice and D only; line_pivots)
    const double* piv;
};
// the pivots and elimination factors of the march's eliminations for every line and slice 0 .. N-2 (L.piv unused)
void line_pivots(const LineView& L, double* piv);
// slices hi-1 ... lo of every k_x line (src/sim_fields.cpp:7-31, src/sim_kernels.hpp:213-313)
void march_lines(const LineView& L, int lo, int hi);
// the pair of x transforms (S1: j_x + i j_y per half-cell plane) unpacked per row, the (1,2,1)/4 y filters of the
// charges, the x filter, the node-0 rule and the mean of j_y at k_x = 0 (src/sim_particles.cpp:160-204);
// jkz holds the x transforms of j_z and is filtered in place
void fh_currents(const cx* S1, cx* jkx, cx* jky, cx* jkz, int planes, const LineView& L, const double* filt, int kf,
                 int keep0, double Ly);
// sponge of the free-hand grid (src/sim_fields.cpp:202-268): bound potential psi per line into S2 (LU factors of
// k_x^2 - Lap_y: pcp, pid), R - R_b packed into S1; then (after the masks) the sponged R with R_b added back
struct FhSpongeArgs {
    const cx* jz;
    const double *kx, *pc, *pd, *hc, *wl, *wu;
    double g2, opb;
    int ns, N, a;
};
void fh_sponge_psi(const FhSpongeArgs& s, const LineView& L, const cx* Rx, const cx* Ry, cx* S1, cx* S2);
void fh_sponge_unpack(const FhSpongeArgs& s, const LineView& L, const cx* S1, const cx* S2, cx* Rx, cx* Ry);
// co-moving field (src/sim_fields.cpp:55-69, 122-147): node values of the half-cell charge to the coefficients of
// the eigenvectors of -Lap_y (evf), back to the nodes (evb), and R of -grad psi on the free-hand grid
void fh_eigen_forward(const cx* jz, const double* evf, cx* e1, int planes, int nx, int ny, long nk);
void fh_eigen_back(const cx* c, const double* evb, cx* out, int planes, int nx, int ny, long nk);
void fh_comoving_R(const cx* psi, const double* kx, const double* iD, double beta, cx* Rx, cx* Ry, cx* Ez, cx* Bz,
                   int planes, int nx, int ny, long nk);
// the co-moving solve along zeta of every (k_x, eigenvalue) pair; ev: the eigenvalue per slot, slot 0 unused
void comoving_solve_ev(const SpecGeom& g, const cx* e1, const double* ev, const double* hcg, const double* wlo,
                       const double* whi, double beta, double gm2, cx* cp, cx* dp);
// one pass of a plasma half-cell per k_x line (kind 0), the front slice (kind 1) or the spectra only into jsn
// (kind 2); X: the per-row unpacked x transforms of the streaming current (from pass_rows_fwd), T: 3 packed planes
// for pass_rows_inv
void pass_lines(const ModeArgs& m, const LineView& L, cx* X, cx* T, cx* jsn, int kind);

}  // namespace g3d
