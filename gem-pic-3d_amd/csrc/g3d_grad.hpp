// The reverse of a march step for gradients (the adjoint of Sim::step on the Fourier-y grid, FP64 form): the
// device pieces used by Sim::gradient (g3d_grad.cpp).  No counterpart in the reference.
#pragma once
#include "g3d_kernels.hpp"

namespace g3d {

// Adjoints of the per-mode state of the plasma march (complex, the reference's layout j nx + ix): the spectral
// fields (ns planes), the currents jk, j0 and js0 (ns + 1 planes) and the bookkeeping planes of a march
// (3 planes each).  lU: the adjoint of the filtered streaming current of a pass, times its filter.
struct PassAdj {
    cx *lRx, *lRy, *lEz, *lBz, *lLx, *lLy;
    cx *ljkx, *ljky, *ljkz, *lj0x, *lj0y, *ljs0x, *ljs0y, *ljs0z;
    cx *lpin, *lsin, *lsi0, *ldl;
    const cx* lT;                     // the 3 packed planes after the forward transform of the real-field adjoint
    cx* lU;
    const double *kx, *ky, *hcg, *wlo, *whi, *kxs, *kys, *filt, *filty;
    double D, sc, wex, qcorr, a2;
    int n, a, N, nx, ny, first, last, top, extrap, bk, keep0;
    long nk;
};

// lT (3 planes per slice, ts apart) = the transpose of E, B = f(R, L, E_z + i B_z) for count slices of lF
// ([6][.][nk] with planes fpl apart, slice l at l nk)
void fields_to_F_adj(const double* lF, long fpl, long nk, int count, cx* lT, long ts);
// one pass of half-cell n (or the front slice, c.top): the transposed bookkeeping and march (lT read, lU written)
void pass_mode_adj(const PassAdj& c);
// J[c] = scale Re X[c] for 3 planes; J[3], J[4] = 0 (the rho planes of the streaming deposit)
void real_part_planes(const cx* X, double* J, double scale, long nk);
// float planes -> double: count planes of n values, sstride / dstride apart
void planes_to_double(const float* src, long sstride, double* dst, long dstride, long n, int count);
// y[c dstride + i] += x[c sstride + i] for count planes of n values
void add_planes(double* y, long dstride, const double* x, long sstride, long n, int count);
// y += x over n values
void add_values(double* y, const double* x, long n);
// a reverse pass's inputs in one launch: F0d, F1d (6 planes nk apart) from the float slices f0, f1 (planes fpl
// apart), gF0 and gF1 (6 nk) cleared
void pass_prep(const float* f0, const float* f1, long fpl, double* F0d, double* F1d, double* gF0, double* gF1, long nk);
// a reverse pass's sums in one launch: acc += gP over m values, dst0 += gF0 (6 planes s0 apart; null: none),
// dst1 += gF1 (planes s1 apart)
void pass_sums(double* acc, const double* gP, long m, double* dst0, long s0, const double* gF0, double* dst1, long s1,
               const double* gF1, long nk);
// a tangent-linear pass's inputs in one launch: F0d, F1d as pass_prep; t0d = t0 (when t0 is given) and t1d = t1
// (6 planes fpl apart into nk apart); J (nJ values) cleared
void tangent_prep(const float* f0, const float* f1, long fpl, double* F0d, double* F1d, const double* t0, double* t0d,
                  const double* t1, double* t1d, double* J, long nJ, long nk);
// up to kMaxSpans device-to-device copies of whole 4-byte words in one launch (the tape's per-slice records)
struct CopySpan {
    void* dst;
    const void* src;
    long words;
};
constexpr int kMaxSpans = 8;
void copy_spans(const CopySpan* spans, int count);

// The sponge and the real fields (Sim::sponge_and_real_fields) reversed: lF (ns slices, consumed) -> the adjoints
// of R, L, E_z, B_z (the sponge rewrites R in place) and of jkz (through the bound field); Fourier y.
struct SpongeAdj {
    SpecGeom g;
    cx *lRx, *lRy, *lEz, *lBz, *lLx, *lLy, *ljkz;
    const double *hc, *wl, *wu, *mx, *mn, *mh;
    double g2, opb;
    int sponge;
};
void sponge_real_fields_adj(const SpongeAdj& c, const FftPlan& p, double* lF, long fpl, cx* T, cx* X, cx* Y,
                            cx* P);

// The beam currents (Sim::currents) reversed: the adjoints of jkx, jky, jkz (ns + 1 planes, consumed) -> those of
// the beam particles' x, y, p_x, p_y, p_z (added to).  ζ moves a particle's whole current from one half-cell to the
// next: its derivative is zero, held fixed.
struct CurrentsAdj {
    SpecGeom g;
    cx *ljkx, *ljky, *ljkz;
    const double* hcg;
    int zf, zu, keep0;
    DepositArgs d;
};
void currents_adj(const CurrentsAdj& c, const FftPlan& p, Parts P, long np, double* lx, double* ly, double* lpx,
                  double* lpy, double* lpz, cx* X, cx* Y, cx* Z, cx* S, double* ljw);

// the loss seed: lpz = 1 / count for the live particles of species sp0 (alive 1), 0 elsewhere
void seed_mean_pz(Parts P, long np, int sp0, double inv_count, double* lpz);
// per slice l < count: out[l] += Re sum_i conj(lRx) X + conj(lRy) Y, (X, Y) the unpacked pair of T (grid laser)
void laser_projection(const SpecGeom& g, const cx* T, const cx* lRx, const cx* lRy, int count, double* out);

// ---------------------------------------------------------------- the tangent-linear march (Sim::tangent)
// the streaming push of one pass on the tangents (g3d_spush.hip): st ([7][np]) in, and out with write; dJ added to
void stream_push_tlm(const g3d_spush_ctx& c, const double* dF0, const double* dF1, double* st, double* dJ, int write);
// the beam push on the tangents (g3d_bpush.hip): dst ([6][stride], x y zeta px py pz) in and out; dF in F's layout
void bunch_push_tlm(const g3d_bpush_ctx& c, const double* dF, double* dst, long stride);
// the tangent of the beam deposit (jw's layout, added to) from the tangents of x, y, p_x, p_y, p_z (zeta: none)
void deposit_tlm(const CurrentsAdj& c, Parts P, long np, const double* dx, const double* dy, const double* dpx,
                 const double* dpy, const double* dpz, double* djw);
// E, B of count slices from the packed planes T (ts apart) into the double planes F (fpl apart)
void fields_to_F_double(const cx* T, long ts, double* F, long fpl, long nk, int count);

}  // namespace g3d
