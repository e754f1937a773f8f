// The gradient of a march (no counterpart in the reference): the mean p_z of the live bunch after a number of
// steps, differentiated in the envelope values of the grid laser and the values of the bunch's current profile,
// by the reverse of every step from the last to the first.
//
// Forward: init(), then steps with a state checkpoint (device copies of the G3DCKPT1 arrays) every `every` steps.
// Reverse, segment by segment from the last: restore the segment's checkpoint, run its steps again with a tape
// (StepTape: the beam before and after its push, the streaming particles before every half-cell, the real fields
// that every pass read), then reverse its steps.  The reverse of a step undoes, in the opposite order, what step()
// did: j0 = jk; the sponge and the real fields; the march of the plasma (jk_z -= js0_z, then the slices rear to
// front and each slice's passes last to first: the real fields of the pass, its per-mode bookkeeping and march,
// the forward transform of its current, the streaming push through the half-cell with its deposit); the beam
// currents; the beam push.  The field side is linear for fixed particles and is reversed by its transpose
// (g3d_grad.hip, g3d_march_adj.hpp); the particle side by the dual-number reverses of the streaming push
// (g3d_spush.hip) and of the beam push (g3d_bpush.hip).  The decisions of the forward (trapping, particle codes,
// the half-cell of each beam particle's deposit, the sub-steps of the beam push) are held fixed; a run in which a
// particle is trapped, re-streamed or lost from the lists is refused.  Then init() in reverse as far as the
// parameters: j0 = jk, the sponge and real fields, the march of D = 0; the laser's spectra are linear in the
// envelope values (the hat functions at each slice), the bunch's zeta in the profile values (the inverse of the
// cumulative distribution, each particle in its trapezoid).
//
// Fourier y, a uniform zeta grid, the whole window on this GPU, electrons only, one bunch, no window motion.  The
// reverse is FP64 in both forms: for the FP32-mixed form it applies the transposes of the same linear field updates
// and the FP64 push reverses at the mixed trajectory (the FP32 roundings of the forward taken as the identity).
#include <algorithm>
#include <cmath>
#include <deque>
#include <unordered_map>

#include "g3d_engine.hpp"
#include "g3d_grad.hpp"
#include "g3d_runtime.hpp"

namespace g3d {

// init() with the profile derivatives of the bunch recorded (when asked), the flag cleared however init ends
void Sim::init_recording(bool dprof) {
    record_dprof_ = dprof;
    dprof_.clear();
    try { init(); } catch (...) { record_dprof_ = false; throw; }
    record_dprof_ = false;
}

void Sim::tape_beam(StepTape& t, bool before) {
    t.np = np;
    DevBuf<double>& d = before ? t.beam0 : t.beam1;
    DevBuf<int>& al = before ? t.alive0 : t.alive1;
    if (d.size() != 8 * size_t(np)) d.resize(8 * size_t(np));
    if (al.size() != size_t(np)) al.resize(size_t(np));
    if (before && t.qm.size() != size_t(np)) t.qm.resize(size_t(np));
    if (!np) return;
    const DevBuf<double>* src[8] = {&beam.x, &beam.y, &beam.zeta, &beam.px, &beam.py, &beam.pz, &beam.z, &beam.rem};
    const size_t b = size_t(np) * sizeof(double);
    for (int c = 0; c < 8; ++c)
        hip_check(hipMemcpyAsync(d.data() + size_t(c) * np, src[c]->data(), b, hipMemcpyDeviceToDevice, stream()), "hipMemcpy D2D");
    hip_check(hipMemcpyAsync(al.data(), beam.alive.data(), size_t(np) * sizeof(int), hipMemcpyDeviceToDevice, stream()), "hipMemcpy D2D");
    if (before) hip_check(hipMemcpyAsync(t.qm.data(), beam.qm.data(), b, hipMemcpyDeviceToDevice, stream()), "hipMemcpy D2D");
}

namespace {

void d2d(void* dst, const void* src, size_t bytes) {
    if (bytes) hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()), "hipMemcpy D2D");
}

// the state of a simulation between steps, in device memory: the arrays of the G3DCKPT1 checkpoint
struct Snapshot {
    long istep = 0, lost = 0, trapped = 0, restreamed = 0;
    double s = 0, X0 = 0, Y0 = 0;
    bool stop = false;
    std::vector<DevBuf<cx>> planes;
    DevBuf<float> F;
    std::vector<g3d_particle> parts, joining;
    std::deque<std::array<double, 3>> cent;
};

// the adjoints of the state and the scratch of the reverse (complex planes in the reference's layout)
struct Adjoints {
    DevBuf<cx> R[6], jk[3], j0[2], js0[3], pin, sin, si0, dl;
    DevBuf<double> F, Fm;                  // real fields after the step / after the march, [6][ns + 2][nk]
    DevBuf<double> W, Wn;                  // beam: x y zeta px py pz, [6][np]
    DevBuf<cx> T3, X, Y, P, S, lT, lU;
    DevBuf<double> jw, Fd, gF0, gF1, F0d, F1d, bJ;
    DevBuf<double> lP, lPacc, gP, zeros;   // streaming particles: x y zeta z px py pz, [7][nstrm]
};

// the beam of a step from its tape: the state before (after = false) or after the push, with the constants (q, w,
// species, ids) of the live beam, whose first t.np particles are the step's
Parts tape_beam_parts(Sim& sm, StepTape& t, bool after) {
    Parts p = sm.beam.view();
    double* d = after ? t.beam1.data() : t.beam0.data();
    p.x = d; p.y = d + t.np; p.zeta = d + 2 * t.np; p.px = d + 3 * t.np; p.py = d + 4 * t.np;
    p.pz = d + 5 * t.np; p.z = d + 6 * t.np; p.rem = d + 7 * t.np;
    p.alive = after ? t.alive1.data() : t.alive0.data();
    p.qm = t.qm.data();
    return p;
}
// the beam push of a step (Sim::push) as the context of its reverse and tangent: the state before the push, the
// fields before the step in FP64 (Fd, F's layout)
g3d_bpush_ctx beam_push_ctx(Sim& sm, StepTape& t, double* Fd, double hmin) {
    const Config& cfg = sm.cfg;
    const Parts b0 = tape_beam_parts(sm, t, false);
    g3d_bpush_ctx bc{};
    bc.nx = cfg.nx; bc.ny = cfg.ny; bc.nsl = sm.ns + 2; bc.dx = sm.dx; bc.dy = sm.dy; bc.zeta0 = cfg.zeta[0];
    bc.ihz = 1 / (cfg.zeta[1] - cfg.zeta[0]);
    bc.F = Fd;
    bc.laser_on = sm.laser.on; bc.laser_pol = sm.laser.pol; bc.a0 = sm.laser.a0; bc.kL = sm.laser.kL;
    bc.tau = sm.laser.tau; bc.cep = sm.laser.cep; bc.eta_c = sm.laser.eta_c; bc.kx = sm.laser.kx; bc.kz = sm.laser.kz;
    bc.X0 = sm.X0; bc.zc = cfg.zeta_center; bc.s0 = t.s0; bc.s1 = t.s0 + cfg.dz; bc.dz_push = cfg.dz_push;
    bc.push_phase = cfg.push_phase; bc.hmin = hmin; bc.zhand = -1e300; bc.zmax = cfg.zeta[sm.N - 1];
    bc.ylo = -1e300; bc.yhi = 1e300; bc.vz_min = cfg.vz_min;
    bc.x = b0.x; bc.y = b0.y; bc.zeta = b0.zeta; bc.px = b0.px; bc.py = b0.py; bc.pz = b0.pz; bc.z = b0.z;
    bc.rem = b0.rem; bc.qm = b0.qm; bc.alive = b0.alive; bc.np = t.np;
    return bc;
}
// the streaming push of a step's half-cells (k_stream_halfcell) as the context of its reverse and tangent: the
// constants; the half-cell, its particles and the fields are set per pass
g3d_spush_ctx stream_push_ctx(Sim& sm, StepTape& t, double ia0, double* F0d, double* F1d) {
    const Config& cfg = sm.cfg;
    const long ns_ = t.nstrm;
    g3d_spush_ctx c{};
    c.nx = cfg.nx; c.ny = cfg.ny; c.dx = sm.dx; c.dy = sm.dy; c.zeta0 = cfg.zeta[0];
    c.ihz = 1 / (cfg.zeta[1] - cfg.zeta[0]);
    c.Delta = cfg.dz; c.s1 = t.s0 + cfg.dz; c.sTrap = c.s1 + (cfg.trap_advance - 1) * cfg.dz;
    c.vz_beam = cfg.vz_beam; c.ia = ia0; c.F0 = F0d; c.F1 = F1d; c.J = nullptr;
    c.q = t.pq.data(); c.qm = t.pq.data() + ns_; c.w = t.pq.data() + 2 * ns_; c.np = ns_;
    return c;
}
// half-cell n of the tape in c: its slice positions, the particles before it and their codes
void stream_push_slot(const Sim& sm, StepTape& t, int n, g3d_spush_ctx& c) {
    const long ns_ = t.nstrm;
    double* st = t.ps.data() + size_t(n) * 7 * size_t(ns_);
    c.n = n; c.zn = sm.cfg.zeta[n]; c.zn1 = sm.cfg.zeta[n + 1];
    c.x = st; c.y = st + ns_; c.zeta = st + 2 * ns_; c.z = st + 3 * ns_; c.px = st + 4 * ns_;
    c.py = st + 5 * ns_; c.pz = st + 6 * ns_; c.alive = t.pa.data() + size_t(n) * ns_;
}
// the beam currents (Sim::currents) as the context of their reverse and tangent
CurrentsAdj currents_args(Sim& sm, const SpecGeom& g) {
    const Config& cfg = sm.cfg;
    CurrentsAdj ca{};
    ca.g = g; ca.hcg = sm.hcg.data(); ca.zf = cfg.zeta_filter; ca.zu = sm.zax.uniform; ca.keep0 = cfg.box_mean;
    ca.d.zax = sm.zax; ca.d.yax = sm.yax; ca.d.hcg = sm.hcg.data(); ca.d.ia = sm.iA.data(); ca.d.id = sm.iD.data();
    ca.d.J = nullptr; ca.d.dx = sm.dx; ca.d.dy = sm.dy; ca.d.nh = sm.ns + 3; ca.d.a = sm.a; ca.d.N = sm.N;
    ca.d.nx = cfg.nx; ca.d.ny = cfg.ny; ca.d.yf = 1; ca.d.nk = sm.nk;
    return ca;
}
// the predictor's extrapolation weight of half-cell n (march_plasma: none for the first half-cell marched)
double pass_wex(const std::vector<double>& cen, int n, int N) {
    return (n >= 1 && n + 1 <= N - 2 && n != N - 2) ? (cen[size_t(n)] - cen[size_t(n - 1)]) / (cen[size_t(n + 1)] - cen[size_t(n)])
                                                     : 0.0;
}

// what the reverse and the tangent-linear run of the march cover
void check_reversible(const Sim& sm, int steps, int every, int nenv, int nprof) {
    const Config& cfg = sm.cfg;
    bool ions = false;
    for (const auto& is : cfg.ions) ions = ions || cfg.loaded(is);
    if (!sm.yf) throw std::runtime_error("gradient: the Fourier y grid only");
    if (sm.a != 0 || sm.b != sm.N) throw std::runtime_error("gradient: the whole window on this GPU");
    if (!cfg.plasma_on || !cfg.self_fields) throw std::runtime_error("gradient: a plasma with self fields");
    if (ions) throw std::runtime_error("gradient: electrons only (no mobile or ionizing ions)");
    if (cfg.move_x || cfg.move_y) throw std::runtime_error("gradient: a fixed window (move_x = move_y = 0)");
    if (!sm.zax.uniform) throw std::runtime_error("gradient: a uniform zeta grid");
    if (cfg.bunches.size() != 1 || cfg.freeze_bunch) throw std::runtime_error("gradient: one bunch, not frozen");
    if (nenv > 0 && (cfg.glasers.empty() || cfg.glasers[0].n_env != nenv))
        throw std::runtime_error("gradient: the first grid laser has no envelope of " + std::to_string(nenv) + " values");
    if (nprof > 0 && int(cfg.prof.size()) != nprof)
        throw std::runtime_error("gradient: the bunch has no profile of " + std::to_string(nprof) + " values");
    if (nprof > 0 && (cfg.bunches[0].emitx > 0 || cfg.bunches[0].emity > 0))
        throw std::runtime_error("gradient: the profile gradient needs a bunch without emittance");
    if (steps < 1 || every < 1) throw std::runtime_error("gradient: steps and every must be positive");
    if (sm.istep != 0 || sm.s != 0 || sm.np != 0) throw std::runtime_error("gradient: runs from init() on a new simulation");

}

}  // namespace

void Sim::gradient(int steps, int every, double* loss, double* genv, int nenv, double* gprof, int nprof) {
    check_reversible(*this, steps, every, nenv, nprof);

    init_recording(nprof > 0);
    const long np0 = np;
    auto check_lists = [&] {
        if (restreamed_total || np != np0 + trapped_total)
            throw std::runtime_error("gradient: beam particles were re-streamed (" + std::to_string(restreamed_total) +
                                     ") or dropped from the beam list (" + std::to_string(np0) + " + " +
                                     std::to_string(trapped_total) + " trapped -> " + std::to_string(np) +
                                     "); the reverse follows the electrons held over to the next step and those "
                                     "trapped into the beam, but not these");
    };
    auto take = [&](Snapshot& sn) {
        ensure_layout(false, false);
        sn.istep = istep; sn.lost = lost_total; sn.trapped = trapped_total; sn.restreamed = restreamed_total;
        sn.s = s; sn.X0 = X0; sn.Y0 = Y0; sn.stop = stop_;
        DevBuf<cx>* src[11] = {&Rx, &Ry, &Ez, &Bz, &Lx, &Ly, &j0x, &j0y, &js0x, &js0y, &js0z};
        sn.planes.clear();
        for (DevBuf<cx>* p : src) {
            sn.planes.emplace_back(p->size());
            d2d(sn.planes.back().data(), p->data(), p->size() * sizeof(cx));
        }
        sn.F.resize(F.size());
        d2d(sn.F.data(), F.data(), F.size() * sizeof(float));
        sn.parts = host_particles();
        sn.joining = joining_;
        sn.cent = cent_;
    };
    auto restore = [&](const Snapshot& sn) {
        ensure_layout(false, false);
        istep = sn.istep; lost_total = sn.lost; trapped_total = sn.trapped; restreamed_total = sn.restreamed;
        s = sn.s; X0 = sn.X0; Y0 = sn.Y0; stop_ = sn.stop;
        DevBuf<cx>* dst[11] = {&Rx, &Ry, &Ez, &Bz, &Lx, &Ly, &j0x, &j0y, &js0x, &js0y, &js0z};
        for (int k = 0; k < 11; ++k) d2d(dst[k]->data(), sn.planes[size_t(k)].data(), dst[k]->size() * sizeof(cx));
        d2d(F.data(), sn.F.data(), F.size() * sizeof(float));
        set_particles(sn.parts);
        joining_ = sn.joining;
        cent_ = sn.cent;
        jw_zero = false;
    };

    // forward, with the checkpoints; the state at the end is kept too, to leave the simulation there
    std::vector<Snapshot> snaps;
    for (int k = 0; k < steps; ++k) {
        if (k % every == 0) { snaps.emplace_back(); take(snaps.back()); }
        step();
        check_lists();
    }
    Snapshot at_end;
    take(at_end);
    {
        double sum = 0;
        long cnt = 0;
        for (const auto& p : host_particles())
            if (p.alive == 1 && p.sp == 0) { sum += p.pz; ++cnt; }
        if (!cnt) throw std::runtime_error("gradient: no live bunch particle at the end");
        *loss = sum / double(cnt);
    }

    // the adjoint state
    const long pf = long(ns) * nk, pj = long(ns + 1) * nk, fpl = long(ns + 2) * nk, Fsz = 6 * fpl;
    Adjoints A;
    for (auto& r : A.R) r.resize(size_t(pf));
    for (auto& r : A.jk) r.resize(size_t(pj));
    for (auto& r : A.j0) r.resize(size_t(pj));
    for (auto& r : A.js0) r.resize(size_t(pj));
    A.pin.resize(3 * size_t(nk)); A.sin.resize(3 * size_t(nk)); A.si0.resize(3 * size_t(nk)); A.dl.resize(3 * size_t(nk));
    A.F.resize(size_t(Fsz)); A.Fm.resize(size_t(Fsz)); A.Fd.resize(size_t(Fsz));
    const long npm = np;                          // the beam only grows (trapped electrons are appended)
    A.W.resize(6 * size_t(npm)); A.Wn.resize(6 * size_t(npm));
    A.T3.resize(3 * size_t(pf)); A.X.resize(size_t(pj)); A.Y.resize(size_t(pj)); A.P.resize(size_t(pj));
    A.S.resize(size_t(pj)); A.lT.resize(3 * size_t(nk)); A.lU.resize(3 * size_t(nk));
    A.jw.resize(3 * size_t(ns + 3) * size_t(nk));
    for (DevBuf<double>* p : {&A.gF0, &A.gF1, &A.F0d, &A.F1d}) p->resize(6 * size_t(nk));
    A.bJ.resize(5 * size_t(nk));
    {                                            // the loss: the mean p_z of the live bunch at the end

        long cnt = 0;
        for (const auto& p : host_particles()) if (p.alive == 1 && p.sp == 0) ++cnt;
        seed_mean_pz(beam.view(), np, 0, 1.0 / double(cnt), A.W.data() + 5 * npm);
    }

    // constants of the reverse
    const int P = cfg.iterations;
    const double D = 0.5 * cfg.dz, h = cfg.zeta[1] - cfg.zeta[0];
    std::vector<double> cen(size_t(N - 1));
    for (int m = 0; m < N - 1; ++m) cen[size_t(m)] = 0.5 * (cfg.zeta[m] + cfg.zeta[m + 1]);
    double ia0 = 0;
    iA.to_host(&ia0, 0, 1);
    double hmin = 1e300;
    for (int m = 0; m + 1 < N; ++m) hmin = std::min(hmin, cfg.zeta[m + 1] - cfg.zeta[m]);
    SpecGeom g = geom();
    g.cm = 0;
    const double sc = 1.0 / (double(cfg.nx) * cfg.ny);

    PassAdj pb{};
    pb.lRx = A.R[0].data(); pb.lRy = A.R[1].data(); pb.lEz = A.R[2].data(); pb.lBz = A.R[3].data();
    pb.lLx = A.R[4].data(); pb.lLy = A.R[5].data();
    pb.ljkx = A.jk[0].data(); pb.ljky = A.jk[1].data(); pb.ljkz = A.jk[2].data();
    pb.lj0x = A.j0[0].data(); pb.lj0y = A.j0[1].data();
    pb.ljs0x = A.js0[0].data(); pb.ljs0y = A.js0[1].data(); pb.ljs0z = A.js0[2].data();
    pb.lpin = A.pin.data(); pb.lsin = A.sin.data(); pb.lsi0 = A.si0.data(); pb.ldl = A.dl.data();
    pb.lT = A.lT.data(); pb.lU = A.lU.data();
    pb.kx = kxd.data(); pb.ky = kyd.data(); pb.hcg = hcg.data(); pb.wlo = wlo.data(); pb.whi = whi.data();
    pb.kxs = kxs.data(); pb.kys = kys.data(); pb.filt = filt.data(); pb.filty = filty.data();
    pb.D = D; pb.sc = sc; pb.a2 = cfg.smooth_length * cfg.smooth_length;
    pb.a = a; pb.N = N; pb.nx = cfg.nx; pb.ny = cfg.ny; pb.extrap = cfg.predictor_extrapolate; pb.keep0 = cfg.box_mean;
    pb.nk = nk;

    SpongeAdj sa{};
    sa.g = g;
    sa.lRx = A.R[0].data(); sa.lRy = A.R[1].data(); sa.lEz = A.R[2].data(); sa.lBz = A.R[3].data();
    sa.lLx = A.R[4].data(); sa.lLy = A.R[5].data(); sa.ljkz = A.jk[2].data();
    sa.hc = hcg.data(); sa.wl = wlo.data(); sa.wu = whi.data(); sa.mx = maskx.data(); sa.mn = masky_n.data();
    sa.mh = masky_h.data(); sa.g2 = gm2_b; sa.opb = 1 + beta_b; sa.sponge = cfg.sponge > 0 || cfg.sponge_y > 0;

    auto sponge_rev = [&] {
        // j0 = jk (x, y) after the fields; then the sponge and the real fields
        axpy_planes(A.jk[0].data(), A.j0[0].data(), 1, pj);
        axpy_planes(A.jk[1].data(), A.j0[1].data(), 1, pj);
        A.j0[0].zero(); A.j0[1].zero();
        sponge_real_fields_adj(sa, fft, A.F.data(), fpl, A.T3.data(), A.X.data(), A.Y.data(), A.P.data());
        A.F.zero();
    };

    std::unordered_map<unsigned long long, std::array<double, 6>> carry;   // x y zeta px py pz by id
    auto march_rev = [&](StepTape& t) {
        const long ns_ = t.nstrm;
        if (A.lP.size() != 7 * size_t(ns_)) {
            A.lP.resize(7 * size_t(ns_)); A.lPacc.resize(7 * size_t(ns_)); A.gP.resize(7 * size_t(ns_));
            A.zeros.resize(7 * size_t(ns_));
        }
        A.Fm.zero(); A.pin.zero(); A.sin.zero(); A.si0.zero(); A.dl.zero();
        // electrons held over to the next step (trapped at z = s1 below the beam threshold) start its list with
        // their state at the end of this march: their adjoints from the next step's reverse, by id, start this
        // one's (z is set to s1 and carries none)
        std::vector<unsigned long long> ids(static_cast<size_t>(ns_));
        std::vector<int> alive0(static_cast<size_t>(ns_));
        t.ids.to_host(ids.data(), 0, size_t(ns_));
        t.pa.to_host(alive0.data(), size_t(N - 2) * size_t(ns_), size_t(ns_));
        {
            std::vector<double> lp(7 * size_t(ns_), 0.0);
            for (long i = 0; i < ns_; ++i) {
                const auto it = carry.find(ids[size_t(i)]);
                if (it == carry.end()) continue;
                const int slot[6] = {0, 1, 2, 4, 5, 6};
                for (int c6 = 0; c6 < 6; ++c6) lp[size_t(slot[c6]) * ns_ + i] = it->second[size_t(c6)];
            }
            A.lP.from_host(lp.data(), 0, lp.size());
        }
        axpy_planes(A.js0[2].data(), A.jk[2].data(), -1, pj);       // jk_z -= js0_z after the slices
        g3d_spush_ctx c = stream_push_ctx(*this, t, ia0, A.F0d.data(), A.F1d.data());
        g3d_spush_adj ad{};
        ad.bJ = A.bJ.data(); ad.gF0 = A.gF0.data(); ad.gF1 = A.gF1.data();
        double* gp = A.gP.data();
        ad.gx = gp; ad.gy = gp + ns_; ad.gzeta = gp + 2 * ns_; ad.gz = gp + 3 * ns_; ad.gpx = gp + 4 * ns_;
        ad.gpy = gp + 5 * ns_; ad.gpz = gp + 6 * ns_;
        for (int n = a; n <= N - 2; ++n) {
            stream_push_slot(*this, t, n, c);
            A.lPacc.zero();
            for (int p = P - 1; p >= 0; --p) {
                // the later pass's field adjoint (gF0 of pass p + 1, kept until this pass clears it below)
                if (p == P - 1) fields_to_F_adj(A.Fm.data() + long(n - a) * nk, fpl, nk, 1, A.lT.data(), nk);
                else fields_to_F_adj(A.gF0.data(), nk, nk, 1, A.lT.data(), nk);
                fft_run(fft, A.lT.data(), 3, false);
                PassAdj pa = pb;
                pa.n = n; pa.first = p == 0; pa.last = p == P - 1; pa.top = 0; pa.bk = P > 1;
                pa.wex = pass_wex(cen, n, N);
                pa.qcorr = D * wlo_h[size_t(n + 1)];
                pass_mode_adj(pa);
                fft_rows(fft, A.lU.data(), 3L * cfg.ny, true, 1.0);
                fft_cols(fft, A.lU.data(), 3, true, 1.0);
                real_part_planes(A.lU.data(), A.bJ.data(), 1.0, nk);
                const float* f0 = p == 0 ? t.F0.data() + long(n - a) * nk
                                         : t.Fmid.data() + size_t(p - 1) * size_t(Fsz) + long(n - a) * nk;
                pass_prep(f0, t.Fmarch.data() + long(n + 1 - a) * nk, fpl, A.F0d.data(), A.F1d.data(), A.gF0.data(),
                          A.gF1.data(), nk);
                const double* bs = p == P - 1 ? A.lP.data() : A.zeros.data();
                ad.bx = bs; ad.by = bs + ns_; ad.bzeta = bs + 2 * ns_; ad.bz = bs + 3 * ns_; ad.bpx = bs + 4 * ns_;
                ad.bpy = bs + 5 * ns_; ad.bpz = bs + 6 * ns_;
                if (g3d_stream_push_vjp(&c, &ad)) throw std::runtime_error(last_error());
                // the pass's particle adjoints summed; its field adjoints into F (the first pass) or, from gF0, into
                // the pass before; into the march's F of the next slice
                pass_sums(A.lPacc.data(), A.gP.data(), 7 * ns_, p == 0 ? A.F.data() + long(n - a) * nk : nullptr, fpl,
                          A.gF0.data(), A.Fm.data() + long(n + 1 - a) * nk, fpl, A.gF1.data(), nk);
            }
            A.lP.swap(A.lPacc);
        }
        // the adjoints of the list's start state: those of the electrons held over from the step before go on
        {
            std::vector<double> lp(7 * size_t(ns_));
            A.lP.to_host(lp.data(), 0, lp.size());
            carry.clear();
            for (long i = 0; i < ns_; ++i) {
                if (alive0[size_t(i)] != 6) continue;
                std::array<double, 6>& v = carry[ids[size_t(i)]];
                const int slot[6] = {0, 1, 2, 4, 5, 6};
                for (int c6 = 0; c6 < 6; ++c6) v[size_t(c6)] = lp[size_t(slot[c6]) * ns_ + i];
            }
        }
        // the front slice
        fields_to_F_adj(A.Fm.data() + long(N - 1 - a) * nk, fpl, nk, 1, A.lT.data(), nk);
        fft_run(fft, A.lT.data(), 3, false);
        PassAdj pa = pb;
        pa.n = N - 1; pa.first = 1; pa.last = 1; pa.top = 1; pa.bk = 0; pa.wex = 0; pa.qcorr = 0;
        pass_mode_adj(pa);
    };

    CurrentsAdj ca = currents_args(*this, g);
    ca.ljkx = A.jk[0].data(); ca.ljky = A.jk[1].data(); ca.ljkz = A.jk[2].data();

    auto step_rev = [&](StepTape& t) {
        sponge_rev();
        // electrons trapped into the beam at the end of the march (appended after the t.np particles of the step):
        // their adjoints start the march's reverse from their streaming state, by id
        if (!t.added.empty()) {
            std::vector<double> w(6 * size_t(npm));
            A.W.to_host(w.data(), 0, w.size());
            for (size_t k = 0; k < t.added.size(); ++k) {
                std::array<double, 6>& v = carry[t.added[k]];
                for (int c6 = 0; c6 < 6; ++c6) v[size_t(c6)] = w[size_t(c6) * npm + size_t(t.np) + k];
            }
        }
        march_rev(t);
        // the beam currents of the pushed beam
        double* W = A.W.data();
        currents_adj(ca, fft, tape_beam_parts(*this, t, true), t.np, W, W + npm, W + 3 * npm, W + 4 * npm, W + 5 * npm,
                     A.X.data(), A.Y.data(), A.P.data(), A.S.data(), A.jw.data());
        // the beam push (src/sim_particles.cpp:6-24) in the fields before the step
        planes_to_double(t.F0.data(), F.size(), A.Fd.data(), Fsz, Fsz, 1);
        const g3d_bpush_ctx bc = beam_push_ctx(*this, t, A.Fd.data(), hmin);
        double* Wn = A.Wn.data();
        g3d_bpush_adj ba{};
        ba.bx = W; ba.by = W + npm; ba.bzeta = W + 2 * npm; ba.bpx = W + 3 * npm; ba.bpy = W + 4 * npm;
        ba.bpz = W + 5 * npm;
        ba.gx = Wn; ba.gy = Wn + npm; ba.gzeta = Wn + 2 * npm; ba.gpx = Wn + 3 * npm; ba.gpy = Wn + 4 * npm;
        ba.gpz = Wn + 5 * npm; ba.gF = A.F.data();
        if (g3d_bunch_push_vjp(&bc, &ba)) throw std::runtime_error(last_error());
        A.W.swap(A.Wn);
    };

    // reverse, segment by segment
    std::vector<StepTape> tapes(size_t(std::min(every, steps)));
    for (int seg = int(snaps.size()) - 1; seg >= 0; --seg) {
        const int k0 = seg * every, k1 = std::min(steps, k0 + every);
        restore(snaps[size_t(seg)]);
        for (int k = k0; k < k1; ++k) {
            StepTape& t = tapes[size_t(k - k0)];
            t.s0 = s;
            tape_beam(t, true);
            if (t.F0.size() != F.size()) t.F0.resize(F.size());
            d2d(t.F0.data(), F.data(), F.size() * sizeof(float));
            tape_ = &t;
            try { step(); } catch (...) { tape_ = nullptr; throw; }
            tape_ = nullptr;
            check_lists();
        }
        for (int k = k1 - 1; k >= k0; --k) step_rev(tapes[size_t(k - k0)]);
    }
    tapes.clear();
    restore(at_end);

    // init() in reverse as far as the parameters: j0 = jk, the sponge and real fields, the march of D = 0
    sponge_rev();
    {
        g3d_march_adj_ctx mc{};
        mc.Rx = reinterpret_cast<g3d_cpx*>(A.R[0].data()); mc.Ry = reinterpret_cast<g3d_cpx*>(A.R[1].data());
        mc.Ez = reinterpret_cast<g3d_cpx*>(A.R[2].data()); mc.Bz = reinterpret_cast<g3d_cpx*>(A.R[3].data());
        mc.Lx = reinterpret_cast<g3d_cpx*>(A.R[4].data()); mc.Ly = reinterpret_cast<g3d_cpx*>(A.R[5].data());
        mc.up = nullptr;
        mc.j0x = reinterpret_cast<g3d_cpx*>(A.j0[0].data()); mc.j0y = reinterpret_cast<g3d_cpx*>(A.j0[1].data());
        mc.jkx = reinterpret_cast<g3d_cpx*>(A.jk[0].data()); mc.jky = reinterpret_cast<g3d_cpx*>(A.jk[1].data());
        mc.jkz = reinterpret_cast<g3d_cpx*>(A.jk[2].data());
        mc.kx = kxd.data(); mc.ky = kyd.data(); mc.hcg = hcg.data(); mc.wlo = wlo.data(); mc.whi = whi.data();
        mc.D = 0; mc.a = a; mc.N = N; mc.nx = cfg.nx; mc.ny = cfg.ny; mc.nk = nk;
        if (g3d_march_adjoint(&mc, a, b)) throw std::runtime_error(last_error());
    }
    // the laser: R += the spectra of the laser field, linear in the envelope at each slice
    if (nenv > 0) {
        const auto& Lz = cfg.glasers[0];
        GridLaserArgs la = grid_laser_args(Lz);
        std::vector<double> ones(size_t(N), 1.0);
        DevBuf<double> envd(static_cast<size_t>(N));
        envd.from_host(ones.data(), 0, size_t(N));
        la.env = envd.data();
        grid_laser_fill(la, T3.data(), a, ns);
        fft_run(fft, T3.data(), ns, false);
        DevBuf<double> proj(static_cast<size_t>(ns));
        laser_projection(g, T3.data(), A.R[0].data(), A.R[1].data(), ns, proj.data());
        std::vector<double> ph(static_cast<size_t>(ns));
        proj.to_host(ph.data(), 0, size_t(ns));
        for (int k = 0; k < nenv; ++k) {
            double gk = 0;
            const double zk = Lz.env_z0 + k * Lz.env_dz;
            for (int n = 0; n < ns; ++n)
                gk += std::max(0.0, 1 - std::fabs(cfg.zeta[a + n] - zk) / Lz.env_dz) * ph[size_t(n)];
            genv[k] = gk;
        }
    }
    // the profile: zeta of each bunch particle
    if (nprof > 0) {
        if (dprof_.size() != size_t(np0) * size_t(nprof)) throw std::runtime_error("gradient: profile derivatives missing");
        std::vector<double> lz(static_cast<size_t>(np0));
        A.W.to_host(lz.data(), 2 * size_t(npm), size_t(np0));
        for (int m = 0; m < nprof; ++m) gprof[m] = 0;
        for (long i = 0; i < np0; ++i)
            for (int m = 0; m < nprof; ++m) gprof[m] += lz[size_t(i)] * dprof_[size_t(i) * nprof + m];
    }
    hip_check(hipStreamSynchronize(stream()), "synchronize");
}

// The tangent-linear run of the same march (no counterpart in the reference): the derivative of the loss of
// Sim::gradient along a direction (denv, dprof) of the parameters, by carrying the tangents of every state array
// forward through the steps beside the primal ones.  The field side is linear for fixed particles, so its tangent
// runs the forward's own kernels on the tangent arrays (the march of D = 0, the pass's row transforms and per-mode
// kernel, the sponge, the currents' filter and transforms) with the real fields kept in FP64; the pushes and the
// beam deposit run on dual numbers of one direction (k_spush_tlm, k_bpush_tlm, k_dep_tlm), reading the primal
// state from the step's tape.  A check of the reverse: <gradient, direction> equals this derivative.
void Sim::tangent(int steps, const double* denv, int nenv, const double* dprof, int nprof, double* loss, double* dloss) {
    check_reversible(*this, steps, 1, nenv, nprof);
    init_recording(nprof > 0);
    const long pf = long(ns) * nk, pj = long(ns + 1) * nk, fpl = long(ns + 2) * nk, Fsz = 6 * fpl, ts = pf;
    const double h = cfg.zeta[1] - cfg.zeta[0], opb = 1 + beta_b, D = 0.5 * cfg.dz;
    const int P = cfg.iterations;
    const bool sp = cfg.sponge > 0 || cfg.sponge_y > 0;
    SpecGeom g = geom();
    g.cm = 0;
    double ia0 = 0;
    iA.to_host(&ia0, 0, 1);
    double hmin = 1e300;
    for (int m = 0; m + 1 < N; ++m) hmin = std::min(hmin, cfg.zeta[m + 1] - cfg.zeta[m]);
    std::vector<double> cen(size_t(N - 1));
    for (int m = 0; m < N - 1; ++m) cen[size_t(m)] = 0.5 * (cfg.zeta[m] + cfg.zeta[m + 1]);

    // the tangent state
    DevBuf<cx> tR[6], tjk[3], tj0[2], tjs0[3], tpin, tsin, tsi0, tdl, tX, tT;
    for (auto& r : tR) r.resize(size_t(pf));
    for (auto& r : tjk) r.resize(size_t(pj));
    for (auto& r : tj0) r.resize(size_t(pj));
    for (auto& r : tjs0) r.resize(size_t(pj));
    for (DevBuf<cx>* q : {&tpin, &tsin, &tsi0, &tdl, &tX, &tT}) q->resize(3 * size_t(nk));
    DevBuf<double> tF(static_cast<size_t>(Fsz)), tFm(static_cast<size_t>(Fsz)), Fd(static_cast<size_t>(Fsz));
    DevBuf<double> tFmid(6 * size_t(nk)), tJ(5 * size_t(nk)), F0d(6 * size_t(nk)), F1d(6 * size_t(nk));
    DevBuf<double> dF0(6 * size_t(nk)), dF1(6 * size_t(nk)), tjw(3 * size_t(ns + 3) * size_t(nk)), tP;
    long cap = np;
    DevBuf<double> tW(6 * size_t(cap));
    if (nprof > 0) {                                  // the bunch's zeta along dprof
        std::vector<double> w(6 * size_t(cap), 0.0);
        for (long i = 0; i < np; ++i) {
            double v = 0;
            for (int m = 0; m < nprof; ++m) v += dprof[m] * dprof_[size_t(i) * nprof + m];
            w[2 * size_t(cap) + size_t(i)] = v;
        }
        tW.from_host(w.data(), 0, w.size());
    }
    auto tmview = [&](double Dm) {
        MarchView v = mview(Dm);
        v.Rx = tR[0].data(); v.Ry = tR[1].data(); v.Ez = tR[2].data(); v.Bz = tR[3].data(); v.Lx = tR[4].data();
        v.Ly = tR[5].data(); v.j0x = tj0[0].data(); v.j0y = tj0[1].data(); v.jkx = tjk[0].data();
        v.jky = tjk[1].data(); v.jkz = tjk[2].data(); v.cm = 0;
        return v;
    };
    auto sponge_tlm = [&] {
        if (sp) {
            sponge_pack(g, tR[0].data(), tR[1].data(), tjk[2].data(), hcg.data(), wlo.data(), whi.data(), gm2_b, opb,
                        S1.data(), S2.data());
            fft_run(fft, S1.data(), ns, true);
            sponge_mask(g, S1.data(), maskx.data(), masky_n.data(), masky_h.data(), ns);
            fft_run(fft, S1.data(), ns, false);
        }
        fields_pack3(g, sp ? S1.data() : nullptr, S2.data(), opb, tR[0].data(), tR[1].data(), tR[4].data(),
                     tR[5].data(), tR[2].data(), tR[3].data(), T3.data(), ts, ns);
        fft_run(fft, T3.data(), 3 * ns, true);
        fields_to_F_double(T3.data(), ts, tF.data(), fpl, nk, ns);
        copy_planes(tj0[0].data(), tjk[0].data(), pj);
        copy_planes(tj0[1].data(), tjk[1].data(), pj);
    };
    // init: the laser along denv (its spectra are linear in the envelope), the march of D = 0, the sponge
    if (nenv > 0) {
        const auto& Lz = cfg.glasers[0];
        GridLaserArgs la = grid_laser_args(Lz);
        std::vector<double> e(static_cast<size_t>(N));
        for (int n = 0; n < N; ++n) e[size_t(n)] = hat_sum(denv, nenv, Lz.env_z0, Lz.env_dz, cfg.zeta[n]);
        DevBuf<double> envd(static_cast<size_t>(N));
        envd.from_host(e.data(), 0, size_t(N));
        la.env = envd.data();
        grid_laser_fill(la, T3.data(), a, ns);
        fft_run(fft, T3.data(), ns, false);
        grid_laser_add(g, T3.data(), tR[0].data(), tR[1].data(), ns, 0);
    }
    march_all(tmview(0.0), a, b);
    sponge_tlm();

    const CurrentsAdj ca = currents_args(*this, g);
    std::unordered_map<unsigned long long, std::array<double, 6>> carry;   // x y zeta px py pz by id
    StepTape t;
    for (int k = 0; k < steps; ++k) {
        t.s0 = s;
        tape_beam(t, true);
        if (t.F0.size() != F.size()) t.F0.resize(F.size());
        d2d(t.F0.data(), F.data(), F.size() * sizeof(float));
        tape_ = &t;
        try { step(); } catch (...) { tape_ = nullptr; throw; }
        tape_ = nullptr;
        if (restreamed_total) throw std::runtime_error("gradient: beam particles were re-streamed");
        const long np_step = t.np;
        // the beam push along the tangent, in the fields before the step
        planes_to_double(t.F0.data(), Fsz, Fd.data(), Fsz, Fsz, 1);
        bunch_push_tlm(beam_push_ctx(*this, t, Fd.data(), hmin), tF.data(), tW.data(), cap);
        // the beam currents of the pushed beam
        tjw.zero();
        double* W = tW.data();
        deposit_tlm(ca, tape_beam_parts(*this, t, true), np_step, W, W + cap, W + 3 * cap, W + 4 * cap, W + 5 * cap, tjw.data());
        currents_zfilter(g, tjw.data(), hcg.data(), cfg.zeta_filter, zax.uniform, S1.data(), tjk[2].data());
        fft_run(fft, S1.data(), ns + 1, false);
        fft_run(fft, tjk[2].data(), ns + 1, false);
        currents_unpack(g, S1.data(), tjk[0].data(), tjk[1].data(), tjk[2].data(), cfg.plasma_on && cfg.box_mean);
        // the march of the plasma
        const long ns_ = t.nstrm;
        if (tP.size() != 7 * size_t(ns_)) tP.resize(7 * size_t(ns_));
        std::vector<unsigned long long> ids(static_cast<size_t>(ns_));
        std::vector<int> alive0(static_cast<size_t>(ns_));
        t.ids.to_host(ids.data(), 0, size_t(ns_));
        t.pa.to_host(alive0.data(), size_t(N - 2) * size_t(ns_), size_t(ns_));
        {
            std::vector<double> tp(7 * size_t(ns_), 0.0);
            const int slot[6] = {0, 1, 2, 4, 5, 6};
            for (long i = 0; i < ns_; ++i) {
                if (alive0[size_t(i)] != 6) continue;
                const auto it = carry.find(ids[size_t(i)]);
                if (it == carry.end()) continue;
                for (int c6 = 0; c6 < 6; ++c6) tp[size_t(slot[c6]) * ns_ + i] = it->second[size_t(c6)];
            }
            tP.from_host(tp.data(), 0, tp.size());
        }
        tpin.zero(); tsin.zero(); tsi0.zero(); tdl.zero();
        ModeArgs m{};
        m.mv = tmview(D);
        m.sin_ = tsin.data(); m.si0 = tsi0.data(); m.pin = tpin.data(); m.dl = tdl.data();
        m.js0x = tjs0[0].data(); m.js0y = tjs0[1].data(); m.js0z = tjs0[2].data();
        m.jkx = tjk[0].data(); m.jky = tjk[1].data(); m.jkz = tjk[2].data();
        m.kxs = kxs.data(); m.kys = kys.data(); m.filt = filt.data(); m.filty = filty.data();
        m.a2 = cfg.smooth_length * cfg.smooth_length; m.extrap = cfg.predictor_extrapolate; m.keep0 = cfg.box_mean;
        m.ny = cfg.ny; m.kf = cfg.k_filter;
        const double sc = 1.0 / (double(cfg.nx) * cfg.ny);
        auto real_fields_t = [&](double* dst, long stride) {
            fft_rows(fft, tT.data(), 3L * cfg.ny, true, sc);
            fields_to_F_double(tT.data(), nk, dst, stride, nk, 1);
        };
        // front: slice N-1
        m.n = N - 1; m.first = 1; m.last = 1; m.top = 1; m.wex = 0; m.qcorr = 0;
        m.bRx = m.bRy = m.bEz = m.bBz = nullptr;
        pass_modes(m, fft, tX.data(), tT.data(), nullptr, 1);
        real_fields_t(tFm.data() + long(N - 1 - a) * nk, fpl);
        g3d_spush_ctx c = stream_push_ctx(*this, t, ia0, F0d.data(), F1d.data());
        for (int n = N - 2; n >= a; --n) {
            stream_push_slot(*this, t, n, c);
            for (int p = 0; p < P; ++p) {
                const bool last = p == P - 1;
                const float* f0 = p == 0 ? t.F0.data() + long(n - a) * nk
                                         : t.Fmid.data() + size_t(p - 1) * size_t(Fsz) + long(n - a) * nk;
                // the fields of the pass and their tangents (the slice's own on the first pass), the tangent
                // currents cleared: one launch
                tangent_prep(f0, t.Fmarch.data() + long(n + 1 - a) * nk, fpl, F0d.data(), F1d.data(),
                             p == 0 ? tF.data() + long(n - a) * nk : nullptr, tFmid.data(),
                             tFm.data() + long(n + 1 - a) * nk, dF1.data(), tJ.data(), 5 * long(nk), nk);
                stream_push_tlm(c, tFmid.data(), dF1.data(), tP.data(), tJ.data(), last);
                pass_rows_fwd(tJ.data(), TileGeom{}, fft, tX.data(), nullptr, nullptr, 0, 0, 0);
                m.n = n; m.first = p == 0; m.last = last; m.top = 0;
                m.wex = pass_wex(cen, n, N);
                m.qcorr = D * wlo_h[size_t(n + 1)];
                // several passes: a pass before the last leaves slice n at s0 (the kernel reads this only as a flag)
                cx* flag = P > 1 ? tX.data() : nullptr;
                m.bRx = m.bRy = m.bEz = m.bBz = flag;
                pass_modes(m, fft, tX.data(), tT.data(), nullptr, 0);
                if (last) real_fields_t(tFm.data() + long(n - a) * nk, fpl);
                else real_fields_t(tFmid.data(), nk);
            }
        }
        axpy_planes(tjk[2].data(), tjs0[2].data(), -1, pj);   // jk_z -= js0_z after the slices
        // the electrons held over to the next step and those trapped into the beam keep their tangents
        {
            std::vector<double> tp(7 * size_t(ns_));
            tP.to_host(tp.data(), 0, tp.size());
            carry.clear();
            const int slot[6] = {0, 1, 2, 4, 5, 6};
            for (long i = 0; i < ns_; ++i) {
                std::array<double, 6>& v = carry[ids[size_t(i)]];
                for (int c6 = 0; c6 < 6; ++c6) v[size_t(c6)] = tp[size_t(slot[c6]) * ns_ + i];
            }
            if (!t.added.empty()) {
                const long ncap = np_step + long(t.added.size());
                std::vector<double> w(6 * size_t(cap)), w2(6 * size_t(ncap), 0.0);
                tW.to_host(w.data(), 0, w.size());
                for (int c6 = 0; c6 < 6; ++c6) {
                    for (long i = 0; i < np_step; ++i) w2[size_t(c6) * ncap + size_t(i)] = w[size_t(c6) * cap + size_t(i)];
                    for (size_t q = 0; q < t.added.size(); ++q) w2[size_t(c6) * ncap + size_t(np_step) + q] = carry[t.added[q]][size_t(c6)];
                }
                cap = ncap;
                tW.resize(6 * size_t(cap));
                tW.from_host(w2.data(), 0, w2.size());
            }
        }
        if (np != cap) throw std::runtime_error("tangent: the beam list changed other than by trapped electrons");
        sponge_tlm();
    }
    // the loss and its derivative along the direction
    const std::vector<g3d_particle> ps = host_particles();
    std::vector<double> tpz(static_cast<size_t>(cap));
    tW.to_host(tpz.data(), 5 * size_t(cap), size_t(cap));
    double sum = 0, dsum = 0;
    long cnt = 0;
    for (long i = 0; i < np; ++i)
        if (ps[size_t(i)].alive == 1 && ps[size_t(i)].sp == 0) { sum += ps[size_t(i)].pz; dsum += tpz[size_t(i)]; ++cnt; }
    if (!cnt) throw std::runtime_error("tangent: no live bunch particle at the end");
    *loss = sum / double(cnt);
    *dloss = dsum / double(cnt);
    hip_check(hipStreamSynchronize(stream()), "synchronize");
}

}  // namespace g3d
