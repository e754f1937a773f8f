// The free-hand y grid: x is Fourier (one x transform per row), y is a staggered finite-difference grid between
// conducting walls, so every k_x is a line along y and each slice couples the points of a line through two
// tridiagonal systems.  One lane runs one line (the eliminations along y are sequential); lanes are adjacent k_x,
// so every load and store of a step along the line is one coalesced row segment.
//
//   march of a slice          src/sim_kernels.hpp:213-313
//   current line filters      src/sim_particles.cpp:160-204
//   sponge bound potential    src/sim_fields.cpp:202-268
//   co-moving field           src/sim_fields.cpp:55-69, 79-147
//   plasma pass per line      src/sim_plasma.cpp:119-184 (filters, smoothing), 300-396 (bookkeeping, march)
#include "g3d_pass.hpp"
#include "g3d_runtime.hpp"

namespace g3d {
namespace {
constexpr int kT = 64;

// ---------------------------------------------------------------- march of one line
// Slice n of line ik at level s1 by the line's workgroup (kLW lanes), with the reference's operations and operand
// order (bit-identical results).  Of the march only the two tridiagonal eliminations (E_z on the nodes 1 .. ny-1,
// B_z on the half-nodes 0 .. ny-1) and their back substitutions are recurrences along y; they run on one lane each
// (lanes 0 and 1 together) over coefficients in LDS.  Everything else is per node and spread over the lanes:
//   1. the explicit parts of R (from slice n at s0) and of L (from slice n+1 at s1);
//   2. from those of the node and its neighbours, the right-hand sides and the fixed parts of the pivots;
//   3. (lanes 0, 1) the eliminations and back substitutions;
//   4. E_z, B_z and R, L of every node.
// Every value of slice n at s0 is read (1) before any is rewritten (4).
constexpr int kLW = 256;
// per node (half-node for B_z): the explicit parts, the right-hand sides (then the eliminated ones, then the
// solution), the fixed pivot parts and the off-diagonals (the upper one then the elimination factor)
struct LineLds {
    cx *sRx, *sRy, *sLx, *sLy, *rE, *rB;
    double *bE, *lE, *uE, *bB, *lB, *uB;
};
constexpr size_t kLineLdsNode = 6 * sizeof(cx) + 6 * sizeof(double);
__device__ LineLds line_lds(void* base, int ny) {
    cx* z = reinterpret_cast<cx*>(base);
    LineLds s;
    s.sRx = z; s.sRy = z + ny; s.sLx = z + 2 * ny; s.sLy = z + 3 * ny; s.rE = z + 4 * ny; s.rB = z + 5 * ny;
    double* d = reinterpret_cast<double*>(z + 6 * ny);
    s.bE = d; s.lE = d + ny; s.uE = d + 2 * ny; s.bB = d + 3 * ny; s.lB = d + 4 * ny; s.uB = d + 5 * ny;
    return s;
}
// the lanes of the workgroup have written global memory that the other lanes read next
__device__ __forceinline__ void line_sync() {
    __threadfence_block();
    __syncthreads();
}

// One tridiagonal elimination and back substitution over LDS on nodes j0 .. n-1: lower off-diagonal lo, pivot
// part bs, upper off-diagonal up, right-hand side v (replaced by the solution); the elimination factors go to cpo
// (which may be up itself).  den = bs - lo c'(j-1), v'(j) = (v(j) - lo v'(j-1)) / den, c'(j) = up / den, then
// v(j) = v'(j) - c'(j) v(j+1).  The operands of the next node are loaded while this one is eliminated, so that the
// recurrence waits on its own arithmetic only.
__device__ __forceinline__ void lds_thomas(cx* v, const double* lo, const double* bs, const double* up, double* cpo,
                                           int j0, int n) {
    if (j0 >= n) return;
    double cv = 0;
    cx dv = mk(0);
    double l = lo[j0], b = bs[j0], u = up[j0];
    cx r = v[j0];
    for (int j = j0; j < n; ++j) {
        const int jn = j + 1 < n ? j + 1 : j;
        const double ln = lo[jn], bn = bs[jn], un = up[jn];
        const cx rn = v[jn];
        const double den = b - l * cv;
        dv = (r - l * dv) / den;
        cv = u / den;
        v[j] = dv; cpo[j] = cv;
        l = ln; b = bn; u = un; r = rn;
    }
    cx nv = mk(0);
    cx rj = v[n - 1];
    double cj = cpo[n - 1];
    for (int j = n - 1; j >= j0; --j) {
        const int jp = j > j0 ? j - 1 : j;
        const cx rp = v[jp];
        const double cp = cpo[jp];
        nv = rj - cj * nv;
        v[j] = nv;
        rj = rp; cj = cp;
    }
}

// The coefficients of the march's systems at node j (src/sim_kernels.hpp:262-279): E_z on the nodes 1 .. ny-1
// (lower, upper off-diagonal, pivot part), B_z on the half-nodes 0 .. ny-1; cc = h/4 (D + h/4), kx2 = k_x^2
__device__ __forceinline__ void coef_E(const double* iA, const double* iD, int ny, double cc, double kx2, int j,
                                       double& lo, double& up, double& bs) {
    lo = j > 1 ? -cc * iA[j] * iD[j - 1] : 0;
    up = j < ny - 1 ? -cc * iA[j] * iD[j] : 0;
    bs = 1 + cc * kx2 + cc * iA[j] * (iD[j - 1] + iD[j]);
}
__device__ __forceinline__ void coef_B(const double* iA, const double* iD, int ny, double cc, double kx2, int j,
                                       double& lo, double& up, double& bs) {
    const double al = j > 0 ? iA[j] : 0, au = j < ny - 1 ? iA[j + 1] : 0;
    lo = -cc * iD[j] * al;
    up = -cc * iD[j] * au;
    bs = 1 + cc * kx2 + cc * iD[j] * (al + au);
}

// lds_thomas with the pivots den(j) = bs - lo c'(j-1) and the factors c'(j) given (bs, cp): the same operations on
// the right-hand side, one complex division per node
__device__ __forceinline__ void lds_thomas_fixed(cx* v, const double* lo, const double* den, const double* cp, int j0,
                                                 int n) {
    if (j0 >= n) return;
    cx dv = mk(0);
    double l = lo[j0], dn = den[j0];
    cx r = v[j0];
    for (int j = j0; j < n; ++j) {
        const int jn = j + 1 < n ? j + 1 : j;
        const double ln = lo[jn], dnn = den[jn];
        const cx rn = v[jn];
        dv = (r - l * dv) / dn;
        v[j] = dv;
        l = ln; dn = dnn; r = rn;
    }
    cx nv = mk(0);
    cx rj = v[n - 1];
    double cj = cp[n - 1];
    for (int j = n - 1; j >= j0; --j) {
        const int jp = j > j0 ? j - 1 : j;
        const cx rp = v[jp];
        const double c = cp[jp];
        nv = rj - cj * nv;
        v[j] = nv;
        rj = rp; cj = c;
    }
}

__device__ void march_line(const LineView& L, int ik, int n, const LineLds& s) {
    const MarchView& c = L.mv;
    const int nx = c.nx, ny = L.ny, t0 = threadIdx.x;
    const long nk = c.nk;
    const long P0 = long(n - c.a) * nk + ik;        // slice n = plane of half-cell n-1 in the currents
    const long P1 = P0 + nk;                        // slice n+1 = half-cell n
    const cx a = mk(0, c.kx[ik]);
    const double kx2 = c.kx[ik] * c.kx[ik], D = c.D, wl = c.wlo[n], wu = c.whi[n];
    const double* iD = L.iD;
    const double* iA = L.iA;
    // explicit part of R at node / half-node j from slice n at level s0 and the currents of both levels
    auto known_R = [&](int j, cx ez0, cx ezu, cx bz0, cx bzm, cx& sx, cx& sy) {
        const long o = P0 + long(j) * nx;
        if (j == 0) sx = mk(0);
        else {
            const cx Jx = wl * (c.j0x[o] + c.jkx[o]) + wu * (c.j0x[o + nk] + c.jkx[o + nk]);
            const cx bm = (bz0 - bzm) * iA[j];
            sx = c.Rx[o] + D * (a * ez0 + bm) - D * Jx;
        }
        const cx Jy = wl * (c.j0y[o] + c.jky[o]) + wu * (c.j0y[o + nk] + c.jky[o + nk]);
        sy = c.Ry[o] + D * ((ezu - ez0) * iD[j] - a * bz0) - D * Jy;
    };
    auto explicit_R = [&](int j) {
        const long o = P0 + long(j) * nx;
        const cx ez0 = c.Ez[o], ezu = j + 1 < ny ? c.Ez[o + nx] : mk(0), bz0 = c.Bz[o];
        const cx bzm = j > 0 ? c.Bz[o - nx] : mk(0);
        known_R(j, ez0, ezu, bz0, bzm, s.sRx[j], s.sRy[j]);
    };
    if (n == c.N - 1) {                              // front of the window: L = E_z = B_z = 0
        for (int j = t0; j < ny; j += kLW) explicit_R(j);
        line_sync();
        for (int j = t0; j < ny; j += kLW) {
            const long o = P0 + long(j) * nx;
            c.Rx[o] = s.sRx[j]; c.Ry[o] = s.sRy[j];
            c.Ez[o] = c.Bz[o] = c.Lx[o] = c.Ly[o] = mk(0);
        }
        line_sync();
        return;
    }
    const double h = c.hcg[n + 2], qh = 0.25 * h, hh = 0.5 * h, cc = qh * (D + qh);
    const long pv = long(n) * 4 * ny * nx + ik;        // this line's pivots of slice n (L.piv)
    // 1. explicit parts
    for (int j = t0; j < ny; j += kLW) {
        explicit_R(j);
        const long u = P1 + long(j) * nx;
        const cx pBz = c.Bz[u], pBzm = j > 0 ? c.Bz[u - nx] : mk(0);
        const cx pEz = j > 0 ? c.Ez[u] : mk(0), pEzu = j + 1 < ny ? c.Ez[u + nx] : mk(0);
        if (j == 0) s.sLx[j] = mk(0);
        else {
            const cx bm = (pBz - pBzm) * iA[j];
            s.sLx[j] = c.Lx[u] - qh * (a * pEz - bm) - hh * c.jkx[u];
        }
        const cx bp = (pEzu - pEz) * iD[j];
        s.sLy[j] = c.Ly[u] - qh * (a * pBz + bp) - hh * c.jky[u];
    }
    __syncthreads();
    // 2. right-hand sides and fixed pivot parts: E_z at node j (j >= 1), B_z at half-node j
    for (int j = t0; j < ny; j += kLW) {
        const long u = P1 + long(j) * nx;
        const cx pRx = c.Rx[u], pRy = c.Ry[u], pLx = c.Lx[u], pLy = c.Ly[u];
        const cx sRx = s.sRx[j], sRy = s.sRy[j], sLx = s.sLx[j], sLy = s.sLy[j];
        if (j >= 1) {
            const cx pRym = c.Ry[u - nx], pLym = c.Ly[u - nx];
            const cx sRym = s.sRy[j - 1], sLym = s.sLy[j - 1];
            const cx pEz = c.Ez[u];
            const cx Qp = 0.5 * (-(a * (pRx - pLx)) + ((pLy - pRy) - (pLym - pRym)) * iA[j]);
            const cx SE = -(a * (sRx - sLx)) + ((sLy - sRy) - (sLym - sRym)) * iA[j];
            s.rE[j] = pEz - hh * Qp - qh * SE - h * c.jkz[u];
            coef_E(iA, iD, ny, cc, kx2, j, s.lE[j], s.uE[j], s.bE[j]);
            if (L.piv) { s.bE[j] = L.piv[pv + long(j) * nx]; s.uE[j] = L.piv[pv + long(ny + j) * nx]; }
        }
        const bool top = j + 1 >= ny;
        const cx fu = top ? mk(0) : c.Rx[u + nx] + c.Lx[u + nx], f0 = pRx + pLx;
        const cx su = top ? mk(0) : s.sRx[j + 1] + s.sLx[j + 1], s0 = sRx + sLx;
        const cx Cp = 0.5 * (a * (pRy + pLy) - (fu - f0) * iD[j]);
        const cx SB = a * (sRy + sLy) - (su - s0) * iD[j];
        s.rB[j] = c.Bz[u] - hh * Cp - qh * SB;
        coef_B(iA, iD, ny, cc, kx2, j, s.lB[j], s.uB[j], s.bB[j]);
        if (L.piv) { s.bB[j] = L.piv[pv + long(2 * ny + j) * nx]; s.uB[j] = L.piv[pv + long(3 * ny + j) * nx]; }
    }
    __syncthreads();
    // 3. the recurrences, side by side in one instruction stream: lane 0 E_z (nodes 1 ..), lane 1 B_z (half-nodes
    //    0 ..); the solution replaces the eliminated right-hand side
    if (t0 < 2) {
        const bool hb = t0 == 1;
        const int j0 = hb ? 0 : 1;
        cx* d = hb ? s.rB : s.rE;
        const double *bs = hb ? s.bB : s.bE, *ls = hb ? s.lB : s.lE;
        double* cp = hb ? s.uB : s.uE;
        if (L.piv) lds_thomas_fixed(d, ls, bs, cp, j0, ny);
        else lds_thomas(d, ls, bs, cp, cp, j0, ny);
        if (!hb) s.rE[0] = mk(0);
    }
    __syncthreads();
    // 4. E_z, B_z and R, L of every node
    for (int j = t0; j < ny; j += kLW) {
        const long o = P0 + long(j) * nx;
        const cx ez = s.rE[j], ezu = j + 1 < ny ? s.rE[j + 1] : mk(0), bz = s.rB[j];
        const cx bzl = j > 0 ? s.rB[j - 1] : mk(0);
        c.Ez[o] = ez; c.Bz[o] = bz;
        if (j == 0) { c.Rx[o] = mk(0); c.Lx[o] = mk(0); }
        else {
            const cx bm = (bz - bzl) * iA[j];
            c.Rx[o] = s.sRx[j] + D * (a * ez + bm);
            c.Lx[o] = s.sLx[j] - qh * (a * ez - bm);
        }
        const cx bp = (ezu - ez) * iD[j];
        c.Ry[o] = s.sRy[j] + D * (bp - a * bz);
        c.Ly[o] = s.sLy[j] - qh * (a * bz + bp);
    }
    line_sync();
}

// the k_x line of workgroup b: workgroup b runs on XCD b mod 8, so with xcd the lines of one XCD are consecutive
// (neighbouring lines share the cache lines of every access strided by nx, and each XCD has its own L2)
__device__ __forceinline__ int line_of(int b, int nb, int xcd) {
    return (xcd && (nb & 7) == 0) ? (b & 7) * (nb >> 3) + (b >> 3) : b;
}

// slices hi-1 ... lo of line line_of(blockIdx.x)
__global__ void __launch_bounds__(kLW) k_march_lines(LineView L, int lo, int hi, int xcd) {
    extern __shared__ double lds_line[];
    const LineLds s = line_lds(lds_line, L.ny);
    const int ik = line_of(blockIdx.x, gridDim.x, xcd);
    for (int n = hi - 1; n >= lo; --n) march_line(L, ik, n, s);
}

// ---------------------------------------------------------------- filters of a line of currents
// (1,2,1)/4 along y of the charges q_j = A_j / w_j, out_j = (q_{j-1}/4 + q_j/2 + q_{j+1}/4) w_j, streamed one
// node per step.  Nodes (w = 1/A): q = 0 beyond the walls, node 0 untouched; half-nodes (w = 1/Delta): mirrored.
struct Filter121 {
    cx qm, q0;
    __device__ cx step(cx qp, double w) {
        const cx v = (0.25 * qm + 0.5 * q0 + 0.25 * qp) * w;
        qm = q0; q0 = qp;
        return v;
    }
};

// one line of the half-cell currents (src/sim_particles.cpp:160-204): unpack of the pair per row, the filters,
// the x filter, node 0 and the k_x = 0 mean of j_y
__global__ void __launch_bounds__(kT) k_fh_currents(const cx* S1, cx* jx, cx* jy, cx* jz, int planes, LineView L,
                                                   const double* fl, int kf, int keep0, double Ly) {
    const long line = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const int nx = L.mv.nx, ny = L.ny;
    const long nk = L.mv.nk;
    if (line >= long(planes) * nx) return;
    const int ixx = int(line % nx), mx = (nx - ixx) % nx;
    const long base = (line / nx) * nk, o = base + ixx;
    const double* ia = L.iA;
    const double* id = L.iD;
    auto unpack = [&](int j, cx& ux, cx& uy) {
        const long k = base + long(j) * nx;
        const cx z = S1[k + ixx], zm = conjg(S1[k + mx]);
        const cx d = z - zm;
        ux = 0.5 * (z + zm);
        uy = 0.5 * mk(d.im, -d.re);
    };
    const double fx = fl[ixx];
    const bool with_mean = ixx == 0 && !keep0;
    cx ax, ay, az = jz[o];
    unpack(0, ax, ay);
    Filter121 fxn{mk(0), mk(0)}, fzn{mk(0), mk(0)}, fyh{mk(0), mk(0)};
    if (kf) { fyh.qm = ay / id[0]; fyh.q0 = fyh.qm; }
    cx mean = mk(0);
    for (int j = 0; j < ny; ++j) {
        const bool more = j + 1 < ny;
        cx bx = mk(0), by = mk(0), bz = mk(0);
        if (more) { unpack(j + 1, bx, by); bz = jz[o + long(j + 1) * nx]; }
        cx vx = ax, vy = ay, vz = az;
        if (kf) {
            vy = fyh.step(more ? by / id[j + 1] : fyh.q0, id[j]);
            if (j == 0) { fxn.q0 = bx / ia[1]; fzn.q0 = bz / ia[1]; }
            else {
                vx = fxn.step(more ? bx / ia[j + 1] : mk(0), ia[j]);
                vz = fzn.step(more ? bz / ia[j + 1] : mk(0), ia[j]);
            }
        }
        const long k = o + long(j) * nx;
        jx[k] = fx * (j == 0 ? mk(0) : vx);
        jz[k] = fx * (j == 0 ? mk(0) : vz);
        if (with_mean) { jy[k] = vy; mean = mean + vy / id[j]; }
        else jy[k] = fx * (vy - mk(0));
        ax = bx; ay = by; az = bz;
    }
    if (with_mean) {
        mean = mean / Ly;
        for (int j = 0; j < ny; ++j) {
            const long k = o + long(j) * nx;
            jy[k] = fx * (jy[k] - mean);
        }
    }
}

// ---------------------------------------------------------------- sponge (src/sim_fields.cpp:202-242, 259-268)
// psi of line (slice l, k_x) = T^-1 (rho + gamma^-2 T^-1 d_zeta^2 rho), T = k_x^2 - Lap_y with its LU factors
// (pc, pd), Dirichlet at the walls; S1 = (R - R_b) packed, R_b = (1 + beta) E_perp of psi
__global__ void __launch_bounds__(kT) k_fh_sponge_psi(FhSpongeArgs s, LineView L, const cx* rx, const cx* ry, cx* s1,
                                                     cx* s2) {
    const int nx = L.mv.nx, ny = L.ny;
    const long nk = L.mv.nk;
    const long line = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (line >= long(s.ns) * nx) return;
    const int ixx = int(line % nx), l = int(line / nx), n = s.a + l;
    const long o = long(l) * nk + ixx;
    const double* ia = L.iA;
    const double* id = L.iD;
    // half-cell planes; beyond the window (no neighbouring blocks on one GPU) the charge is zero
    auto half = [&](int hk, long jo) -> cx { return (hk < 0 || hk > s.ns) ? mk(0) : s.jz[long(hk) * nk + ixx + jo]; };
    auto rho_at = [&](int dl, long jo) -> cx {
        const int nn = n + dl;
        if (nn < 0 || nn > s.N - 1) return mk(0);
        return s.wl[nn] * half(l + dl, jo) + s.wu[nn] * half(l + dl + 1, jo);
    };
    const double hm = s.hc[n + 1], hp = s.hc[n + 2];
    const double cm = 2 / ((hm + hp) * hm), cp_ = 2 / ((hm + hp) * hp);
    for (int second = 0; second < 2; ++second) {
        cx dpv = mk(0);
        for (int j = 1; j < ny; ++j) {
            const long jo = long(j) * nx, k = o + jo;
            cx r;
            if (!second) r = s.g2 * (cp_ * rho_at(1, jo) - (cm + cp_) * rho_at(0, jo) + cm * rho_at(-1, jo));
            else r = rho_at(0, jo) + s2[k];
            const double lo = j > 1 ? -ia[j] * id[j - 1] : 0;
            dpv = (r - lo * dpv) * s.pd[long(j) * nx + ixx];
            s2[k] = dpv;
        }
        cx next = mk(0);
        for (int j = ny - 1; j >= 1; --j) {
            next = s2[o + long(j) * nx] - s.pc[long(j) * nx + ixx] * next;
            s2[o + long(j) * nx] = next;
        }
    }
    s2[o] = mk(0);
    const cx a = mk(0, s.kx[ixx]);
    for (int j = 0; j < ny; ++j) {
        const long k = o + long(j) * nx;
        const cx pu = j + 1 < ny ? s2[k + nx] : mk(0);
        const cx bx = -s.opb * (a * s2[k]), by = -s.opb * ((pu - s2[k]) * id[j]);
        const cx u = rx[k] - bx, v = ry[k] - by;
        s1[k] = mk(u.re - v.im, u.im + v.re);
    }
}
// the sponged R - R_b unpacked per row, R_b added back (node 0: R_x = 0)
__global__ void k_fh_sponge_unpack(FhSpongeArgs s, LineView L, const cx* s1, const cx* s2, cx* rx, cx* ry, long tot) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const int nx = L.mv.nx, ny = L.ny;
    const long i = idx % L.mv.nk;
    const int j = int(i / nx), ixx = int(i % nx);
    const long m = idx - ixx + (nx - ixx) % nx;
    const cx z = s1[idx], zm = conjg(s1[m]);
    const cx pu = j + 1 < ny ? s2[idx + nx] : mk(0);
    const cx bx = -s.opb * (mk(0, s.kx[ixx]) * s2[idx]), by = -s.opb * ((pu - s2[idx]) * L.iD[j]);
    const cx d = z - zm;
    rx[idx] = j == 0 ? mk(0) : 0.5 * (z + zm) + bx;
    ry[idx] = 0.5 * mk(d.im, -d.re) + by;
}

// ---------------------------------------------------------------- co-moving field
// coefficient s of the eigenvector expansion of the node values (slot 0 unused: zero)
__global__ void k_fh_eig_fwd(const cx* jz, const double* ef, cx* e1, long tot, int nx, int ny, long nk) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % nk, base = idx - i;
    const int s = int(i / nx), ixx = int(i % nx);
    cx v = mk(0);
    if (s > 0)
        for (int j = 1; j < ny; ++j) v = v + ef[long(s) * ny + j] * jz[base + long(j) * nx + ixx];
    e1[idx] = v;
}
// node value j from the coefficients (node 0 on the wall: zero)
__global__ void k_fh_eig_back(const cx* c, const double* eb, cx* out, long tot, int nx, int ny, long nk) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % nk, base = idx - i;
    const int j = int(i / nx), ixx = int(i % nx);
    cx v = mk(0);
    if (j > 0)
        for (int s = 1; s < ny; ++s) v = v + eb[long(j) * ny + s] * c[base + long(s) * nx + ixx];
    out[idx] = v;
}
__global__ void k_fh_comoving_R(const cx* psi, const double* kx, const double* id, double beta, cx* rx, cx* ry, cx* ez,
                                cx* bz, long tot, int nx, int ny, long nk) {
    const long idx = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= tot) return;
    const long i = idx % nk;
    const int j = int(i / nx), ixx = int(i % nx);
    rx[idx] = -(1 + beta) * (mk(0, kx[ixx]) * psi[idx]);
    const cx up = j + 1 < ny ? psi[idx + nx] : mk(0);
    ry[idx] = -(1 + beta) * ((up - psi[idx]) * id[j]);
    ez[idx] = mk(0); bz[idx] = mk(0);
}

// ---------------------------------------------------------------- one pass of a plasma half-cell, per line
// Per line ik (one workgroup of kLW lanes): the y filters, the x filter, node 0 and the smoothing
// (1 + a^2 (k_x^2 - Lap_y))^-1 of the streaming current of this pass (X: the per-row unpacked x transforms, modified
// in place); kind 2 stops there (jsn).  Otherwise per point the bookkeeping of the pass and the pass backup of slice
// n, the march of the line, on the last pass the correction of R at slice n+1, and the three packed planes of
// slice n for the inverse transforms.  The per-node work is spread over the lanes; the smoothing eliminations are
// recurrences along y and run one system per lane (three lanes).
__global__ void __launch_bounds__(kLW) k_pass_lines(ModeArgs m, LineView L, cx* X, cx* T, cx* jsn, int kind,
                                                    int xcd) {
    extern __shared__ double lds_line[];
    const LineLds s = line_lds(lds_line, L.ny);
    const int ik = line_of(blockIdx.x, gridDim.x, xcd), t0 = threadIdx.x;
    const MarchView& c = m.mv;
    const int nx = c.nx, ny = L.ny;
    const long nk = c.nk;
    const int n = m.n;
    const double* ia = L.iA;
    const double* id = L.iD;
    cx* ox = X;
    cx* oy = X + nk;
    cx* oz = X + 2 * nk;
    const long o = ik;
    auto at = [&](cx* A, int j) -> cx& { return A[o + long(j) * nx]; };
    if (kind != 1) {
        const double fx = m.filt[ik];
        if (m.kf) {
            // (1,2,1)/4 of the charges q = value / weight: the charges of every node first (they are read by the
            // neighbours), q = 0 beyond the walls for the nodes, mirrored for the half-nodes
            cx* qx = s.sRx;
            cx* qy = s.sRy;
            cx* qz = s.sLx;
            for (int j = t0; j < ny; j += kLW) {
                qy[j] = at(oy, j) / id[j];
                if (j >= 1) { qx[j] = at(ox, j) / ia[j]; qz[j] = at(oz, j) / ia[j]; }
            }
            __syncthreads();
            for (int j = t0; j < ny; j += kLW) {
                const cx ym = j > 0 ? qy[j - 1] : qy[0], yp = j + 1 < ny ? qy[j + 1] : qy[j];
                at(oy, j) = fx * ((0.25 * ym + 0.5 * qy[j] + 0.25 * yp) * id[j]);
                if (j >= 1) {
                    const cx xm = j > 1 ? qx[j - 1] : mk(0), xp = j + 1 < ny ? qx[j + 1] : mk(0);
                    const cx zm = j > 1 ? qz[j - 1] : mk(0), zp = j + 1 < ny ? qz[j + 1] : mk(0);
                    at(ox, j) = fx * ((0.25 * xm + 0.5 * qx[j] + 0.25 * xp) * ia[j]);
                    at(oz, j) = fx * ((0.25 * zm + 0.5 * qz[j] + 0.25 * zp) * ia[j]);
                }
            }
        } else
            for (int j = t0; j < ny; j += kLW) {
                if (j >= 1) { at(ox, j) = fx * at(ox, j); at(oz, j) = fx * at(oz, j); }
                at(oy, j) = fx * at(oy, j);
            }
        if (t0 == 0) { at(ox, 0) = fx * mk(0); at(oz, 0) = fx * mk(0); }
        line_sync();
        const double a2 = m.a2;
        if (a2 > 0) {
            // the three columns and the coefficients of their systems into LDS, over the lanes (the filters' charge
            // arrays are free again): node coefficients for j_x and j_z (zero at the walls, as E_z in the march),
            // half-node ones for j_y (as B_z in the march)
            const double kx2 = m.kxs[ik] * m.kxs[ik];
            cx *vx = s.rE, *vz = s.rB, *vy = s.sLy;
            for (int j = t0; j < ny; j += kLW) {
                vx[j] = at(ox, j); vz[j] = at(oz, j); vy[j] = at(oy, j);
                if (j >= 1) {
                    s.lE[j] = j > 1 ? -a2 * ia[j] * id[j - 1] : 0;
                    s.uE[j] = j < ny - 1 ? -a2 * ia[j] * id[j] : 0;
                    s.bE[j] = 1 + a2 * kx2 + a2 * ia[j] * (id[j - 1] + id[j]);
                }
                const double al = j > 0 ? ia[j] : 0, au = j < ny - 1 ? ia[j + 1] : 0;
                s.lB[j] = -a2 * id[j] * al;
                s.uB[j] = -a2 * id[j] * au;
                s.bB[j] = 1 + a2 * kx2 + a2 * id[j] * (al + au);
            }
            __syncthreads();
            // lane 0: j_x, lane 1: j_z (nodes 1 .. ny-1, the elimination factors of each in its own array), lane 2:
            // j_y (half-nodes 0 .. ny-1)
            if (t0 < 2)
                lds_thomas(t0 ? vz : vx, s.lE, s.bE, s.uE, reinterpret_cast<double*>(t0 ? s.sRy : s.sRx), 1, ny);
            else if (t0 == 2)
                lds_thomas(vy, s.lB, s.bB, s.uB, s.uB, 0, ny);
            __syncthreads();
            for (int j = t0; j < ny; j += kLW) { at(ox, j) = vx[j]; at(oz, j) = vz[j]; at(oy, j) = vy[j]; }
            line_sync();
        }
        if (kind == 2) {
            for (int j = t0; j < ny; j += kLW) {
                const long i = o + long(j) * nx;
                jsn[i] = ox[i]; jsn[nk + i] = oy[i]; jsn[2 * nk + i] = oz[i];
            }
            return;
        }
    }
    for (int j = t0; j < ny; j += kLW) {
        const long i = o + long(j) * nx, so = long(n - c.a) * nk + i;
        cx u[3] = {mk(0), mk(0), mk(0)};
        if (kind != 1) { u[0] = ox[i]; u[1] = oy[i]; u[2] = oz[i]; }
        pass_currents(m, i, u);
        if (m.bRx) {                                 // slice n at s0: saved on the first pass, restored after
            if (m.first) { m.bRx[i] = c.Rx[so]; m.bRy[i] = c.Ry[so]; m.bEz[i] = c.Ez[so]; m.bBz[i] = c.Bz[so]; }
            else { c.Rx[so] = m.bRx[i]; c.Ry[so] = m.bRy[i]; c.Ez[so] = m.bEz[i]; c.Bz[so] = m.bBz[i]; }
        }
    }
    line_sync();
    march_line(L, ik, n, s);
    for (int j = t0; j < ny; j += kLW) {
        const long i = o + long(j) * nx, so = long(n - c.a) * nk + i;
        if (m.last && !m.top) {
            const cx u[3] = {ox[i], oy[i], oz[i]};
            pass_correct(m, i, u);
        }
        cx t[3];
        pack3(t, c.Rx[so], c.Ry[so], c.Lx[so], c.Ly[so], c.Ez[so], c.Bz[so]);
        T[i] = t[0]; T[nk + i] = t[1]; T[2 * nk + i] = t[2];
    }
}

// thread (ik, n): the pivots and elimination factors of line ik, slice n, as march_line's eliminations form them
__global__ void __launch_bounds__(64) k_line_pivots(LineView L, double* piv) {
    const MarchView& c = L.mv;
    const int ik = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y, nx = c.nx, ny = L.ny;
    if (ik >= nx) return;
    const double h = c.hcg[n + 2], qh = 0.25 * h, cc = qh * (c.D + qh), kx2 = c.kx[ik] * c.kx[ik];
    double* p = piv + long(n) * 4 * ny * nx + ik;
    double cv = 0;
    for (int j = 1; j < ny; ++j) {
        double lo, up, bs;
        coef_E(L.iA, L.iD, ny, cc, kx2, j, lo, up, bs);
        const double den = bs - lo * cv;
        cv = up / den;
        p[long(j) * nx] = den; p[long(ny + j) * nx] = cv;
    }
    cv = 0;
    for (int j = 0; j < ny; ++j) {
        double lo, up, bs;
        coef_B(L.iA, L.iD, ny, cc, kx2, j, lo, up, bs);
        const double den = bs - lo * cv;
        cv = up / den;
        p[long(2 * ny + j) * nx] = den; p[long(3 * ny + j) * nx] = cv;
    }
}

template <class K, class... A> void launch(K k, long n, int t, const char* what, A... a) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k, dim3(blocks_for(n, t)), dim3(t), 0, stream(), a...);
    launch_check(what);
}
}  // namespace

// one workgroup per line with the line's LDS
static size_t line_lds_bytes(int ny) {
    const size_t b = size_t(ny) * kLineLdsNode;
    if (b > 64 * 1024) throw std::runtime_error("free-hand lines: ny " + std::to_string(ny) + " above the LDS of a line");
    return b;
}
void line_pivots(const LineView& L, double* piv) {
    if (L.mv.N < 2 || L.mv.nx <= 0) return;
    hipLaunchKernelGGL(k_line_pivots, dim3(unsigned((L.mv.nx + 63) / 64), unsigned(L.mv.N - 1)), dim3(64), 0, stream(),
                       L, piv);
    launch_check("line pivots");
}
void march_lines(const LineView& L, int lo, int hi) {
    if (L.mv.nx <= 0 || hi <= lo) return;
    hipLaunchKernelGGL(k_march_lines, dim3(unsigned(L.mv.nx)), dim3(kLW), line_lds_bytes(L.ny), stream(), L, lo, hi,
                       xcd_rows());
    launch_check("march lines");
}
void fh_currents(const cx* S1, cx* jkx, cx* jky, cx* jkz, int planes, const LineView& L, const double* filt, int kf,
                 int keep0, double Ly) {
    launch(k_fh_currents, long(planes) * L.mv.nx, kT, "free-hand currents", S1, jkx, jky, jkz, planes, L, filt, kf,
           keep0, Ly);
}
void fh_sponge_psi(const FhSpongeArgs& s, const LineView& L, const cx* Rx, const cx* Ry, cx* S1, cx* S2) {
    launch(k_fh_sponge_psi, long(s.ns) * L.mv.nx, kT, "free-hand sponge", s, L, Rx, Ry, S1, S2);
}
void fh_sponge_unpack(const FhSpongeArgs& s, const LineView& L, const cx* S1, const cx* S2, cx* Rx, cx* Ry) {
    const long tot = long(s.ns) * L.mv.nk;
    launch(k_fh_sponge_unpack, tot, 256, "free-hand sponge unpack", s, L, S1, S2, Rx, Ry, tot);
}
void fh_eigen_forward(const cx* jz, const double* evf, cx* e1, int planes, int nx, int ny, long nk) {
    const long tot = long(planes) * nk;
    launch(k_fh_eig_fwd, tot, 256, "eigen forward", jz, evf, e1, tot, nx, ny, nk);
}
void fh_eigen_back(const cx* cf, const double* evb, cx* out, int planes, int nx, int ny, long nk) {
    const long tot = long(planes) * nk;
    launch(k_fh_eig_back, tot, 256, "eigen back", cf, evb, out, tot, nx, ny, nk);
}
void fh_comoving_R(const cx* psi, const double* kx, const double* iD, double beta, cx* Rx, cx* Ry, cx* Ez, cx* Bz,
                   int planes, int nx, int ny, long nk) {
    const long tot = long(planes) * nk;
    launch(k_fh_comoving_R, tot, 256, "free-hand comoving R", psi, kx, iD, beta, Rx, Ry, Ez, Bz, tot, nx, ny, nk);
}
void pass_lines(const ModeArgs& m, const LineView& L, cx* X, cx* T, cx* jsn, int kind) {
    if (m.mv.nx <= 0) return;
    hipLaunchKernelGGL(k_pass_lines, dim3(unsigned(m.mv.nx)), dim3(kLW), line_lds_bytes(L.ny), stream(), m, L, X, T,
                       jsn, kind, xcd_rows());
    launch_check("pass lines");
}

}  // namespace g3d
