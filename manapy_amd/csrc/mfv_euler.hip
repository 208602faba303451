// 3D compressible Euler, first-order Rusanov on tetrahedra (solvers/euler/fvm_utils3d.py).
//
// State layout: [6][C] = rho, P, rhou, rhov, rhow, rhoE. Ghost layout: [6][F].
// quirk = 1 reproduces the reference's two slips on this path:
//   Q1 compute_flux_euler_3d_rusanov (:842): qr = rhou_R n0 + rhov_L n1 + rhow_L n2;
//   Q2 ghost_value_TubeSchok3D (:493-494): rhovg is overwritten by rho*w_g, rhowg never set
//      (it keeps its previous content: zero for zero-initialised ghost arrays).
// quirk = 0 uses the right state in qr and the mirrored velocity in both components.
#include <stdlib.h>

#include <utility>
#include <vector>

#include "mfv_common.hpp"

using namespace mfv;

namespace {

struct Prim6 {
    double rho, P, ru, rv, rw, rE;
};

__device__ __forceinline__ void euler_flux(const Prim6 &L, const Prim6 &R, const double *n, double mesure,
                                           double gamma, int quirk, double (&fl)[5])
{
    const double rhol = L.rho, Pl = L.P, rhoul = L.ru, rhovl = L.rv, rhowl = L.rw, rhoEl = L.rE;
    const double rhor = R.rho, Pr = R.P, rhour = R.ru, rhovr = R.rv, rhowr = R.rw, rhoEr = R.rE;
    double ql = (rhoul * n[0] + rhovl * n[1] + rhowl * n[2]);
    double qr = quirk ? (rhour * n[0] + rhovl * n[1] + rhowl * n[2]) : (rhour * n[0] + rhovr * n[1] + rhowr * n[2]);
    double cl = sqrt(gamma * Pl / rhol);
    double cr = sqrt(gamma * Pr / rhor);
    double fl_rho = (rhoul * n[0] + rhovl * n[1] + rhowl * n[2]);
    double fl_rhou = ((rhoul * rhoul / rhol + Pl) * n[0] + (rhoul * rhovl / rhol) * n[1] + (rhoul * rhowl / rhol) * n[2]);
    double fl_rhov = ((rhovl * rhoul / rhol) * n[0] + (rhovl * rhovl / rhol + Pl) * n[1] + (rhovl * rhowl / rhol) * n[2]);
    double fl_rhow = ((rhoul * rhowl / rhol) * n[0] + (rhovl * rhowl / rhol) * n[1] + (rhowl * rhowl / rhol + Pl) * n[2]);
    double fl_rhoE =
        (rhoul / rhol * (rhoEl + Pl) * n[0] + rhovl / rhol * (rhoEl + Pl) * n[1] + rhowl / rhol * (rhoEl + Pl) * n[2]);
    double fr_rho = (rhour * n[0] + rhovr * n[1] + rhowr * n[2]);
    double fr_rhou = ((rhour * rhour / rhor + Pr) * n[0] + (rhour * rhovr / rhor) * n[1] + (rhour * rhowr / rhor) * n[2]);
    double fr_rhov = ((rhovr * rhour / rhor) * n[0] + (rhovr * rhovr / rhor + Pr) * n[1] + (rhovr * rhowr / rhor) * n[2]);
    double fr_rhow = ((rhour * rhowr / rhor) * n[0] + (rhovr * rhowr / rhor) * n[1] + (rhowr * rhowr / rhor + Pr) * n[2]);
    double fr_rhoE =
        (rhour / rhor * (rhoEr + Pr) * n[0] + rhovr / rhor * (rhoEr + Pr) * n[1] + rhowr / rhor * (rhoEr + Pr) * n[2]);
    double l1 = fabs((ql) / rhol - cl), l2 = fabs((ql) / rhol), l3 = fabs((ql) / rhol + cl);
    double r1 = fabs((qr) / rhor - cr), r2 = fabs((qr) / rhor), r3 = fabs((qr) / rhor + cr);
    double Ll = pmax(pmax(l1, l2), l3);
    double Lr = pmax(pmax(r1, r2), r3);
    double S = (Ll > Lr) ? Ll : Lr;
    fl[0] = (0.5 * (fl_rho + fr_rho) - 0.5 * S * (rhor - rhol)) * mesure;
    fl[1] = (0.5 * (fl_rhou + fr_rhou) - 0.5 * S * (rhour - rhoul)) * mesure;
    fl[2] = (0.5 * (fl_rhov + fr_rhov) - 0.5 * S * (rhovr - rhovl)) * mesure;
    fl[3] = (0.5 * (fl_rhow + fr_rhow) - 0.5 * S * (rhowr - rhowl)) * mesure;
    fl[4] = (0.5 * (fl_rhoE + fr_rhoE) - 0.5 * S * (rhoEr - rhoEl)) * mesure;
}

// Where the k-th field of a state (or residual) lives: either [k][n] rows of one array
// (Rows) or k separate arrays, as the reference passes them (Arrays). Both give field k of
// element c at f(k)[c].
struct Rows {
    double *p;
    size_t n;
    __device__ __forceinline__ double *f(int k) const { return p + (size_t)k * n; }
};
template <int N>
struct Arrays {
    double *p[N];
    __device__ __forceinline__ double *f(int k) const { return p[k]; }
};

template <class A>
__device__ __forceinline__ Prim6 load6(const A &U, int c)
{
    return Prim6{U.f(0)[c], U.f(1)[c], U.f(2)[c], U.f(3)[c], U.f(4)[c], U.f(5)[c]};
}

__device__ __forceinline__ Prim6 load6(const double *__restrict__ U, size_t n, int c)
{
    Prim6 p;
    p.rho = U[c];
    p.P = U[n + c];
    p.ru = U[2 * n + c];
    p.rv = U[3 * n + c];
    p.rw = U[4 * n + c];
    p.rE = U[5 * n + c];
    return p;
}

// ghost_value_TubeSchok3D for one face from its owner state. rhowg_prev is the ghost
// array's previous content, kept under quirk 1 on slip faces.
__device__ __forceinline__ Prim6 euler_ghost(const Prim6 &c, const double *n, int name, int quirk,
                                             double rhowg_prev)
{
    Prim6 g = c;
    if (name == 3 || name == 4 || name == 5 || name == 6) {
        double u_i = c.ru / c.rho, v_i = c.rv / c.rho, w_i = c.rw / c.rho;
        double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        double s0 = n[0] / nn, s1 = n[1] / nn, s2 = n[2] / nn;
        double un = u_i * s0 + v_i * s1 + w_i * s2;
        double u_g = u_i - 2 * un * s0;
        double v_g = v_i - 2 * un * s1;
        double w_g = w_i - 2 * un * s2;
        g.ru = c.rho * u_g;
        if (quirk) {
            g.rv = c.rho * w_g;
            g.rw = rhowg_prev;
        } else {
            g.rv = c.rho * v_g;
            g.rw = c.rho * w_g;
        }
    }
    return g;
}

// vel (optional, Arrays<3> of [F]): the ghost velocities ug, vg, wg the reference also writes,
// each the ghost momentum over the ghost density.
template <class A>
__global__ void __launch_bounds__(kBlock) k_euler_ghost(Mesh m, const int32_t *__restrict__ bnd, int nb, A U, A G,
                                                         int quirk, Arrays<3> vel)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    int f = bnd[k];
    int name = m.face_name[f];
    if (!(name >= 1 && name <= 6)) return;
    Prim6 c = load6(U, m.face_cells[2 * f]);
    Prim6 g = euler_ghost(c, m.face_normal + 3 * (size_t)f, name, quirk, G.f(4)[f]);
    G.f(0)[f] = g.rho;
    G.f(1)[f] = g.P;
    G.f(2)[f] = g.ru;
    G.f(3)[f] = g.rv;
    G.f(4)[f] = g.rw;
    G.f(5)[f] = g.rE;
    if (vel.p[0]) {
        vel.p[0][f] = g.ru / g.rho;
        vel.p[1][f] = g.rv / g.rho;
        vel.p[2][f] = g.rw / g.rho;
    }
}

// time_step_euler_3d (fvm_utils3d.py:809-825): first four faces of the cell.
__device__ __forceinline__ double euler_cell_dt(const Mesh &m, int i, double rho, double P, double ru, double rv,
                                                double rw, double cfl, double gamma)
{
    double lam = 0.;
    double velson = sqrt(gamma * fabs(P / rho));
    for (int j = 0; j < 4; j++) {
        int f = m.cell_faces[(size_t)j * m.C + i];
        const double *n = m.face_normal + 3 * (size_t)f;
        double u_n = fabs((ru * n[0] + rv * n[1] + rw * n[2]) / rho) / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        double lam_convect = u_n + velson;
        lam += lam_convect * m.face_mesure[f];
    }
    return cfl * m.cell_volume[i] / lam;
}

__global__ void __launch_bounds__(kBlock) k_euler_dt(Mesh m, const double *__restrict__ U, double cfl, double gamma,
                                                      double *__restrict__ dtpart)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = 1e300;
    if (i < m.C) {
        const size_t C = m.C;
        d = euler_cell_dt(m, i, U[i], U[C + i], U[2 * C + i], U[3 * C + i], U[4 * C + i], cfl, gamma);
    }
    wave_min_part(dtpart, (size_t)i, d);
}

// time_step_euler_3d in the reference's form: the per-cell dt_c array (the caller takes the min).
__global__ void __launch_bounds__(kBlock) k_euler_dt_cells(Mesh m, Arrays<5> U, double cfl, double gamma,
                                                            double *__restrict__ dt_c)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    dt_c[i] = euler_cell_dt(m, i, U.f(0)[i], U.f(1)[i], U.f(2)[i], U.f(3)[i], U.f(4)[i], cfl, gamma);
}

// Residual of cell i, faces in increasing id (the single all-faces loop of
// explicitscheme_euler_3d_rusanov). GHOST_ARRAY: right state of boundary faces from G; else
// computed from the owner with euler_ghost (rhowg_prev = 0 under quirk).
template <bool GHOST_ARRAY, class A>
__device__ __forceinline__ void euler_cell_residual(const Mesh &m, int i, const A &U, const A &G, double gamma,
                                                    int quirk, double (&rez)[5])
{
    for (int v = 0; v < 5; v++) rez[v] = 0.;
    const size_t C = m.C;
    const int nf = m.cell_nfaces[i];
    for (int r = 0; r < nf; r++) {
        int s = m.order_sorted[(size_t)r * C + i];
        int f = m.cell_faces[(size_t)s * C + i];
        int a = m.face_cells[2 * f];
        const double mesu = m.face_mesure[f];
        const double *nrm = m.face_normal + 3 * (size_t)f;
        double n[3] = {nrm[0] / mesu, nrm[1] / mesu, nrm[2] / mesu};
        Prim6 L = load6(U, a);
        Prim6 R;
        int name = m.face_name[f];
        if (name == 0) {
            R = load6(U, m.face_cells[2 * f + 1]);
        } else if (GHOST_ARRAY) {
            R = load6(G, f);
        } else {
            R = euler_ghost(L, nrm, name, quirk, 0.);
        }
        double fl[5];
        euler_flux(L, R, n, mesu, gamma, quirk, fl);
        if (a == i) {
            for (int v = 0; v < 5; v++) rez[v] -= fl[v];
        } else {
            for (int v = 0; v < 5; v++) rez[v] += fl[v];
        }
    }
}

template <class A, class R>
__global__ void __launch_bounds__(kBlock) k_euler_rusanov(Mesh m, A U, A G, double gamma, int quirk, R rez)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    double r[5];
    euler_cell_residual<true>(m, i, U, G, gamma, quirk, r);
    for (int v = 0; v < 5; v++) rez.f(v)[i] = r[v];
}

__device__ __forceinline__ void euler_update_cell(double (&u)[6], const double (&r)[5], double gamma, double dt,
                                                  double vol)
{
    // u = rho, P, rhou, rhov, rhow, rhoE ; r = rho, rhou, rhov, rhow, rhoE
    u[0] += dt * ((r[0]) / vol);
    u[2] += dt * ((r[1]) / vol);
    u[3] += dt * ((r[2]) / vol);
    u[4] += dt * ((r[3]) / vol);
    u[5] += dt * ((r[4]) / vol);
    u[1] = (gamma - 1) * (u[5] - 0.5 * (u[2] * u[2] + u[3] * u[3] + u[4] * u[4]) / u[0]);
}

// dt from dt_dev[0] (device scalar), or from dt_host when dt_dev is NULL.
template <class A, class R>
__global__ void __launch_bounds__(kBlock) k_euler_update(int C, A U, R rez, double gamma,
                                                          const double *__restrict__ dt_dev, double dt_host,
                                                          const double *__restrict__ vol)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    double u[6], r[5];
    for (int v = 0; v < 6; v++) u[v] = U.f(v)[i];
    for (int v = 0; v < 5; v++) r[v] = rez.f(v)[i];
    euler_update_cell(u, r, gamma, dt_dev ? dt_dev[0] : dt_host, vol[i]);
    for (int v = 0; v < 6; v++) U.f(v)[i] = u[v];
}

__global__ void __launch_bounds__(kBlock) k_euler_step(Mesh m, const double *__restrict__ Uin,
                                                        double *__restrict__ Uout, double gamma, double cfl,
                                                        int quirk, DtRing ring, double *__restrict__ dtpart)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    ring_advance(ring);
    double d = 1e300;
    if (i < m.C) {
        const size_t C = m.C;
        double r[5];
        const Rows rows{const_cast<double *>(Uin), C};
        euler_cell_residual<false>(m, i, rows, rows, gamma, quirk, r);
        double u[6];
        for (int v = 0; v < 6; v++) u[v] = Uin[v * C + i];
        euler_update_cell(u, r, gamma, ring.cur[0], m.cell_volume[i]);
        for (int v = 0; v < 6; v++) Uout[v * C + i] = u[v];
        d = euler_cell_dt(m, i, u[0], u[1], u[2], u[3], u[4], cfl, gamma);
    }
    wave_min_part(dtpart, (size_t)i, d);
}

// The fused step on tetrahedra (4 faces per cell): the arithmetic of k_euler_step, with the
// loads batched: the cell's face ids and residual order first, then the faces in residual
// order NB at a time (their geometry, then the neighbour states), so that a cell waits for a
// few memory latencies instead of three per face.
template <int NB>
__device__ __forceinline__ void euler_faces(const Mesh &m, int i, const double *__restrict__ U, size_t C,
                                            const Prim6 &own, const int *f, double gamma, int quirk, double (&r)[5])
{
    int fa[NB], fb[NB], nm[NB];
    double ms[NB], nx[NB], ny[NB], nz[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        fa[k] = m.face_cells[2 * f[k]];
        fb[k] = m.face_cells[2 * f[k] + 1];
        nm[k] = m.face_name[f[k]];
        ms[k] = m.face_mesure[f[k]];
        nx[k] = m.face_normal[3 * (size_t)f[k]];
        ny[k] = m.face_normal[3 * (size_t)f[k] + 1];
        nz[k] = m.face_normal[3 * (size_t)f[k] + 2];
    }
    Prim6 nb[NB];  // the cell across each face (the own cell on a boundary face: not used)
#pragma unroll
    for (int k = 0; k < NB; k++) nb[k] = load6(U, C, nm[k] == 0 ? (fa[k] == i ? fb[k] : fa[k]) : i);
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const double nrm[3] = {nx[k], ny[k], nz[k]};
        const double mesu = ms[k];
        const double n[3] = {nrm[0] / mesu, nrm[1] / mesu, nrm[2] / mesu};
        const bool own_L = fa[k] == i;
        const Prim6 L = own_L ? own : nb[k];
        Prim6 R;
        if (nm[k] == 0)
            R = own_L ? nb[k] : own;
        else
            R = euler_ghost(L, nrm, nm[k], quirk, 0.);
        double fl[5];
        euler_flux(L, R, n, mesu, gamma, quirk, fl);
        // rez -= fl (own = L) is rez += -fl
#pragma unroll
        for (int v = 0; v < 5; v++) r[v] += own_L ? -fl[v] : fl[v];
    }
}

// Faces per load batch: two keep the kernel at 144 VGPRs (3 waves per SIMD) and measured
// fastest on the 6M-tetrahedra Sod case (0.69 ms against 0.81 for one face per latency chain,
// 0.74 for all four faces at once at 210 VGPRs); the batch loop stays rolled.
constexpr int kEulerBatch = 2;

__global__ void __launch_bounds__(kBlock) k_euler_step_tet(Mesh m, const double *__restrict__ Uin,
                                                            double *__restrict__ Uout, double gamma, double cfl,
                                                            int quirk, DtRing ring, double *__restrict__ dtpart)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    ring_advance(ring);
    double d = 1e300;
    if (i < m.C) {
        const size_t C = m.C;
        slot_i fs;
        int fk[4];
        unsigned ord = 0;
#pragma unroll
        for (int s = 0; s < 4; s++) {
            fs[s] = m.cell_faces[(size_t)s * C + i];
            ord |= (unsigned)m.order_sorted[(size_t)s * C + i] << (2 * s);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) fk[k] = fs[(ord >> (2 * k)) & 3];  // faces in increasing id
        const Prim6 own = load6(Uin, C, i);
        double r[5] = {0., 0., 0., 0., 0.};
#pragma unroll 1
        for (int k = 0; k < 4; k += kEulerBatch) euler_faces<kEulerBatch>(m, i, Uin, C, own, fk + k, gamma, quirk, r);
        double u[6] = {own.rho, own.P, own.ru, own.rv, own.rw, own.rE};
        euler_update_cell(u, r, gamma, ring.cur[0], m.cell_volume[i]);
#pragma unroll
        for (int v = 0; v < 6; v++) Uout[v * C + i] = u[v];
        // time_step_euler_3d of the new state: euler_cell_dt with the face geometry loaded
        // together
        double nx[4], ny[4], nz[4], ms[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int f = fs[j];
            nx[j] = m.face_normal[3 * (size_t)f];
            ny[j] = m.face_normal[3 * (size_t)f + 1];
            nz[j] = m.face_normal[3 * (size_t)f + 2];
            ms[j] = m.face_mesure[f];
        }
        double lam = 0.;
        const double velson = sqrt(gamma * fabs(u[1] / u[0]));
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double u_n = fabs((u[2] * nx[j] + u[3] * ny[j] + u[4] * nz[j]) / u[0]) /
                               sqrt(nx[j] * nx[j] + ny[j] * ny[j] + nz[j] * nz[j]);
            const double lam_convect = u_n + velson;
            lam += lam_convect * ms[j];
        }
        d = cfl * m.cell_volume[i] / lam;
    }
    wave_min_part(dtpart, (size_t)i, d);
}

// The fused step on packed records (mfv_euler_tet): the arithmetic and summation orders of
// k_euler_step_tet; a neighbour is one 64-byte record [C][8]; the cell's across-codes and
// face geometry are stored per cell in the residual order, so a wave reads them with
// contiguous loads and gathers only the neighbours' records. The geometry is parked in LDS
// by slot for the time step, which sums over the faces in slot order.
struct EulerTetK {
    const int4 *__restrict__ cell_nb;
    const uint8_t *__restrict__ cell_rank;
    const double2 *__restrict__ cell_geo;
    size_t C;
};

constexpr int kNbBoundary = int(0x80000000u);
constexpr int kNbOwnRight = 1 << 30;
constexpr int kNbIndex = (1 << 30) - 1;

__device__ __forceinline__ Prim6 load_rec(const double *__restrict__ R, int c)
{
    const double2 *p = reinterpret_cast<const double2 *>(R + 8 * (size_t)c);
    const double2 a = p[0], b = p[1], e = p[2];
    return Prim6{a.x, a.y, b.x, b.y, e.x, e.y};
}

// Faces k0 and k0 + 1 of the residual order (across codes v0, v1): their geometry and the
// neighbour records load together, the geometry is parked in LDS at the faces' slots
// (sg[slot][component][thread]), then the fluxes are added to r in that order.
__device__ __forceinline__ void euler_face_pair_packed(const EulerTetK &t, int i, const double *__restrict__ R,
                                                       const Prim6 &own, int k0, int v0, int v1, unsigned ord,
                                                       double gamma, int quirk, double (&r)[5], double *sg)
{
    const int v[2] = {v0, v1};
    double2 g01[2], g23[2];
    Prim6 nb[2];  // the cell across each face (the own cell on a boundary face: not used)
#pragma unroll
    for (int k = 0; k < 2; k++) {
        g01[k] = t.cell_geo[(size_t)(2 * (k0 + k)) * t.C + i];
        g23[k] = t.cell_geo[(size_t)(2 * (k0 + k) + 1) * t.C + i];
        nb[k] = load_rec(R, v[k] >= 0 ? (v[k] & kNbIndex) : i);
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int so = ((ord >> (2 * (k0 + k))) & 3) * 4 * kBlock;
        sg[so] = g01[k].x;
        sg[so + kBlock] = g01[k].y;
        sg[so + 2 * kBlock] = g23[k].x;
        sg[so + 3 * kBlock] = g23[k].y;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const double nrm[3] = {g01[k].x, g01[k].y, g23[k].x};
        const double mesu = g23[k].y;
        const double n[3] = {nrm[0] / mesu, nrm[1] / mesu, nrm[2] / mesu};
        const bool own_L = (v[k] & kNbOwnRight) == 0;
        const Prim6 L = own_L ? own : nb[k];
        Prim6 Rs;
        if (v[k] >= 0)
            Rs = own_L ? nb[k] : own;
        else
            Rs = euler_ghost(L, nrm, v[k] & kNbIndex, quirk, 0.);
        double fl[5];
        euler_flux(L, Rs, n, mesu, gamma, quirk, fl);
#pragma unroll
        for (int q = 0; q < 5; q++) r[q] += own_L ? -fl[q] : fl[q];
    }
}

__global__ void __launch_bounds__(kBlock) k_euler_step_packed(EulerTetK t, const double *__restrict__ Rin,
                                                               double *__restrict__ Rout, double gamma, double cfl,
                                                               int quirk, DtRing ring, double *__restrict__ dtpart)
{
    __shared__ double sgeo[16 * kBlock];
    int i = xcd_block() * blockDim.x + threadIdx.x;
    ring_advance(ring);
    double d = 1e300;
    if (i < (int)t.C) {
        const int4 v4 = t.cell_nb[i];
        const unsigned ord = t.cell_rank[i];
        const double2 *rp = reinterpret_cast<const double2 *>(Rin + 8 * (size_t)i);
        const double2 a = rp[0], b = rp[1], e = rp[2], vv = rp[3];
        const Prim6 own{a.x, a.y, b.x, b.y, e.x, e.y};
        const double vol = vv.x;
        double *sg = sgeo + threadIdx.x;
        double r[5] = {0., 0., 0., 0., 0.};
#pragma unroll 1
        for (int k = 0; k < 4; k += 2)
            euler_face_pair_packed(t, i, Rin, own, k, k ? v4.z : v4.x, k ? v4.w : v4.y, ord, gamma, quirk, r, sg);
        double u[6] = {own.rho, own.P, own.ru, own.rv, own.rw, own.rE};
        euler_update_cell(u, r, gamma, ring.cur[0], vol);
        double2 *wp = reinterpret_cast<double2 *>(Rout + 8 * (size_t)i);
        wp[0] = make_double2(u[0], u[1]);
        wp[1] = make_double2(u[2], u[3]);
        wp[2] = make_double2(u[4], u[5]);
        wp[3] = make_double2(vol, 0.);
        // time_step_euler_3d of the new state, faces in slot order
        double lam = 0.;
        const double velson = sqrt(gamma * fabs(u[1] / u[0]));
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double nx = sg[j * 4 * kBlock], ny = sg[j * 4 * kBlock + kBlock];
            const double nz = sg[j * 4 * kBlock + 2 * kBlock], ms = sg[j * 4 * kBlock + 3 * kBlock];
            const double u_n = fabs((u[2] * nx + u[3] * ny + u[4] * nz) / u[0]) / sqrt(nx * nx + ny * ny + nz * nz);
            const double lam_convect = u_n + velson;
            lam += lam_convect * ms;
        }
        d = cfl * vol / lam;
    }
    wave_min_part(dtpart, (size_t)i, d);
}

// The packed step on face tiles: stage the tile cells' records (one contiguous read per
// thread) in LDS, evaluate each face entry once from its own contiguous entry, with the
// records of tile cells from LDS and those of halo cells read where the entry needs them (the
// host puts a tile's entries with a halo cell last, so few waves mix both), then every tile
// cell adds its faces' fluxes in the residual order, updates its record and takes its next-dt
// candidate (face geometry from the tile's entries, still in cache; terms summed in slot
// order). LDS holds 48 B per tile cell and 40 B per face entry: 39 KiB for the 256-cell tiles of
// a Hilbert-ordered tetrahedral mesh, four blocks per CU. Same expressions and orders as
// k_euler_step_packed, so the same bits.
struct EulerTilesK {
    const int32_t *__restrict__ cell0;
    const int32_t *__restrict__ ncells;
    const int32_t *__restrict__ hptr;
    const int32_t *__restrict__ halo;
    const int32_t *__restrict__ fptr;
    const double2 *__restrict__ fgeo;
    const int2 *__restrict__ flr;
    const int4 *__restrict__ code;
    int MI, MFc;  // max items, max face entries per tile
};

template <int T>
__global__ void __launch_bounds__(T) k_euler_step_packed_tiled(EulerTetK t, EulerTilesK tt,
                                                                const double *__restrict__ Rin,
                                                                double *__restrict__ Rout, double gamma, double cfl,
                                                                int quirk, DtRing ring, double *__restrict__ dtpart)
{
    extern __shared__ double lds[];
    double2 *sR = reinterpret_cast<double2 *>(lds);  // [T][3]: the tile cells' six values
    double *sF = lds + 6 * T;                        // [5][MFc]
    const int tile = xcd_block();
    const int c0 = tt.cell0[tile], nc = tt.ncells[tile];
    const int h0 = tt.hptr[tile];
    const int f0 = tt.fptr[tile], nfc = tt.fptr[tile + 1] - f0;
    ring_advance(ring);
    double vol = 0.;  // the thread's cell's volume, kept from its record for the update
    if ((int)threadIdx.x < nc) {
        const double2 *rp = reinterpret_cast<const double2 *>(Rin + 8 * (size_t)(c0 + threadIdx.x));
        const double2 a = rp[0], b = rp[1], e2 = rp[2], vv = rp[3];
        sR[3 * threadIdx.x] = a;
        sR[3 * threadIdx.x + 1] = b;
        sR[3 * threadIdx.x + 2] = e2;
        vol = vv.x;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nfc; k += T) {
        const double2 g01 = tt.fgeo[2 * (size_t)(f0 + k)], g23 = tt.fgeo[2 * (size_t)(f0 + k) + 1];
        const int2 lr = tt.flr[f0 + k];
        const bool hl = lr.x >= nc, hr = lr.y >= nc;
        const int cl = hl ? tt.halo[h0 + (lr.x - nc)] : 0;
        const int cr = hr ? tt.halo[h0 + (lr.y - nc)] : 0;
        Prim6 L, R;
        if (hl) {
            L = load_rec(Rin, cl);
        } else {
            const double2 la = sR[3 * lr.x], lb = sR[3 * lr.x + 1], lc = sR[3 * lr.x + 2];
            L = Prim6{la.x, la.y, lb.x, lb.y, lc.x, lc.y};
        }
        const double nrm[3] = {g01.x, g01.y, g23.x};
        const double mesu = g23.y;
        const double n[3] = {nrm[0] / mesu, nrm[1] / mesu, nrm[2] / mesu};
        if (hr) {
            R = load_rec(Rin, cr);
        } else if (lr.y >= 0) {
            const double2 ra = sR[3 * lr.y], rb = sR[3 * lr.y + 1], rc = sR[3 * lr.y + 2];
            R = Prim6{ra.x, ra.y, rb.x, rb.y, rc.x, rc.y};
        } else {
            R = euler_ghost(L, nrm, -lr.y, quirk, 0.);
        }
        double fl[5];
        euler_flux(L, R, n, mesu, gamma, quirk, fl);
#pragma unroll
        for (int v = 0; v < 5; v++) sF[v * tt.MFc + k] = fl[v];
    }
    __syncthreads();
    const int q = threadIdx.x;
    if ((q & ~63) >= nc) return;  // the whole wave lies past the tile: it owns no partial
    double d = 1e300;
    if (q < nc) {
        const int i = c0 + q;
        const int4 c4 = tt.code[i];
        const unsigned ord = t.cell_rank[i];
        const int cd[4] = {c4.x, c4.y, c4.z, c4.w};
        double2 g01[4], g23[4];  // the faces' geometry from the tile's entries (in cache since the face phase)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            g01[k] = tt.fgeo[2 * (size_t)(f0 + (cd[k] >> 1))];
            g23[k] = tt.fgeo[2 * (size_t)(f0 + (cd[k] >> 1)) + 1];
        }
        double r[5] = {0., 0., 0., 0., 0.};
#pragma unroll
        for (int k = 0; k < 4; k++) {  // faces in increasing original id
            const int ph = cd[k] >> 1;
            const bool own_R = cd[k] & 1;
            // rez -= fl (own = L) is rez += -fl
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const double x = sF[v * tt.MFc + ph];
                r[v] += own_R ? x : -x;
            }
        }
        const double2 oa = sR[3 * q], ob = sR[3 * q + 1], oc = sR[3 * q + 2];
        double u[6] = {oa.x, oa.y, ob.x, ob.y, oc.x, oc.y};
        euler_update_cell(u, r, gamma, ring.cur[0], vol);
        double2 *wp = reinterpret_cast<double2 *>(Rout + 8 * (size_t)i);
        wp[0] = make_double2(u[0], u[1]);
        wp[1] = make_double2(u[2], u[3]);
        wp[2] = make_double2(u[4], u[5]);
        wp[3] = make_double2(vol, 0.);
        // time_step_euler_3d of the new state: the terms in residual order, summed in slot order
        const double velson = sqrt(gamma * fabs(u[1] / u[0]));
        double tm[4];
        int sl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double nx = g01[k].x, ny = g01[k].y, nz = g23[k].x;
            const double u_n = fabs((u[2] * nx + u[3] * ny + u[4] * nz) / u[0]) / sqrt(nx * nx + ny * ny + nz * nz);
            const double lam_convect = u_n + velson;
            tm[k] = lam_convect * g23[k].y;
            sl[k] = (ord >> (2 * k)) & 3;
        }
        double lam = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double tj = tm[0];
#pragma unroll
            for (int k = 1; k < 4; k++) tj = sl[k] == j ? tm[k] : tj;
            lam += tj;
        }
        d = cfl * vol / lam;
    }
    wave_min_part(dtpart, (size_t)c0 + q, d);
}

// The packed step on face tiles with a single cell phase: the tile's items' records are staged
// in LDS as above, then every tile cell evaluates the fluxes of its own four faces from its
// faces' entries (geometry and item pair) and the staged records, in the residual order, and
// updates its record. An inner face of two tile cells is evaluated by both, with the same
// arguments, so both get the same bits as the face phase of k_euler_step_packed_tiled; the LDS
// holds only the records (48 B per item), which lets tiles of 512 cells keep several blocks
// per CU and cuts the halo share of the staged records.
template <int T>
__global__ void __launch_bounds__(T) k_euler_step_cells_tiled(EulerTetK t, EulerTilesK tt,
                                                               const double *__restrict__ Rin,
                                                               double *__restrict__ Rout, double gamma, double cfl,
                                                               int quirk, DtRing ring, double *__restrict__ dtpart)
{
    extern __shared__ double lds[];
    double2 *sR = reinterpret_cast<double2 *>(lds);  // [MI][3]: six values per item
    const int tile = xcd_block();
    const int c0 = tt.cell0[tile], nc = tt.ncells[tile];
    const int h0 = tt.hptr[tile], ni = nc + (tt.hptr[tile + 1] - h0);
    const int f0 = tt.fptr[tile];
    ring_advance(ring);
    for (int e = threadIdx.x; e < ni; e += T) {
        const int c = (e < nc) ? c0 + e : tt.halo[h0 + (e - nc)];
        const double2 *rp = reinterpret_cast<const double2 *>(Rin + 8 * (size_t)c);
        const double2 a = rp[0], b = rp[1], e2 = rp[2];
        sR[3 * e] = a;
        sR[3 * e + 1] = b;
        sR[3 * e + 2] = e2;
    }
    __syncthreads();
    const int q = threadIdx.x;
    if ((q & ~63) >= nc) return;  // the whole wave lies past the tile: it owns no partial
    double d = 1e300;
    if (q < nc) {
        const int i = c0 + q;
        const int4 c4 = tt.code[i];
        const unsigned ord = t.cell_rank[i];
        const double vol = Rin[8 * (size_t)i + 6];
        const int cd[4] = {c4.x, c4.y, c4.z, c4.w};
        double r[5] = {0., 0., 0., 0., 0.};
        // the next face's entry is loaded while the current flux is evaluated
        size_t ent = (size_t)(f0 + (cd[0] >> 1));
        double2 n01 = tt.fgeo[2 * ent], n23 = tt.fgeo[2 * ent + 1];
        int2 nlr = tt.flr[ent];
#pragma unroll 1
        for (int k = 0; k < 4; k++) {  // faces in increasing original id
            const bool own_R = cd[k] & 1;
            const double2 g01 = n01, g23 = n23;
            const int2 lr = nlr;
            if (k < 3) {
                ent = (size_t)(f0 + (cd[k + 1] >> 1));
                n01 = tt.fgeo[2 * ent];
                n23 = tt.fgeo[2 * ent + 1];
                nlr = tt.flr[ent];
            }
            const double nrm[3] = {g01.x, g01.y, g23.x};
            const double mesu = g23.y;
            const double n[3] = {nrm[0] / mesu, nrm[1] / mesu, nrm[2] / mesu};
            const double2 la = sR[3 * lr.x], lb = sR[3 * lr.x + 1], lc = sR[3 * lr.x + 2];
            const Prim6 L{la.x, la.y, lb.x, lb.y, lc.x, lc.y};
            Prim6 R;
            if (lr.y >= 0) {
                const double2 ra = sR[3 * lr.y], rb = sR[3 * lr.y + 1], rc = sR[3 * lr.y + 2];
                R = Prim6{ra.x, ra.y, rb.x, rb.y, rc.x, rc.y};
            } else {
                R = euler_ghost(L, nrm, -lr.y, quirk, 0.);
            }
            double fl[5];
            euler_flux(L, R, n, mesu, gamma, quirk, fl);
            // rez -= fl (own = L) is rez += -fl
#pragma unroll
            for (int v = 0; v < 5; v++) r[v] += own_R ? fl[v] : -fl[v];
        }
        const double2 oa = sR[3 * q], ob = sR[3 * q + 1], oc = sR[3 * q + 2];
        double u[6] = {oa.x, oa.y, ob.x, ob.y, oc.x, oc.y};
        euler_update_cell(u, r, gamma, ring.cur[0], vol);
        double2 *wp = reinterpret_cast<double2 *>(Rout + 8 * (size_t)i);
        wp[0] = make_double2(u[0], u[1]);
        wp[1] = make_double2(u[2], u[3]);
        wp[2] = make_double2(u[4], u[5]);
        wp[3] = make_double2(vol, 0.);
        // time_step_euler_3d of the new state: the terms in residual order, summed in slot order
        const double velson = sqrt(gamma * fabs(u[1] / u[0]));
        double tm[4];
        int sl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {  // the faces' geometry again (in cache since the flux loop)
            const size_t ent = (size_t)(f0 + (cd[k] >> 1));
            const double2 g01 = tt.fgeo[2 * ent], g23 = tt.fgeo[2 * ent + 1];
            const double nx = g01.x, ny = g01.y, nz = g23.x;
            const double u_n = fabs((u[2] * nx + u[3] * ny + u[4] * nz) / u[0]) / sqrt(nx * nx + ny * ny + nz * nz);
            const double lam_convect = u_n + velson;
            tm[k] = lam_convect * g23.y;
            sl[k] = (ord >> (2 * k)) & 3;
        }
        double lam = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double tj = tm[0];
#pragma unroll
            for (int k = 1; k < 4; k++) tj = sl[k] == j ? tm[k] : tj;
            lam += tj;
        }
        d = cfl * vol / lam;
    }
    wave_min_part(dtpart, (size_t)c0 + q, d);
}

// ---------------------------------------------------------------------------------------------
// FP32-bulk step: float32 records [C][8] = (rho, P, rhou, rhov, rhow, rhoE, |K|, 0) and float32
// face geometry per tile entry (nx, ny, nz, |f|); float32 face fluxes (pressure of the momentum
// flux in deviation form, 1/2 (P_L - P_R) n, added by both cells); float64 per-cell gather in the
// residual order, float64 update, float32 store; float64 dt and conservation sums. The CPU
// model oracle/euler32.c makes the same operations in the same order.

struct Rusanov32 {
    float F[5];  // mass, momentum without pressure, energy (x |f|)
    float pd;    // 1/2 (P_L - P_R) |f|
};

__device__ __forceinline__ float pmax32(float a, float b) { return (b > a) ? b : a; }

__device__ __forceinline__ Rusanov32 flux32(const float (&L)[6], const float (&R)[6], float4 g, float gamma,
                                            int quirk)
{
    const float mes = g.w;
    const float n0 = g.x / mes, n1 = g.y / mes, n2 = g.z / mes;
    const float rl = L[0], Pl = L[1], ul = L[2], vl = L[3], wl = L[4], El = L[5];
    const float rr = R[0], Pr = R[1], ur = R[2], vr = R[3], wr = R[4], Er = R[5];
    const float ql = ul * n0 + vl * n1 + wl * n2;
    const float qr = quirk ? (ur * n0 + vl * n1 + wl * n2) : (ur * n0 + vr * n1 + wr * n2);
    const float qrt = ur * n0 + vr * n1 + wr * n2;
    const float cl = sqrtf(gamma * Pl / rl), cr = sqrtf(gamma * Pr / rr);
    const float al = ql / rl, ar = qr / rr;
    const float Ll = pmax32(pmax32(fabsf(al - cl), fabsf(al)), fabsf(al + cl));
    const float Lr = pmax32(pmax32(fabsf(ar - cr), fabsf(ar)), fabsf(ar + cr));
    const float S = (Ll > Lr) ? Ll : Lr;
    const float sl = ql / rl, sr = qrt / rr;
    Rusanov32 o;
    o.F[0] = (0.5f * (ql + qrt) - 0.5f * S * (rr - rl)) * mes;
    o.F[1] = (0.5f * (ul * sl + ur * sr) - 0.5f * S * (ur - ul)) * mes;
    o.F[2] = (0.5f * (vl * sl + vr * sr) - 0.5f * S * (vr - vl)) * mes;
    o.F[3] = (0.5f * (wl * sl + wr * sr) - 0.5f * S * (wr - wl)) * mes;
    o.F[4] = (0.5f * ((El + Pl) * sl + (Er + Pr) * sr) - 0.5f * S * (Er - El)) * mes;
    o.pd = 0.5f * (Pl - Pr) * mes;
    return o;
}

// ghost_value_TubeSchok3D in float32 (quirk: rhovg = rho w_g, rhowg = 0)
__device__ __forceinline__ void ghost32(const float (&c)[6], float4 g, int name, int quirk, float (&o)[6])
{
#pragma unroll
    for (int k = 0; k < 6; k++) o[k] = c[k];
    if (name >= 3 && name <= 6) {
        const float u = c[2] / c[0], v = c[3] / c[0], w = c[4] / c[0];
        const float nn = sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
        const float s0 = g.x / nn, s1 = g.y / nn, s2 = g.z / nn;
        const float un = u * s0 + v * s1 + w * s2;
        const float ug = u - 2.f * un * s0, vg = v - 2.f * un * s1, wg = w - 2.f * un * s2;
        o[2] = c[0] * ug;
        if (quirk) {
            o[3] = c[0] * wg;
            o[4] = 0.f;
        } else {
            o[3] = c[0] * vg;
            o[4] = c[0] * wg;
        }
    }
}

// Block sum of v[5] in a fixed tree (so the sums do not depend on timing); result in lane 0 of
// wave 0. red: 5 * (T / 64) doubles of LDS.
template <int T>
__device__ __forceinline__ void block_sum5(double (&v)[5], double *red)
{
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < 5; k++) red[k * (T / 64) + w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 5; k++) {
            double a = 0.;
            for (int j = 0; j < T / 64; j++) a += red[k * (T / 64) + j];
            v[k] = a;
        }
}

template <int T>
__global__ void __launch_bounds__(T) k_euler_step32_tiled(EulerTetK t, EulerTilesK tt, const float4 *__restrict__ geo,
                                                           const float *__restrict__ Rin, float *__restrict__ Rout,
                                                           float gamma32, double gamma, double cfl, int quirk,
                                                           DtRing ring, double *__restrict__ dtpart,
                                                           double *__restrict__ conspart)
{
    extern __shared__ float lds32[];
    float2 *sR = reinterpret_cast<float2 *>(lds32);  // [MI][3]: six values per item
    float *sF = lds32 + 6 * tt.MI;                   // [6][MFc]
    const int tile = xcd_block();
    const int c0 = tt.cell0[tile], nc = tt.ncells[tile];
    const int h0 = tt.hptr[tile], ni = nc + (tt.hptr[tile + 1] - h0);
    const int f0 = tt.fptr[tile], nfc = tt.fptr[tile + 1] - f0;
    ring_advance(ring);
    for (int e = threadIdx.x; e < ni; e += T) {
        const int c = (e < nc) ? c0 + e : tt.halo[h0 + (e - nc)];
        const float4 *rp = reinterpret_cast<const float4 *>(Rin + 8 * (size_t)c);
        const float4 a = rp[0], b = rp[1];
        sR[3 * e] = make_float2(a.x, a.y);
        sR[3 * e + 1] = make_float2(a.z, a.w);
        sR[3 * e + 2] = make_float2(b.x, b.y);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nfc; k += T) {
        const float4 g = geo[f0 + k];
        const int2 lr = tt.flr[f0 + k];
        const float2 la = sR[3 * lr.x], lb = sR[3 * lr.x + 1], lc = sR[3 * lr.x + 2];
        const float L[6] = {la.x, la.y, lb.x, lb.y, lc.x, lc.y};
        float R[6];
        if (lr.y >= 0) {
            const float2 ra = sR[3 * lr.y], rb = sR[3 * lr.y + 1], rc = sR[3 * lr.y + 2];
            R[0] = ra.x; R[1] = ra.y; R[2] = rb.x; R[3] = rb.y; R[4] = rc.x; R[5] = rc.y;
        } else {
            ghost32(L, g, -lr.y, quirk, R);
        }
        const Rusanov32 o = flux32(L, R, g, gamma32, quirk);
#pragma unroll
        for (int v = 0; v < 5; v++) sF[v * tt.MFc + k] = o.F[v];
        sF[5 * tt.MFc + k] = o.pd;
    }
    __syncthreads();
    const int q = threadIdx.x;
    double d = 1e300;
    double cs[5] = {0., 0., 0., 0., 0.};
    if (q < nc) {
        const int i = c0 + q;
        const int4 c4 = tt.code[i];
        const unsigned ord = t.cell_rank[i];
        const float vol32 = Rin[8 * (size_t)i + 6];
        const int cd[4] = {c4.x, c4.y, c4.z, c4.w};
        float4 g[4];  // the faces' geometry from the tile's entries (in cache since the face phase)
#pragma unroll
        for (int k = 0; k < 4; k++) g[k] = geo[f0 + (cd[k] >> 1)];
        double r[5] = {0., 0., 0., 0., 0.};
#pragma unroll
        for (int k = 0; k < 4; k++) {  // faces in increasing original id
            const int ph = cd[k] >> 1;
            const bool own_R = cd[k] & 1;
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const double x = (double)sF[v * tt.MFc + ph];
                r[v] += own_R ? x : -x;
            }
            const float pd = sF[5 * tt.MFc + ph];
            r[1] += (double)(pd * (g[k].x / g[k].w));
            r[2] += (double)(pd * (g[k].y / g[k].w));
            r[3] += (double)(pd * (g[k].z / g[k].w));
        }
        const float2 oa = sR[3 * q], ob = sR[3 * q + 1], oc = sR[3 * q + 2];
        const double vol = (double)vol32;
        const double dv = ring.cur[0] / vol;
        const double rho = (double)oa.x + r[0] * dv;
        const double mx = (double)ob.x + r[1] * dv;
        const double my = (double)ob.y + r[2] * dv;
        const double mz = (double)oc.x + r[3] * dv;
        const double E = (double)oc.y + r[4] * dv;
        const double P = (gamma - 1.) * (E - 0.5 * (mx * mx + my * my + mz * mz) / rho);
        const float w0 = (float)rho, w1 = (float)P, w2 = (float)mx, w3 = (float)my, w4 = (float)mz, w5 = (float)E;
        float4 *wp = reinterpret_cast<float4 *>(Rout + 8 * (size_t)i);
        wp[0] = make_float4(w0, w1, w2, w3);
        wp[1] = make_float4(w4, w5, vol32, 0.f);
        cs[0] = (double)w0 * vol;
        cs[1] = (double)w2 * vol;
        cs[2] = (double)w3 * vol;
        cs[3] = (double)w4 * vol;
        cs[4] = (double)w5 * vol;
        // time_step_euler_3d of the stored state in float64, terms summed in slot order
        const double ur = w0, uP = w1, ux = w2, uy = w3, uz = w4;
        const double velson = sqrt(gamma * fabs(uP / ur));
        double tm[4];
        int sl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const double nx = g[k].x, ny = g[k].y, nz = g[k].z;
            const double u_n = fabs((ux * nx + uy * ny + uz * nz) / ur) / sqrt(nx * nx + ny * ny + nz * nz);
            tm[k] = (u_n + velson) * (double)g[k].w;
            sl[k] = (ord >> (2 * k)) & 3;
        }
        double lam = 0.;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double tj = tm[0];
#pragma unroll
            for (int k = 1; k < 4; k++) tj = sl[k] == j ? tm[k] : tj;
            lam += tj;
        }
        d = cfl * vol / lam;
    }
    if ((q & ~63) < nc) wave_min_part(dtpart, (size_t)c0 + q, d);
    if (conspart) {
        __syncthreads();  // the flux LDS is free: reuse it for the block sums
        block_sum5<T>(cs, reinterpret_cast<double *>(lds32));
        if (q == 0)
#pragma unroll
            for (int k = 0; k < 5; k++) conspart[(size_t)k * gridDim.x + tile] = cs[k];
    }
}

// cons[k] = sum over tiles of the block partials (one block, fixed order per thread, tree fold)
__global__ void __launch_bounds__(256) k_cons_fold(const double *__restrict__ part, int n, double *__restrict__ cons)
{
    __shared__ double sh[256];
    for (int k = 0; k < 5; k++) {
        double a = 0.;
        for (int j = threadIdx.x; j < n; j += 256) a += part[(size_t)k * n + j];
        sh[threadIdx.x] = a;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) cons[k] = sh[0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) k_euler32_pack(int C, const double *__restrict__ U,
                                                          const double *__restrict__ vol, float *__restrict__ rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const size_t n = C;
    float4 *w = reinterpret_cast<float4 *>(rec + 8 * (size_t)i);
    w[0] = make_float4((float)U[i], (float)U[n + i], (float)U[2 * n + i], (float)U[3 * n + i]);
    w[1] = make_float4((float)U[4 * n + i], (float)U[5 * n + i], (float)vol[i], 0.f);
}

__global__ void __launch_bounds__(kBlock) k_euler32_unpack(int C, const float *__restrict__ rec,
                                                            double *__restrict__ U)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const size_t n = C;
    const float4 *p = reinterpret_cast<const float4 *>(rec + 8 * (size_t)i);
    const float4 a = p[0], b = p[1];
    U[i] = a.x;
    U[n + i] = a.y;
    U[2 * n + i] = a.z;
    U[3 * n + i] = a.w;
    U[4 * n + i] = b.x;
    U[5 * n + i] = b.y;
}

__global__ void __launch_bounds__(kBlock) k_euler32_geo(int n, const double2 *__restrict__ g64,
                                                         float4 *__restrict__ g32)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double2 a = g64[2 * (size_t)k], b = g64[2 * (size_t)k + 1];
    g32[k] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
}

__global__ void __launch_bounds__(kBlock) k_euler_pack(int C, const double *__restrict__ U,
                                                        const double *__restrict__ vol, double *__restrict__ rec)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const size_t n = C;
    double2 *w = reinterpret_cast<double2 *>(rec + 8 * (size_t)i);
    w[0] = make_double2(U[i], U[n + i]);
    w[1] = make_double2(U[2 * n + i], U[3 * n + i]);
    w[2] = make_double2(U[4 * n + i], U[5 * n + i]);
    w[3] = make_double2(vol[i], 0.);
}

__global__ void __launch_bounds__(kBlock) k_euler_unpack(int C, const double *__restrict__ rec,
                                                          double *__restrict__ U)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const size_t n = C;
    const double2 *p = reinterpret_cast<const double2 *>(rec + 8 * (size_t)i);
    const double2 a = p[0], b = p[1], e = p[2];
    U[i] = a.x;
    U[n + i] = a.y;
    U[2 * n + i] = b.x;
    U[3 * n + i] = b.y;
    U[4 * n + i] = e.x;
    U[5 * n + i] = e.y;
}

// The fused step on face tiles (mfv_face_tiles): the block stages the states of its cells and
// halo in LDS ([6][max_items]), evaluates every face of its tile once into LDS ([5][max_faces]),
// then each cell adds its faces' fluxes in the residual order (increasing face id), updates
// and takes its next-dt candidate. Same expressions and order as k_euler_step, so the same
// bits; a face is evaluated once per tile instead of once per cell side.
struct FaceTilesK {
    const int32_t *__restrict__ cell0;
    const int32_t *__restrict__ ncells;
    const int32_t *__restrict__ hptr;
    const int32_t *__restrict__ halo;
    const int32_t *__restrict__ fptr;
    const int32_t *__restrict__ fid;
    const int32_t *__restrict__ fl;
    const int32_t *__restrict__ fr;
    const int32_t *__restrict__ lface;
    int MI, MFc;  // max items, max faces per tile
};

template <int T>
__global__ void __launch_bounds__(T) k_euler_step_tiled(Mesh m, FaceTilesK t, const double *__restrict__ Uin,
                                                         double *__restrict__ Uout, double gamma, double cfl,
                                                         int quirk, DtRing ring, double *__restrict__ dtpart)
{
    extern __shared__ double lds[];
    double *sU = lds;               // [6][MI]
    double *sF = lds + 6 * t.MI;    // [5][MFc]
    const int tile = xcd_block();
    const int c0 = t.cell0[tile], nc = t.ncells[tile];
    const int h0 = t.hptr[tile], ni = nc + (t.hptr[tile + 1] - h0);
    const int f0 = t.fptr[tile], nfc = t.fptr[tile + 1] - f0;
    const size_t C = m.C;
    ring_advance(ring);
    for (int e = threadIdx.x; e < ni; e += T) {
        const int c = (e < nc) ? c0 + e : t.halo[h0 + (e - nc)];
#pragma unroll
        for (int v = 0; v < 6; v++) sU[v * t.MI + e] = Uin[v * C + c];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nfc; k += T) {
        const int f = t.fid[f0 + k], pl = t.fl[f0 + k], pr = t.fr[f0 + k];
        const int name = m.face_name[f];
        const double mesu = m.face_mesure[f];
        const double nrm[3] = {m.face_normal[3 * (size_t)f], m.face_normal[3 * (size_t)f + 1],
                               m.face_normal[3 * (size_t)f + 2]};
        const double n[3] = {nrm[0] / mesu, nrm[1] / mesu, nrm[2] / mesu};
        const Prim6 L = {sU[pl], sU[t.MI + pl], sU[2 * t.MI + pl], sU[3 * t.MI + pl], sU[4 * t.MI + pl],
                         sU[5 * t.MI + pl]};
        Prim6 R;
        if (name == 0) {
            R = {sU[pr], sU[t.MI + pr], sU[2 * t.MI + pr], sU[3 * t.MI + pr], sU[4 * t.MI + pr], sU[5 * t.MI + pr]};
        } else {
            R = euler_ghost(L, nrm, name, quirk, 0.);
        }
        double fl[5];
        euler_flux(L, R, n, mesu, gamma, quirk, fl);
#pragma unroll
        for (int v = 0; v < 5; v++) sF[v * t.MFc + k] = fl[v];
    }
    __syncthreads();
    const int q = threadIdx.x;
    if ((q & ~63) >= nc) return;  // the whole wave lies past the tile: it owns no partial
    double d = 1e300;
    if (q < nc) {
        const int i = c0 + q;
        slot_i lf;
        unsigned ord = 0;
        const int nf = m.cell_nfaces[i];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            lf[s] = t.lface[(size_t)s * C + i];
            ord |= (unsigned)m.order_sorted[(size_t)s * C + i] << (2 * s);
        }
        double nx[4], ny[4], nz[4], ms[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int f = m.cell_faces[(size_t)j * C + i];
            nx[j] = m.face_normal[3 * (size_t)f];
            ny[j] = m.face_normal[3 * (size_t)f + 1];
            nz[j] = m.face_normal[3 * (size_t)f + 2];
            ms[j] = m.face_mesure[f];
        }
        double r[5] = {0., 0., 0., 0., 0.};
#pragma unroll
        for (int k = 0; k < 4; k++) {  // faces in increasing id (order_sorted)
            if (k >= nf) break;
            const int code = lf[(ord >> (2 * k)) & 3];
            const int ph = code >> 1;
            const bool own_R = code & 1;
            // rez -= fl (own = L) is rez += -fl
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const double x = sF[v * t.MFc + ph];
                r[v] += own_R ? x : -x;
            }
        }
        double u[6];
#pragma unroll
        for (int v = 0; v < 6; v++) u[v] = sU[v * t.MI + q];
        euler_update_cell(u, r, gamma, ring.cur[0], m.cell_volume[i]);
#pragma unroll
        for (int v = 0; v < 6; v++) Uout[v * C + i] = u[v];
        // time_step_euler_3d of the new state (euler_cell_dt, first four faces)
        double lam = 0.;
        const double velson = sqrt(gamma * fabs(u[1] / u[0]));
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double u_n = fabs((u[2] * nx[j] + u[3] * ny[j] + u[4] * nz[j]) / u[0]) /
                               sqrt(nx[j] * nx[j] + ny[j] * ny[j] + nz[j] * nz[j]);
            const double lam_convect = u_n + velson;
            lam += lam_convect * ms[j];
        }
        d = cfl * m.cell_volume[i] / lam;
    }
    wave_min_part(dtpart, (size_t)c0 + q, d);
}

}  // namespace

extern "C" int mfv_euler3d_ghost(const mfv_mesh *m, const double *U, double *G, int quirk, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 3, "Euler 3D needs a 3D mesh");
    MFV_REQUIRE(U && G, "NULL array");
    if (m->nb_bnd == 0) return MFV_OK;
    MFV_REQUIRE(m->bnd_faces, "mesh descriptor lacks bnd_faces");
    const Rows u{const_cast<double *>(U), (size_t)m->nb_cells}, g{G, (size_t)m->nb_faces};
    hipLaunchKernelGGL(k_euler_ghost<Rows>, dim3(grid_for(m->nb_bnd)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), m->bnd_faces, m->nb_bnd, u, g, quirk, Arrays<3>{{nullptr, nullptr, nullptr}});
    return check_launch("k_euler_ghost");
}

static Arrays<6> arrays6(const mfv_euler_arrays *a)
{
    return Arrays<6>{{a->rho, a->P, a->rhou, a->rhov, a->rhow, a->rhoE}};
}

static bool complete(const mfv_euler_arrays *a, bool need_P = true)
{
    return a && a->rho && (a->P || !need_P) && a->rhou && a->rhov && a->rhow && a->rhoE;
}

extern "C" int mfv_ghost_value_tubeschok3d(const mfv_mesh *m, const mfv_euler_arrays *ghosts, double *ug, double *vg,
                                           double *wg, const mfv_euler_arrays *cells, int quirk, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 3, "Euler 3D needs a 3D mesh");
    MFV_REQUIRE(complete(ghosts) && complete(cells), "NULL Euler array");
    MFV_REQUIRE((ug && vg && wg) || (!ug && !vg && !wg), "ug, vg, wg must be all given or all NULL");
    if (m->nb_bnd == 0) return MFV_OK;
    MFV_REQUIRE(m->bnd_faces, "mesh descriptor lacks bnd_faces");
    hipLaunchKernelGGL(k_euler_ghost<Arrays<6>>, dim3(grid_for(m->nb_bnd)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), m->bnd_faces, m->nb_bnd, arrays6(cells), arrays6(ghosts), quirk,
                       Arrays<3>{{ug, vg, wg}});
    return check_launch("k_euler_ghost");
}

extern "C" int mfv_time_step_euler_3d(const mfv_mesh *m, const mfv_euler_arrays *cells, double cfl, double gamma,
                                      double *dt_c, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 3 && m->max_cell_faces >= 4, "Euler 3D time step needs tetrahedra");
    MFV_REQUIRE(complete(cells, true) && dt_c, "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    const Arrays<5> u{{cells->rho, cells->P, cells->rhou, cells->rhov, cells->rhow}};
    hipLaunchKernelGGL(k_euler_dt_cells, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), u, cfl, gamma, dt_c);
    return check_launch("k_euler_dt_cells");
}

extern "C" int mfv_explicitscheme_euler_3d_rusanov(const mfv_mesh *m, const mfv_euler_rez *rez,
                                                   const mfv_euler_arrays *cells, const mfv_euler_arrays *ghosts,
                                                   double gamma, int quirk, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 1)) return rc;
    MFV_REQUIRE(m->dim == 3, "Euler 3D needs a 3D mesh");
    MFV_REQUIRE(rez && rez->rho && rez->rhou && rez->rhov && rez->rhow && rez->rhoE, "NULL residual array");
    MFV_REQUIRE(complete(cells) && complete(ghosts), "NULL Euler array");
    if (m->nb_cells == 0) return MFV_OK;
    const Arrays<5> r{{rez->rho, rez->rhou, rez->rhov, rez->rhow, rez->rhoE}};
    hipLaunchKernelGGL((k_euler_rusanov<Arrays<6>, Arrays<5>>), dim3(grid_for(m->nb_cells)), dim3(kBlock), 0,
                       as_stream(stream), to_kernel(m), arrays6(cells), arrays6(ghosts), gamma, quirk, r);
    return check_launch("k_euler_rusanov");
}

extern "C" int mfv_update_euler_3d_fvc(const mfv_mesh *m, const mfv_euler_arrays *cells, const mfv_euler_rez *rez,
                                       double gamma, double dtime, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(rez && rez->rho && rez->rhou && rez->rhov && rez->rhow && rez->rhoE, "NULL residual array");
    MFV_REQUIRE(complete(cells), "NULL Euler array");
    if (m->nb_cells == 0) return MFV_OK;
    const Arrays<5> r{{rez->rho, rez->rhou, rez->rhov, rez->rhow, rez->rhoE}};
    hipLaunchKernelGGL((k_euler_update<Arrays<6>, Arrays<5>>), dim3(grid_for(m->nb_cells)), dim3(kBlock), 0,
                       as_stream(stream), m->nb_cells, arrays6(cells), r, gamma, nullptr, dtime, m->cell_volume);
    return check_launch("k_euler_update");
}

extern "C" int mfv_euler3d_time_step(const mfv_mesh *m, const double *U, double cfl, double gamma, double *dt_dev,
                                     double *dt_host, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 3 && m->max_cell_faces >= 4, "Euler 3D time step needs tetrahedra");
    MFV_REQUIRE(U && dt_dev, "NULL array");
    if (int rc = mfv_dt_reset(dt_dev, stream)) return rc;
    if (m->nb_cells) {
        const size_t nw = waves_for(m->nb_cells);
        double *part = dt_scratch(nw, as_stream(stream));
        if (!part) return MFV_ENOMEM;
        hipLaunchKernelGGL(k_euler_dt, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m),
                           U, cfl, gamma, part);
        if (int rc = dt_finish(part, nw, dt_dev, as_stream(stream))) return rc;
    }
    if (int rc = check_launch("k_euler_dt")) return rc;
    if (dt_host) {
        hipError_t e = hipMemcpyAsync(dt_host, dt_dev, sizeof(double), hipMemcpyDeviceToHost, as_stream(stream));
        if (e == hipSuccess) e = hipStreamSynchronize(as_stream(stream));
        if (e != hipSuccess) {
            set_error("dt readback: %s", hipGetErrorString(e));
            return MFV_EHIP;
        }
    }
    return MFV_OK;
}

extern "C" int mfv_euler3d_rusanov(const mfv_mesh *m, const double *U, const double *G, double gamma, int quirk,
                                   double *rez, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 1)) return rc;
    MFV_REQUIRE(m->dim == 3, "Euler 3D needs a 3D mesh");
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(U && G && rez, "NULL array");
    const Rows u{const_cast<double *>(U), (size_t)m->nb_cells}, g{const_cast<double *>(G), (size_t)m->nb_faces};
    hipLaunchKernelGGL((k_euler_rusanov<Rows, Rows>), dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), u, g, gamma, quirk, Rows{rez, (size_t)m->nb_cells});
    return check_launch("k_euler_rusanov");
}

extern "C" int mfv_euler3d_update(const mfv_mesh *m, double *U, const double *rez, double gamma, const double *dt_dev,
                                  void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(U && rez && dt_dev, "NULL array");
    hipLaunchKernelGGL((k_euler_update<Rows, Rows>), dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       m->nb_cells, Rows{U, (size_t)m->nb_cells}, Rows{const_cast<double *>(rez), (size_t)m->nb_cells},
                       gamma, dt_dev, 0., m->cell_volume);
    return check_launch("k_euler_update");
}

extern "C" int mfv_euler3d_step_fused(const mfv_mesh *m, const double *U_in, double *U_out, double gamma, double cfl,
                                      int quirk, double *dt_ring_p, int step, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 1)) return rc;
    MFV_REQUIRE(m->dim == 3 && m->max_cell_faces >= 4, "Euler 3D needs tetrahedra");
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(U_in && U_out && U_in != U_out && dt_ring_p, "NULL or aliased array");
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = waves_for(m->nb_cells);
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    hipLaunchKernelGGL(m->max_cell_faces == 4 ? k_euler_step_tet : k_euler_step, dim3(grid_for(m->nb_cells)),
                       dim3(kBlock), 0, as_stream(stream), to_kernel(m), U_in, U_out, gamma, cfl, quirk, r, part);
    if (int rc = check_launch("k_euler_step")) return rc;
    return dt_finish(part, nw, r.next, as_stream(stream));
}

static int check_face_tiles(const mfv_mesh *m, const mfv_face_tiles *t)
{
    MFV_REQUIRE(t != nullptr, "face tiles descriptor is NULL");
    MFV_REQUIRE(t->tile_size == 128 || t->tile_size == 256 || t->tile_size == 512, "tile_size %d not in {128, 256, "
                "512}", t->tile_size);
    MFV_REQUIRE(t->nb_tiles == (m->nb_cells + t->tile_size - 1) / t->tile_size,
                "face tiles were built for another mesh (%d tiles of %d cells for %d cells)", t->nb_tiles,
                t->tile_size, m->nb_cells);
    MFV_REQUIRE(t->max_items >= 1 && t->max_faces >= 1, "bad tile counts");
    MFV_REQUIRE(t->tile_cell0 && t->tile_ncells && t->halo_ptr && t->halo_cells && t->face_ptr && t->face_id &&
                    t->face_l && t->face_r && t->lface,
                "face tiles descriptor has NULL arrays");
    const size_t lds = 48 * (size_t)t->max_items + 40 * (size_t)t->max_faces;
    MFV_REQUIRE(lds <= 64 * 1024, "face tiles of %d cells need %zu bytes of LDS (> 64 KiB): use smaller tiles or "
                "renumber the cells", t->tile_size, lds);
    return MFV_OK;
}

extern "C" int mfv_euler3d_step_tiled(const mfv_mesh *m, const mfv_face_tiles *t, const double *U_in, double *U_out,
                                      double gamma, double cfl, int quirk, double *dt_ring_p, int step, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 1)) return rc;
    MFV_REQUIRE(m->dim == 3 && m->max_cell_faces == 4, "the tiled Euler step needs tetrahedra");
    if (int rc = check_face_tiles(m, t)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(U_in && U_out && U_in != U_out && dt_ring_p, "NULL or aliased array");
    DtRing r = dt_ring(dt_ring_p, step);
    // Tiles start on 64-cell boundaries and waves past a tile's cells return early, so exactly
    // ceil(C/64) partials are written: fold those and no stale slot beyond them.
    const size_t nw = ((size_t)m->nb_cells + 63) / 64;
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    const FaceTilesK tk{t->tile_cell0, t->tile_ncells, t->halo_ptr, t->halo_cells, t->face_ptr, t->face_id,
                        t->face_l,    t->face_r,      t->lface,    t->max_items, t->max_faces};
    const size_t lds = 48 * (size_t)t->max_items + 40 * (size_t)t->max_faces;
    hipStream_t st = as_stream(stream);
    switch (t->tile_size) {
    case 128:
        hipLaunchKernelGGL(k_euler_step_tiled<128>, dim3(t->nb_tiles), dim3(128), lds, st, to_kernel(m), tk, U_in,
                           U_out, gamma, cfl, quirk, r, part);
        break;
    case 256:
        hipLaunchKernelGGL(k_euler_step_tiled<256>, dim3(t->nb_tiles), dim3(256), lds, st, to_kernel(m), tk, U_in,
                           U_out, gamma, cfl, quirk, r, part);
        break;
    default:
        hipLaunchKernelGGL(k_euler_step_tiled<512>, dim3(t->nb_tiles), dim3(512), lds, st, to_kernel(m), tk, U_in,
                           U_out, gamma, cfl, quirk, r, part);
        break;
    }
    if (int rc = check_launch("k_euler_step_tiled")) return rc;
    return dt_finish(part, nw, r.next, st);
}

extern "C" void mfv_euler_tet_host_free(mfv_euler_tet_host *t)
{
    if (!t) return;
    free(t->cell_nb);
    free(t->cell_rank);
    free(t->cell_geo);
    *t = mfv_euler_tet_host{};
}

extern "C" int mfv_euler_tet_build(int C, int F, const int32_t *cell_faces, const int32_t *face_cells,
                                   const int32_t *face_name, const double *face_normal, const double *face_mesure,
                                   mfv_euler_tet_host *out)
{
    MFV_REQUIRE(out, "out is NULL");
    *out = mfv_euler_tet_host{};
    MFV_REQUIRE(C >= 0 && F >= 0 && C <= kNbIndex, "sizes out of range (C %d, F %d)", C, F);
    MFV_REQUIRE((cell_faces || C == 0) && ((face_cells && face_name && face_normal && face_mesure) || F == 0),
                "NULL table");
    mfv_euler_tet_host h{};
    h.nb_cells = C;
    h.nb_faces = F;
    const size_t c1 = (size_t)(C > 0 ? C : 1);
    h.cell_nb = static_cast<int32_t *>(malloc(sizeof(int32_t) * 4 * c1));
    h.cell_rank = static_cast<uint8_t *>(malloc(c1));
    h.cell_geo = static_cast<double *>(malloc(sizeof(double) * 16 * c1));
    auto fail = [&](int rc) {
        mfv_euler_tet_host_free(&h);
        return rc;
    };
    if (!h.cell_nb || !h.cell_rank || !h.cell_geo) {
        set_error("mfv_euler_tet_build: out of host memory");
        return fail(MFV_ENOMEM);
    }
    for (int f = 0; f < F; f++) {
        const int32_t L = face_cells[2 * (size_t)f], R = face_cells[2 * (size_t)f + 1], nm = face_name[f];
        if (L < 0 || L >= C || nm < 0 || nm > kNbIndex || R >= C || (nm == 0 && R < 0)) {
            set_error("face %d: cells (%d, %d), name %d: an inner face (name 0) needs two cells, a boundary face "
                      "a left cell and a name in [1, 2^30)", f, L, R, nm);
            return fail(MFV_EINVAL);
        }
    }
    const size_t CC = (size_t)C;
    for (int i = 0; i < C; i++) {
        int32_t o[4];
        for (int s = 0; s < 4; s++) {
            o[s] = cell_faces[(size_t)s * C + i];
            if (o[s] < 0 || o[s] >= F) {
                set_error("cell %d slot %d: face %d outside [0, %d) (the packed Euler step needs four faces per "
                          "cell)", i, s, o[s], F);
                return fail(MFV_EINVAL);
            }
            const int32_t L = face_cells[2 * (size_t)o[s]], R = face_cells[2 * (size_t)o[s] + 1];
            if (L != i && R != i) {
                set_error("cell %d slot %d: the cell is no side of its face %d (cells %d, %d)", i, s, o[s], L, R);
                return fail(MFV_EINVAL);
            }
        }
        int sl[4] = {0, 1, 2, 3};  // slots by increasing original face id (insertion sort, stable)
        for (int a = 1; a < 4; a++)
            for (int b = a; b > 0 && o[sl[b - 1]] > o[sl[b]]; b--) std::swap(sl[b - 1], sl[b]);
        for (int a = 1; a < 4; a++)
            if (o[sl[a - 1]] == o[sl[a]]) {
                set_error("cell %d holds face %d twice", i, o[sl[a]]);
                return fail(MFV_EINVAL);
            }
        h.cell_rank[i] = (uint8_t)(sl[0] | (sl[1] << 2) | (sl[2] << 4) | (sl[3] << 6));
        for (int k = 0; k < 4; k++) {
            const size_t f = (size_t)o[sl[k]];
            const int32_t L = face_cells[2 * f], R = face_cells[2 * f + 1], nm = face_name[f];
            const int32_t right = L == i ? 0 : kNbOwnRight;
            h.cell_nb[4 * (size_t)i + k] = nm == 0 ? ((L == i ? R : L) | right) : (kNbBoundary | right | nm);
            double *g0 = h.cell_geo + 2 * ((size_t)(2 * k) * CC + i);
            double *g1 = h.cell_geo + 2 * ((size_t)(2 * k + 1) * CC + i);
            g0[0] = face_normal[3 * f];
            g0[1] = face_normal[3 * f + 1];
            g1[0] = face_normal[3 * f + 2];
            g1[1] = face_mesure[f];
        }
    }
    *out = h;
    return MFV_OK;
}

static int check_euler_tet(const mfv_mesh *m, const mfv_euler_tet *t)
{
    MFV_REQUIRE(t != nullptr, "packed Euler tables are NULL");
    MFV_REQUIRE(t->nb_cells == m->nb_cells && t->nb_faces == m->nb_faces,
                "packed Euler tables were built for another mesh (%d cells, %d faces; the mesh has %d, %d)",
                t->nb_cells, t->nb_faces, m->nb_cells, m->nb_faces);
    MFV_REQUIRE(t->cell_nb && t->cell_rank && t->cell_geo, "packed Euler tables have NULL arrays");
    MFV_REQUIRE(((uintptr_t)t->cell_nb & 15) == 0 && ((uintptr_t)t->cell_geo & 15) == 0,
                "packed Euler tables are misaligned");
    return MFV_OK;
}

extern "C" int mfv_euler3d_pack(const mfv_mesh *m, const double *U, double *rec, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(U && rec && ((uintptr_t)rec & 15) == 0, "NULL or misaligned array");
    hipLaunchKernelGGL(k_euler_pack, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), m->nb_cells,
                       U, m->cell_volume, rec);
    return check_launch("k_euler_pack");
}

extern "C" int mfv_euler3d_unpack(const mfv_mesh *m, const double *rec, double *U, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(U && rec && ((uintptr_t)rec & 15) == 0, "NULL or misaligned array");
    hipLaunchKernelGGL(k_euler_unpack, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       m->nb_cells, rec, U);
    return check_launch("k_euler_unpack");
}

extern "C" int mfv_euler3d_step_packed(const mfv_mesh *m, const mfv_euler_tet *t, const double *rec_in,
                                       double *rec_out, double gamma, double cfl, int quirk, double *dt_ring_p,
                                       int step, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 3, "the packed Euler step needs a 3D mesh");
    if (int rc = check_euler_tet(m, t)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(rec_in && rec_out && rec_in != rec_out && dt_ring_p, "NULL or aliased array");
    MFV_REQUIRE(((uintptr_t)rec_in & 15) == 0 && ((uintptr_t)rec_out & 15) == 0, "misaligned records");
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = waves_for(m->nb_cells);
    hipStream_t st = as_stream(stream);
    double *part = dt_scratch(nw, st);
    if (!part) return MFV_ENOMEM;
    const EulerTetK tk{reinterpret_cast<const int4 *>(t->cell_nb), t->cell_rank,
                       reinterpret_cast<const double2 *>(t->cell_geo), (size_t)m->nb_cells};
    hipLaunchKernelGGL(k_euler_step_packed, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, st, tk, rec_in, rec_out,
                       gamma, cfl, quirk, r, part);
    if (int rc = check_launch("k_euler_step_packed")) return rc;
    return dt_finish(part, nw, r.next, st);
}

// lds: the bytes of LDS a block of the kernel takes on these tiles
static int check_euler_tiles(const mfv_mesh *m, const mfv_euler_tet_tiles *tt, size_t lds)
{
    MFV_REQUIRE(tt != nullptr, "packed Euler tiles are NULL");
    MFV_REQUIRE(tt->tile_size == 128 || tt->tile_size == 256 || tt->tile_size == 512,
                "tile_size %d not in {128, 256, 512}", tt->tile_size);
    MFV_REQUIRE(tt->nb_tiles == (m->nb_cells + tt->tile_size - 1) / tt->tile_size,
                "packed Euler tiles were built for another mesh (%d tiles of %d cells for %d cells)", tt->nb_tiles,
                tt->tile_size, m->nb_cells);
    MFV_REQUIRE(tt->max_items >= 1 && tt->max_faces >= 1, "bad tile counts");
    MFV_REQUIRE(tt->tile_cell0 && tt->tile_ncells && tt->halo_ptr && tt->halo_cells && tt->face_ptr && tt->face_geo &&
                    tt->face_lr && tt->cell_code,
                "packed Euler tiles have NULL arrays");
    MFV_REQUIRE(((uintptr_t)tt->face_geo & 15) == 0 && ((uintptr_t)tt->face_lr & 7) == 0 &&
                    ((uintptr_t)tt->cell_code & 15) == 0,
                "packed Euler tiles are misaligned");
    MFV_REQUIRE(lds <= 64 * 1024, "packed Euler tiles of %d cells need %zu bytes of LDS (> 64 KiB): use smaller "
                "tiles or renumber the cells", tt->tile_size, lds);
    return MFV_OK;
}

static int euler_step_packed_tiled(bool cells, const mfv_mesh *m, const mfv_euler_tet *t,
                                   const mfv_euler_tet_tiles *tt, const double *rec_in, double *rec_out, double gamma,
                                   double cfl, int quirk, double *dt_ring_p, int step, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 3, "the packed Euler step needs a 3D mesh");
    if (int rc = check_euler_tet(m, t)) return rc;
    if (!tt) return check_euler_tiles(m, tt, 0);
    // the single-phase step stages the halo's records too; the face-phase step holds the tile
    // cells' records and the face fluxes, and reads the halo's records in the face phase
    const size_t lds = cells ? 48 * (size_t)tt->max_items : 48 * (size_t)tt->tile_size + 40 * (size_t)tt->max_faces;
    if (int rc = check_euler_tiles(m, tt, lds)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(rec_in && rec_out && rec_in != rec_out && dt_ring_p, "NULL or aliased array");
    MFV_REQUIRE(((uintptr_t)rec_in & 15) == 0 && ((uintptr_t)rec_out & 15) == 0, "misaligned records");
    DtRing r = dt_ring(dt_ring_p, step);
    // One partial per 64-cell wave that holds cells (tiles are whole multiples of 64 cells).
    const size_t nw = ((size_t)m->nb_cells + 63) / 64;
    hipStream_t st = as_stream(stream);
    double *part = dt_scratch(nw, st);
    if (!part) return MFV_ENOMEM;
    const EulerTetK tk{reinterpret_cast<const int4 *>(t->cell_nb), t->cell_rank,
                       reinterpret_cast<const double2 *>(t->cell_geo), (size_t)m->nb_cells};
    const EulerTilesK ek{tt->tile_cell0,
                         tt->tile_ncells,
                         tt->halo_ptr,
                         tt->halo_cells,
                         tt->face_ptr,
                         reinterpret_cast<const double2 *>(tt->face_geo),
                         reinterpret_cast<const int2 *>(tt->face_lr),
                         reinterpret_cast<const int4 *>(tt->cell_code),
                         tt->max_items,
                         tt->max_faces};
#define MFV_EULER_TILED(K, TS)                                                                                        \
    hipLaunchKernelGGL(K<TS>, dim3(tt->nb_tiles), dim3(TS), lds, st, tk, ek, rec_in, rec_out, gamma, cfl, quirk, r, \
                       part)
    switch (tt->tile_size) {
    case 128:
        if (cells) MFV_EULER_TILED(k_euler_step_cells_tiled, 128);
        else MFV_EULER_TILED(k_euler_step_packed_tiled, 128);
        break;
    case 256:
        if (cells) MFV_EULER_TILED(k_euler_step_cells_tiled, 256);
        else MFV_EULER_TILED(k_euler_step_packed_tiled, 256);
        break;
    default:
        if (cells) MFV_EULER_TILED(k_euler_step_cells_tiled, 512);
        else MFV_EULER_TILED(k_euler_step_packed_tiled, 512);
        break;
    }
#undef MFV_EULER_TILED
    if (int rc = check_launch(cells ? "k_euler_step_cells_tiled" : "k_euler_step_packed_tiled")) return rc;
    return dt_finish(part, nw, r.next, st);
}

extern "C" int mfv_euler3d_step_packed_tiled(const mfv_mesh *m, const mfv_euler_tet *t, const mfv_euler_tet_tiles *tt,
                                             const double *rec_in, double *rec_out, double gamma, double cfl,
                                             int quirk, double *dt_ring_p, int step, void *stream)
{
    return euler_step_packed_tiled(false, m, t, tt, rec_in, rec_out, gamma, cfl, quirk, dt_ring_p, step, stream);
}

extern "C" int mfv_euler3d_step_cells_tiled(const mfv_mesh *m, const mfv_euler_tet *t, const mfv_euler_tet_tiles *tt,
                                            const double *rec_in, double *rec_out, double gamma, double cfl,
                                            int quirk, double *dt_ring_p, int step, void *stream)
{
    return euler_step_packed_tiled(true, m, t, tt, rec_in, rec_out, gamma, cfl, quirk, dt_ring_p, step, stream);
}

extern "C" int mfv_euler32_pack(const mfv_mesh *m, const double *U, float *rec, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(U && rec, "NULL array");
    MFV_REQUIRE(((uintptr_t)rec & 15) == 0, "misaligned records");
    if (m->nb_cells == 0) return MFV_OK;
    hipLaunchKernelGGL(k_euler32_pack, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), m->nb_cells,
                       U, m->cell_volume, rec);
    return check_launch("k_euler32_pack");
}

extern "C" int mfv_euler32_unpack(const mfv_mesh *m, const float *rec, double *U, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(U && rec, "NULL array");
    MFV_REQUIRE(((uintptr_t)rec & 15) == 0, "misaligned records");
    if (m->nb_cells == 0) return MFV_OK;
    hipLaunchKernelGGL(k_euler32_unpack, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       m->nb_cells, rec, U);
    return check_launch("k_euler32_unpack");
}

extern "C" int mfv_euler32_tile_geometry(const mfv_euler_tet_tiles *tt, int n_entries, float *geo32, void *stream)
{
    MFV_REQUIRE(tt && tt->face_geo && geo32 && n_entries >= 0, "NULL argument");
    MFV_REQUIRE(((uintptr_t)geo32 & 15) == 0 && ((uintptr_t)tt->face_geo & 15) == 0, "misaligned geometry");
    if (n_entries == 0) return MFV_OK;
    hipLaunchKernelGGL(k_euler32_geo, dim3(grid_for(n_entries)), dim3(kBlock), 0, as_stream(stream), n_entries,
                       reinterpret_cast<const double2 *>(tt->face_geo), reinterpret_cast<float4 *>(geo32));
    return check_launch("k_euler32_geo");
}

extern "C" int mfv_euler3d_step32_tiled(const mfv_mesh *m, const mfv_euler_tet *t, const mfv_euler_tet_tiles *tt,
                                        const float *geo32, const float *rec_in, float *rec_out, double gamma,
                                        double cfl, int quirk, double *dt_ring_p, int step, double *cons,
                                        void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 3, "the Euler step needs a 3D mesh");
    if (int rc = check_euler_tet(m, t)) return rc;
    if (int rc = check_euler_tiles(m, tt, tt ? 24 * (size_t)tt->max_items + 24 * (size_t)tt->max_faces : 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(geo32 && rec_in && rec_out && rec_in != rec_out && dt_ring_p, "NULL or aliased array");
    MFV_REQUIRE(((uintptr_t)rec_in & 15) == 0 && ((uintptr_t)rec_out & 15) == 0 && ((uintptr_t)geo32 & 15) == 0,
                "misaligned records or geometry");
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = ((size_t)m->nb_cells + 63) / 64;  // waves with cells (tiles are multiples of 64 cells)
    hipStream_t st = as_stream(stream);
    double *part = dt_scratch(nw + 5 * (size_t)tt->nb_tiles, st);
    if (!part) return MFV_ENOMEM;
    double *conspart = cons ? part + nw : nullptr;
    const EulerTetK tk{reinterpret_cast<const int4 *>(t->cell_nb), t->cell_rank,
                       reinterpret_cast<const double2 *>(t->cell_geo), (size_t)m->nb_cells};
    const EulerTilesK ek{tt->tile_cell0,
                         tt->tile_ncells,
                         tt->halo_ptr,
                         tt->halo_cells,
                         tt->face_ptr,
                         reinterpret_cast<const double2 *>(tt->face_geo),
                         reinterpret_cast<const int2 *>(tt->face_lr),
                         reinterpret_cast<const int4 *>(tt->cell_code),
                         tt->max_items,
                         tt->max_faces};
    const size_t lds = 24 * (size_t)tt->max_items + 24 * (size_t)tt->max_faces;
    MFV_REQUIRE(lds >= 5 * 8 * 8, "tile too small for the conservation sums");
    const float4 *g = reinterpret_cast<const float4 *>(geo32);
    const float g32 = (float)gamma;
    switch (tt->tile_size) {
    case 128:
        hipLaunchKernelGGL(k_euler_step32_tiled<128>, dim3(tt->nb_tiles), dim3(128), lds, st, tk, ek, g, rec_in,
                           rec_out, g32, gamma, cfl, quirk, r, part, conspart);
        break;
    case 256:
        hipLaunchKernelGGL(k_euler_step32_tiled<256>, dim3(tt->nb_tiles), dim3(256), lds, st, tk, ek, g, rec_in,
                           rec_out, g32, gamma, cfl, quirk, r, part, conspart);
        break;
    default:
        hipLaunchKernelGGL(k_euler_step32_tiled<512>, dim3(tt->nb_tiles), dim3(512), lds, st, tk, ek, g, rec_in,
                           rec_out, g32, gamma, cfl, quirk, r, part, conspart);
        break;
    }
    if (int rc = check_launch("k_euler_step32_tiled")) return rc;
    if (int rc = dt_finish(part, nw, r.next, st)) return rc;
    if (cons) {
        hipLaunchKernelGGL(k_cons_fold, dim3(1), dim3(256), 0, st, conspart, tt->nb_tiles, cons);
        return check_launch("k_cons_fold");
    }
    return MFV_OK;
}
