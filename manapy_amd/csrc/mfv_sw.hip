// Shallow water on triangles/quads: second-order MUSCL with least-squares gradients,
// Barth-Jespersen limiting, hydrostatic reconstruction and the Rusanov flux.
//
// Wall boundaries as in manapy's shallow-water driver: h Neumann, hu/hv "nonslip"
// ghosts (w_ghost = -w of the owner cell, boundary/bc_utils.py:45-55), Z Neumann. The
// gradients use these cell ghosts; the wall flux applies the same mirror to the
// reconstructed face state, which keeps the walls exactly mass-conserving at second
// order. Ghosts are computed from the owner on the fly, so no ghost arrays are stored.
//
// State layout: [3][C] = h, hu, hv. Gradient layout: [9][C] = (x, y, psi) of eta = h+Z,
// hu, hv. Z gradient: [2][C], unlimited, computed once per mesh.
#include <mutex>

#include "mfv_common.hpp"

using namespace mfv;

namespace {

constexpr double kDry = 1e-12;

// Hydrostatic-reconstruction Rusanov flux of one face (same expression order as the
// oracle's orc_sw_face). flux: from L to R, |n| included. bL/bR: bathymetry term of each
// side, to be multiplied by that side's outward normal and subtracted.
template <typename R>
__device__ __forceinline__ R pminT(R a, R b) { return (b < a) ? b : a; }
template <typename R>
__device__ __forceinline__ R pmaxT(R a, R b) { return (b > a) ? b : a; }

template <typename R>
__device__ __forceinline__ void sw_face_t(R etaL, R huL, R hvL, R zfL, R hcL, R zcL, R etaR, R huR, R hvR, R zfR,
                                          R hcR, R zcR, R n0, R n1, R mes, R g, R &F0, R &F1, R &F2, R &bL, R &bR)
{
    const R zero = R(0), half = R(0.5);
    R hL = etaL - zfL;
    if (hL < zero) hL = zero;
    R hR = etaR - zfR;
    if (hR < zero) hR = zero;
    R uL = zero, vL = zero, uR = zero, vR = zero;
    if (hL > R(kDry)) {
        uL = huL / hL;
        vL = hvL / hL;
    }
    if (hR > R(kDry)) {
        uR = huR / hR;
        vR = hvR / hR;
    }
    R zs = pmaxT(zfL, zfR);
    R hsL = hL + zfL - zs;
    if (hsL < zero) hsL = zero;
    R hsR = hR + zfR - zs;
    if (hsR < zero) hsR = zero;
    R qL = uL * n0 + vL * n1;
    R qR = uR * n0 + vR * n1;
    R pL = half * g * hsL * hsL;
    R pR = half * g * hsR * hsR;
    R fL0 = hsL * qL, fL1 = hsL * uL * qL + pL * n0, fL2 = hsL * vL * qL + pL * n1;
    R fR0 = hsR * qR, fR1 = hsR * uR * qR + pR * n0, fR2 = hsR * vR * qR + pR * n1;
    R sL = fabs(qL) / mes + sqrt(g * hsL);
    R sR = fabs(qR) / mes + sqrt(g * hsR);
    R s = pmaxT(sL, sR) * mes;
    F0 = half * (fL0 + fR0) - half * s * (hsR - hsL);
    F1 = half * (fL1 + fR1) - half * s * (hsR * uR - hsL * uL);
    F2 = half * (fL2 + fR2) - half * s * (hsR * vR - hsL * vL);
    bL = half * g * (hL * hL - hsL * hsL) + g * (hL + hcL) * half * (zfL - zcL);
    bR = half * g * (hR * hR - hsR * hsR) + g * (hR + hcR) * half * (zfR - zcR);
}

__device__ __forceinline__ void sw_face(double etaL, double huL, double hvL, double zfL, double hcL, double zcL,
                                        double etaR, double huR, double hvR, double zfR, double hcR, double zcR,
                                        double n0, double n1, double mes, double g, double &F0, double &F1,
                                        double &F2, double &bL, double &bR)
{
    sw_face_t<double>(etaL, huL, hvL, zfL, hcL, zcL, etaR, huR, hvR, zfR, hcR, zcR, n0, n1, mes, g, F0, F1, F2, bL,
                      bR);
}

// time_step_SW (shallowater/fvm_utils.py:16-42): per-cell candidate, faces in faceid order.
__device__ __forceinline__ double sw_cell_dt(const Mesh &m, int i, double h, double hu, double hv, double cfl,
                                             double grav, double Dxx, double Dyy)
{
    const double vol = m.cell_volume[i];
    const double velson = sqrt(grav * h);
    const int nf = m.cell_nfaces[i];
    double lam = 0.;
    for (int j = 0; j < nf; j++) {
        int f = m.cell_faces[(size_t)j * m.C + i];
        const double *n = m.face_normal + 3 * (size_t)f;
        double mesu = m.face_mesure[f];
        double u_n = fabs(hu / h * n[0] + hv / h * n[1]);
        double lam_convect = u_n / mesu + velson;
        lam += lam_convect * mesu;
        double mes = sqrt(n[0] * n[0] + n[1] * n[1]);
        double lam_diff = Dxx * (mes * mes) + Dyy * (mes * mes);
        lam += lam_diff / vol;
    }
    return cfl * vol / lam;
}

__global__ void __launch_bounds__(kBlock) k_sw_time_step(Mesh m, const double *__restrict__ h,
                                                          const double *__restrict__ hu,
                                                          const double *__restrict__ hv, double cfl, double grav,
                                                          double Dxx, double Dyy, double *__restrict__ dtpart)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double d = (i < m.C) ? sw_cell_dt(m, i, h[i], hu[i], hv[i], cfl, grav, Dxx, Dyy) : 1e300;
    wave_min_part(dtpart, (size_t)i, d);
}

// cell_gradient_2d + barthlimiter_2d on eta = h + Z, hu, hv with wall ghosts.
__global__ void __launch_bounds__(kBlock) k_sw_gradients(Mesh m, const double *__restrict__ h,
                                                          const double *__restrict__ hu,
                                                          const double *__restrict__ hv,
                                                          const double *__restrict__ Z,
                                                          double *__restrict__ grad)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const size_t C = m.C;
    const double cx = m.cell_center[3 * i], cy = m.cell_center[3 * i + 1];
    const double ei = h[i] + Z[i], ui = hu[i], vi = hv[i];
    double ixx = 0., iyy = 0., ixy = 0.;
    double je0 = 0., je1 = 0., ju0 = 0., ju1 = 0., jv0 = 0., jv1 = 0.;
    const int nn = m.lsq_ncells[i];
    for (int j = 0; j < nn; j++) {
        int c = m.lsq_cells[(size_t)j * C + i];
        const double *p = m.cell_center + 3 * (size_t)c;
        double jx = p[0] - cx, jy = p[1] - cy;
        ixx += jx * jx;
        iyy += jy * jy;
        ixy += (jx * jy);
        double de = (h[c] + Z[c]) - ei, du = hu[c] - ui, dv = hv[c] - vi;
        je0 += (jx * de);
        je1 += (jy * de);
        ju0 += (jx * du);
        ju1 += (jy * du);
        jv0 += (jx * dv);
        jv1 += (jy * dv);
    }
    const int ng = m.lsq_nghosts[i];
    for (int j = 0; j < ng; j++) {
        int f = m.lsq_ghosts[(size_t)j * C + i];
        int o = m.face_cells[2 * f];
        const double *p = m.face_ghost + 3 * (size_t)f;
        double jx = p[0] - cx, jy = p[1] - cy;
        ixx += jx * jx;
        iyy += jy * jy;
        ixy += (jx * jy);
        double de = (h[o] + Z[o]) - ei, du = (-1 * hu[o]) - ui, dv = (-1 * hv[o]) - vi;
        je0 += (jx * de);
        je1 += (jy * de);
        ju0 += (jx * du);
        ju1 += (jy * du);
        jv0 += (jx * dv);
        jv1 += (jy * dv);
    }
    const double dia = ixx * iyy - ixy * ixy;
    const double ex = (iyy * je0 - ixy * je1) / dia, ey = (ixx * je1 - ixy * je0) / dia;
    const double ux = (iyy * ju0 - ixy * ju1) / dia, uy = (ixx * ju1 - ixy * ju0) / dia;
    const double vx = (iyy * jv0 - ixy * jv1) / dia, vy = (ixx * jv1 - ixy * jv0) / dia;

    // limiter bounds over the face neighbours (barthlimiter_2d, functions2d.py:289-304)
    double emax = ei, emin = ei, umax = ui, umin = ui, vmax = vi, vmin = vi;
    const int nf = m.cell_nfaces[i];
    for (int j = 0; j < nf; j++) {
        int f = m.cell_faces[(size_t)j * C + i];
        int a = m.face_cells[2 * f], b = m.face_cells[2 * f + 1];
        double ea = h[a] + Z[a], ua = hu[a], va = hv[a];
        double eb, ub, vb;
        if (b >= 0) {
            eb = h[b] + Z[b];
            ub = hu[b];
            vb = hv[b];
        } else {
            eb = h[a] + Z[a];
            ub = -1 * hu[a];
            vb = -1 * hv[a];
        }
        emax = pmax(pmax(emax, ea), eb);
        emin = pmin(pmin(emin, ea), eb);
        umax = pmax(pmax(umax, ua), ub);
        umin = pmin(pmin(umin, ua), ub);
        vmax = pmax(pmax(vmax, va), vb);
        vmin = pmin(pmin(vmin, va), vb);
    }
    double pe = 1., pu = 1., pv = 1., qe = 1., qu = 1., qv = 1.;
    for (int j = 0; j < nf; j++) {
        int f = m.cell_faces[(size_t)j * C + i];
        const double *fc = m.face_center + 3 * (size_t)f;
        double r1 = fc[0] - cx, r2 = fc[1] - cy;
        double d;
        d = ex * r1 + ey * r2;
        if (fabs(d) < 1e-8) qe = 1.;
        else {
            if (d > 0. || d < 0.) qe = pmin(1., ((d > 0.) ? emax - ei : emin - ei) / d);  // one division
        }
        pe = pmin(pe, qe);
        d = ux * r1 + uy * r2;
        if (fabs(d) < 1e-8) qu = 1.;
        else {
            if (d > 0. || d < 0.) qu = pmin(1., ((d > 0.) ? umax - ui : umin - ui) / d);  // one division
        }
        pu = pmin(pu, qu);
        d = vx * r1 + vy * r2;
        if (fabs(d) < 1e-8) qv = 1.;
        else {
            if (d > 0. || d < 0.) qv = pmin(1., ((d > 0.) ? vmax - vi : vmin - vi) / d);  // one division
        }
        pv = pmin(pv, qv);
    }
    grad[0 * C + i] = ex;
    grad[1 * C + i] = ey;
    grad[2 * C + i] = pe;
    grad[3 * C + i] = ux;
    grad[4 * C + i] = uy;
    grad[5 * C + i] = pu;
    grad[6 * C + i] = vx;
    grad[7 * C + i] = vy;
    grad[8 * C + i] = pv;
}

struct SwSide {
    double e, u, v, z;
};

// MUSCL value of cell c at face centre fc (same expressions as orc_sw_residuals).
__device__ __forceinline__ SwSide sw_reconstruct(const Mesh &m, int c, const double *fc,
                                                 const double *__restrict__ h, const double *__restrict__ hu,
                                                 const double *__restrict__ hv, const double *__restrict__ Z,
                                                 const double *__restrict__ grad,
                                                 const double *__restrict__ zgrad, double om)
{
    const size_t C = m.C;
    const double *cc = m.cell_center + 3 * (size_t)c;
    double r0 = fc[0] - cc[0], r1 = fc[1] - cc[1];
    SwSide s;
    s.e = (h[c] + Z[c]) + om * grad[2 * C + c] * (grad[0 * C + c] * r0 + grad[1 * C + c] * r1);
    s.u = hu[c] + om * grad[5 * C + c] * (grad[3 * C + c] * r0 + grad[4 * C + c] * r1);
    s.v = hv[c] + om * grad[8 * C + c] * (grad[6 * C + c] * r0 + grad[7 * C + c] * r1);
    s.z = Z[c] + om * (zgrad[c] * r0 + zgrad[C + c] * r1);
    return s;
}

// Residual of cell i (convective rez and bathymetry src), face contributions added in the
// order of the reference's face loops (inner faces by id, then boundary faces by id).
__device__ __forceinline__ void sw_cell_residual(const Mesh &m, int i, const double *__restrict__ h,
                                                 const double *__restrict__ hu, const double *__restrict__ hv,
                                                 const double *__restrict__ Z, const double *__restrict__ grad,
                                                 const double *__restrict__ zgrad, double g, double om,
                                                 double (&rez)[3], double (&src)[3])
{
    rez[0] = rez[1] = rez[2] = 0.;
    src[0] = src[1] = src[2] = 0.;
    const int nf = m.cell_nfaces[i];
    for (int r = 0; r < nf; r++) {
        int s = m.order_split[(size_t)r * m.C + i];
        int f = m.cell_faces[(size_t)s * m.C + i];
        int a = m.face_cells[2 * f], b = m.face_cells[2 * f + 1];
        const double *fc = m.face_center + 3 * (size_t)f;
        const double *n = m.face_normal + 3 * (size_t)f;
        const double mes = m.face_mesure[f];
        SwSide L = sw_reconstruct(m, a, fc, h, hu, hv, Z, grad, zgrad, om);
        double F0, F1, F2, bL, bR;
        if (b >= 0) {
            SwSide R = sw_reconstruct(m, b, fc, h, hu, hv, Z, grad, zgrad, om);
            sw_face(L.e, L.u, L.v, L.z, h[a], Z[a], R.e, R.u, R.v, R.z, h[b], Z[b], n[0], n[1], mes, g, F0, F1, F2,
                    bL, bR);
        } else {
            // wall: Neumann (eta, Z) / nonslip (hu, hv) mirror of the reconstructed face state
            sw_face(L.e, L.u, L.v, L.z, h[a], Z[a], L.e, -1 * L.u, -1 * L.v, L.z, h[a], Z[a], n[0], n[1], mes, g, F0,
                    F1, F2, bL, bR);
        }
        if (a == i) {
            rez[0] -= F0;
            rez[1] -= F1;
            rez[2] -= F2;
            src[1] -= bL * n[0];
            src[2] -= bL * n[1];
        } else {
            rez[0] += F0;
            rez[1] += F1;
            rez[2] += F2;
            src[1] -= -(bR * n[0]);
            src[2] -= -(bR * n[1]);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_sw_residuals(Mesh m, const double *__restrict__ h,
                                                          const double *__restrict__ hu,
                                                          const double *__restrict__ hv,
                                                          const double *__restrict__ Z,
                                                          const double *__restrict__ grad,
                                                          const double *__restrict__ zgrad, double g, double om,
                                                          double *__restrict__ rez, double *__restrict__ src)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    double r[3], s[3];
    sw_cell_residual(m, i, h, hu, hv, Z, grad, zgrad, g, om, r, s);
    const size_t C = m.C;
    for (int k = 0; k < 3; k++) {
        rez[k * C + i] = r[k];
        src[k * C + i] = s[k];
    }
}

__global__ void __launch_bounds__(kBlock) k_update_sw(int C, double *__restrict__ h, double *__restrict__ hu,
                                                       double *__restrict__ hv, const double *__restrict__ rez,
                                                       const double *__restrict__ src,
                                                       const double *__restrict__ dt_dev,
                                                       const double *__restrict__ vol)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double dt = dt_dev[0];
    const size_t n = C;
    h[i] += dt * (rez[i] + src[i]) / vol[i];
    hu[i] += dt * ((rez[n + i] + src[n + i]) / vol[i] + 0. + 0.);
    hv[i] += dt * ((rez[2 * n + i] + src[2 * n + i]) / vol[i] + 0. + 0.);
}

__global__ void __launch_bounds__(kBlock) k_sw_step(Mesh m, const double *__restrict__ Sin,
                                                     double *__restrict__ Sout, const double *__restrict__ Z,
                                                     const double *__restrict__ grad,
                                                     const double *__restrict__ zgrad, double g, double cfl,
                                                     double om, DtRing ring, double *__restrict__ dtpart)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    ring_advance(ring);
    double dloc = 1e300;
    if (i < m.C) {
        const size_t C = m.C;
        const double *h = Sin, *hu = Sin + C, *hv = Sin + 2 * C;
        double r[3], s[3];
        sw_cell_residual(m, i, h, hu, hv, Z, grad, zgrad, g, om, r, s);
        const double dt = ring.cur[0];
        const double vol = m.cell_volume[i];
        double hn = h[i] + dt * (r[0] + s[0]) / vol;
        double un = hu[i] + dt * ((r[1] + s[1]) / vol + 0. + 0.);
        double vn = hv[i] + dt * ((r[2] + s[2]) / vol + 0. + 0.);
        Sout[i] = hn;
        Sout[C + i] = un;
        Sout[2 * C + i] = vn;
        dloc = sw_cell_dt(m, i, hn, un, vn, cfl, g, 0., 0.);
    }
    wave_min_part(dtpart, (size_t)i, dloc);
}

}  // namespace

extern "C" int mfv_time_step_sw(const mfv_mesh *m, const double *h, const double *hu, const double *hv,
                                double cfl, double grav, double Dxx, double Dyy, double *dt_dev, double *dt_host,
                                void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(h && hu && hv && dt_dev, "NULL array");
    if (int rc = mfv_dt_reset(dt_dev, stream)) return rc;
    if (m->nb_cells) {
        const size_t nw = waves_for(m->nb_cells);
        double *part = dt_scratch(nw, as_stream(stream));
        if (!part) return MFV_ENOMEM;
        hipLaunchKernelGGL(k_sw_time_step, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                           to_kernel(m), h, hu, hv, cfl, grav, Dxx, Dyy, part);
        if (int rc = dt_finish(part, nw, dt_dev, as_stream(stream))) return rc;
    }
    return finish_dt(dt_dev, dt_host, stream);
}

extern "C" int mfv_sw_gradients(const mfv_mesh *m, const double *h, const double *hu, const double *hv,
                                const double *Z, double *grad, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(h && hu && hv && Z && grad, "NULL array");
    hipLaunchKernelGGL(k_sw_gradients, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), h, hu, hv, Z, grad);
    return check_launch("k_sw_gradients");
}

extern "C" int mfv_sw_residuals(const mfv_mesh *m, const double *h, const double *hu, const double *hv,
                                const double *Z, const double *grad, const double *zgrad, double grav, int order,
                                double *rez, double *src, void *stream)
{
    if (int rc = validate_mesh(m, 0, 1, 0)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(h && hu && hv && Z && grad && zgrad && rez && src, "NULL array");
    hipLaunchKernelGGL(k_sw_residuals, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), h, hu, hv, Z, grad, zgrad, grav, (double)(order - 1), rez, src);
    return check_launch("k_sw_residuals");
}

extern "C" int mfv_update_sw(const mfv_mesh *m, double *h, double *hu, double *hv, const double *rez,
                             const double *src, const double *dt_dev, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(h && hu && hv && rez && src && dt_dev, "NULL array");
    hipLaunchKernelGGL(k_update_sw, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       m->nb_cells, h, hu, hv, rez, src, dt_dev, m->cell_volume);
    return check_launch("k_update_sw");
}

extern "C" int mfv_sw_step_fused(const mfv_mesh *m, const double *state_in, double *state_out, const double *Z,
                                 const double *grad, const double *zgrad, double grav, double cfl, int order,
                                 double *dt_ring_p, int step, void *stream)
{
    if (int rc = validate_mesh(m, 0, 1, 0)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(state_in && state_out && state_in != state_out && Z && grad && zgrad && dt_ring_p,
                "NULL or aliased array");
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = waves_for(m->nb_cells);
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    hipLaunchKernelGGL(k_sw_step, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m),
                       state_in, state_out, Z, grad, zgrad, grav, cfl, (double)(order - 1), r, part);
    if (int rc = check_launch("k_sw_step")) return rc;
    return dt_finish(part, nw, r.next, as_stream(stream));
}

// =============================================================================================
// Packed shallow-water path.
//
// Per-cell and per-cell-face records so that each gather is one or two contiguous 16-byte
// loads instead of one scattered load per array:
//   cellrec [C][8]      cx, cy, eta, hu, hv, h, Z, vol   (64 bytes, one record per cell)
//   facerec [C][MF][4]  eta_f, hu_f, hv_f, z_f           (MUSCL values at the face centres)
//   cfrec   [C][MF][4]  rx, ry, z_f, z_f across          (face centre - cell centre, the face's
//                                                     bathymetry on this side and on the
//                                                     neighbour's side (own on a wall); static)
//   cfgeo   [C][MF][4]  n0, n1, mesure, 0                (faces.normal / mesure; static)
//   nbr     [MF][C]     (j << 3) | (k << 1) | side, or -1 on a wall: j = cell across the face,
//                       k = the slot of the face in j's faceid row, side = 1 when the cell is
//                       faces.cellid[f][1]
//   order   [C]         residual order: slot of the r-th contribution in bits 2r..2r+1
// Slots follow cells.faceid order. The arithmetic is the unpacked kernels' expression for
// expression, so both paths give the same bits.
namespace {

template <int MF>
__global__ void __launch_bounds__(kBlock) k_sw_pack(Mesh m, mfv_sw_packed p, const double *__restrict__ h,
                                                     const double *__restrict__ hu, const double *__restrict__ hv,
                                                     const double *__restrict__ Z, const double *__restrict__ zgrad,
                                                     double om)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const size_t C = m.C;
    double *r = p.cellrec + 8 * (size_t)i;
    r[0] = m.cell_center[3 * (size_t)i];
    r[1] = m.cell_center[3 * (size_t)i + 1];
    r[2] = h[i] + Z[i];
    r[3] = hu[i];
    r[4] = hv[i];
    r[5] = h[i];
    r[6] = Z[i];
    r[7] = m.cell_volume[i];
    const int nf = m.cell_nfaces[i];
    for (int s = 0; s < MF && s < nf; s++) {
        double *q = p.cfrec + 4 * (MF * (size_t)i + s);
        q[2] = Z[i] + om * (zgrad[i] * q[0] + zgrad[C + i] * q[1]);
    }
}

// cfrec[i][s][3] = z_f of the neighbour's side of face slot s (its cfrec[j][k][2]), own z_f on
// a wall; after k_sw_pack has written every z_f.
template <int MF>
__global__ void __launch_bounds__(kBlock) k_sw_pack_zf_across(Mesh m, mfv_sw_packed p)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const size_t C = m.C;
    const int nf = m.cell_nfaces[i];
    for (int s = 0; s < MF && s < nf; s++) {
        double *q = p.cfrec + 4 * (MF * (size_t)i + s);
        const int code = p.nbr[(size_t)s * C + i];
        q[3] = (code >= 0) ? p.cfrec[4 * (MF * (size_t)(code >> 3) + ((code >> 1) & 3)) + 2] : q[2];
    }
}

template <int MF>
__global__ void __launch_bounds__(kBlock) k_sw_unpack(int C, const double *__restrict__ cellrec,
                                                       double *__restrict__ h, double *__restrict__ hu,
                                                       double *__restrict__ hv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double *r = cellrec + 8 * (size_t)i;
    h[i] = r[5];
    hu[i] = r[3];
    hv[i] = r[4];
}

__device__ __forceinline__ void load_ehv(const double *__restrict__ cellrec, int c, double &cx, double &cy,
                                         double &e, double &u, double &v)
{
    const double2 *r = reinterpret_cast<const double2 *>(cellrec + 8 * (size_t)c);
    double2 a = r[0], b = r[1], d = r[2];
    cx = a.x;
    cy = a.y;
    e = b.x;
    u = b.y;
    v = d.x;
}

// Gradients, limiter and MUSCL face values (k_sw_gradients + the reconstruction of
// sw_reconstruct, written per face slot). KMAX > 0: all stencil indices are loaded first and
// the neighbour records in independent groups of four, so a cell waits for a few memory
// latencies instead of two per stencil entry; the sums keep the stencil order. KMAX = 0:
// plain loop (stencils wider than the largest KMAX).
// Where a reconstruction reads cell records: by cell index in the record array, or by
// position in a block's LDS copy of the records its tile reads (stencil, ghost owners and
// face neighbours then come from the tile's tables).
struct PackedCells {
    const Mesh &m;
    const double *__restrict__ rec;
    int i;
    __device__ __forceinline__ int self() const { return i; }
    __device__ __forceinline__ void ehv(int h, double &cx, double &cy, double &e, double &u, double &v) const
    {
        load_ehv(rec, h, cx, cy, e, u, v);
    }
    __device__ __forceinline__ int stencil(int j) const { return m.lsq_cells[(size_t)j * m.C + i]; }
    template <int KM>
    __device__ __forceinline__ void stencil_counts(const Mesh &mm, int c, int (&h)[KM], int &nn, int &ng) const
    {
        nn = mm.lsq_ncells[c];
        ng = mm.lsq_nghosts[c];
#pragma unroll
        for (int j = 0; j < KM; j++) h[j] = (j < mm.K) ? stencil(j) : 0;
    }
    __device__ __forceinline__ int ghost_owner(int, int f) const { return m.face_cells[2 * f]; }
    // the neighbour code of face slot s: nbr's (global cell << 3 | k << 1 | side, -1 on walls)
    __device__ __forceinline__ int slot_code(const int32_t *__restrict__ nbr, size_t C, int s) const
    {
        return nbr[(size_t)s * C + i];
    }
};

struct PackedStaged {
    const double2 *xy;  // LDS copy of the tile's ext list: (cx, cy), (eta, hu), hv
    const double2 *eu;
    const double *hv;
    const uint16_t *__restrict__ lst;
    const uint16_t *__restrict__ lgh;
    const int32_t *__restrict__ lfn;
    size_t row, I;
    int q, kp;
    __device__ __forceinline__ int self() const { return q; }
    __device__ __forceinline__ void ehv(int h, double &cx, double &cy, double &e, double &u, double &v) const
    {
        const double2 a = xy[h], b = eu[h];
        cx = a.x;
        cy = a.y;
        e = b.x;
        u = b.y;
        v = hv[h];
    }
    __device__ __forceinline__ int stencil(int j) const { return lst[row * kp + j]; }
    template <int KM>
    __device__ __forceinline__ void stencil_counts(const Mesh &mm, int, int (&h)[KM], int &nn, int &ng) const
    {
        stencil_row<KM>(lst, row, kp, mm.K, h, nn, ng);
    }
    __device__ __forceinline__ int ghost_owner(int j, int) const { return lgh[(size_t)j * I + row]; }
    // the neighbour code of face slot s with the cell replaced by its ext-list position
    __device__ __forceinline__ int slot_code(const int32_t *__restrict__, size_t, int s) const
    {
        return lfn[(size_t)s * I + row];
    }
};

// Reconstruction of cell i into out[2 s], out[2 s + 1] = (eta_f, hu_f), (hv_f, z_f) for its nf
// face slots; returns nf.
template <int MF, int KMAX, class Acc>
__device__ __forceinline__ int sw_recon_core(const Mesh &m, const mfv_sw_packed &p, const Acc &A, int i,
                                             double om, double2 *out, double2 *zpair = nullptr, int *lcode = nullptr)
{
    const size_t C = m.C;
    // Loads are issued in batches ahead of the arithmetic: the per-cell counts, the stencil
    // handles of the K rows every cell has, and the face slots of the [MF]-wide tables, which
    // are in bounds whatever the cell's face count (loads behind the count-dependent control
    // flow would each cost a memory latency of their own).
    constexpr int KA = KMAX > 0 ? KMAX : 8;
    int cidx[KA];  // stencil handles (KMAX > 0) and the stencil and ghost counts, in one batch
    int nn, ng;
    A.template stencil_counts<KA>(m, i, cidx, nn, ng);
    const int nf = m.cell_nfaces[i];
    int codes[MF], nbh[MF];
#pragma unroll
    for (int s = 0; s < MF; s++) codes[s] = A.slot_code(p.nbr, C, s);
#pragma unroll
    for (int s = 0; s < MF; s++) nbh[s] = codes[s] >> 3;  // the cell across, in the accessor's numbering
    double cx, cy, ei, ui, vi;
    A.ehv(A.self(), cx, cy, ei, ui, vi);
    double ixx = 0., iyy = 0., ixy = 0.;
    double je0 = 0., je1 = 0., ju0 = 0., ju1 = 0., jv0 = 0., jv1 = 0.;
    auto accum = [&](double px, double py, double ec, double uc, double vc) {
        double jx = px - cx, jy = py - cy;
        ixx += jx * jx;
        iyy += jy * jy;
        ixy += (jx * jy);
        double de = ec - ei, du = uc - ui, dv = vc - vi;
        je0 += (jx * de);
        je1 += (jy * de);
        ju0 += (jx * du);
        ju1 += (jy * du);
        jv0 += (jx * dv);
        jv1 += (jy * dv);
    };
    if (KMAX > 0) {
#pragma unroll
        for (int j = 0; j < KMAX; j++) cidx[j] = (j < nn) ? cidx[j] : A.self();
        // Every group's entries are loaded (entries past nn read the own cell, a valid handle)
        // and only the accumulation depends on nn: conditional loads made the compiler copy
        // registers for the entries they may skip. Groups past the wave's largest stencil are
        // skipped; the scheduling barrier keeps the next group's loads from being hoisted
        // (which would hold all of them in registers).
        const int nn_max = wave_max_int(nn);
#pragma unroll
        for (int j0 = 0; j0 < KMAX; j0 += 4) {
            if (j0 >= nn_max) break;
            double px[4], py[4], ec[4], uc[4], vc[4];
#pragma unroll
            for (int k = 0; k < 4; k++) A.ehv(cidx[j0 + k], px[k], py[k], ec[k], uc[k], vc[k]);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (j0 + k < nn) accum(px[k], py[k], ec[k], uc[k], vc[k]);
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        for (int j = 0; j < nn; j++) {
            int c = A.stencil(j);
            double px, py, ec, uc, vc;
            A.ehv(c, px, py, ec, uc, vc);
            accum(px, py, ec, uc, vc);
        }
    }
    double2 cfa[MF], cfb[MF];  // rx, ry, z_f of each slot
#pragma unroll
    for (int s = 0; s < MF; s++) {
        const double2 *q = reinterpret_cast<const double2 *>(p.cfrec + 4 * (MF * (size_t)i + s));
        cfa[s] = q[0];
        cfb[s] = q[1];
    }
    for (int j = 0; j < ng; j++) {
        int f = m.lsq_ghosts[(size_t)j * C + i];
        int o = A.ghost_owner(j, f);
        const double *g = m.face_ghost + 3 * (size_t)f;
        double px, py, eo, uo, vo;
        A.ehv(o, px, py, eo, uo, vo);
        double jx = g[0] - cx, jy = g[1] - cy;
        ixx += jx * jx;
        iyy += jy * jy;
        ixy += (jx * jy);
        double de = eo - ei, du = (-1 * uo) - ui, dv = (-1 * vo) - vi;
        je0 += (jx * de);
        je1 += (jy * de);
        ju0 += (jx * du);
        ju1 += (jy * du);
        jv0 += (jx * dv);
        jv1 += (jy * dv);
    }
    // face neighbours' values (the own cell's on a wall or a missing slot: not used)
    double ne[MF], nu[MF], nv[MF];
#pragma unroll
    for (int s = 0; s < MF; s++) {
        double px, py;
        A.ehv(codes[s] >= 0 ? nbh[s] : A.self(), px, py, ne[s], nu[s], nv[s]);
    }
    const double dia = ixx * iyy - ixy * ixy;
    const double ex = (iyy * je0 - ixy * je1) / dia, ey = (ixx * je1 - ixy * je0) / dia;
    const double ux = (iyy * ju0 - ixy * ju1) / dia, uy = (ixx * ju1 - ixy * ju0) / dia;
    const double vx = (iyy * jv0 - ixy * jv1) / dia, vy = (ixx * jv1 - ixy * jv0) / dia;

    double emax = ei, emin = ei, umax = ui, umin = ui, vmax = vi, vmin = vi;
#pragma unroll
    for (int s = 0; s < MF; s++) {
        if (s >= nf) break;
        const int code = codes[s];
        double ea, ua, va, eb, ub, vb;
        if (code >= 0) {
            const double ej = ne[s], uj = nu[s], vj = nv[s];
            if (code & 1) {
                ea = ej; ua = uj; va = vj;
                eb = ei; ub = ui; vb = vi;
            } else {
                ea = ei; ua = ui; va = vi;
                eb = ej; ub = uj; vb = vj;
            }
        } else {
            ea = ei; ua = ui; va = vi;
            eb = ei; ub = -1 * ui; vb = -1 * vi;
        }
        emax = pmax(pmax(emax, ea), eb);
        emin = pmin(pmin(emin, ea), eb);
        umax = pmax(pmax(umax, ua), ub);
        umin = pmin(pmin(umin, ua), ub);
        vmax = pmax(pmax(vmax, va), vb);
        vmin = pmin(pmin(vmin, va), vb);
    }
    double pe = 1., pu = 1., pv = 1., qe = 1., qu = 1., qv = 1.;
#pragma unroll
    for (int s = 0; s < MF; s++) {
        if (s >= nf) break;
        double r1 = cfa[s].x, r2 = cfa[s].y, d;
        d = ex * r1 + ey * r2;
        if (fabs(d) < 1e-8) qe = 1.;
        else {
            if (d > 0. || d < 0.) qe = pmin(1., ((d > 0.) ? emax - ei : emin - ei) / d);  // one division
        }
        pe = pmin(pe, qe);
        d = ux * r1 + uy * r2;
        if (fabs(d) < 1e-8) qu = 1.;
        else {
            if (d > 0. || d < 0.) qu = pmin(1., ((d > 0.) ? umax - ui : umin - ui) / d);  // one division
        }
        pu = pmin(pu, qu);
        d = vx * r1 + vy * r2;
        if (fabs(d) < 1e-8) qv = 1.;
        else {
            if (d > 0. || d < 0.) qv = pmin(1., ((d > 0.) ? vmax - vi : vmin - vi) / d);  // one division
        }
        pv = pmin(pv, qv);
    }
#pragma unroll
    for (int s = 0; s < MF; s++) {
        if (s >= nf) break;
        double2 a, b;
        a.x = ei + om * pe * (ex * cfa[s].x + ey * cfa[s].y);
        a.y = ui + om * pu * (ux * cfa[s].x + uy * cfa[s].y);
        b.x = vi + om * pv * (vx * cfa[s].x + vy * cfa[s].y);
        b.y = cfb[s].x;
        out[2 * s] = a;
        out[2 * s + 1] = b;
    }
    if (zpair) {  // z_f of both sides of every slot (cfrec[i][s][2..3]), for a caller that keeps them
#pragma unroll
        for (int s = 0; s < MF; s++) zpair[s] = cfb[s];
    }
    if (lcode) {  // the neighbour codes in the accessor's numbering: (position << 3) | (k << 1) | side
#pragma unroll
        for (int s = 0; s < MF; s++) lcode[s] = codes[s];
    }
    return nf;
}

template <int MF, int KMAX>
__global__ void __launch_bounds__(kBlock) k_sw_recon(Mesh m, mfv_sw_packed p, double om)
{
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    double2 rec[2 * MF];
    const int nf = sw_recon_core<MF, KMAX>(m, p, PackedCells{m, p.cellrec, i}, i, om, rec);
    double2 *o = reinterpret_cast<double2 *>(p.facerec + 4 * (MF * (size_t)i));
#pragma unroll
    for (int s = 0; s < MF; s++)
        if (s < nf) {
            o[2 * s] = rec[2 * s];
            o[2 * s + 1] = rec[2 * s + 1];
        }
}

// Face values of the flux gather, addressed by face-value index: own_idx(s) is this cell's
// slot s, nb_idx(code) the other side of a face (code = its neighbour code in the numbering
// the values are stored in); vals(k) gives (eta_f, hu_f) and hv_f of index k; zpair(s, code)
// the z_f of both sides of slot s (own z_f on a wall). GlobalFaceVals: the facerec array
// (codes are global); StagedFaceVals: a block's LDS face values without z_f (codes are
// block-local), z_f of both sides from the static cfrec.
struct GlobalFaceVals {
    const double *__restrict__ fr;
    int i;
    template <int MF>
    __device__ __forceinline__ int own_idx(int s) const { return MF * i + s; }
    template <int MF>
    __device__ __forceinline__ int nb_idx(int code) const { return MF * (code >> 3) + ((code >> 1) & 3); }
    __device__ __forceinline__ void vals(int k, double2 &eu, double &v) const
    {
        const double2 *q = reinterpret_cast<const double2 *>(fr + 4 * (size_t)k);
        eu = q[0];
        v = q[1].x;
    }
    template <int MF>
    __device__ __forceinline__ double2 zpair(int s, int code) const
    {
        const double zo = fr[4 * (MF * (size_t)i + s) + 3];
        return make_double2(zo, code >= 0 ? fr[4 * (size_t)nb_idx<MF>(code) + 3] : zo);
    }
};

struct StagedFaceVals {
    const double2 *eu;  // [items][MF] (eta_f, hu_f)
    const double *v;    // [items][MF] hv_f
    slot_d zo, zn;      // z_f of both sides of the cell's slots, kept in registers since phase A
    int q;
    template <int MF>
    __device__ __forceinline__ int own_idx(int s) const { return MF * q + s; }
    template <int MF>
    __device__ __forceinline__ int nb_idx(int code) const { return MF * (code >> 3) + ((code >> 1) & 3); }
    __device__ __forceinline__ void vals(int k, double2 &a, double &b) const
    {
        a = eu[k];
        b = v[k];
    }
    template <int MF>
    __device__ __forceinline__ double2 zpair(int s, int) const
    {
        return make_double2(zo[s], zn[s]);
    }
};

// The float64 cell values of the update: hu, hv, h, Z, vol of the cell record.
struct CellVals {
    double hu, hv, h, Z, vol;
};

__device__ __forceinline__ CellVals cell_vals(const double *__restrict__ rec)
{
    const double2 *r2 = reinterpret_cast<const double2 *>(rec);
    const double2 c1 = r2[1], c2 = r2[2], c3 = r2[3];
    return CellVals{c1.y, c2.x, c2.y, c3.x, c3.y};
}

// One cell of the step: fluxes + bathymetry terms + update_SW from the cell values cv, new eta,
// hu, hv, h written into the record rout; returns the next-dt candidate. fv: where the face
// values come from; code_of(s): the neighbour code of slot s in fv's numbering. The slot data
// of the [MF]-wide tables (in bounds whatever nf) is loaded in one batch before the arithmetic:
// in the residual order for the fluxes and, for the dt, the geometry again in faceid order.
template <int MF, class FV, class Codes>
__device__ __forceinline__ double sw_flux_packed_core(const mfv_sw_packed &p, int i, const CellVals cv,
                                                      double *rout, const FV &fv, Codes code_of, int nf,
                                                      unsigned ord, double g, double cfl, double dt)
{
    int sl[MF], cd[MF];
    double gn0[MF], gn1[MF], gms[MF], zo[MF], zn[MF];
#pragma unroll
    for (int r = 0; r < MF; r++) {  // slot s of the r-th contribution is in bits 2r..2r+1
        const int s = min((int)((ord >> (2 * r)) & 3), MF - 1);
        sl[r] = s;
        cd[r] = code_of(s);
        const double2 *gq = reinterpret_cast<const double2 *>(p.cfgeo + 4 * (MF * (size_t)i + s));
        const double2 ga = gq[0], gb = gq[1];
        gn0[r] = ga.x;
        gn1[r] = ga.y;
        gms[r] = gb.x;
    }
#pragma unroll
    for (int r = 0; r < MF; r++) {
        const double2 z = fv.template zpair<MF>(sl[r], cd[r]);
        zo[r] = z.x;
        zn[r] = z.y;
    }
    // the dt's geometry in faceid order: slot s's values are those loaded for any r with
    // sl[r] == s (the first nf residual positions hold each of the cell's slots once), picked
    // from registers instead of loaded again
    double dn0[MF], dn1[MF], dms[MF];
#pragma unroll
    for (int s = 0; s < MF; s++) {
        double a = gn0[0], b = gn1[0], c = gms[0];
#pragma unroll
        for (int r = 1; r < MF; r++) {
            a = sl[r] == s ? gn0[r] : a;
            b = sl[r] == s ? gn1[r] : b;
            c = sl[r] == s ? gms[r] : c;
        }
        dn0[s] = a;
        dn1[s] = b;
        dms[s] = c;
    }
    const double hu_i = cv.hu, hv_i = cv.hv, h_i = cv.h, Z_i = cv.Z, vol = cv.vol;
    double r0 = 0., r1 = 0., rr2 = 0., s1 = 0., s2 = 0.;
#pragma unroll
    for (int r = 0; r < MF; r++) {
        if (r >= nf) break;
        const double n0 = gn0[r], n1 = gn1[r], mes = gms[r];
        const int code = cd[r];
        // One evaluation in the face's orientation whichever side the cell is on (a wave holds
        // cells of both sides, so two branches would each run for the whole wave): the L and R
        // face values are read at selected indices, and on a wall R is the own value with the
        // momenta negated (-1 * x, as the two-sided form). The own side's bathymetry term is
        // bL (own = L) or bR (own = R); the signs are exact multiplications by -1 or 1.
        const bool wall = code < 0, own_R = !wall && (code & 1);
        const int ko = fv.template own_idx<MF>(sl[r]);
        const int kn = wall ? ko : fv.template nb_idx<MF>(code);
        double2 eL, eR;
        double vL, vR;
        fv.vals(own_R ? kn : ko, eL, vL);
        fv.vals(own_R ? ko : kn, eR, vR);
        const double mR = wall ? -1. : 1.;
        const double zO = zo[r], zN = zn[r];
        double F0, F1, F2, bL, bR;
        sw_face(eL.x, eL.y, vL, own_R ? zN : zO, h_i, Z_i, eR.x, mR * eR.y, mR * vR, own_R ? zO : zN, h_i, Z_i, n0,
                n1, mes, g, F0, F1, F2, bL, bR);
        const double sg = own_R ? 1. : -1.;  // rez -= F (own = L) is rez += -1 * F
        r0 += sg * F0;
        r1 += sg * F1;
        rr2 += sg * F2;
        const double b = own_R ? bR : bL;  // s -= b n (own = L), s -= -(b n) (own = R)
        s1 -= -sg * (b * n0);
        s2 -= -sg * (b * n1);
    }
    const double s0 = 0.;
    double hn = h_i + dt * (r0 + s0) / vol;
    double un = hu_i + dt * ((r1 + s1) / vol + 0. + 0.);
    double vn = hv_i + dt * ((rr2 + s2) / vol + 0. + 0.);
    double2 w1, w2;
    w1.x = hn + Z_i;
    w1.y = un;
    w2.x = vn;
    w2.y = hn;
    reinterpret_cast<double2 *>(rout)[1] = w1;
    reinterpret_cast<double2 *>(rout)[2] = w2;
    // time_step_SW of the new state, faces in faceid order
    const double velson = sqrt(g * hn);
    double lam = 0.;
#pragma unroll
    for (int s = 0; s < MF; s++) {
        if (s >= nf) break;
        const double n0 = dn0[s], n1 = dn1[s], mesu = dms[s];
        double u_n = fabs(un / hn * n0 + vn / hn * n1);
        double lam_convect = u_n / mesu + velson;
        lam += lam_convect * mesu;
        // time_step_SW then adds (Dxx |n|^2 + Dyy |n|^2) / vol with Dxx = Dyy = 0: +0 for finite
        // normals and volumes, which leaves lam (>= 0) unchanged
    }
    return cfl * vol / lam;
}

// Fluxes + bathymetry terms + update_SW + next dt, in place on cellrec.
template <int MF>
__global__ void __launch_bounds__(kBlock) k_sw_flux_packed(Mesh m, mfv_sw_packed p, double g, double cfl,
                                                            DtRing ring, double *__restrict__ dtpart)
{
    int i = xcd_block() * blockDim.x + threadIdx.x;
    ring_advance(ring);
    double dloc = 1e300;
    if (i < m.C) {
        const size_t C = m.C;
        double *rec = p.cellrec + 8 * (size_t)i;
        const int32_t *__restrict__ nbr = p.nbr;
        auto code_of = [nbr, C, i](int s) { return nbr[(size_t)s * C + i]; };
        dloc = sw_flux_packed_core<MF>(p, i, cell_vals(rec), rec, GlobalFaceVals{p.facerec, i}, code_of, m.cell_nfaces[i],
                                       p.order[i], g, cfl, ring.cur[0]);
    }
    wave_min_part(dtpart, (size_t)i, dloc);
}

// Fused float64 step on tiles (the layout of mfv_sw_tiles, built from the same neighbour
// codes): the block copies the fields the reconstruction reads (cx, cy, eta, hu, hv) of every
// cell its tile reads into LDS, reconstructs its cells and halo into LDS face values, then
// updates its cells into cellrec_out (the step is double-buffered: neighbouring tiles still
// read the old records). The per-cell functions are those of the two-kernel path, so the bits
// are the same. LDS: 40 bytes per ext cell + 24 per item face slot (z_f comes from cfrec),
// so that four blocks of 256 threads fit a CU on Hilbert-ordered triangle meshes.
struct TilesP {
    int E;
    const int32_t *__restrict__ cell0;
    const int32_t *__restrict__ ncells;
    const int32_t *__restrict__ nhalo;
    const int32_t *__restrict__ eptr;
    const int32_t *__restrict__ ext;
    const int32_t *__restrict__ iptr;
    const int32_t *__restrict__ lnbr;
    const uint16_t *__restrict__ lst;
    const uint16_t *__restrict__ lgh;
    const int32_t *__restrict__ lfn;
    int I;
};

// LDS bytes of the fused float64 step: ext cells (E rounded up to even, so the arrays after
// the odd-sized hv array stay 16-byte aligned) and item face values.
inline size_t packed_tiled_lds(int max_ext, int max_items, int MF)
{
    const size_t E = ((size_t)max_ext + 1) & ~(size_t)1;
    return E * 40 + (size_t)max_items * MF * 24;
}

template <int MF, int T, int KMAX>
__global__ void __launch_bounds__(T) k_sw_step_packed_tiled(Mesh m, mfv_sw_packed p, TilesP t,
                                                            double *__restrict__ rec_out, double om, double g,
                                                            double cfl, DtRing ring, double *__restrict__ dtpart)
{
    extern __shared__ double2 lds2[];
    const int E = (t.E + 1) & ~1;
    double2 *s_xy = lds2;                                            // [E]
    double2 *s_eu = lds2 + E;                                        // [E]
    double *s_v = reinterpret_cast<double *>(lds2 + 2 * E);          // [E]
    double2 *f_eu = lds2 + 2 * E + E / 2;                            // [items][MF]
    const int tile = xcd_block();
    const int c0 = t.cell0[tile], nc = t.ncells[tile], nh = t.nhalo[tile];
    const int e0 = t.eptr[tile], ne = t.eptr[tile + 1] - e0;
    double *f_v = reinterpret_cast<double *>(f_eu + MF * (nc + nh));  // [items][MF]
    const size_t r0 = t.iptr[tile];
    ring_advance(ring);
    for (int e = threadIdx.x; e < nc + ne; e += 2 * T) {  // two cells per pass: one latency for both
        const int2 c = tile_ext_pair(t.ext, c0, nc, e0, ne, e, T);
        const int eb = (e + T < nc + ne) ? e + T : e;
        const double2 *sa = reinterpret_cast<const double2 *>(p.cellrec + 8 * (size_t)c.x);
        const double2 *sb = reinterpret_cast<const double2 *>(p.cellrec + 8 * (size_t)c.y);
        const double2 a0 = sa[0], a1 = sa[1], a2 = sa[2], b0 = sb[0], b1 = sb[1], b2 = sb[2];
        s_xy[e] = a0;
        s_eu[e] = a1;
        s_v[e] = a2.x;
        s_xy[eb] = b0;
        s_eu[eb] = b1;
        s_v[eb] = b2.x;
    }
    __syncthreads();
    // Phase A. A tile cell is reconstructed and updated by the same thread (q = threadIdx.x is
    // in the first pass), which keeps the cell's z_f pairs and its h, Z, vol (loaded while the
    // staging reads the same record lines) in registers for phase B.
    const int q = threadIdx.x;
    double2 zp[MF];
    double hZv[3] = {0., 0., 0.};
    unsigned ord_q = 0;  // the residual order, loaded here so that phase B starts with the slot loads
    if (q < nc) {
        const double *rc = p.cellrec + 8 * (size_t)(c0 + q);
        hZv[0] = rc[5];
        hZv[1] = rc[6];
        hZv[2] = rc[7];
        ord_q = p.order[c0 + q];
    }
    int lc[MF];  // the tile cell's block-local neighbour codes (t.lnbr's values), from phase A
    int nf_q = 0;
    auto recon_item = [&](int qq, double2 *keep, int *lkeep) {
        const int c = (qq < nc) ? c0 + qq : t.ext[e0 + (qq - nc)];
        double2 rec[2 * MF];
        const PackedStaged acc{s_xy, s_eu, s_v, t.lst, t.lgh, t.lfn, r0 + qq, (size_t)t.I, qq, MFV_SW_TILES_KP(m.K)};
        const int nf = sw_recon_core<MF, KMAX>(m, p, acc, c, om, rec, keep, lkeep);
        if (keep) nf_q = nf;
#pragma unroll
        for (int s = 0; s < MF; s++)
            if (s < nf) {
                f_eu[MF * qq + s] = rec[2 * s];
                f_v[MF * qq + s] = rec[2 * s + 1].x;
            }
    };
    if (q < nc + nh) recon_item(q, zp, lc);
    for (int qq = q + T; qq < nc + nh; qq += T) recon_item(qq, nullptr, nullptr);  // items beyond one per thread
    __syncthreads();
    if ((q & ~63) >= nc) return;  // the whole wave lies past the tile: it owns no partial
    double dloc = 1e300;
    if (q < nc) {
        const int i = c0 + q;
        // the codes of t.lnbr (the tile cell's lfn row) kept from the reconstruction; a slot index
        // picks one by selects (an array indexed by a lane-dependent slot would live in scratch)
        const int l0 = lc[0], l1 = lc[1], l2 = lc[2], l3 = lc[MF - 1];
        auto code_of = [l0, l1, l2, l3](int s) { return s == 0 ? l0 : s == 1 ? l1 : s == 2 ? l2 : l3; };
        StagedFaceVals fv{f_eu, f_v, 0., 0., q};
#pragma unroll
        for (int s = 0; s < MF; s++) {
            fv.zo[s] = zp[s].x;
            fv.zn[s] = zp[s].y;
        }
        const CellVals cv{s_eu[q].y, s_v[q], hZv[0], hZv[1], hZv[2]};
        dloc = sw_flux_packed_core<MF>(p, i, cv, rec_out + 8 * (size_t)i, fv, code_of, nf_q, ord_q, g, cfl,
                                       ring.cur[0]);
    }
    wave_min_part(dtpart, (size_t)c0 + q, dloc);
}

}  // namespace

static int check_packed(const mfv_mesh *m, const mfv_sw_packed *p)
{
    MFV_REQUIRE(p != nullptr, "packed descriptor is NULL");
    MFV_REQUIRE(p->nb_cells == m->nb_cells && p->max_cell_faces == m->max_cell_faces,
                "packed descriptor does not match the mesh");
    MFV_REQUIRE(m->max_cell_faces == 3 || m->max_cell_faces == 4, "packed path supports 3 or 4 faces per cell");
    MFV_REQUIRE(p->cellrec && p->facerec && p->cfrec && p->cfgeo && p->nbr && p->order, "packed descriptor "
                "has NULL arrays");
    return MFV_OK;
}

extern "C" int mfv_sw_pack(const mfv_mesh *m, const mfv_sw_packed *p, const double *h, const double *hu,
                           const double *hv, const double *Z, const double *zgrad, int order, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_packed(m, p)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(h && hu && hv && Z && zgrad, "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    const bool tri = m->max_cell_faces == 3;
    hipLaunchKernelGGL(tri ? k_sw_pack<3> : k_sw_pack<4>, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0,
                       as_stream(stream), to_kernel(m), *p, h, hu, hv, Z, zgrad, (double)(order - 1));
    if (int rc = check_launch("k_sw_pack")) return rc;
    hipLaunchKernelGGL(tri ? k_sw_pack_zf_across<3> : k_sw_pack_zf_across<4>, dim3(grid_for(m->nb_cells)),
                       dim3(kBlock), 0, as_stream(stream), to_kernel(m), *p);
    return check_launch("k_sw_pack_zf_across");
}

extern "C" int mfv_sw_unpack(const mfv_mesh *m, const mfv_sw_packed *p, double *h, double *hu, double *hv,
                             void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_packed(m, p)) return rc;
    MFV_REQUIRE(h && hu && hv, "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    hipLaunchKernelGGL(k_sw_unpack<3>, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       m->nb_cells, p->cellrec, h, hu, hv);
    return check_launch("k_sw_unpack");
}

extern "C" int mfv_sw_recon(const mfv_mesh *m, const mfv_sw_packed *p, int order, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_packed(m, p)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    if (m->nb_cells == 0) return MFV_OK;
    const bool tri = m->max_cell_faces == 3;
    const int K = m->lsq_width;
    auto k = (K <= 12)   ? (tri ? k_sw_recon<3, 12> : k_sw_recon<4, 12>)
             : (K <= 16) ? (tri ? k_sw_recon<3, 16> : k_sw_recon<4, 16>)
             : (K <= 24) ? (tri ? k_sw_recon<3, 24> : k_sw_recon<4, 24>)
                         : (tri ? k_sw_recon<3, 0> : k_sw_recon<4, 0>);
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), *p,
                       (double)(order - 1));
    return check_launch("k_sw_recon");
}

static int check_tiles_packed(const mfv_mesh *m, const mfv_sw_tiles *t)
{
    MFV_REQUIRE(t != nullptr, "tiles descriptor is NULL");
    MFV_REQUIRE(t->tile_size == 128 || t->tile_size == 256 || t->tile_size == 512,
                "tile_size %d not in {128, 256, 512}", t->tile_size);
    MFV_REQUIRE(t->nb_tiles >= 1 && t->max_ext >= 1 && t->nb_items >= 1 && t->max_items >= 1, "bad tile counts");
    MFV_REQUIRE(t->tile_cell0 && t->tile_ncells && t->tile_nhalo && t->ext_ptr && t->ext_cells && t->item_ptr &&
                    t->lnbr && t->lst && t->lgh && t->lfn,
                "tiles descriptor has NULL arrays");
    const size_t lds = packed_tiled_lds(t->max_ext, t->max_items, m->max_cell_faces);
    MFV_REQUIRE(lds <= 64 * 1024, "tiles of %d cells need %zu bytes of LDS for the float64 step (> 64 KiB): use "
                "smaller tiles", t->tile_size, lds);
    return MFV_OK;
}

template <int MF, int T>
static void launch_packed_tiled(const mfv_mesh *m, const mfv_sw_packed *p, const mfv_sw_tiles *t, double *out,
                                double om, double g, double cfl, DtRing r, double *part, hipStream_t st)
{
    const TilesP tp{t->max_ext, t->tile_cell0, t->tile_ncells, t->tile_nhalo, t->ext_ptr, t->ext_cells,
                    t->item_ptr, t->lnbr, t->lst, t->lgh, t->lfn, t->nb_items};
    const size_t lds = packed_tiled_lds(t->max_ext, t->max_items, MF);
    if (m->lsq_width <= 16)
        hipLaunchKernelGGL((k_sw_step_packed_tiled<MF, T, 16>), dim3(t->nb_tiles), dim3(T), lds, st, to_kernel(m),
                           *p, tp, out, om, g, cfl, r, part);
    else
        hipLaunchKernelGGL((k_sw_step_packed_tiled<MF, T, 0>), dim3(t->nb_tiles), dim3(T), lds, st, to_kernel(m),
                           *p, tp, out, om, g, cfl, r, part);
}

extern "C" int mfv_sw_step_packed_tiled(const mfv_mesh *m, const mfv_sw_packed *p, const mfv_sw_tiles *t,
                                        double *cellrec_out, double grav, double cfl, int order, double *dt_ring_p,
                                        int step, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_packed(m, p)) return rc;
    if (int rc = check_tiles_packed(m, t)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(dt_ring_p && cellrec_out && cellrec_out != p->cellrec, "cellrec_out must be a second record buffer");
    if (m->nb_cells == 0) return MFV_OK;
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = ((size_t)m->nb_cells + 63) / 64;
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    const double om = (double)(order - 1);
    hipStream_t st = as_stream(stream);
    const bool tri = m->max_cell_faces == 3;
    switch (t->tile_size) {
    case 128:
        if (tri) launch_packed_tiled<3, 128>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        else launch_packed_tiled<4, 128>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        break;
    case 256:
        if (tri) launch_packed_tiled<3, 256>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        else launch_packed_tiled<4, 256>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        break;
    default:
        if (tri) launch_packed_tiled<3, 512>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        else launch_packed_tiled<4, 512>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        break;
    }
    if (int rc = check_launch("k_sw_step_packed_tiled")) return rc;
    return dt_finish(part, nw, r.next, st);
}

// ---------------------------------------------------------------------------------------------
// The fused float64 step without device allocation. Its only scratch is the per-wave dt
// partials; here they live in a static device array (one copy per GPU, for the life of the
// process) instead of the per-stream buffer that dt_scratch grows with hipMalloc. Host code
// never passes the array's name to a kernel (on the host it only names the symbol): the
// device address comes from hipGetSymbolAddress, once per device, and is kept per device.
// Every wave with cells overwrites its slot before the fold reads it, so nothing needs
// clearing between calls; calls that overlap on one GPU (several streams or threads) share
// the array and race.
static __device__ __align__(256) double scratch[MFV_NOALLOC_MAX_WAVES];

static std::mutex g_scratch_mu;
static double *g_scratch_dev[64];

static double *noalloc_partials()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("no current device");
        return nullptr;
    }
    if (dev < 0 || dev >= 64) {
        set_error("device ordinal %d out of range", dev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_dev[dev] = static_cast<double *>(ptr);
    }
    return g_scratch_dev[dev];
}

extern "C" int mfv_sw_step_packed_tiled_noalloc(const mfv_mesh *m, const mfv_sw_packed *p, const mfv_sw_tiles *t,
                                                double *cellrec_out, double grav, double cfl, int order,
                                                double *dt_ring_p, int step, void *stream)
{
    MFV_REQUIRE(m != nullptr, "mesh descriptor is NULL");
    MFV_REQUIRE(m->nb_cells >= 0 && ((size_t)m->nb_cells + 63) / 64 <= (size_t)MFV_NOALLOC_MAX_WAVES,
                "%d cells need %zu dt partials: the allocation-free step holds at most %d (%d cells); use "
                "mfv_sw_step_packed_tiled",
                m->nb_cells, ((size_t)m->nb_cells + 63) / 64, MFV_NOALLOC_MAX_WAVES, 64 * MFV_NOALLOC_MAX_WAVES);
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_packed(m, p)) return rc;
    if (int rc = check_tiles_packed(m, t)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(dt_ring_p && cellrec_out && cellrec_out != p->cellrec, "cellrec_out must be a second record buffer");
    if (m->nb_cells == 0) return MFV_OK;
    double *part = noalloc_partials();
    if (!part) return MFV_EHIP;
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = ((size_t)m->nb_cells + 63) / 64;
    const double om = (double)(order - 1);
    hipStream_t st = as_stream(stream);
    const bool tri = m->max_cell_faces == 3;
    switch (t->tile_size) {
    case 128:
        if (tri) launch_packed_tiled<3, 128>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        else launch_packed_tiled<4, 128>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        break;
    case 256:
        if (tri) launch_packed_tiled<3, 256>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        else launch_packed_tiled<4, 256>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        break;
    default:
        if (tri) launch_packed_tiled<3, 512>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        else launch_packed_tiled<4, 512>(m, p, t, cellrec_out, om, grav, cfl, r, part, st);
        break;
    }
    if (int rc = check_launch("k_sw_step_packed_tiled")) return rc;
    return dt_finish(part, nw, r.next, st);
}

extern "C" int mfv_sw_flux_packed(const mfv_mesh *m, const mfv_sw_packed *p, double grav, double cfl,
                                  double *dt_ring_p, int step, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_packed(m, p)) return rc;
    MFV_REQUIRE(dt_ring_p, "dt ring is NULL");
    if (m->nb_cells == 0) return MFV_OK;
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = waves_for(m->nb_cells);
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    auto k = (m->max_cell_faces == 3) ? k_sw_flux_packed<3> : k_sw_flux_packed<4>;
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), *p, grav,
                       cfl, r, part);
    if (int rc = check_launch("k_sw_flux_packed")) return rc;
    return dt_finish(part, nw, r.next, as_stream(stream));
}
