// libmfv shallow water with the reference's own numerical flux and source: the SRNH
// Roe-type flux on five fields (h, hu, hv, hc, Z), MUSCL on h and hc limited by hc's psi,
// the well-balanced bathymetry source of triangles, and update_SW.
//
// The residual is a cell-centric gather: each cell evaluates the flux of each of its faces
// in the face's own orientation (left = faces.cellid[f][0]) and adds -F (left) or +F
// (right) in the order of the reference's face loops (inner faces by id, then boundary
// faces by id), so the float64 sums follow the reference's scatter term by term. The flux
// uses acos, cos and pow of the device math library; those may differ from the host
// libm in the last bit, so parity with the CPU restatement is to rounding, not bitwise.
#include "mfv_common.hpp"

namespace mfv {

struct Fields5 {
    const double *__restrict__ f[5];
};
struct Out5 {
    double *__restrict__ f[5];
};

// Sign matrix of the Roe state times w_dif, row by row (shallowater/fvm_utils.py:220-327;
// As = 0). out[k] = sum_j r[k][j] * w_dif[j], summed j = 0..4 from 0.
__device__ __forceinline__ void srnh_sign_apply(double hroe, double uroe, double vroe, double croe, double grav,
                                                const double *w_dif, double *out)
{
    const double As = 0.;
    const double xi = 1. / (1. - 0.4);
    const double PI = 3.141592653589793;
    const double d = As * xi * (3. * (uroe * uroe) + vroe * vroe);
    const double sound = sqrt(grav * hroe);
    const double Q = -((uroe * uroe) + 3. * grav * (hroe + d)) / 9.;
    const double R = uroe * (9. * grav * (2. * hroe - d) - 2. * (uroe * uroe)) / 54.;
    const double theta = acos(R / sqrt(-pow(Q, 3.)));
    const double sq = 2. * sqrt(-Q);
    const double l1 = sq * cos(theta / 3.) + (2. / 3.) * uroe;
    const double l2 = sq * cos((theta + 2. * PI) / 3.) + (2. / 3.) * uroe;
    const double l3 = sq * cos((theta + 4. * PI) / 3.) + (2. / 3.) * uroe;
    const double l4 = uroe, l5 = uroe;
    const double a1 = l1 - uroe, a2 = l2 - uroe, a3 = l3 - uroe;
    const double beta = 2. * As * xi * vroe / hroe;
    const double ss = sound * sound, uu = uroe * uroe;
    const double g1 = ss - uu + l2 * l3 - beta * a2 * a3 * vroe;
    const double g2 = ss - uu + l1 * l3 - beta * a1 * a3 * vroe;
    const double g3 = ss - uu + l1 * l2 - beta * a1 * a2 * vroe;
    const double sg1 = -a1 * a2 + a2 * a3 - a1 * a3 + a1 * a1;
    const double sg2 = a1 * a2 + a2 * a3 - a1 * a3 - a2 * a2;
    const double sg3 = a1 * a2 - a2 * a3 - a1 * a3 + a3 * a3;
    const double eps = 1e-10;
    const double s1 = (fabs(l1) < eps) ? 0. : l1 / fabs(l1);
    const double s2 = (fabs(l2) < eps) ? 0. : l2 / fabs(l2);
    const double s3 = (fabs(l3) < eps) ? 0. : l3 / fabs(l3);
    const double s4 = (fabs(l4) < eps) ? 0. : l4 / fabs(l4);
    const double s5 = (fabs(l5) < eps) ? 0. : l5 / fabs(l5);
    const double G1 = g1 / sg1, G2 = g2 / sg2, G3 = g3 / sg3;
    const double A1 = a1 * a1 / ss - 1., A2 = a2 * a2 / ss - 1., A3 = a3 * a3 / ss - 1.;
    const double v = vroe, c = croe, u = uroe;
    double r[5][5];
    r[0][0] = s1 * G1 - s2 * G2 + s3 * G3 + s5 * (beta * v);
    r[1][0] = l1 * s1 * G1 - l2 * s2 * G2 + l3 * s3 * G3 + s5 * (beta * u * v);
    r[2][0] = v * s1 * G1 - v * s2 * G2 + v * s3 * G3 - v * s5 * (1. - beta * v);
    r[3][0] = c * s1 * G1 - c * s2 * G2 + c * s3 * G3 - c * s4 + c * s5 * beta * v;
    r[4][0] = A1 * s1 * G1 - A2 * s2 * G2 + A3 * s3 * G3 - s5 * (beta * v);
    r[0][1] = -s1 * (a2 + a3) / sg1 + s2 * (a1 + a3) / sg2 - s3 * (a1 + a2) / sg3;
    r[1][1] = -l1 * s1 * (a2 + a3) / sg1 + l2 * s2 * (a1 + a3) / sg2 - l3 * s3 * (a1 + a2) / sg3;
    r[2][1] = -v * s1 * (a2 + a3) / sg1 + v * s2 * (a1 + a3) / sg2 - v * s3 * (a1 + a2) / sg3;
    r[3][1] = -c * s1 * (a2 + a3) / sg1 + c * s2 * (a1 + a3) / sg2 - c * s3 * (a1 + a2) / sg3;
    r[4][1] = -A1 * s1 * (a2 + a3) / sg1 + A2 * s2 * (a1 + a3) / sg2 - A3 * s3 * (a1 + a2) / sg3;
    r[0][2] = s1 * beta * a2 * a3 / sg1 - s2 * beta * a1 * a3 / sg2 + s3 * beta * a1 * a2 / sg3 - s5 * beta;
    r[1][2] = l1 * s1 * beta * a2 * a3 / sg1 - l2 * s2 * beta * a1 * a3 / sg2 + l3 * s3 * beta * a1 * a2 / sg3 -
              s5 * beta * u;
    r[2][2] = v * s1 * beta * a2 * a3 / sg1 - v * s2 * beta * a1 * a3 / sg2 + v * s3 * beta * a1 * a2 / sg3 +
              s5 * (1. - beta * v);
    r[3][2] = c * s1 * beta * a2 * a3 / sg1 - c * s2 * beta * a1 * a3 / sg2 + c * s3 * beta * a1 * a2 / sg3 -
              c * s5 * beta;
    r[4][2] = A1 * s1 * beta * a2 * a3 / sg1 - A2 * s2 * beta * a1 * a3 / sg2 + A3 * s3 * beta * a1 * a2 / sg3 +
              s5 * beta;
    r[0][3] = 0.;
    r[1][3] = 0.;
    r[2][3] = 0.;
    r[3][3] = s4;
    r[4][3] = 0.;
    r[0][4] = s1 * ss / sg1 - s2 * ss / sg2 + s3 * ss / sg3;
    r[1][4] = l1 * s1 * ss / sg1 - l2 * s2 * ss / sg2 + l3 * s3 * ss / sg3;
    r[2][4] = v * s1 * ss / sg1 - v * s2 * ss / sg2 + v * s3 * ss / sg3;
    r[3][4] = c * s1 * ss / sg1 - c * s2 * ss / sg2 + c * s3 * ss / sg3;
    r[4][4] = A1 * s1 * ss / sg1 - A2 * s2 * ss / sg2 + A3 * s3 * ss / sg3;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        double s = 0.;
#pragma unroll
        for (int j = 0; j < 5; j++) s += r[k][j] * w_dif[j];
        out[k] = s;
    }
}

// srnh_scheme (shallowater/fvm_utils.py:152-342): flux in the face orientation, |n| = mes.
__device__ __forceinline__ void srnh_flux(double hu_l, double hu_r, double hv_l, double hv_r, double h_l, double h_r,
                                          double hc_l, double hc_r, double Z_l, double Z_r, double n0, double n1,
                                          double mes, double grav, double *flux)
{
    const double As = 0.;
    const double xi = 1. / (1. - 0.4);
    const double ninv0 = -1. * n1, ninv1 = n0;
    const double sl = sqrt(h_l), sr = sqrt(h_r);
    const double u_h = (hu_l / h_l * sl + hu_r / h_r * sr) / (sl + sr);
    const double v_h = (hv_l / h_l * sl + hv_r / h_r * sr) / (sl + sr);
    const double c_h = (hc_l / h_l * sl + hc_r / h_r * sr) / (sl + sr);
    const double un_h = (u_h * n0 + v_h * n1) / mes;
    const double vn_h = (u_h * ninv0 + v_h * ninv1) / mes;
    const double hroe = (h_l + h_r) / 2.;
    const double uleft = (hu_l * n0 + hv_l * n1) / mes;
    const double vleft = (hu_l * ninv0 + hv_l * ninv1) / mes;
    const double uright = (hu_r * n0 + hv_r * n1) / mes;
    const double vright = (hu_r * ninv0 + hv_r * ninv1) / mes;
    double w_lr[5], w_dif[5], corr[5];
    w_lr[0] = (h_l + h_r) / 2.;
    w_lr[1] = (uleft + uright) / 2.;
    w_lr[2] = (vleft + vright) / 2.;
    w_lr[3] = (hc_l + hc_r) / 2.;
    w_lr[4] = (Z_l + Z_r) / 2.;
    w_dif[0] = h_r - h_l;
    w_dif[1] = uright - uleft;
    w_dif[2] = vright - vleft;
    w_dif[3] = hc_r - hc_l;
    w_dif[4] = Z_r - Z_l;
    srnh_sign_apply(hroe, un_h, vn_h, c_h, grav, w_dif, corr);
#pragma unroll
    for (int k = 0; k < 5; k++) w_lr[k] = w_lr[k] - corr[k] / 2.;
    const double unew = (w_lr[1] * n0 + w_lr[2] * -1. * n1) / mes;
    const double vnew = (w_lr[1] * -1. * ninv0 + w_lr[2] * ninv1) / mes;
    const double q_s = n0 * unew + n1 * vnew;
    const double w_h = w_lr[0];
    flux[0] = q_s;
    flux[1] = q_s * unew / w_h + 0.5 * grav * w_h * w_h * n0;
    flux[2] = q_s * vnew / w_h + 0.5 * grav * w_h * w_h * n1;
    flux[3] = q_s * w_lr[3] / w_h;
    const double k2 = unew * unew + vnew * vnew, w3 = pow(w_h, 3.);
    flux[4] = As * xi * n0 * unew * k2 / w3 + As * xi * n1 * vnew * k2 / w3;
}

// explicitscheme_convective_SW (shallowater/fvm_utils.py:344-496), one cell per thread.
__global__ void __launch_bounds__(kBlock) k_srnh_convective(Mesh m, Out5 rez, Fields5 cell, Fields5 ghost,
                                                            const double *__restrict__ h_x,
                                                            const double *__restrict__ h_y,
                                                            const double *__restrict__ hc_x,
                                                            const double *__restrict__ hc_y,
                                                            const double *__restrict__ psi, double om, double grav)
{
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const double *h = cell.f[0], *hu = cell.f[1], *hv = cell.f[2], *hc = cell.f[3], *Z = cell.f[4];
    const int nf = m.cell_nfaces[i];
    double acc[5] = {0., 0., 0., 0., 0.};
    for (int r = 0; r < nf; r++) {
        const int s = m.order_split[(size_t)r * m.C + i];
        const int f = m.cell_faces[(size_t)s * m.C + i];
        const int L = m.face_cells[2 * f], R = m.face_cells[2 * f + 1];
        const double *cf = m.face_center + 3 * (size_t)f;
        const double *cl = m.cell_center + 3 * (size_t)L;
        const double *nrm = m.face_normal + 3 * (size_t)f;
        const double mes = m.face_mesure[f];
        const double rl0 = cf[0] - cl[0], rl1 = cf[1] - cl[1];
        const double h_l = h[L] + om * psi[L] * (h_x[L] * rl0 + h_y[L] * rl1);
        const double hc_l = hc[L] + om * psi[L] * (hc_x[L] * rl0 + hc_y[L] * rl1);
        double fl[5];
        if (R >= 0) {
            const double *cr = m.cell_center + 3 * (size_t)R;
            const double rr0 = cf[0] - cr[0], rr1 = cf[1] - cr[1];
            const double h_r = h[R] + om * psi[R] * (h_x[R] * rr0 + h_y[R] * rr1);
            const double hc_r = hc[R] + om * psi[R] * (hc_x[R] * rr0 + hc_y[R] * rr1);
            srnh_flux(hu[L], hu[R], hv[L], hv[R], h_l, h_r, hc_l, hc_r, Z[L], Z[R], nrm[0], nrm[1], mes, grav, fl);
        } else {
            srnh_flux(hu[L], ghost.f[1][f], hv[L], ghost.f[2][f], h_l, ghost.f[0][f], hc_l, ghost.f[3][f], Z[L],
                      ghost.f[4][f], nrm[0], nrm[1], mes, grav, fl);
        }
        if (L == i) {
#pragma unroll
            for (int k = 0; k < 5; k++) acc[k] -= fl[k];
        } else {
#pragma unroll
            for (int k = 0; k < 5; k++) acc[k] += fl[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 5; k++) rez.f[k][i] = acc[k];
}

// term_source_srnh_SW (shallowater/fvm_utils.py:44-148) for triangles. With quirk the
// neighbour of face slot j is cellfid[i][j] of the compacted neighbour list, negative
// entries wrapping to the end as Python indexing does; without, the cell across face j.
__global__ void __launch_bounds__(kBlock) k_srnh_source(Mesh m, Out5 src, const double *__restrict__ h,
                                                        const double *__restrict__ Z,
                                                        const double *__restrict__ h_ghost,
                                                        const double *__restrict__ Z_ghost,
                                                        const int32_t *__restrict__ nodeid, int nw,
                                                        const int32_t *__restrict__ cellfid, int fw,
                                                        const double *__restrict__ nf,
                                                        const double *__restrict__ vertex, double grav, int quirk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const double *G = m.cell_center + 3 * (size_t)i;
    double ss[3][2], zv[3], mata[3], matb[3];
    double c_1 = 0., c_2 = 0.;
    const double h_1p = h[i], z_1p = Z[i];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int f = m.cell_faces[(size_t)j * m.C + i];
        ss[j][0] = nf[(size_t)i * 9 + j * 3 + 0];
        ss[j][1] = nf[(size_t)i * 9 + j * 3 + 1];
        double h_p1, z_p1;
        if (m.face_name[f] == 0) {
            int nb;
            if (quirk) {
                nb = cellfid[(size_t)i * fw + j];
                if (nb < 0) nb += m.C;
            } else {
                const int L = m.face_cells[2 * f];
                nb = (L == i) ? m.face_cells[2 * f + 1] : L;
            }
            h_p1 = h[nb];
            z_p1 = Z[nb];
        } else {
            h_p1 = h_ghost[f];
            z_p1 = Z_ghost[f];
        }
        zv[j] = z_p1;
        mata[j] = h_p1 * ss[j][0];
        matb[j] = h_p1 * ss[j][1];
        c_1 = c_1 + (0.5 * (h_1p + h_p1) * 0.5 * (h_1p + h_p1)) * ss[j][0];
        c_2 = c_2 + (0.5 * (h_1p + h_p1) * 0.5 * (h_1p + h_p1)) * ss[j][1];
    }
    const double c_3 = 3.0 * h_1p;
    const double delta = (mata[1] * matb[2] - mata[2] * matb[1]) - (mata[0] * matb[2] - matb[0] * mata[2]) +
                         (mata[0] * matb[1] - matb[0] * mata[1]);
    const double deltax = c_3 * (mata[1] * matb[2] - mata[2] * matb[1]) - (c_1 * matb[2] - c_2 * mata[2]) +
                          (c_1 * matb[1] - c_2 * mata[1]);
    const double deltay = (c_1 * matb[2] - c_2 * mata[2]) - c_3 * (mata[0] * matb[2] - matb[0] * mata[2]) +
                          (mata[0] * c_2 - matb[0] * c_1);
    const double deltaz = (mata[1] * c_2 - matb[1] * c_1) - (mata[0] * c_2 - matb[0] * c_1) +
                          c_3 * (mata[0] * matb[1] - matb[0] * mata[1]);
    const double h_1 = deltax / delta, h_2 = deltay / delta, h_3 = deltaz / delta;
    const double z_1 = z_1p + h_1p - h_1;
    const double z_2 = z_1p + h_1p - h_2;
    const double z_3 = z_1p + h_1p - h_3;
    const double *b = vertex + 4 * (size_t)nodeid[(size_t)i * nw + 1];
    double ns[3][2];
    ns[0][0] = G[1] - b[1];
    ns[0][1] = -(G[0] - b[0]);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        ns[1][k] = ns[0][k] - ss[1][k];
        ns[2][k] = ns[0][k] + ss[0][k];
    }
    double s_1[2], s_2[2], s_3[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        s_1[k] = 0.5 * h_1 * (zv[0] * ss[0][k] + z_2 * ns[0][k] + z_3 * -1. * ns[2][k]);
        s_2[k] = 0.5 * h_2 * (zv[1] * ss[1][k] + z_1 * -1. * ns[0][k] + z_3 * ns[1][k]);
        s_3[k] = 0.5 * h_3 * (zv[2] * ss[2][k] + z_1 * ns[2][k] + z_2 * -1. * ns[1][k]);
    }
    src.f[0][i] = 0.;
    src.f[1][i] = -grav * (s_1[0] + s_2[0] + s_3[0]);
    src.f[2][i] = -grav * (s_1[1] + s_2[1] + s_3[1]);
    src.f[3][i] = 0.;
    src.f[4][i] = 0.;
}

// update_SW (shallowater/fvm_utils.py:3-14), all five fields.
__global__ void __launch_bounds__(kBlock) k_update_sw5(int C, Out5 st, Fields5 rez, Fields5 src,
                                                       const double *__restrict__ dissip_hc,
                                                       const double *__restrict__ corio_hu,
                                                       const double *__restrict__ corio_hv, double wind_hu,
                                                       double wind_hv, double dtime,
                                                       const double *__restrict__ dt_dev,
                                                       const double *__restrict__ vol)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double dt = dt_dev ? dt_dev[0] : dtime;
    const double v = vol[i];
    st.f[0][i] += dt * (rez.f[0][i] + src.f[0][i]) / v;
    st.f[1][i] += dt * ((rez.f[1][i] + src.f[1][i]) / v + corio_hu[i] + wind_hu);
    st.f[2][i] += dt * ((rez.f[2][i] + src.f[2][i]) / v + corio_hv[i] + wind_hv);
    st.f[3][i] += dt * (rez.f[3][i] + src.f[3][i] - dissip_hc[i]) / v;
    st.f[4][i] += dt * (rez.f[4][i] + src.f[4][i]) / v;
}

}  // namespace mfv

using namespace mfv;

extern "C" int mfv_explicitscheme_convective_sw(const mfv_mesh *m, double *rez_h, double *rez_hu, double *rez_hv,
                                                double *rez_hc, double *rez_Z, const double *h_c,
                                                const double *hu_c, const double *hv_c, const double *hc_c,
                                                const double *Z_c, const double *h_ghost, const double *hu_ghost,
                                                const double *hv_ghost, const double *hc_ghost,
                                                const double *Z_ghost, const double *h_x, const double *h_y,
                                                const double *hc_x, const double *hc_y, const double *psi,
                                                double grav, int order, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(rez_h && rez_hu && rez_hv && rez_hc && rez_Z && h_c && hu_c && hv_c && hc_c && Z_c && h_ghost &&
                    hu_ghost && hv_ghost && hc_ghost && Z_ghost && h_x && h_y && hc_x && hc_y && psi,
                "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    Out5 rez{{rez_h, rez_hu, rez_hv, rez_hc, rez_Z}};
    Fields5 cell{{h_c, hu_c, hv_c, hc_c, Z_c}};
    Fields5 ghost{{h_ghost, hu_ghost, hv_ghost, hc_ghost, Z_ghost}};
    hipLaunchKernelGGL(k_srnh_convective, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), rez, cell, ghost, h_x, h_y, hc_x, hc_y, psi, (double)(order - 1), grav);
    return check_launch("k_srnh_convective");
}

extern "C" int mfv_term_source_srnh_sw(const mfv_mesh *m, double *src_h, double *src_hu, double *src_hv,
                                       double *src_hc, double *src_Z, const double *h_c, const double *Z_c,
                                       const double *h_ghost, const double *Z_ghost, const int32_t *nodeid, int nw,
                                       const int32_t *cellfid, int fw, const double *nf, const double *vertex,
                                       double grav, int quirk, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(m->dim == 2 && m->max_cell_faces == 3,
                "term_source_srnh_SW is defined for triangles only (2D, 3 faces per cell)");
    MFV_REQUIRE(nw >= 3 && fw >= 3, "nodeid / cellfid widths must be >= 3 (got %d, %d)", nw, fw);
    MFV_REQUIRE(src_h && src_hu && src_hv && src_hc && src_Z && h_c && Z_c && h_ghost && Z_ghost && nodeid &&
                    cellfid && nf && vertex,
                "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    Out5 src{{src_h, src_hu, src_hv, src_hc, src_Z}};
    hipLaunchKernelGGL(k_srnh_source, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m),
                       src, h_c, Z_c, h_ghost, Z_ghost, nodeid, nw, cellfid, fw, nf, vertex, grav, quirk);
    return check_launch("k_srnh_source");
}

extern "C" int mfv_update_sw5(const mfv_mesh *m, double *h_c, double *hu_c, double *hv_c, double *hc_c, double *Z_c,
                              const double *rez_h, const double *rez_hu, const double *rez_hv, const double *rez_hc,
                              const double *rez_Z, const double *src_h, const double *src_hu, const double *src_hv,
                              const double *src_hc, const double *src_Z, const double *dissip_hc,
                              const double *corio_hu, const double *corio_hv, double wind_hu, double wind_hv,
                              double dtime, const double *dt_dev, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(h_c && hu_c && hv_c && hc_c && Z_c && rez_h && rez_hu && rez_hv && rez_hc && rez_Z && src_h &&
                    src_hu && src_hv && src_hc && src_Z && dissip_hc && corio_hu && corio_hv,
                "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    Out5 st{{h_c, hu_c, hv_c, hc_c, Z_c}};
    Fields5 rez{{rez_h, rez_hu, rez_hv, rez_hc, rez_Z}};
    Fields5 src{{src_h, src_hu, src_hv, src_hc, src_Z}};
    hipLaunchKernelGGL(k_update_sw5, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), m->nb_cells,
                       st, rez, src, dissip_hc, corio_hu, corio_hv, wind_hu, wind_hv, dtime, dt_dev,
                       m->cell_volume);
    return check_launch("k_update_sw5");
}

// ---------------------------------------------------------------------------------------------
// term_friction_SW, term_coriolis_SW, term_wind_SW (shallowater/fvm_utils.py:498-544)

__global__ void __launch_bounds__(kBlock) k_term_friction_sw(int C, const double *__restrict__ h,
                                                              double *__restrict__ hu, double *__restrict__ hv,
                                                              double grav, double eta, double time)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double ufric = hu[i] / h[i];
    const double vfric = hv[i] / h[i];
    const double hfric = h[i];
    const double A = 1 + time * grav * (eta * eta) * (sqrt(ufric * ufric + vfric * vfric) / pow(hfric, 4. / 3.));
    const double hutild = hu[i] / A;
    const double hvtild = hv[i] / A;
    hu[i] = hutild;
    hv[i] = hvtild;
}

__global__ void __launch_bounds__(kBlock) k_term_coriolis_sw(int C, const double *__restrict__ hu,
                                                              const double *__restrict__ hv,
                                                              double *__restrict__ corio_hu,
                                                              double *__restrict__ corio_hv,
                                                              const double *__restrict__ f_c, double f_const)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double f = f_c ? f_c[i] : f_const;
    corio_hu[i] = f * hu[i];
    corio_hv[i] = -f * hv[i];
}

__global__ void __launch_bounds__(kBlock) k_term_wind_sw(int C, double *__restrict__ Tx, double *__restrict__ Ty,
                                                          const double *__restrict__ wx,
                                                          const double *__restrict__ wy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double ro_a = 1.25;
    const double u = wx[i], v = wy[i];
    const double nor = sqrt(u * u + v * v);
    const double Cw = ro_a * (0.75 + 0.067 * nor) * 1e-03;
    Tx[i] = Cw * u * nor;
    Ty[i] = Cw * v * nor;
}

extern "C" int mfv_term_friction_sw(int C, const double *h, double *hu, double *hv, double grav, double eta,
                                    double time, void *stream)
{
    MFV_REQUIRE(C >= 0, "negative cell count");
    if (C == 0) return MFV_OK;
    MFV_REQUIRE(h && hu && hv, "NULL array");
    hipLaunchKernelGGL(k_term_friction_sw, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), C, h, hu, hv, grav,
                       eta, time);
    return check_launch("k_term_friction_sw");
}

extern "C" int mfv_term_coriolis_sw(int C, const double *hu, const double *hv, double *corio_hu, double *corio_hv,
                                    const double *f_c, double f_const, void *stream)
{
    MFV_REQUIRE(C >= 0, "negative cell count");
    if (C == 0) return MFV_OK;
    MFV_REQUIRE(hu && hv && corio_hu && corio_hv, "NULL array");
    hipLaunchKernelGGL(k_term_coriolis_sw, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), C, hu, hv,
                       corio_hu, corio_hv, f_c, f_const);
    return check_launch("k_term_coriolis_sw");
}

extern "C" int mfv_term_wind_sw(int C, double *Tx, double *Ty, const double *wind_wx, const double *wind_wy,
                                void *stream)
{
    MFV_REQUIRE(C >= 0, "negative cell count");
    if (C == 0) return MFV_OK;
    MFV_REQUIRE(Tx && Ty && wind_wx && wind_wy, "NULL array");
    hipLaunchKernelGGL(k_term_wind_sw, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), C, Tx, Ty, wind_wx,
                       wind_wy);
    return check_launch("k_term_wind_sw");
}
