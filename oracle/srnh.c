/*
 * ORACLE -- test infrastructure only. Never linked into the product library.
 *
 * Strict-IEEE C restatement (built with -O2 -ffp-contract=off -fno-fast-math) of the
 * reference's shallow-water path with its own numerical flux:
 *   srnh_scheme                  solvers/shallowater/fvm_utils.py:152-342
 *   explicitscheme_convective_SW solvers/shallowater/fvm_utils.py:344-496
 *   term_source_srnh_SW          solvers/shallowater/fvm_utils.py:44-148
 *   update_SW                    solvers/shallowater/fvm_utils.py:3-14
 * Five conserved fields (h, hu, hv, hc, Z), single subdomain (the halo loops of the
 * reference are empty there). Expressions keep the reference's operand order; Python's
 * x**2 is x*x and x**3 is pow(x, 3). The sign matrix is built by orc_srnh_roe so that the
 * tests can check it on its own (against sign(A) of the rotated Jacobian).
 */
#include <math.h>
#include <stdint.h>

#define ROW(a, w, i) ((a) + (int64_t)(i) * (w))

/* Roe state -> eigenvalues lam[5] and the 5x5 sign matrix r[5*row+col]
 * (fvm_utils.py:220-315, with As = 0). */
void orc_srnh_roe(double hroe, double uroe, double vroe, double croe, double grav, double *lam, double *r)
{
    const double As = 0.;
    const double p = 0.4;
    const double xi = 1. / (1. - p);
    const double PI = 3.141592653589793;
    double d = As * xi * (3. * (uroe * uroe) + vroe * vroe);
    double sound = sqrt(grav * hroe);
    double Q = -((uroe * uroe) + 3. * grav * (hroe + d)) / 9.;
    double R = uroe * (9. * grav * (2. * hroe - d) - 2. * (uroe * uroe)) / 54.;
    double theta = acos(R / sqrt(-pow(Q, 3.)));
    double sq = 2. * sqrt(-Q);
    double l1 = sq * cos(theta / 3.) + (2. / 3.) * uroe;
    double l2 = sq * cos((theta + 2. * PI) / 3.) + (2. / 3.) * uroe;
    double l3 = sq * cos((theta + 4. * PI) / 3.) + (2. / 3.) * uroe;
    double l4 = uroe, l5 = uroe;
    double a1 = l1 - uroe, a2 = l2 - uroe, a3 = l3 - uroe;
    double beta = 2. * As * xi * vroe / hroe;
    double ss = sound * sound, uu = uroe * uroe;
    double g1 = ss - uu + l2 * l3 - beta * a2 * a3 * vroe;
    double g2 = ss - uu + l1 * l3 - beta * a1 * a3 * vroe;
    double g3 = ss - uu + l1 * l2 - beta * a1 * a2 * vroe;
    double sg1 = -a1 * a2 + a2 * a3 - a1 * a3 + a1 * a1;
    double sg2 = a1 * a2 + a2 * a3 - a1 * a3 - a2 * a2;
    double sg3 = a1 * a2 - a2 * a3 - a1 * a3 + a3 * a3;
    const double eps = 1e-10;
    double s1 = (fabs(l1) < eps) ? 0. : l1 / fabs(l1);
    double s2 = (fabs(l2) < eps) ? 0. : l2 / fabs(l2);
    double s3 = (fabs(l3) < eps) ? 0. : l3 / fabs(l3);
    double s4 = (fabs(l4) < eps) ? 0. : l4 / fabs(l4);
    double s5 = (fabs(l5) < eps) ? 0. : l5 / fabs(l5);
    double G1 = g1 / sg1, G2 = g2 / sg2, G3 = g3 / sg3;
    /* (alpha_k^2 / sound^2 - 1) */
    double A1 = a1 * a1 / ss - 1., A2 = a2 * a2 / ss - 1., A3 = a3 * a3 / ss - 1.;
    double v = vroe, c = croe, u = uroe;
    lam[0] = l1; lam[1] = l2; lam[2] = l3; lam[3] = l4; lam[4] = l5;
    /* column 0 */
    r[0 * 5 + 0] = s1 * G1 - s2 * G2 + s3 * G3 + s5 * (beta * v);
    r[1 * 5 + 0] = l1 * s1 * G1 - l2 * s2 * G2 + l3 * s3 * G3 + s5 * (beta * u * v);
    r[2 * 5 + 0] = v * s1 * G1 - v * s2 * G2 + v * s3 * G3 - v * s5 * (1. - beta * v);
    r[3 * 5 + 0] = c * s1 * G1 - c * s2 * G2 + c * s3 * G3 - c * s4 + c * s5 * beta * v;
    r[4 * 5 + 0] = A1 * s1 * G1 - A2 * s2 * G2 + A3 * s3 * G3 - s5 * (beta * v);
    /* column 1 */
    r[0 * 5 + 1] = -s1 * (a2 + a3) / sg1 + s2 * (a1 + a3) / sg2 - s3 * (a1 + a2) / sg3;
    r[1 * 5 + 1] = -l1 * s1 * (a2 + a3) / sg1 + l2 * s2 * (a1 + a3) / sg2 - l3 * s3 * (a1 + a2) / sg3;
    r[2 * 5 + 1] = -v * s1 * (a2 + a3) / sg1 + v * s2 * (a1 + a3) / sg2 - v * s3 * (a1 + a2) / sg3;
    r[3 * 5 + 1] = -c * s1 * (a2 + a3) / sg1 + c * s2 * (a1 + a3) / sg2 - c * s3 * (a1 + a2) / sg3;
    r[4 * 5 + 1] = -A1 * s1 * (a2 + a3) / sg1 + A2 * s2 * (a1 + a3) / sg2 - A3 * s3 * (a1 + a2) / sg3;
    /* column 2 */
    r[0 * 5 + 2] = s1 * beta * a2 * a3 / sg1 - s2 * beta * a1 * a3 / sg2 + s3 * beta * a1 * a2 / sg3 - s5 * beta;
    r[1 * 5 + 2] = l1 * s1 * beta * a2 * a3 / sg1 - l2 * s2 * beta * a1 * a3 / sg2 +
                   l3 * s3 * beta * a1 * a2 / sg3 - s5 * beta * u;
    r[2 * 5 + 2] = v * s1 * beta * a2 * a3 / sg1 - v * s2 * beta * a1 * a3 / sg2 + v * s3 * beta * a1 * a2 / sg3 +
                   s5 * (1. - beta * v);
    r[3 * 5 + 2] = c * s1 * beta * a2 * a3 / sg1 - c * s2 * beta * a1 * a3 / sg2 + c * s3 * beta * a1 * a2 / sg3 -
                   c * s5 * beta;
    r[4 * 5 + 2] = A1 * s1 * beta * a2 * a3 / sg1 - A2 * s2 * beta * a1 * a3 / sg2 +
                   A3 * s3 * beta * a1 * a2 / sg3 + s5 * beta;
    /* column 3 */
    r[0 * 5 + 3] = 0.; r[1 * 5 + 3] = 0.; r[2 * 5 + 3] = 0.; r[3 * 5 + 3] = s4; r[4 * 5 + 3] = 0.;
    /* column 4 */
    r[0 * 5 + 4] = s1 * ss / sg1 - s2 * ss / sg2 + s3 * ss / sg3;
    r[1 * 5 + 4] = l1 * s1 * ss / sg1 - l2 * s2 * ss / sg2 + l3 * s3 * ss / sg3;
    r[2 * 5 + 4] = v * s1 * ss / sg1 - v * s2 * ss / sg2 + v * s3 * ss / sg3;
    r[3 * 5 + 4] = c * s1 * ss / sg1 - c * s2 * ss / sg2 + c * s3 * ss / sg3;
    r[4 * 5 + 4] = A1 * s1 * ss / sg1 - A2 * s2 * ss / sg2 + A3 * s3 * ss / sg3;
}

/* srnh_scheme (fvm_utils.py:152-342): flux[5] in the face orientation, |n| = mesure. */
void orc_srnh_flux(double hu_l, double hu_r, double hv_l, double hv_r, double h_l, double h_r, double hc_l,
                   double hc_r, double Z_l, double Z_r, const double *normal, double mesure, double grav,
                   double *flux)
{
    const double As = 0.;
    const double xi = 1. / (1. - 0.4);
    double ninv0 = -1. * normal[1], ninv1 = normal[0];
    double sl = sqrt(h_l), sr = sqrt(h_r);
    double u_h = (hu_l / h_l * sl + hu_r / h_r * sr) / (sl + sr);
    double v_h = (hv_l / h_l * sl + hv_r / h_r * sr) / (sl + sr);
    double c_h = (hc_l / h_l * sl + hc_r / h_r * sr) / (sl + sr);
    double un_h = (u_h * normal[0] + v_h * normal[1]) / mesure;
    double vn_h = (u_h * ninv0 + v_h * ninv1) / mesure;
    double hroe = (h_l + h_r) / 2.;
    double uleft = (hu_l * normal[0] + hv_l * normal[1]) / mesure;
    double vleft = (hu_l * ninv0 + hv_l * ninv1) / mesure;
    double uright = (hu_r * normal[0] + hv_r * normal[1]) / mesure;
    double vright = (hu_r * ninv0 + hv_r * ninv1) / mesure;
    double w_lr[5], w_dif[5], lam[5], r[25];
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
    orc_srnh_roe(hroe, un_h, vn_h, c_h, grav, lam, r);
    for (int k = 0; k < 5; k++) {
        double s = 0.;
        for (int j = 0; j < 5; j++) s += r[k * 5 + j] * w_dif[j];
        w_lr[k] = w_lr[k] - s / 2.;
    }
    double unew = (w_lr[1] * normal[0] + w_lr[2] * -1. * normal[1]) / mesure;
    double vnew = (w_lr[1] * -1. * ninv0 + w_lr[2] * ninv1) / mesure;
    double q_s = normal[0] * unew + normal[1] * vnew;
    double w_h = w_lr[0];
    flux[0] = q_s;
    flux[1] = q_s * unew / w_h + 0.5 * grav * w_h * w_h * normal[0];
    flux[2] = q_s * vnew / w_h + 0.5 * grav * w_h * w_h * normal[1];
    flux[3] = q_s * w_lr[3] / w_h;
    double k2 = unew * unew + vnew * vnew, w3 = pow(w_h, 3.);
    flux[4] = As * xi * normal[0] * unew * k2 / w3 + As * xi * normal[1] * vnew * k2 / w3;
}

/* explicitscheme_convective_SW (fvm_utils.py:344-496) without halo faces. rez[5*C]
 * (h, hu, hv, hc, Z blocks of C), state[5*C] likewise, ghost[5*F]. MUSCL on h and hc
 * only, both limited by psi (= hc.psi); boundary faces take the raw ghost state. */
void orc_convective_srnh(double *rez, const double *state, const double *ghost, const double *h_x,
                         const double *h_y, const double *hc_x, const double *hc_y, const double *psi,
                         const double *centerc, const double *centerf, const int32_t *cellidf,
                         const double *mesuref, const double *normalf, const int32_t *innerfaces, int n_inner,
                         const int32_t *boundaryfaces, int n_bnd, double grav, int order, int C, int F)
{
    const double *h = state, *hu = state + C, *hv = state + 2 * (int64_t)C, *hc = state + 3 * (int64_t)C,
                 *Z = state + 4 * (int64_t)C;
    for (int64_t k = 0; k < 5 * (int64_t)C; k++) rez[k] = 0.;
    double flux[5];
    for (int t = 0; t < n_inner; t++) {
        int f = innerfaces[t];
        int L = cellidf[2 * f], R = cellidf[2 * f + 1];
        const double *cl = centerc + 3 * (int64_t)L, *cr = centerc + 3 * (int64_t)R, *cf = centerf + 3 * (int64_t)f;
        double rl0 = cf[0] - cl[0], rr0 = cf[0] - cr[0];
        double rl1 = cf[1] - cl[1], rr1 = cf[1] - cr[1];
        double h_l = h[L] + (order - 1) * psi[L] * (h_x[L] * rl0 + h_y[L] * rl1);
        double h_r = h[R] + (order - 1) * psi[R] * (h_x[R] * rr0 + h_y[R] * rr1);
        double hc_l = hc[L] + (order - 1) * psi[L] * (hc_x[L] * rl0 + hc_y[L] * rl1);
        double hc_r = hc[R] + (order - 1) * psi[R] * (hc_x[R] * rr0 + hc_y[R] * rr1);
        orc_srnh_flux(hu[L], hu[R], hv[L], hv[R], h_l, h_r, hc_l, hc_r, Z[L], Z[R], normalf + 3 * (int64_t)f,
                      mesuref[f], grav, flux);
        for (int k = 0; k < 5; k++) {
            rez[k * (int64_t)C + L] -= flux[k];
            rez[k * (int64_t)C + R] += flux[k];
        }
    }
    for (int t = 0; t < n_bnd; t++) {
        int f = boundaryfaces[t];
        int L = cellidf[2 * f];
        const double *cl = centerc + 3 * (int64_t)L, *cf = centerf + 3 * (int64_t)f;
        double rl0 = cf[0] - cl[0], rl1 = cf[1] - cl[1];
        double h_l = h[L] + (order - 1) * psi[L] * (h_x[L] * rl0 + h_y[L] * rl1);
        double hc_l = hc[L] + (order - 1) * psi[L] * (hc_x[L] * rl0 + hc_y[L] * rl1);
        orc_srnh_flux(hu[L], ghost[1 * (int64_t)F + f], hv[L], ghost[2 * (int64_t)F + f], h_l, ghost[f], hc_l,
                      ghost[3 * (int64_t)F + f], Z[L], ghost[4 * (int64_t)F + f], normalf + 3 * (int64_t)f,
                      mesuref[f], grav, flux);
        for (int k = 0; k < 5; k++) rez[k * (int64_t)C + L] -= flux[k];
    }
}

/* term_source_srnh_SW (fvm_utils.py:44-148), triangles. src_hu, src_hv [C]; the other
 * three source fields are zero. nodeid [C][nw], faceid/cellfid [C][fw] (padded, count
 * last), nf [C][3][3] (cells.nf), vertex [N][4], names [F].
 *
 * quirk != 0 reproduces the reference exactly: the neighbour across face slot j is read
 * as cellfid[i][j], but cells.cellfid lists the neighbours compacted (boundary faces
 * skipped, -1 padding; ddm_utils2d.py:32-53), so in a cell with a boundary face the
 * slots after it read the wrong cell, and a -1 entry wraps to the last cell as Python and
 * Numba indexing do. quirk == 0 takes the cell on the other side of face j
 * (faces.cellid), which is what the scheme intends. */
void orc_source_srnh(double *src_hu, double *src_hv, const double *h, const double *Z, const double *h_ghost,
                     const double *Z_ghost, const int32_t *nodeid, int nw, const int32_t *faceid,
                     const int32_t *cellfid, int fw, const int32_t *cellidf, const double *centerc,
                     const double *nf, const int32_t *name, const double *vertex, double grav, int quirk, int C)
{
    for (int i = 0; i < C; i++) {
        const double *G = centerc + 3 * (int64_t)i;
        double ss[3][3], ns[3][3], zv[3], mata[3], matb[3], hi_p[3], zi_p[3];
        double c_1 = 0., c_2 = 0., h_1p = 0.;
        for (int j = 0; j < 3; j++) {
            int f = ROW(faceid, fw, i)[j];
            for (int k = 0; k < 3; k++) ss[j][k] = nf[(int64_t)i * 9 + j * 3 + k];
            double z_1p, h_p1, z_p1;
            h_1p = h[i];
            z_1p = Z[i];
            if (name[f] == 0) {
                int nb;
                if (quirk) {
                    nb = ROW(cellfid, fw, i)[j];
                    if (nb < 0) nb += C;
                } else {
                    nb = (cellidf[2 * f] == i) ? cellidf[2 * f + 1] : cellidf[2 * f];
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
            hi_p[j] = h_1p;
            zi_p[j] = z_1p;
        }
        double c_3 = 3.0 * h_1p;
        double delta = (mata[1] * matb[2] - mata[2] * matb[1]) - (mata[0] * matb[2] - matb[0] * mata[2]) +
                       (mata[0] * matb[1] - matb[0] * mata[1]);
        double deltax = c_3 * (mata[1] * matb[2] - mata[2] * matb[1]) - (c_1 * matb[2] - c_2 * mata[2]) +
                        (c_1 * matb[1] - c_2 * mata[1]);
        double deltay = (c_1 * matb[2] - c_2 * mata[2]) - c_3 * (mata[0] * matb[2] - matb[0] * mata[2]) +
                        (mata[0] * c_2 - matb[0] * c_1);
        double deltaz = (mata[1] * c_2 - matb[1] * c_1) - (mata[0] * c_2 - matb[0] * c_1) +
                        c_3 * (mata[0] * matb[1] - matb[0] * mata[1]);
        double h_1 = deltax / delta, h_2 = deltay / delta, h_3 = deltaz / delta;
        double z_1 = zi_p[0] + hi_p[0] - h_1;
        double z_2 = zi_p[1] + hi_p[1] - h_2;
        double z_3 = zi_p[2] + hi_p[2] - h_3;
        const double *b = vertex + 4 * (int64_t)ROW(nodeid, nw, i)[1];
        ns[0][0] = G[1] - b[1];
        ns[0][1] = -(G[0] - b[0]);
        ns[0][2] = 0.;
        for (int k = 0; k < 3; k++) {
            ns[1][k] = ns[0][k] - ss[1][k];
            ns[2][k] = ns[0][k] + ss[0][k];
        }
        double s_1[2], s_2[2], s_3[2];
        for (int k = 0; k < 2; k++) {
            s_1[k] = 0.5 * h_1 * (zv[0] * ss[0][k] + z_2 * ns[0][k] + z_3 * -1. * ns[2][k]);
            s_2[k] = 0.5 * h_2 * (zv[1] * ss[1][k] + z_1 * -1. * ns[0][k] + z_3 * ns[1][k]);
            s_3[k] = 0.5 * h_3 * (zv[2] * ss[2][k] + z_1 * ns[2][k] + z_2 * -1. * ns[1][k]);
        }
        src_hu[i] = -grav * (s_1[0] + s_2[0] + s_3[0]);
        src_hv[i] = -grav * (s_1[1] + s_2[1] + s_3[1]);
    }
}

/* update_SW (fvm_utils.py:3-14), all five fields; state/rez/src are [5][C] blocks. */
void orc_update_sw5(double *state, const double *rez, const double *src, const double *dissip_hc,
                    const double *corio_hu, const double *corio_hv, double wind_hu, double wind_hv, double dtime,
                    const double *vol, int C)
{
    double *h = state, *hu = state + C, *hv = state + 2 * (int64_t)C, *hc = state + 3 * (int64_t)C,
           *Z = state + 4 * (int64_t)C;
    const double *r[5], *s[5];
    for (int k = 0; k < 5; k++) {
        r[k] = rez + k * (int64_t)C;
        s[k] = src + k * (int64_t)C;
    }
    for (int i = 0; i < C; i++) {
        h[i] += dtime * (r[0][i] + s[0][i]) / vol[i];
        hu[i] += dtime * ((r[1][i] + s[1][i]) / vol[i] + corio_hu[i] + wind_hu);
        hv[i] += dtime * ((r[2][i] + s[2][i]) / vol[i] + corio_hv[i] + wind_hv);
        hc[i] += dtime * (r[3][i] + s[3][i] - dissip_hc[i]) / vol[i];
        Z[i] += dtime * (r[4][i] + s[4][i]) / vol[i];
    }
}

/* term_friction_SW (shallowater/fvm_utils.py:512-529): implicit Manning friction of the cell
 * momenta, hu /= A, hv /= A with A = 1 + time g eta^2 |u| / h^(4/3). */
void orc_term_friction_sw(const double *h, double *hu, double *hv, double grav, double eta, double time, int C)
{
    for (int i = 0; i < C; i++) {
        double ufric = hu[i] / h[i];
        double vfric = hv[i] / h[i];
        double hfric = h[i];
        double A = 1 + time * grav * (eta * eta) * (sqrt(ufric * ufric + vfric * vfric) / pow(hfric, 4. / 3.));
        double hutild = hu[i] / A;
        double hvtild = hv[i] / A;
        hu[i] = hutild;
        hv[i] = hvtild;
    }
}

/* term_coriolis_SW (fvm_utils.py:498-503), with its signs: corio_hu = f hu, corio_hv = -f hv. */
void orc_term_coriolis_sw(const double *hu, const double *hv, double *corio_hu, double *corio_hv, const double *f_c,
                          int C)
{
    for (int i = 0; i < C; i++) {
        corio_hu[i] = f_c[i] * hu[i];
        corio_hv[i] = -f_c[i] * hv[i];
    }
}

/* term_wind_SW (fvm_utils.py:531-544): wind stress C |w| w with C = rho_air (0.75 + 0.067 |w|) 1e-3. */
void orc_term_wind_sw(double *Tx, double *Ty, const double *wx, const double *wy, int C)
{
    const double ro_a = 1.25;
    for (int i = 0; i < C; i++) {
        double u = wx[i], v = wy[i];
        double nor = sqrt(u * u + v * v);
        double Cw = ro_a * (0.75 + 0.067 * nor) * 1e-03;
        Tx[i] = Cw * u * nor;
        Ty[i] = Cw * v * nor;
    }
}
