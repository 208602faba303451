/*
 * ORACLE -- test infrastructure only. Never linked into the product library.
 *
 * Plain-C, strict-IEEE (built with -O2 -ffp-contract=off -fno-fast-math)
 * restatement of manapy's CPU loops for the explicit finite-volume step, in the
 * reference's iteration order: face-ordered scatters (inner faces, then
 * boundary faces), cell loops in cell order, per-cell sums in stencil order.
 * Only the single-subdomain, non-periodic case exists (halo and periodic
 * loops of the reference are empty there).
 *
 * Array conventions follow manapy (ddm/domain.py): padded connectivity
 * [rows, width+1] with the count in the last column, int32 indices, float64
 * values, normals [F,3] non-unit with |n| = face length/area.
 *
 * What pins each function is listed in its comment and tested in tests/.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define ROW(a, w, i) ((a) + (int64_t)(i) * (w))

static double pmin(double a, double b) { return (b < a) ? b : a; } /* Python min(a, b) */
static double pmax(double a, double b) { return (b > a) ? b : a; } /* Python max(a, b) */

/* ---------------------------------------------------------------------------------------- */
/* advection: solvers/advec/fvm_utils.py                                                    */

/* time_step (fvm_utils.py:293-309). Pinned: closed form on a single triangle. */
double orc_time_step(const double *u, const double *v, const double *w, double cfl, const double *normal,
                     const double *volume, const int32_t *faceid, int mf1, int C)
{
    double dt = 1e6;
    for (int i = 0; i < C; i++) {
        const int32_t *fr = ROW(faceid, mf1, i);
        double lam = 0.;
        for (int j = 0; j < fr[mf1 - 1]; j++) {
            const double *n = normal + 3 * (int64_t)fr[j];
            lam += fabs(u[i] * n[0] + v[i] * n[1] + w[i] * n[2]);
        }
        dt = pmin(dt, cfl * volume[i] / lam);
    }
    return dt;
}

/* update_new_value (fvm_utils.py:311-315). Pinned: algebraic identity. */
void orc_update_new_value(double *ne_c, const double *rez, const double *dissip, const double *src, double dtime,
                          const double *vol, int C)
{
    for (int i = 0; i < C; i++)
        ne_c[i] += dtime * ((rez[i] + dissip[i]) / vol[i] + src[i]);
}

/* facetocell (ast/ast_utils.py:8-18). */
void orc_facetocell(const double *u_face, double *u_c, const int32_t *faceid, int mf1, int C)
{
    for (int i = 0; i < C; i++) {
        const int32_t *fr = ROW(faceid, mf1, i);
        u_c[i] = 0.;
        for (int j = 0; j < fr[mf1 - 1]; j++) u_c[i] += u_face[fr[j]];
    }
    for (int i = 0; i < C; i++) u_c[i] /= (double)ROW(faceid, mf1, i)[mf1 - 1];
}

/* ghost_value_{neumann,nonslip,dirichlet} (boundary/bc_utils.py:45-68); kind 0/1/2. */
void orc_ghost_value(int kind, const double *w_c_or_value, double *w_ghost, const int32_t *cellid,
                     const int32_t *faces, int nf)
{
    for (int k = 0; k < nf; k++) {
        int f = faces[k];
        if (kind == 0) w_ghost[f] = w_c_or_value[cellid[2 * f]];
        else if (kind == 1) w_ghost[f] = -1 * w_c_or_value[cellid[2 * f]];
        else w_ghost[f] = w_c_or_value[f];
    }
}

/* explicitscheme_convective_2d / _3d (fvm_utils.py:19-152, 156-290), upwind flux
 * compute_upwind_flux (:5-17). dim selects the reconstruction dot product.
 * Pinned: closed-form flux per face, global conservation, constant states. */
void orc_explicitscheme_convective(int dim, double *rez_w, const double *w_c, const double *w_ghost,
                                   const double *u_face, const double *v_face, const double *w_face,
                                   const double *w_x, const double *w_y, const double *w_z, const double *psi,
                                   const double *centerc, const double *centerf, const int32_t *cellidf,
                                   const double *normalf, const int32_t *innerfaces, int n_inner,
                                   const int32_t *boundaryfaces, int n_bnd, int order, int C)
{
    for (int i = 0; i < C; i++) rez_w[i] = 0.;
    for (int k = 0; k < n_inner; k++) {
        int f = innerfaces[k];
        int c0 = cellidf[2 * f], c1 = cellidf[2 * f + 1];
        const double *n = normalf + 3 * (int64_t)f;
        double w_l = w_c[c0], w_r = w_c[c1];
        double rl0 = centerf[3 * f] - centerc[3 * c0], rr0 = centerf[3 * f] - centerc[3 * c1];
        double rl1 = centerf[3 * f + 1] - centerc[3 * c0 + 1], rr1 = centerf[3 * f + 1] - centerc[3 * c1 + 1];
        if (dim == 2) {
            w_l = w_l + (double)(order - 1) * psi[c0] * (w_x[c0] * rl0 + w_y[c0] * rl1);
            w_r = w_r + (double)(order - 1) * psi[c1] * (w_x[c1] * rr0 + w_y[c1] * rr1);
        } else {
            double rl2 = centerf[3 * f + 2] - centerc[3 * c0 + 2], rr2 = centerf[3 * f + 2] - centerc[3 * c1 + 2];
            w_l = w_l + (double)(order - 1) * psi[c0] * (w_x[c0] * rl0 + w_y[c0] * rl1 + w_z[c0] * rl2);
            w_r = w_r + (double)(order - 1) * psi[c1] * (w_x[c1] * rr0 + w_y[c1] * rr1 + w_z[c1] * rr2);
        }
        double sign = u_face[f] * n[0] + v_face[f] * n[1] + w_face[f] * n[2];
        double sol = (sign >= 0) ? w_l : w_r;
        double flux = sign * sol;
        rez_w[c0] -= flux;
        rez_w[c1] += flux;
    }
    for (int k = 0; k < n_bnd; k++) {
        int f = boundaryfaces[k];
        int c0 = cellidf[2 * f];
        const double *n = normalf + 3 * (int64_t)f;
        double w_l = w_c[c0], w_r = w_ghost[f];
        double rl0 = centerf[3 * f] - centerc[3 * c0];
        double rl1 = centerf[3 * f + 1] - centerc[3 * c0 + 1];
        if (dim == 2) {
            w_l = w_l + (double)(order - 1) * psi[c0] * (w_x[c0] * rl0 + w_y[c0] * rl1);
        } else {
            double rl2 = centerf[3 * f + 2] - centerc[3 * c0 + 2];
            w_l = w_l + (double)(order - 1) * psi[c0] * (w_x[c0] * rl0 + w_y[c0] * rl1 + w_z[c0] * rl2);
        }
        double sign = u_face[f] * n[0] + v_face[f] * n[1] + w_face[f] * n[2];
        double sol = (sign >= 0) ? w_l : w_r;
        double flux = sign * sol;
        rez_w[c0] -= flux;
    }
}

/* ---------------------------------------------------------------------------------------- */
/* gradient + limiter: ast/functions2d.py                                                   */

/* cell_gradient_2d (functions2d.py:5-104), one-rank non-periodic terms only.
 * Pinned: exact for affine fields; zero for constants. */
void orc_cell_gradient_2d_range(const double *w_c, const double *w_ghost, const double *centerc,
                                const int32_t *cellnid, int k1, const int32_t *ghostnid, int g1,
                                const double *centergf, int gcw, double *w_x, double *w_y, double *w_z, int i0,
                                int i1);
void orc_cell_gradient_2d(const double *w_c, const double *w_ghost, const double *centerc,
                          const int32_t *cellnid, int k1, const int32_t *ghostnid, int g1,
                          const double *centergf, int gcw, double *w_x, double *w_y, double *w_z, int C)
{
    orc_cell_gradient_2d_range(w_c, w_ghost, centerc, cellnid, k1, ghostnid, g1, centergf, gcw, w_x, w_y, w_z, 0, C);
}

/* the loop of orc_cell_gradient_2d over cells [i0, i1) */
void orc_cell_gradient_2d_range(const double *w_c, const double *w_ghost, const double *centerc,
                                const int32_t *cellnid, int k1, const int32_t *ghostnid, int g1,
                                const double *centergf, int gcw, double *w_x, double *w_y, double *w_z, int i0,
                                int i1)
{
    for (int i = i0; i < i1; i++) {
        double i_xx = 0., i_yy = 0., i_xy = 0., j_xw = 0., j_yw = 0.;
        const int32_t *nr = ROW(cellnid, k1, i);
        for (int j = 0; j < nr[k1 - 1]; j++) {
            int c = nr[j];
            double jx = centerc[3 * c] - centerc[3 * i];
            double jy = centerc[3 * c + 1] - centerc[3 * i + 1];
            i_xx += jx * jx;
            i_yy += jy * jy;
            i_xy += (jx * jy);
            j_xw += (jx * (w_c[c] - w_c[i]));
            j_yw += (jy * (w_c[c] - w_c[i]));
        }
        const int32_t *gr = ROW(ghostnid, g1, i);
        for (int j = 0; j < gr[g1 - 1]; j++) {
            int f = gr[j];
            double jx = centergf[(int64_t)gcw * f] - centerc[3 * i];
            double jy = centergf[(int64_t)gcw * f + 1] - centerc[3 * i + 1];
            i_xx += jx * jx;
            i_yy += jy * jy;
            i_xy += (jx * jy);
            j_xw += (jx * (w_ghost[f] - w_c[i]));
            j_yw += (jy * (w_ghost[f] - w_c[i]));
        }
        double dia = i_xx * i_yy - i_xy * i_xy;
        w_x[i] = (i_yy * j_xw - i_xy * j_yw) / dia;
        w_y[i] = (i_xx * j_yw - i_xy * j_xw) / dia;
        w_z[i] = 0.;
    }
}

/* barthlimiter_2d (functions2d.py:280-325), one-rank branches (inner / names 1-4).
 * Pinned: psi in [0,1]; psi = 1 for constants; brute force on tiny meshes. */
void orc_barthlimiter_2d_range(const double *w_c, const double *w_ghost, const double *w_x, const double *w_y,
                               double *psi, const int32_t *cellid, const int32_t *faceid, int mf1,
                               const int32_t *namef, const double *centerc, const double *centerf, int i0, int i1);
void orc_barthlimiter_2d(const double *w_c, const double *w_ghost, const double *w_x, const double *w_y,
                         double *psi, const int32_t *cellid, const int32_t *faceid, int mf1,
                         const int32_t *namef, const double *centerc, const double *centerf, int C)
{
    orc_barthlimiter_2d_range(w_c, w_ghost, w_x, w_y, psi, cellid, faceid, mf1, namef, centerc, centerf, 0, C);
}

/* the loops of orc_barthlimiter_2d over cells [i0, i1) */
void orc_barthlimiter_2d_range(const double *w_c, const double *w_ghost, const double *w_x, const double *w_y,
                               double *psi, const int32_t *cellid, const int32_t *faceid, int mf1,
                               const int32_t *namef, const double *centerc, const double *centerf, int i0, int i1)
{
    for (int i = i0; i < i1; i++) psi[i] = 1.;
    for (int i = i0; i < i1; i++) {
        double w_max = w_c[i], w_min = w_c[i];
        const int32_t *fr = ROW(faceid, mf1, i);
        int nfc = fr[mf1 - 1];
        for (int j = 0; j < nfc; j++) {
            int f = fr[j];
            int nm = namef[f];
            double a = w_c[cellid[2 * f]];
            double b = (nm == 0 || nm > 10) ? w_c[cellid[2 * f + 1]] : w_ghost[f];
            w_max = pmax(pmax(w_max, a), b);
            w_min = pmin(pmin(w_min, a), b);
        }
        double psi_ij = 1.;
        for (int j = 0; j < nfc; j++) {
            int f = fr[j];
            double r1 = centerf[3 * f] - centerc[3 * i];
            double r2 = centerf[3 * f + 1] - centerc[3 * i + 1];
            double delta2 = w_x[i] * r1 + w_y[i] * r2;
            if (fabs(delta2) < 1e-8) {
                psi_ij = 1.;
            } else {
                if (delta2 > 0.) psi_ij = pmin(1., (w_max - w_c[i]) / delta2);
                if (delta2 < 0.) psi_ij = pmin(1., (w_min - w_c[i]) / delta2);
            }
            psi[i] = pmin(psi[i], psi_ij);
        }
    }
}

/* cell_gradient_3d (functions3d.py:4-120), one-rank non-periodic terms. x**2 is x*x. */
void orc_cell_gradient_3d(const double *w_c, const double *w_ghost, const double *centerc,
                          const int32_t *cellnid, int k1, const int32_t *ghostnid, int g1,
                          const double *centergf, int gcw, double *w_x, double *w_y, double *w_z, int C)
{
    for (int i = 0; i < C; i++) {
        double i_xx = 0., i_yy = 0., i_zz = 0., i_xy = 0., i_xz = 0., i_yz = 0., j_x = 0., j_y = 0., j_z = 0.;
        const int32_t *nr = ROW(cellnid, k1, i);
        const int32_t *gr = ROW(ghostnid, g1, i);
        int nn = nr[k1 - 1], ng = gr[g1 - 1];
        for (int j = 0; j < nn + ng; j++) {
            const double *p;
            double wv;
            if (j < nn) { p = centerc + 3 * (int64_t)nr[j]; wv = w_c[nr[j]]; }
            else { p = centergf + (int64_t)gcw * gr[j - nn]; wv = w_ghost[gr[j - nn]]; }
            double jx = p[0] - centerc[3 * i], jy = p[1] - centerc[3 * i + 1], jz = p[2] - centerc[3 * i + 2];
            i_xx += jx * jx; i_yy += jy * jy; i_zz += jz * jz;
            i_xy += (jx * jy); i_xz += (jx * jz); i_yz += (jy * jz);
            j_x += (jx * (wv - w_c[i])); j_y += (jy * (wv - w_c[i])); j_z += (jz * (wv - w_c[i]));
        }
        double dia = i_xx * i_yy * i_zz + 2. * i_xy * i_xz * i_yz - i_xx * (i_yz * i_yz) - i_yy * (i_xz * i_xz)
                     - i_zz * (i_xy * i_xy);
        w_x[i] = ((i_yy * i_zz - i_yz * i_yz) * j_x + (i_xz * i_yz - i_xy * i_zz) * j_y + (i_xy * i_yz - i_xz * i_yy) * j_z) / dia;
        w_y[i] = ((i_xz * i_yz - i_xy * i_zz) * j_x + (i_xx * i_zz - i_xz * i_xz) * j_y + (i_xy * i_xz - i_yz * i_xx) * j_z) / dia;
        w_z[i] = ((i_xy * i_yz - i_xz * i_yy) * j_x + (i_xy * i_xz - i_yz * i_xx) * j_y + (i_xx * i_yy - i_xy * i_xy) * j_z) / dia;
    }
}

/* barthlimiter_3d (functions3d.py:364-408), one-rank branches. */
void orc_barthlimiter_3d(const double *w_c, const double *w_ghost, const double *w_x, const double *w_y,
                         const double *w_z, double *psi, const int32_t *cellid, const int32_t *faceid, int mf1,
                         const int32_t *namef, const double *centerc, const double *centerf, int C)
{
    for (int i = 0; i < C; i++) psi[i] = 1.;
    for (int i = 0; i < C; i++) {
        double w_max = w_c[i], w_min = w_c[i];
        const int32_t *fr = ROW(faceid, mf1, i);
        int nfc = fr[mf1 - 1];
        for (int j = 0; j < nfc; j++) {
            int f = fr[j];
            int nm = namef[f];
            double a = w_c[cellid[2 * f]];
            double b = (nm == 0 || nm > 10) ? w_c[cellid[2 * f + 1]] : w_ghost[f];
            w_max = pmax(pmax(w_max, a), b);
            w_min = pmin(pmin(w_min, a), b);
        }
        double psi_ij = 1.;
        for (int j = 0; j < nfc; j++) {
            int f = fr[j];
            double r1 = centerf[3 * f] - centerc[3 * i];
            double r2 = centerf[3 * f + 1] - centerc[3 * i + 1];
            double r3 = centerf[3 * f + 2] - centerc[3 * i + 2];
            double delta2 = w_x[i] * r1 + w_y[i] * r2 + w_z[i] * r3;
            if (fabs(delta2) < 1e-10) {
                psi_ij = 1.;
            } else {
                if (delta2 > 0.) psi_ij = pmin(1., (w_max - w_c[i]) / delta2);
                if (delta2 < 0.) psi_ij = pmin(1., (w_min - w_c[i]) / delta2);
            }
            psi[i] = pmin(psi[i], psi_ij);
        }
    }
}

/* ---------------------------------------------------------------------------------------- */
/* shallow water                                                                            */

/* time_step_SW (solvers/shallowater/fvm_utils.py:16-42). Pinned: closed form. */
double orc_time_step_sw_range(const double *h, const double *hu, const double *hv, double cfl,
                              const double *normal, const double *mesure, const double *volume,
                              const int32_t *faceid, int mf1, double grav, double Dxx, double Dyy, int i0, int i1);
double orc_time_step_sw(const double *h, const double *hu, const double *hv, double cfl, const double *normal,
                        const double *mesure, const double *volume, const int32_t *faceid, int mf1, double grav,
                        double Dxx, double Dyy, int C)
{
    return orc_time_step_sw_range(h, hu, hv, cfl, normal, mesure, volume, faceid, mf1, grav, Dxx, Dyy, 0, C);
}

/* the minimum of orc_time_step_sw over cells [i0, i1) (from the reference's initial 1e6) */
double orc_time_step_sw_range(const double *h, const double *hu, const double *hv, double cfl,
                              const double *normal, const double *mesure, const double *volume,
                              const int32_t *faceid, int mf1, double grav, double Dxx, double Dyy, int i0, int i1)
{
    double dt = 1e6;
    for (int i = i0; i < i1; i++) {
        double velson = sqrt(grav * h[i]);
        double lam = 0.;
        const int32_t *fr = ROW(faceid, mf1, i);
        for (int j = 0; j < fr[mf1 - 1]; j++) {
            int f = fr[j];
            const double *n = normal + 3 * (int64_t)f;
            double u_n = fabs(hu[i] / h[i] * n[0] + hv[i] / h[i] * n[1]);
            double lam_convect = u_n / mesure[f] + velson;
            lam += lam_convect * mesure[f];
            double mes = sqrt(n[0] * n[0] + n[1] * n[1]);
            double lam_diff = Dxx * (mes * mes) + Dyy * (mes * mes);
            lam += lam_diff / volume[i];
        }
        dt = pmin(dt, cfl * volume[i] / lam);
    }
    return dt;
}

/* update_SW (fvm_utils.py:3-14) for h, hu, hv (Coriolis and wind terms zero). */
void orc_update_sw_range(double *h, double *hu, double *hv, const double *rez_h, const double *rez_hu,
                         const double *rez_hv, const double *src_h, const double *src_hu, const double *src_hv,
                         double dtime, const double *vol, int i0, int i1);
void orc_update_sw(double *h, double *hu, double *hv, const double *rez_h, const double *rez_hu,
                   const double *rez_hv, const double *src_h, const double *src_hu, const double *src_hv,
                   double dtime, const double *vol, int C)
{
    orc_update_sw_range(h, hu, hv, rez_h, rez_hu, rez_hv, src_h, src_hu, src_hv, dtime, vol, 0, C);
}

/* the loop of orc_update_sw over cells [i0, i1) */
void orc_update_sw_range(double *h, double *hu, double *hv, const double *rez_h, const double *rez_hu,
                         const double *rez_hv, const double *src_h, const double *src_hu, const double *src_hv,
                         double dtime, const double *vol, int i0, int i1)
{
    for (int i = i0; i < i1; i++) {
        h[i] += dtime * (rez_h[i] + src_h[i]) / vol[i];
        hu[i] += dtime * ((rez_hu[i] + src_hu[i]) / vol[i] + 0. + 0.);
        hv[i] += dtime * ((rez_hv[i] + src_hv[i]) / vol[i] + 0. + 0.);
    }
}

/* Shallow-water face flux with hydrostatic reconstruction (Audusse et al., SIAM J.
 * Sci. Comput. 25 (2004) 2050-2065) and the Rusanov (local Lax-Friedrichs) flux.
 * NOT in the reference (its SW flux is SRNH): parity unpinned against manapy;
 * pinned by consistency F(U,U) = F_phys(U).n, lake-at-rest balance, mass
 * conservation and left/right antisymmetry (tests/test_oracle_sw.py).
 *
 * Inputs are the reconstructed face values on each side: eta (= h + Z), hu, hv,
 * Z at the face, and the cell-average h and Z of the owning cell (for the
 * centred bathymetry term). Outputs: flux[3] in the face orientation (from L to
 * R, |n| included) and the bathymetry source of each side, srcL[2] / srcR[2],
 * written as the quantity to SUBTRACT from that side's residual with the
 * face normal oriented OUT of that side. */
#define SW_DRY 1e-12
void orc_sw_face(double etaL, double huL, double hvL, double zfL, double hcL, double zcL,
                 double etaR, double huR, double hvR, double zfR, double hcR, double zcR,
                 double n0, double n1, double mes, double g, double *flux, double *srcL, double *srcR)
{
    double hL = etaL - zfL; if (hL < 0.) hL = 0.;
    double hR = etaR - zfR; if (hR < 0.) hR = 0.;
    double uL = 0., vL = 0., uR = 0., vR = 0.;
    if (hL > SW_DRY) { uL = huL / hL; vL = hvL / hL; }
    if (hR > SW_DRY) { uR = huR / hR; vR = hvR / hR; }
    double zs = pmax(zfL, zfR);
    double hsL = hL + zfL - zs; if (hsL < 0.) hsL = 0.;
    double hsR = hR + zfR - zs; if (hsR < 0.) hsR = 0.;
    double qL = uL * n0 + vL * n1;
    double qR = uR * n0 + vR * n1;
    double pL = 0.5 * g * hsL * hsL;
    double pR = 0.5 * g * hsR * hsR;
    double fL0 = hsL * qL, fL1 = hsL * uL * qL + pL * n0, fL2 = hsL * vL * qL + pL * n1;
    double fR0 = hsR * qR, fR1 = hsR * uR * qR + pR * n0, fR2 = hsR * vR * qR + pR * n1;
    double sL = fabs(qL) / mes + sqrt(g * hsL);
    double sR = fabs(qR) / mes + sqrt(g * hsR);
    double s = pmax(sL, sR) * mes;
    flux[0] = 0.5 * (fL0 + fR0) - 0.5 * s * (hsR - hsL);
    flux[1] = 0.5 * (fL1 + fR1) - 0.5 * s * (hsR * uR - hsL * uL);
    flux[2] = 0.5 * (fL2 + fR2) - 0.5 * s * (hsR * vR - hsL * vL);
    /* per side: g/2 (h_f^2 - h*_f^2) + g (h_f + h_cell)/2 (Z_f - Z_cell), times the outward normal */
    double bL = 0.5 * g * (hL * hL - hsL * hsL) + g * (hL + hcL) * 0.5 * (zfL - zcL);
    double bR = 0.5 * g * (hR * hR - hsR * hsR) + g * (hR + hcR) * 0.5 * (zfR - zcR);
    srcL[0] = bL * n0; srcL[1] = bL * n1;
    srcR[0] = -(bR * n0); srcR[1] = -(bR * n1);
}

/* SW MUSCL step residuals, face-ordered scatter over inner faces then boundary
 * faces (the loop structure of explicitscheme_convective_SW,
 * shallowater/fvm_utils.py:344-496). Walls: h Neumann, hu/hv nonslip, Z Neumann,
 * applied to the reconstructed face state (ghost arrays h_g.. are accepted for the
 * interface but the wall face only reads the reconstructed owner values).
 * eta = h + Z is reconstructed with its own limited gradient; Z with an
 * unlimited gradient (zx, zy). rez_* are the convective residuals, src_* the
 * bathymetry terms; both are zeroed first. */
void orc_sw_face_terms(int f, int inner, const double *h, const double *hu, const double *hv, const double *Z,
                       const double *ex, const double *ey, const double *epsi, const double *ux, const double *uy,
                       const double *upsi, const double *vx, const double *vy, const double *vpsi,
                       const double *zx, const double *zy, const double *centerc, const double *centerf,
                       const int32_t *cellidf, const double *normalf, const double *mesuref, double grav, double om,
                       double *flux, double *sL, double *sR);

void orc_sw_residuals(double *rez_h, double *rez_hu, double *rez_hv, double *src_h, double *src_hu, double *src_hv,
                      const double *h, const double *hu, const double *hv, const double *Z,
                      const double *h_g, const double *hu_g, const double *hv_g, const double *Z_g,
                      const double *ex, const double *ey, const double *epsi,
                      const double *ux, const double *uy, const double *upsi,
                      const double *vx, const double *vy, const double *vpsi,
                      const double *zx, const double *zy,
                      const double *centerc, const double *centerf, const int32_t *cellidf,
                      const double *normalf, const double *mesuref,
                      const int32_t *innerfaces, int n_inner, const int32_t *boundaryfaces, int n_bnd,
                      double grav, int order, int C)
{
    double om = (double)(order - 1);
    for (int i = 0; i < C; i++) {
        rez_h[i] = 0.; rez_hu[i] = 0.; rez_hv[i] = 0.;
        src_h[i] = 0.; src_hu[i] = 0.; src_hv[i] = 0.;
    }
    double flux[3], sL[2], sR[2];
    for (int k = 0; k < n_inner; k++) {
        int f = innerfaces[k];
        int a = cellidf[2 * f], b = cellidf[2 * f + 1];
        orc_sw_face_terms(f, 1, h, hu, hv, Z, ex, ey, epsi, ux, uy, upsi, vx, vy, vpsi, zx, zy, centerc, centerf,
                          cellidf, normalf, mesuref, grav, om, flux, sL, sR);
        rez_h[a] -= flux[0]; rez_hu[a] -= flux[1]; rez_hv[a] -= flux[2];
        src_hu[a] -= sL[0]; src_hv[a] -= sL[1];
        rez_h[b] += flux[0]; rez_hu[b] += flux[1]; rez_hv[b] += flux[2];
        src_hu[b] -= sR[0]; src_hv[b] -= sR[1];
    }
    for (int k = 0; k < n_bnd; k++) {
        int f = boundaryfaces[k];
        int a = cellidf[2 * f];
        orc_sw_face_terms(f, 0, h, hu, hv, Z, ex, ey, epsi, ux, uy, upsi, vx, vy, vpsi, zx, zy, centerc, centerf,
                          cellidf, normalf, mesuref, grav, om, flux, sL, sR);
        rez_h[a] -= flux[0]; rez_hu[a] -= flux[1]; rez_hv[a] -= flux[2];
        src_hu[a] -= sL[0]; src_hv[a] -= sL[1];
    }
    (void)h_g; (void)hu_g; (void)hv_g; (void)Z_g;
}

/* The MUSCL face values of face f and its flux and bathymetry terms (orc_sw_face): inner faces
 * reconstruct from both cells; a boundary (wall) face uses the Neumann (eta, Z) / nonslip
 * (hu, hv) mirror of the reconstructed owner state as its right side, which at order 1 is the
 * cell ghost h_g, -hu, -hv, Z_g of bc_utils.py:45-55. */
void orc_sw_face_terms(int f, int inner, const double *h, const double *hu, const double *hv, const double *Z,
                       const double *ex, const double *ey, const double *epsi, const double *ux, const double *uy,
                       const double *upsi, const double *vx, const double *vy, const double *vpsi,
                       const double *zx, const double *zy, const double *centerc, const double *centerf,
                       const int32_t *cellidf, const double *normalf, const double *mesuref, double grav, double om,
                       double *flux, double *sL, double *sR)
{
    int a = cellidf[2 * f];
    double ra0 = centerf[3 * f] - centerc[3 * a], ra1 = centerf[3 * f + 1] - centerc[3 * a + 1];
    double etaA = h[a] + Z[a];
    double eL = etaA + om * epsi[a] * (ex[a] * ra0 + ey[a] * ra1);
    double uL = hu[a] + om * upsi[a] * (ux[a] * ra0 + uy[a] * ra1);
    double vL = hv[a] + om * vpsi[a] * (vx[a] * ra0 + vy[a] * ra1);
    double zL = Z[a] + om * (zx[a] * ra0 + zy[a] * ra1);
    const double *n = normalf + 3 * (int64_t)f;
    if (inner) {
        int b = cellidf[2 * f + 1];
        double rb0 = centerf[3 * f] - centerc[3 * b], rb1 = centerf[3 * f + 1] - centerc[3 * b + 1];
        double etaB = h[b] + Z[b];
        double eR = etaB + om * epsi[b] * (ex[b] * rb0 + ey[b] * rb1);
        double uR = hu[b] + om * upsi[b] * (ux[b] * rb0 + uy[b] * rb1);
        double vR = hv[b] + om * vpsi[b] * (vx[b] * rb0 + vy[b] * rb1);
        double zR = Z[b] + om * (zx[b] * rb0 + zy[b] * rb1);
        orc_sw_face(eL, uL, vL, zL, h[a], Z[a], eR, uR, vR, zR, h[b], Z[b], n[0], n[1], mesuref[f], grav,
                    flux, sL, sR);
    } else {
        orc_sw_face(eL, uL, vL, zL, h[a], Z[a], eL, -1 * uL, -1 * vL, zL, h[a], Z[a], n[0], n[1],
                    mesuref[f], grav, flux, sL, sR);
    }
}

/* ---------------------------------------------------------------------------------------- */
/* 3D Euler: solvers/euler/fvm_utils3d.py                                                   */

/* ghost_value_TubeSchok3D (fvm_utils3d.py:441-500). quirk = 1 keeps the reference's
 * overwrite of rhovg by rho*w_g with rhowg left unset (Q2, :493-494); quirk = 0 sets
 * rhovg = rho v_g and rhowg = rho w_g. Arrays are [F]; only boundary names act. */
void orc_ghost_tubeschok3d(double *rhog, double *Pg, double *rhoug, double *rhovg, double *rhowg, double *rhoEg,
                           const double *rhoc, const double *Pc, const double *rhouc, const double *rhovc,
                           const double *rhowc, const double *rhoEc, const int32_t *cellid, const int32_t *name,
                           const double *normal, int F, int quirk)
{
    for (int i = 0; i < F; i++) {
        int nm = name[i];
        int c = cellid[2 * i];
        if (nm == 1 || nm == 2) {
            rhog[i] = rhoc[c]; rhoug[i] = rhouc[c]; rhovg[i] = rhovc[c]; rhowg[i] = rhowc[c];
            rhoEg[i] = rhoEc[c]; Pg[i] = Pc[c];
        } else if (nm == 3 || nm == 4 || nm == 5 || nm == 6) {
            double u_i = rhouc[c] / rhoc[c], v_i = rhovc[c] / rhoc[c], w_i = rhowc[c] / rhoc[c];
            const double *n = normal + 3 * (int64_t)i;
            double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double s0 = n[0] / nn, s1 = n[1] / nn, s2 = n[2] / nn;
            double un = u_i * s0 + v_i * s1 + w_i * s2;
            double u_g = u_i - 2 * un * s0;
            double v_g = v_i - 2 * un * s1;
            double w_g = w_i - 2 * un * s2;
            rhog[i] = rhoc[c];
            rhoug[i] = rhoc[c] * u_g;
            if (quirk) {
                rhovg[i] = rhoc[c] * w_g;
            } else {
                rhovg[i] = rhoc[c] * v_g;
                rhowg[i] = rhoc[c] * w_g;
            }
            rhoEg[i] = rhoEc[c];
            Pg[i] = Pc[c];
        }
    }
}

/* time_step_euler_3d (fvm_utils3d.py:809-825): per-cell dt over the first 4 faces. */
void orc_time_step_euler_3d(const double *rho, const double *P, const double *rhou, const double *rhov,
                            const double *rhow, double cfl, const double *normal, const double *mesure,
                            const double *volume, const int32_t *faceid, int mf1, double gamma, double *dt_c, int C)
{
    for (int i = 0; i < C; i++) {
        double lam = 0.;
        double velson = sqrt(gamma * fabs(P[i] / rho[i]));
        const int32_t *fr = ROW(faceid, mf1, i);
        for (int j = 0; j < 4; j++) {
            const double *n = normal + 3 * (int64_t)fr[j];
            double u_n = fabs((rhou[i] * n[0] + rhov[i] * n[1] + rhow[i] * n[2]) / rho[i])
                         / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            double lam_convect = u_n + velson;
            lam += lam_convect * mesure[fr[j]];
        }
        dt_c[i] = cfl * volume[i] / lam;
    }
}

/* compute_flux_euler_3d_rusanov (fvm_utils3d.py:838-886). quirk = 1 keeps
 * qr = rhou_R n0 + rhov_L n1 + rhow_L n2 (Q1, :842); quirk = 0 uses the right state. */
void orc_flux_euler_rusanov(double rhol, double Pl, double rhoul, double rhovl, double rhowl, double rhoEl,
                            double rhor, double Pr, double rhour, double rhovr, double rhowr, double rhoEr,
                            const double *n, double mesure, double gamma, int quirk, double *fl)
{
    double ql = (rhoul * n[0] + rhovl * n[1] + rhowl * n[2]);
    double qr = quirk ? (rhour * n[0] + rhovl * n[1] + rhowl * n[2]) : (rhour * n[0] + rhovr * n[1] + rhowr * n[2]);
    double cl = sqrt(gamma * Pl / rhol);
    double cr = sqrt(gamma * Pr / rhor);
    double fl_rho = (rhoul * n[0] + rhovl * n[1] + rhowl * n[2]);
    double fl_rhou = ((rhoul * rhoul / rhol + Pl) * n[0] + (rhoul * rhovl / rhol) * n[1] + (rhoul * rhowl / rhol) * n[2]);
    double fl_rhov = ((rhovl * rhoul / rhol) * n[0] + (rhovl * rhovl / rhol + Pl) * n[1] + (rhovl * rhowl / rhol) * n[2]);
    double fl_rhow = ((rhoul * rhowl / rhol) * n[0] + (rhovl * rhowl / rhol) * n[1] + (rhowl * rhowl / rhol + Pl) * n[2]);
    double fl_rhoE = (rhoul / rhol * (rhoEl + Pl) * n[0] + rhovl / rhol * (rhoEl + Pl) * n[1] + rhowl / rhol * (rhoEl + Pl) * n[2]);
    double fr_rho = (rhour * n[0] + rhovr * n[1] + rhowr * n[2]);
    double fr_rhou = ((rhour * rhour / rhor + Pr) * n[0] + (rhour * rhovr / rhor) * n[1] + (rhour * rhowr / rhor) * n[2]);
    double fr_rhov = ((rhovr * rhour / rhor) * n[0] + (rhovr * rhovr / rhor + Pr) * n[1] + (rhovr * rhowr / rhor) * n[2]);
    double fr_rhow = ((rhour * rhowr / rhor) * n[0] + (rhovr * rhowr / rhor) * n[1] + (rhowr * rhowr / rhor + Pr) * n[2]);
    double fr_rhoE = (rhour / rhor * (rhoEr + Pr) * n[0] + rhovr / rhor * (rhoEr + Pr) * n[1] + rhowr / rhor * (rhoEr + Pr) * n[2]);
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

/* explicitscheme_euler_3d_rusanov (fvm_utils3d.py:1083-1170): one loop over all
 * faces; name 0 scatters to both sides, any other name uses the ghost state.
 * U = [6][C] cell states (rho, P, rhou, rhov, rhow, rhoE), G = [6][F] ghosts,
 * rez = [5][C]. */
void orc_euler3d_rusanov(double *rez, const double *U, const double *G, const int32_t *cellidf,
                         const double *normal, const double *mesurf, const int32_t *name, double gamma,
                         int quirk, int C, int F)
{
    for (int k = 0; k < 5 * C; k++) rez[k] = 0.;
    const double *rho = U, *P = U + C, *ru = U + 2 * (int64_t)C, *rv = U + 3 * (int64_t)C, *rw = U + 4 * (int64_t)C,
                 *rE = U + 5 * (int64_t)C;
    const double *grho = G, *gP = G + F, *gru = G + 2 * (int64_t)F, *grv = G + 3 * (int64_t)F,
                 *grw = G + 4 * (int64_t)F, *grE = G + 5 * (int64_t)F;
    double fl[5], nrm[3];
    for (int i = 0; i < F; i++) {
        nrm[0] = normal[3 * (int64_t)i] / mesurf[i];
        nrm[1] = normal[3 * (int64_t)i + 1] / mesurf[i];
        nrm[2] = normal[3 * (int64_t)i + 2] / mesurf[i];
        int a = cellidf[2 * i];
        if (name[i] == 0) {
            int b = cellidf[2 * i + 1];
            orc_flux_euler_rusanov(rho[a], P[a], ru[a], rv[a], rw[a], rE[a], rho[b], P[b], ru[b], rv[b], rw[b], rE[b],
                                   nrm, mesurf[i], gamma, quirk, fl);
            for (int v = 0; v < 5; v++) {
                rez[(int64_t)v * C + a] -= fl[v];
                rez[(int64_t)v * C + b] += fl[v];
            }
        } else {
            orc_flux_euler_rusanov(rho[a], P[a], ru[a], rv[a], rw[a], rE[a], grho[i], gP[i], gru[i], grv[i], grw[i],
                                   grE[i], nrm, mesurf[i], gamma, quirk, fl);
            for (int v = 0; v < 5; v++) rez[(int64_t)v * C + a] -= fl[v];
        }
    }
}

/* update_euler_3d_fvc (fvm_utils3d.py:1316-1327). */
void orc_update_euler_3d(double *U, const double *rez, double gamma, double dtime, const double *vol, int C)
{
    double *rho = U, *P = U + C, *ru = U + 2 * (int64_t)C, *rv = U + 3 * (int64_t)C, *rw = U + 4 * (int64_t)C,
           *rE = U + 5 * (int64_t)C;
    for (int i = 0; i < C; i++) {
        rho[i] += dtime * ((rez[i]) / vol[i]);
        ru[i] += dtime * ((rez[C + i]) / vol[i]);
        rv[i] += dtime * ((rez[2 * (int64_t)C + i]) / vol[i]);
        rw[i] += dtime * ((rez[3 * (int64_t)C + i]) / vol[i]);
        rE[i] += dtime * ((rez[4 * (int64_t)C + i]) / vol[i]);
        P[i] = (gamma - 1) * (rE[i] - 0.5 * (ru[i] * ru[i] + rv[i] * rv[i] + rw[i] * rw[i]) / rho[i]);
    }
}
