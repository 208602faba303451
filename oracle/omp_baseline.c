/*
 * ORACLE (test infrastructure and CPU baseline only; never linked into the product library).
 *
 * The all-cores CPU baseline of the shallow-water step: the oracle's own per-cell and
 * per-face functions (mfv_oracle.c, included below, so the arithmetic is the same code),
 * run with OpenMP over cells and faces. The reference's face-ordered scatter (inner faces,
 * then boundary faces; shallowater/fvm_utils.py:344-496) becomes a per-face pass that forms
 * each face's flux once and a per-cell gather that applies them in the scatter's order, so
 * every cell sums the same terms in the same order and the result equals the serial oracle
 * bit for bit.
 */
#include "mfv_oracle.c"

#include <omp.h>
#include <stdlib.h>

int orc_omp_threads(void) { return omp_get_max_threads(); }

/* One step of oracle.sw_step (time_step_SW -> wall ghosts -> limited LSQ gradients of eta, hu,
 * hv -> MUSCL/Rusanov residuals -> update_SW), in place on h, hu, hv; returns the dt used.
 * scratch: at least 26 C + 13 F doubles, caller-owned (kept across steps). */
double orc_sw_step_omp(int C, int F, double *h, double *hu, double *hv, const double *Z, const double *zx,
                       const double *zy, const double *centerc, const double *centerf, const int32_t *cellidf,
                       const double *normalf, const double *mesuref, const int32_t *faceid, int mf1,
                       const int32_t *namef, const int32_t *cellnid, int k1, const int32_t *ghostnid, int g1,
                       const double *centergf, int gcw, const int32_t *innerfaces, int n_inner,
                       const int32_t *boundaryfaces, int n_bnd, const double *vol, double grav, double cfl,
                       int order, double *scratch)
{
    double *eta = scratch, *grad = eta + C, *rez = grad + 9 * (size_t)C, *src = rez + 3 * (size_t)C;
    double *gh = src + 3 * (size_t)C;          /* ghosts: eta, hu, hv, and h, Z [5][F] */
    double *fl = gh + 5 * (size_t)F;           /* per face: flux[3], srcL[2], srcR[2], key */
    double *wz = fl + 8 * (size_t)F;           /* [C] zeros (w_z of the 2D gradient) */
    double dt = 1e6;
    const double om = (double)(order - 1);
#pragma omp parallel
    {
        const int nt = omp_get_num_threads(), t = omp_get_thread_num();
        const int c0 = (int)((int64_t)C * t / nt), c1 = (int)((int64_t)C * (t + 1) / nt);
        /* time_step_SW: the minimum is the same in any order */
        const double dmin = orc_time_step_sw_range(h, hu, hv, cfl, normalf, mesuref, vol, faceid, mf1, grav, 0., 0.,
                                                   c0, c1);
#pragma omp critical
        if (dmin < dt) dt = dmin;
        for (int i = c0; i < c1; i++) {
            eta[i] = h[i] + Z[i];
            wz[i] = 0.;
        }
#pragma omp for
        for (int f = 0; f < F; f++) {
            gh[f] = 0.; gh[F + f] = 0.; gh[2 * (size_t)F + f] = 0.;
        }
#pragma omp for
        for (int k = 0; k < n_bnd; k++) {  /* walls: h, Z Neumann; hu, hv nonslip */
            const int f = boundaryfaces[k], a = cellidf[2 * f];
            gh[f] = h[a] + Z[a];
            gh[F + f] = -hu[a];
            gh[2 * (size_t)F + f] = -hv[a];
        }
        /* limited LSQ gradients, cell ranges of the oracle's own loops */
        if (order == 2) {
            const double *w[3] = {eta, hu, hv};
            for (int q = 0; q < 3; q++) {
                double *gx = grad + 3 * q * (size_t)C, *gy = gx + C;
                orc_cell_gradient_2d_range(w[q], gh + q * (size_t)F, centerc, cellnid, k1, ghostnid, g1, centergf,
                                           gcw, gx, gy, wz, c0, c1);
            }
#pragma omp barrier
            for (int q = 0; q < 3; q++) {
                double *gx = grad + 3 * q * (size_t)C, *gy = gx + C, *ps = gy + C;
                orc_barthlimiter_2d_range(w[q], gh + q * (size_t)F, gx, gy, ps, cellidf, faceid, mf1, namef,
                                          centerc, centerf, c0, c1);
            }
        } else {
            for (size_t j = 0; j < 9; j++)
                for (int i = c0; i < c1; i++) grad[j * C + i] = 0.;
        }
#pragma omp barrier
        /* every face once: its flux and bathymetry terms, and its position in the scatter */
#pragma omp for
        for (int k = 0; k < n_inner + n_bnd; k++) {
            const int inner = k < n_inner;
            const int f = inner ? innerfaces[k] : boundaryfaces[k - n_inner];
            double *o = fl + 8 * (size_t)f;
            orc_sw_face_terms(f, inner, h, hu, hv, Z, grad, grad + C, grad + 2 * (size_t)C, grad + 3 * (size_t)C,
                              grad + 4 * (size_t)C, grad + 5 * (size_t)C, grad + 6 * (size_t)C,
                              grad + 7 * (size_t)C, grad + 8 * (size_t)C, zx, zy, centerc, centerf, cellidf, normalf,
                              mesuref, grav, om, o, o + 3, o + 5);
            o[7] = (double)k;
        }
        /* each cell applies its faces' terms in the scatter's order */
        for (int i = c0; i < c1; i++) {
            const int32_t *fr = ROW(faceid, mf1, i);
            const int nf = fr[mf1 - 1];
            int ord[8];
            for (int j = 0; j < nf; j++) {
                int x = fr[j], p = j;
                while (p > 0 && fl[8 * (size_t)ord[p - 1] + 7] > fl[8 * (size_t)x + 7]) {
                    ord[p] = ord[p - 1];
                    p--;
                }
                ord[p] = x;
            }
            double r0 = 0., r1 = 0., r2 = 0., s1 = 0., s2 = 0.;
            for (int j = 0; j < nf; j++) {
                const int f = ord[j];
                const double *o = fl + 8 * (size_t)f;
                if (cellidf[2 * f] == i) {
                    r0 -= o[0]; r1 -= o[1]; r2 -= o[2];
                    s1 -= o[3]; s2 -= o[4];
                } else {
                    r0 += o[0]; r1 += o[1]; r2 += o[2];
                    s1 -= o[5]; s2 -= o[6];
                }
            }
            rez[i] = r0; rez[C + i] = r1; rez[2 * (size_t)C + i] = r2;
            src[i] = 0.; src[C + i] = s1; src[2 * (size_t)C + i] = s2;
        }
#pragma omp barrier
        orc_update_sw_range(h, hu, hv, rez, rez + C, rez + 2 * (size_t)C, src, src + C, src + 2 * (size_t)C, dt, vol,
                            c0, c1);
    }
    return dt;
}
