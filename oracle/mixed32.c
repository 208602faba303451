/*
 * ORACLE (test infrastructure only): CPU restatement, operation for operation, of the
 * float32 arithmetic of the mixed-precision shallow-water kernels (k_sw_pack32,
 * k_sw_recon32, k_sw_flux32 in manapy_amd/csrc/mfv_sw_mixed.hip). Compiled with
 * -ffp-contract=off, float operations round like the GPU's (IEEE division and square
 * root; fused multiply-adds only where both sides call fmaf), so a GPU run must reproduce
 * these results bit for bit; mixed32 results are in turn compared with the float64
 * restatement of the reference scheme (mfv_oracle.c: orc_sw_step) within the stated error.
 *
 * Arrays follow the GPU layout: state [C][4] (h, hu, hv, Z), geoA [C][MF][4], geoB
 * [C][MF][4] (rx, ry, |n|, 1/|n|), facerec [C][MF][4], per-cell tables column-major
 * [slot][cell].
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

static float pmaxf_(float a, float b) { return (b > a) ? b : a; }

static void face_dev(float hb, float aL, float aR, float dz, float huL, float hvL, float huR, float hvR, float dzo,
                     float n0, float n1, float mes, float im, float g, float shift, int own_is_L, float *F0, float *F1,
                     float *F2, float *bo)
{
    float hL = hb + aL, hR = hb + aR;
    const float cL = pmaxf_(0.f, dz), cR = pmaxf_(0.f, -dz);
    const int wetL = hL > 0.f, wetR = hR > 0.f;
    float eL = aL - cL, eR = aR - cR;
    float hsL = hb + eL, hsR = hb + eR;
    /* h - hs of each side from the deviations while both are positive (the difference of the
       two rounded depths would lose the bathymetry step to cancellation) */
    const float dhL = wetL ? ((hsL > 0.f) ? aL - eL : hb + aL) : 0.f;
    const float dhR = wetR ? ((hsR > 0.f) ? aR - eR : hb + aR) : 0.f;
    if (!wetL) hL = 0.f;
    if (!(hsL > 0.f)) {
        hsL = 0.f;
        eL = -hb;
    }
    if (!wetR) hR = 0.f;
    if (!(hsR > 0.f)) {
        hsR = 0.f;
        eR = -hb;
    }
    float uL = 0.f, vL = 0.f, uR = 0.f, vR = 0.f;
    if (hL > 1e-12f) {
        const float r = 1.f / hL;
        uL = huL * r;
        vL = hvL * r;
    }
    if (hR > 1e-12f) {
        const float r = 1.f / hR;
        uR = huR * r;
        vR = hvR * r;
    }
    const float qL = uL * n0 + vL * n1, qR = uR * n0 + vR * n1;
    const float sL = fabsf(qL) * im + sqrtf(g * hsL), sR = fabsf(qR) * im + sqrtf(g * hsR);
    const float s = pmaxf_(sL, sR) * mes;
    const float ho = hb - shift;
    const float dL = eL + shift, dR = eR + shift;
    const float PL = 0.5f * g * dL * (2.f * ho + dL);
    const float PR = 0.5f * g * dR * (2.f * ho + dR);
    *F0 = 0.5f * (hsL * qL + hsR * qR) - 0.5f * s * (eR - eL);
    *F1 = 0.5f * (hsL * uL * qL + hsR * uR * qR) + 0.5f * (PL + PR) * n0 - 0.5f * s * (hsR * uR - hsL * uL);
    *F2 = 0.5f * (hsL * vL * qL + hsR * vR * qR) + 0.5f * (PL + PR) * n1 - 0.5f * s * (hsR * vR - hsL * vL);
    if (own_is_L) {
        *bo = 0.5f * g * dhL * (hL + hsL) + g * (hL + ho) * 0.5f * dzo;
    } else {
        *bo = 0.5f * g * dhR * (hR + hsR) + g * (hR + ho) * 0.5f * dzo;
    }
}

void orc_mixed_pack(int C, int MF, const int32_t *cell_faces, const int32_t *cell_nfaces, const int32_t *face_cells,
                    const double *face_normal, const double *face_center, const double *cell_center, const double *h,
                    const double *hu, const double *hv, const double *Z, const double *zgrad, double om,
                    double *state, float *geoA, float *geoB)
{
    for (int i = 0; i < C; i++) {
        state[4 * (size_t)i] = h[i];
        state[4 * (size_t)i + 1] = hu[i];
        state[4 * (size_t)i + 2] = hv[i];
        state[4 * (size_t)i + 3] = Z[i];
        const double cx = cell_center[3 * (size_t)i], cy = cell_center[3 * (size_t)i + 1];
        for (int s = 0; s < MF && s < cell_nfaces[i]; s++) {
            const int f = cell_faces[(size_t)s * C + i];
            const double fx = face_center[3 * (size_t)f], fy = face_center[3 * (size_t)f + 1];
            const double rx = fx - cx, ry = fy - cy;
            const double zo = Z[i] + om * (zgrad[i] * rx + zgrad[C + i] * ry);
            const int L = face_cells[2 * f], R = face_cells[2 * f + 1];
            double dz = 0.;
            if (R >= 0) {
                const int j = (L == i) ? R : L;
                const double rjx = fx - cell_center[3 * (size_t)j], rjy = fy - cell_center[3 * (size_t)j + 1];
                const double zt = Z[j] + om * (zgrad[j] * rjx + zgrad[C + j] * rjy);
                dz = (L == i) ? zt - zo : zo - zt;
            }
            float *ga = geoA + 4 * ((size_t)MF * i + s);
            ga[0] = (float)face_normal[3 * (size_t)f];
            ga[1] = (float)face_normal[3 * (size_t)f + 1];
            ga[2] = (float)dz;
            ga[3] = (float)(zo - Z[i]);
            float *gb = geoB + 4 * ((size_t)MF * i + s);
            gb[0] = (float)rx;
            gb[1] = (float)ry;
            gb[2] = sqrtf(ga[0] * ga[0] + ga[1] * ga[1]);
            gb[3] = 1.f / gb[2];
        }
    }
}

static void accum(float jx, float jy, float de, float du, float dv, float *a)
{
    a[0] = fmaf(jx, jx, a[0]);
    a[1] = fmaf(jy, jy, a[1]);
    a[2] = fmaf(jx, jy, a[2]);
    a[3] = fmaf(jx, de, a[3]);
    a[4] = fmaf(jy, de, a[4]);
    a[5] = fmaf(jx, du, a[5]);
    a[6] = fmaf(jy, du, a[6]);
    a[7] = fmaf(jx, dv, a[7]);
    a[8] = fmaf(jy, dv, a[8]);
}

void orc_mixed_recon(int C, int MF, const double *state, const double *cell_center, const int32_t *lsq_cells,
                     const int32_t *lsq_ncells, const int32_t *lsq_ghosts, const int32_t *lsq_nghosts,
                     const int32_t *face_cells, const double *face_ghost, const int32_t *cell_nfaces,
                     const int32_t *nbr, const float *geoA, const float *geoB, float om, float *facerec)
{
    for (int i = 0; i < C; i++) {
        const double *si = state + 4 * (size_t)i;
        const double hi = si[0], ei = si[0] + si[3], ui = si[1], vi = si[2];
        const double cx = cell_center[3 * (size_t)i], cy = cell_center[3 * (size_t)i + 1];
        float a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = 0; j < lsq_ncells[i]; j++) {
            const int c = lsq_cells[(size_t)j * C + i];
            const double *sc = state + 4 * (size_t)c;
            accum((float)(cell_center[3 * (size_t)c] - cx), (float)(cell_center[3 * (size_t)c + 1] - cy),
                  (float)((sc[0] + sc[3]) - ei), (float)(sc[1] - ui), (float)(sc[2] - vi), a);
        }
        for (int j = 0; j < lsq_nghosts[i]; j++) {
            const int f = lsq_ghosts[(size_t)j * C + i];
            const double *so = state + 4 * (size_t)face_cells[2 * f];
            const double *gc = face_ghost + 3 * (size_t)f;
            accum((float)(gc[0] - cx), (float)(gc[1] - cy), (float)((so[0] + so[3]) - ei), (float)(-so[1] - ui),
                  (float)(-so[2] - vi), a);
        }
        const float ixx = a[0], iyy = a[1], ixy = a[2];
        const float dia = ixx * iyy - ixy * ixy;
        const float idia = 1.f / dia;
        const float ex = (iyy * a[3] - ixy * a[4]) * idia, ey = (ixx * a[4] - ixy * a[3]) * idia;
        const float ux = (iyy * a[5] - ixy * a[6]) * idia, uy = (ixx * a[6] - ixy * a[5]) * idia;
        const float vx = (iyy * a[7] - ixy * a[8]) * idia, vy = (ixx * a[8] - ixy * a[7]) * idia;
        const int nf = cell_nfaces[i];
        float emax = 0.f, emin = 0.f, umax = 0.f, umin = 0.f, vmax = 0.f, vmin = 0.f;
        float dh[8], rx[8], ry[8];
        for (int s = 0; s < MF; s++) {
            dh[s] = 0.f;
            if (s >= nf) continue;
            const int code = nbr[(size_t)s * C + i];
            float eb, ub, vb;
            if (code >= 0) {
                const double *sb = state + 4 * (size_t)(code >> 3);
                eb = (float)((sb[0] + sb[3]) - ei);
                ub = (float)(sb[1] - ui);
                vb = (float)(sb[2] - vi);
                if (code & 1) dh[s] = (float)(hi - sb[0]);
            } else {
                eb = 0.f;
                ub = (float)(-2. * ui);
                vb = (float)(-2. * vi);
            }
            emax = fmaxf(emax, eb);
            emin = fminf(emin, eb);
            umax = fmaxf(umax, ub);
            umin = fminf(umin, ub);
            vmax = fmaxf(vmax, vb);
            vmin = fminf(vmin, vb);
        }
        float pe = 1.f, pu = 1.f, pv = 1.f;
        for (int s = 0; s < MF; s++) {
            rx[s] = geoB[4 * ((size_t)MF * i + s)];
            ry[s] = geoB[4 * ((size_t)MF * i + s) + 1];
            if (s >= nf) continue;
            float d = ex * rx[s] + ey * ry[s];
            if (!(fabsf(d) < 1e-8f)) pe = fminf(pe, d > 0.f ? emax / d : emin / d);
            d = ux * rx[s] + uy * ry[s];
            if (!(fabsf(d) < 1e-8f)) pu = fminf(pu, d > 0.f ? umax / d : umin / d);
            d = vx * rx[s] + vy * ry[s];
            if (!(fabsf(d) < 1e-8f)) pv = fminf(pv, d > 0.f ? vmax / d : vmin / d);
        }
        const float hu32 = (float)ui, hv32 = (float)vi, h32 = (float)hi;
        for (int s = 0; s < MF && s < nf; s++) {
            const float dzo = geoA[4 * ((size_t)MF * i + s) + 3];
            const float de = om * pe * (ex * rx[s] + ey * ry[s]);
            const int code = nbr[(size_t)s * C + i];
            const int right = code >= 0 && (code & 1);
            float *o = facerec + 4 * ((size_t)MF * i + s);
            o[0] = right ? dh[s] + (de - dzo) : de - dzo;
            o[1] = hu32 + om * pu * (ux * rx[s] + uy * ry[s]);
            o[2] = hv32 + om * pv * (vx * rx[s] + vy * ry[s]);
            o[3] = right ? dh[s] : h32;
        }
    }
}

/* One step (fluxes + float64 update) of every cell in place; returns the next dt (the minimum
 * of the per-cell candidates, as the GPU's wavefront minimum). */
double orc_mixed_flux(int C, int MF, double *state, const double *vol, const float *facerec, const float *geoA,
                      const float *geoB, const int32_t *cell_nfaces, const int32_t *nbr, const uint8_t *order,
                      float g, double cfl, double dt)
{
    double dmin = 1e300;
    for (int i = 0; i < C; i++) {
        double *si = state + 4 * (size_t)i;
        const int nf = cell_nfaces[i];
        const unsigned ord = order[i];
        double r0 = 0., r1 = 0., r2 = 0., s1 = 0., s2 = 0.;
        for (int r = 0; r < MF && r < nf; r++) {
            const int s = (ord >> (2 * r)) & 3;
            const float *ga = geoA + 4 * ((size_t)MF * i + s);
            const float *gb = geoB + 4 * ((size_t)MF * i + s);
            const float mes = gb[2], im = gb[3];
            const float *own = facerec + 4 * ((size_t)MF * i + s);
            const int code = nbr[(size_t)s * C + i];
            float F0, F1, F2, bo;
            if (code < 0) {
                face_dev(own[3], own[0], own[0], 0.f, own[1], own[2], -own[1], -own[2], ga[3], ga[0], ga[1], mes, im, g,
                         0.f, 1, &F0, &F1, &F2, &bo);
            } else {
                const float *oth = facerec + 4 * ((size_t)MF * (code >> 3) + ((code >> 1) & 3));
                if ((code & 1) == 0)
                    face_dev(own[3], own[0], oth[0], ga[2], own[1], own[2], oth[1], oth[2], ga[3], ga[0], ga[1], mes,
                             im, g, 0.f, 1, &F0, &F1, &F2, &bo);
                else
                    face_dev(oth[3], oth[0], own[0], ga[2], oth[1], oth[2], own[1], own[2], ga[3], ga[0], ga[1], mes,
                             im, g, -own[3], 0, &F0, &F1, &F2, &bo);
            }
            if (code < 0 || (code & 1) == 0) {
                r0 += (double)-F0;
                r1 += (double)-F1;
                r2 += (double)-F2;
                s1 += (double)-(bo * ga[0]);
                s2 += (double)-(bo * ga[1]);
            } else {
                r0 += (double)F0;
                r1 += (double)F1;
                r2 += (double)F2;
                s1 += (double)(bo * ga[0]);
                s2 += (double)(bo * ga[1]);
            }
        }
        const double v = vol[i];
        const double R0 = r0;
        const double R1 = r1 + s1;
        const double R2 = r2 + s2;
        const double dv = dt / v;
        const double hn = si[0] + R0 * dv;
        const double un = si[1] + R1 * dv;
        const double vn = si[2] + R2 * dv;
        si[0] = hn;
        si[1] = un;
        si[2] = vn;
        const double velson = sqrt((double)g * hn);
        const double ih = 1. / hn;
        double lam = 0.;
        for (int s = 0; s < MF && s < nf; s++) {
            const float *ga = geoA + 4 * ((size_t)MF * i + s);
            const double n0 = ga[0], n1 = ga[1], mesu = (double)sqrtf(ga[0] * ga[0] + ga[1] * ga[1]);
            lam += fabs((un * n0 + vn * n1) * ih) + velson * mesu;
        }
        const double dl = cfl * v / lam;
        if (dl < dmin || !(dmin == dmin)) dmin = dl;
    }
    return dmin;
}

/* time_step_SW of the current state with the float32 geometry (the ensembles' first dt). */
double orc_mixed_dt(int C, int MF, const double *state, const double *vol, const float *geoA,
                    const int32_t *cell_nfaces, float g, double cfl)
{
    double dmin = 1e300;
    for (int i = 0; i < C; i++) {
        const double hn = state[4 * (size_t)i], un = state[4 * (size_t)i + 1], vn = state[4 * (size_t)i + 2];
        const double velson = sqrt((double)g * hn);
        const double ih = 1. / hn;
        double lam = 0.;
        for (int s = 0; s < MF && s < cell_nfaces[i]; s++) {
            const float *ga = geoA + 4 * ((size_t)MF * i + s);
            const double n0 = ga[0], n1 = ga[1], mesu = (double)sqrtf(ga[0] * ga[0] + ga[1] * ga[1]);
            lam += fabs((un * n0 + vn * n1) * ih) + velson * mesu;
        }
        const double dl = cfl * vol[i] / lam;
        if (dl < dmin) dmin = dl;
    }
    return dmin;
}
