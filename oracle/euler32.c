/*
 * ORACLE (test infrastructure only): CPU model, operation for operation, of the float32-bulk
 * Euler step (k_euler_step32_tiled in manapy_amd/csrc/mfv_euler.hip), the FP32-bulk variant
 * of explicitscheme_euler_3d_rusanov + update_euler_3d_fvc + time_step_euler_3d
 * (solvers/euler/fvm_utils3d.py:838-886, 1083-1170, 1316-1327, 809-825).
 *
 * Storage is float32: the state record [C][8] = (rho, P, rhou, rhov, rhow, rhoE, |K|, 0) and
 * the face geometry [F][4] = (nx, ny, nz, |f|) (the float32 roundings of the float64 mesh).
 * Each face's Rusanov flux is evaluated once in float32 from its two cells' records (or the
 * slip/transmissive ghost of its owner); the pressure part of the momentum flux is kept in
 * deviation form, 1/2 (P_L - P_R) n, which both cells add (the same scheme in exact
 * arithmetic, since sum_f P_K n_f = 0 over the closed cell K, and exactly zero where the
 * pressure is uniform, where float32 normals would otherwise leave a rounding-sized force).
 * Each cell adds its faces' terms in float64 in increasing face id (the reference's face
 * order), updates in float64 and stores float32; the next dt is formed in float64 from the
 * stored state. Compiled with -ffp-contract=off, so a GPU run reproduces these bits.
 */
#include <math.h>
#include <stdint.h>

static float pmaxf_(float a, float b) { return (b > a) ? b : a; }

/* Rusanov flux of one face, canonical orientation L -> R, n the float32 normal and mes its
 * float32 measure. F[0..4]: mass, momentum without pressure, energy; *pd: 1/2 (P_L - P_R). */
static void flux32(const float *L, const float *R, const float *g, float gamma, int quirk, float *F, float *pd)
{
    const float mes = g[3];
    const float n0 = g[0] / mes, n1 = g[1] / mes, n2 = g[2] / mes;
    const float rl = L[0], Pl = L[1], ul = L[2], vl = L[3], wl = L[4], El = L[5];
    const float rr = R[0], Pr = R[1], ur = R[2], vr = R[3], wr = R[4], Er = R[5];
    const float ql = ul * n0 + vl * n1 + wl * n2;
    const float qr = quirk ? (ur * n0 + vl * n1 + wl * n2) : (ur * n0 + vr * n1 + wr * n2);
    const float qrt = ur * n0 + vr * n1 + wr * n2;
    const float cl = sqrtf(gamma * Pl / rl), cr = sqrtf(gamma * Pr / rr);
    const float al = ql / rl, ar = qr / rr;
    const float Ll = pmaxf_(pmaxf_(fabsf(al - cl), fabsf(al)), fabsf(al + cl));
    const float Lr = pmaxf_(pmaxf_(fabsf(ar - cr), fabsf(ar)), fabsf(ar + cr));
    const float S = (Ll > Lr) ? Ll : Lr;
    const float sl = ql / rl, sr = qrt / rr;  /* normal velocities */
    F[0] = (0.5f * (ql + qrt) - 0.5f * S * (rr - rl)) * mes;
    F[1] = (0.5f * (ul * sl + ur * sr) - 0.5f * S * (ur - ul)) * mes;
    F[2] = (0.5f * (vl * sl + vr * sr) - 0.5f * S * (vr - vl)) * mes;
    F[3] = (0.5f * (wl * sl + wr * sr) - 0.5f * S * (wr - wl)) * mes;
    F[4] = (0.5f * ((El + Pl) * sl + (Er + Pr) * sr) - 0.5f * S * (Er - El)) * mes;
    *pd = 0.5f * (Pl - Pr) * mes;
}

/* ghost_value_TubeSchok3D of one boundary face from its owner (names 1, 2 transmissive; 3-6
 * slip walls, quirk 1: rhovg = rho w_g, rhowg = 0 (zero-initialised ghosts)). */
static void ghost32(const float *c, const float *g, int name, int quirk, float *o)
{
    for (int k = 0; k < 6; k++) o[k] = c[k];
    if (name >= 3 && name <= 6) {
        const float u = c[2] / c[0], v = c[3] / c[0], w = c[4] / c[0];
        const float nn = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const float s0 = g[0] / nn, s1 = g[1] / nn, s2 = g[2] / nn;
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

/* Round the float64 state and geometry to the float32 records. */
void orc_euler32_pack(int C, int F, const double *U, const double *vol, const double *normal, const double *mesure,
                      float *rec, float *geo)
{
    for (int i = 0; i < C; i++) {
        for (int k = 0; k < 6; k++) rec[8 * (int64_t)i + k] = (float)U[(int64_t)k * C + i];
        rec[8 * (int64_t)i + 6] = (float)vol[i];
        rec[8 * (int64_t)i + 7] = 0.f;
    }
    for (int f = 0; f < F; f++) {
        for (int k = 0; k < 3; k++) geo[4 * (int64_t)f + k] = (float)normal[3 * (int64_t)f + k];
        geo[4 * (int64_t)f + 3] = (float)mesure[f];
    }
}

/* per-cell time_step_euler_3d of a float32 record with the float32 face geometry, in float64 */
static double cell_lam(const float *u, const float *const *g, double gamma)
{
    const double r = u[0], P = u[1], mx = u[2], my = u[3], mz = u[4];
    const double velson = sqrt(gamma * fabs(P / r));
    double lam = 0.;
    for (int j = 0; j < 4; j++) {
        const double nx = g[j][0], ny = g[j][1], nz = g[j][2];
        const double u_n = fabs((mx * nx + my * ny + mz * nz) / r) / sqrt(nx * nx + ny * ny + nz * nz);
        lam += (u_n + velson) * (double)g[j][3];
    }
    return lam;
}

/* One step: in -> out (float32 records), fluxes of every face once, per-cell float64 gather
 * in increasing face id, float64 update, float32 store. faceid: cells.faceid [C][mf1] (slot
 * order, for the dt sum); sorted: the cell's faces in increasing id [C][4]; flux scratch
 * [F][6]. cons[5]: float64 sums of rho, rhou, rhov, rhow, rhoE times |K| over the new state.
 * Returns the next dt. */
double orc_euler32_step(int C, int F, const float *in, float *out, const float *geo, const int32_t *face_cells,
                        const int32_t *face_name, const int32_t *faceid, int mf1, const int32_t *sorted,
                        double gamma, double cfl, double dt, int quirk, float *flx, double *cons)
{
    const float g32 = (float)gamma;
    for (int f = 0; f < F; f++) {
        const float *L = in + 8 * (int64_t)face_cells[2 * f];
        const int b = face_cells[2 * f + 1];
        float Rg[6];
        const float *R;
        if (b >= 0) {
            R = in + 8 * (int64_t)b;
        } else {
            ghost32(L, geo + 4 * (int64_t)f, face_name[f], quirk, Rg);
            R = Rg;
        }
        flux32(L, R, geo + 4 * (int64_t)f, g32, quirk, flx + 6 * (int64_t)f, flx + 6 * (int64_t)f + 5);
    }
    for (int v = 0; v < 5; v++) cons[v] = 0.;
    double dmin = 1e300;
    for (int i = 0; i < C; i++) {
        const float *u32 = in + 8 * (int64_t)i;
        double r[5] = {0., 0., 0., 0., 0.};
        for (int k = 0; k < 4; k++) {
            const int f = sorted[4 * (int64_t)i + k];
            const float *o = flx + 6 * (int64_t)f;
            const float *g = geo + 4 * (int64_t)f;
            const int own_L = face_cells[2 * f] == i;
            const float pd = o[5];
            if (own_L) {
                r[0] -= (double)o[0];
                r[1] -= (double)o[1];
                r[2] -= (double)o[2];
                r[3] -= (double)o[3];
                r[4] -= (double)o[4];
            } else {
                r[0] += (double)o[0];
                r[1] += (double)o[1];
                r[2] += (double)o[2];
                r[3] += (double)o[3];
                r[4] += (double)o[4];
            }
            r[1] += (double)(pd * (g[0] / g[3]));
            r[2] += (double)(pd * (g[1] / g[3]));
            r[3] += (double)(pd * (g[2] / g[3]));
        }
        const double vol = (double)u32[6];
        const double dv = dt / vol;
        const double rho = (double)u32[0] + r[0] * dv;
        const double mx = (double)u32[2] + r[1] * dv;
        const double my = (double)u32[3] + r[2] * dv;
        const double mz = (double)u32[4] + r[3] * dv;
        const double E = (double)u32[5] + r[4] * dv;
        const double P = (gamma - 1.) * (E - 0.5 * (mx * mx + my * my + mz * mz) / rho);
        float *w = out + 8 * (int64_t)i;
        w[0] = (float)rho;
        w[1] = (float)P;
        w[2] = (float)mx;
        w[3] = (float)my;
        w[4] = (float)mz;
        w[5] = (float)E;
        w[6] = u32[6];
        w[7] = 0.f;
        const double vw = (double)w[6];
        cons[0] += (double)w[0] * vw;
        cons[1] += (double)w[2] * vw;
        cons[2] += (double)w[3] * vw;
        cons[3] += (double)w[4] * vw;
        cons[4] += (double)w[5] * vw;
        const float *g[4];
        for (int j = 0; j < 4; j++) g[j] = geo + 4 * (int64_t)faceid[(int64_t)i * mf1 + j];
        const double d = cfl * vw / cell_lam(w, g, gamma);
        if (d < dmin) dmin = d;
    }
    return dmin;
}
