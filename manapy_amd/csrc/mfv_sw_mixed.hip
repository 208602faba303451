// libmfv mixed-precision shallow-water path and its ensembles.
//
// Same scheme as the float64 paths (MUSCL on eta, hu, hv with LSQ gradients and the
// Barth-Jespersen limiter, hydrostatic reconstruction, Rusanov flux, centred bathymetry term,
// mirrored walls), arranged so that float32 arithmetic only ever sees small quantities:
//   * the cell states (h, hu, hv, Z) stay in float64; the reconstruction takes float64
//     differences to its neighbours and runs the least-squares gradient and limiter in float32;
//   * a face record holds the face depth as a deviation from the depth of the face's left cell
//     (the face's canonical base), the face momenta, and that base (left side) or the deviation
//     of the own cell from it (right side): both cells of a face evaluate the face from the
//     same float32 numbers, so the mass flux is the same bits on both sides;
//   * pressures enter the momentum fluxes as deviations from the own cell's hydrostatic
//     pressure 1/2 g h_K^2 (whose contribution sum_f n_f vanishes on a closed cell);
//   * the residual gathers add the float32 face terms in float64 (float64 adds cost what
//     float32 adds do on this GPU, and a sum of a few float32 terms is then exact or nearly),
//     so the float64 update conserves mass to round-off.
//   * the least-squares sums use fused multiply-adds (fmaf), as the CPU model does.
// dt and the conservation totals are float64.
#include "mfv_common.hpp"

namespace {

using namespace mfv;

// Python max(a, b) on floats.
__device__ __forceinline__ float pmaxf(float a, float b) { return (b > a) ? b : a; }

// One face in its canonical orientation L -> R (normal n out of L, |n| = mes).
//   hb          depth of the L cell (float32)
//   aL, aR      face depths (eta_f - z_f) minus hb
//   dz          z_f(R) - z_f(L)
//   huL..hvR    face momenta
//   dzo         z_f - Z of the own side
//   shift       hb - h_own (0 when the own cell is L)
// Outputs: mass and momentum fluxes L -> R (momentum with pressures relative to the own cell)
// and the own side's bathymetry term.
__device__ __forceinline__ void face_dev(float hb, float aL, float aR, float dz, float huL, float hvL, float huR,
                                         float hvR, float dzo, float n0, float n1, float mes, float im, float g,
                                         float shift, bool own_is_L, float &F0, float &F1, float &F2, float &bo)
{
    float hL = hb + aL, hR = hb + aR;
    const float cL = pmaxf(0.f, dz), cR = pmaxf(0.f, -dz);
    const bool wetL = hL > 0.f, wetR = hR > 0.f;
    float eL = aL - cL, eR = aR - cR;  // hydrostatic depths minus hb
    float hsL = hb + eL, hsR = hb + eR;
    // h - hs of each side from the deviations (= the bathymetry step c up to rounding) while both
    // are positive: the difference of the two rounded depths would lose the step to cancellation
    // (an error of ~1e-7 h in the bathymetry term that breaks the lake-at-rest balance)
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
    const float s = pmaxf(sL, sR) * mes;
    const float ho = hb - shift;
    const float dL = eL + shift, dR = eR + shift;  // hs - h_own
    const float PL = 0.5f * g * dL * (2.f * ho + dL);
    const float PR = 0.5f * g * dR * (2.f * ho + dR);
    F0 = 0.5f * (hsL * qL + hsR * qR) - 0.5f * s * (eR - eL);
    F1 = 0.5f * (hsL * uL * qL + hsR * uR * qR) + 0.5f * (PL + PR) * n0 - 0.5f * s * (hsR * uR - hsL * uL);
    F2 = 0.5f * (hsL * vL * qL + hsR * vR * qR) + 0.5f * (PL + PR) * n1 - 0.5f * s * (hsR * vR - hsL * vL);
    const float hO = own_is_L ? hL : hR, hsO = own_is_L ? hsL : hsR;
    const float dh = own_is_L ? dhL : dhR;
    bo = 0.5f * g * dh * (hO + hsO) + g * (hO + ho) * 0.5f * dzo;
}

// Static face geometry and the constant bathymetry differences (Z gradient from zgrad [3][C]):
//   geoA[i][s] = (n0, n1, zf_R - zf_L, zf_own - Z_i), geoB[i][s] = (rx, ry, |n|, 1/|n|) with |n|
//   the float32 length of the float32 normal (what the flux and dt would compute every step)
template <int MF>
__global__ void __launch_bounds__(kBlock) k_sw_pack32(Mesh m, mfv_sw_mixed p, const double *__restrict__ h,
                                                       const double *__restrict__ hu,
                                                       const double *__restrict__ hv, const double *__restrict__ Z,
                                                       const double *__restrict__ zgrad, double om)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.C) return;
    const size_t C = m.C;
    reinterpret_cast<double4 *>(p.state)[i] = make_double4(h[i], hu[i], hv[i], Z[i]);
    const int nf = m.cell_nfaces[i];
    const double cx = m.cell_center[3 * (size_t)i], cy = m.cell_center[3 * (size_t)i + 1];
    for (int s = 0; s < MF && s < nf; s++) {
        const int f = m.cell_faces[(size_t)s * C + i];
        const double fx = m.face_center[3 * (size_t)f], fy = m.face_center[3 * (size_t)f + 1];
        const double rx = fx - cx, ry = fy - cy;
        const double zo = Z[i] + om * (zgrad[i] * rx + zgrad[C + i] * ry);
        const int L = m.face_cells[2 * f], R = m.face_cells[2 * f + 1];
        double dz = 0.;
        if (R >= 0) {
            const int j = (L == i) ? R : L;
            const double rjx = fx - m.cell_center[3 * (size_t)j], rjy = fy - m.cell_center[3 * (size_t)j + 1];
            const double zt = Z[j] + om * (zgrad[j] * rjx + zgrad[C + j] * rjy);
            dz = (L == i) ? zt - zo : zo - zt;
        }
        const double *n = m.face_normal + 3 * (size_t)f;
        const float n0 = (float)n[0], n1 = (float)n[1];
        reinterpret_cast<float4 *>(p.geoA)[MF * (size_t)i + s] = make_float4(n0, n1, (float)dz, (float)(zo - Z[i]));
        const float mes = sqrtf(n0 * n0 + n1 * n1);
        reinterpret_cast<float4 *>(p.geoB)[MF * (size_t)i + s] = make_float4((float)rx, (float)ry, mes, 1.f / mes);
    }
}

__device__ __forceinline__ void accum_grad(float jx, float jy, float de, float du, float dv, float &ixx, float &iyy,
                                           float &ixy, float &je0, float &je1, float &ju0, float &ju1, float &jv0,
                                           float &jv1)
{
    ixx = fmaf(jx, jx, ixx);
    iyy = fmaf(jy, jy, iyy);
    ixy = fmaf(jx, jy, ixy);
    je0 = fmaf(jx, de, je0);
    je1 = fmaf(jy, de, je1);
    ju0 = fmaf(jx, du, ju0);
    ju1 = fmaf(jy, du, ju1);
    jv0 = fmaf(jx, dv, jv0);
    jv1 = fmaf(jy, dv, jv1);
}

// Static geometry of a cell's face slots in faceid order (geoA and the |n|, 1/|n| of geoB), as
// native vectors: element s with a lane-dependent s is a select chain.
struct SlotGeo {
    slot_f nx, ny, dz, dzo, len, ilen;
    int lc0, lc1, lc2, lc3, nf;  // neighbour codes in the accessor's numbering, face count
};

// LEN: |n| and 1/|n| from geoB, else computed from the normal (the same float32 expressions).
template <int MF, bool LEN>
__device__ __forceinline__ SlotGeo load_slot_geo(const mfv_sw_mixed &p, int i)
{
    SlotGeo G{0.f, 0.f, 0.f, 0.f, 1.f, 1.f, -1, -1, -1, -1, 0};
#pragma unroll
    for (int s = 0; s < MF; s++) {
        const float4 q = reinterpret_cast<const float4 *>(p.geoA)[MF * (size_t)i + s];
        G.nx[s] = q.x;
        G.ny[s] = q.y;
        G.dz[s] = q.z;
        G.dzo[s] = q.w;
        if (LEN) {
            const float4 b = reinterpret_cast<const float4 *>(p.geoB)[MF * (size_t)i + s];
            G.len[s] = b.z;
            G.ilen[s] = b.w;
        } else {
            G.len[s] = sqrtf(q.x * q.x + q.y * q.y);
            G.ilen[s] = 1.f / G.len[s];
        }
    }
    return G;
}

// Reconstruction of cell i: float32 LSQ gradients of eta, hu, hv from float64 differences,
// Barth-Jespersen limiter on deviations, face records (a, hu_f, hv_f, w) of its nf faces in
// out[0..nf) -- see the file comment. st is the state the step starts from.
// Where the reconstruction of one cell reads its inputs: cell states and centres by a
// handle (a global cell index, or a position in a block's LDS copy), the stencil, ghost
// owners and face neighbours as handles.
struct GlobalCells {
    static constexpr int kChunk = 4;  // neighbour states loaded per group (global memory latency)
    const Mesh &m;
    const double4 *__restrict__ st;
    int i;
    __device__ __forceinline__ int self() const { return i; }
    __device__ __forceinline__ double4 state(int h) const { return st[h]; }
    __device__ __forceinline__ double2 center(int h) const
    {
        return make_double2(m.cell_center[3 * (size_t)h], m.cell_center[3 * (size_t)h + 1]);
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

// KMAX > 0: the stencil handles are loaded first and the neighbour states in independent
// groups of Cells::kChunk (a few memory latencies per cell instead of two per stencil entry);
// the sums keep the stencil order. KMAX = 0: plain loop, for stencils wider than KMAX.
template <int MF, int KMAX, class Cells>
__device__ __forceinline__ int recon32_core(const Mesh &m, const mfv_sw_mixed &p, const Cells &a, int i, float om,
                                            float4 *out, SlotGeo *keep = nullptr)
{
    const size_t C = m.C;
    // Loads in batches ahead of the arithmetic: the counts, the stencil handles of the K rows
    // every cell has, and every slot of the [MF]-wide tables (in bounds whatever the face
    // count; loads behind count-dependent control flow would each wait a memory latency).
    constexpr int KA = KMAX > 0 ? KMAX : 8;
    int h[KA];  // stencil handles (KMAX > 0) and the stencil and ghost counts, in one batch
    int nn, ng;
    a.template stencil_counts<KA>(m, i, h, nn, ng);
    const int nf = m.cell_nfaces[i];
    int codes[MF], nbh[MF];
#pragma unroll
    for (int s = 0; s < MF; s++) codes[s] = a.slot_code(p.nbr, C, s);
#pragma unroll
    for (int s = 0; s < MF; s++) nbh[s] = codes[s] >> 3;  // the cell across, in the accessor's numbering
    const double4 si = a.state(a.self());
    const double hi = si.x, ei = si.x + si.w, ui = si.y, vi = si.z;
    const double2 ci = a.center(a.self());
    const double cx = ci.x, cy = ci.y;
    float ixx = 0.f, iyy = 0.f, ixy = 0.f, je0 = 0.f, je1 = 0.f, ju0 = 0.f, ju1 = 0.f, jv0 = 0.f, jv1 = 0.f;
    auto accum = [&](const double4 sc, const double2 cc) {
        const float jx = (float)(cc.x - cx);
        const float jy = (float)(cc.y - cy);
        accum_grad(jx, jy, (float)((sc.x + sc.w) - ei), (float)(sc.y - ui), (float)(sc.z - vi), ixx, iyy, ixy, je0,
                   je1, ju0, ju1, jv0, jv1);
    };
    if (KMAX > 0) {

#pragma unroll
        for (int j = 0; j < KMAX; j++) h[j] = (j < nn) ? h[j] : a.self();
        constexpr int CH = Cells::kChunk;
        // Every group's entries are loaded (entries past nn read the own cell, a valid handle)
        // and only the accumulation depends on nn: conditional loads made the compiler copy
        // registers for the entries they may skip. Groups past the wave's largest stencil are
        // skipped; the scheduling barrier keeps the next group's loads from being hoisted.
        const int nn_max = wave_max_int(nn);
#pragma unroll
        for (int j0 = 0; j0 < KMAX; j0 += CH) {
            if (j0 >= nn_max) break;
            double4 sc[CH];
            double2 cc[CH];
#pragma unroll
            for (int k = 0; k < CH; k++) {
                sc[k] = a.state(h[j0 + k]);
                cc[k] = a.center(h[j0 + k]);
            }
#pragma unroll
            for (int k = 0; k < CH; k++)
                if (j0 + k < nn) accum(sc[k], cc[k]);
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        for (int j = 0; j < nn; j++) {
            const int c = a.stencil(j);
            accum(a.state(c), a.center(c));
        }
    }
    float2 gb[MF];  // face centre - cell centre of each slot
    float dzo[MF];  // z_f - Z of each slot
#pragma unroll
    for (int s = 0; s < MF; s++) {
        const float4 q = reinterpret_cast<const float4 *>(p.geoB)[MF * (size_t)i + s];
        const float4 ga = reinterpret_cast<const float4 *>(p.geoA)[MF * (size_t)i + s];
        gb[s] = make_float2(q.x, q.y);
        dzo[s] = ga.w;
        if (keep) {  // the slot geometry of the flux step, for a caller that keeps it
            keep->nx[s] = ga.x;
            keep->ny[s] = ga.y;
            keep->dz[s] = ga.z;
            keep->dzo[s] = ga.w;
            keep->len[s] = q.z;
            keep->ilen[s] = q.w;
        }
    }
    if (keep) {  // the codes (position << 3) | (k << 1) | side in the accessor's numbering
        keep->lc0 = codes[0];
        keep->lc1 = codes[1];
        keep->lc2 = codes[2];
        keep->lc3 = codes[MF - 1];
        keep->nf = nf;
    }
    for (int j = 0; j < ng; j++) {
        const int f = m.lsq_ghosts[(size_t)j * C + i];
        const double4 so = a.state(a.ghost_owner(j, f));
        const double *gc = m.face_ghost + 3 * (size_t)f;
        // wall ghost: eta and Z mirrored as Neumann, momenta as nonslip (-owner)
        accum_grad((float)(gc[0] - cx), (float)(gc[1] - cy), (float)((so.x + so.w) - ei), (float)(-so.y - ui),
                   (float)(-so.z - vi), ixx, iyy, ixy, je0, je1, ju0, ju1, jv0, jv1);
    }
    // face neighbours' states (the own state on a wall or a missing slot: not used)
    double4 sn[MF];
#pragma unroll
    for (int s = 0; s < MF; s++) sn[s] = a.state(codes[s] >= 0 ? nbh[s] : a.self());
    const float idia = 1.f / (ixx * iyy - ixy * ixy);
    const float ex = (iyy * je0 - ixy * je1) * idia, ey = (ixx * je1 - ixy * je0) * idia;
    const float ux = (iyy * ju0 - ixy * ju1) * idia, uy = (ixx * ju1 - ixy * ju0) * idia;
    const float vx = (iyy * jv0 - ixy * jv1) * idia, vy = (ixx * jv1 - ixy * jv0) * idia;
    float emax = 0.f, emin = 0.f, umax = 0.f, umin = 0.f, vmax = 0.f, vmin = 0.f;  // bounds as deviations
    float dh[MF];
#pragma unroll
    for (int s = 0; s < MF; s++) {
        dh[s] = 0.f;
        if (s >= nf) continue;
        const int code = codes[s];
        float eb, ub, vb;
        if (code >= 0) {
            const double4 sb = sn[s];
            eb = (float)((sb.x + sb.w) - ei);
            ub = (float)(sb.y - ui);
            vb = (float)(sb.z - vi);
            if (code & 1) dh[s] = (float)(hi - sb.x);  // own cell is the face's R: deviation from the L depth
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
#pragma unroll
    for (int s = 0; s < MF; s++) {
        if (s >= nf) continue;
        const float rx = gb[s].x, ry = gb[s].y;
        float d = ex * rx + ey * ry;
        if (!(fabsf(d) < 1e-8f)) pe = fminf(pe, (d > 0.f ? emax : emin) / d);
        d = ux * rx + uy * ry;
        if (!(fabsf(d) < 1e-8f)) pu = fminf(pu, (d > 0.f ? umax : umin) / d);
        d = vx * rx + vy * ry;
        if (!(fabsf(d) < 1e-8f)) pv = fminf(pv, (d > 0.f ? vmax : vmin) / d);
    }
    const float hu32 = (float)ui, hv32 = (float)vi, h32 = (float)hi;
#pragma unroll
    for (int s = 0; s < MF; s++) {
        if (s >= nf) break;
        const float rx = gb[s].x, ry = gb[s].y;
        const float de = om * pe * (ex * rx + ey * ry);
        const bool right = codes[s] >= 0 && (codes[s] & 1);
        float4 o;
        o.x = right ? dh[s] + (de - dzo[s]) : de - dzo[s];
        o.y = hu32 + om * pu * (ux * rx + uy * ry);
        o.z = hv32 + om * pv * (vx * rx + vy * ry);
        o.w = right ? dh[s] : h32;
        out[s] = o;
    }
    return nf;
}

template <int MF, int KMAX>
__device__ __forceinline__ int recon32_cell(const Mesh &m, const mfv_sw_mixed &p, const double4 *__restrict__ st,
                                            int i, float om, float4 *out)
{
    return recon32_core<MF, KMAX>(m, p, GlobalCells{m, st, i}, i, om, out);
}

// Stencil width a reconstruction kernel is compiled for: 16 covers the vertex stencils of
// triangle and quadrilateral meshes; wider ones take the plain loop.
inline int kmax_for(int K) { return K <= 16 ? 16 : 0; }

template <int MF, int BS, int KMAX>
__global__ void __launch_bounds__(BS) k_sw_recon32(Mesh m, mfv_sw_mixed p, float om)
{
    const int i = xcd_block() * BS + threadIdx.x;
    if (i >= m.C) return;
    float4 rec[MF];
    const int nf = recon32_cell<MF, KMAX>(m, p, reinterpret_cast<const double4 *>(p.state), i, om, rec);
    float4 *o = reinterpret_cast<float4 *>(p.facerec) + MF * (size_t)i;
#pragma unroll
    for (int s = 0; s < MF; s++)
        if (s < nf) o[s] = rec[s];
}

// time_step_SW of a float64 state with the float32 face normals (mesure = |n|), faces in
// faceid order; (gx[s], gy[s]) = the normal of slot s (geoA), gm[s] = its length (geoB).
template <int MF>
__device__ __forceinline__ double sw_dt32(const slot_f gx, const slot_f gy, const slot_f gm, int nf, double hn,
                                          double un, double vn, double vol, float g, double cfl)
{
    const double velson = sqrt((double)g * hn);
    const double ih = 1. / hn;
    double lam = 0.;
#pragma unroll
    for (int s = 0; s < MF; s++) {
        if (s >= nf) break;
        const double n0 = gx[s], n1 = gy[s], mesu = (double)gm[s];
        // (|u.n| / |n| + c) |n| of time_step_SW, with one division per cell
        lam += fabs((un * n0 + vn * n1) * ih) + velson * mesu;
    }
    return cfl * vol / lam;
}

// Face records of the flux gather: rec(k) is record k (= MF * cell + slot) of the array the
// neighbour codes index: the global facerec, or a block's LDS copy with block-local codes.
struct GlobalRecs {
    const float4 *__restrict__ fr;
    __device__ __forceinline__ float4 operator()(size_t k) const { return fr[k]; }
};

// One cell of the step: compensated float32 residual, float64 update into *out, next-dt
// candidate. own: this cell's index in the record array; code_of(s): neighbour code of slot s
// in the record array's numbering. The slot data of the [MF]-wide tables is loaded in one
// batch before the arithmetic (in bounds whatever nf).
// G: the cell's slot geometry (load_slot_geo, or kept from the reconstruction).
template <int MF, class Recs, class Codes>
__device__ __forceinline__ double flux32_core(const mfv_sw_mixed &p, int i, int own, const double4 si, int nf,
                                              unsigned ord, Recs rec, Codes code_of, const SlotGeo &G, float g,
                                              double cfl, double dt, double4 *out)
{
    // slots and codes in the residual order (slot s of the r-th contribution in bits 2r..2r+1)
    int sl[MF], cd[MF];
#pragma unroll
    for (int r = 0; r < MF; r++) {
        const int s = min((int)((ord >> (2 * r)) & 3), MF - 1);
        sl[r] = s;
        cd[r] = code_of(s);
    }
    const double vol = p.vol[i];
    double r0 = 0., r1 = 0., r2 = 0., s1 = 0., s2 = 0.;
#pragma unroll
    for (int r = 0; r < MF; r++) {
        if (r >= nf) break;
        const int s = sl[r];
        const float4 ga = make_float4(G.nx[s], G.ny[s], G.dz[s], G.dzo[s]);
        const float mes = G.len[s], im = G.ilen[s];
        const int code = cd[r];
        // One evaluation in the face's orientation whichever side the cell is on (a wave holds
        // cells of both sides, so per-side branches would each run for the whole wave): the L
        // and R records are read at selected indices. On a wall the other side is the own
        // record with the momenta negated (-1 * x) and dz = 0 (geoA holds 0 there). The signs
        // of the contributions are exact multiplications by -1 or 1.
        const bool wall = code < 0, own_R = !wall && (code & 1);
        const size_t ko = MF * (size_t)own + s;
        const size_t kn = wall ? ko : MF * (size_t)(code >> 3) + ((code >> 1) & 3);
        const float4 L = rec(own_R ? kn : ko), R = rec(own_R ? ko : kn);
        const float mR = wall ? -1.f : 1.f;
        float F0, F1, F2, bo;
        face_dev(L.w, L.x, R.x, ga.z, L.y, L.z, mR * R.y, mR * R.z, ga.w, ga.x, ga.y, mes, im, g,
                 own_R ? -R.w : 0.f, !own_R, F0, F1, F2, bo);
        const float sg = own_R ? 1.f : -1.f;
        r0 += (double)(sg * F0);
        r1 += (double)(sg * F1);
        r2 += (double)(sg * F2);
        s1 += (double)(sg * (bo * ga.x));
        s2 += (double)(sg * (bo * ga.y));
    }
    const double R0 = r0;
    const double R1 = r1 + s1;
    const double R2 = r2 + s2;
    const double dv = dt / vol;
    const double hn = si.x + R0 * dv;
    const double un = si.y + R1 * dv;
    const double vn = si.z + R2 * dv;
    *out = make_double4(hn, un, vn, si.w);
    return sw_dt32<MF>(G.nx, G.ny, G.len, nf, hn, un, vn, vol, g, cfl);
}

// The cell step on the global records, in place on p.state.
template <int MF>
__device__ __forceinline__ double sw_flux32_cell(const Mesh &m, const mfv_sw_mixed &p, float g, double cfl, int i,
                                                 double dt)
{
    const size_t C = m.C;
    double4 *st = reinterpret_cast<double4 *>(p.state);
    const int32_t *__restrict__ nbr = p.nbr;
    auto code_of = [nbr, C, i](int s) { return nbr[(size_t)s * C + i]; };
    return flux32_core<MF>(p, i, i, st[i], m.cell_nfaces[i], p.order[i],
                           GlobalRecs{reinterpret_cast<const float4 *>(p.facerec)}, code_of,
                           load_slot_geo<MF, false>(p, i), g, cfl, dt, st + i);
}

template <int MF>
__global__ void __launch_bounds__(kBlock) k_sw_flux32(Mesh m, mfv_sw_mixed p, float g, double cfl, DtRing ring,
                                                       double *__restrict__ dtpart)
{
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    ring_advance(ring);
    double dloc = 1e300;
    if (i < m.C) dloc = sw_flux32_cell<MF>(m, p, g, cfl, i, ring.cur[0]);
    wave_min_part(dtpart, (size_t)i, dloc);
}

__global__ void __launch_bounds__(kBlock) k_sw_unpack32(int C, const double *__restrict__ state,
                                                         double *__restrict__ h, double *__restrict__ hu,
                                                         double *__restrict__ hv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const double4 s = reinterpret_cast<const double4 *>(state)[i];
    h[i] = s.x;
    hu[i] = s.y;
    hv[i] = s.z;
}

// Deterministic float64 sums of (h, hu, hv) * vol: fixed-shape block tree, then a single-block
// pass over the block partials in index order.
constexpr int kSumBlock = 256;

__device__ __forceinline__ void block_tree3(double (*sh)[kSumBlock])
{
    __syncthreads();
    for (int w = kSumBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kSumBlock) k_sw_sums_partial(int C, const double *__restrict__ state,
                                                               const double *__restrict__ vol,
                                                               double *__restrict__ partial)
{
    __shared__ double sh[3][kSumBlock];
    const int i = blockIdx.x * kSumBlock + threadIdx.x;
    double a = 0., b = 0., c = 0.;
    if (i < C) {
        const double4 s = reinterpret_cast<const double4 *>(state)[i];
        a = s.x * vol[i];
        b = s.y * vol[i];
        c = s.z * vol[i];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    sh[2][threadIdx.x] = c;
    block_tree3(sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) partial[3 * (size_t)blockIdx.x + k] = sh[k][0];
}

__global__ void __launch_bounds__(kSumBlock) k_sw_sums_final(int nb, const double *__restrict__ partial,
                                                             double *__restrict__ out)
{
    __shared__ double sh[3][kSumBlock];
    double a = 0., b = 0., c = 0.;
    for (int j = threadIdx.x; j < nb; j += kSumBlock) {
        a += partial[3 * (size_t)j];
        b += partial[3 * (size_t)j + 1];
        c += partial[3 * (size_t)j + 2];
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    sh[2][threadIdx.x] = c;
    block_tree3(sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[k] = sh[k][0];
}

// ---------------------------------------------------------------------------------------------
// Ensembles: many independent instances concatenated into one mesh, each instance's cells
// starting at a multiple of 64 so that every wavefront belongs to one instance. The step
// kernels are the ones above; only the dt ring is per instance.

struct BatchK {
    int I, W;
    const int32_t *__restrict__ wave_inst;
    const int32_t *__restrict__ inst_cell0;
    const int32_t *__restrict__ inst_ncells;
    double *__restrict__ rings;
};

// Each wave stores its dt minimum at part[global wave] (wave w covers cells 64w..64w+63);
// k_batch_dt_fold then reduces the waves of each instance.
template <int MF, int BS>
__global__ void __launch_bounds__(BS) k_sw_flux32_batch(Mesh m, mfv_sw_mixed p, float g, double cfl, BatchK b,
                                                        int kcur, double *__restrict__ part)
{
    const int i = xcd_block() * BS + threadIdx.x;
    // C is a multiple of 64: all lanes of a wave agree on i < C and on the instance
    if (i >= m.C) return;
    const int inst = b.wave_inst[i >> 6];
    if (inst < 0) return;
    const double dloc = sw_flux32_cell<MF>(m, p, g, cfl, i, b.rings[4 * (size_t)inst + kcur]);
    wave_min_part(part, (size_t)i, dloc);
}

// One wave per instance: ring[knext] = min(ring[knext], min over the instance's waves).
// With advance != 0 it also closes the step: ring[kspare] = 1e6, ring[3] += ring[kcur].
__global__ void __launch_bounds__(256) k_batch_dt_fold(const double *__restrict__ part, BatchK b, int kcur,
                                                       int knext, int kspare, int advance)
{
    const int inst = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (inst >= b.I) return;
    const int w0 = b.inst_cell0[inst] >> 6;
    const int w1 = (b.inst_cell0[inst] + b.inst_ncells[inst] + 63) >> 6;
    // four partials in flight per lane (the minimum of a set does not depend on the order)
    double m0 = __longlong_as_double(0x7ff0000000000000LL), m1 = m0, m2 = m0, m3 = m0;
    int w = w0 + lane;
    for (; w + 192 < w1; w += 256) {
        const double a = part[w], c = part[w + 64], e = part[w + 128], f = part[w + 192];
        m0 = (a < m0) ? a : m0;
        m1 = (c < m1) ? c : m1;
        m2 = (e < m2) ? e : m2;
        m3 = (f < m3) ? f : m3;
    }
    for (; w < w1; w += 64) {
        const double v = part[w];
        m0 = (v < m0) ? v : m0;
    }
    m0 = (m1 < m0) ? m1 : m0;
    m2 = (m3 < m2) ? m3 : m2;
    double m = (m2 < m0) ? m2 : m0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o < m) ? o : m;
    }
    if (lane == 0) {
        double *ring = b.rings + 4 * (size_t)inst;
        if (m < ring[knext]) ring[knext] = m;
        if (advance) {
            ring[kspare] = 1e6;
            ring[3] += ring[kcur];
        }
    }
}

__global__ void k_ring_reset_batch(double *__restrict__ rings, int I)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= I) return;
    rings[4 * (size_t)t + 0] = 1e6;
    rings[4 * (size_t)t + 1] = 1e6;
    rings[4 * (size_t)t + 2] = 1e6;
    rings[4 * (size_t)t + 3] = 0.;
}

// dt of the current state into ring slot 0 (after k_ring_reset_batch)
template <int MF>
__global__ void __launch_bounds__(kBlock) k_sw_dt_batch(Mesh m, mfv_sw_mixed p, float g, double cfl, BatchK b,
                                                        double *__restrict__ part)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m.C) return;
    const int inst = b.wave_inst[i >> 6];
    if (inst < 0) return;
    const double4 s = reinterpret_cast<const double4 *>(p.state)[i];
    slot_f gx = 0.f, gy = 0.f, gm = 1.f;
#pragma unroll
    for (int k = 0; k < MF; k++) {
        const float4 q = reinterpret_cast<const float4 *>(p.geoA)[MF * (size_t)i + k];
        gx[k] = q.x;
        gy[k] = q.y;
        gm[k] = reinterpret_cast<const float4 *>(p.geoB)[MF * (size_t)i + k].z;
    }
    const double dloc = sw_dt32<MF>(gx, gy, gm, m.cell_nfaces[i], s.x, s.y, s.z, p.vol[i], g, cfl);
    wave_min_part(part, (size_t)i, dloc);
}

// Per-instance float64 totals sum (h, hu, hv) vol: one block per instance, fixed order.
__global__ void __launch_bounds__(kSumBlock) k_sw_sums_batch(const double *__restrict__ state,
                                                             const double *__restrict__ vol, BatchK b,
                                                             double *__restrict__ out)
{
    __shared__ double sh[3][kSumBlock];
    const int inst = blockIdx.x;
    const int c0 = b.inst_cell0[inst], n = b.inst_ncells[inst];
    double a = 0., bb = 0., c = 0.;
    for (int j = threadIdx.x; j < n; j += kSumBlock) {
        const double4 s = reinterpret_cast<const double4 *>(state)[c0 + j];
        const double v = vol[c0 + j];
        a += s.x * v;
        bb += s.y * v;
        c += s.z * v;
    }
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = bb;
    sh[2][threadIdx.x] = c;
    block_tree3(sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[3 * (size_t)inst + k] = sh[k][0];
}

inline unsigned fold_grid(int I) { return (unsigned)((I + 3) / 4); }  // 4 instances per 256-thread block

template <int MF, int BS>
void launch_batch_step(const Mesh &m, const mfv_sw_mixed &p, const BatchK &b, float g, double cfl, float om,
                       int step, double *part, hipStream_t st)
{
    const int k = ((step % 3) + 3) % 3;
    const unsigned grid = (unsigned)((m.C + BS - 1) / BS);
    if (kmax_for(m.K) == 16)
        hipLaunchKernelGGL((k_sw_recon32<MF, BS, 16>), dim3(grid), dim3(BS), 0, st, m, p, om);
    else
        hipLaunchKernelGGL((k_sw_recon32<MF, BS, 0>), dim3(grid), dim3(BS), 0, st, m, p, om);
    hipLaunchKernelGGL((k_sw_flux32_batch<MF, BS>), dim3(grid), dim3(BS), 0, st, m, p, g, cfl, b, k, part);
    hipLaunchKernelGGL(k_batch_dt_fold, dim3(fold_grid(b.I)), dim3(256), 0, st, part, b, k, (k + 1) % 3,
                       (k + 2) % 3, 1);
}

template <int MF>
int batch_step_bs(const Mesh &m, const mfv_sw_mixed &p, const BatchK &b, float g, double cfl, float om, int step,
                  int block, double *part, hipStream_t st)
{
    switch (block) {
    case 64: launch_batch_step<MF, 64>(m, p, b, g, cfl, om, step, part, st); break;
    case 128: launch_batch_step<MF, 128>(m, p, b, g, cfl, om, step, part, st); break;
    case 256: launch_batch_step<MF, 256>(m, p, b, g, cfl, om, step, part, st); break;
    case 512: launch_batch_step<MF, 512>(m, p, b, g, cfl, om, step, part, st); break;
    default: set_error("block size %d not in {64, 128, 256, 512}", block); return MFV_EINVAL;
    }
    return check_launch("ensemble step");
}

int check_mixed(const mfv_mesh *m, const mfv_sw_mixed *p)
{
    MFV_REQUIRE(p != nullptr, "mixed descriptor is NULL");
    MFV_REQUIRE(p->nb_cells == m->nb_cells && p->max_cell_faces == m->max_cell_faces,
                "mixed descriptor does not match the mesh");
    MFV_REQUIRE(m->max_cell_faces == 3 || m->max_cell_faces == 4, "mixed path supports 3 or 4 faces per cell");
    MFV_REQUIRE(p->state && p->vol && p->facerec && p->geoA && p->geoB && p->nbr && p->order,
                "mixed descriptor has NULL arrays");
    return MFV_OK;
}

int check_batch(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b)
{
    if (int rc = check_mixed(m, p)) return rc;
    MFV_REQUIRE(b != nullptr, "batch descriptor is NULL");
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(m->nb_cells % 64 == 0, "batched mesh must have a multiple of 64 cells, got %d", m->nb_cells);
    MFV_REQUIRE(b->nb_waves == m->nb_cells / 64, "nb_waves %d != cells/64 %d", b->nb_waves, m->nb_cells / 64);
    MFV_REQUIRE(b->nb_inst >= 1, "nb_inst must be positive");
    MFV_REQUIRE(b->wave_inst && b->inst_cell0 && b->inst_ncells && b->rings, "batch descriptor has NULL arrays");
    return MFV_OK;
}

BatchK to_batch(const mfv_sw_batch *b)
{
    return BatchK{b->nb_inst, b->nb_waves, b->wave_inst, b->inst_cell0, b->inst_ncells, b->rings};
}

// ---------------------------------------------------------------------------------------------
// Fused step on tiles. A block takes one tile (a run of at most T cells) and first
// reconstructs the face records of its cells and of its halo (the face neighbours outside the
// tile) into LDS, then evaluates the fluxes and the update of its cells from LDS. The face
// records never go to HBM; the halo's reconstruction is computed by both tiles that need it.
// Every record is computed by recon32_cell and every update by flux32_core, the functions of
// the two-kernel path, so the bits are the same. The state is double-buffered (the step reads
// p.state and writes out) because neighbouring tiles read the old state while a tile writes
// the new one.

struct TilesK {
    int n, T, H, E, I;
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
};

// Reconstruction inputs from a block's LDS copy: handles are positions in the tile's ext
// list (its cells, its halo, then every other cell its reconstructions read).
struct StagedCells {
    // LDS latency is short: stencil groups of two hold fewer registers than groups of four
    static constexpr int kChunk = 2;
    const double4 *st;
    const double2 *cc;
    const uint16_t *__restrict__ lst;
    const uint16_t *__restrict__ lgh;
    const int32_t *__restrict__ lfn;
    size_t row, I;
    int q, kp;
    __device__ __forceinline__ int self() const { return q; }
    __device__ __forceinline__ double4 state(int h) const { return st[h]; }
    __device__ __forceinline__ double2 center(int h) const { return cc[h]; }
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

// STAGED = false: face records in LDS, cell states read from global memory.
// STAGED = true: first every cell state and centre the tile reads is copied into LDS (one
// gather per cell and tile instead of one per stencil entry), then both phases read LDS.
template <int MF, int T, bool BATCH, bool STAGED, int KMAX>
__global__ void __launch_bounds__(T) k_sw_step32_tiled(Mesh m, mfv_sw_mixed p, TilesK t, double4 *__restrict__ out,
                                                       float g, double cfl, float om, DtRing ring, BatchK b,
                                                       int kcur, double *__restrict__ part)
{
    extern __shared__ double4 lds4[];
    double4 *s_st = lds4;                                       // [E] when STAGED
    double2 *s_cc = reinterpret_cast<double2 *>(lds4 + (STAGED ? t.E : 0));
    float4 *lfr = reinterpret_cast<float4 *>(s_cc + (STAGED ? t.E : 0));  // [max items * MF]
    const int tile = xcd_block();
    const int c0 = t.cell0[tile], nc = t.ncells[tile], nh = t.nhalo[tile];
    const int e0 = t.eptr[tile], ne = t.eptr[tile + 1] - e0;
    const size_t r0 = t.iptr[tile];
    const double4 *__restrict__ st = reinterpret_cast<const double4 *>(p.state);
    if (!BATCH) ring_advance(ring);
    if (STAGED) {
        for (int e = threadIdx.x; e < nc + ne; e += 2 * T) {  // two cells per pass: one latency for both
            const int2 c = tile_ext_pair(t.ext, c0, nc, e0, ne, e, T);
            const int eb = (e + T < nc + ne) ? e + T : e;
            const double4 sa = st[c.x], sb = st[c.y];
            const double *ca = m.cell_center + 3 * (size_t)c.x, *cb = m.cell_center + 3 * (size_t)c.y;
            const double2 pa = make_double2(ca[0], ca[1]), pb = make_double2(cb[0], cb[1]);
            s_st[e] = sa;
            s_cc[e] = pa;
            s_st[eb] = sb;
            s_cc[eb] = pb;
        }
        __syncthreads();
    }
    // A tile cell is reconstructed and updated by the same thread (q = threadIdx.x is in the
    // first pass), which keeps the cell's slot geometry in registers for the flux phase.
    const int q = threadIdx.x;
    // the residual order, loaded here so that the flux phase starts with the slot loads
    const unsigned ord_q = (q < nc) ? p.order[c0 + q] : 0u;
    SlotGeo G{0.f, 0.f, 0.f, 0.f, 1.f, 1.f, -1, -1, -1, -1, 0};
    auto recon_item = [&](int qq, SlotGeo *keep) {
        const int c = (qq < nc) ? c0 + qq : t.ext[e0 + (qq - nc)];
        float4 rec[MF];
        int nf;
        if (STAGED) {
            const StagedCells acc{s_st, s_cc, t.lst, t.lgh, t.lfn, r0 + qq, (size_t)t.I, qq, MFV_SW_TILES_KP(m.K)};
            nf = recon32_core<MF, KMAX>(m, p, acc, c, om, rec, keep);
        } else {
            nf = recon32_core<MF, KMAX>(m, p, GlobalCells{m, st, c}, c, om, rec, keep);
        }
#pragma unroll
        for (int s = 0; s < MF; s++)
            if (s < nf) lfr[MF * qq + s] = rec[s];
    };
    if (q < nc + nh) recon_item(q, &G);
    for (int qq = q + T; qq < nc + nh; qq += T) recon_item(qq, nullptr);  // items beyond one per thread
    __syncthreads();
    if ((q & ~63) >= nc) return;  // the whole wave lies past the tile: it owns no partial
    double dloc = 1e300;
    if (q < nc) {
        const int i = c0 + q;
        const double dt = BATCH ? b.rings[4 * (size_t)b.wave_inst[i >> 6] + kcur] : ring.cur[0];
        const int32_t *__restrict__ lnbr = t.lnbr;
        const size_t C = m.C;
        // staged: the codes of t.lnbr (the tile cell's lfn row) and the face count kept from the
        // reconstruction of the same cell; unstaged: the reconstruction numbered the neighbours
        // globally, so the block-local codes are loaded
        const int l0 = G.lc0, l1 = G.lc1, l2 = G.lc2, l3 = G.lc3;
        auto code_of = [lnbr, C, i, l0, l1, l2, l3](int s) {
            if (STAGED) return s == 0 ? l0 : s == 1 ? l1 : s == 2 ? l2 : l3;
            return lnbr[(size_t)s * C + i];
        };
        auto recs = [lfr](size_t k) { return lfr[k]; };
        const double4 si = STAGED ? s_st[q] : st[i];
        const int nf = STAGED ? G.nf : m.cell_nfaces[i];
        dloc = flux32_core<MF>(p, i, q, si, nf, ord_q, recs, code_of, G, g, cfl, dt, out + i);
    }
    wave_min_part(part, (size_t)c0 + q, dloc);
}

size_t tiles_lds_bytes(const mfv_sw_tiles *t, int MF)
{
    const size_t rec = (size_t)t->max_items * MF * sizeof(float4);
    return t->stage_states ? rec + (size_t)t->max_ext * (sizeof(double4) + sizeof(double2)) : rec;
}

int check_tiles(const mfv_mesh *m, const mfv_sw_tiles *t)
{
    MFV_REQUIRE(t != nullptr, "tiles descriptor is NULL");
    MFV_REQUIRE(t->tile_size == 128 || t->tile_size == 256 || t->tile_size == 512,
                "tile_size %d not in {128, 256, 512}", t->tile_size);
    MFV_REQUIRE(t->nb_tiles >= 1 && t->max_halo >= 0 && t->max_ext >= 0 && t->nb_items >= 1 &&
                    t->max_items >= 1,
                "bad tile counts");
    MFV_REQUIRE(t->tile_cell0 && t->tile_ncells && t->tile_nhalo && t->ext_ptr && t->ext_cells && t->item_ptr &&
                    t->lnbr && t->lst && t->lgh && t->lfn,
                "tiles descriptor has NULL arrays");
    const size_t lds = tiles_lds_bytes(t, m->max_cell_faces);
    MFV_REQUIRE(lds <= 64 * 1024, "tiles of %d cells need %zu bytes of LDS (> 64 KiB): renumber the mesh, use "
                "smaller tiles or do not stage the states", t->tile_size, lds);
    return MFV_OK;
}

TilesK to_tiles(const mfv_sw_tiles *t)
{
    return TilesK{t->nb_tiles, t->tile_size, t->max_halo, t->max_ext, t->nb_items, t->tile_cell0,
                  t->tile_ncells, t->tile_nhalo, t->ext_ptr, t->ext_cells, t->item_ptr, t->lnbr,
                  t->lst, t->lgh, t->lfn};
}

template <int MF, int T, bool BATCH, bool STAGED>
void launch_tiled(const Mesh &mk, const mfv_sw_mixed &p, const TilesK &tk, size_t lds, double *out, float g,
                  double cfl, float om, DtRing r, const BatchK &bk, int kcur, double *part, hipStream_t st)
{
    if (kmax_for(mk.K) == 16)
        hipLaunchKernelGGL((k_sw_step32_tiled<MF, T, BATCH, STAGED, 16>), dim3(tk.n), dim3(T), lds, st, mk, p, tk,
                           reinterpret_cast<double4 *>(out), g, cfl, om, r, bk, kcur, part);
    else
        hipLaunchKernelGGL((k_sw_step32_tiled<MF, T, BATCH, STAGED, 0>), dim3(tk.n), dim3(T), lds, st, mk, p, tk,
                           reinterpret_cast<double4 *>(out), g, cfl, om, r, bk, kcur, part);
}

template <int MF, bool BATCH, bool STAGED>
void launch_tiled_t(const Mesh &mk, const mfv_sw_mixed &p, const TilesK &tk, size_t lds, double *out, float g,
                    double cfl, float om, DtRing r, const BatchK &bk, int kcur, double *part, hipStream_t st)
{
    switch (tk.T) {
    case 128: launch_tiled<MF, 128, BATCH, STAGED>(mk, p, tk, lds, out, g, cfl, om, r, bk, kcur, part, st); break;
    case 256: launch_tiled<MF, 256, BATCH, STAGED>(mk, p, tk, lds, out, g, cfl, om, r, bk, kcur, part, st); break;
    default: launch_tiled<MF, 512, BATCH, STAGED>(mk, p, tk, lds, out, g, cfl, om, r, bk, kcur, part, st); break;
    }
}

template <bool BATCH>
int tiled_step(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_tiles *t, double *out, float g, double cfl,
               float om, DtRing r, const BatchK &bk, int kcur, double *part, hipStream_t st)
{
    const Mesh mk = to_kernel(m);
    const TilesK tk = to_tiles(t);
    const size_t lds = tiles_lds_bytes(t, m->max_cell_faces);
    const bool four = m->max_cell_faces == 4;
    if (t->stage_states) {
        if (four) launch_tiled_t<4, BATCH, true>(mk, *p, tk, lds, out, g, cfl, om, r, bk, kcur, part, st);
        else launch_tiled_t<3, BATCH, true>(mk, *p, tk, lds, out, g, cfl, om, r, bk, kcur, part, st);
    } else {
        if (four) launch_tiled_t<4, BATCH, false>(mk, *p, tk, lds, out, g, cfl, om, r, bk, kcur, part, st);
        else launch_tiled_t<3, BATCH, false>(mk, *p, tk, lds, out, g, cfl, om, r, bk, kcur, part, st);
    }
    return check_launch("k_sw_step32_tiled");
}

}  // namespace

extern "C" int mfv_sw_pack_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, const double *h, const double *hu,
                                 const double *hv, const double *Z, const double *zgrad, int order, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_mixed(m, p)) return rc;
    MFV_REQUIRE(m->dim == 2, "shallow water needs a 2D mesh");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(h && hu && hv && Z && zgrad, "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    auto k = (m->max_cell_faces == 3) ? k_sw_pack32<3> : k_sw_pack32<4>;
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), *p, h, hu,
                       hv, Z, zgrad, (double)(order - 1));
    return check_launch("k_sw_pack32");
}

extern "C" int mfv_sw_unpack_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, double *h, double *hu, double *hv,
                                   void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_mixed(m, p)) return rc;
    MFV_REQUIRE(h && hu && hv, "NULL array");
    if (m->nb_cells == 0) return MFV_OK;
    hipLaunchKernelGGL(k_sw_unpack32, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), m->nb_cells,
                       p->state, h, hu, hv);
    return check_launch("k_sw_unpack32");
}

extern "C" int mfv_sw_recon_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, int order, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_mixed(m, p)) return rc;
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    if (m->nb_cells == 0) return MFV_OK;
    const bool k16 = kmax_for(m->lsq_width) == 16;
    auto k = (m->max_cell_faces == 3) ? (k16 ? k_sw_recon32<3, kBlock, 16> : k_sw_recon32<3, kBlock, 0>)
                                      : (k16 ? k_sw_recon32<4, kBlock, 16> : k_sw_recon32<4, kBlock, 0>);
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), *p,
                       (float)(order - 1));
    return check_launch("k_sw_recon32");
}

extern "C" int mfv_sw_flux_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, double grav, double cfl,
                                 double *dt_ring_p, int step, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_mixed(m, p)) return rc;
    MFV_REQUIRE(dt_ring_p, "dt ring is NULL");
    if (m->nb_cells == 0) return MFV_OK;
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = waves_for(m->nb_cells);
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    auto k = (m->max_cell_faces == 3) ? k_sw_flux32<3> : k_sw_flux32<4>;
    hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream), to_kernel(m), *p,
                       (float)grav, cfl, r, part);
    if (int rc = check_launch("k_sw_flux32")) return rc;
    return dt_finish(part, nw, r.next, as_stream(stream));
}

extern "C" int mfv_sw_conservation(const mfv_mesh *m, const mfv_sw_mixed *p, double *partial, int n_partial,
                                   double *out3, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_mixed(m, p)) return rc;
    const int nb = (m->nb_cells + kSumBlock - 1) / kSumBlock;
    MFV_REQUIRE(partial && out3, "NULL array");
    MFV_REQUIRE(n_partial >= 3 * nb, "partial needs 3*ceil(C/256) = %d doubles, got %d", 3 * nb, n_partial);
    if (nb == 0) return MFV_OK;
    hipLaunchKernelGGL(k_sw_sums_partial, dim3(nb), dim3(kSumBlock), 0, as_stream(stream), m->nb_cells, p->state,
                       p->vol, partial);
    if (int rc = check_launch("k_sw_sums_partial")) return rc;
    hipLaunchKernelGGL(k_sw_sums_final, dim3(1), dim3(kSumBlock), 0, as_stream(stream), nb, partial, out3);
    return check_launch("k_sw_sums_final");
}

extern "C" int mfv_sw_batch_dt_init(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b, double grav,
                                    double cfl, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_batch(m, p, b)) return rc;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_ring_reset_batch, dim3((b->nb_inst + 255) / 256), dim3(256), 0, st, b->rings, b->nb_inst);
    if (m->nb_cells) {
        const size_t nw = waves_for(m->nb_cells);
        double *part = dt_scratch(nw, st);
        if (!part) return MFV_ENOMEM;
        auto k = (m->max_cell_faces == 3) ? k_sw_dt_batch<3> : k_sw_dt_batch<4>;
        hipLaunchKernelGGL(k, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, st, to_kernel(m), *p, (float)grav, cfl,
                           to_batch(b), part);
        hipLaunchKernelGGL(k_batch_dt_fold, dim3(fold_grid(b->nb_inst)), dim3(256), 0, st, part, to_batch(b), 0, 0,
                           0, 0);
    }
    return check_launch("k_sw_dt_batch");
}

extern "C" int mfv_sw_batch_step(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b, double grav,
                                 double cfl, int order, int step, int block, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_batch(m, p, b)) return rc;
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    if (m->nb_cells == 0) return MFV_OK;
    const Mesh mk = to_kernel(m);
    const BatchK bk = to_batch(b);
    const float om = (float)(order - 1);
    double *part = dt_scratch((size_t)m->nb_cells / 64, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    if (m->max_cell_faces == 3)
        return batch_step_bs<3>(mk, *p, bk, (float)grav, cfl, om, step, block, part, as_stream(stream));
    return batch_step_bs<4>(mk, *p, bk, (float)grav, cfl, om, step, block, part, as_stream(stream));
}

extern "C" int mfv_sw_batch_conservation(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b,
                                         double *out, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    if (int rc = check_batch(m, p, b)) return rc;
    MFV_REQUIRE(out, "NULL array");
    hipLaunchKernelGGL(k_sw_sums_batch, dim3(b->nb_inst), dim3(kSumBlock), 0, as_stream(stream), p->state, p->vol,
                       to_batch(b), out);
    return check_launch("k_sw_sums_batch");
}

extern "C" int mfv_sw_batch_graph(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b, double grav,
                                  double cfl, int order, int nsteps, int block, void **graph_out)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_batch(m, p, b)) return rc;
    MFV_REQUIRE(graph_out, "graph_out is NULL");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(nsteps >= 3 && nsteps % 3 == 0 && nsteps <= 3000, "nsteps %d must be a multiple of 3 in [3, 3000]",
                nsteps);
    MFV_REQUIRE(block == 64 || block == 128 || block == 256 || block == 512, "block size %d not in {64..512}",
                block);
    *graph_out = nullptr;
    MfvGraph *g = new MfvGraph();
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&g->part), ((size_t)m->nb_cells / 64 + 1) * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("graph capture setup: %s", hipGetErrorString(e));
        graph_free(g);
        return MFV_EHIP;
    }
    const Mesh mk = to_kernel(m);
    const BatchK bk = to_batch(b);
    int rc = MFV_OK;
    for (int k = 0; k < nsteps && rc == MFV_OK; k++) {
        rc = (m->max_cell_faces == 3)
                 ? batch_step_bs<3>(mk, *p, bk, (float)grav, cfl, (float)(order - 1), k, block, g->part, g->stream)
                 : batch_step_bs<4>(mk, *p, bk, (float)grav, cfl, (float)(order - 1), k, block, g->part, g->stream);
    }
    e = hipStreamEndCapture(g->stream, &g->graph);
    if (rc != MFV_OK || e != hipSuccess) {
        if (rc == MFV_OK) set_error("graph capture: %s", hipGetErrorString(e));
        graph_free(g);
        return rc != MFV_OK ? rc : MFV_EHIP;
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        set_error("graph instantiate: %s", hipGetErrorString(e));
        graph_free(g);
        return MFV_EHIP;
    }
    *graph_out = g;
    return MFV_OK;
}

extern "C" int mfv_graph_launch(void *graph, void *stream)
{
    MFV_REQUIRE(graph, "graph is NULL");
    hipError_t e = hipGraphLaunch(static_cast<MfvGraph *>(graph)->exec, as_stream(stream));
    if (e != hipSuccess) {
        set_error("hipGraphLaunch: %s", hipGetErrorString(e));
        return MFV_EHIP;
    }
    return MFV_OK;
}

extern "C" int mfv_graph_destroy(void *graph)
{
    graph_free(static_cast<MfvGraph *>(graph));
    return MFV_OK;
}

extern "C" int mfv_sw_step_mixed_tiled(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_tiles *t,
                                       double *state_out, double grav, double cfl, int order, double *dt_ring_p,
                                       int step, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_mixed(m, p)) return rc;
    if (int rc = check_tiles(m, t)) return rc;
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(dt_ring_p && state_out && state_out != p->state, "state_out must be a second state buffer");
    if (m->nb_cells == 0) return MFV_OK;
    DtRing r = dt_ring(dt_ring_p, step);
    const size_t nw = ((size_t)m->nb_cells + 63) / 64;
    double *part = dt_scratch(nw, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    const BatchK none{0, 0, nullptr, nullptr, nullptr, nullptr};
    if (int rc = tiled_step<false>(m, p, t, state_out, (float)grav, cfl, (float)(order - 1), r, none, 0, part,
                                   as_stream(stream)))
        return rc;
    return dt_finish(part, nw, r.next, as_stream(stream));
}

static int batch_tiled_launch(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b,
                              const mfv_sw_tiles *t, double *out, float g, double cfl, float om, int step,
                              double *part, hipStream_t st)
{
    const int k = ((step % 3) + 3) % 3;
    const BatchK bk = to_batch(b);
    const DtRing none{};
    if (int rc = tiled_step<true>(m, p, t, out, g, cfl, om, none, bk, k, part, st)) return rc;
    hipLaunchKernelGGL(k_batch_dt_fold, dim3(fold_grid(b->nb_inst)), dim3(256), 0, st, part, bk, k, (k + 1) % 3,
                       (k + 2) % 3, 1);
    return check_launch("k_batch_dt_fold");
}

extern "C" int mfv_sw_batch_step_tiled(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b,
                                       const mfv_sw_tiles *t, double *state_out, double grav, double cfl,
                                       int order, int step, void *stream)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_batch(m, p, b)) return rc;
    if (int rc = check_tiles(m, t)) return rc;
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(state_out && state_out != p->state, "state_out must be a second state buffer");
    if (m->nb_cells == 0) return MFV_OK;
    double *part = dt_scratch((size_t)m->nb_cells / 64, as_stream(stream));
    if (!part) return MFV_ENOMEM;
    return batch_tiled_launch(m, p, b, t, state_out, (float)grav, cfl, (float)(order - 1), step, part,
                              as_stream(stream));
}

extern "C" int mfv_sw_batch_graph_tiled(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b,
                                        const mfv_sw_tiles *t, double *state2, double grav, double cfl, int order,
                                        int nsteps, void **graph_out)
{
    if (int rc = validate_mesh(m, 1, 0, 0)) return rc;
    if (int rc = check_batch(m, p, b)) return rc;
    if (int rc = check_tiles(m, t)) return rc;
    MFV_REQUIRE(graph_out, "graph_out is NULL");
    MFV_REQUIRE(order == 1 || order == 2, "order must be 1 or 2");
    MFV_REQUIRE(state2 && state2 != p->state, "state2 must be a second state buffer");
    MFV_REQUIRE(nsteps >= 6 && nsteps % 6 == 0 && nsteps <= 3000, "nsteps %d must be a multiple of 6 in [6, 3000]",
                nsteps);
    *graph_out = nullptr;
    MfvGraph *g = new MfvGraph();
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&g->part), ((size_t)m->nb_cells / 64 + 1) * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        set_error("graph capture setup: %s", hipGetErrorString(e));
        graph_free(g);
        return MFV_EHIP;
    }
    // steps alternate the two state buffers; nsteps is even, so a replay ends in p->state
    mfv_sw_mixed pa = *p, pb = *p;
    pb.state = state2;
    int rc = MFV_OK;
    for (int k = 0; k < nsteps && rc == MFV_OK; k++) {
        const mfv_sw_mixed &pin = (k % 2 == 0) ? pa : pb;
        double *out = (k % 2 == 0) ? state2 : p->state;
        rc = batch_tiled_launch(m, &pin, b, t, out, (float)grav, cfl, (float)(order - 1), k, g->part, g->stream);
    }
    e = hipStreamEndCapture(g->stream, &g->graph);
    if (rc != MFV_OK || e != hipSuccess) {
        if (rc == MFV_OK) set_error("graph capture: %s", hipGetErrorString(e));
        graph_free(g);
        return rc != MFV_OK ? rc : MFV_EHIP;
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        set_error("graph instantiate: %s", hipGetErrorString(e));
        graph_free(g);
        return MFV_EHIP;
    }
    *graph_out = g;
    return MFV_OK;
}
