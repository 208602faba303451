// libmfv diamond scheme (2D and 3D): matrix assembly into ELL slices, Dirichlet lifting of
// the right-hand side, node interpolation, diamond face gradients, the dissipative residual
// and the diffusion / advection-diffusion time step.
//
// Every output element is produced by one thread that gathers its terms in the order the
// reference adds them (faces in the reference loop order, nodes in face order, cells and
// ghost centres in node-list order), so sums are bit-identical to the reference's
// sequential loops and do not depend on the launch shape. Nothing uses atomics.
#include "mfv_common.hpp"

namespace mfv {

struct Diamond {
    int C, F, N, MF, dim;
    const int32_t *__restrict__ cell_faces;
    const int32_t *__restrict__ cell_nfaces;
    const uint8_t *__restrict__ order_sorted;
    const double *__restrict__ vol;
    const int32_t *__restrict__ face_cells;
    const int32_t *__restrict__ face_nodes;
    const double *__restrict__ param;
    const double *__restrict__ fk;
    const double *__restrict__ air;
    const double *__restrict__ normal;
    const int32_t *__restrict__ nc_ptr;
    const int32_t *__restrict__ nc;
    const double *__restrict__ nc_alpha;
    const int32_t *__restrict__ ng_ptr;
    const int32_t *__restrict__ ng_face;
    const int32_t *__restrict__ ng_owner;
    const double *__restrict__ ng_alpha;
};

static int validate_diamond(const mfv_diamond *d)
{
    MFV_REQUIRE(d != nullptr, "diamond descriptor is NULL");
    MFV_REQUIRE(d->nb_cells >= 0 && d->nb_faces >= 0 && d->nb_nodes >= 0, "negative diamond sizes");
    MFV_REQUIRE(d->max_cell_faces >= 1 && d->max_cell_faces <= 8, "max_cell_faces %d out of [1, 8]",
                d->max_cell_faces);
    MFV_REQUIRE(d->dim == 2 || d->dim == 3, "diamond dim %d not in {2, 3}", d->dim);
    MFV_REQUIRE(d->cell_faces && d->cell_nfaces && d->order_sorted && d->cell_volume && d->face_cells &&
                    d->face_nodes && d->face_param && d->face_fk && d->face_air && d->face_normal &&
                    d->node_cell_ptr && d->node_cell && d->node_cell_alpha && d->node_ghost_ptr &&
                    d->node_ghost_face && d->node_ghost_owner && d->node_ghost_alpha,
                "diamond descriptor has NULL pointers");
    return MFV_OK;
}

static Diamond to_kernel(const mfv_diamond *d)
{
    Diamond k;
    k.C = d->nb_cells;
    k.F = d->nb_faces;
    k.N = d->nb_nodes;
    k.MF = d->max_cell_faces;
    k.dim = d->dim;
    k.cell_faces = d->cell_faces;
    k.cell_nfaces = d->cell_nfaces;
    k.order_sorted = d->order_sorted;
    k.vol = d->cell_volume;
    k.face_cells = d->face_cells;
    k.face_nodes = d->face_nodes;
    k.param = d->face_param;
    k.fk = d->face_fk;
    k.air = d->face_air;
    k.normal = d->face_normal;
    k.nc_ptr = d->node_cell_ptr;
    k.nc = d->node_cell;
    k.nc_alpha = d->node_cell_alpha;
    k.ng_ptr = d->node_ghost_ptr;
    k.ng_face = d->node_ghost_face;
    k.ng_owner = d->node_ghost_owner;
    k.ng_alpha = d->node_ghost_alpha;
    return k;
}

// Face of cell K at sorted position jj (ascending face id).
__device__ __forceinline__ int sorted_face(const Diamond &d, int K, int jj)
{
    int slot = d.order_sorted[(size_t)jj * d.C + K];
    return d.cell_faces[(size_t)slot * d.C + K];
}

// Dirichlet faces of K in the order of the variable's dirichletfaces list (rank >= 0);
// returns the face with the smallest rank above `after`, or -1.
__device__ __forceinline__ int next_dirichlet_face(const Diamond &d, int K, int nf,
                                                   const int32_t *__restrict__ face_dir_rank, int after,
                                                   int &rank_out)
{
    int best = -1, best_rank = 0x7fffffff;
    for (int j = 0; j < nf; j++) {
        int f = d.cell_faces[(size_t)j * d.C + K];
        int r = face_dir_rank[f];
        if (r > after && r < best_rank) {
            best_rank = r;
            best = f;
        }
    }
    rank_out = best_rank;
    return best;
}

// ---------------------------------------------------------------------------------------------
// matrix rows (get_triplet_2d, ast/functions2d.py:328-580; compute_2dmatrix_size :582-664)
//
// Row K collects, in the reference's insertion order, the entries of the matrix-inner faces
// of K (ascending face id: the sorted matrixinnerfaces loop) and then the entries of K's
// Dirichlet faces (dirichletfaces list order). Duplicate (row, col) entries are added in that
// order, which is what summing the reference's COO triplets in insertion order gives.

// unique columns per row (vertex-neighbour stencil + self): a triangle's stencil stays under
// 32; a tetrahedron's reaches about 110 on unstructured meshes
constexpr int kRowCap2 = 32;
constexpr int kRowCap3 = 128;

struct RowAcc {
    int32_t (*col)[64];
    double (*val)[64];
    int lane;
    int n;
    int cap;
    bool overflow;
    __device__ __forceinline__ void add(int c, double v)
    {
        for (int j = 0; j < n; j++)
            if (col[j][lane] == c) {
                val[j][lane] += v;
                return;
            }
        if (n < cap) {
            col[n][lane] = c;
            val[n][lane] = v;
            n++;
        } else {
            overflow = true;
        }
    }
};

__device__ void diamond_row2(const Diamond &d, int K, const uint8_t *__restrict__ node_dir,
                             const int32_t *__restrict__ face_dir_rank, RowAcc &acc)
{
    const int nf = d.cell_nfaces[K];
    const double volK = d.vol[K];
    for (int jj = 0; jj < nf; jj++) {
        const int f = sorted_face(d, K, jj);
        const int L = d.face_cells[2 * f];
        const int R = d.face_cells[2 * f + 1];
        if (R < 0) continue;
        const bool left = (L == K);
        const double p1 = d.param[4 * f + 0], p2 = d.param[4 * f + 1];
        const double p3 = d.param[4 * f + 2], p4 = d.param[4 * f + 3];
        if (left) acc.add(K, p1 / volK);
        // node 0 takes param4, node 1 param2 (parameters[cmptparam], cmptparam 0 then 1)
        for (int s = 0; s < 2; s++) {
            const int n = d.face_nodes[2 * f + s];
            if (node_dir[n]) continue;
            const double p = (s == 0) ? p4 : p2;
            for (int q = d.nc_ptr[n]; q < d.nc_ptr[n + 1]; q++) {
                const double a = d.nc_alpha[q];
                acc.add(d.nc[q], left ? (a / volK * p) : (-1. * a / volK * p));
            }
            for (int q = d.ng_ptr[n]; q < d.ng_ptr[n + 1]; q++) {
                const double a = d.ng_alpha[q];
                acc.add(d.ng_owner[q], left ? (a / volK * p) : (-1. * a / volK * p));
            }
        }
        if (left) {
            acc.add(R, p3 / volK);
        } else {
            acc.add(L, -1. * p1 / volK);
            acc.add(K, -1. * p3 / volK);
        }
    }
    int rank = -1;
    for (;;) {
        int nr;
        const int f = next_dirichlet_face(d, K, nf, face_dir_rank, rank, nr);
        if (f < 0) break;
        rank = nr;
        acc.add(K, d.param[4 * f + 0] / volK);
        acc.add(K, -1. * d.param[4 * f + 2] / volK);
    }
}

// get_triplet_3d (ast/functions3d.py:498-782), row K: for a matrix-inner face the left row
// takes (K,K) -param3/|K|, the node entries and (K,R) param3/|K|; the right row takes the node
// entries, (K,L) param3/|K| and (K,K) -param3/|K|. The face's nodes A, B, C, D (D = C on a
// triangle) carry param1, param2, -param1, -param2. A Dirichlet face adds -param3/|K| to the
// diagonal twice, as the reference does.
__device__ void diamond_row3(const Diamond &d, int K, const uint8_t *__restrict__ node_dir,
                             const int32_t *__restrict__ face_dir_rank, RowAcc &acc)
{
    const int nf = d.cell_nfaces[K];
    const double volK = d.vol[K];
    for (int jj = 0; jj < nf; jj++) {
        const int f = sorted_face(d, K, jj);
        const int L = d.face_cells[2 * f];
        const int R = d.face_cells[2 * f + 1];
        if (R < 0) continue;
        const bool left = (L == K);
        const double p1 = d.param[4 * f + 0], p2 = d.param[4 * f + 1], p3 = d.param[4 * f + 2];
        const double prm[4] = {p1, p2, -1. * p1, -1. * p2};
        if (left) acc.add(K, -1. * p3 / volK);
        for (int s = 0; s < 4; s++) {
            const int n = d.face_nodes[4 * f + s];
            if (node_dir[n]) continue;
            const double p = prm[s];
            for (int q = d.nc_ptr[n]; q < d.nc_ptr[n + 1]; q++) {
                const double a = d.nc_alpha[q];
                acc.add(d.nc[q], left ? (a / volK * p) : (-1. * a / volK * p));
            }
            for (int q = d.ng_ptr[n]; q < d.ng_ptr[n + 1]; q++) {
                const double a = d.ng_alpha[q];
                acc.add(d.ng_owner[q], left ? (a / volK * p) : (-1. * a / volK * p));
            }
        }
        if (left) {
            acc.add(R, p3 / volK);
        } else {
            acc.add(L, p3 / volK);
            acc.add(K, -1. * p3 / volK);
        }
    }
    int rank = -1;
    for (;;) {
        int nr;
        const int f = next_dirichlet_face(d, K, nf, face_dir_rank, rank, nr);
        if (f < 0) break;
        rank = nr;
        acc.add(K, -1. * d.param[4 * f + 2] / volK);
        acc.add(K, -1. * d.param[4 * f + 2] / volK);
    }
}

template <int DIM>
__device__ __forceinline__ void diamond_row(const Diamond &d, int K, const uint8_t *__restrict__ node_dir,
                                            const int32_t *__restrict__ face_dir_rank, RowAcc &acc)
{
    if (DIM == 2) diamond_row2(d, K, node_dir, face_dir_rank, acc);
    else diamond_row3(d, K, node_dir, face_dir_rank, acc);
}

// One wave = one 64-row slice. Writes the used width of each slice (max unique columns of
// its rows) and raises *overflow if a row exceeds CAP. The 3D rows' accumulators take 96 KiB
// of LDS (gfx950 gives a workgroup up to 160 KiB).
template <int DIM, int CAP>
__global__ void __launch_bounds__(64) k_diamond_width(Diamond d, const uint8_t *__restrict__ node_dir,
                                                      const int32_t *__restrict__ face_dir_rank,
                                                      int32_t *__restrict__ slice_w, int32_t *__restrict__ info)
{
    __shared__ int32_t s_col[CAP][64];
    __shared__ double s_val[CAP][64];
    const int lane = threadIdx.x;
    const int K = blockIdx.x * 64 + lane;
    RowAcc acc{s_col, s_val, lane, 0, CAP, false};
    if (K < d.C) diamond_row<DIM>(d, K, node_dir, face_dir_rank, acc);
    int w = acc.n;
    int ov = acc.overflow ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        w = max(w, __shfl_xor(w, off, 64));
        ov |= __shfl_xor(ov, off, 64);
    }
    if (lane == 0) {
        slice_w[blockIdx.x] = w;
        if (ov) atomicOr(&info[0], 1);
        atomicMax(&info[1], w);
    }
}

// Fills the ELL slots of the rows: entry j of local row K sits at ((K/64)*width + j)*64 + K%64,
// columns ascending, global column = row0 + local column; slots past the row's length (up to
// the slice width) hold (row0 + K, 0). Rows past C (padding of the last slice) are all padding
// with diag 0.
template <int DIM, int CAP>
__global__ void __launch_bounds__(64) k_diamond_fill(Diamond d, const uint8_t *__restrict__ node_dir,
                                                     const int32_t *__restrict__ face_dir_rank, int row0,
                                                     int width, const int32_t *__restrict__ slice_w,
                                                     int32_t *__restrict__ cols, double *__restrict__ vals,
                                                     float *__restrict__ vals32, double *__restrict__ diag)
{
    __shared__ int32_t s_col[CAP][64];
    __shared__ double s_val[CAP][64];
    const int lane = threadIdx.x;
    const int slice = blockIdx.x;
    const int K = slice * 64 + lane;
    RowAcc acc{s_col, s_val, lane, 0, CAP, false};
    if (K < d.C) diamond_row<DIM>(d, K, node_dir, face_dir_rank, acc);
    const int n = acc.n;
    // insertion sort of this lane's entries by column
    for (int i = 1; i < n; i++) {
        int c = s_col[i][lane];
        double v = s_val[i][lane];
        int j = i - 1;
        while (j >= 0 && s_col[j][lane] > c) {
            s_col[j + 1][lane] = s_col[j][lane];
            s_val[j + 1][lane] = s_val[j][lane];
            j--;
        }
        s_col[j + 1][lane] = c;
        s_val[j + 1][lane] = v;
    }
    double dg = 0.;
    const int w = slice_w[slice];
    const size_t base = (size_t)slice * width * 64 + lane;
    for (int j = 0; j < w; j++) {
        int c;
        double v;
        if (j < n) {
            c = s_col[j][lane];
            v = s_val[j][lane];
            if (c == K) dg = v;
        } else {
            c = K;
            v = 0.;
        }
        cols[base + (size_t)j * 64] = row0 + c;
        vals[base + (size_t)j * 64] = v;
        vals32[base + (size_t)j * 64] = (float)v;
    }
    diag[K] = dg;
}

// ---------------------------------------------------------------------------------------------
// right-hand side (get_rhs_glob_2d, ast/functions2d.py:740-817)

__global__ void __launch_bounds__(kBlock) k_diamond_rhs(Diamond d, const uint8_t *__restrict__ node_dir,
                                                         const int32_t *__restrict__ face_dir_rank,
                                                         const double *__restrict__ Pn,
                                                         const double *__restrict__ Pf, double *__restrict__ rhs)
{
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= d.C) return;
    const int nf = d.cell_nfaces[K];
    const double volK = d.vol[K];
    double r = 0.;
    for (int jj = 0; jj < nf; jj++) {
        const int f = sorted_face(d, K, jj);
        const int L = d.face_cells[2 * f];
        if (d.face_cells[2 * f + 1] < 0) continue;
        const bool left = (L == K);
        const int n1 = d.face_nodes[2 * f], n2 = d.face_nodes[2 * f + 1];
        const double p2 = d.param[4 * f + 1], p4 = d.param[4 * f + 3];
        if (node_dir[n1]) {
            const double V = Pn[n1];
            r += left ? (-1. * V * p4 / volK) : (V * p4 / volK);
        }
        if (node_dir[n2]) {
            const double V = Pn[n2];
            r += left ? (-1. * V * p2 / volK) : (V * p2 / volK);
        }
    }
    int rank = -1;
    for (;;) {
        int nr;
        const int f = next_dirichlet_face(d, K, nf, face_dir_rank, rank, nr);
        if (f < 0) break;
        rank = nr;
        const int n1 = d.face_nodes[2 * f], n2 = d.face_nodes[2 * f + 1];
        // centergn[node][0][2] != -1: the node has a ghost centre (a boundary node)
        if (d.ng_ptr[n1 + 1] > d.ng_ptr[n1]) r += -1. * Pn[n1] * d.param[4 * f + 3] / volK;
        if (d.ng_ptr[n2 + 1] > d.ng_ptr[n2]) r += -1. * Pn[n2] * d.param[4 * f + 1] / volK;
        r += -2. * d.param[4 * f + 2] / volK * Pf[f];
    }
    rhs[K] = r;
}

// get_rhs_glob_3d (ast/functions3d.py:865-946): the Dirichlet nodes of the matrix-inner faces
// (parameters param1, param2, -param1, -param2 on nodes A, B, C, D), then per Dirichlet face
// its boundary nodes (centergn[node][0][3] != -1) and -2 param3/|K| V_K.
__global__ void __launch_bounds__(kBlock) k_diamond_rhs3(Diamond d, const uint8_t *__restrict__ node_dir,
                                                          const int32_t *__restrict__ face_dir_rank,
                                                          const double *__restrict__ Pn,
                                                          const double *__restrict__ Pf, double *__restrict__ rhs)
{
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= d.C) return;
    const int nf = d.cell_nfaces[K];
    const double volK = d.vol[K];
    double r = 0.;
    for (int jj = 0; jj < nf; jj++) {
        const int f = sorted_face(d, K, jj);
        const int L = d.face_cells[2 * f];
        if (d.face_cells[2 * f + 1] < 0) continue;
        const bool left = (L == K);
        const double p1 = d.param[4 * f + 0], p2 = d.param[4 * f + 1];
        const double prm[4] = {p1, p2, -1. * p1, -1. * p2};
        for (int s = 0; s < 4; s++) {
            const int n = d.face_nodes[4 * f + s];
            if (node_dir[n]) {
                const double V = Pn[n];
                r += left ? (-1. * V * prm[s] / volK) : (V * prm[s] / volK);
            }
        }
    }
    int rank = -1;
    for (;;) {
        int nr;
        const int f = next_dirichlet_face(d, K, nf, face_dir_rank, rank, nr);
        if (f < 0) break;
        rank = nr;
        const double p1 = d.param[4 * f + 0], p2 = d.param[4 * f + 1];
        const double prm[4] = {p1, p2, -1. * p1, -1. * p2};
        for (int s = 0; s < 4; s++) {
            const int n = d.face_nodes[4 * f + s];
            if (d.ng_ptr[n + 1] > d.ng_ptr[n]) r += -1. * Pn[n] * prm[s] / volK;
        }
        r += -2. * d.param[4 * f + 2] / volK * Pf[f];
    }
    rhs[K] = r;
}

// ---------------------------------------------------------------------------------------------
// node values (centertovertex_2d, ast/functions2d.py:199-277)

__global__ void __launch_bounds__(kBlock) k_centertovertex(Diamond d, const double *__restrict__ w_c,
                                                            const double *__restrict__ w_ghost,
                                                            double *__restrict__ w_n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.N) return;
    double s = 0.;
    for (int q = d.nc_ptr[i]; q < d.nc_ptr[i + 1]; q++) s += d.nc_alpha[q] * w_c[d.nc[q]];
    for (int q = d.ng_ptr[i]; q < d.ng_ptr[i + 1]; q++) s += d.ng_alpha[q] * w_ghost[d.ng_face[q]];
    w_n[i] = s;
}

// ---------------------------------------------------------------------------------------------
// diamond face gradients: kind 0 = face_gradient_2d (ast/functions2d.py:108-196), kind 1 =
// compute_P_gradient_2d_diamond (:820-936). face_bc[f] on boundary faces: 0 = not listed
// (left untouched), 1 = listed Neumann face, 2 = listed Dirichlet face.

__global__ void __launch_bounds__(kBlock) k_face_gradient(Diamond d, int kind, const double *__restrict__ w_c,
                                                           const double *__restrict__ w_ghost,
                                                           const double *__restrict__ w_node,
                                                           const int8_t *__restrict__ face_bc,
                                                           const uint8_t *__restrict__ node_dir,
                                                           const double *__restrict__ Pn,
                                                           const double *__restrict__ Pf, double *__restrict__ wx,
                                                           double *__restrict__ wy)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= d.F) return;
    const int L = d.face_cells[2 * f];
    const int R = d.face_cells[2 * f + 1];
    const int bc = (R >= 0) ? 0 : face_bc[f];
    if (R < 0 && bc == 0) return;
    const int n1 = d.face_nodes[2 * f], n2 = d.face_nodes[2 * f + 1];
    double vi1 = w_node[n1];
    double vi2 = w_node[n2];
    const double vv1 = w_c[L];
    double vv2 = (R >= 0) ? w_c[R] : w_ghost[f];
    if (kind == 1) {
        if (bc == 2) {
            vi1 = Pn[n1];
            vi2 = Pn[n2];
            vv2 = 2. * Pf[f] - vv1;
        } else {
            if (node_dir[n1]) vi1 = Pn[n1];
            if (node_dir[n2]) vi2 = Pn[n2];
        }
    }
    const double *k = d.fk + 8 * f;
    const double sx = (vi1 + vv1) * k[1] + (vv1 + vi2) * k[3] + (vi2 + vv2) * k[5] + (vv2 + vi1) * k[7];
    const double sy = (vi1 + vv1) * k[0] + (vv1 + vi2) * k[2] + (vi2 + vv2) * k[4] + (vv2 + vi1) * k[6];
    const double a2 = 2 * d.air[f];
    if (kind == 0) {
        wx[f] = 1 / a2 * sx;
        wy[f] = -1 / a2 * sy;
    } else {
        wx[f] = -1 / a2 * sx;
        wy[f] = 1 / a2 * sy;
    }
}

// 3D: kind 0 = face_gradient_3d (ast/functions3d.py:122-248), kind 1 =
// compute_P_gradient_3d_diamond (:949-1108): (n1 (V_A - V_C) + n2 (V_B - V_D) + n (V_R - V_L))
// / airDiamond per component (negated for kind 1); face_fk holds n1 and n2 (faces.f_1, f_2).
__global__ void __launch_bounds__(kBlock) k_face_gradient3(Diamond d, int kind, const double *__restrict__ w_c,
                                                            const double *__restrict__ w_ghost,
                                                            const double *__restrict__ w_node,
                                                            const int8_t *__restrict__ face_bc,
                                                            const uint8_t *__restrict__ node_dir,
                                                            const double *__restrict__ Pn,
                                                            const double *__restrict__ Pf, double *__restrict__ wx,
                                                            double *__restrict__ wy, double *__restrict__ wz)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= d.F) return;
    const int L = d.face_cells[2 * f];
    const int R = d.face_cells[2 * f + 1];
    const int bc = (R >= 0) ? 0 : face_bc[f];
    if (R < 0 && bc == 0) return;
    int nd[4];
    double V[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        nd[s] = d.face_nodes[4 * f + s];
        V[s] = w_node[nd[s]];
    }
    const double VL = w_c[L];
    double VR = (R >= 0) ? w_c[R] : w_ghost[f];
    if (kind == 1) {
        if (bc == 2) {
#pragma unroll
            for (int s = 0; s < 4; s++) V[s] = Pn[nd[s]];
            VR = 2. * Pf[f] - VL;
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++)
                if (node_dir[nd[s]]) V[s] = Pn[nd[s]];
        }
    }
    const double *k = d.fk + 6 * f;
    const double *nr = d.normal + 3 * f;
    const double air = d.air[f];
    double g[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double sum = k[c] * (V[0] - V[2]) + k[3 + c] * (V[1] - V[3]) + nr[c] * (VR - VL);
        g[c] = (kind == 0) ? sum / air : -1. * sum / air;
    }
    wx[f] = g[0];
    wy[f] = g[1];
    wz[f] = g[2];
}

// ---------------------------------------------------------------------------------------------
// dissipative residual (explicitscheme_dissipative, solvers/diffusion/fvm_utils.py:5-26):
// all faces in id order, +q to cellid[0], -q to cellid[1] on inner faces.

__global__ void __launch_bounds__(kBlock) k_dissipative(Mesh m, const double *__restrict__ wx,
                                                         const double *__restrict__ wy,
                                                         const double *__restrict__ wz, double Dxx, double Dyy,
                                                         double Dzz, double *__restrict__ dis)
{
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= m.C) return;
    const int nf = m.cell_nfaces[K];
    double s = 0.;
    for (int jj = 0; jj < nf; jj++) {
        const int slot = m.order_sorted[(size_t)jj * m.C + K];
        const int f = m.cell_faces[(size_t)slot * m.C + K];
        const double *n = m.face_normal + 3 * f;
        const double z = wz ? wz[f] : 0.;
        const double q = Dxx * wx[f] * n[0] + Dyy * wy[f] * n[1] + Dzz * z * n[2];
        if (m.face_name[f] == 0 && m.face_cells[2 * f + 1] == K) s -= q;
        else s += q;
    }
    dis[K] = s;
}

// ---------------------------------------------------------------------------------------------
// time steps with diffusion: solvers/diffusion/fvm_utils.py:30-49 (u == NULL) and
// solvers/advecdiff/fvm_utils.py:317-340 (u, v, w cell velocities; cells with lam == 0 skipped)

__global__ void __launch_bounds__(kBlock) k_time_step_diff(Mesh m, const double *__restrict__ u,
                                                            const double *__restrict__ v,
                                                            const double *__restrict__ w, double cfl,
                                                            double Dxx, double Dyy, double Dzz,
                                                            double *__restrict__ dtpart)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dt = 1e6;
    if (i < m.C) {
        const double vol = m.cell_volume[i];
        const int nf = m.cell_nfaces[i];
        double lam = 0.;
        for (int j = 0; j < nf; j++) {
            const double *n = m.face_normal + 3 * (size_t)m.cell_faces[(size_t)j * m.C + i];
            if (u) lam += fabs(u[i] * n[0] + v[i] * n[1] + w[i] * n[2]);
            const double mes = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            const double lam_diff = Dxx * (mes * mes) + Dyy * (mes * mes) + Dzz * (mes * mes);
            lam += lam_diff / vol;
        }
        if (!u || lam != 0) dt = pmin(dt, cfl * vol / lam);
    }
    wave_min_part(dtpart, (size_t)i, dt);
}

}  // namespace mfv

using namespace mfv;

extern "C" int mfv_diamond_widths(const mfv_diamond *d, const uint8_t *node_dir, const int32_t *face_dir_rank,
                                  int32_t *slice_w, int32_t *info, void *stream)
{
    if (int rc = validate_diamond(d)) return rc;
    MFV_REQUIRE(node_dir && face_dir_rank && slice_w && info, "NULL array");
    const int S = (d->nb_cells + 63) / 64;
    if (S == 0) return MFV_OK;
    if (d->dim == 2)
        hipLaunchKernelGGL((k_diamond_width<2, kRowCap2>), dim3(S), dim3(64), 0, as_stream(stream), to_kernel(d),
                           node_dir, face_dir_rank, slice_w, info);
    else
        hipLaunchKernelGGL((k_diamond_width<3, kRowCap3>), dim3(S), dim3(64), 0, as_stream(stream), to_kernel(d),
                           node_dir, face_dir_rank, slice_w, info);
    return check_launch("k_diamond_width");
}

extern "C" int mfv_diamond_assemble(const mfv_diamond *d, const uint8_t *node_dir, const int32_t *face_dir_rank,
                                    int row0, int width, const int32_t *slice_w, int32_t *cols, double *vals,
                                    float *vals32, double *diag, void *stream)
{
    if (int rc = validate_diamond(d)) return rc;
    MFV_REQUIRE(node_dir && face_dir_rank && slice_w && cols && vals && vals32 && diag, "NULL array");
    const int cap = d->dim == 2 ? kRowCap2 : kRowCap3;
    MFV_REQUIRE(width >= 1 && width <= cap, "ELL width %d out of [1, %d]", width, cap);
    MFV_REQUIRE(row0 >= 0 && row0 % 64 == 0, "row0 %d must be a non-negative multiple of 64", row0);
    const int S = (d->nb_cells + 63) / 64;
    if (S == 0) return MFV_OK;
    if (d->dim == 2)
        hipLaunchKernelGGL((k_diamond_fill<2, kRowCap2>), dim3(S), dim3(64), 0, as_stream(stream), to_kernel(d),
                           node_dir, face_dir_rank, row0, width, slice_w, cols, vals, vals32, diag);
    else
        hipLaunchKernelGGL((k_diamond_fill<3, kRowCap3>), dim3(S), dim3(64), 0, as_stream(stream), to_kernel(d),
                           node_dir, face_dir_rank, row0, width, slice_w, cols, vals, vals32, diag);
    return check_launch("k_diamond_fill");
}

extern "C" int mfv_diamond_rhs(const mfv_diamond *d, const uint8_t *node_dir, const int32_t *face_dir_rank,
                               const double *Pbordnode, const double *Pbordface, double *rhs, void *stream)
{
    if (int rc = validate_diamond(d)) return rc;
    if (d->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(node_dir && face_dir_rank && Pbordnode && Pbordface && rhs, "NULL array");
    if (d->dim == 2)
        hipLaunchKernelGGL(k_diamond_rhs, dim3(grid_for(d->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                           to_kernel(d), node_dir, face_dir_rank, Pbordnode, Pbordface, rhs);
    else
        hipLaunchKernelGGL(k_diamond_rhs3, dim3(grid_for(d->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                           to_kernel(d), node_dir, face_dir_rank, Pbordnode, Pbordface, rhs);
    return check_launch("k_diamond_rhs");
}

extern "C" int mfv_centertovertex(const mfv_diamond *d, const double *w_c, const double *w_ghost, double *w_node,
                                  void *stream)
{
    if (int rc = validate_diamond(d)) return rc;
    if (d->nb_nodes == 0) return MFV_OK;
    MFV_REQUIRE(w_c && w_ghost && w_node, "NULL array");
    hipLaunchKernelGGL(k_centertovertex, dim3(grid_for(d->nb_nodes)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(d), w_c, w_ghost, w_node);
    return check_launch("k_centertovertex");
}

extern "C" int mfv_face_gradient(const mfv_diamond *d, int kind, const double *w_c, const double *w_ghost,
                                 const double *w_node, const int8_t *face_bc, const uint8_t *node_dir,
                                 const double *Pbordnode, const double *Pbordface, double *wx, double *wy,
                                 void *stream)
{
    if (int rc = validate_diamond(d)) return rc;
    MFV_REQUIRE(d->dim == 2, "mfv_face_gradient is the 2D scheme: use mfv_face_gradient_3d");
    MFV_REQUIRE(kind == 0 || kind == 1, "face gradient kind %d not in {0, 1}", kind);
    if (d->nb_faces == 0) return MFV_OK;
    MFV_REQUIRE(w_c && w_ghost && w_node && face_bc && wx && wy, "NULL array");
    if (kind == 1) MFV_REQUIRE(node_dir && Pbordnode && Pbordface, "kind 1 needs node_dir, Pbordnode, Pbordface");
    hipLaunchKernelGGL(k_face_gradient, dim3(grid_for(d->nb_faces)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(d), kind, w_c, w_ghost, w_node, face_bc, node_dir, Pbordnode, Pbordface, wx, wy);
    return check_launch("k_face_gradient");
}

extern "C" int mfv_face_gradient_3d(const mfv_diamond *d, int kind, const double *w_c, const double *w_ghost,
                                    const double *w_node, const int8_t *face_bc, const uint8_t *node_dir,
                                    const double *Pbordnode, const double *Pbordface, double *wx, double *wy,
                                    double *wz, void *stream)
{
    if (int rc = validate_diamond(d)) return rc;
    MFV_REQUIRE(d->dim == 3, "mfv_face_gradient_3d needs a 3D diamond descriptor");
    MFV_REQUIRE(kind == 0 || kind == 1, "face gradient kind %d not in {0, 1}", kind);
    if (d->nb_faces == 0) return MFV_OK;
    MFV_REQUIRE(w_c && w_ghost && w_node && face_bc && wx && wy && wz, "NULL array");
    if (kind == 1) MFV_REQUIRE(node_dir && Pbordnode && Pbordface, "kind 1 needs node_dir, Pbordnode, Pbordface");
    hipLaunchKernelGGL(k_face_gradient3, dim3(grid_for(d->nb_faces)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(d), kind, w_c, w_ghost, w_node, face_bc, node_dir, Pbordnode, Pbordface, wx, wy, wz);
    return check_launch("k_face_gradient3");
}

extern "C" int mfv_explicitscheme_dissipative(const mfv_mesh *m, const double *wx, const double *wy,
                                              const double *wz, double Dxx, double Dyy, double Dzz,
                                              double *dissip, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 1)) return rc;
    if (m->nb_cells == 0) return MFV_OK;
    MFV_REQUIRE(wx && wy && dissip, "NULL array");
    hipLaunchKernelGGL(k_dissipative, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                       to_kernel(m), wx, wy, wz, Dxx, Dyy, Dzz, dissip);
    return check_launch("k_dissipative");
}

extern "C" int mfv_time_step_diffusion(const mfv_mesh *m, const double *u, const double *v, const double *w,
                                       double cfl, double Dxx, double Dyy, double Dzz, double *dt_dev,
                                       double *dt_host, void *stream)
{
    if (int rc = validate_mesh(m, 0, 0, 0)) return rc;
    MFV_REQUIRE(dt_dev, "NULL dt_dev");
    MFV_REQUIRE((u && v && w) || (!u && !v && !w), "u, v, w must all be given or all be NULL");
    if (int rc = mfv_dt_reset(dt_dev, stream)) return rc;
    if (m->nb_cells) {
        const size_t nw = waves_for(m->nb_cells);
        double *part = dt_scratch(nw, as_stream(stream));
        if (!part) return MFV_ENOMEM;
        hipLaunchKernelGGL(k_time_step_diff, dim3(grid_for(m->nb_cells)), dim3(kBlock), 0, as_stream(stream),
                           to_kernel(m), u, v, w, cfl, Dxx, Dyy, Dzz, part);
        if (int rc = dt_finish(part, nw, dt_dev, as_stream(stream))) return rc;
    }
    return finish_dt(dt_dev, dt_host, stream);
}
