// libmfv host-side mesh topology: the connectivity that ddm/domain.py:130-612 builds with
// per-element Python loops (create_2dfaces / create_3dfaces, the sort + np.unique face
// numbering of domain.py:314-315, create_cell_faceid, create_cellsOfFace,
// create_NeighborCellByFace, create_node_cellid; ddm_utils2d.py:12-97, 482-505,
// ddm_utils3d.py:766-810), in C++ with sorts and counting passes instead of Python lists.
// Same numbering and orderings as the reference:
//  * a face's node row is its nodes sorted in descending order (-1 padding last) followed by
//    its node count (2 in 2D); faces are numbered in the lexicographic order of these rows
//    (what np.unique(axis=0) gives);
//  * faces.cellid[f][0] is the first cell (lowest id) listing f, [1] the last other one;
//  * cells.cellfid lists, in cells.faceid order, the cell across each inner face;
//  * nodes.cellid lists each node's cells in ascending order, cells.cellnid each cell's
//    vertex neighbours (cells sharing a node, itself excluded) in ascending order.
// Padded tables follow manapy's convention: [rows][width + 1], count in the last column,
// padding -1.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <thread>
#include <vector>

#include "mfv_common.hpp"

using namespace mfv;

namespace {

struct LocalFaces {
    int n;
    int nodes[6][4];
    int len[6];
};

// create_2dfaces (ddm_utils2d.py:482-505), create_3dfaces (ddm_utils3d.py:766-810)
static bool local_faces(int dim, int nv, LocalFaces &lf)
{
    static const int t2[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    static const int q2[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    static const int tet[4][3] = {{0, 1, 2}, {2, 3, 0}, {0, 1, 3}, {3, 1, 2}};
    static const int pyr[5][4] = {{0, 1, 4, -1}, {1, 2, 4, -1}, {2, 3, 4, -1}, {0, 3, 4, -1}, {0, 1, 2, 3}};
    static const int hex[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {1, 2, 5, 6}, {2, 3, 6, 7}, {0, 3, 4, 7}, {4, 5, 6, 7}};
    memset(&lf, -1, sizeof(lf));
    if (dim == 2 && (nv == 3 || nv == 4)) {
        lf.n = nv;
        for (int j = 0; j < nv; j++) {
            lf.len[j] = 2;
            for (int k = 0; k < 2; k++) lf.nodes[j][k] = nv == 3 ? t2[j][k] : q2[j][k];
        }
        return true;
    }
    if (dim == 3 && nv == 4) {
        lf.n = 4;
        for (int j = 0; j < 4; j++) {
            lf.len[j] = 3;
            for (int k = 0; k < 3; k++) lf.nodes[j][k] = tet[j][k];
        }
        return true;
    }
    if (dim == 3 && nv == 5) {
        lf.n = 5;
        for (int j = 0; j < 5; j++) {
            lf.len[j] = pyr[j][3] < 0 ? 3 : 4;
            for (int k = 0; k < lf.len[j]; k++) lf.nodes[j][k] = pyr[j][k];
        }
        return true;
    }
    if (dim == 3 && nv == 8) {
        lf.n = 6;
        for (int j = 0; j < 6; j++) {
            lf.len[j] = 4;
            for (int k = 0; k < 4; k++) lf.nodes[j][k] = hex[j][k];
        }
        return true;
    }
    return false;
}

template <class T>
T *copy_out(const std::vector<T> &v)
{
    T *p = static_cast<T *>(malloc(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (p && !v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

extern "C" void mfv_domain_topology_free(mfv_domain_topology *t)
{
    if (!t) return;
    free(t->faces_nodeid);
    free(t->cells_faceid);
    free(t->faces_cellid);
    free(t->cells_cellfid);
    free(t->nodes_cellid);
    free(t->cells_cellnid);
    memset(t, 0, sizeof(*t));
}

extern "C" int mfv_domain_topology_build(int dim, int C, int ew, const int32_t *elements, int N, int want_cellnid,
                                         mfv_domain_topology *out)
{
    MFV_REQUIRE(out, "out is NULL");
    memset(out, 0, sizeof(*out));
    MFV_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    MFV_REQUIRE(C > 0 && N > 0 && ew >= 3 && ew <= 9 && elements, "bad mesh sizes C=%d N=%d width=%d", C, N, ew);
    // local faces of every cell, rows in cell-major order (the k counter of create_*faces)
    std::vector<int> nf(C);
    int maxcellfid = 0, maxfacenid = dim == 2 ? 2 : 0;
    for (int c = 0; c < C; c++) {
        const int nv = elements[(size_t)c * ew + ew - 1];
        LocalFaces lf;
        MFV_REQUIRE(nv >= 1 && nv <= ew - 1 && local_faces(dim, nv, lf), "cell %d: unsupported %d-vertex cell in %dD",
                    c, nv, dim);
        nf[c] = lf.n;
        maxcellfid = std::max(maxcellfid, lf.n);
        if (dim == 3)
            for (int j = 0; j < lf.n; j++) maxfacenid = std::max(maxfacenid, lf.len[j]);
        for (int k = 0; k < nv; k++) {
            const int v = elements[(size_t)c * ew + k];
            MFV_REQUIRE(v >= 0 && v < N, "cell %d: vertex %d out of [0, %d)", c, v, N);
        }
    }
    const int fw = maxfacenid + 1;  // face row width: nodes (descending, -1 padding) + count
    std::vector<size_t> off(C + 1, 0);
    for (int c = 0; c < C; c++) off[c + 1] = off[c] + nf[c];
    const size_t P = off[C];
    std::vector<int32_t> rows(P * fw);
    for (int c = 0; c < C; c++) {
        const int nv = elements[(size_t)c * ew + ew - 1];
        LocalFaces lf;
        local_faces(dim, nv, lf);
        for (int j = 0; j < lf.n; j++) {
            int32_t *r = &rows[(off[c] + j) * fw];
            for (int k = 0; k < maxfacenid; k++) r[k] = k < lf.len[j] ? elements[(size_t)c * ew + lf.nodes[j][k]] : -1;
            std::sort(r, r + maxfacenid, [](int32_t a, int32_t b) { return a > b; });
            r[maxfacenid] = dim == 2 ? 2 : lf.len[j];
        }
    }
    // face numbering: lexicographic order of the rows (np.unique(axis=0)); ties keep row order.
    // When every face has maxfacenid nodes and N^maxfacenid < 2^62 the node columns fold into
    // one integer key with the same order (a radix-N number), sorted with its row index.
    std::vector<uint32_t> ord(P);
    double span = 1.;
    for (int k = 0; k < maxfacenid; k++) span *= (double)N;
    bool keyed = span < 4.0e18;
    for (size_t i = 0; keyed && i < P; i++)
        if (rows[i * fw + maxfacenid - 1] < 0) keyed = false;
    if (keyed) {
        std::vector<std::pair<uint64_t, uint32_t>> kv(P);
        for (size_t i = 0; i < P; i++) {
            uint64_t key = 0;
            for (int k = 0; k < maxfacenid; k++) key = key * (uint64_t)N + (uint64_t)rows[i * fw + k];
            kv[i] = {key, (uint32_t)i};
        }
        std::sort(kv.begin(), kv.end());
        for (size_t i = 0; i < P; i++) ord[i] = kv[i].second;
    } else {
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
            const int32_t *x = &rows[(size_t)a * fw], *y = &rows[(size_t)b * fw];
            for (int k = 0; k < fw; k++)
                if (x[k] != y[k]) return x[k] < y[k];
            return false;
        });
    }
    std::vector<int32_t> inv(P);
    std::vector<int32_t> fnode;
    int F = 0;
    for (size_t i = 0; i < P; i++) {
        const int32_t *r = &rows[(size_t)ord[i] * fw];
        if (i == 0 || memcmp(r, &rows[(size_t)ord[i - 1] * fw], fw * sizeof(int32_t)) != 0) {
            fnode.insert(fnode.end(), r, r + fw);
            F++;
        }
        inv[ord[i]] = F - 1;
    }
    // cells.faceid (create_cell_faceid)
    std::vector<int32_t> faceid((size_t)C * (maxcellfid + 1), -1);
    for (int c = 0; c < C; c++) {
        for (int j = 0; j < nf[c]; j++) faceid[(size_t)c * (maxcellfid + 1) + j] = inv[off[c] + j];
        faceid[(size_t)c * (maxcellfid + 1) + maxcellfid] = nf[c];
    }
    // faces.cellid (create_cellsOfFace): first visitor, then the last other visitor
    std::vector<int32_t> cellid((size_t)F * 2, -1);
    std::vector<int32_t> cnt(F, 0);
    for (int c = 0; c < C; c++)
        for (int j = 0; j < nf[c]; j++) {
            const int f = inv[off[c] + j];
            if (cnt[f] == 0) cellid[2 * (size_t)f] = c;
            else cellid[2 * (size_t)f + 1] = c;
            cnt[f]++;
        }
    // cells.cellfid (create_NeighborCellByFace): the cell across each inner face, faceid order
    std::vector<int32_t> cellfid((size_t)C * (maxcellfid + 1), -1);
    for (int c = 0; c < C; c++) {
        int n = 0;
        for (int j = 0; j < nf[c]; j++) {
            const int f = inv[off[c] + j];
            const int a = cellid[2 * (size_t)f], b = cellid[2 * (size_t)f + 1];
            if (b == -1) continue;
            cellfid[(size_t)c * (maxcellfid + 1) + n++] = (a == c) ? b : a;
        }
        cellfid[(size_t)c * (maxcellfid + 1) + maxcellfid] = n;
    }
    // nodes.cellid (create_node_cellid): cells of each node, ascending
    std::vector<int32_t> ncount(N, 0);
    for (int c = 0; c < C; c++) {
        const int nv = elements[(size_t)c * ew + ew - 1];
        for (int k = 0; k < nv; k++) ncount[elements[(size_t)c * ew + k]]++;
    }
    const int nw = std::max(1, *std::max_element(ncount.begin(), ncount.end()));
    std::vector<int32_t> ncell((size_t)N * (nw + 1), -1);
    std::vector<int32_t> fill(N, 0);
    for (int c = 0; c < C; c++) {
        const int nv = elements[(size_t)c * ew + ew - 1];
        for (int k = 0; k < nv; k++) {
            const int v = elements[(size_t)c * ew + k];
            ncell[(size_t)v * (nw + 1) + fill[v]++] = c;
        }
    }
    for (int v = 0; v < N; v++) ncell[(size_t)v * (nw + 1) + nw] = ncount[v];
    // cells.cellnid: union of the cells of the cell's nodes, itself excluded, ascending (only
    // when asked: at millions of tetrahedra it is large, and a first-order step never reads it)
    std::vector<int32_t> cellnid;
    int kw = -1;
    if (want_cellnid) {
        std::vector<int32_t> nb_ptr(C + 1, 0), nb;
        nb.reserve((size_t)C * 16);
        std::vector<int32_t> mark(C, -1), tmp;
        kw = 0;
        for (int c = 0; c < C; c++) {
            tmp.clear();
            const int nv = elements[(size_t)c * ew + ew - 1];
            for (int k = 0; k < nv; k++) {
                const int v = elements[(size_t)c * ew + k];
                for (int q = 0; q < ncount[v]; q++) {
                    const int o = ncell[(size_t)v * (nw + 1) + q];
                    if (o != c && mark[o] != c) {
                        mark[o] = c;
                        tmp.push_back(o);
                    }
                }
            }
            std::sort(tmp.begin(), tmp.end());
            nb.insert(nb.end(), tmp.begin(), tmp.end());
            nb_ptr[c + 1] = (int32_t)nb.size();
            kw = std::max(kw, (int)tmp.size());
        }
        cellnid.assign((size_t)C * (kw + 1), -1);
        for (int c = 0; c < C; c++) {
            const int n = nb_ptr[c + 1] - nb_ptr[c];
            memcpy(&cellnid[(size_t)c * (kw + 1)], &nb[nb_ptr[c]], n * sizeof(int32_t));
            cellnid[(size_t)c * (kw + 1) + kw] = n;
        }
    }
    out->nb_faces = F;
    out->max_cell_faces = maxcellfid;
    out->max_face_nodes = maxfacenid;
    out->node_cells_width = nw;
    out->cellnid_width = kw;
    out->faces_nodeid = copy_out(fnode);
    out->cells_faceid = copy_out(faceid);
    out->faces_cellid = copy_out(cellid);
    out->cells_cellfid = copy_out(cellfid);
    out->nodes_cellid = copy_out(ncell);
    out->cells_cellnid = copy_out(cellnid);
    if (!out->faces_nodeid || !out->cells_faceid || !out->faces_cellid || !out->cells_cellfid || !out->nodes_cellid ||
        !out->cells_cellnid) {
        mfv_domain_topology_free(out);
        set_error("out of host memory");
        return MFV_ENOMEM;
    }
    return MFV_OK;
}

// ---------------------------------------------------------------------------------------------
// 2D geometry, ghost mirrors and diamond coefficients (ddm_utils2d.py:122-223, 433-479, 507-705,
// 716-744), the same floating-point expressions in the same order as the NumPy construction, so
// the arrays are identical; loops over cells, faces and nodes run on host threads.

namespace {

template <class F>
void parallel_for(int n, F body)
{
    const int hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min(hw, n / 16384 + 1));
    if (nt == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) {
        const int a = (int)((int64_t)n * t / nt), b = (int)((int64_t)n * (t + 1) / nt);
        th.emplace_back([=] { body(a, b); });
    }
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" void mfv_domain_geometry_2d_free(mfv_domain_geometry_2d *g)
{
    if (!g) return;
    double **d[] = {&g->cell_center, &g->cell_volume, &g->face_normal, &g->face_mesure, &g->face_center,
                    &g->cell_nf, &g->face_ghostcenter, &g->dist_ortho, &g->airDiamond, &g->param1, &g->param2,
                    &g->param3, &g->param4, &g->f_1, &g->f_2, &g->f_3, &g->f_4, &g->lambda_x, &g->lambda_y,
                    &g->R_x, &g->R_y, &g->node_ghostcenter, &g->node_ghostfaceinfo};
    for (double **p : d) free(*p);
    free(g->face_name);
    free(g->number);
    free(g->node_ghostid);
    free(g->cell_ghostnid);
    memset(g, 0, sizeof(*g));
}

extern "C" int mfv_domain_geometry_2d_build(int C, int ew, const int32_t *nodeid, int N, const double *vertex, int F,
                                      const int32_t *faces_nodeid, int mf1, const int32_t *cells_faceid,
                                      const int32_t *faces_cellid, const int32_t *node_cellid, int ncw1,
                                      mfv_domain_geometry_2d *out)
{
    MFV_REQUIRE(out, "out is NULL");
    memset(out, 0, sizeof(*out));
    MFV_REQUIRE(C > 0 && N > 0 && F > 0 && nodeid && vertex && faces_nodeid && cells_faceid && faces_cellid &&
                    node_cellid && mf1 >= 2 && ncw1 >= 2,
                "bad arguments");
    const int mf = mf1 - 1;
    auto X = [&](int n) { return vertex[4 * (size_t)n]; };
    auto Y = [&](int n) { return vertex[4 * (size_t)n + 1]; };
    std::vector<double> cc((size_t)C * 3, 0.), vol(C, 0.);
    // Compute_2dcentervolumeOfCell (ddm_utils2d.py:433-479)
    parallel_for(C, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            const int32_t *s = nodeid + (size_t)i * ew;
            const int nv = s[ew - 1];
            if (nv == 3) {
                const double x1 = X(s[0]), y1 = Y(s[0]), x2 = X(s[1]), y2 = Y(s[1]), x3 = X(s[2]), y3 = Y(s[2]);
                cc[3 * (size_t)i] = 1. / 3 * ((x1 + x2) + x3);
                cc[3 * (size_t)i + 1] = 1. / 3 * ((y1 + y2) + y3);
                vol[i] = (1. / 2) * fabs((x1 - x2) * (y1 - y3) - (x1 - x3) * (y1 - y2));
            } else {
                double xs[4], ys[4];
                for (int k = 0; k < 4; k++) {
                    xs[k] = X(s[k]);
                    ys[k] = Y(s[k]);
                }
                const double cx = (((xs[0] + xs[1]) + xs[2]) + xs[3]) / 4;
                const double cy = (((ys[0] + ys[1]) + ys[2]) + ys[3]) / 4;
                cc[3 * (size_t)i] = cx;
                cc[3 * (size_t)i + 1] = cy;
                double v = 0.;
                for (int k = 0; k < 4; k++) {
                    const double x1 = xs[k], y1 = ys[k], x2 = xs[(k + 1) % 4], y2 = ys[(k + 1) % 4];
                    v += (1. / 2) * fabs((x1 - x2) * (y1 - cy) - (x1 - cx) * (y1 - y2));
                }
                vol[i] = v;
            }
        }
    });
    // create_info_2dfaces (ddm_utils2d.py:507-551): names, outward non-unit normals, lengths
    std::vector<int32_t> name(F, 0);
    std::vector<double> nrm((size_t)F * 3, 0.), mes(F), cf((size_t)F * 3);
    parallel_for(F, [&](int a, int b) {
        for (int f = a; f < b; f++) {
            const int n0 = faces_nodeid[(size_t)f * 3], n1 = faces_nodeid[(size_t)f * 3 + 1];
            if (faces_cellid[2 * (size_t)f + 1] == -1) {
                const int ta = (int)vertex[4 * (size_t)n0 + 3], tb = (int)vertex[4 * (size_t)n1 + 3];
                int nm = 100;
                if (ta == tb) {
                    nm = ta;
                } else {
                    const int tags[4] = {3, 4, 33, 44};
                    for (int t : tags)
                        if ((ta == t && tb != 0) || (ta != 0 && tb == t)) {
                            nm = t;
                            break;
                        }
                }
                name[f] = nm;
            }
            const double normx = Y(n0) - Y(n1), normy = X(n1) - X(n0);
            for (int k = 0; k < 3; k++)
                cf[3 * (size_t)f + k] = 0.5 * (vertex[4 * (size_t)n0 + k] + vertex[4 * (size_t)n1 + k]);
            const int c0 = faces_cellid[2 * (size_t)f];
            const double sx = cc[3 * (size_t)c0] - cf[3 * (size_t)f], sy = cc[3 * (size_t)c0 + 1] - cf[3 * (size_t)f + 1];
            const bool flip = (sx * normx + sy * normy) > 0;
            const double nx = flip ? -1 * normx : normx, ny = flip ? -1 * normy : normy;
            nrm[3 * (size_t)f] = nx;
            nrm[3 * (size_t)f + 1] = ny;
            mes[f] = sqrt(nx * nx + ny * ny);
        }
    });
    // create_NormalFacesOfCell (ddm_utils2d.py:99-120)
    std::vector<double> nf((size_t)C * mf * 3, 0.);
    parallel_for(C, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            const int32_t *fr = cells_faceid + (size_t)i * mf1;
            for (int j = 0; j < fr[mf]; j++) {
                const int f = fr[j];
                const double d0 = cc[3 * (size_t)i] - cf[3 * (size_t)f], d1 = cc[3 * (size_t)i + 1] - cf[3 * (size_t)f + 1],
                             d2 = cc[3 * (size_t)i + 2] - cf[3 * (size_t)f + 2];
                const double *n = &nrm[3 * (size_t)f];
                const bool inward = ((d0 * n[0] + d1 * n[1]) + d2 * n[2]) < 0.;
                for (int k = 0; k < 3; k++) nf[((size_t)i * mf + j) * 3 + k] = inward ? n[k] : -1.0 * n[k];
            }
        }
    });
    // boundary faces (names 1-4, ascending) and face_info_2d ghost mirrors (ddm_utils2d.py:163-223)
    std::vector<int32_t> bf;
    for (int f = 0; f < F; f++)
        if (name[f] >= 1 && name[f] <= 4) bf.push_back(f);
    const int nb = (int)bf.size();
    std::vector<double> gc((size_t)F * 3, -1.);
    parallel_for(nb, [&](int a, int b) {
        for (int q = a; q < b; q++) {
            const int f = bf[q];
            const int nod1 = faces_nodeid[(size_t)f * 3 + 1], nod2 = faces_nodeid[(size_t)f * 3];
            const double x10 = X(nod1), x11 = Y(nod1), x20 = X(nod2), x21 = Y(nod2);
            const int c0 = faces_cellid[2 * (size_t)f];
            const double v10 = cc[3 * (size_t)c0], v11 = cc[3 * (size_t)c0 + 1];
            const double gamma = ((v10 - x20) * (x10 - x20) + (v11 - x21) * (x11 - x21)) /
                                 ((x10 - x20) * (x10 - x20) + (x11 - x21) * (x11 - x21));
            const double kk0 = gamma * x10 + (1 - gamma) * x20, kk1 = gamma * x11 + (1 - gamma) * x21;
            gc[3 * (size_t)f] = 2 * kk0 + (-1 * v10);
            gc[3 * (size_t)f + 1] = 2 * kk1 + (-1 * v11);
            gc[3 * (size_t)f + 2] = gamma;
        }
    });
    // per-node ghost lists, by boundary face then by the node's slot in it (nod1, nod2)
    std::vector<int32_t> gcount(N, 0);
    for (int q = 0; q < nb; q++)
        for (int s = 0; s < 2; s++) gcount[faces_nodeid[(size_t)bf[q] * 3 + (s == 0 ? 1 : 0)]]++;
    int g = 1;
    for (int v = 0; v < N; v++) g = std::max(g, (int)gcount[v]);
    std::vector<double> gcn((size_t)N * g * 5, -1.), gfi((size_t)N * g * 4, -1.);
    std::vector<int32_t> gid((size_t)N * (g + 1), -1), gfill(N, 0);
    for (int q = 0; q < nb; q++) {
        const int f = bf[q];
        for (int s = 0; s < 2; s++) {
            const int v = faces_nodeid[(size_t)f * 3 + (s == 0 ? 1 : 0)];
            const int p = gfill[v]++;
            double *r = &gcn[((size_t)v * g + p) * 5];
            r[0] = gc[3 * (size_t)f];
            r[1] = gc[3 * (size_t)f + 1];
            r[2] = (double)faces_cellid[2 * (size_t)f];
            r[3] = (double)name[f];
            r[4] = (double)f;
            double *t = &gfi[((size_t)v * g + p) * 4];
            t[0] = cf[3 * (size_t)f];
            t[1] = cf[3 * (size_t)f + 1];
            t[2] = nrm[3 * (size_t)f];
            t[3] = nrm[3 * (size_t)f + 1];
            gid[(size_t)v * (g + 1) + p] = f;
        }
    }
    for (int v = 0; v < N; v++) gid[(size_t)v * (g + 1) + g] = gcount[v];
    // cells.ghostnid: the sorted union of the ghost faces of the cell's nodes
    std::vector<std::vector<int32_t>> cg(C);
    int gw = 0;
    parallel_for(C, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            const int32_t *s = nodeid + (size_t)i * ew;
            std::vector<int32_t> &l = cg[i];
            for (int k = 0; k < s[ew - 1]; k++)
                for (int p = 0; p < gcount[s[k]]; p++) l.push_back(gid[(size_t)s[k] * (g + 1) + p]);
            std::sort(l.begin(), l.end());
            l.erase(std::unique(l.begin(), l.end()), l.end());
        }
    });
    for (int i = 0; i < C; i++) gw = std::max(gw, (int)cg[i].size());
    std::vector<int32_t> gnid((size_t)C * (gw + 1), -1);
    for (int i = 0; i < C; i++) {
        std::copy(cg[i].begin(), cg[i].end(), &gnid[(size_t)i * (gw + 1)]);
        gnid[(size_t)i * (gw + 1) + gw] = (int32_t)cg[i].size();
    }
    // variables_2d (ddm_utils2d.py:608-705): the node's cells, then its ghost mirror centres
    std::vector<double> lx(N), ly(N), Rx(N), Ry(N);
    std::vector<int32_t> number(N);
    parallel_for(N, [&](int a, int b) {
        for (int v = a; v < b; v++) {
            double Ixx = 0., Iyy = 0., Ixy = 0., rx = 0., ry = 0.;
            int num = 0;
            const double vx = X(v), vy = Y(v);
            const int nc = node_cellid[(size_t)v * ncw1 + ncw1 - 1];
            for (int j = 0; j < nc; j++) {
                const int c = node_cellid[(size_t)v * ncw1 + j];
                const double dx = cc[3 * (size_t)c] - vx, dy = cc[3 * (size_t)c + 1] - vy;
                Ixx += dx * dx;
                Iyy += dy * dy;
                Ixy += dx * dy;
                rx += dx;
                ry += dy;
                num++;
            }
            for (int j = 0; j < gcount[v]; j++) {
                const int f = gid[(size_t)v * (g + 1) + j];
                const double dx = gc[3 * (size_t)f] - vx, dy = gc[3 * (size_t)f + 1] - vy;
                Ixx += dx * dx;
                Iyy += dy * dy;
                Ixy += dx * dy;
                rx += dx;
                ry += dy;
                num++;
            }
            const double D = Ixx * Iyy - Ixy * Ixy;
            lx[v] = (Ixy * ry - Iyy * rx) / D;
            ly[v] = (Ixy * rx - Ixx * ry) / D;
            Rx[v] = rx;
            Ry[v] = ry;
            number[v] = num;
        }
    });
    // face_gradient_info_2d (ddm_utils2d.py:553-605) and dist_ortho (ddm_utils2d.py:716-744)
    std::vector<double> air(F), p1(F), p2(F), p3(F), p4(F), f1((size_t)F * 2), f2((size_t)F * 2), f3((size_t)F * 2),
        f4((size_t)F * 2), dort(F, 0.);
    parallel_for(F, [&](int a, int b) {
        for (int f = a; f < b; f++) {
            const int n0 = faces_nodeid[(size_t)f * 3], n1 = faces_nodeid[(size_t)f * 3 + 1];
            const double xy1[2] = {X(n0), Y(n0)}, xy2[2] = {X(n1), Y(n1)};
            // a boundary face left unnamed (its nodes untagged) keeps name 0 with no second cell:
            // the NumPy construction's index -1 then takes the last cell, and so does this one
            const int c0 = faces_cellid[2 * (size_t)f], c1r = faces_cellid[2 * (size_t)f + 1];
            const int c1 = c1r < 0 ? C + c1r : c1r;
            const double v1[2] = {cc[3 * (size_t)c0], cc[3 * (size_t)c0 + 1]};
            double v2[2] = {gc[3 * (size_t)f], gc[3 * (size_t)f + 1]};
            if (name[f] == 0) {
                v2[0] = cc[3 * (size_t)c1];
                v2[1] = cc[3 * (size_t)c1 + 1];
            }
            const double F1[2] = {v1[0] - xy1[0], v1[1] - xy1[1]}, F2[2] = {xy2[0] - v1[0], xy2[1] - v1[1]},
                         F3[2] = {v2[0] - xy2[0], v2[1] - xy2[1]}, F4[2] = {xy1[0] - v2[0], xy1[1] - v2[1]};
            const double nx = nrm[3 * (size_t)f], ny = nrm[3 * (size_t)f + 1];
            const double ar = 0.5 * ((xy2[0] - xy1[0]) * (v2[1] - v1[1]) + (v1[0] - v2[0]) * (xy2[1] - xy1[1]));
            const double k = 1. / (2. * ar);
            air[f] = ar;
            p1[f] = k * ((F1[1] + F2[1]) * nx - (F1[0] + F2[0]) * ny);
            p2[f] = k * ((F2[1] + F3[1]) * nx - (F2[0] + F3[0]) * ny);
            p3[f] = k * ((F3[1] + F4[1]) * nx - (F3[0] + F4[0]) * ny);
            p4[f] = k * ((F4[1] + F1[1]) * nx - (F4[0] + F1[0]) * ny);
            for (int q = 0; q < 2; q++) {
                f1[2 * (size_t)f + q] = F1[q];
                f2[2 * (size_t)f + q] = F2[q];
                f3[2 * (size_t)f + q] = F3[q];
                f4[2 * (size_t)f + q] = F4[q];
            }
            auto proj = [&](int K) {
                const double vx = cc[3 * (size_t)K] - cf[3 * (size_t)f], vy = cc[3 * (size_t)K + 1] - cf[3 * (size_t)f + 1];
                const double s = vx * nx + vy * ny;
                const double px = cc[3 * (size_t)K] - s * nx, py = cc[3 * (size_t)K + 1] - s * ny;
                const double dx = cc[3 * (size_t)K] - px, dy = cc[3 * (size_t)K + 1] - py;
                return sqrt(dx * dx + dy * dy);
            };
            if (name[f] == 0) dort[f] = proj(c0) + proj(c1);
            else if (name[f] >= 1 && name[f] <= 4) dort[f] = 2 * proj(c0);
        }
    });
    out->ghost_width = g;
    out->ghostnid_width = gw;
    out->cell_center = copy_out(cc);
    out->cell_volume = copy_out(vol);
    out->face_name = copy_out(name);
    out->face_normal = copy_out(nrm);
    out->face_mesure = copy_out(mes);
    out->face_center = copy_out(cf);
    out->cell_nf = copy_out(nf);
    out->face_ghostcenter = copy_out(gc);
    out->dist_ortho = copy_out(dort);
    out->airDiamond = copy_out(air);
    out->param1 = copy_out(p1);
    out->param2 = copy_out(p2);
    out->param3 = copy_out(p3);
    out->param4 = copy_out(p4);
    out->f_1 = copy_out(f1);
    out->f_2 = copy_out(f2);
    out->f_3 = copy_out(f3);
    out->f_4 = copy_out(f4);
    out->lambda_x = copy_out(lx);
    out->lambda_y = copy_out(ly);
    out->R_x = copy_out(Rx);
    out->R_y = copy_out(Ry);
    out->number = copy_out(number);
    out->node_ghostcenter = copy_out(gcn);
    out->node_ghostfaceinfo = copy_out(gfi);
    out->node_ghostid = copy_out(gid);
    out->cell_ghostnid = copy_out(gnid);
    return MFV_OK;
}

// ---------------------------------------------------------------------------------------------
// 3D geometry, face orientation, ghost mirrors and diamond coefficients (ddm_utils3d.py:87-157,
// 362-403, 459-527, 530-764, 813-981; ddm_utils2d.py:122-161, 716-744), the NumPy construction's
// expressions in its order, so the arrays are identical; loops run on host threads.

namespace {

struct V3 {
    double x, y, z;
};

// outputs are malloc'ed (freed by mfv_domain_geometry_3d_free); a V3 array is a [n][3] double array
template <typename T>
T *alloc_out(size_t n)
{
    return static_cast<T *>(malloc((n ? n : 1) * sizeof(T)));
}

V3 *alloc_v3(double *&slot, size_t n)
{
    slot = alloc_out<double>(3 * n);
    return reinterpret_cast<V3 *>(slot);
}

inline V3 v3sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 v3add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 v3scale(double s, V3 v) { return {s * v.x, s * v.y, s * v.z}; }
inline V3 v3cross(V3 u, V3 v) { return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x}; }
inline double v3dot(V3 u, V3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }

// create_info_3dfaces' boundary naming (ddm_utils3d.py:549-593, 615-669)
int name3d(const int *nm, int k)
{
    bool same = true;
    for (int j = 1; j < k; j++) same = same && nm[j] == nm[0];
    if (same) return nm[0];
    const int tags[8] = {5, 6, 3, 4, 55, 66, 33, 44};
    for (int tag : tags)
        for (int j = 0; j < k; j++) {
            if (nm[j] != tag) continue;
            bool others = true;
            for (int l = 0; l < k; l++)
                if (l != j) others = others && nm[l] != 0;
            if (others) return tag;
        }
    return k == 3 ? 100 : 200;
}

}  // namespace

extern "C" void mfv_domain_geometry_3d_free(mfv_domain_geometry_3d *g)
{
    if (!g) return;
    double **d[] = {&g->cell_center, &g->cell_volume, &g->face_normal, &g->face_mesure, &g->face_center,
                    &g->face_tangent, &g->face_binormal, &g->face_ghostcenter, &g->dist_ortho, &g->airDiamond,
                    &g->param1, &g->param2, &g->param3, &g->f_1, &g->f_2, &g->lambda_x, &g->lambda_y,
                    &g->lambda_z, &g->R_x, &g->R_y, &g->R_z, &g->node_ghostcenter, &g->node_ghostfaceinfo};
    for (double **p : d) free(*p);
    free(g->face_name);
    free(g->face_nodeid);
    free(g->number);
    free(g->node_ghostid);
    free(g->cell_ghostnid);
    memset(g, 0, sizeof(*g));
}

extern "C" int mfv_domain_geometry_3d_build(int C, int ew, const int32_t *nodeid, int N, const double *vertex, int F,
                                            int fw, const int32_t *faces_nodeid, const int32_t *faces_cellid,
                                            const int32_t *node_cellid, int ncw1, mfv_domain_geometry_3d *out)
{
    MFV_REQUIRE(out, "out is NULL");
    memset(out, 0, sizeof(*out));
    MFV_REQUIRE(C > 0 && N > 0 && F > 0 && nodeid && vertex && faces_nodeid && faces_cellid && node_cellid &&
                    ew >= 5 && (fw == 4 || fw == 5) && ncw1 >= 2,
                "bad arguments");
    auto P = [&](int n) { return V3{vertex[4 * (size_t)n], vertex[4 * (size_t)n + 1], vertex[4 * (size_t)n + 2]}; };
    auto tag = [&](int n) { return (int)vertex[4 * (size_t)n + 3]; };
    // Compute_3dcentervolumeOfCell (ddm_utils3d.py:686-764): tetrahedra, pyramids and hexahedra
    for (int i = 0; i < C; i++) {
        const int nv = nodeid[(size_t)i * ew + ew - 1];
        MFV_REQUIRE(nv == 4 || nv == 5 || nv == 8, "cell %d has %d vertices: tetrahedra, pyramids and hexahedra only",
                    i, nv);
    }
    // outputs of known size are allocated up front and filled in place by the threads
    V3 *cc = alloc_v3(out->cell_center, C);
    V3 *nrm = alloc_v3(out->face_normal, F), *cf = alloc_v3(out->face_center, F);
    V3 *tan = alloc_v3(out->face_tangent, F), *bin = alloc_v3(out->face_binormal, F);
    V3 *n1v = alloc_v3(out->f_1, F), *n2v = alloc_v3(out->f_2, F);
    double *vol = out->cell_volume = alloc_out<double>(C);
    int32_t *name = out->face_name = alloc_out<int32_t>(F);
    int32_t *fn = out->face_nodeid = alloc_out<int32_t>((size_t)F * fw);
    double *mes = out->face_mesure = alloc_out<double>(F);
    double *gc = out->face_ghostcenter = alloc_out<double>((size_t)F * 4);
    double *air = out->airDiamond = alloc_out<double>(F);
    double *p1 = out->param1 = alloc_out<double>(F), *p2 = out->param2 = alloc_out<double>(F),
           *p3 = out->param3 = alloc_out<double>(F);
    double *dort = out->dist_ortho = alloc_out<double>(F);
    double *lx = out->lambda_x = alloc_out<double>(N), *ly = out->lambda_y = alloc_out<double>(N),
           *lz = out->lambda_z = alloc_out<double>(N);
    double *Rx = out->R_x = alloc_out<double>(N), *Ry = out->R_y = alloc_out<double>(N),
           *Rz = out->R_z = alloc_out<double>(N);
    int32_t *number = out->number = alloc_out<int32_t>(N);
    if (!cc || !vol || !name || !fn || !nrm || !cf || !tan || !bin || !n1v || !n2v || !mes || !gc || !air || !p1 ||
        !p2 || !p3 || !dort || !lx || !ly || !lz || !Rx || !Ry || !Rz || !number) {
        mfv_domain_geometry_3d_free(out);
        set_error("out of host memory");
        return MFV_ENOMEM;
    }
    parallel_for(C, [&](int a, int b) {
        static const int hexr[6][4] = {{0, 1, 3, 4}, {1, 3, 4, 5}, {4, 5, 3, 7}, {1, 3, 5, 2}, {3, 7, 5, 2}, {5, 7, 6, 2}};
        static const int pyr[2][4] = {{0, 1, 2, 4}, {0, 2, 3, 4}};
        for (int i = a; i < b; i++) {
            const int32_t *s = nodeid + (size_t)i * ew;
            const int nv = s[ew - 1];
            V3 vt[8];
            V3 acc{0., 0., 0.};
            for (int k = 0; k < nv; k++) {
                vt[k] = P(s[k]);
                acc = v3add(acc, vt[k]);
            }
            cc[i] = V3{acc.x / nv, acc.y / nv, acc.z / nv};
            auto tet = [&](int r0, int r1, int r2, int r3) {
                const V3 u = v3sub(vt[r1], vt[r0]), v = v3sub(vt[r2], vt[r0]), w = v3sub(vt[r3], vt[r0]);
                return fabs(v3dot(u, v3cross(v, w))) / 6.0;
            };
            if (nv == 4) {
                vol[i] = tet(0, 1, 2, 3);
            } else {
                double v = 0.;
                const int nr = nv == 8 ? 6 : 2;
                for (int r = 0; r < nr; r++) {
                    const int *q = nv == 8 ? hexr[r] : pyr[r];
                    v += tet(q[0], q[1], q[2], q[3]);
                }
                vol[i] = v;
            }
        }
    });
    // create_info_3dfaces (ddm_utils3d.py:530-684): centres, areas, names, outward normals,
    // tangents and binormals; then oriente_3dfacenodeid (:362-403) on a copy of the face table
    parallel_for(F, [&](int a, int b) {
        for (int f = a; f < b; f++) {
            int32_t *r = &fn[(size_t)f * fw];
            for (int j = 0; j < fw; j++) r[j] = faces_nodeid[(size_t)f * fw + j];
            name[f] = 0;
            const int k = r[fw - 1];
            V3 v[4];
            for (int j = 0; j < k; j++) v[j] = P(r[j]);
            V3 norm;
            if (k == 3) {
                const V3 s3 = v3add(v3add(v[0], v[1]), v[2]);
                cf[f] = v3scale(1. / 3, s3);
                norm = v3scale(0.5, v3cross(v3sub(v[1], v[0]), v3sub(v[2], v[0])));
                mes[f] = sqrt(norm.x * norm.x + norm.y * norm.y + norm.z * norm.z);
            } else {
                const V3 s4 = v3add(v3add(v3add(v[0], v[1]), v[2]), v[3]);
                cf[f] = v3scale(1. / 4, s4);
                const V3 a1 = v3scale(0.5, v3cross(v3sub(v[1], v[0]), v3sub(v[2], v[0])));
                norm = v3add(a1, a1);  // area2 = area1 (the reference's quadrangle area, ddm_utils3d.py:605)
                const double mag = sqrt(a1.x * a1.x + a1.y * a1.y + a1.z * a1.z);
                mes[f] = mag + mag;
            }
            if (faces_cellid[2 * (size_t)f + 1] == -1) {
                int nm[4];
                for (int j = 0; j < k; j++) nm[j] = tag(r[j]);
                name[f] = name3d(nm, k);
            }
            const V3 snorm = v3sub(cc[faces_cellid[2 * (size_t)f]], cf[f]);
            const V3 n = v3dot(snorm, norm) > 0 ? v3scale(-1, norm) : norm;
            nrm[f] = n;
            const V3 t = v3sub(P(r[1]), P(r[0]));
            tan[f] = t;
            bin[f] = v3scale(0.5, v3cross(t, n));
            if (k == 3) {
                const int n1 = r[0], n2 = r[1], n3 = r[2];
                const V3 e1 = v3sub(P(n2), P(n1)), e2 = v3sub(P(n3), P(n1));
                if (v3dot(v3cross(e1, e2), n) < 0) {
                    r[1] = n3;
                    r[2] = n2;
                }
            } else {
                const int s[4] = {r[0], r[1], r[2], r[3]};
                auto cz = [&](int p, int q) { return P(p).x * P(q).y - P(p).y * P(q).x; };
                auto orient = [&](int p, int q, int w) { return cz(p, q) + cz(q, w) + cz(w, p); };
                const bool o123 = orient(s[0], s[1], s[2]) < 0, o134 = orient(s[0], s[2], s[3]) < 0,
                           o124 = orient(s[0], s[1], s[3]) < 0;
                if (o123 && !o134) {
                    if (o124) {
                        r[2] = s[3];
                        r[3] = s[2];
                    } else {
                        r[0] = s[3];
                        r[3] = s[0];
                    }
                } else if (!o123 && o134) {
                    if (o124) {
                        r[1] = s[2];
                        r[2] = s[1];
                    } else {
                        r[0] = s[1];
                        r[1] = s[0];
                    }
                } else if (!o123 && !o134) {
                    r[0] = s[2];
                    r[2] = s[0];
                }
            }
        }
    });
    // face_info_3d (ddm_utils3d.py:87-157): mirror centres of the boundary faces (names 1-6)
    std::vector<int32_t> bf;
    for (int f = 0; f < F; f++)
        if (name[f] >= 1 && name[f] <= 6) bf.push_back(f);
    const int nb = (int)bf.size();
    parallel_for(F, [&](int a, int b) {
        for (size_t q = 4 * (size_t)a; q < 4 * (size_t)b; q++) gc[q] = -1.;
    });
    parallel_for(nb, [&](int a, int b) {
        for (int q = a; q < b; q++) {
            const int f = bf[q];
            const V3 n{nrm[f].x / mes[f], nrm[f].y / mes[f], nrm[f].z / mes[f]};
            const V3 v1 = cc[faces_cellid[2 * (size_t)f]];
            const double gamma = v3dot(v3sub(cf[f], v1), n);
            const V3 kk = v3add(v1, v3scale(gamma, n));
            const V3 v2 = v3add(v3scale(2, kk), v3scale(-1, v1));
            gc[4 * (size_t)f] = v2.x;
            gc[4 * (size_t)f + 1] = v2.y;
            gc[4 * (size_t)f + 2] = v2.z;
            gc[4 * (size_t)f + 3] = gamma;
        }
    });
    // per-node ghost lists, by boundary face then by the node's slot in it (nodes 1, 0, 2, 3)
    const int slot_node[4] = {1, 0, 2, 3};
    auto nslots = [&](int f) { return fn[(size_t)f * fw + fw - 1] == 4 ? 4 : 3; };
    std::vector<int32_t> gcount(N, 0);
    for (int q = 0; q < nb; q++)
        for (int s = 0; s < nslots(bf[q]); s++) gcount[fn[(size_t)bf[q] * fw + slot_node[s]]]++;
    int g = 1;
    for (int v = 0; v < N; v++) g = std::max(g, (int)gcount[v]);
    double *gcn = alloc_out<double>((size_t)N * g * 6), *gfi = alloc_out<double>((size_t)N * g * 6);
    int32_t *gid = alloc_out<int32_t>((size_t)N * (g + 1));
    out->ghost_width = g;
    out->node_ghostcenter = gcn;
    out->node_ghostfaceinfo = gfi;
    out->node_ghostid = gid;
    if (!gcn || !gfi || !gid) {
        mfv_domain_geometry_3d_free(out);
        set_error("out of host memory");
        return MFV_ENOMEM;
    }
    parallel_for(N, [&](int a, int b) {
        std::fill(gcn + (size_t)a * g * 6, gcn + (size_t)b * g * 6, -1.);
        std::fill(gfi + (size_t)a * g * 6, gfi + (size_t)b * g * 6, -1.);
        std::fill(gid + (size_t)a * (g + 1), gid + (size_t)b * (g + 1), -1);
    });
    std::vector<int32_t> gfill(N, 0);
    for (int q = 0; q < nb; q++) {
        const int f = bf[q];
        for (int s = 0; s < nslots(f); s++) {
            const int v = fn[(size_t)f * fw + slot_node[s]];
            const int p = gfill[v]++;
            double *r = &gcn[((size_t)v * g + p) * 6];
            r[0] = gc[4 * (size_t)f];
            r[1] = gc[4 * (size_t)f + 1];
            r[2] = gc[4 * (size_t)f + 2];
            r[3] = (double)faces_cellid[2 * (size_t)f];
            r[4] = (double)name[f];
            r[5] = (double)f;
            double *t = &gfi[((size_t)v * g + p) * 6];
            t[0] = cf[f].x;
            t[1] = cf[f].y;
            t[2] = cf[f].z;
            t[3] = nrm[f].x;
            t[4] = nrm[f].y;
            t[5] = nrm[f].z;
            gid[(size_t)v * (g + 1) + p] = f;
        }
    }
    for (int v = 0; v < N; v++) gid[(size_t)v * (g + 1) + g] = gcount[v];
    // cells.ghostnid: the sorted union of the ghost faces of the cell's nodes
    std::vector<std::vector<int32_t>> cg(C);
    parallel_for(C, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            const int32_t *s = nodeid + (size_t)i * ew;
            std::vector<int32_t> &l = cg[i];
            for (int k = 0; k < s[ew - 1]; k++)
                for (int p = 0; p < gcount[s[k]]; p++) l.push_back(gid[(size_t)s[k] * (g + 1) + p]);
            std::sort(l.begin(), l.end());
            l.erase(std::unique(l.begin(), l.end()), l.end());
        }
    });
    int gw = 0;
    for (int i = 0; i < C; i++) gw = std::max(gw, (int)cg[i].size());
    int32_t *gnid = alloc_out<int32_t>((size_t)C * (gw + 1));
    out->ghostnid_width = gw;
    out->cell_ghostnid = gnid;
    if (!gnid) {
        mfv_domain_geometry_3d_free(out);
        set_error("out of host memory");
        return MFV_ENOMEM;
    }
    parallel_for(C, [&](int a, int b) {
        for (int i = a; i < b; i++) {
            int32_t *row = gnid + (size_t)i * (gw + 1);
            std::fill(row, row + gw, -1);
            std::copy(cg[i].begin(), cg[i].end(), row);
            row[gw] = (int32_t)cg[i].size();
        }
    });
    // variables_3d (ddm_utils3d.py:813-981): the node's cells, then its ghost mirror centres
    parallel_for(N, [&](int a, int b) {
        for (int v = a; v < b; v++) {
            double xx = 0., yy = 0., zz = 0., xy = 0., xz = 0., yz = 0., rx = 0., ry = 0., rz = 0.;
            int num = 0;
            const V3 vt = P(v);
            auto add = [&](V3 p) {
                const double dx = p.x - vt.x, dy = p.y - vt.y, dz = p.z - vt.z;
                xx += dx * dx;
                yy += dy * dy;
                zz += dz * dz;
                xy += dx * dy;
                xz += dx * dz;
                yz += dy * dz;
                rx += dx;
                ry += dy;
                rz += dz;
                num++;
            };
            const int nc = node_cellid[(size_t)v * ncw1 + ncw1 - 1];
            for (int j = 0; j < nc; j++) add(cc[node_cellid[(size_t)v * ncw1 + j]]);
            for (int j = 0; j < gcount[v]; j++) {
                const int f = gid[(size_t)v * (g + 1) + j];
                add(V3{gc[4 * (size_t)f], gc[4 * (size_t)f + 1], gc[4 * (size_t)f + 2]});
            }
            const double D = xx * yy * zz + 2 * xy * xz * yz - xx * yz * yz - yy * xz * xz - zz * xy * xy;
            lx[v] = ((yz * yz - yy * zz) * rx + (xy * zz - xz * yz) * ry + (xz * yy - xy * yz) * rz) / D;
            ly[v] = ((xy * zz - xz * yz) * rx + (xz * xz - xx * zz) * ry + (yz * xx - xz * xy) * rz) / D;
            lz[v] = ((xz * yy - xy * yz) * rx + (yz * xx - xz * xy) * ry + (xy * xy - xx * yy) * rz) / D;
            Rx[v] = rx;
            Ry[v] = ry;
            Rz[v] = rz;
            number[v] = num;
        }
    });
    // face_gradient_info_3d (ddm_utils3d.py:459-527) and dist_ortho (ddm_utils2d.py:716-744)
    parallel_for(F, [&](int a, int b) {
        for (int f = a; f < b; f++) {
            const int32_t *r = &fn[(size_t)f * fw];
            const int i1 = r[0], i2 = r[1], i3 = r[2];
            const int i4 = (fw > 4 && r[fw - 1] == 4) ? r[3] : i3;
            // an unnamed boundary face: the NumPy construction's index -1 takes the last cell
            const int c0 = faces_cellid[2 * (size_t)f], c1r = faces_cellid[2 * (size_t)f + 1];
            const int c1 = c1r < 0 ? C + c1r : c1r;
            const V3 v1 = cc[c0];
            const V3 v2 = name[f] == 0 ? cc[c1] : V3{gc[4 * (size_t)f], gc[4 * (size_t)f + 1], gc[4 * (size_t)f + 2]};
            const V3 s1 = v3sub(v2, P(i2)), s2 = v3sub(P(i4), P(i2)), s3 = v3sub(v1, P(i2));
            const V3 n1 = v3add(v3scale(0.5, v3cross(s1, s2)), v3scale(0.5, v3cross(s2, s3)));
            const V3 s4 = v3sub(v2, P(i3)), s5 = v3sub(P(i1), P(i3)), s6 = v3sub(v1, P(i3));
            const V3 n2 = v3add(v3scale(0.5, v3cross(s4, s5)), v3scale(0.5, v3cross(s5, s6)));
            const V3 n = nrm[f];
            const double ar = v3dot(n, v3sub(v2, v1));
            air[f] = ar;
            p1[f] = v3dot(n1, n) / ar;
            p2[f] = v3dot(n2, n) / ar;
            p3[f] = v3dot(n, n) / ar;
            n1v[f] = n1;
            n2v[f] = n2;
            auto proj = [&](int K) {
                const double vx = cc[K].x - cf[f].x, vy = cc[K].y - cf[f].y;
                const double s = vx * n.x + vy * n.y;
                const double px = cc[K].x - s * n.x, py = cc[K].y - s * n.y;
                const double dx = cc[K].x - px, dy = cc[K].y - py;
                return sqrt(dx * dx + dy * dy);
            };
            if (name[f] == 0) dort[f] = proj(c0) + proj(c1);
            else if (name[f] >= 1 && name[f] <= 6) dort[f] = 2 * proj(c0);
            else dort[f] = 0.;
        }
    });
    return MFV_OK;
}
