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
