// libmfv host-side builder of the fused mixed step's tiles (mfv_sw_tiles).
//
// Tiles are runs of whole 64-cell groups inside each cell range (the mesh, or one ensemble
// instance padded to 64 cells), grown greedily along the cell order: a group joins the current
// tile while the tile holds at most N cells and its cells plus halo (face neighbours outside
// it) fit the block of T threads, so the reconstruction phase runs in one pass. For every
// tile the builder lists the halo and the other cells the reconstructions read (stencil
// cells, ghost owners, face neighbours) and writes the per-item tables of positions in that
// list. Marker arrays indexed by cell make every lookup O(1); tiles are built one after the
// other, so a marker equal to the tile id means "already listed for this tile".
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mfv_common.hpp"

namespace {

struct TileLists {
    int c0, nc;
    std::vector<int32_t> halo, extra;
};

template <class T>
T *copy_out(const std::vector<T> &v)
{
    T *p = static_cast<T *>(malloc(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (p && !v.empty()) memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

extern "C" void mfv_sw_tiles_host_free(mfv_sw_tiles_host *t)
{
    if (!t) return;
    free(t->tile_cell0);
    free(t->tile_ncells);
    free(t->tile_nhalo);
    free(t->ext_ptr);
    free(t->ext_cells);
    free(t->item_ptr);
    free(t->lnbr);
    free(t->lst);
    free(t->lgh);
    free(t->lfn);
    memset(t, 0, sizeof(*t));
}

extern "C" int mfv_sw_tiles_build(int C, int MF, const int32_t *nbr, int K, const int32_t *lsq_cells,
                                  const int32_t *lsq_ncells, int G, const int32_t *lsq_ghosts,
                                  const int32_t *lsq_nghosts, const int32_t *face_cells, int nranges,
                                  const int32_t *range_cell0, const int32_t *range_ncells, int T, int N,
                                  mfv_sw_tiles_host *out)
{
    MFV_REQUIRE(out, "out is NULL");
    memset(out, 0, sizeof(*out));
    MFV_REQUIRE(C > 0 && MF >= 1 && MF <= 4 && K >= 0 && G >= 0, "bad sizes C=%d MF=%d K=%d G=%d", C, MF, K, G);
    MFV_REQUIRE(nbr && lsq_ncells && lsq_nghosts && face_cells && range_cell0 && range_ncells, "NULL array");
    for (int c = 0; c < C; c++)
        MFV_REQUIRE(lsq_ncells[c] >= 0 && lsq_ncells[c] <= K && lsq_ncells[c] < 256 && lsq_nghosts[c] >= 0 &&
                        lsq_nghosts[c] <= G && lsq_nghosts[c] < 256,
                    "cell %d: stencil and ghost counts %d, %d out of [0, %d], [0, %d] or above 255", c,
                    lsq_ncells[c], lsq_nghosts[c], K, G);
    MFV_REQUIRE((K == 0 || lsq_cells) && (G == 0 || lsq_ghosts), "NULL stencil table");
    MFV_REQUIRE(T == 128 || T == 256 || T == 512, "block size %d not in {128, 256, 512}", T);
    MFV_REQUIRE(N >= 64 && N <= T && N % 64 == 0, "cells per tile %d must be a multiple of 64 in [64, %d]", N, T);
    MFV_REQUIRE(nranges >= 1, "no cell range");
    for (int r = 0; r < nranges; r++)
        MFV_REQUIRE(range_cell0[r] >= 0 && range_ncells[r] >= 0 && (int64_t)range_cell0[r] + range_ncells[r] <= C,
                    "range %d out of the mesh", r);
    auto code_at = [&](int s, int c) { return nbr[(size_t)s * C + c]; };

    // 1. tiles: greedy growth by 64-cell groups under the cell and item budgets
    std::vector<int32_t> mark(C, -1), hmark(C, -1);
    std::vector<TileLists> tiles;
    for (int r = 0; r < nranges; r++) {
        const int a = range_cell0[r], n = range_ncells[r];
        int p = a;
        while (p < a + n) {
            const int t = (int)tiles.size();
            int nc = 0, nh = 0;
            std::vector<int32_t> undo_add, undo_del;
            while (p + nc < a + n && nc < N) {
                const int g = std::min(64, a + n - (p + nc));
                undo_add.clear();
                undo_del.clear();
                int nh2 = nh;
                for (int c = p + nc; c < p + nc + g; c++) {
                    mark[c] = t;
                    if (hmark[c] == t) {
                        hmark[c] = -1;
                        undo_del.push_back(c);
                        nh2--;
                    }
                }
                for (int c = p + nc; c < p + nc + g; c++)
                    for (int s = 0; s < MF; s++) {
                        const int code = code_at(s, c);
                        if (code < 0) continue;
                        const int j = code >> 3;
                        if (mark[j] != t && hmark[j] != t) {
                            hmark[j] = t;
                            undo_add.push_back(j);
                            nh2++;
                        }
                    }
                if (nc > 0 && nc + g + nh2 > T) {  // the group does not fit: undo it, close the tile
                    for (int j : undo_add) hmark[j] = -1;
                    for (int c : undo_del) hmark[c] = t;
                    for (int c = p + nc; c < p + nc + g; c++) mark[c] = -1;
                    break;
                }
                nc += g;
                nh = nh2;
            }
            TileLists tl;
            tl.c0 = p;
            tl.nc = nc;
            tiles.push_back(tl);
            p += nc;
        }
    }
    const int nt = (int)tiles.size();
    MFV_REQUIRE(nt >= 1, "no tiles");

    // 2. per tile: halo and extra cells, ascending; positions through pos[] (valid where mark == t)
    std::fill(mark.begin(), mark.end(), -1);
    std::vector<int32_t> pos(C, 0);
    int64_t n_items = 0;
    for (int t = 0; t < nt; t++) {
        TileLists &tl = tiles[t];
        for (int q = 0; q < tl.nc; q++) {
            mark[tl.c0 + q] = t;
            pos[tl.c0 + q] = q;
        }
        for (int q = 0; q < tl.nc; q++)
            for (int s = 0; s < MF; s++) {
                const int code = code_at(s, tl.c0 + q);
                if (code < 0) continue;
                const int j = code >> 3;
                if (mark[j] != t) {
                    mark[j] = t;
                    tl.halo.push_back(j);
                }
            }
        std::sort(tl.halo.begin(), tl.halo.end());
        for (size_t k = 0; k < tl.halo.size(); k++) pos[tl.halo[k]] = tl.nc + (int)k;
        const int ni = tl.nc + (int)tl.halo.size();
        auto want = [&](int c) {
            if (mark[c] != t) {
                mark[c] = t;
                tl.extra.push_back(c);
            }
        };
        for (int q = 0; q < ni; q++) {
            const int c = (q < tl.nc) ? tl.c0 + q : tl.halo[q - tl.nc];
            for (int j = 0; j < lsq_ncells[c]; j++) want(lsq_cells[(size_t)j * C + c]);
            for (int j = 0; j < lsq_nghosts[c]; j++) want(face_cells[2 * (size_t)lsq_ghosts[(size_t)j * C + c]]);
            for (int s = 0; s < MF; s++) {
                const int code = code_at(s, c);
                if (code >= 0) want(code >> 3);
            }
        }
        std::sort(tl.extra.begin(), tl.extra.end());
        for (size_t k = 0; k < tl.extra.size(); k++) pos[tl.extra[k]] = ni + (int)k;
        MFV_REQUIRE(ni + (int64_t)tl.extra.size() < 65535, "tile %d reads more than 65535 cells", t);
        n_items += ni;
    }
    MFV_REQUIRE(n_items < (int64_t)1 << 31, "too many tile items");
    const int I = (int)n_items;

    // 3. tables
    std::vector<int32_t> cell0(nt), ncells(nt), nhalo(nt), eptr(nt + 1, 0), iptr(nt + 1, 0), ext;
    std::vector<int32_t> lnbr((size_t)MF * C, -1);
    const int KP = MFV_SW_TILES_KP(K);  // stencil row stride: whole 16-byte vectors, counts last
    std::vector<uint16_t> lst((size_t)KP * std::max(I, 1), 0);
    std::vector<uint16_t> lgh((size_t)std::max(G, 1) * std::max(I, 1), 0);
    std::vector<int32_t> lfn((size_t)MF * std::max(I, 1), -1);
    int max_halo = 0, max_ext = 0, max_items = 0;
    std::fill(mark.begin(), mark.end(), -1);
    for (int t = 0; t < nt; t++) {
        const TileLists &tl = tiles[t];
        const int nh = (int)tl.halo.size(), ni = tl.nc + nh, ne = ni + (int)tl.extra.size();
        cell0[t] = tl.c0;
        ncells[t] = tl.nc;
        nhalo[t] = nh;
        max_halo = std::max(max_halo, nh);
        max_items = std::max(max_items, ni);
        max_ext = std::max(max_ext, ne);
        eptr[t + 1] = eptr[t] + nh + (int)tl.extra.size();
        iptr[t + 1] = iptr[t] + ni;
        ext.insert(ext.end(), tl.halo.begin(), tl.halo.end());
        ext.insert(ext.end(), tl.extra.begin(), tl.extra.end());
        // positions of this tile's ext list
        for (int q = 0; q < tl.nc; q++) {
            mark[tl.c0 + q] = t;
            pos[tl.c0 + q] = q;
        }
        for (int k = 0; k < nh; k++) {
            mark[tl.halo[k]] = t;
            pos[tl.halo[k]] = tl.nc + k;
        }
        for (size_t k = 0; k < tl.extra.size(); k++) {
            mark[tl.extra[k]] = t;
            pos[tl.extra[k]] = ni + (int)k;
        }
        for (int q = 0; q < ni; q++) {
            const int c = (q < tl.nc) ? tl.c0 + q : tl.halo[q - tl.nc];
            const size_t row = (size_t)iptr[t] + q;
            for (int j = 0; j < lsq_ncells[c]; j++)
                lst[row * KP + j] = (uint16_t)pos[lsq_cells[(size_t)j * C + c]];
            lst[row * KP + KP - 1] = (uint16_t)(lsq_ncells[c] | lsq_nghosts[c] << 8);
            for (int j = 0; j < lsq_nghosts[c]; j++)
                lgh[(size_t)j * I + row] = (uint16_t)pos[face_cells[2 * (size_t)lsq_ghosts[(size_t)j * C + c]]];
            for (int s = 0; s < MF; s++) {
                const int code = code_at(s, c);
                if (code >= 0) lfn[(size_t)s * I + row] = (pos[code >> 3] << 3) | (code & 7);
                if (q < tl.nc && code >= 0) lnbr[(size_t)s * C + c] = (pos[code >> 3] << 3) | (code & 7);
            }
        }
    }
    out->nb_tiles = nt;
    out->tile_size = T;
    out->max_halo = max_halo;
    out->max_ext = max_ext;
    out->nb_items = I;
    out->max_items = max_items;
    out->ext_len = (int32_t)ext.size();
    out->MF = MF;
    out->K = K;
    out->G = G;
    out->tile_cell0 = copy_out(cell0);
    out->tile_ncells = copy_out(ncells);
    out->tile_nhalo = copy_out(nhalo);
    out->ext_ptr = copy_out(eptr);
    out->ext_cells = copy_out(ext);
    out->item_ptr = copy_out(iptr);
    out->lnbr = copy_out(lnbr);
    out->lst = copy_out(lst);
    out->lgh = copy_out(lgh);
    out->lfn = copy_out(lfn);
    if (!out->tile_cell0 || !out->tile_ncells || !out->tile_nhalo || !out->ext_ptr || !out->ext_cells ||
        !out->item_ptr || !out->lnbr || !out->lst || !out->lgh || !out->lfn) {
        mfv_sw_tiles_host_free(out);
        mfv::set_error("tiles: host allocation failed");
        return MFV_ENOMEM;
    }
    return MFV_OK;
}

// ---------------------------------------------------------------------------------------------
// Face tiles (mfv_face_tiles): tiles of T consecutive cells, their halo (face neighbours outside
// the tile, ascending) and the faces touching their cells (ascending), so that a block
// evaluates each such face once.

extern "C" void mfv_face_tiles_host_free(mfv_face_tiles_host *t)
{
    if (!t) return;
    free(t->tile_cell0);
    free(t->tile_ncells);
    free(t->halo_ptr);
    free(t->halo_cells);
    free(t->face_ptr);
    free(t->face_id);
    free(t->face_l);
    free(t->face_r);
    free(t->lface);
    memset(t, 0, sizeof(*t));
}

extern "C" int mfv_face_tiles_build(int C, int MF, const int32_t *cell_faces, const int32_t *cell_nfaces, int F,
                                    const int32_t *face_cells, int T, mfv_face_tiles_host *out)
{
    MFV_REQUIRE(out, "out is NULL");
    memset(out, 0, sizeof(*out));
    MFV_REQUIRE(C > 0 && F > 0 && MF >= 1 && MF <= 8, "bad sizes C=%d F=%d MF=%d", C, F, MF);
    MFV_REQUIRE(cell_faces && cell_nfaces && face_cells, "NULL array");
    MFV_REQUIRE(T == 128 || T == 256 || T == 512, "tile size %d not in {128, 256, 512}", T);
    for (int c = 0; c < C; c++) {
        MFV_REQUIRE(cell_nfaces[c] >= 0 && cell_nfaces[c] <= MF, "cell %d has %d faces (max %d)", c, cell_nfaces[c],
                    MF);
        for (int s = 0; s < cell_nfaces[c]; s++) {
            const int f = cell_faces[(size_t)s * C + c];
            MFV_REQUIRE(f >= 0 && f < F, "cell %d slot %d: face %d out of range", c, s, f);
            const int a = face_cells[2 * (size_t)f], b = face_cells[2 * (size_t)f + 1];
            MFV_REQUIRE((a == c || b == c) && a >= 0 && a < C && b < C, "face %d does not list cell %d", f, c);
        }
    }
    const int nt = (C + T - 1) / T;
    std::vector<int32_t> cell0(nt), ncells(nt), hptr(nt + 1, 0), fptr(nt + 1, 0), halo, fid, fl, fr;
    std::vector<int32_t> lface((size_t)MF * C, -1);
    std::vector<int32_t> cmark(C, -1), pos(C, 0), fmark(F, -1), fpos(F, 0);
    std::vector<int32_t> th, tf;
    int max_items = 0, max_faces = 0;
    for (int t = 0; t < nt; t++) {
        const int c0 = t * T, nc = std::min(T, C - c0);
        cell0[t] = c0;
        ncells[t] = nc;
        for (int q = 0; q < nc; q++) {
            cmark[c0 + q] = t;
            pos[c0 + q] = q;
        }
        th.clear();
        tf.clear();
        for (int q = 0; q < nc; q++) {
            const int c = c0 + q;
            for (int s = 0; s < cell_nfaces[c]; s++) {
                const int f = cell_faces[(size_t)s * C + c];
                if (fmark[f] != t) {
                    fmark[f] = t;
                    tf.push_back(f);
                }
                const int a = face_cells[2 * (size_t)f], b = face_cells[2 * (size_t)f + 1];
                const int o = (a == c) ? b : a;
                if (o >= 0 && cmark[o] != t) {
                    cmark[o] = t;
                    th.push_back(o);
                }
            }
        }
        std::sort(th.begin(), th.end());
        std::sort(tf.begin(), tf.end());
        for (size_t k = 0; k < th.size(); k++) pos[th[k]] = nc + (int)k;
        for (size_t k = 0; k < tf.size(); k++) {
            const int f = tf[k];
            fpos[f] = (int)k;
            const int a = face_cells[2 * (size_t)f], b = face_cells[2 * (size_t)f + 1];
            fid.push_back(f);
            fl.push_back(pos[a]);
            fr.push_back(b >= 0 ? pos[b] : -1);
        }
        for (int q = 0; q < nc; q++) {
            const int c = c0 + q;
            for (int s = 0; s < cell_nfaces[c]; s++) {
                const int f = cell_faces[(size_t)s * C + c];
                lface[(size_t)s * C + c] = (fpos[f] << 1) | (face_cells[2 * (size_t)f + 1] == c ? 1 : 0);
            }
        }
        halo.insert(halo.end(), th.begin(), th.end());
        hptr[t + 1] = (int32_t)halo.size();
        fptr[t + 1] = (int32_t)fid.size();
        max_items = std::max(max_items, nc + (int)th.size());
        max_faces = std::max(max_faces, (int)tf.size());
    }
    out->nb_tiles = nt;
    out->tile_size = T;
    out->max_items = max_items;
    out->max_faces = max_faces;
    out->halo_len = (int32_t)halo.size();
    out->face_len = (int32_t)fid.size();
    out->MF = MF;
    out->C = C;
    out->tile_cell0 = copy_out(cell0);
    out->tile_ncells = copy_out(ncells);
    out->halo_ptr = copy_out(hptr);
    out->halo_cells = copy_out(halo);
    out->face_ptr = copy_out(fptr);
    out->face_id = copy_out(fid);
    out->face_l = copy_out(fl);
    out->face_r = copy_out(fr);
    out->lface = copy_out(lface);
    if (!out->tile_cell0 || !out->tile_ncells || !out->halo_ptr || !out->halo_cells || !out->face_ptr ||
        !out->face_id || !out->face_l || !out->face_r || !out->lface) {
        mfv_face_tiles_host_free(out);
        mfv::set_error("face tiles: host allocation failed");
        return MFV_ENOMEM;
    }
    return MFV_OK;
}
