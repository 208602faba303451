/*
 * mfv.h -- C ABI of libmfv.so, the MI355X (gfx950) kernels of manapy's explicit
 * finite-volume step.
 *
 * Conventions
 *   - Every pointer argument is a DEVICE pointer unless its name ends in _host.
 *   - Indices are int32, floating-point values float64 unless a name says f32.
 *   - Lengths are element counts. No C++ types, no exceptions cross the ABI.
 *   - Every entry point returns 0 on success or a negative MFV_E* code; the
 *     message of the last failure on the calling thread is mfv_last_error().
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream). All
 *     calls are asynchronous on that stream except those with a *_host output,
 *     which synchronise it.
 *   - A mesh descriptor (mfv_mesh) holds device pointers owned by the caller; the
 *     library never frees them.
 *
 * Each kernel-level entry point replaces the reference function cited beside
 * it, with the reference's argument order and in-place semantics (outputs are
 * arguments). The mesh arrays of the reference signature are taken from the
 * descriptor instead of being passed one by one.
 */
#ifndef MFV_H
#define MFV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFV_OK 0
#define MFV_EINVAL (-1)
#define MFV_ENOMEM (-2)
#define MFV_EHIP (-3)
#define MFV_ENOTSUP (-4)

/* Device-resident mesh, manapy layout (ddm/domain.py) regrouped for the GPU:
 * per-cell tables are column-major [slot][cell] so that a wavefront reading slot
 * s of 64 consecutive cells reads 256 contiguous bytes. */
typedef struct mfv_mesh {
    int32_t dim;            /* 2 or 3 */
    int32_t nb_cells;       /* C */
    int32_t nb_faces;       /* F */
    int32_t max_cell_faces; /* MF: columns of cell_faces */
    int32_t lsq_width;      /* K: columns of lsq_cells */
    int32_t ghost_width;    /* G: columns of lsq_ghosts */
    const int32_t *face_cells;   /* [F][2]   faces.cellid (-1 on boundary faces) */
    const int32_t *face_name;    /* [F]      faces.name */
    const double *face_normal;   /* [F][3]   faces.normal, non-unit, out of face_cells[f][0] */
    const double *face_mesure;   /* [F]      faces.mesure */
    const double *face_center;   /* [F][3]   faces.center */
    const double *face_ghost;    /* [F][3]   faces.ghostcenter (x, y[, z]) of boundary faces */
    const double *cell_center;   /* [C][3]   cells.center */
    const double *cell_volume;   /* [C]      cells.volume */
    const int32_t *cell_faces;   /* [MF][C]  cells.faceid (reference order), -1 padded */
    const int32_t *cell_nfaces;  /* [C]      cells.faceid[:, -1] */
    const uint8_t *order_split;  /* [MF][C]  residual order of inner-then-boundary face loops:
                                             slot r -> column of cell_faces */
    const uint8_t *order_sorted; /* [MF][C]  residual order of all-faces-in-id-order loops */
    const int32_t *lsq_cells;    /* [K][C]   cells.cellnid */
    const int32_t *lsq_ncells;   /* [C]      cells.cellnid[:, -1] */
    const int32_t *lsq_ghosts;   /* [G][C]   cells.ghostnid */
    const int32_t *lsq_nghosts;  /* [C]      cells.ghostnid[:, -1] */
    const int32_t *bnd_faces;    /* [nb_bnd] domain.boundaryfaces */
    int32_t nb_bnd;
    int32_t pad_;
} mfv_mesh;

const char *mfv_version(void);
const char *mfv_last_error(void);
int mfv_device_count(int *count_host);

/* ---------------- generic operators ----------------------------------------------- */

/* ghost_value_neumann / _nonslip / _dirichlet (boundary/bc_utils.py:45-68).
 * kind 0: w_ghost[f] = src[cellid[f][0]]; 1: w_ghost[f] = -src[cellid[f][0]];
 * 2: w_ghost[f] = src[f] (src = BC value array). Acts on the faces listed. */
int mfv_ghost_value(const mfv_mesh *m, int kind, const double *src, double *w_ghost,
                    const int32_t *faces, int nfaces, void *stream);

/* facetocell (ast/ast_utils.py:8-18). */
int mfv_facetocell(const mfv_mesh *m, const double *u_face, double *u_c, void *stream);

/* cell_gradient_2d / cell_gradient_3d (ast/functions2d.py:5-104, functions3d.py:4-120),
 * single-subdomain non-periodic terms. */
int mfv_cell_gradient(const mfv_mesh *m, const double *w_c, const double *w_ghost,
                      double *w_x, double *w_y, double *w_z, void *stream);

/* barthlimiter_2d / barthlimiter_3d (ast/functions2d.py:280-325, functions3d.py:364-408). */
int mfv_barthlimiter(const mfv_mesh *m, const double *w_c, const double *w_ghost,
                     const double *w_x, const double *w_y, const double *w_z, double *psi, void *stream);

/* ---------------- advection (solvers/advec) --------------------------------------- */

/* time_step (solvers/advec/fvm_utils.py:293-309): dt = min_K cfl |K| / sum_f |u_K.n_f|.
 * Writes the minimum to dt_dev[0] (device) and, if dt_host != NULL, to *dt_host (sync). */
int mfv_time_step(const mfv_mesh *m, const double *u, const double *v, const double *w, double cfl,
                  double *dt_dev, double *dt_host, void *stream);

/* explicitscheme_convective_2d / _3d (solvers/advec/fvm_utils.py:19-152, :156-290) with the
 * residual formed by a deterministic cell-centric gather in the reference's face order. */
int mfv_explicitscheme_convective(const mfv_mesh *m, double *rez_w, const double *w_c,
                                  const double *w_ghost, const double *u_face, const double *v_face,
                                  const double *w_face, const double *w_x, const double *w_y,
                                  const double *w_z, const double *psi, int order, void *stream);

/* update_new_value (solvers/advec/fvm_utils.py:311-315); dtime read from dt_dev[0]
 * when dt_dev != NULL, else the value argument. */
int mfv_update_new_value(const mfv_mesh *m, double *ne_c, const double *rez, const double *dissip,
                         const double *src, double dtime, const double *dt_dev, void *stream);

/* Fused steps keep dt on the device in a ring of four doubles: step k uses dt_ring[k % 3],
 * min-reduces the next step's dt into dt_ring[(k+1) % 3] (which must hold 1e6 when step k
 * starts), sets dt_ring[(k+2) % 3] = 1e6 for step k+1 and adds its dt to dt_ring[3] (the
 * simulated time). Initialise the ring to {dt_0, 1e6, 1e6, 0} (mfv_dt_ring_init) and pass
 * k = 0, 1, 2, ... */
int mfv_dt_ring_init(double *dt_ring, const double *dt0, void *stream);

/* One fused explicit upwind step for linear advection (time_step + ghost Neumann +
 * explicitscheme_convective(order) + update_new_value, order 1 or 2 with LSQ gradient and
 * Barth-Jespersen limiter). w_in -> w_out. scratch: (dim+1)*C doubles (gradient and limiter),
 * unused at order 1. */
int mfv_advection_step(const mfv_mesh *m, const double *w_in, double *w_out, const double *u_c,
                       const double *v_c, const double *w_c, const double *u_face,
                       const double *v_face, const double *w_face, double cfl, int order,
                       double *dt_ring, int step, double *scratch, void *stream);

/* ---------------- shallow water (solvers/shallowater) ----------------------------- */

/* time_step_SW (solvers/shallowater/fvm_utils.py:16-42). Output as mfv_time_step. */
int mfv_time_step_sw(const mfv_mesh *m, const double *h, const double *hu, const double *hv,
                     double cfl, double grav, double Dxx, double Dyy, double *dt_dev,
                     double *dt_host, void *stream);

/* Limited least-squares gradients of eta = h + Z, hu, hv for the MUSCL shallow-water step:
 * cell_gradient_2d + barthlimiter_2d applied to the three fields with wall ghosts
 * (h Neumann, hu/hv nonslip, Z Neumann; boundary/bc_utils.py:45-55).
 * grad: [9][C] = eta_x, eta_y, eta_psi, hu_x, hu_y, hu_psi, hv_x, hv_y, hv_psi. */
int mfv_sw_gradients(const mfv_mesh *m, const double *h, const double *hu, const double *hv,
                     const double *Z, double *grad, void *stream);

/* Residuals of the MUSCL/Rusanov shallow-water step with hydrostatic reconstruction
 * (the loop structure of explicitscheme_convective_SW, shallowater/fvm_utils.py:344-496;
 * the reference's flux is SRNH, this one is Rusanov). rez/src: [3][C] (h, hu, hv);
 * zgrad: [2][C] unlimited LSQ gradient of Z. */
int mfv_sw_residuals(const mfv_mesh *m, const double *h, const double *hu, const double *hv,
                     const double *Z, const double *grad, const double *zgrad, double grav,
                     int order, double *rez, double *src, void *stream);

/* update_SW (solvers/shallowater/fvm_utils.py:3-14) for h, hu, hv; dt from dt_dev[0]. */
int mfv_update_sw(const mfv_mesh *m, double *h, double *hu, double *hv, const double *rez,
                  const double *src, const double *dt_dev, void *stream);

/* The reference's own shallow-water scheme on five fields (h, hu, hv, hc, Z).
 *
 * explicitscheme_convective_SW + srnh_scheme (solvers/shallowater/fvm_utils.py:344-496,
 * 152-342): SRNH Roe-type flux; MUSCL on h and hc only, both limited by psi (= hc.psi);
 * boundary faces take the ghost state. Residuals are overwritten. */
int mfv_explicitscheme_convective_sw(const mfv_mesh *m, double *rez_h, double *rez_hu, double *rez_hv,
                                     double *rez_hc, double *rez_Z, const double *h_c, const double *hu_c,
                                     const double *hv_c, const double *hc_c, const double *Z_c,
                                     const double *h_ghost, const double *hu_ghost, const double *hv_ghost,
                                     const double *hc_ghost, const double *Z_ghost, const double *h_x,
                                     const double *h_y, const double *hc_x, const double *hc_y,
                                     const double *psi, double grav, int order, void *stream);

/* term_source_srnh_SW (solvers/shallowater/fvm_utils.py:44-148), triangles only.
 * nodeid [C][nw], cellfid [C][fw] (cells.nodeid / cells.cellfid, row-major, count in
 * the last column), nf [C][3][3] (cells.nf), vertex [N][4] (nodes.vertex).
 * quirk = 1 reads the neighbour of face slot j as cellfid[i][j] like the reference
 * (cellfid is compacted, so boundary cells read a shifted neighbour and -1 wraps to the
 * last cell); quirk = 0 uses the cell across face j. src_h, src_hc, src_Z are set to 0. */
int mfv_term_source_srnh_sw(const mfv_mesh *m, double *src_h, double *src_hu, double *src_hv,
                            double *src_hc, double *src_Z, const double *h_c, const double *Z_c,
                            const double *h_ghost, const double *Z_ghost, const int32_t *nodeid, int nw,
                            const int32_t *cellfid, int fw, const double *nf, const double *vertex,
                            double grav, int quirk, void *stream);

/* update_SW (solvers/shallowater/fvm_utils.py:3-14), all five fields with the Coriolis,
 * wind and hc dissipation terms. dt = dt_dev[0] if dt_dev is not NULL, else dtime. */
int mfv_update_sw5(const mfv_mesh *m, double *h_c, double *hu_c, double *hv_c, double *hc_c, double *Z_c,
                   const double *rez_h, const double *rez_hu, const double *rez_hv, const double *rez_hc,
                   const double *rez_Z, const double *src_h, const double *src_hu, const double *src_hv,
                   const double *src_hc, const double *src_Z, const double *dissip_hc,
                   const double *corio_hu, const double *corio_hv, double wind_hu, double wind_hv,
                   double dtime, const double *dt_dev, void *stream);

/* Fused MUSCL/Rusanov step: residuals + update_SW + the next step's time_step_SW.
 * state_in/state_out: [3][C] (h, hu, hv); grad from mfv_sw_gradients(state_in); dt ring as
 * for mfv_advection_step. */
int mfv_sw_step_fused(const mfv_mesh *m, const double *state_in, double *state_out,
                      const double *Z, const double *grad, const double *zgrad, double grav,
                      double cfl, int order, double *dt_ring, int step, void *stream);

/* Packed shallow-water path: the same step as mfv_sw_gradients + mfv_sw_step_fused with
 * the state held in one 64-byte record per cell and the MUSCL values stored per face slot,
 * so that every neighbour access is a contiguous load. Bit-identical results. */
typedef struct mfv_sw_packed {
    int32_t nb_cells;       /* C */
    int32_t max_cell_faces; /* MF (3 or 4), equal to the mesh's */
    double *cellrec;        /* [C][8]     cx, cy, eta, hu, hv, h, Z, vol */
    double *facerec;        /* [C][MF][4] eta_f, hu_f, hv_f, z_f at each face centre */
    double *cfrec;          /* [C][MF][4] rx, ry (face centre - cell centre), z_f, z_f of the
                                          neighbour's side (own z_f on a wall) */
    const double *cfgeo;    /* [C][MF][4] n0, n1, mesure, 0 of the face (faces.normal orientation) */
    const int32_t *nbr;     /* [MF][C]    (j << 3) | (k << 1) | side, -1 on walls: j = cell across
                                          the face, k = the face's slot in j's faceid row,
                                          side = 1 when the cell is faces.cellid[f][1] */
    const uint8_t *order;   /* [C]        slot of the r-th residual contribution in bits 2r..2r+1 */
} mfv_sw_packed;

/* Fills cellrec from (h, hu, hv, Z), and z_f = Z + (order-1)(Z_x rx + Z_y ry) of both sides
 * of every face slot into cfrec. */
int mfv_sw_pack(const mfv_mesh *m, const mfv_sw_packed *p, const double *h, const double *hu,
                const double *hv, const double *Z, const double *zgrad, int order, void *stream);

/* Copies h, hu, hv out of cellrec. */
int mfv_sw_unpack(const mfv_mesh *m, const mfv_sw_packed *p, double *h, double *hu, double *hv,
                  void *stream);

/* Gradients + limiter + MUSCL face values into facerec (as mfv_sw_gradients). */
int mfv_sw_recon(const mfv_mesh *m, const mfv_sw_packed *p, int order, void *stream);

/* Fluxes + update_SW + next dt, in place on cellrec (as mfv_sw_step_fused). */
int mfv_sw_flux_packed(const mfv_mesh *m, const mfv_sw_packed *p, double grav, double cfl,
                       double *dt_ring, int step, void *stream);

/* Mixed-precision shallow-water path: the scheme of the float64 paths with float64 cell
 * states, float32 reconstruction, face records and flux arithmetic on deviations from the cell
 * states, TwoSum-compensated float32 residual gathers and float64 updates, dt and totals.
 * The face record of slot s of cell i is (a, hu_f, hv_f, w): a = face depth minus the depth of
 * the face's left cell, w = that depth (i is the left cell) or h_i minus it (i is the right cell). */
typedef struct mfv_sw_mixed {
    int32_t nb_cells;       /* C */
    int32_t max_cell_faces; /* MF (3 or 4) */
    double *state;          /* [C][4]     h, hu, hv, Z (float64) */
    const double *vol;      /* [C]        cell volumes (float64) */
    float *facerec;         /* [C][MF][4] face records (mfv_sw_recon_mixed) */
    float *geoA;            /* [C][MF][4] n0, n1, zf_R - zf_L, zf_own - Z_i (mfv_sw_pack_mixed) */
    float *geoB;            /* [C][MF][4] face centre - cell centre, |n|, 1/|n| (mfv_sw_pack_mixed) */
    const int32_t *nbr;     /* [MF][C]    as mfv_sw_packed */
    const uint8_t *order;   /* [C]        as mfv_sw_packed */
} mfv_sw_mixed;

/* Loads (h, hu, hv, Z) and builds geoA/geoB from the mesh and the bathymetry gradient zgrad [3][C]. */
int mfv_sw_pack_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, const double *h, const double *hu,
                      const double *hv, const double *Z, const double *zgrad, int order, void *stream);
int mfv_sw_unpack_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, double *h, double *hu, double *hv,
                        void *stream);
int mfv_sw_recon_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, int order, void *stream);
int mfv_sw_flux_mixed(const mfv_mesh *m, const mfv_sw_mixed *p, double grav, double cfl,
                      double *dt_ring, int step, void *stream);

/* Deterministic float64 totals sum_K (h, hu, hv)_K |K| of the mixed state into out3[0..2];
 * partial: workspace of at least 3*ceil(C/256) doubles. */
int mfv_sw_conservation(const mfv_mesh *m, const mfv_sw_mixed *p, double *partial, int n_partial,
                        double *out3, void *stream);

/* Tiles of the fused mixed step. Tile t is the cell run tile_cell0[t] .. + tile_ncells[t] - 1
 * (at most tile_size cells; a multiple of 64 cells except possibly the mesh's last tile),
 * processed by one block of tile_size threads (a tile may hold fewer cells than the block
 * has threads, so that its cells and halo are reconstructed in one pass). Its items are its cells followed by its halo
 * (the tile_nhalo[t] face neighbours outside it); its ext list is its items followed by every
 * other cell their reconstructions read (stencil cells, ghost owners, face neighbours):
 * the halo and the extra cells are ext_cells[ext_ptr[t] .. ext_ptr[t+1]). Item q of tile t
 * is row item_ptr[t] + q of the per-item tables lst [nb_items][KP] (stencil, one row per item;
 * KP = K + 1 rounded up to a multiple of 8 so that a row is whole 16-byte vectors; entry KP - 1
 * holds the item's stencil and ghost counts, lsq_ncells | lsq_nghosts << 8; other unused
 * entries 0) and lgh [G][nb_items] (ghost owners), which hold positions in the ext list, and
 * lfn [MF][nb_items]: nbr of the item's cell with the cell part replaced by the neighbour's
 * position in the ext list ((position << 3) | (k << 1) | side, -1 on walls). lnbr [MF][C] is nbr with the cell part replaced by the
 * neighbour's position in the item list. LDS per block: max_items * MF * 16 bytes, plus
 * max_ext * 48 bytes with stage_states; at most 64 KiB. */
typedef struct mfv_sw_tiles {
    int32_t nb_tiles;
    int32_t tile_size;          /* 128, 256 or 512 */
    int32_t max_halo;
    int32_t max_ext;
    int32_t nb_items;
    int32_t stage_states;       /* 1: copy every cell state a tile reads into LDS first */
    int32_t max_items;          /* max over tiles of cells + halo */
    int32_t pad_;
    const int32_t *tile_cell0;  /* [nb_tiles] */
    const int32_t *tile_ncells; /* [nb_tiles] */
    const int32_t *tile_nhalo;  /* [nb_tiles] */
    const int32_t *ext_ptr;     /* [nb_tiles + 1] */
    const int32_t *ext_cells;   /* [ext_ptr[nb_tiles]] */
    const int32_t *item_ptr;    /* [nb_tiles + 1] */
    const int32_t *lnbr;        /* [MF][C] */
    const uint16_t *lst;        /* [nb_items][MFV_SW_TILES_KP(K)] */
    const uint16_t *lgh;        /* [G][nb_items] */
    const int32_t *lfn;         /* [MF][nb_items] */
} mfv_sw_tiles;

/* Row stride of the tiles' stencil table for a stencil width K: K + 1 (the handles and the
 * counts entry) rounded up to a multiple of 8. */
#define MFV_SW_TILES_KP(K) (((K) + 8) & ~7)

/* Host-side construction of the tiles (no GPU needed). The cell ranges are the mesh
 * ({0}, {C}) or the ensemble instances (each starting at a multiple of 64 and padded to a
 * multiple of 64). Tiles are grown along the cell order by 64-cell groups while they hold at
 * most N cells and their cells + halo fit the T threads of a block. nbr [MF][C] is
 * mfv_sw_mixed.nbr; the stencil tables are those of mfv_mesh (host copies). The arrays of out
 * are malloc'd; release them with mfv_sw_tiles_host_free. No counterpart in the reference. */
typedef struct mfv_sw_tiles_host {
    int32_t nb_tiles, tile_size, max_halo, max_ext, nb_items, max_items, ext_len, MF, K, G;
    int32_t *tile_cell0, *tile_ncells, *tile_nhalo, *ext_ptr, *ext_cells, *item_ptr, *lnbr;
    uint16_t *lst, *lgh;
    int32_t *lfn;
} mfv_sw_tiles_host;

int mfv_sw_tiles_build(int C, int MF, const int32_t *nbr, int K, const int32_t *lsq_cells,
                       const int32_t *lsq_ncells, int G, const int32_t *lsq_ghosts,
                       const int32_t *lsq_nghosts, const int32_t *face_cells, int nranges,
                       const int32_t *range_cell0, const int32_t *range_ncells, int T, int N,
                       mfv_sw_tiles_host *out);
void mfv_sw_tiles_host_free(mfv_sw_tiles_host *t);

/* The float64 packed step fused on tiles (the same mfv_sw_tiles, built from the packed
 * neighbour codes): reads p->cellrec, writes cellrec_out (a second [C][8] buffer holding the
 * same static fields); the fields a tile's reconstructions read and its face values stay in
 * LDS (40 * (max_ext rounded up to even) + 24 * MF * max_items bytes <= 64 KiB). Gives the bits
 * of mfv_sw_recon followed by mfv_sw_flux_packed. No counterpart call in the reference: it
 * replaces explicitscheme_convective_SW + update_SW + time_step_SW in one launch. */
int mfv_sw_step_packed_tiled(const mfv_mesh *m, const mfv_sw_packed *p, const mfv_sw_tiles *t,
                             double *cellrec_out, double grav, double cfl, int order, double *dt_ring,
                             int step, void *stream);

/* mfv_sw_step_packed_tiled with no device allocation of its own. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Its scratch (the per-wave dt partials) is a
 * static device array of MFV_NOALLOC_MAX_WAVES doubles (8 MiB), one per GPU, held until the
 * process ends, so it takes at most 64 * MFV_NOALLOC_MAX_WAVES cells (larger meshes are
 * rejected with MFV_EINVAL). Same bits as mfv_sw_step_packed_tiled. Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other threads) race. */
#define MFV_NOALLOC_MAX_WAVES (1 << 20)
int mfv_sw_step_packed_tiled_noalloc(const mfv_mesh *m, const mfv_sw_packed *p, const mfv_sw_tiles *t,
                                     double *cellrec_out, double grav, double cfl, int order, double *dt_ring,
                                     int step, void *stream);
/* One fused mixed step (reconstruction, fluxes, update, next dt) from p->state into
 * state_out (a second [C][4] buffer); the face records stay in LDS. Gives the same bits as
 * mfv_sw_recon_mixed followed by mfv_sw_flux_mixed. No counterpart call in the reference:
 * it replaces the pair above (explicitscheme_convective_SW + update_SW in one launch). */
int mfv_sw_step_mixed_tiled(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_tiles *t,
                            double *state_out, double grav, double cfl, int order, double *dt_ring,
                            int step, void *stream);

/* Ensemble of independent shallow-water instances concatenated into one mesh (mfv_mesh +
 * mfv_sw_mixed over all cells). Instance i owns cells inst_cell0[i] .. + inst_ncells[i] - 1 with
 * inst_cell0[i] a multiple of 64; cells up to the next multiple of 64 are padding (no faces,
 * volume 1, zero state). Each instance has its own dt ring rings[i][0..3] (as mfv_dt_ring_init). */
typedef struct mfv_sw_batch {
    int32_t nb_inst;
    int32_t nb_waves;             /* nb_cells / 64 */
    const int32_t *wave_inst;     /* [nb_waves] instance of each 64-cell group */
    const int32_t *inst_cell0;    /* [I] */
    const int32_t *inst_ncells;   /* [I] */
    double *rings;                /* [I][4] */
} mfv_sw_batch;

/* rings[i] = {dt_i, 1e6, 1e6, 0} with dt_i the time_step_SW of instance i's packed state. */
int mfv_sw_batch_dt_init(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b, double grav,
                         double cfl, void *stream);
/* One step of every instance (reconstruction + fluxes + update + next dt, per-instance rings);
 * block = threads per block of the two cell kernels (64, 128, 256 or 512). */
int mfv_sw_batch_step(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b, double grav,
                      double cfl, int order, int step, int block, void *stream);
/* Deterministic float64 totals per instance: out[i][0..2] = sum (h, hu, hv) |K|. */
int mfv_sw_batch_conservation(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b,
                              double *out, void *stream);
/* Captures nsteps (a multiple of 3) ensemble steps into an executable graph (kernel nodes only);
 * each mfv_graph_launch advances every instance by nsteps steps. */
int mfv_sw_batch_graph(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b, double grav,
                       double cfl, int order, int nsteps, int block, void **graph_out);
/* Fused tiled ensemble step (tiles never straddle instances) from p->state into state_out,
 * and a graph of nsteps (a multiple of 6) such steps alternating p->state and state2, which
 * leaves the result in p->state. */
int mfv_sw_batch_step_tiled(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b,
                            const mfv_sw_tiles *t, double *state_out, double grav, double cfl,
                            int order, int step, void *stream);
int mfv_sw_batch_graph_tiled(const mfv_mesh *m, const mfv_sw_mixed *p, const mfv_sw_batch *b,
                             const mfv_sw_tiles *t, double *state2, double grav, double cfl, int order,
                             int nsteps, void **graph_out);
int mfv_graph_launch(void *graph, void *stream);
int mfv_graph_destroy(void *graph);

/* term_friction_SW (shallowater/fvm_utils.py:512-529): hu, hv /= 1 + time g eta^2 |u| / h^(4/3)
 * (implicit Manning friction; the reference calls it with the step's dt as time). */
int mfv_term_friction_sw(int C, const double *h, double *hu, double *hv, double grav, double eta, double time,
                         void *stream);
/* term_coriolis_SW (fvm_utils.py:498-503), its signs kept: corio_hu = f hu, corio_hv = -f hv, with
 * f = f_c[i], or f_const in every cell when f_c is NULL. */
int mfv_term_coriolis_sw(int C, const double *hu, const double *hv, double *corio_hu, double *corio_hv,
                         const double *f_c, double f_const, void *stream);
/* term_wind_SW (fvm_utils.py:531-544): Tx, Ty = 1.25 (0.75 + 0.067 |w|) 1e-3 |w| (wx, wy). */
int mfv_term_wind_sw(int C, double *Tx, double *Ty, const double *wind_wx, const double *wind_wy, void *stream);

/* Sets dt[0] = 1e6, the initial value of the reference's dt min-loops. */
int mfv_dt_reset(double *dt, void *stream);

/* ---------------- 3D Euler (solvers/euler/fvm_utils3d.py) ------------------------- */

/* ghost_value_TubeSchok3D (fvm_utils3d.py:441-500); quirk=1 keeps the reference's
 * rhovg overwrite (rhowg untouched), quirk=0 the corrected slip wall.
 * U: [6][C] (rho, P, rhou, rhov, rhow, rhoE); G: [6][F] ghosts. */
int mfv_euler3d_ghost(const mfv_mesh *m, const double *U, double *G, int quirk, void *stream);

/* time_step_euler_3d (fvm_utils3d.py:809-825) min-reduced: dt_dev[0] = min_K dt_c[K]. */
int mfv_euler3d_time_step(const mfv_mesh *m, const double *U, double cfl, double gamma,
                          double *dt_dev, double *dt_host, void *stream);

/* explicitscheme_euler_3d_rusanov (fvm_utils3d.py:1083-1170) + compute_flux_euler_3d_rusanov
 * (:838-886), cell-centric gather in face-id order. rez: [5][C]. */
int mfv_euler3d_rusanov(const mfv_mesh *m, const double *U, const double *G, double gamma,
                        int quirk, double *rez, void *stream);

/* update_euler_3d_fvc (fvm_utils3d.py:1316-1327); dt from dt_dev[0]. */
int mfv_euler3d_update(const mfv_mesh *m, double *U, const double *rez, double gamma,
                       const double *dt_dev, void *stream);

/* The reference's own argument form of the four Euler kernels: the state as six separate
 * device arrays (cells [C], ghosts [F]) and the residual as five, so a caller that holds the
 * reference's arrays passes them without packing them into [6][C]. */
typedef struct mfv_euler_arrays {
    double *rho, *P, *rhou, *rhov, *rhow, *rhoE;
} mfv_euler_arrays;
typedef struct mfv_euler_rez {
    double *rho, *rhou, *rhov, *rhow, *rhoE;
} mfv_euler_rez;

/* ghost_value_TubeSchok3D (fvm_utils3d.py:441-500), with its argument list: ghosts <- cells on
 * the boundary faces (names 1-6); ug, vg, wg (optional, all or none) <- ghost momentum / rho. */
int mfv_ghost_value_tubeschok3d(const mfv_mesh *m, const mfv_euler_arrays *ghosts, double *ug, double *vg,
                                double *wg, const mfv_euler_arrays *cells, int quirk, void *stream);
/* time_step_euler_3d (fvm_utils3d.py:809-825) as the reference computes it: dt_c[K] per cell
 * (cells->rhoE is not read). */
int mfv_time_step_euler_3d(const mfv_mesh *m, const mfv_euler_arrays *cells, double cfl, double gamma,
                           double *dt_c, void *stream);
/* explicitscheme_euler_3d_rusanov (fvm_utils3d.py:1083-1170): same bits as mfv_euler3d_rusanov. */
int mfv_explicitscheme_euler_3d_rusanov(const mfv_mesh *m, const mfv_euler_rez *rez, const mfv_euler_arrays *cells,
                                        const mfv_euler_arrays *ghosts, double gamma, int quirk, void *stream);
/* update_euler_3d_fvc (fvm_utils3d.py:1316-1327) with the host dt the reference passes. */
int mfv_update_euler_3d_fvc(const mfv_mesh *m, const mfv_euler_arrays *cells, const mfv_euler_rez *rez,
                            double gamma, double dtime, void *stream);

/* Fused step: ghosts on the fly + Rusanov fluxes + update + next dt.
 * U_in -> U_out ([6][C]); dt ring as for mfv_advection_step. */
int mfv_euler3d_step_fused(const mfv_mesh *m, const double *U_in, double *U_out, double gamma,
                           double cfl, int quirk, double *dt_ring, int step, void *stream);

/* Face tiles of a first-order cell-centred scheme: tiles of consecutive cells, each with its
 * halo (the face neighbours outside it) and the list of faces that touch its cells, so that a
 * block evaluates every such face once. Items of tile t: its ncells cells from cell0, then its
 * halo halo_cells[halo_ptr[t] ..]. Faces of tile t: face_id[face_ptr[t] ..], ascending, with
 * face_l = item position of faces.cellid[f][0] and face_r = that of faces.cellid[f][1] (-1 on
 * a boundary face). lface[s][c] = (position of face slot s of cell c in its tile's face list)
 * << 1 | (c is faces.cellid[f][1]). No counterpart in the reference: the layout of the tiled
 * steps below. */
typedef struct mfv_face_tiles {
    int32_t nb_tiles;
    int32_t tile_size;          /* cells per tile (and threads per block): 128, 256 or 512 */
    int32_t max_items;          /* max over tiles of cells + halo */
    int32_t max_faces;          /* max over tiles of the face count */
    const int32_t *tile_cell0;  /* [nb_tiles] */
    const int32_t *tile_ncells; /* [nb_tiles] */
    const int32_t *halo_ptr;    /* [nb_tiles + 1] */
    const int32_t *halo_cells;  /* [halo_ptr[nb_tiles]] */
    const int32_t *face_ptr;    /* [nb_tiles + 1] */
    const int32_t *face_id;     /* [face_ptr[nb_tiles]] */
    const int32_t *face_l;      /* [face_ptr[nb_tiles]] */
    const int32_t *face_r;      /* [face_ptr[nb_tiles]] */
    const int32_t *lface;       /* [MF][C] */
} mfv_face_tiles;

/* Host-side construction (no GPU needed) from host copies of cells.faceid ([MF][C] slots,
 * cell_nfaces) and faces.cellid ([F][2]); tiles of T consecutive cells. The arrays of out are
 * malloc'd; release them with mfv_face_tiles_host_free. */
typedef struct mfv_face_tiles_host {
    int32_t nb_tiles, tile_size, max_items, max_faces, halo_len, face_len, MF, C;
    int32_t *tile_cell0, *tile_ncells, *halo_ptr, *halo_cells, *face_ptr, *face_id, *face_l, *face_r, *lface;
} mfv_face_tiles_host;

int mfv_face_tiles_build(int C, int MF, const int32_t *cell_faces, const int32_t *cell_nfaces, int F,
                         const int32_t *face_cells, int T, mfv_face_tiles_host *out);
void mfv_face_tiles_host_free(mfv_face_tiles_host *t);

/* The fused Euler step on face tiles (tetrahedra): every face that touches a tile is evaluated
 * once in LDS and both of its cells in the tile take the flux from there, in the residual
 * order; same bits as mfv_euler3d_step_fused. LDS: 48 * max_items + 40 * max_faces bytes
 * (<= 64 KiB). Replaces explicitscheme_euler_3d_rusanov (fvm_utils3d.py:1083-1170) +
 * update_euler_3d + time_step_euler_3d in one launch. */
int mfv_euler3d_step_tiled(const mfv_mesh *m, const mfv_face_tiles *t, const double *U_in, double *U_out,
                           double gamma, double cfl, int quirk, double *dt_ring, int step, void *stream);

/* Tables of the packed Euler step on tetrahedra, per cell and in the order of the cell's
 * faces by increasing original id (the reference's face-loop order, the residual's summation
 * order): what lies across each face and its geometry, so that a wave reads them with
 * contiguous loads and only the neighbours' records are gathered; plus the slot of each face
 * in cells.faceid (the time step sums over the faces in slot order). */
typedef struct mfv_euler_tet {
    int32_t nb_cells, nb_faces;
    const int32_t *cell_nb;   /* [C][4] k-th face: bit 31 boundary face, bit 30 the cell is the face's right
                               * side, bits 0-29 the cell across (inner face) or faces.name (boundary face) */
    const uint8_t *cell_rank; /* [C] bits 2k..2k+1: slot of the k-th face in cells.faceid */
    const double *cell_geo;   /* [4][2][C][2] k-th face: (normal x, normal y), (normal z, mesure) */
} mfv_euler_tet;

/* Host-side construction (no GPU needed) from host copies of the mfv_mesh tables: cell_faces
 * [4][C], face_cells [F][2], face_name [F], face_normal [F][3], face_mesure [F]. Every cell
 * needs four distinct faces it is a side of; an inner face (name 0) two cells; names are in
 * [0, 2^30) and C < 2^30. The arrays of out are malloc'd; release them with
 * mfv_euler_tet_host_free. */
typedef struct mfv_euler_tet_host {
    int32_t nb_cells, nb_faces;
    int32_t *cell_nb;
    uint8_t *cell_rank;
    double *cell_geo;
} mfv_euler_tet_host;

int mfv_euler_tet_build(int C, int F, const int32_t *cell_faces, const int32_t *face_cells, const int32_t *face_name,
                        const double *face_normal, const double *face_mesure, mfv_euler_tet_host *out);
void mfv_euler_tet_host_free(mfv_euler_tet_host *t);

/* Cell records of the packed Euler step: rec[c][0..5] = rho, P, rhou, rhov, rhow, rhoE of
 * U [6][C], rec[c][6] = the cell volume, rec[c][7] = 0 (64 bytes per cell). */
int mfv_euler3d_pack(const mfv_mesh *m, const double *U, double *rec, void *stream);
int mfv_euler3d_unpack(const mfv_mesh *m, const double *rec, double *U, void *stream);

/* The fused Euler step on packed records (tetrahedra): each neighbour is one 64-byte record
 * and the cell's face tables are read contiguously; the volume travels in the record
 * (rec_out[c][6] = rec_in[c][6]). Same expressions and order as mfv_euler3d_step_fused, so
 * the same bits. Replaces explicitscheme_euler_3d_rusanov (fvm_utils3d.py:1083-1170) +
 * update_euler_3d_fvc (:1316-1327) + time_step_euler_3d (:809-825) in one launch. */
int mfv_euler3d_step_packed(const mfv_mesh *m, const mfv_euler_tet *t, const double *rec_in, double *rec_out,
                            double gamma, double cfl, int quirk, double *dt_ring, int step, void *stream);

/* Face tiles of the packed Euler step (tiles of T consecutive cells, as mfv_face_tiles): the
 * tile's face entries carry their own geometry and item positions, so a tile reads its faces
 * with contiguous loads. */
typedef struct mfv_euler_tet_tiles {
    int32_t nb_tiles;
    int32_t tile_size;          /* cells per tile (and threads per block): 128, 256 or 512 */
    int32_t max_items;          /* max over tiles of cells + halo */
    int32_t max_faces;          /* max over tiles of the face entries */
    const int32_t *tile_cell0;  /* [nb_tiles] */
    const int32_t *tile_ncells; /* [nb_tiles] */
    const int32_t *halo_ptr;    /* [nb_tiles + 1] */
    const int32_t *halo_cells;  /* [halo_ptr[nb_tiles]] */
    const int32_t *face_ptr;    /* [nb_tiles + 1] */
    const double *face_geo;     /* [face_ptr[nb_tiles]][4] faces.normal (3), faces.mesure of each entry */
    const int32_t *face_lr;     /* [face_ptr[nb_tiles]][2] left item; right item (inner face) or -faces.name */
    const int32_t *cell_code;   /* [C][4] k-th face of the cell (increasing original id): entry << 1 | the cell
                                 * is the face's right side */
} mfv_euler_tet_tiles;

/* The packed Euler step on face tiles: the block stages its cells' records in LDS with
 * contiguous loads, evaluates every face entry of the tile once into LDS (a halo cell's record
 * is read by the entry that needs it; any entry order works, the host puts the entries with a
 * halo cell last), then each cell adds its faces' fluxes in the residual order, updates its
 * record and takes its next-dt candidate; same bits as mfv_euler3d_step_packed. LDS:
 * 48 * tile_size + 40 * max_faces bytes (<= 64 KiB). Replaces explicitscheme_euler_3d_rusanov
 * (fvm_utils3d.py:1083-1170) + update_euler_3d_fvc + time_step_euler_3d in one launch. */
int mfv_euler3d_step_packed_tiled(const mfv_mesh *m, const mfv_euler_tet *t, const mfv_euler_tet_tiles *tt,
                                  const double *rec_in, double *rec_out, double gamma, double cfl, int quirk,
                                  double *dt_ring, int step, void *stream);
/* The same step on the same tiles with one cell phase: after staging the records, each cell
 * evaluates its own four faces' fluxes (an inner face of two tile cells is evaluated by both,
 * from the same arguments), so the LDS holds only the records (48 * max_items bytes <= 64 KiB)
 * and tiles of 512 cells fit. Same bits as mfv_euler3d_step_packed_tiled. Replaces the same
 * reference functions. */
int mfv_euler3d_step_cells_tiled(const mfv_mesh *m, const mfv_euler_tet *t, const mfv_euler_tet_tiles *tt,
                                 const double *rec_in, double *rec_out, double gamma, double cfl, int quirk,
                                 double *dt_ring, int step, void *stream);

/* FP32-bulk Euler (the float32 variant of explicitscheme_euler_3d_rusanov + update_euler_3d_fvc
 * + time_step_euler_3d, fvm_utils3d.py:838-886, 1083-1170, 1316-1327, 809-825): float32 records
 * [C][8] = (rho, P, rhou, rhov, rhow, rhoE, |K|, 0) and float32 face geometry per tile entry
 * [entries][4] = (nx, ny, nz, |f|); float32 face fluxes on the same face tiles (the pressure of
 * the momentum flux as 1/2 (P_L - P_R) n, added by both cells: the same scheme in exact
 * arithmetic, since sum_f P_K n_f = 0 on a closed cell, and zero where the pressure is uniform);
 * float64 gathers in the residual order, float64 update, float32 store; float64 dt ring and, when
 * cons is not NULL, float64 conservation sums cons[5] = sum_K (rho, rhou, rhov, rhow, rhoE) |K|
 * of the new state (block partials folded in a fixed order). The CPU model oracle/euler32.c
 * gives the same records and dt bit for bit. LDS: 24 * max_items + 24 * max_faces bytes. */
int mfv_euler32_pack(const mfv_mesh *m, const double *U, float *rec, void *stream);
int mfv_euler32_unpack(const mfv_mesh *m, const float *rec, double *U, void *stream);
int mfv_euler32_tile_geometry(const mfv_euler_tet_tiles *tt, int n_entries, float *geo32, void *stream);
int mfv_euler3d_step32_tiled(const mfv_mesh *m, const mfv_euler_tet *t, const mfv_euler_tet_tiles *tt,
                             const float *geo32, const float *rec_in, float *rec_out, double gamma, double cfl,
                             int quirk, double *dt_ring, int step, double *cons, void *stream);

/* ---------------- diamond scheme, 2D (ast/functions2d.py, solvers/diffusion) --------- */

/* Diamond-scheme geometry. The pseudo-Laplacian weight of a node incidence is
 *   alpha = (1 + lambda_x (x - x_n) + lambda_y (y - y_n)) / (number + lambda_x R_x + lambda_y R_y)
 * (ast/functions2d.py:221, :362), evaluated once per incidence with that expression order. */
typedef struct mfv_diamond {
    int32_t nb_cells, nb_faces, nb_nodes, max_cell_faces;
    int32_t dim;                    /* 2 or 3 */
    const int32_t *cell_faces;      /* [MF][C] as mfv_mesh */
    const int32_t *cell_nfaces;     /* [C] */
    const uint8_t *order_sorted;    /* [MF][C] as mfv_mesh (faces of a cell by ascending id) */
    const double *cell_volume;      /* [C] */
    const int32_t *face_cells;      /* [F][2] */
    const int32_t *face_nodes;      /* 2D [F][2] faces.nodeid[:, :2]; 3D [F][4] nodes A, B, C, D (D = C
                                     * on a triangle) */
    const double *face_param;       /* [F][4] faces.param1..param4 */
    const double *face_fk;          /* 2D [F][8] faces.f_1..f_4 (x, y); 3D [F][6] faces.f_1, f_2 (n1, n2) */
    const double *face_air;         /* [F]    faces.airDiamond */
    const double *face_normal;      /* [F][3] */
    const int32_t *node_cell_ptr;   /* [N+1]  CSR of nodes.cellid */
    const int32_t *node_cell;       /*        incident cells, reference order */
    const double *node_cell_alpha;  /*        alpha of each incident cell centre */
    const int32_t *node_ghost_ptr;  /* [N+1]  CSR of the valid nodes.ghostcenter entries */
    const int32_t *node_ghost_face; /*        boundary face of the ghost centre */
    const int32_t *node_ghost_owner;/*        its owner cell (ghostcenter[..][2]; [..][3] in 3D) */
    const double *node_ghost_alpha; /*        alpha of the ghost centre */
} mfv_diamond;

/* Boundary data of a variable for the diamond scheme:
 *   node_dir[N]      1 if nodes.oldname is in BCdirichlet (search_element(BCdirichlet, ..) == 1)
 *   face_dir_rank[F] position of the face in the variable's dirichletfaces list, -1 if absent */

/* get_triplet_2d + compute_2dmatrix_size (ast/functions2d.py:328-664), or in 3D get_triplet_3d
 * + compute_3dmatrix_size (ast/functions3d.py:410-782), first pass: for each 64-row slice the
 * largest number of distinct columns in its rows -> slice_w[ceil(C/64)]; info[0] |= 1 if a row
 * has more than 32 (2D) or 128 (3D) distinct columns, info[1] = max(info[1], width). */
int mfv_diamond_widths(const mfv_diamond *d, const uint8_t *node_dir, const int32_t *face_dir_rank,
                       int32_t *slice_w, int32_t *info, void *stream);

/* get_triplet_2d / get_triplet_3d, second pass: the matrix rows with the reference's duplicate triplets summed
 * in insertion order, written as ELL slices (entry j of row K at ((K/64)*width + j)*64 + K%64,
 * columns ascending and offset by row0, unused slots (row0+K, 0)), vals32 = (float)vals,
 * diag[K] = A[K][K]. Rows C..64*ceil(C/64)-1 are padding (no entries, diag 0). */
int mfv_diamond_assemble(const mfv_diamond *d, const uint8_t *node_dir, const int32_t *face_dir_rank,
                         int row0, int width, const int32_t *slice_w, int32_t *cols, double *vals,
                         float *vals32, double *diag, void *stream);

/* get_rhs_glob_2d (ast/functions2d.py:740-817) or get_rhs_glob_3d (ast/functions3d.py:865-946):
 * Dirichlet lifting rhs[C]. */
int mfv_diamond_rhs(const mfv_diamond *d, const uint8_t *node_dir, const int32_t *face_dir_rank,
                    const double *Pbordnode, const double *Pbordface, double *rhs, void *stream);

/* centertovertex_2d (ast/functions2d.py:199-277) or centertovertex_3d (ast/functions3d.py:251-362):
 * w_node[N]. */
int mfv_centertovertex(const mfv_diamond *d, const double *w_c, const double *w_ghost, double *w_node,
                       void *stream);

/* kind 0: face_gradient_2d (ast/functions2d.py:108-196); kind 1: compute_P_gradient_2d_diamond
 * (:820-936). face_bc[F] on boundary faces: 0 not in the face lists (left untouched), 1 in the
 * Neumann list, 2 in the Dirichlet list. node_dir/Pbord* are read by kind 1 only. */
int mfv_face_gradient(const mfv_diamond *d, int kind, const double *w_c, const double *w_ghost,
                      const double *w_node, const int8_t *face_bc, const uint8_t *node_dir,
                      const double *Pbordnode, const double *Pbordface, double *wx, double *wy,
                      void *stream);

/* 3D: kind 0 face_gradient_3d (ast/functions3d.py:122-248); kind 1 compute_P_gradient_3d_diamond
 * (:949-1108). face_bc, node_dir, Pbord* as mfv_face_gradient. */
int mfv_face_gradient_3d(const mfv_diamond *d, int kind, const double *w_c, const double *w_ghost,
                         const double *w_node, const int8_t *face_bc, const uint8_t *node_dir,
                         const double *Pbordnode, const double *Pbordface, double *wx, double *wy, double *wz,
                         void *stream);

/* explicitscheme_dissipative (solvers/diffusion/fvm_utils.py:5-26); wz may be NULL (zeros).
 * Needs mesh->order_sorted. */
int mfv_explicitscheme_dissipative(const mfv_mesh *m, const double *wx, const double *wy,
                                   const double *wz, double Dxx, double Dyy, double Dzz,
                                   double *dissip, void *stream);

/* time_step of solvers/diffusion/fvm_utils.py:30-49 (u = v = w = NULL) or of
 * solvers/advecdiff/fvm_utils.py:317-340 (cell velocities). Output as mfv_time_step. */
int mfv_time_step_diffusion(const mfv_mesh *m, const double *u, const double *v, const double *w,
                            double cfl, double Dxx, double Dyy, double Dzz, double *dt_dev,
                            double *dt_host, void *stream);

/* ---------------- batched sparse solves (replaces solvers/ls/system.py:624-744) ------- */

/* Many independent systems (sigma_i I + tau_i A_i) x_i = b_i in one block-diagonal ELL matrix
 * of 64-row slices. Instance i owns the contiguous slices inst_slice0[i] ..
 * inst_slice0[i] + inst_nslices[i] - 1; its columns are global row indices inside its block.
 * live_slices / live_inst list the slices and instances the kernels visit (their counts are
 * device scalars so that a captured graph stays valid as instances join and leave). */
typedef struct mfv_ell_batch {
    int32_t nb_slices;        /* slice capacity S (rows = 64 S) */
    int32_t width;            /* ELL slots per row */
    int32_t nb_inst;          /* instance capacity I */
    int32_t pad_;
    const int32_t *slice_inst;    /* [S] owning instance, -1 if free */
    const int32_t *slice_w;       /* [S] slots used in the slice */
    const int32_t *cols;          /* [S][width][64] */
    const double *vals;           /* [S][width][64] */
    const float *vals32;          /* [S][width][64] */
    const double *diag;           /* [64 S] */
    const int32_t *inst_slice0;   /* [I] */
    const int32_t *inst_nslices;  /* [I] */
    const int32_t *live_slices;   /* [S] */
    const int32_t *n_live_slices; /* device scalar */
    const int32_t *live_inst;     /* [I] */
    const int32_t *n_live_inst;   /* device scalar */
    const int16_t *cols16;        /* [S][width][64] column - row of the narrow slices, or NULL */
    const uint8_t *slice_narrow;  /* [S] 1: the slice's columns are read from cols16, or NULL */
} mfv_ell_batch;

/* Per-instance solver state: status[i] and scal[i][MFV_KS_NSCAL]. The caller sets SIGMA, TAU,
 * TOL_OUT, TOL_IN, MAXIT_IN, MAXIT_OUT, fills x (initial guess) and b (zero on padding rows)
 * and sets status START (which restarts the refinement count OUTER; ITTOT keeps adding up).
 * Ticks then move it through INNER/CORRECT to DONE (||b - op(x)|| <= TOL_OUT ||b||) or FAILED
 * (more than MAXIT_OUT refinements, or a residual or right-hand side that is not finite). */
#define MFV_KS_IDLE 0
#define MFV_KS_START 1
#define MFV_KS_INNER 2
#define MFV_KS_CORRECT 3
#define MFV_KS_DONE 4
#define MFV_KS_FAILED 5

#define MFV_KS_SIGMA 0
#define MFV_KS_TAU 1
#define MFV_KS_TOL_OUT 2
#define MFV_KS_TOL_IN 3
#define MFV_KS_MAXIT_IN 4
#define MFV_KS_MAXIT_OUT 5
#define MFV_KS_RHO 6
#define MFV_KS_ALPHA 7
#define MFV_KS_OMEGA 8
#define MFV_KS_BETA 9
#define MFV_KS_R0N2 10
#define MFV_KS_RN2 11
#define MFV_KS_BN2 12
#define MFV_KS_RESN2 13
#define MFV_KS_IT 14
#define MFV_KS_OUTER 15
#define MFV_KS_ITTOT 16
#define MFV_KS_NSCAL 24

typedef struct mfv_krylov_work {
    double *x, *b;                                  /* [64 S] FP64 solution and right-hand side */
    float *d, *r, *rh, *p, *v, *yz, *t, *dinv;      /* [64 S] FP32 inner-solve vectors */
    double *partial;                                /* [S][4] per-slice partial sums */
    double *scal;                                   /* [I][MFV_KS_NSCAL] */
    int32_t *status;                                /* [I] */
} mfv_krylov_work;

/* One tick: refinement step (x += d, r = b - op(x) in FP64, convergence test) for instances in
 * START/CORRECT, then n_iter BiCGStab iterations (FP32, Jacobi) for instances in INNER.
 * grid_slices / grid_inst are the persistent grid sizes (blocks of 256) of the per-slice and
 * per-instance kernels. Asynchronous; poll status[] between ticks. */
int mfv_krylov_tick(const mfv_ell_batch *e, const mfv_krylov_work *w, int n_iter, int grid_slices,
                    int grid_inst, void *stream);

/* The same tick captured in a graph (replay with mfv_graph_launch, release with
 * mfv_graph_destroy). The descriptors hold capacities and device pointers only and the kernels
 * read the live lists and counts from device memory, so one graph serves every later state
 * of the arena; the tick's result does not depend on the grid sizes. No counterpart in the
 * reference: it lets a serving loop advance all solves with one launch. */
int mfv_krylov_graph(const mfv_ell_batch *e, const mfv_krylov_work *w, int n_iter, int grid_slices,
                     int grid_inst, void **graph_out);

/* 16-bit columns for the slices [slice0, slice0 + nslices) of an assembled ELL batch (no
 * counterpart in the reference: a storage form of the matrix of solvers/ls/system.py:154-171).
 * A slice is narrow when every used slot's column lies within +-32767 of its row (a banded
 * numbering, such as a mesh's row-major or Cuthill-McKee order, keeps most slices narrow); its
 * cols16 = column - row and slice_narrow = 1, and the SpMVs then read 2 bytes per slot instead
 * of 4 (same columns, same results). Other slices get slice_narrow = 0. */
int mfv_ell_narrow_columns(const mfv_ell_batch *e, int slice0, int nslices, int16_t *cols16, uint8_t *slice_narrow,
                           void *stream);

/* y = A x (FP64) on the live slices. */
int mfv_ell_spmv(const mfv_ell_batch *e, const double *x, double *y, int grid_slices, void *stream);

/* ---------------- mesh topology on the host (ddm/domain.py:130-612) ------------------- */

/* The connectivity of a Domain built in C++ (create_2dfaces / create_3dfaces, the face
 * numbering of domain.py:314-315, create_cell_faceid, create_cellsOfFace,
 * create_NeighborCellByFace, create_node_cellid; ddm_utils2d.py:12-97, 482-505,
 * ddm_utils3d.py:766-810) with the reference's numbering and orderings: face rows = nodes in
 * descending order, -1 padding, then the node count (2 in 2D), faces numbered in the
 * lexicographic order of their rows; faces_cellid[f][0] = first (lowest) cell listing f;
 * cellfid in faceid order; nodes_cellid and cells_cellnid ascending. Tables are padded
 * [rows][width + 1] with the count last and -1 padding. Host arrays, malloc'ed; free them with
 * mfv_domain_topology_free. elements: [C][ew] with the vertex count in column ew - 1. The
 * vertex-neighbour table is built only when want_cellnid is set (cellnid_width = -1 otherwise). */
typedef struct mfv_domain_topology {
    int32_t nb_faces, max_cell_faces, max_face_nodes, node_cells_width, cellnid_width;
    int32_t *faces_nodeid;  /* [F][max_face_nodes + 1] */
    int32_t *cells_faceid;  /* [C][max_cell_faces + 1] */
    int32_t *faces_cellid;  /* [F][2] */
    int32_t *cells_cellfid; /* [C][max_cell_faces + 1] */
    int32_t *nodes_cellid;  /* [N][node_cells_width + 1] */
    int32_t *cells_cellnid; /* [C][cellnid_width + 1] */
} mfv_domain_topology;

int mfv_domain_topology_build(int dim, int C, int ew, const int32_t *elements, int N, int want_cellnid,
                              mfv_domain_topology *out);
void mfv_domain_topology_free(mfv_domain_topology *t);

/* The 2D geometry of a Domain from its topology (Compute_2dcentervolumeOfCell,
 * create_info_2dfaces, create_NormalFacesOfCell, face_info_2d ghost mirrors and per-node ghost
 * lists, create_node_ghostid, variables_2d, face_gradient_info_2d, dist_ortho_function;
 * ddm_utils2d.py:99-223, 433-479, 507-744), the same floating-point expressions in the same
 * order as the reference's loops, on host threads. Boundary faces are those named 1-4, in
 * ascending order. Host arrays, malloc'ed; free them with mfv_domain_geometry_2d_free. */
typedef struct mfv_domain_geometry_2d {
    int32_t ghost_width, ghostnid_width;
    double *cell_center;        /* [C][3] */
    double *cell_volume;        /* [C] */
    int32_t *face_name;         /* [F] */
    double *face_normal;        /* [F][3] */
    double *face_mesure;        /* [F] */
    double *face_center;        /* [F][3] */
    double *cell_nf;            /* [C][max_cell_faces][3] */
    double *face_ghostcenter;   /* [F][3] (x, y, gamma; -1 off the boundary) */
    double *dist_ortho;         /* [F] */
    double *airDiamond, *param1, *param2, *param3, *param4; /* [F] */
    double *f_1, *f_2, *f_3, *f_4; /* [F][2] */
    double *lambda_x, *lambda_y, *R_x, *R_y; /* [N] */
    int32_t *number;            /* [N] */
    double *node_ghostcenter;   /* [N][ghost_width][5] */
    double *node_ghostfaceinfo; /* [N][ghost_width][4] */
    int32_t *node_ghostid;      /* [N][ghost_width + 1] */
    int32_t *cell_ghostnid;     /* [C][ghostnid_width + 1] */
} mfv_domain_geometry_2d;

int mfv_domain_geometry_2d_build(int C, int ew, const int32_t *nodeid, int N, const double *vertex, int F,
                                 const int32_t *faces_nodeid, int mf1, const int32_t *cells_faceid,
                                 const int32_t *faces_cellid, const int32_t *node_cellid, int ncw1,
                                 mfv_domain_geometry_2d *out);
void mfv_domain_geometry_2d_free(mfv_domain_geometry_2d *g);

/* The 3D geometry of a Domain from its topology (Compute_3dcentervolumeOfCell,
 * create_info_3dfaces, oriente_3dfacenodeid, face_info_3d ghost mirrors and per-node ghost
 * lists, create_node_ghostid, variables_3d, face_gradient_info_3d, dist_ortho_function;
 * ddm_utils3d.py:87-157, 362-403, 459-527, 530-764, 813-981, ddm_utils2d.py:122-161, 716-744),
 * the same floating-point expressions in the same order as the reference's loops, on host
 * threads. Cells: tetrahedra, pyramids, hexahedra. Boundary faces are those named 1-6, in
 * ascending order. face_nodeid is the face table with each face's nodes re-oriented (the
 * caller's table is not changed). Host arrays, malloc'ed; free them with
 * mfv_domain_geometry_3d_free. */
typedef struct mfv_domain_geometry_3d {
    int32_t ghost_width, ghostnid_width;
    double *cell_center;        /* [C][3] */
    double *cell_volume;        /* [C] */
    int32_t *face_name;         /* [F] */
    double *face_normal;        /* [F][3] */
    double *face_mesure;        /* [F] */
    double *face_center;        /* [F][3] */
    double *face_tangent;       /* [F][3] */
    double *face_binormal;      /* [F][3] */
    int32_t *face_nodeid;       /* [F][fw] oriented */
    double *face_ghostcenter;   /* [F][4] (x, y, z, gamma; -1 off the boundary) */
    double *dist_ortho;         /* [F] */
    double *airDiamond, *param1, *param2, *param3; /* [F] */
    double *f_1, *f_2;          /* [F][3] */
    double *lambda_x, *lambda_y, *lambda_z, *R_x, *R_y, *R_z; /* [N] */
    int32_t *number;            /* [N] */
    double *node_ghostcenter;   /* [N][ghost_width][6] */
    double *node_ghostfaceinfo; /* [N][ghost_width][6] */
    int32_t *node_ghostid;      /* [N][ghost_width + 1] */
    int32_t *cell_ghostnid;     /* [C][ghostnid_width + 1] */
} mfv_domain_geometry_3d;

int mfv_domain_geometry_3d_build(int C, int ew, const int32_t *nodeid, int N, const double *vertex, int F, int fw,
                                 const int32_t *faces_nodeid, const int32_t *faces_cellid, const int32_t *node_cellid,
                                 int ncw1, mfv_domain_geometry_3d *out);
void mfv_domain_geometry_3d_free(mfv_domain_geometry_3d *g);

#ifdef __cplusplus
}
#endif
#endif /* MFV_H */
