"""ctypes binding of libmfv.so (declarations in include/mfv.h).

The library is loaded from the package directory only; there is no fallback. A
missing or unloadable library raises ``MfvError`` at first use, so a run can
never silently compute on the CPU.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libmfv.so')

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
P = ctypes.c_void_p  # every array argument is a raw device address


class MfvError(RuntimeError):
    pass


class MfvMesh(ctypes.Structure):
    """Mirror of ``struct mfv_mesh`` (include/mfv.h)."""
    _fields_ = [
        ('dim', ctypes.c_int32), ('nb_cells', ctypes.c_int32), ('nb_faces', ctypes.c_int32),
        ('max_cell_faces', ctypes.c_int32), ('lsq_width', ctypes.c_int32), ('ghost_width', ctypes.c_int32),
        ('face_cells', P), ('face_name', P), ('face_normal', P), ('face_mesure', P), ('face_center', P),
        ('face_ghost', P), ('cell_center', P), ('cell_volume', P), ('cell_faces', P), ('cell_nfaces', P),
        ('order_split', P), ('order_sorted', P), ('lsq_cells', P), ('lsq_ncells', P), ('lsq_ghosts', P),
        ('lsq_nghosts', P), ('bnd_faces', P), ('nb_bnd', ctypes.c_int32), ('pad_', ctypes.c_int32),
    ]


class MfvSwTiles(ctypes.Structure):
    """Mirror of ``struct mfv_sw_tiles`` (include/mfv.h)."""
    _fields_ = [('nb_tiles', ctypes.c_int32), ('tile_size', ctypes.c_int32), ('max_halo', ctypes.c_int32),
                ('max_ext', ctypes.c_int32), ('nb_items', ctypes.c_int32), ('stage_states', ctypes.c_int32),
                ('max_items', ctypes.c_int32), ('pad_', ctypes.c_int32),
                ('tile_cell0', P), ('tile_ncells', P), ('tile_nhalo', P), ('ext_ptr', P), ('ext_cells', P),
                ('item_ptr', P), ('lnbr', P), ('lst', P), ('lgh', P), ('lfn', P)]


class MfvSwTilesHost(ctypes.Structure):
    """Mirror of ``struct mfv_sw_tiles_host`` (include/mfv.h)."""
    _fields_ = [(k, ctypes.c_int32) for k in ('nb_tiles', 'tile_size', 'max_halo', 'max_ext', 'nb_items', 'max_items',
                                              'ext_len', 'MF', 'K', 'G')] + \
        [(k, P) for k in ('tile_cell0', 'tile_ncells', 'tile_nhalo', 'ext_ptr', 'ext_cells', 'item_ptr', 'lnbr', 'lst',
                          'lgh', 'lfn')]


class MfvFaceTiles(ctypes.Structure):
    """Mirror of ``struct mfv_face_tiles`` (include/mfv.h)."""
    _fields_ = [(k, ctypes.c_int32) for k in ('nb_tiles', 'tile_size', 'max_items', 'max_faces')] + \
        [(k, P) for k in ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'face_ptr', 'face_id', 'face_l',
                          'face_r', 'lface')]


class MfvFaceTilesHost(ctypes.Structure):
    """Mirror of ``struct mfv_face_tiles_host`` (include/mfv.h)."""
    _fields_ = [(k, ctypes.c_int32) for k in ('nb_tiles', 'tile_size', 'max_items', 'max_faces', 'halo_len',
                                              'face_len', 'MF', 'C')] + \
        [(k, P) for k in ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'face_ptr', 'face_id', 'face_l',
                          'face_r', 'lface')]


class MfvEulerTet(ctypes.Structure):
    """Mirror of ``struct mfv_euler_tet`` (include/mfv.h)."""
    _fields_ = [('nb_cells', ctypes.c_int32), ('nb_faces', ctypes.c_int32)] + \
        [(k, P) for k in ('cell_nb', 'cell_rank', 'cell_geo')]


class MfvEulerTetHost(ctypes.Structure):
    """Mirror of ``struct mfv_euler_tet_host`` (include/mfv.h)."""
    _fields_ = [('nb_cells', ctypes.c_int32), ('nb_faces', ctypes.c_int32)] + \
        [(k, P) for k in ('cell_nb', 'cell_rank', 'cell_geo')]


class MfvEulerTetTiles(ctypes.Structure):
    """Mirror of ``struct mfv_euler_tet_tiles`` (include/mfv.h)."""
    _fields_ = [(k, ctypes.c_int32) for k in ('nb_tiles', 'tile_size', 'max_items', 'max_faces')] + \
        [(k, P) for k in ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'face_ptr', 'face_geo', 'face_lr',
                          'cell_code')]


class MfvEulerArrays(ctypes.Structure):
    """struct mfv_euler_arrays: six device arrays rho, P, rhou, rhov, rhow, rhoE."""
    _fields_ = [(k, P) for k in ('rho', 'P', 'rhou', 'rhov', 'rhow', 'rhoE')]


class MfvEulerRez(ctypes.Structure):
    """struct mfv_euler_rez: five device residual arrays rho, rhou, rhov, rhow, rhoE."""
    _fields_ = [(k, P) for k in ('rho', 'rhou', 'rhov', 'rhow', 'rhoE')]


class MfvDomainTopology(ctypes.Structure):
    """struct mfv_domain_topology: host tables of the mesh connectivity (malloc'ed)."""
    _fields_ = [(k, ctypes.c_int32) for k in ('nb_faces', 'max_cell_faces', 'max_face_nodes', 'node_cells_width',
                                              'cellnid_width')] + \
               [(k, P) for k in ('faces_nodeid', 'cells_faceid', 'faces_cellid', 'cells_cellfid', 'nodes_cellid',
                                 'cells_cellnid')]


class MfvDomainGeometry2d(ctypes.Structure):
    """struct mfv_domain_geometry_2d: host arrays of the 2D Domain geometry (malloc'ed)."""
    _fields_ = [('ghost_width', ctypes.c_int32), ('ghostnid_width', ctypes.c_int32)] + \
               [(k, P) for k in ('cell_center', 'cell_volume', 'face_name', 'face_normal', 'face_mesure',
                                 'face_center', 'cell_nf', 'face_ghostcenter', 'dist_ortho', 'airDiamond', 'param1',
                                 'param2', 'param3', 'param4', 'f_1', 'f_2', 'f_3', 'f_4', 'lambda_x', 'lambda_y',
                                 'R_x', 'R_y', 'number', 'node_ghostcenter', 'node_ghostfaceinfo', 'node_ghostid',
                                 'cell_ghostnid')]


class MfvDomainGeometry3d(ctypes.Structure):
    """struct mfv_domain_geometry_3d: host arrays of the 3D Domain geometry (malloc'ed)."""
    _fields_ = [('ghost_width', ctypes.c_int32), ('ghostnid_width', ctypes.c_int32)] + \
               [(k, P) for k in ('cell_center', 'cell_volume', 'face_name', 'face_normal', 'face_mesure',
                                 'face_center', 'face_tangent', 'face_binormal', 'face_nodeid', 'face_ghostcenter',
                                 'dist_ortho', 'airDiamond', 'param1', 'param2', 'param3', 'f_1', 'f_2', 'lambda_x',
                                 'lambda_y', 'lambda_z', 'R_x', 'R_y', 'R_z', 'number', 'node_ghostcenter',
                                 'node_ghostfaceinfo', 'node_ghostid', 'cell_ghostnid')]


class MfvSwPacked(ctypes.Structure):
    """Mirror of ``struct mfv_sw_packed`` (include/mfv.h)."""
    _fields_ = [('nb_cells', ctypes.c_int32), ('max_cell_faces', ctypes.c_int32), ('cellrec', P), ('facerec', P),
                ('cfrec', P), ('cfgeo', P), ('nbr', P), ('order', P)]


class MfvSwMixed(ctypes.Structure):
    """Mirror of ``struct mfv_sw_mixed`` (include/mfv.h)."""
    _fields_ = [('nb_cells', ctypes.c_int32), ('max_cell_faces', ctypes.c_int32), ('state', P), ('vol', P),
                ('facerec', P), ('geoA', P), ('geoB', P), ('nbr', P), ('order', P)]


class MfvDiamond(ctypes.Structure):
    """Mirror of ``struct mfv_diamond`` (include/mfv.h)."""
    _fields_ = [('nb_cells', ctypes.c_int32), ('nb_faces', ctypes.c_int32), ('nb_nodes', ctypes.c_int32),
                ('max_cell_faces', ctypes.c_int32), ('dim', ctypes.c_int32), ('cell_faces', P), ('cell_nfaces', P),
                ('order_sorted', P),
                ('cell_volume', P), ('face_cells', P), ('face_nodes', P), ('face_param', P), ('face_fk', P),
                ('face_air', P), ('face_normal', P), ('node_cell_ptr', P), ('node_cell', P),
                ('node_cell_alpha', P), ('node_ghost_ptr', P), ('node_ghost_face', P), ('node_ghost_owner', P),
                ('node_ghost_alpha', P)]


class MfvEllBatch(ctypes.Structure):
    """Mirror of ``struct mfv_ell_batch`` (include/mfv.h)."""
    _fields_ = [('nb_slices', ctypes.c_int32), ('width', ctypes.c_int32), ('nb_inst', ctypes.c_int32),
                ('pad_', ctypes.c_int32), ('slice_inst', P), ('slice_w', P), ('cols', P), ('vals', P),
                ('vals32', P), ('diag', P), ('inst_slice0', P), ('inst_nslices', P), ('live_slices', P),
                ('n_live_slices', P), ('live_inst', P), ('n_live_inst', P), ('cols16', P), ('slice_narrow', P)]


class MfvKrylovWork(ctypes.Structure):
    """Mirror of ``struct mfv_krylov_work`` (include/mfv.h)."""
    _fields_ = [('x', P), ('b', P), ('d', P), ('r', P), ('rh', P), ('p', P), ('v', P), ('yz', P), ('t', P),
                ('dinv', P), ('partial', P), ('scal', P), ('status', P)]


class MfvSwBatch(ctypes.Structure):
    """Mirror of ``struct mfv_sw_batch`` (include/mfv.h)."""
    _fields_ = [('nb_inst', ctypes.c_int32), ('nb_waves', ctypes.c_int32), ('wave_inst', P), ('inst_cell0', P),
                ('inst_ncells', P), ('rings', P)]


# per-instance solver states and scalar slots (MFV_KS_* in include/mfv.h)
KS_IDLE, KS_START, KS_INNER, KS_CORRECT, KS_DONE, KS_FAILED = 0, 1, 2, 3, 4, 5
KS_SIGMA, KS_TAU, KS_TOL_OUT, KS_TOL_IN, KS_MAXIT_IN, KS_MAXIT_OUT = 0, 1, 2, 3, 4, 5
KS_RHO, KS_ALPHA, KS_OMEGA, KS_BETA, KS_R0N2, KS_RN2, KS_BN2, KS_RESN2 = 6, 7, 8, 9, 10, 11, 12, 13
KS_IT, KS_OUTER, KS_ITTOT, KS_NSCAL = 14, 15, 16, 24

MESH_P = ctypes.POINTER(MfvMesh)
DIA_P = ctypes.POINTER(MfvDiamond)
ELL_P = ctypes.POINTER(MfvEllBatch)
KW_P = ctypes.POINTER(MfvKrylovWork)
SWB_P = ctypes.POINTER(MfvSwBatch)
SWP_P = ctypes.POINTER(MfvSwPacked)
SWM_P = ctypes.POINTER(MfvSwMixed)
SWT_P = ctypes.POINTER(MfvSwTiles)
FT_P = ctypes.POINTER(MfvFaceTiles)
ET_P = ctypes.POINTER(MfvEulerTet)
ETT_P = ctypes.POINTER(MfvEulerTetTiles)
DP = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (restype is int unless listed in _RESTYPE)
_SIGNATURES = {
    'mfv_version': [],
    'mfv_last_error': [],
    'mfv_device_count': [ctypes.POINTER(c_int)],
    'mfv_ghost_value': [MESH_P, c_int, P, P, P, c_int, P],
    'mfv_facetocell': [MESH_P, P, P, P],
    'mfv_cell_gradient': [MESH_P, P, P, P, P, P, P],
    'mfv_barthlimiter': [MESH_P, P, P, P, P, P, P, P],
    'mfv_time_step': [MESH_P, P, P, P, c_double, P, DP, P],
    'mfv_explicitscheme_convective': [MESH_P, P, P, P, P, P, P, P, P, P, P, c_int, P],
    'mfv_update_new_value': [MESH_P, P, P, P, P, c_double, P, P],
    'mfv_dt_ring_init': [P, P, P],
    'mfv_advection_step': [MESH_P, P, P, P, P, P, P, P, P, c_double, c_int, P, c_int, P, P],
    'mfv_time_step_sw': [MESH_P, P, P, P, c_double, c_double, c_double, c_double, P, DP, P],
    'mfv_sw_gradients': [MESH_P, P, P, P, P, P, P],
    'mfv_sw_residuals': [MESH_P, P, P, P, P, P, P, c_double, c_int, P, P, P],
    'mfv_update_sw': [MESH_P, P, P, P, P, P, P, P],
    'mfv_explicitscheme_convective_sw': [MESH_P] + [P] * 20 + [c_double, c_int, P],
    'mfv_term_source_srnh_sw': [MESH_P, P, P, P, P, P, P, P, P, P, P, c_int, P, c_int, P, P, c_double, c_int, P],
    'mfv_update_sw5': [MESH_P] + [P] * 18 + [c_double, c_double, c_double, P, P],
    'mfv_sw_step_fused': [MESH_P, P, P, P, P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_dt_reset': [P, P],
    'mfv_term_friction_sw': [c_int, P, P, P, c_double, c_double, c_double, P],
    'mfv_term_coriolis_sw': [c_int, P, P, P, P, P, c_double, P],
    'mfv_term_wind_sw': [c_int, P, P, P, P, P],
    'mfv_sw_pack': [MESH_P, SWP_P, P, P, P, P, P, c_int, P],
    'mfv_sw_unpack': [MESH_P, SWP_P, P, P, P, P],
    'mfv_sw_recon': [MESH_P, SWP_P, c_int, P],
    'mfv_sw_flux_packed': [MESH_P, SWP_P, c_double, c_double, P, c_int, P],
    'mfv_sw_pack_mixed': [MESH_P, SWM_P, P, P, P, P, P, c_int, P],
    'mfv_sw_unpack_mixed': [MESH_P, SWM_P, P, P, P, P],
    'mfv_sw_recon_mixed': [MESH_P, SWM_P, c_int, P],
    'mfv_sw_flux_mixed': [MESH_P, SWM_P, c_double, c_double, P, c_int, P],
    'mfv_sw_conservation': [MESH_P, SWM_P, P, c_int, P, P],
    'mfv_euler3d_ghost': [MESH_P, P, P, c_int, P],
    'mfv_euler3d_time_step': [MESH_P, P, c_double, c_double, P, DP, P],
    'mfv_euler3d_rusanov': [MESH_P, P, P, c_double, c_int, P, P],
    'mfv_euler3d_update': [MESH_P, P, P, c_double, P, P],
    'mfv_ghost_value_tubeschok3d': [MESH_P, ctypes.POINTER(MfvEulerArrays), P, P, P, ctypes.POINTER(MfvEulerArrays),
                                    c_int, P],
    'mfv_time_step_euler_3d': [MESH_P, ctypes.POINTER(MfvEulerArrays), c_double, c_double, P, P],
    'mfv_explicitscheme_euler_3d_rusanov': [MESH_P, ctypes.POINTER(MfvEulerRez), ctypes.POINTER(MfvEulerArrays),
                                            ctypes.POINTER(MfvEulerArrays), c_double, c_int, P],
    'mfv_update_euler_3d_fvc': [MESH_P, ctypes.POINTER(MfvEulerArrays), ctypes.POINTER(MfvEulerRez), c_double,
                                c_double, P],
    'mfv_euler3d_step_fused': [MESH_P, P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_euler3d_step_tiled': [MESH_P, FT_P, P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_face_tiles_build': [c_int, c_int, P, P, c_int, P, c_int, ctypes.POINTER(MfvFaceTilesHost)],
    'mfv_face_tiles_host_free': [ctypes.POINTER(MfvFaceTilesHost)],
    'mfv_euler_tet_build': [c_int, c_int, P, P, P, P, P, ctypes.POINTER(MfvEulerTetHost)],
    'mfv_euler_tet_host_free': [ctypes.POINTER(MfvEulerTetHost)],
    'mfv_euler3d_pack': [MESH_P, P, P, P],
    'mfv_euler3d_unpack': [MESH_P, P, P, P],
    'mfv_euler3d_step_packed': [MESH_P, ET_P, P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_euler3d_step_packed_tiled': [MESH_P, ET_P, ETT_P, P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_euler3d_step_cells_tiled': [MESH_P, ET_P, ETT_P, P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_euler32_pack': [MESH_P, P, P, P],
    'mfv_euler32_unpack': [MESH_P, P, P, P],
    'mfv_euler32_tile_geometry': [ETT_P, c_int, P, P],
    'mfv_euler3d_step32_tiled': [MESH_P, ET_P, ETT_P, P, P, P, c_double, c_double, c_int, P, c_int, P, P],
    'mfv_diamond_widths': [DIA_P, P, P, P, P, P],
    'mfv_diamond_assemble': [DIA_P, P, P, c_int, c_int, P, P, P, P, P, P],
    'mfv_diamond_rhs': [DIA_P, P, P, P, P, P, P],
    'mfv_centertovertex': [DIA_P, P, P, P, P],
    'mfv_face_gradient': [DIA_P, c_int, P, P, P, P, P, P, P, P, P, P],
    'mfv_face_gradient_3d': [DIA_P, c_int, P, P, P, P, P, P, P, P, P, P, P],
    'mfv_explicitscheme_dissipative': [MESH_P, P, P, P, c_double, c_double, c_double, P, P],
    'mfv_time_step_diffusion': [MESH_P, P, P, P, c_double, c_double, c_double, c_double, P, DP, P],
    'mfv_sw_batch_dt_init': [MESH_P, SWM_P, SWB_P, c_double, c_double, P],
    'mfv_sw_batch_step': [MESH_P, SWM_P, SWB_P, c_double, c_double, c_int, c_int, c_int, P],
    'mfv_sw_step_mixed_tiled': [MESH_P, SWM_P, SWT_P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_sw_step_packed_tiled': [MESH_P, SWP_P, SWT_P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_sw_step_packed_tiled_noalloc': [MESH_P, SWP_P, SWT_P, P, c_double, c_double, c_int, P, c_int, P],
    'mfv_sw_tiles_build': [c_int, c_int, P, c_int, P, P, c_int, P, P, P, c_int, P, P, c_int, c_int,
                           ctypes.POINTER(MfvSwTilesHost)],
    'mfv_sw_tiles_host_free': [ctypes.POINTER(MfvSwTilesHost)],
    'mfv_sw_batch_step_tiled': [MESH_P, SWM_P, SWB_P, SWT_P, P, c_double, c_double, c_int, c_int, P],
    'mfv_sw_batch_graph_tiled': [MESH_P, SWM_P, SWB_P, SWT_P, P, c_double, c_double, c_int, c_int,
                                 ctypes.POINTER(c_void_p)],
    'mfv_sw_batch_conservation': [MESH_P, SWM_P, SWB_P, P, P],
    'mfv_sw_batch_graph': [MESH_P, SWM_P, SWB_P, c_double, c_double, c_int, c_int, c_int, ctypes.POINTER(c_void_p)],
    'mfv_graph_launch': [c_void_p, P],
    'mfv_graph_destroy': [c_void_p],
    'mfv_domain_topology_build': [c_int, c_int, c_int, P, c_int, c_int, ctypes.POINTER(MfvDomainTopology)],
    'mfv_domain_topology_free': [ctypes.POINTER(MfvDomainTopology)],
    'mfv_domain_geometry_2d_build': [c_int, c_int, P, c_int, P, c_int, P, c_int, P, P, P, c_int,
                                     ctypes.POINTER(MfvDomainGeometry2d)],
    'mfv_domain_geometry_2d_free': [ctypes.POINTER(MfvDomainGeometry2d)],
    'mfv_domain_geometry_3d_build': [c_int, c_int, P, c_int, P, c_int, c_int, P, P, P, c_int,
                                     ctypes.POINTER(MfvDomainGeometry3d)],
    'mfv_domain_geometry_3d_free': [ctypes.POINTER(MfvDomainGeometry3d)],
    'mfv_krylov_tick': [ELL_P, KW_P, c_int, c_int, c_int, P],
    'mfv_krylov_graph': [ELL_P, KW_P, c_int, c_int, c_int, ctypes.POINTER(c_void_p)],
    'mfv_ell_spmv': [ELL_P, P, P, c_int, P],
    'mfv_ell_narrow_columns': [ELL_P, c_int, c_int, P, P, P],
}
_RESTYPE = {'mfv_version': c_char_p, 'mfv_last_error': c_char_p, 'mfv_sw_tiles_host_free': None,
            'mfv_face_tiles_host_free': None, 'mfv_euler_tet_host_free': None, 'mfv_domain_topology_free': None,
            'mfv_domain_geometry_2d_free': None, 'mfv_domain_geometry_3d_free': None}

_lib = None


def exported_symbols():
    return sorted(_SIGNATURES)


def load():
    """Load libmfv.so and bind every entry point of include/mfv.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MfvError("libmfv.so not built (%s): run __graft_entry__.build() or "
                       "python manapy_amd/build.py" % LIB_PATH)
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as exc:
        raise MfvError("cannot load %s: %s" % (LIB_PATH, exc))
    for name, args in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, c_int)
    _lib = lib
    return lib


def call(name, *args):
    """Call an int-returning entry point; raise MfvError with mfv_last_error() on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if _RESTYPE.get(name, c_int) is None:  # void entry points
        return 0
    if rc != 0:
        msg = lib.mfv_last_error().decode(errors='replace')
        raise MfvError("%s failed (%d): %s" % (name, rc, msg))
    return rc
