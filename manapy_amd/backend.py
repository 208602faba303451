"""``HIPBackend``: the drop-in for manapy's ``Backend.compile`` seam.

manapy binds every kernel once with ``backend.compile(func, signature=...)``
(``backends/cpu/backend.py:69-93``) and calls the result with the kernel's own
positional arguments. ``HIPBackend.compile`` looks the function up by name and
returns a callable with the same positional signature that runs the libmfv entry
point on the GPU. Mesh arrays in the argument list (centres, normals, connectivity,
face lists) are taken from the Domain's device copy; their shapes are checked
against the Domain. Field arguments may be device tensors (used in place) or NumPy
arrays (copied to the device, outputs copied back in place), so a reference driver
runs unchanged.

A kernel with no device implementation raises ``NotImplementedError``: there is no
CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from .device import stream_ptr


class _Args(object):
    """Converts reference arguments to device tensors and writes outputs back."""

    def __init__(self, device):
        self.device = device
        self.back = []

    def dev(self, a, out=False):
        import torch
        if torch.is_tensor(a):
            if not a.is_cuda or not a.is_contiguous() or a.dtype != torch.float64:
                raise _lib.MfvError("tensors must be contiguous float64 device tensors (or pass numpy arrays)")
            return a
        arr = np.ascontiguousarray(a, dtype=np.float64)
        t = torch.as_tensor(arr, device=self.device)
        if out:
            self.back.append((a, t))
        return t

    def finish(self):
        import torch
        if self.back:
            torch.cuda.current_stream().synchronize()
        for host, t in self.back:
            host[...] = t.cpu().numpy()
        self.back = []


class HIPBackend(object):
    """GPU backend of a Domain (``backends/base.py:4-6`` interface)."""

    _registry = {}

    def __init__(self, domain):
        self.domain = domain
        self.backend = "hip"
        self.float_precision = 'f8'
        self.int_precision = 'i4'

    @classmethod
    def for_domain(cls, domain):
        return cls(domain)

    @classmethod
    def register(cls, *names):
        def deco(fn):
            for n in names:
                cls._registry[n] = fn
            return fn
        return deco

    def compile(self, func, signature=False, forcedbackend=None):
        name = func if isinstance(func, str) else getattr(func, '__name__', None)
        # kernels that share a name across solver modules are keyed '<module>.<name>'
        # (time_step of solvers/diffusion and solvers/advecdiff differ from solvers/advec's)
        module = '' if isinstance(func, str) else str(getattr(func, '__module__', '') or '')
        parts = module.split('.')
        qual = (parts[-2] + '.' + name) if len(parts) >= 2 else None
        if qual in self._registry:
            name = qual
        impl = self._registry.get(name)
        if impl is None:
            raise NotImplementedError("no device kernel for %r (the GPU backend covers the explicit finite-volume "
                                      "step: %s)" % (name, ', '.join(sorted(self._registry))))
        domain = self.domain

        def kernel(*args, **kwargs):
            return impl(domain, *args, **kwargs)

        kernel.__name__ = name
        kernel.__doc__ = impl.__doc__
        return kernel

    def supported(self):
        return sorted(self._registry)


def _m(domain):
    return domain.device_mesh


def _check_len(a, n, what):
    if len(a) != n:
        raise ValueError("%s has %d rows, the Domain has %d" % (what, len(a), n))


@HIPBackend.register('explicitscheme_convective_2d', 'explicitscheme_convective_3d')
def _explicitscheme_convective(domain, rez_w, w_c, w_ghost, w_halo, u_face, v_face, w_face, w_x, w_y, w_z,
                               wx_halo, wy_halo, wz_halo, psi, psi_halo, centerc, centerf, centerh, centerg,
                               cellidf, normalf, halofid, name, innerfaces, halofaces, boundaryfaces,
                               periodicboundaryfaces, shift, order):
    """solvers/advec/fvm_utils.py:19-152 / 156-290."""
    m = _m(domain)
    _check_len(centerc, m.C, "centerc")
    _check_len(normalf, m.F, "normalf")
    if len(periodicboundaryfaces) or len(halofaces):
        raise NotImplementedError("one subdomain, non-periodic: halo/periodic face lists must be empty")
    A = _Args(m.device)
    rez = A.dev(rez_w, out=True)
    ts = [A.dev(x) for x in (w_c, w_ghost, u_face, v_face, w_face, w_x, w_y, w_z, psi)]
    _lib.call('mfv_explicitscheme_convective', m.ptr, rez.data_ptr(), *[t.data_ptr() for t in ts], int(order),
              stream_ptr())
    A.finish()


@HIPBackend.register('time_step')
def _time_step(domain, u, v, w, cfl, normal, mesure, volume, faceid, dim):
    """solvers/advec/fvm_utils.py:293-309; returns dt (float)."""
    import torch
    m = _m(domain)
    _check_len(faceid, m.C, "faceid")
    A = _Args(m.device)
    ts = [A.dev(x) for x in (u, v, w)]
    dt = torch.zeros(1, dtype=torch.float64, device=m.device)
    h = ctypes.c_double(0.)
    _lib.call('mfv_time_step', m.ptr, *[t.data_ptr() for t in ts], float(cfl), dt.data_ptr(), ctypes.byref(h),
              stream_ptr())
    return h.value


@HIPBackend.register('update_new_value')
def _update_new_value(domain, ne_c, rez_ne, dissip_ne, src_ne, dtime, vol):
    """solvers/advec/fvm_utils.py:311-315."""
    m = _m(domain)
    A = _Args(m.device)
    ne = A.dev(ne_c, out=True)
    ts = [A.dev(x) for x in (rez_ne, dissip_ne, src_ne)]
    _lib.call('mfv_update_new_value', m.ptr, ne.data_ptr(), *[t.data_ptr() for t in ts], float(dtime), None,
              stream_ptr())
    A.finish()


@HIPBackend.register('cell_gradient_2d', 'cell_gradient_3d')
def _cell_gradient(domain, w_c, w_ghost, w_halo, w_haloghost, centerc, cellnid, ghostnid, haloghostnid, halonid,
                   nodecid, periodicn, periodic, centergf, halocenterg, vertexn, centerh, shift, w_x, w_y, w_z):
    """ast/functions2d.py:5-104, functions3d.py:4-120."""
    m = _m(domain)
    m.ensure_lsq()
    _check_len(cellnid, m.C, "cellnid")
    A = _Args(m.device)
    wc, wg = A.dev(w_c), A.dev(w_ghost)
    outs = [A.dev(x, out=True) for x in (w_x, w_y, w_z)]
    _lib.call('mfv_cell_gradient', m.ptr, wc.data_ptr(), wg.data_ptr(), *[t.data_ptr() for t in outs], stream_ptr())
    A.finish()


@HIPBackend.register('barthlimiter_2d', 'barthlimiter_3d')
def _barthlimiter(domain, w_c, w_ghost, w_halo, w_x, w_y, w_z, psi, cellid, faceid, namef, halofid, centerc,
                  centerf):
    """ast/functions2d.py:280-325, functions3d.py:364-408."""
    m = _m(domain)
    _check_len(faceid, m.C, "faceid")
    A = _Args(m.device)
    ts = [A.dev(x) for x in (w_c, w_ghost, w_x, w_y, w_z)]
    p = A.dev(psi, out=True)
    _lib.call('mfv_barthlimiter', m.ptr, *[t.data_ptr() for t in ts], p.data_ptr(), stream_ptr())
    A.finish()


@HIPBackend.register('facetocell')
def _facetocell(domain, u_face, u_c, faceidc, dim):
    """ast/ast_utils.py:8-18."""
    m = _m(domain)
    A = _Args(m.device)
    uf = A.dev(u_face)
    uc = A.dev(u_c, out=True)
    _lib.call('mfv_facetocell', m.ptr, uf.data_ptr(), uc.data_ptr(), stream_ptr())
    A.finish()


def _ghost(kind):
    def run(domain, w_c, w_ghost, cellid, faces, cst, dist):
        import torch
        m = _m(domain)
        A = _Args(m.device)
        src = A.dev(w_c)
        g = A.dev(w_ghost, out=True)
        f = torch.as_tensor(np.asarray(faces, dtype=np.int32) if not torch.is_tensor(faces) else faces,
                            dtype=torch.int32, device=m.device)
        _lib.call('mfv_ghost_value', m.ptr, kind, src.data_ptr(), g.data_ptr(), f.data_ptr(), len(f), stream_ptr())
        A.finish()
    run.__doc__ = "boundary/bc_utils.py:45-68"
    return run


HIPBackend.register('ghost_value_neumann')(_ghost(0))
HIPBackend.register('ghost_value_nonslip')(_ghost(1))
HIPBackend.register('ghost_value_dirichlet')(_ghost(2))


@HIPBackend.register('time_step_SW')
def _time_step_sw(domain, h_c, hu_c, hv_c, cfl, normal, mesure, volume, faceid, grav, Dxx, Dyy):
    """solvers/shallowater/fvm_utils.py:16-42; returns dt (float)."""
    import torch
    m = _m(domain)
    A = _Args(m.device)
    ts = [A.dev(x) for x in (h_c, hu_c, hv_c)]
    dt = torch.zeros(1, dtype=torch.float64, device=m.device)
    h = ctypes.c_double(0.)
    _lib.call('mfv_time_step_sw', m.ptr, *[t.data_ptr() for t in ts], float(cfl), float(grav), float(Dxx), float(Dyy),
              dt.data_ptr(), ctypes.byref(h), stream_ptr())
    return h.value


@HIPBackend.register('update_SW')
def _update_sw(domain, h_c, hu_c, hv_c, hc_c, Z_c, rez_h, rez_hu, rez_hv, rez_hc, rez_Z, src_h, src_hu, src_hv,
               src_hc, src_Z, dissip_hc, corio_hu, corio_hv, wind_hu, wind_hv, dtime, vol):
    """solvers/shallowater/fvm_utils.py:3-14, all five fields with the Coriolis, wind and hc
    dissipation terms (``vol`` is the mesh's cell volume)."""
    m = _m(domain)
    A = _Args(m.device)
    st = [A.dev(x, out=True) for x in (h_c, hu_c, hv_c, hc_c, Z_c)]
    ins = [A.dev(x) for x in (rez_h, rez_hu, rez_hv, rez_hc, rez_Z, src_h, src_hu, src_hv, src_hc, src_Z,
                              dissip_hc, corio_hu, corio_hv)]
    _lib.call('mfv_update_sw5', m.ptr, *[t.data_ptr() for t in st], *[t.data_ptr() for t in ins], float(wind_hu),
              float(wind_hv), float(dtime), None, stream_ptr())
    A.finish()


@HIPBackend.register('term_friction_SW')
def _friction_sw(domain, h_c, hu_c, hv_c, grav, eta, time):
    """solvers/shallowater/fvm_utils.py:512-529 (in place on hu_c, hv_c)."""
    m = _m(domain)
    A = _Args(m.device)
    h = A.dev(h_c)
    hu, hv = A.dev(hu_c, out=True), A.dev(hv_c, out=True)
    _lib.call('mfv_term_friction_sw', int(h.numel()), h.data_ptr(), hu.data_ptr(), hv.data_ptr(), float(grav),
              float(eta), float(time), stream_ptr())
    A.finish()


@HIPBackend.register('term_coriolis_SW')
def _coriolis_sw(domain, hu_c, hv_c, corio_hu, corio_hv, f_c):
    """solvers/shallowater/fvm_utils.py:498-503. ``f_c`` per cell, as the kernel reads it (a
    scalar, as ShallowWaterSolver passes, is taken in every cell)."""
    m = _m(domain)
    A = _Args(m.device)
    hu, hv = A.dev(hu_c), A.dev(hv_c)
    cu, cv = A.dev(corio_hu, out=True), A.dev(corio_hv, out=True)
    if np.ndim(f_c) == 0 and not hasattr(f_c, 'data_ptr'):
        _lib.call('mfv_term_coriolis_sw', int(hu.numel()), hu.data_ptr(), hv.data_ptr(), cu.data_ptr(), cv.data_ptr(),
                  None, float(f_c), stream_ptr())
    else:
        f = A.dev(f_c)
        _lib.call('mfv_term_coriolis_sw', int(hu.numel()), hu.data_ptr(), hv.data_ptr(), cu.data_ptr(), cv.data_ptr(),
                  f.data_ptr(), 0., stream_ptr())
    A.finish()


@HIPBackend.register('term_wind_SW')
def _wind_sw(domain, Tx_wind, Ty_wind, wind_wx, wind_wy):
    """solvers/shallowater/fvm_utils.py:531-544 (the kernel's own signature)."""
    m = _m(domain)
    A = _Args(m.device)
    tx, ty = A.dev(Tx_wind, out=True), A.dev(Ty_wind, out=True)
    wx, wy = A.dev(wind_wx), A.dev(wind_wy)
    _lib.call('mfv_term_wind_sw', int(tx.numel()), tx.data_ptr(), ty.data_ptr(), wx.data_ptr(), wy.data_ptr(),
              stream_ptr())
    A.finish()


def _no_halo(halofaces):
    if len(halofaces) != 0:
        raise NotImplementedError("halo faces: one GPU holds one subdomain (no MPI halo path)")


@HIPBackend.register('explicitscheme_convective_SW')
def _convective_sw(domain, rez_h, rez_hu, rez_hv, rez_hc, rez_Z, h_c, hu_c, hv_c, hc_c, Z_c, h_ghost, hu_ghost,
                   hv_ghost, hc_ghost, Z_ghost, h_halo, hu_halo, hv_halo, hc_halo, Z_halo, h_x, h_y, hx_halo, hy_halo,
                   hc_x, hc_y, hcx_halo, hcy_halo, psi, psi_halo, centerc, centerf, centerh, centerg, cellidf,
                   mesuref, normalf, halofid, innerfaces, halofaces, boundaryfaces, grav, order):
    """solvers/shallowater/fvm_utils.py:344-496 (SRNH flux, srnh_scheme :152-342). The mesh
    arrays are the Domain's own (its device copy is used)."""
    _no_halo(halofaces)
    m = _m(domain)
    A = _Args(m.device)
    rez = [A.dev(x, out=True) for x in (rez_h, rez_hu, rez_hv, rez_hc, rez_Z)]
    ins = [A.dev(x) for x in (h_c, hu_c, hv_c, hc_c, Z_c, h_ghost, hu_ghost, hv_ghost, hc_ghost, Z_ghost, h_x, h_y,
                              hc_x, hc_y, psi)]
    _lib.call('mfv_explicitscheme_convective_sw', m.ptr, *[t.data_ptr() for t in rez], *[t.data_ptr() for t in ins],
              float(grav), int(order), stream_ptr())
    A.finish()


@HIPBackend.register('term_source_srnh_SW')
def _source_srnh(domain, src_h, src_hu, src_hv, src_hc, src_Z, h_c, hu_c, hv_c, hc_c, Z_c, h_ghost, hu_ghost,
                 hv_ghost, hc_ghost, Z_ghost, h_halo, hu_halo, hv_halo, hc_halo, Z_halo, h_x, h_y, psi, hx_halo,
                 hy_halo, psi_halo, nodeidc, faceidc, cellidc, cellidf, centerc, normalc, namef, centerf, centerh,
                 vertexn, halofid, grav, order, quirk=1):
    """solvers/shallowater/fvm_utils.py:44-148 (triangles). ``quirk=1`` reads the neighbour of
    face slot j from ``cellidc`` (cells.cellfid) exactly as the reference does."""
    import torch
    m = _m(domain)
    A = _Args(m.device)
    out = [A.dev(x, out=True) for x in (src_h, src_hu, src_hv, src_hc, src_Z)]
    ins = [A.dev(x) for x in (h_c, Z_c, h_ghost, Z_ghost)]
    ti = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=m.device)
    nid, cfid = ti(nodeidc), ti(cellidc)
    nf = A.dev(np.ascontiguousarray(normalc, dtype=np.float64))
    vx = A.dev(np.ascontiguousarray(vertexn, dtype=np.float64))
    if nf.dim() != 3 or tuple(nf.shape[1:]) != (3, 3) or vx.dim() != 2 or vx.shape[1] != 4:
        raise _lib.MfvError("normalc must be [C][3][3] and vertexn [N][4]")
    _lib.call('mfv_term_source_srnh_sw', m.ptr, *[t.data_ptr() for t in out], *[t.data_ptr() for t in ins],
              nid.data_ptr(), int(nid.shape[1]), cfid.data_ptr(), int(cfid.shape[1]), nf.data_ptr(), vx.data_ptr(),
              float(grav), int(quirk), stream_ptr())
    A.finish()


def _euler_arrays(A, arrays, out=False):
    """mfv_euler_arrays over six reference arrays (device tensors used in place; NumPy arrays
    copied, and written back by A.finish() when ``out``)."""
    ts = [A.dev(x, out=out) for x in arrays]
    return _lib.MfvEulerArrays(*[t.data_ptr() for t in ts]), ts


def _euler_rez(A, arrays, out=False):
    ts = [A.dev(x, out=out) for x in arrays]
    return _lib.MfvEulerRez(*[t.data_ptr() for t in ts]), ts


@HIPBackend.register('explicitscheme_euler_3d_rusanov')
def _euler_rusanov(domain, rez_rho, rez_rhou, rez_rhov, rez_rhow, rez_rhoE, rho_c, P_c, rhou_c, rhov_c, rhow_c,
                   rhoE_c, rho_g, P_g, rhou_g, rhov_g, rhow_g, rhoE_g, rho_h, P_h, rhou_h, rhov_h, rhow_h, rhoE_h,
                   cellidf, halofid, normal, mesurf, name, gamma, tangent, binormal, quirk=1):
    """solvers/euler/fvm_utils3d.py:1083-1170 (quirk=1: the reference's flux, Q1)."""
    m = _m(domain)
    _check_len(cellidf, m.F, "cellidf")
    A = _Args(m.device)
    rez, keep_r = _euler_rez(A, (rez_rho, rez_rhou, rez_rhov, rez_rhow, rez_rhoE), out=True)
    U, keep_u = _euler_arrays(A, (rho_c, P_c, rhou_c, rhov_c, rhow_c, rhoE_c))
    G, keep_g = _euler_arrays(A, (rho_g, P_g, rhou_g, rhov_g, rhow_g, rhoE_g))
    _lib.call('mfv_explicitscheme_euler_3d_rusanov', m.ptr, ctypes.byref(rez), ctypes.byref(U), ctypes.byref(G),
              float(gamma), int(quirk), stream_ptr())
    A.finish()
    del keep_r, keep_u, keep_g


@HIPBackend.register('time_step_euler_3d')
def _euler_time_step(domain, rho, P, rhou, rhov, rhow, cfl, normal, mesure, volume, faceid, gamma, dt_c):
    """solvers/euler/fvm_utils3d.py:809-825: dt_c[K] per cell (the caller takes the min)."""
    m = _m(domain)
    _check_len(faceid, m.C, "faceid")
    A = _Args(m.device)
    ts = [A.dev(x) for x in (rho, P, rhou, rhov, rhow)]
    U = _lib.MfvEulerArrays(*[t.data_ptr() for t in ts], ts[0].data_ptr())
    out = A.dev(dt_c, out=True)
    _lib.call('mfv_time_step_euler_3d', m.ptr, ctypes.byref(U), float(cfl), float(gamma), out.data_ptr(),
              stream_ptr())
    A.finish()
    del ts


@HIPBackend.register('update_euler_3d_fvc')
def _euler_update(domain, rho_c, P_c, rhou_c, rhov_c, rhow_c, rhoE_c, rez_rho, rez_rhou, rez_rhov, rez_rhow,
                  rez_rhoE, gamma, dtime, vol):
    """solvers/euler/fvm_utils3d.py:1316-1327."""
    m = _m(domain)
    _check_len(vol, m.C, "vol")
    A = _Args(m.device)
    U, keep_u = _euler_arrays(A, (rho_c, P_c, rhou_c, rhov_c, rhow_c, rhoE_c), out=True)
    rez, keep_r = _euler_rez(A, (rez_rho, rez_rhou, rez_rhov, rez_rhow, rez_rhoE))
    _lib.call('mfv_update_euler_3d_fvc', m.ptr, ctypes.byref(U), ctypes.byref(rez), float(gamma), float(dtime),
              stream_ptr())
    A.finish()
    del keep_u, keep_r


@HIPBackend.register('ghost_value_TubeSchok3D')
def _euler_ghost(domain, rhog, Pg, rhoug, rhovg, rhowg, ug, vg, wg, rhoEg, rhoc, Pc, rhouc, rhovc, rhowc, rhoEc,
                 cellid, name, normal, mesure, quirk=1):
    """solvers/euler/fvm_utils3d.py:441-500 (quirk=1: the reference's slip ghost, Q2)."""
    m = _m(domain)
    _check_len(cellid, m.F, "cellid")
    A = _Args(m.device)
    G, keep_g = _euler_arrays(A, (rhog, Pg, rhoug, rhovg, rhowg, rhoEg), out=True)
    vel = [A.dev(x, out=True) for x in (ug, vg, wg)]
    U, keep_u = _euler_arrays(A, (rhoc, Pc, rhouc, rhovc, rhowc, rhoEc))
    _lib.call('mfv_ghost_value_tubeschok3d', m.ptr, ctypes.byref(G), *[t.data_ptr() for t in vel], ctypes.byref(U),
              int(quirk), stream_ptr())
    A.finish()
    del keep_g, keep_u


# ---------------------------------------------------------------------------------------------
# diamond scheme (ast/functions2d.py, ast/functions3d.py, solvers/diffusion, solvers/advecdiff);
# the 3D functions share the 2D signatures

def _i32dev(a, device):
    import torch
    if torch.is_tensor(a):
        return a.to(device=device, dtype=torch.int32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)


def _bc_dev(domain, BCdirichlet, dirichletfaces, neumannfaces):
    import torch
    from .diamond import bc_tables
    dev = _m(domain).device
    return tuple(torch.from_numpy(a).to(dev) for a in bc_tables(domain, np.asarray(BCdirichlet),
                                                                 np.asarray(dirichletfaces), np.asarray(neumannfaces)))


def _no_halo(*lists):
    if any(len(x) for x in lists):
        raise NotImplementedError("one subdomain, non-periodic: halo/periodic face lists must be empty")


@HIPBackend.register('compute_2dmatrix_size', 'compute_3dmatrix_size')
def _matrix_size(domain, nodeidf, halofid, cellnid, halonid, periodicnid, centergn, halocentergn, oldnamen,
                 BCdirichlet, matrixinnerfaces, halofaces, dirichletfaces):
    """ast/functions2d.py:582-664 / functions3d.py:410-495 (an index count, evaluated on the
    host)."""
    from .solvers.ls import compute_2dmatrix_size, compute_3dmatrix_size
    _no_halo(halofaces)
    size = compute_2dmatrix_size if int(domain.dim) == 2 else compute_3dmatrix_size
    return size(domain, BCdirichlet, matrixinnerfaces, dirichletfaces)


@HIPBackend.register('get_triplet_2d', 'get_triplet_3d')
def _get_triplet_2d(domain, cellfid, nodeidf, vertexn, halofid, haloext, oldnamen, volume, cellnid, centerc,
                    centerh, halonid, periodicnid, centergn, halocentergn, airDiamond, lambda_x, lambda_y,
                    lambda_z, number, R_x, R_y, R_z, param1, param2, param3, param4, shift, nbelements,
                    loctoglob, BCdirichlet, a_loc, irn_loc, jcn_loc, matrixinnerfaces, halofaces,
                    dirichletfaces):
    """ast/functions2d.py:328-580 / functions3d.py:498-782. The triplets come back with
    duplicates already summed (in
    the reference's insertion order), one per distinct (row, col); the remaining slots of the
    caller's arrays are (0, 0, 0.0), so any sparse constructor that sums duplicates builds the
    same matrix."""
    import torch
    from .krylov import EllBatch
    _no_halo(halofaces)
    m = _m(domain)
    C = m.C
    if not np.array_equal(np.sort(np.asarray(matrixinnerfaces)), np.sort(np.asarray(domain.innerfaces))):
        raise NotImplementedError("matrixinnerfaces must be the domain's inner faces (no periodic faces)")
    node_dir, rank, _ = _bc_dev(domain, BCdirichlet, dirichletfaces, [])
    W = 32 if int(domain.dim) == 2 else domain.device_diamond.width(node_dir, rank)
    batch = EllBatch((C + 63) // 64, max_inst=1, width=W, device=m.device)
    inst = batch.alloc(C)
    batch.assemble_diamond(inst, domain.device_diamond, node_dir, rank)
    W = batch.W
    ns = (C + 63) // 64
    sw = batch.slice_w[:ns].cpu().numpy()
    cols = batch.cols[:ns * W * 64].cpu().numpy().reshape(ns, W, 64)
    vals = batch.vals[:ns * W * 64].cpu().numpy().reshape(ns, W, 64)
    rows = np.broadcast_to(np.arange(ns)[:, None, None] * 64 + np.arange(64)[None, None, :], cols.shape)
    is_diag = cols == rows
    keep = (np.arange(W)[None, :, None] < sw[:, None, None]) & (rows < C) & \
        (~is_diag | (np.cumsum(is_diag, axis=1) == 1))
    v, r, c = vals[keep], rows[keep], cols[keep]
    n = len(v)
    if n > len(a_loc):
        raise ValueError("a_loc holds %d entries, the matrix has %d" % (len(a_loc), n))
    l2g = np.asarray(loctoglob)
    a_loc[:] = 0.
    irn_loc[:] = 0
    jcn_loc[:] = 0
    a_loc[:n] = v
    irn_loc[:n] = l2g[r]
    jcn_loc[:n] = l2g[c]
    return n


@HIPBackend.register('get_rhs_glob_2d', 'get_rhs_glob_3d')
def _get_rhs_glob_2d(domain, cellfid, nodeidf, oldname, volume, centergn, loctoglob, param1, param2, param3, param4,
                     Pbordnode, Pbordface, rhs, BCdirichlet, centergf, matrixinnerfaces, halofaces, dirichletfaces):
    """ast/functions2d.py:740-817 / functions3d.py:865-946."""
    _no_halo(halofaces)
    m = _m(domain)
    dd = domain.device_diamond
    node_dir, rank, _ = _bc_dev(domain, BCdirichlet, dirichletfaces, [])
    A = _Args(m.device)
    pn, pf = A.dev(Pbordnode), A.dev(Pbordface)
    out = A.dev(rhs, out=True)
    _lib.call('mfv_diamond_rhs', dd.ptr, node_dir.data_ptr(), rank.data_ptr(), pn.data_ptr(), pf.data_ptr(),
              out.data_ptr(), stream_ptr())
    A.finish()


@HIPBackend.register('centertovertex_2d', 'centertovertex_3d')
def _centertovertex_2d(domain, w_c, w_ghost, w_halo, w_haloghost, centerc, centerh, cellid, ghostid, haloghostid,
                       periodicid, haloid, vertexn, centergf, halocenterg, R_x, R_y, R_z, lambda_x, lambda_y,
                       lambda_z, number, shift, w_n):
    """ast/functions2d.py:199-277 / functions3d.py:251-362."""
    m = _m(domain)
    dd = domain.device_diamond
    A = _Args(m.device)
    wc, wg = A.dev(w_c), A.dev(w_ghost)
    out = A.dev(w_n, out=True)
    _lib.call('mfv_centertovertex', dd.ptr, wc.data_ptr(), wg.data_ptr(), out.data_ptr(), stream_ptr())
    A.finish()


@HIPBackend.register('face_gradient_2d', 'face_gradient_3d')
def _face_gradient_2d(domain, w_c, w_ghost, w_halo, w_node, cellidf, nodeidf, centergf, halofid, centerc, centerh,
                      vertexn, airDiamond, normalf, f_1, f_2, f_3, f_4, shift, wx_face, wy_face, wz_face, innerfaces,
                      halofaces, dirichletfaces, neumann, periodicfaces):
    """ast/functions2d.py:108-196 / functions3d.py:122-248."""
    _no_halo(halofaces, periodicfaces)
    m = _m(domain)
    dd = domain.device_diamond
    _, _, face_bc = _bc_dev(domain, [], dirichletfaces, neumann)
    A = _Args(m.device)
    wc, wg, wn = A.dev(w_c), A.dev(w_ghost), A.dev(w_node)
    wx, wy = A.dev(wx_face, out=True), A.dev(wy_face, out=True)
    args = [dd.ptr, 0, wc.data_ptr(), wg.data_ptr(), wn.data_ptr(), face_bc.data_ptr(), None, None, None,
            wx.data_ptr(), wy.data_ptr()]
    if int(domain.dim) == 2:
        _lib.call('mfv_face_gradient', *args, stream_ptr())
    else:
        wz = A.dev(wz_face, out=True)
        _lib.call('mfv_face_gradient_3d', *args, wz.data_ptr(), stream_ptr())
    A.finish()


@HIPBackend.register('compute_P_gradient_2d_diamond', 'compute_P_gradient_3d_diamond')
def _p_gradient_2d(domain, P_c, P_ghost, P_halo, P_node, cellidf, nodeidf, centergf, halofid, centerc, centerh,
                   oldname, airDiamond, f_1, f_2, f_3, f_4, normalf, shift, Pbordnode, Pbordface, Px_face, Py_face,
                   Pz_face, BCdirichlet, innerfaces, halofaces, neumannfaces, dirichletfaces, periodicfaces):
    """ast/functions2d.py:820-936 / functions3d.py:949-1108 (f_1, f_2 are the 3D n1, n2)."""
    _no_halo(halofaces, periodicfaces)
    m = _m(domain)
    dd = domain.device_diamond
    node_dir, _, face_bc = _bc_dev(domain, BCdirichlet, dirichletfaces, neumannfaces)
    A = _Args(m.device)
    pc, pg, pn = A.dev(P_c), A.dev(P_ghost), A.dev(P_node)
    bn, bf = A.dev(Pbordnode), A.dev(Pbordface)
    px, py = A.dev(Px_face, out=True), A.dev(Py_face, out=True)
    args = [dd.ptr, 1, pc.data_ptr(), pg.data_ptr(), pn.data_ptr(), face_bc.data_ptr(), node_dir.data_ptr(),
            bn.data_ptr(), bf.data_ptr(), px.data_ptr(), py.data_ptr()]
    if int(domain.dim) == 2:
        _lib.call('mfv_face_gradient', *args, stream_ptr())
    else:
        pz = A.dev(Pz_face, out=True)
        _lib.call('mfv_face_gradient_3d', *args, pz.data_ptr(), stream_ptr())
    A.finish()


@HIPBackend.register('explicitscheme_dissipative')
def _dissipative(domain, wx_face, wy_face, wz_face, cellidf, normalf, namef, dissip_w, Dxx, Dyy, Dzz):
    """solvers/diffusion/fvm_utils.py:5-26 (same function in solvers/advecdiff)."""
    m = _m(domain)
    A = _Args(m.device)
    wx, wy, wz = A.dev(wx_face), A.dev(wy_face), A.dev(wz_face)
    out = A.dev(dissip_w, out=True)
    _lib.call('mfv_explicitscheme_dissipative', m.ptr, wx.data_ptr(), wy.data_ptr(), wz.data_ptr(), float(Dxx),
              float(Dyy), float(Dzz), out.data_ptr(), stream_ptr())
    A.finish()


def _time_step_diff(with_velocity):
    def run(domain, u, v, w, cfl, normal, mesure, volume, faceid, dim, Dxx, Dyy, Dzz):
        import torch
        m = _m(domain)
        A = _Args(m.device)
        # the device copies must outlive the call: the caching allocator would hand a freed
        # temporary's memory to the next allocation before the kernel has read it
        ts = [A.dev(x) for x in (u, v, w)] if with_velocity else []
        uvw = [t.data_ptr() for t in ts] if with_velocity else [None, None, None]
        dt = torch.zeros(1, dtype=torch.float64, device=m.device)
        h = ctypes.c_double(0.)
        _lib.call('mfv_time_step_diffusion', m.ptr, *uvw, float(cfl), float(Dxx), float(Dyy), float(Dzz),
                  dt.data_ptr(), ctypes.byref(h), stream_ptr())
        return h.value
    run.__doc__ = ("solvers/advecdiff/fvm_utils.py:317-340" if with_velocity
                   else "solvers/diffusion/fvm_utils.py:30-49") + "; returns dt (float)."
    return run


HIPBackend.register('diffusion.time_step')(_time_step_diff(False))
HIPBackend.register('advecdiff.time_step')(_time_step_diff(True))
HIPBackend.register('diffusion.update_new_value', 'advecdiff.update_new_value')(_update_new_value)
HIPBackend.register('advecdiff.explicitscheme_convective_2d')(_explicitscheme_convective)
