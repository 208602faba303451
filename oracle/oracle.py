"""ORACLE (test infrastructure only): numpy front-end of liboracle.so.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg use this module, and only as the checker / CPU baseline -- the product path
(``manapy_amd``) never imports it.

Each wrapper takes the reference's arrays (``Domain`` attributes and numpy
fields) and runs the C restatement in ``oracle/mfv_oracle.c``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle.so')
_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from .build import build_oracle
            build_oracle()
        _lib = ctypes.CDLL(LIB)
        _lib.orc_time_step.restype = _D
        _lib.orc_time_step_sw.restype = _D
        _lib.orc_mixed_flux.restype = _D
        _lib.orc_mixed_dt.restype = _D
        _lib.orc_euler32_step.restype = _D
    return _lib


def _p(a, dtype):
    assert a.dtype == dtype and a.flags['C_CONTIGUOUS'], (a.dtype, dtype)
    return a.ctypes.data_as(_P)


def f64(a):
    return _p(a, np.float64)


def i32(a):
    return _p(a, np.int32)


def c3(a):
    """[n, >=3] float array -> contiguous [n, 3]."""
    return np.ascontiguousarray(a[:, :3], dtype=np.float64)


# ---------------------------------------------------------------------------------------------
def time_step(d, u, v, w, cfl):
    fid = d.cells.faceid
    return lib().orc_time_step(f64(u), f64(v), f64(w), _D(cfl), f64(c3(d.faces.normal)), f64(d.cells.volume),
                               i32(fid), _I(fid.shape[1]), _I(int(d.nbcells)))


def update_new_value(ne_c, rez, dissip, src, dtime, vol):
    lib().orc_update_new_value(f64(ne_c), f64(rez), f64(dissip), f64(src), _D(dtime), f64(vol), _I(len(ne_c)))


def facetocell(d, u_face, u_c):
    fid = d.cells.faceid
    lib().orc_facetocell(f64(u_face), f64(u_c), i32(fid), _I(fid.shape[1]), _I(int(d.nbcells)))


def ghost_value(d, kind, src, w_ghost, faces):
    faces = np.ascontiguousarray(faces, dtype=np.int32)
    lib().orc_ghost_value(_I(kind), f64(src), f64(w_ghost), i32(d.faces.cellid), i32(faces), _I(len(faces)))


def explicitscheme_convective(d, rez, w_c, w_ghost, uf, vf, wf, wx, wy, wz, psi, order):
    lib().orc_explicitscheme_convective(
        _I(int(d.dim)), f64(rez), f64(w_c), f64(w_ghost), f64(uf), f64(vf), f64(wf), f64(wx), f64(wy), f64(wz),
        f64(psi), f64(c3(d.cells.center)), f64(c3(d.faces.center)), i32(d.faces.cellid), f64(c3(d.faces.normal)),
        i32(d.innerfaces), _I(len(d.innerfaces)), i32(d.boundaryfaces), _I(len(d.boundaryfaces)), _I(order),
        _I(int(d.nbcells)))


def cell_gradient(d, w_c, w_ghost, wx, wy, wz):
    cn = d.cells.cellnid
    gn = d.cells.ghostnid
    gc = np.ascontiguousarray(d.faces.ghostcenter, dtype=np.float64)
    fn = lib().orc_cell_gradient_2d if int(d.dim) == 2 else lib().orc_cell_gradient_3d
    fn(f64(w_c), f64(w_ghost), f64(c3(d.cells.center)), i32(cn), _I(cn.shape[1]), i32(gn), _I(gn.shape[1]),
       f64(gc), _I(gc.shape[1]), f64(wx), f64(wy), f64(wz), _I(int(d.nbcells)))


def barthlimiter(d, w_c, w_ghost, wx, wy, wz, psi):
    fid = d.cells.faceid
    if int(d.dim) == 2:
        lib().orc_barthlimiter_2d(f64(w_c), f64(w_ghost), f64(wx), f64(wy), f64(psi), i32(d.faces.cellid), i32(fid),
                                  _I(fid.shape[1]), i32(d.faces.name), f64(c3(d.cells.center)),
                                  f64(c3(d.faces.center)), _I(int(d.nbcells)))
    else:
        lib().orc_barthlimiter_3d(f64(w_c), f64(w_ghost), f64(wx), f64(wy), f64(wz), f64(psi), i32(d.faces.cellid),
                                  i32(fid), _I(fid.shape[1]), i32(d.faces.name), f64(c3(d.cells.center)),
                                  f64(c3(d.faces.center)), _I(int(d.nbcells)))


# ---------------------------------------------------------------------------------------------
def time_step_sw(d, h, hu, hv, cfl, grav, Dxx=0., Dyy=0.):
    fid = d.cells.faceid
    return lib().orc_time_step_sw(f64(h), f64(hu), f64(hv), _D(cfl), f64(c3(d.faces.normal)), f64(d.faces.mesure),
                                  f64(d.cells.volume), i32(fid), _I(fid.shape[1]), _D(grav), _D(Dxx), _D(Dyy),
                                  _I(int(d.nbcells)))


def update_sw(h, hu, hv, rez, src, dt, vol):
    lib().orc_update_sw(f64(h), f64(hu), f64(hv), f64(rez[0]), f64(rez[1]), f64(rez[2]), f64(src[0]), f64(src[1]),
                        f64(src[2]), _D(dt), f64(vol), _I(len(h)))


def sw_face(etaL, huL, hvL, zfL, hcL, zcL, etaR, huR, hvR, zfR, hcR, zcR, n0, n1, mes, g):
    fl = np.zeros(3)
    sL = np.zeros(2)
    sR = np.zeros(2)
    lib().orc_sw_face(*[_D(x) for x in (etaL, huL, hvL, zfL, hcL, zcL, etaR, huR, hvR, zfR, hcR, zcR, n0, n1, mes, g)],
                      f64(fl), f64(sL), f64(sR))
    return fl, sL, sR


def sw_ghosts(d, h, hu, hv, Z):
    """Wall ghosts: h Neumann, hu/hv nonslip, Z Neumann, on every boundary face."""
    F = int(d.nbfaces)
    bf = d.boundaryfaces
    out = []
    for kind, w in ((0, h), (1, hu), (1, hv), (0, Z)):
        g = np.zeros(F)
        ghost_value(d, kind, w, g, bf)
        out.append(g)
    return out


def sw_gradients(d, h, hu, hv, Z):
    """Limited LSQ gradients of eta = h + Z, hu, hv ([9][C]) with wall ghosts."""
    C = int(d.nbcells)
    hg, hug, hvg, zg = sw_ghosts(d, h, hu, hv, Z)
    eta = h + Z
    etag = hg + zg
    grad = np.zeros((9, C))
    wz = np.zeros(C)
    for k, (w, wg) in enumerate(((eta, etag), (hu, hug), (hv, hvg))):
        gx = np.zeros(C)
        gy = np.zeros(C)
        psi = np.zeros(C)
        cell_gradient(d, w, wg, gx, gy, wz)
        barthlimiter(d, w, wg, gx, gy, wz, psi)
        grad[3 * k], grad[3 * k + 1], grad[3 * k + 2] = gx, gy, psi
    return grad


def z_gradient(d, Z):
    C = int(d.nbcells)
    zg = np.zeros(int(d.nbfaces))
    ghost_value(d, 0, Z, zg, d.boundaryfaces)
    gx, gy, gz = np.zeros(C), np.zeros(C), np.zeros(C)
    cell_gradient(d, Z, zg, gx, gy, gz)
    return np.stack([gx, gy])


def sw_residuals(d, h, hu, hv, Z, grad, zgrad, grav, order):
    C = int(d.nbcells)
    hg, hug, hvg, zg = sw_ghosts(d, h, hu, hv, Z)
    rez = np.zeros((3, C))
    src = np.zeros((3, C))
    g = [np.ascontiguousarray(x) for x in grad]
    zx, zy = np.ascontiguousarray(zgrad[0]), np.ascontiguousarray(zgrad[1])
    lib().orc_sw_residuals(
        f64(rez[0]), f64(rez[1]), f64(rez[2]), f64(src[0]), f64(src[1]), f64(src[2]), f64(h), f64(hu), f64(hv),
        f64(Z), f64(hg), f64(hug), f64(hvg), f64(zg), *[f64(x) for x in g], f64(zx), f64(zy),
        f64(c3(d.cells.center)), f64(c3(d.faces.center)), i32(d.faces.cellid), f64(c3(d.faces.normal)),
        f64(d.faces.mesure), i32(d.innerfaces), _I(len(d.innerfaces)), i32(d.boundaryfaces),
        _I(len(d.boundaryfaces)), _D(grav), _I(order), _I(C))
    return rez, src


def sw_step(d, h, hu, hv, Z, zgrad, grav, cfl, order):
    """One step of the driver sequence stepper -> compute_fluxes -> compute_new_val
    (examples/2D/shallowwater2d.py:99-107), in place. Returns the dt used."""
    dt = time_step_sw(d, h, hu, hv, cfl, grav)
    if order == 2:
        grad = sw_gradients(d, h, hu, hv, Z)
    else:
        grad = np.zeros((9, int(d.nbcells)))
    rez, src = sw_residuals(d, h, hu, hv, Z, grad, zgrad, grav, order)
    update_sw(h, hu, hv, rez, src, dt, d.cells.volume)
    return dt


class Euler32Model(object):
    """CPU model of the FP32-bulk Euler step (euler32.c) on a Domain of tetrahedra: float32
    records and geometry, float32 face fluxes, float64 gathers / update / dt / sums."""

    def __init__(self, d, U, quirk=0, gamma=1.4):
        C, F = int(d.nbcells), int(d.nbfaces)
        self.d, self.C, self.F, self.quirk, self.gamma = d, C, F, int(quirk), float(gamma)
        self.rec = [np.zeros((C, 8), np.float32) for _ in range(2)]
        self.geo = np.zeros((F, 4), np.float32)
        lib().orc_euler32_pack(_I(C), _I(F), f64(np.ascontiguousarray(U, dtype=np.float64)), f64(d.cells.volume),
                               f64(c3(d.faces.normal)), f64(d.faces.mesure), _p(self.rec[0], np.float32),
                               _p(self.geo, np.float32))
        self.fid = np.ascontiguousarray(d.cells.faceid, dtype=np.int32)
        self.sorted = np.ascontiguousarray(np.sort(self.fid[:, :4], axis=1), dtype=np.int32)
        self.fc = np.ascontiguousarray(d.faces.cellid, dtype=np.int32)
        self.name = np.ascontiguousarray(d.faces.name, dtype=np.int32)
        self.flx = np.zeros((F, 6), np.float32)
        self.cons = np.zeros(5)
        self.k = 0

    @property
    def state(self):
        """[C][8] float32 records of the current state."""
        return self.rec[self.k % 2]

    def step(self, cfl, dt):
        """One step with the given dt; returns the next dt."""
        a, b = self.rec[self.k % 2], self.rec[(self.k + 1) % 2]
        nxt = lib().orc_euler32_step(_I(self.C), _I(self.F), _p(a, np.float32), _p(b, np.float32),
                                     _p(self.geo, np.float32), i32(self.fc), i32(self.name), i32(self.fid),
                                     _I(self.fid.shape[1]), i32(self.sorted), _D(self.gamma), _D(cfl), _D(dt),
                                     _I(self.quirk), _p(self.flx, np.float32), f64(self.cons))
        self.k += 1
        return nxt


class OmpSwStep(object):
    """The all-cores CPU baseline of ``sw_step`` (omp_baseline.c): the oracle's own per-cell
    and per-face code run with OpenMP, bit-identical to ``sw_step``. Holds the mesh arrays and
    the scratch across steps."""

    def __init__(self, d, Z):
        C, F = int(d.nbcells), int(d.nbfaces)
        self.d, self.C, self.F = d, C, F
        self.lib = ctypes.CDLL(os.path.join(HERE, 'liboracle_omp.so'))
        self.lib.orc_sw_step_omp.restype = _D
        self.Z = np.ascontiguousarray(Z, dtype=np.float64)
        zg = z_gradient(d, self.Z)
        self.zx, self.zy = np.ascontiguousarray(zg[0]), np.ascontiguousarray(zg[1])
        self.gc = np.ascontiguousarray(d.faces.ghostcenter, dtype=np.float64)
        self.keep = [c3(d.cells.center), c3(d.faces.center), c3(d.faces.normal)]
        self.scratch = np.zeros(26 * C + 13 * F)

    def threads(self):
        return int(self.lib.orc_omp_threads())

    def step(self, h, hu, hv, grav, cfl, order):
        d = self.d
        cc, fc, nf = self.keep
        return self.lib.orc_sw_step_omp(
            _I(self.C), _I(self.F), f64(h), f64(hu), f64(hv), f64(self.Z), f64(self.zx), f64(self.zy), f64(cc),
            f64(fc), i32(d.faces.cellid), f64(nf), f64(d.faces.mesure), i32(d.cells.faceid),
            _I(d.cells.faceid.shape[1]), i32(d.faces.name), i32(d.cells.cellnid), _I(d.cells.cellnid.shape[1]),
            i32(d.cells.ghostnid), _I(d.cells.ghostnid.shape[1]), f64(self.gc), _I(self.gc.shape[1]),
            i32(d.innerfaces), _I(len(d.innerfaces)), i32(d.boundaryfaces), _I(len(d.boundaryfaces)),
            f64(d.cells.volume), _D(grav), _D(cfl), _I(order), f64(self.scratch))


# --- reference shallow-water path: SRNH flux + well-balanced source (oracle/srnh.c) ----------
def srnh_roe(hroe, uroe, vroe, croe, grav):
    """Eigenvalues (5,) and sign matrix (5, 5) of fvm_utils.py:220-315."""
    lam = np.zeros(5)
    r = np.zeros(25)
    lib().orc_srnh_roe(_D(hroe), _D(uroe), _D(vroe), _D(croe), _D(grav), f64(lam), f64(r))
    return lam, r.reshape(5, 5)


def srnh_flux(hu_l, hu_r, hv_l, hv_r, h_l, h_r, hc_l, hc_r, Z_l, Z_r, normal, mesure, grav):
    fl = np.zeros(5)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    lib().orc_srnh_flux(*[_D(x) for x in (hu_l, hu_r, hv_l, hv_r, h_l, h_r, hc_l, hc_r, Z_l, Z_r)], f64(n),
                        _D(mesure), _D(grav), f64(fl))
    return fl


SRNH_WALLS = (0, 1, 1, 0, 0)  # ghost kinds of h, hu, hv, hc, Z: Neumann / nonslip


def srnh_ghosts(d, S, kinds=SRNH_WALLS):
    F = int(d.nbfaces)
    G = np.zeros((5, F))
    for k in range(5):
        ghost_value(d, kinds[k], np.ascontiguousarray(S[k]), G[k], d.boundaryfaces)
    return G


def srnh_convective(d, S, G, grads, grav, order):
    """explicitscheme_convective_SW; grads = (h_x, h_y, hc_x, hc_y, psi). Returns rez [5][C]."""
    C, F = int(d.nbcells), int(d.nbfaces)
    rez = np.zeros((5, C))
    g = [np.ascontiguousarray(x, dtype=np.float64) for x in grads]
    lib().orc_convective_srnh(
        f64(rez), f64(np.ascontiguousarray(S)), f64(np.ascontiguousarray(G)), *[f64(x) for x in g],
        f64(c3(d.cells.center)), f64(c3(d.faces.center)), i32(d.faces.cellid), f64(d.faces.mesure),
        f64(c3(d.faces.normal)), i32(d.innerfaces), _I(len(d.innerfaces)), i32(d.boundaryfaces),
        _I(len(d.boundaryfaces)), _D(grav), _I(order), _I(C), _I(F))
    return rez


def srnh_source(d, S, G, grav, quirk=True):
    """term_source_srnh_SW. Returns src [5][C] (only hu, hv non-zero). quirk=True reads the
    neighbour of face slot j from the compacted cells.cellfid as the reference does."""
    C = int(d.nbcells)
    src = np.zeros((5, C))
    nid, fid, cfid = d.cells.nodeid, d.cells.faceid, d.cells.cellfid
    nf = np.ascontiguousarray(d.cells.nf, dtype=np.float64)
    vx = np.ascontiguousarray(d.nodes.vertex, dtype=np.float64)
    assert vx.shape[1] == 4 and nf.shape[1:] == (3, 3) and fid.shape[1] == cfid.shape[1]
    lib().orc_source_srnh(f64(src[1]), f64(src[2]), f64(np.ascontiguousarray(S[0])), f64(np.ascontiguousarray(S[4])),
                          f64(np.ascontiguousarray(G[0])), f64(np.ascontiguousarray(G[4])), i32(nid),
                          _I(nid.shape[1]), i32(fid), i32(cfid), _I(fid.shape[1]), i32(d.faces.cellid),
                          f64(c3(d.cells.center)), f64(nf), i32(d.faces.name), f64(vx), _D(grav), _I(int(quirk)),
                          _I(C))
    return src


def srnh_update(d, S, rez, src, dt, dissip_hc=None, corio_hu=None, corio_hv=None, wind=(0., 0.)):
    C = int(d.nbcells)
    z = np.zeros(C)
    dh = z if dissip_hc is None else np.ascontiguousarray(dissip_hc)
    cu = z if corio_hu is None else np.ascontiguousarray(corio_hu)
    cv = z if corio_hv is None else np.ascontiguousarray(corio_hv)
    lib().orc_update_sw5(f64(S), f64(np.ascontiguousarray(rez)), f64(np.ascontiguousarray(src)), f64(dh), f64(cu),
                         f64(cv), _D(wind[0]), _D(wind[1]), _D(dt), f64(d.cells.volume), _I(C))


def srnh_gradients(d, S, G):
    """h and hc gradients, each limited by its own Barth-Jespersen psi
    (Variable.compute_cell_gradient, ast/core.py:442-456); the scheme uses hc.psi for both."""
    C = int(d.nbcells)
    wz = np.zeros(C)
    out = []
    for k in (0, 3):
        gx, gy, psi = np.zeros(C), np.zeros(C), np.zeros(C)
        w = np.ascontiguousarray(S[k])
        cell_gradient(d, w, G[k], gx, gy, wz)
        barthlimiter(d, w, G[k], gx, gy, wz, psi)
        out.append((gx, gy, psi))
    (hx, hy, _), (hcx, hcy, hcpsi) = out
    return hx, hy, hcx, hcy, hcpsi


def term_friction_sw(h, hu, hv, grav, eta, time):
    """term_friction_SW in place on hu, hv."""
    lib().orc_term_friction_sw(f64(h), f64(hu), f64(hv), _D(grav), _D(eta), _D(time), _I(len(h)))


def term_coriolis_sw(hu, hv, corio_hu, corio_hv, f_c):
    lib().orc_term_coriolis_sw(f64(hu), f64(hv), f64(corio_hu), f64(corio_hv), f64(f_c), _I(len(hu)))


def term_wind_sw(Tx, Ty, wx, wy):
    lib().orc_term_wind_sw(f64(Tx), f64(Ty), f64(wx), f64(wy), _I(len(Tx)))


def srnh_step(d, S, grav, cfl, order, kinds=SRNH_WALLS, quirk=True, Mann=0., fc=0.):
    """ShallowWaterSolver: stepper -> compute_fluxes -> compute_new_val
    (solvers/shallowater/system.py:186-281) with Dxx = Dyy = 0, in place on S [5][C] (h, hu,
    hv, hc, Z). Mann != 0: friction with the driver's dt, then compute_fluxes' second stepper
    call, whose dt the update uses; fc != 0: term_coriolis_SW with f_c = fc in every cell
    (the reference passes the scalar, which the kernel indexes). Returns the dt of the update."""
    C = int(d.nbcells)
    dt = time_step_sw(d, np.ascontiguousarray(S[0]), np.ascontiguousarray(S[1]), np.ascontiguousarray(S[2]), cfl,
                      grav)
    if Mann != 0:
        term_friction_sw(S[0], S[1], S[2], grav, Mann, dt)
        dt = time_step_sw(d, S[0], S[1], S[2], cfl, grav)
    G = srnh_ghosts(d, S, kinds)
    if order == 2:
        grads = srnh_gradients(d, S, G)
    else:
        grads = tuple(np.zeros(C) for _ in range(5))
    rez = srnh_convective(d, S, G, grads, grav, order)
    src = srnh_source(d, S, G, grav, quirk)
    corio_hu, corio_hv = np.zeros(C), np.zeros(C)
    if fc != 0:
        term_coriolis_sw(S[1], S[2], corio_hu, corio_hv, np.full(C, float(fc)))
    srnh_update(d, S, rez, src, dt, corio_hu=corio_hu, corio_hv=corio_hv)
    return dt


def advection_step(d, w, u_c, v_c, w_c, uf, vf, wf, cfl, order):
    """AdvectionSolver.stepper -> compute_fluxes -> compute_new_val with all-Neumann BCs
    (solvers/advec/system.py:129-147), in place. Returns dt."""
    C = int(d.nbcells)
    dt = time_step(d, u_c, v_c, w_c, cfl)
    ghost = np.zeros(int(d.nbfaces))
    ghost_value(d, 0, w, ghost, d.boundaryfaces)
    wx, wy, wz, psi = np.zeros(C), np.zeros(C), np.zeros(C), np.zeros(C)
    if order == 2:
        cell_gradient(d, w, ghost, wx, wy, wz)
        barthlimiter(d, w, ghost, wx, wy, wz, psi)
    rez = np.zeros(C)
    explicitscheme_convective(d, rez, w, ghost, uf, vf, wf, wx, wy, wz, psi, order)
    z = np.zeros(C)
    update_new_value(w, rez, z, z, dt, d.cells.volume)
    return dt


# ---------------------------------------------------------------------------------------------
def euler_ghost(d, U, G, quirk):
    C = int(d.nbcells)
    F = int(d.nbfaces)
    Uv = [U[k] for k in range(6)]
    Gv = [G[k] for k in range(6)]
    # argument order of the C function: rhog, Pg, rhoug, rhovg, rhowg, rhoEg, then cell arrays
    lib().orc_ghost_tubeschok3d(*[f64(g) for g in Gv], *[f64(u) for u in Uv], i32(d.faces.cellid),
                                i32(d.faces.name.astype(np.int32)), f64(c3(d.faces.normal)), _I(F), _I(quirk))
    del C


def euler_dt(d, U, cfl, gamma):
    C = int(d.nbcells)
    fid = d.cells.faceid
    dtc = np.zeros(C)
    lib().orc_time_step_euler_3d(f64(U[0]), f64(U[1]), f64(U[2]), f64(U[3]), f64(U[4]), _D(cfl),
                                 f64(c3(d.faces.normal)), f64(d.faces.mesure), f64(d.cells.volume), i32(fid),
                                 _I(fid.shape[1]), _D(gamma), f64(dtc), _I(C))
    return dtc


def euler_rusanov(d, U, G, gamma, quirk):
    C = int(d.nbcells)
    rez = np.zeros((5, C))
    lib().orc_euler3d_rusanov(f64(rez), f64(np.ascontiguousarray(U)), f64(np.ascontiguousarray(G)),
                              i32(d.faces.cellid), f64(c3(d.faces.normal)), f64(d.faces.mesure),
                              i32(d.faces.name.astype(np.int32)), _D(gamma), _I(quirk), _I(C), _I(int(d.nbfaces)))
    return rez


def euler_update(d, U, rez, gamma, dt):
    lib().orc_update_euler_3d(f64(U), f64(np.ascontiguousarray(rez)), _D(gamma), _D(dt), f64(d.cells.volume),
                              _I(int(d.nbcells)))


def euler_step(d, U, G, cfl, gamma, quirk):
    """ghost -> dt (min of dt_c) -> Rusanov residual -> update, in place on U [6][C] and G [6][F]."""
    euler_ghost(d, U, G, quirk)
    dt = float(np.min(np.concatenate([[1e6], euler_dt(d, U, cfl, gamma)])))
    rez = euler_rusanov(d, U, G, gamma, quirk)
    euler_update(d, U, rez, gamma, dt)
    return dt


# ---------------------------------------------------------------------------------------------
# mixed-precision shallow-water kernels restated in C float arithmetic (mixed32.c)

class MixedModel(object):
    """CPU model of the mixed-precision shallow-water path on a Domain (layout as the GPU's)."""

    def __init__(self, d, h, hu, hv, Z, zgrad, order=2):
        from manapy_amd.device import gpu_layout, sw_packed_layout
        L = gpu_layout(d, lsq=True)
        Pk = sw_packed_layout(d)
        self.L, self.Pk = L, Pk
        self.C, self.MF = L['C'], L['MF']
        C, MF = self.C, self.MF
        self.state = np.zeros((C, 4))
        self.geoA = np.zeros((C, MF, 4), np.float32)
        self.geoB = np.zeros((C, MF, 4), np.float32)
        self.facerec = np.zeros((C, MF, 4), np.float32)
        self.vol = np.ascontiguousarray(d.cells.volume, dtype=np.float64)
        self.om = float(order - 1)
        zg = np.ascontiguousarray(zgrad, dtype=np.float64)
        lib().orc_mixed_pack(_I(C), _I(MF), i32(L['cell_faces']), i32(L['cell_nfaces']), i32(L['face_cells']),
                             f64(L['face_normal']), f64(L['face_center']), f64(L['cell_center']), f64(h), f64(hu),
                             f64(hv), f64(Z), f64(zg), _D(self.om), f64(self.state), _p(self.geoA, np.float32),
                             _p(self.geoB, np.float32))

    def dt(self, g, cfl):
        """time_step_SW of the current state with the float32 face normals."""
        return lib().orc_mixed_dt(_I(self.C), _I(self.MF), f64(self.state), f64(self.vol), _p(self.geoA, np.float32),
                                  i32(self.L['cell_nfaces']), ctypes.c_float(g), _D(cfl))

    def step(self, g, cfl, dt):
        """One step; returns the next dt."""
        L, Pk = self.L, self.Pk
        lib().orc_mixed_recon(_I(self.C), _I(self.MF), f64(self.state), f64(L['cell_center']), i32(L['lsq_cells']),
                              i32(L['lsq_ncells']), i32(L['lsq_ghosts']), i32(L['lsq_nghosts']), i32(L['face_cells']),
                              f64(L['face_ghost']), i32(L['cell_nfaces']), i32(Pk['nbr']), _p(self.geoA, np.float32),
                              _p(self.geoB, np.float32), ctypes.c_float(self.om), _p(self.facerec, np.float32))
        return lib().orc_mixed_flux(_I(self.C), _I(self.MF), f64(self.state), f64(self.vol),
                                    _p(self.facerec, np.float32), _p(self.geoA, np.float32),
                                    _p(self.geoB, np.float32), i32(L['cell_nfaces']),
                                    i32(Pk['nbr']), _p(Pk['order'], np.uint8), ctypes.c_float(g), _D(cfl), _D(dt))
