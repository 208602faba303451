"""ORACLE (test infrastructure only): loop restatement of manapy's 2D diamond scheme.

Follows, loop for loop and in insertion order:
  * compute_2dmatrix_size      ast/functions2d.py:582-664
  * get_triplet_2d             ast/functions2d.py:328-580   (incl. the `cmptparam =+1` line, :466)
  * get_rhs_glob_2d            ast/functions2d.py:740-817
  * LinearSolver.update_ghost_values  solvers/ls/system.py:179-189
  * centertovertex_2d          ast/functions2d.py:199-277
  * face_gradient_2d           ast/functions2d.py:108-196
  * explicitscheme_dissipative solvers/diffusion/fvm_utils.py:5-26
  * time_step (diffusion)      solvers/diffusion/fvm_utils.py:30-49
  * time_step (advecdiff)      solvers/advecdiff/fvm_utils.py:317-340
  * compute_P_gradient_2d_diamond  ast/functions2d.py:820-936
  * insertion-order duplicate sum of the COO triplets (what summing them one by one gives)
One subdomain, no periodic faces: the halo/periodic loops are empty.

The 3D functions at the end of the file follow functions3d.py the same way.

Pinned in tests/test_diamond.py: the assembled operator is exact on affine fields
(the diamond scheme's defining property), rows of the pure-Neumann operator sum to zero,
the Dirichlet problem with affine data is solved exactly by SuperLU (the reference
ScipySolver's algorithm), node interpolation and face gradients are exact for affine
fields.
"""
import numpy as np


def _alpha(d, nod, cx, cy):
    n = d.nodes
    xdiff = cx - n.vertex[nod][0]
    ydiff = cy - n.vertex[nod][1]
    return (1. + n.lambda_x[nod] * xdiff + n.lambda_y[nod] * ydiff) / (
        n.number[nod] + n.lambda_x[nod] * n.R_x[nod] + n.lambda_y[nod] * n.R_y[nod])


def matrixinnerfaces(d):
    """``solvers/ls/system.py:124-127`` (no periodic faces)."""
    return np.sort(np.asarray(d.innerfaces, dtype=np.int32))


def compute_2dmatrix_size(d, BCdirichlet, mif, dirichletfaces):
    ncell = d.nodes.cellid
    gcn = d.nodes.ghostcenter
    cmpt = 0
    for i in mif:
        cmpt += 1
        for nod in d.faces.nodeid[i][:d.faces.nodeid[i][-1]]:
            if d.nodes.oldname[nod] not in BCdirichlet:
                cmpt += 2 * ncell[nod][-1]
                cmpt += 2 * int(np.sum(gcn[nod][:, -1] != -1))
        cmpt += 3
    cmpt += 2 * len(dirichletfaces)
    return cmpt


def get_triplet_2d(d, BCdirichlet, mif, dirichletfaces):
    """COO entries (data, row, col) in the reference's insertion order."""
    vol = d.cells.volume
    cid = d.faces.cellid
    ncell = d.nodes.cellid
    gcn = d.nodes.ghostcenter
    p1, p2, p3, p4 = d.faces.param1, d.faces.param2, d.faces.param3, d.faces.param4
    data, row, col = [], [], []

    def add(r, c, v):
        row.append(r)
        col.append(c)
        data.append(v)

    for i in mif:
        cl, cr = cid[i][0], cid[i][1]
        params = (p4[i], p2[i])
        add(cl, cl, p1[i] / vol[cl])
        cmptparam = 0
        for nod in d.faces.nodeid[i][:d.faces.nodeid[i][-1]]:
            if d.nodes.oldname[nod] not in BCdirichlet:
                for j in range(ncell[nod][-1]):
                    c = ncell[nod][j]
                    a = _alpha(d, nod, d.cells.center[c][0], d.cells.center[c][1])
                    add(cl, c, a / vol[cl] * params[cmptparam])
                    add(cr, c, -1. * a / vol[cr] * params[cmptparam])
                for j in range(len(gcn[nod])):
                    if gcn[nod][j][-1] != -1:
                        a = _alpha(d, nod, gcn[nod][j][0], gcn[nod][j][1])
                        index = np.int32(gcn[nod][j][2])
                        add(cl, index, a / vol[cl] * params[cmptparam])
                        add(cr, index, -1. * a / vol[cr] * params[cmptparam])
            cmptparam = 1  # `cmptparam =+1` (functions2d.py:466): set to 1, not incremented
        add(cl, cr, p3[i] / vol[cl])
        add(cr, cl, -1. * p1[i] / vol[cr])
        add(cr, cr, -1. * p3[i] / vol[cr])
    for i in dirichletfaces:
        cl = cid[i][0]
        add(cl, cl, p1[i] / vol[cl])
        add(cl, cl, -1. * p3[i] / vol[cl])
    return np.array(data), np.array(row, dtype=np.int32), np.array(col, dtype=np.int32)


def update_ghost_values(d, bcs, values):
    """``LinearSolver.update_ghost_values`` (solvers/ls/system.py:179-189): Pbordface/Pbordnode
    from Dirichlet values (constants or lambdas of x, y, z), 1.0 on Neumann nodes."""
    Pbordnode = np.zeros(int(d.nbnodes))
    Pbordface = np.zeros(int(d.nbfaces))
    index = {"in": 1, "out": 2, "upper": 3, "bottom": 4, "front": 5, "back": 6}
    faces = {"in": d.infaces, "out": d.outfaces, "upper": d.upperfaces, "bottom": d.bottomfaces}
    if int(d.dim) == 3:
        faces.update({"front": d.frontfaces, "back": d.backfaces})
    for loc, bct in bcs.items():
        nodes = np.nonzero(d.nodes.oldname == index[loc])[0]
        if bct == "dirichlet":
            v = values[loc]
            fc = d.faces.center
            vx = d.nodes.vertex
            for f in faces[loc]:
                Pbordface[f] = v(fc[f][0], fc[f][1], fc[f][2]) if callable(v) else float(v)
            for n in nodes:
                Pbordnode[n] = v(vx[n][0], vx[n][1], vx[n][2]) if callable(v) else float(v)
        elif bct == "neumann":
            Pbordnode[nodes] = 1.
    return Pbordnode, Pbordface


def get_rhs_glob_2d(d, Pbordnode, Pbordface, BCdirichlet, mif, dirichletfaces):
    vol = d.cells.volume
    cid = d.faces.cellid
    fn = d.faces.nodeid
    on = d.nodes.oldname
    gcn = d.nodes.ghostcenter
    p2, p3, p4 = d.faces.param2, d.faces.param3, d.faces.param4
    rhs = np.zeros(int(d.nbcells))
    for i in mif:
        cl, cr = cid[i][0], cid[i][1]
        i_1, i_2 = fn[i][0], fn[i][1]
        if on[i_1] in BCdirichlet:
            V = Pbordnode[i_1]
            rhs[cl] += -1. * V * p4[i] / vol[cl]
            rhs[cr] += V * p4[i] / vol[cr]
        if on[i_2] in BCdirichlet:
            V = Pbordnode[i_2]
            rhs[cl] += -1. * V * p2[i] / vol[cl]
            rhs[cr] += V * p2[i] / vol[cr]
    for i in dirichletfaces:
        cl = cid[i][0]
        i_1, i_2 = fn[i][0], fn[i][1]
        if gcn[i_1][0][2] != -1:
            rhs[cl] += -1. * Pbordnode[i_1] * p4[i] / vol[cl]
        if gcn[i_2][0][2] != -1:
            rhs[cl] += -1. * Pbordnode[i_2] * p2[i] / vol[cl]
        rhs[cl] += -2. * p3[i] / vol[cl] * Pbordface[i]
    return rhs


def centertovertex_2d(d, w_c, w_ghost):
    n = d.nodes
    w_n = np.zeros(int(d.nbnodes))
    for i in range(int(d.nbnodes)):
        den = n.number[i] + n.lambda_x[i] * n.R_x[i] + n.lambda_y[i] * n.R_y[i]
        for j in range(n.cellid[i][-1]):
            c = n.cellid[i][j]
            xd = d.cells.center[c][0] - n.vertex[i][0]
            yd = d.cells.center[c][1] - n.vertex[i][1]
            w_n[i] += (1. + n.lambda_x[i] * xd + n.lambda_y[i] * yd) / den * w_c[c]
        for j in range(n.ghostid[i][-1]):
            f = n.ghostid[i][j]
            xd = d.faces.ghostcenter[f][0] - n.vertex[i][0]
            yd = d.faces.ghostcenter[f][1] - n.vertex[i][1]
            w_n[i] += (1. + n.lambda_x[i] * xd + n.lambda_y[i] * yd) / den * w_ghost[f]
    return w_n


def face_gradient_2d(d, w_c, w_ghost, w_node, faces_boundary):
    """Inner faces and the given boundary faces (dirichlet + neumann lists)."""
    F = int(d.nbfaces)
    wx, wy = np.zeros(F), np.zeros(F)
    cid = d.faces.cellid
    fn = d.faces.nodeid
    f1, f2, f3, f4 = d.faces.f_1, d.faces.f_2, d.faces.f_3, d.faces.f_4
    air = d.faces.airDiamond

    def one(i, vv2):
        vi1, vi2 = w_node[fn[i][0]], w_node[fn[i][1]]
        vv1 = w_c[cid[i][0]]
        wx[i] = 1 / (2 * air[i]) * ((vi1 + vv1) * f1[i][1] + (vv1 + vi2) * f2[i][1] + (vi2 + vv2) * f3[i][1]
                                    + (vv2 + vi1) * f4[i][1])
        wy[i] = -1 / (2 * air[i]) * ((vi1 + vv1) * f1[i][0] + (vv1 + vi2) * f2[i][0] + (vi2 + vv2) * f3[i][0]
                                     + (vv2 + vi1) * f4[i][0])

    for i in d.innerfaces:
        one(i, w_c[cid[i][1]])
    for i in faces_boundary:
        one(i, w_ghost[i])
    return wx, wy


def explicitscheme_dissipative(d, wx, wy, Dxx, Dyy, wz=None, Dzz=0.):
    """wz = None: the 2D call (zero z gradient, Dzz = 0)."""
    dis = np.zeros(int(d.nbcells))
    cid = d.faces.cellid
    nrm = d.faces.normal
    wz = np.zeros(int(d.nbfaces)) if wz is None else wz
    for i in range(int(d.nbfaces)):
        q = Dxx * wx[i] * nrm[i][0] + Dyy * wy[i] * nrm[i][1] + Dzz * wz[i] * nrm[i][2]
        if d.faces.name[i] == 0:
            dis[cid[i][0]] += q
            dis[cid[i][1]] -= q
        else:
            dis[cid[i][0]] += q
    return dis


def time_step_diffusion(d, Dxx, Dyy, cfl, Dzz=0.):
    """``solvers/diffusion/fvm_utils.py:30-49`` (2D: Dzz = 0)."""
    dt = 1e6
    nrm = d.faces.normal
    for i in range(int(d.nbcells)):
        lam = 0.
        for j in range(d.cells.faceid[i][-1]):
            n = nrm[d.cells.faceid[i][j]]
            mes = np.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
            lam += (Dxx * (mes * mes) + Dyy * (mes * mes) + Dzz * (mes * mes)) / d.cells.volume[i]
        dt = min(dt, cfl * d.cells.volume[i] / lam)
    return dt


def time_step_advdiff(d, u, v, w, Dxx, Dyy, cfl):
    """``solvers/advecdiff/fvm_utils.py:317-340`` (2D: Dzz = 0)."""
    dt = 1e6
    nrm = d.faces.normal
    for i in range(int(d.nbcells)):
        lam = 0.
        for j in range(d.cells.faceid[i][-1]):
            n = nrm[d.cells.faceid[i][j]]
            lam += abs(u[i] * n[0] + v[i] * n[1] + w[i] * n[2])
            mes = np.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
            lam += (Dxx * (mes * mes) + Dyy * (mes * mes) + 0. * (mes * mes)) / d.cells.volume[i]
        if lam != 0:
            dt = min(dt, cfl * d.cells.volume[i] / lam)
    return dt


def p_gradient_2d(d, P_c, P_ghost, P_node, Pbordnode, Pbordface, BCdirichlet, neumannfaces, dirichletfaces):
    """``compute_P_gradient_2d_diamond``: inner faces, Neumann faces (ghost values) and
    Dirichlet faces (mirror 2 V_K - P_c, Dirichlet node values)."""
    F = int(d.nbfaces)
    px, py = np.zeros(F), np.zeros(F)
    cid = d.faces.cellid
    fn = d.faces.nodeid
    on = d.nodes.oldname
    f1, f2, f3, f4 = d.faces.f_1, d.faces.f_2, d.faces.f_3, d.faces.f_4
    air = d.faces.airDiamond

    def nodeval(n):
        return Pbordnode[n] if on[n] in BCdirichlet else P_node[n]

    def one(i, vi1, vi2, vv1, vv2):
        px[i] = -1 / (2 * air[i]) * ((vi1 + vv1) * f1[i][1] + (vv1 + vi2) * f2[i][1] + (vi2 + vv2) * f3[i][1]
                                     + (vv2 + vi1) * f4[i][1])
        py[i] = 1 / (2 * air[i]) * ((vi1 + vv1) * f1[i][0] + (vv1 + vi2) * f2[i][0] + (vi2 + vv2) * f3[i][0]
                                    + (vv2 + vi1) * f4[i][0])

    for i in d.innerfaces:
        one(i, nodeval(fn[i][0]), nodeval(fn[i][1]), P_c[cid[i][0]], P_c[cid[i][1]])
    for i in neumannfaces:
        one(i, nodeval(fn[i][0]), nodeval(fn[i][1]), P_c[cid[i][0]], P_ghost[i])
    for i in dirichletfaces:
        vv1 = P_c[cid[i][0]]
        one(i, Pbordnode[fn[i][0]], Pbordnode[fn[i][1]], vv1, 2. * Pbordface[i] - vv1)
    return px, py


def coo_sum_in_order(data, row, col):
    """{(row, col): value} with duplicates added in insertion order."""
    out = {}
    for v, r, c in zip(data.tolist(), row.tolist(), col.tolist()):
        k = (r, c)
        out[k] = out[k] + v if k in out else v
    return out


# ---------------------------------------------------------------------------------------------
# 3D diamond scheme (one subdomain: the halo, haloghost and periodic loops are empty)
#   * compute_3dmatrix_size        ast/functions3d.py:410-495
#   * get_triplet_3d               ast/functions3d.py:498-782
#   * get_rhs_glob_3d              ast/functions3d.py:865-946
#   * centertovertex_3d            ast/functions3d.py:251-362
#   * face_gradient_3d             ast/functions3d.py:122-248
#   * compute_P_gradient_3d_diamond  ast/functions3d.py:949-1108
# A face has 3 or 4 nodes; a triangle's fourth node is its third (nodes[3] = nodeidf[i][2]).

def _alpha3(d, nod, cx, cy, cz):
    n = d.nodes
    xdiff = cx - n.vertex[nod][0]
    ydiff = cy - n.vertex[nod][1]
    zdiff = cz - n.vertex[nod][2]
    return (1. + n.lambda_x[nod] * xdiff + n.lambda_y[nod] * ydiff + n.lambda_z[nod] * zdiff) / (
        n.number[nod] + n.lambda_x[nod] * n.R_x[nod] + n.lambda_y[nod] * n.R_y[nod] + n.lambda_z[nod] * n.R_z[nod])


def face_nodes4(d, i):
    fn = d.faces.nodeid[i]
    return [int(fn[0]), int(fn[1]), int(fn[2]), int(fn[3]) if fn[-1] == 4 else int(fn[2])]


def compute_3dmatrix_size(d, BCdirichlet, mif, dirichletfaces):
    ncell = d.nodes.cellid
    gcn = d.nodes.ghostcenter
    cmpt = 0
    for i in mif:
        cmpt += 1
        for nod in face_nodes4(d, i):
            if d.nodes.oldname[nod] not in BCdirichlet:
                cmpt += 2 * ncell[nod][-1]
                cmpt += 2 * int(np.sum(gcn[nod][:, -1] != -1))
        cmpt += 3
    cmpt += 2 * len(dirichletfaces)
    return cmpt


def get_triplet_3d(d, BCdirichlet, mif, dirichletfaces):
    """COO entries (data, row, col) in the reference's insertion order."""
    vol = d.cells.volume
    cid = d.faces.cellid
    ncell = d.nodes.cellid
    gcn = d.nodes.ghostcenter
    cc = d.cells.center
    p1, p2, p3 = d.faces.param1, d.faces.param2, d.faces.param3
    data, row, col = [], [], []

    def add(r, c, v):
        row.append(r)
        col.append(c)
        data.append(v)

    for i in mif:
        cl, cr = cid[i][0], cid[i][1]
        params = (p1[i], p2[i], -1. * p1[i], -1. * p2[i])
        add(cl, cl, -1 * p3[i] / vol[cl])
        for cmptparam, nod in enumerate(face_nodes4(d, i)):
            if d.nodes.oldname[nod] not in BCdirichlet:
                for j in range(ncell[nod][-1]):
                    c = ncell[nod][j]
                    a = _alpha3(d, nod, cc[c][0], cc[c][1], cc[c][2])
                    add(cl, c, a / vol[cl] * params[cmptparam])
                    add(cr, c, -1. * a / vol[cr] * params[cmptparam])
                for j in range(len(gcn[nod])):
                    if gcn[nod][j][-1] != -1:
                        a = _alpha3(d, nod, gcn[nod][j][0], gcn[nod][j][1], gcn[nod][j][2])
                        index = np.int32(gcn[nod][j][3])
                        add(cl, index, a / vol[cl] * params[cmptparam])
                        add(cr, index, -1. * a / vol[cr] * params[cmptparam])
        add(cl, cr, p3[i] / vol[cl])
        add(cr, cl, p3[i] / vol[cr])
        add(cr, cr, -1. * p3[i] / vol[cr])
    for i in dirichletfaces:
        cl = cid[i][0]
        add(cl, cl, -1 * p3[i] / vol[cl])
        add(cl, cl, -1. * p3[i] / vol[cl])
    return np.array(data), np.array(row, dtype=np.int32), np.array(col, dtype=np.int32)


def get_rhs_glob_3d(d, Pbordnode, Pbordface, BCdirichlet, mif, dirichletfaces):
    vol = d.cells.volume
    cid = d.faces.cellid
    on = d.nodes.oldname
    gcn = d.nodes.ghostcenter
    p1, p2, p3 = d.faces.param1, d.faces.param2, d.faces.param3
    rhs = np.zeros(int(d.nbcells))
    for i in mif:
        cl, cr = cid[i][0], cid[i][1]
        params = (p1[i], p2[i], -1. * p1[i], -1. * p2[i])
        for cmpt, nod in enumerate(face_nodes4(d, i)):
            if on[nod] in BCdirichlet:
                V = Pbordnode[nod]
                rhs[cl] += -1. * V * params[cmpt] / vol[cl]
                rhs[cr] += V * params[cmpt] / vol[cr]
    for i in dirichletfaces:
        cl = cid[i][0]
        params = (p1[i], p2[i], -1. * p1[i], -1. * p2[i])
        for cmpt, nod in enumerate(face_nodes4(d, i)):
            if gcn[nod][0][3] != -1:
                V = Pbordnode[nod]
                rhs[cl] += -1. * V * params[cmpt] / vol[cl]
        rhs[cl] += -2. * p3[i] / vol[cl] * Pbordface[i]
    return rhs


def centertovertex_3d(d, w_c, w_ghost):
    n = d.nodes
    cc = d.cells.center
    gc = d.faces.ghostcenter
    w_n = np.zeros(int(d.nbnodes))
    for i in range(int(d.nbnodes)):
        for j in range(n.cellid[i][-1]):
            c = n.cellid[i][j]
            w_n[i] += _alpha3(d, i, cc[c][0], cc[c][1], cc[c][2]) * w_c[c]
        for j in range(n.ghostid[i][-1]):
            f = n.ghostid[i][j]
            w_n[i] += _alpha3(d, i, gc[f][0], gc[f][1], gc[f][2]) * w_ghost[f]
    return w_n


def face_gradient_3d(d, w_c, w_ghost, w_node, faces_boundary):
    """Inner faces and the given boundary faces (dirichlet + neumann lists)."""
    F = int(d.nbfaces)
    out = (np.zeros(F), np.zeros(F), np.zeros(F))
    cid = d.faces.cellid
    f1, f2, nrm, air = d.faces.f_1, d.faces.f_2, d.faces.normal, d.faces.airDiamond

    def one(i, VR):
        VA, VB, VC, VD = (w_node[k] for k in face_nodes4(d, i))
        VL = w_c[cid[i][0]]
        for k in range(3):
            out[k][i] = (f1[i][k] * (VA - VC) + f2[i][k] * (VB - VD) + nrm[i][k] * (VR - VL)) / air[i]

    for i in d.innerfaces:
        one(i, w_c[cid[i][1]])
    for i in faces_boundary:
        one(i, w_ghost[i])
    return out


def p_gradient_3d(d, P_c, P_ghost, P_node, Pbordnode, Pbordface, BCdirichlet, neumannfaces, dirichletfaces):
    """``compute_P_gradient_3d_diamond``: inner faces, Neumann faces (ghost values), Dirichlet
    faces (the four Dirichlet node values, mirror 2 V_K - P_c)."""
    F = int(d.nbfaces)
    out = (np.zeros(F), np.zeros(F), np.zeros(F))
    cid = d.faces.cellid
    on = d.nodes.oldname
    f1, f2, nrm, air = d.faces.f_1, d.faces.f_2, d.faces.normal, d.faces.airDiamond

    def nodeval(n):
        return Pbordnode[n] if on[n] in BCdirichlet else P_node[n]

    def one(i, VA, VB, VC, VD, VL, VR):
        for k in range(3):
            out[k][i] = -1. * (f1[i][k] * (VA - VC) + f2[i][k] * (VB - VD) + nrm[i][k] * (VR - VL)) / air[i]

    for i in d.innerfaces:
        one(i, *[nodeval(k) for k in face_nodes4(d, i)], P_c[cid[i][0]], P_c[cid[i][1]])
    for i in neumannfaces:
        one(i, *[nodeval(k) for k in face_nodes4(d, i)], P_c[cid[i][0]], P_ghost[i])
    for i in dirichletfaces:
        VL = P_c[cid[i][0]]
        one(i, *[Pbordnode[k] for k in face_nodes4(d, i)], VL, 2. * Pbordface[i] - VL)
    return out
