"""ORACLE (test infrastructure only -- never imported by the product path).

Loop-for-loop restatement of manapy's 2D/3D domain construction for small
meshes. Each function follows the cited reference function's iteration order
and arithmetic with plain Python loops, so it pins the vectorised builder in
``manapy_amd/ddm.py`` (which must produce identical arrays).

Pinned by: comparison with ``manapy_amd.ddm`` in ``tests/test_domain.py`` and by
the geometric invariants tested there (outward normals, |n| = edge length /
face area, sum of a cell's outward normals = 0, volumes summing to the box
measure, mirror centres at twice the distance to the face line).
"""
import math

import numpy as np


def build_domain_loops(elements, vertex, dim):
    """Return a dict of the reference arrays for a one-rank, non-periodic mesh."""
    nodeid = np.asarray(elements, dtype=np.int32)
    vertex = np.asarray(vertex, dtype=np.float64)
    C = len(nodeid)
    N = len(vertex)
    out = {}

    # Compute_2dcentervolumeOfCell (ddm_utils2d.py:433-479) / Compute_3dcentervolumeOfCell (ddm_utils3d.py:686-764)
    center = np.zeros((C, 3))
    volume = np.zeros(C)
    for i in range(C):
        nv = nodeid[i][-1]
        if dim == 2 and nv == 3:
            x1, y1 = vertex[nodeid[i][0]][0], vertex[nodeid[i][0]][1]
            x2, y2 = vertex[nodeid[i][1]][0], vertex[nodeid[i][1]][1]
            x3, y3 = vertex[nodeid[i][2]][0], vertex[nodeid[i][2]][1]
            center[i][0] = 1. / 3 * (x1 + x2 + x3)
            center[i][1] = 1. / 3 * (y1 + y2 + y3)
            volume[i] = (1. / 2) * abs((x1 - x2) * (y1 - y3) - (x1 - x3) * (y1 - y2))
        elif dim == 2 and nv == 4:
            xs = [vertex[nodeid[i][k]][0] for k in range(4)]
            ys = [vertex[nodeid[i][k]][1] for k in range(4)]
            sx = 0.
            sy = 0.
            for k in range(4):
                sx += xs[k]
                sy += ys[k]
            cx, cy = sx / 4, sy / 4
            center[i][0] = cx
            center[i][1] = cy
            for k in range(4):
                x1, y1 = xs[k], ys[k]
                x2, y2 = xs[(k + 1) % 4], ys[(k + 1) % 4]
                volume[i] += (1. / 2) * abs((x1 - x2) * (y1 - cy) - (x1 - cx) * (y1 - y2))
        elif dim == 3 and nv == 4:
            vs = [vertex[nodeid[i][k]][:3] for k in range(4)]
            for k in range(3):
                s = 0.
                for j in range(4):
                    s += vs[j][k]
                center[i][k] = s / 4
            u = [vs[1][k] - vs[0][k] for k in range(3)]
            v = [vs[2][k] - vs[0][k] for k in range(3)]
            w = [vs[3][k] - vs[0][k] for k in range(3)]
            wedge = [v[1] * w[2] - v[2] * w[1], v[2] * w[0] - v[0] * w[2], v[0] * w[1] - v[1] * w[0]]
            volume[i] = abs(u[0] * wedge[0] + u[1] * wedge[1] + u[2] * wedge[2]) / 6.0
        else:
            raise ValueError("oracle handles triangles, quads and tetrahedra")
    out['center'] = center
    out['volume'] = volume

    # create_node_cellid (ddm_utils2d.py:55-97)
    tmp = [[] for _ in range(N)]
    for i in range(C):
        for j in range(nodeid[i][-1]):
            tmp[nodeid[i][j]].append(i)
    maxlongn = max(len(t) for t in tmp)
    ncell = -np.ones((N, maxlongn + 1), dtype=np.int32)
    for i in range(N):
        for j, c in enumerate(tmp[i]):
            ncell[i][j] = c
        ncell[i][-1] = len(tmp[i])
    out['nodes_cellid'] = ncell
    tmp2 = [[] for _ in range(C)]
    for i in range(C):
        for j in range(nodeid[i][-1]):
            for c in tmp[nodeid[i][j]]:
                if c not in tmp2[i] and c != i:
                    tmp2[i].append(c)
        tmp2[i].sort()
    maxlongc = max(len(t) for t in tmp2)
    cellnid = -np.ones((C, maxlongc + 1), dtype=np.int32)
    for i in range(C):
        for j, c in enumerate(tmp2[i]):
            cellnid[i][j] = c
        cellnid[i][-1] = len(tmp2[i])
    out['cellnid'] = cellnid

    # create_2dfaces / create_3dfaces + sort/unique (domain.py:294-335)
    if dim == 2:
        loc = {3: [[0, 1], [1, 2], [2, 0]], 4: [[0, 1], [1, 2], [2, 3], [3, 0]]}
    else:
        loc = {4: [[0, 1, 2], [2, 3, 0], [0, 1, 3], [3, 1, 2]]}
    maxfacenid = 2 if dim == 2 else 3
    maxcellfid = max(len(loc[nv]) for nv in set(int(x) for x in nodeid[:, -1]))
    rows = []
    cellf = -np.ones((C, maxcellfid), dtype=np.int32)
    for i in range(C):
        for j, lf in enumerate(loc[int(nodeid[i][-1])]):
            cellf[i][j] = len(rows)
            rows.append([nodeid[i][v] for v in lf] + [maxfacenid])
    p_faces = np.array(rows, dtype=np.int32)
    p_faces[:, :-1] = np.sort(p_faces[:, :-1], axis=1)[:, ::-1]
    fnodeid, inv = np.unique(p_faces, axis=0, return_inverse=True)
    inv = np.asarray(inv).reshape(-1)
    F = len(fnodeid)
    faceid = -np.ones((C, maxcellfid + 1), dtype=np.int32)
    for i in range(C):
        for j in range(maxcellfid):
            if cellf[i][j] != -1:
                faceid[i][j] = inv[cellf[i][j]]
        faceid[i][-1] = int(np.sum(faceid[i] != -1))
    # create_cellsOfFace (ddm_utils2d.py:12-20)
    cellid = -np.ones((F, 2), dtype=np.int32)
    for i in range(C):
        for j in range(faceid[i][-1]):
            f = faceid[i][j]
            if cellid[f][0] == -1:
                cellid[f][0] = i
            if cellid[f][0] != i:
                cellid[f][1] = i
    # create_NeighborCellByFace (ddm_utils2d.py:32-53)
    cfid = [[] for _ in range(C)]
    for i in range(C):
        for j in range(faceid[i][-1]):
            f = faceid[i][j]
            if cellid[f][1] != -1:
                cfid[i].append(cellid[f][1] if i == cellid[f][0] else cellid[f][0])
    cellfid = -np.ones((C, maxcellfid + 1), dtype=np.int32)
    for i in range(C):
        for j, c in enumerate(cfid[i]):
            cellfid[i][j] = c
        cellfid[i][-1] = len(cfid[i])
    out['faces_nodeid_unoriented'] = fnodeid.copy()
    out['faceid'] = faceid
    out['faces_cellid'] = cellid
    out['cellfid'] = cellfid

    namen = [int(vertex[i][3]) for i in range(N)]
    namef = np.zeros(F, dtype=np.int32)
    normal = np.zeros((F, 3))
    mesure = np.zeros(F)
    centerf = np.zeros((F, 3))
    if dim == 2:
        # create_info_2dfaces (ddm_utils2d.py:507-551)
        for i in range(F):
            a, b = namen[fnodeid[i][0]], namen[fnodeid[i][1]]
            if cellid[i][1] == -1:
                if a == b:
                    namef[i] = a
                elif (a == 3 and b != 0) or (a != 0 and b == 3):
                    namef[i] = 3
                elif (a == 4 and b != 0) or (a != 0 and b == 4):
                    namef[i] = 4
                elif (a == 33 and b != 0) or (a != 0 and b == 33):
                    namef[i] = 33
                elif (a == 44 and b != 0) or (a != 0 and b == 44):
                    namef[i] = 44
                else:
                    namef[i] = 100
            p0 = vertex[fnodeid[i][0]]
            p1 = vertex[fnodeid[i][1]]
            nx = p0[1] - p1[1]
            ny = p1[0] - p0[0]
            for k in range(3):
                centerf[i][k] = 0.5 * (p0[k] + p1[k])
            sx = center[cellid[i][0]][0] - centerf[i][0]
            sy = center[cellid[i][0]][1] - centerf[i][1]
            if sx * nx + sy * ny > 0:
                normal[i][0], normal[i][1] = -1 * nx, -1 * ny
            else:
                normal[i][0], normal[i][1] = nx, ny
            mesure[i] = math.sqrt(normal[i][0] * normal[i][0] + normal[i][1] * normal[i][1])
    else:
        # create_info_3dfaces, triangle branch (ddm_utils3d.py:541-593, 673-678)
        for i in range(F):
            p = [vertex[fnodeid[i][k]][:3] for k in range(3)]
            for k in range(3):
                centerf[i][k] = 1. / 3 * (p[0][k] + p[1][k] + p[2][k])
            u = [p[1][k] - p[0][k] for k in range(3)]
            v = [p[2][k] - p[0][k] for k in range(3)]
            nrm = [0.5 * (u[1] * v[2] - u[2] * v[1]), 0.5 * (u[2] * v[0] - u[0] * v[2]), 0.5 * (u[0] * v[1] - u[1] * v[0])]
            mesure[i] = math.sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2])
            if cellid[i][1] == -1:
                nm = [namen[fnodeid[i][k]] for k in range(3)]
                if nm[0] == nm[1] and nm[0] == nm[2]:
                    namef[i] = nm[0]
                else:
                    namef[i] = 100
                    for tag in (5, 6, 3, 4, 55, 66, 33, 44):
                        hit = any(nm[j] == tag and all(nm[l] != 0 for l in range(3) if l != j) for j in range(3))
                        if hit:
                            namef[i] = tag
                            break
            s = [center[cellid[i][0]][k] - centerf[i][k] for k in range(3)]
            if s[0] * nrm[0] + s[1] * nrm[1] + s[2] * nrm[2] > 0:
                normal[i] = [-1 * nrm[0], -1 * nrm[1], -1 * nrm[2]]
            else:
                normal[i] = nrm
    out['faces_name'] = namef
    out['faces_normal'] = normal
    out['faces_mesure'] = mesure
    out['faces_center'] = centerf

    # face lists (domain.py:401-458)
    inner = [i for i in range(F) if namef[i] == 0]
    tags = (1, 2, 4, 3) if dim == 2 else (1, 2, 4, 3, 6, 5)
    bnd = sorted(i for t in tags for i in range(F) if namef[i] == t)
    out['innerfaces'] = np.array(inner, dtype=np.int32)
    out['boundaryfaces'] = np.array(bnd, dtype=np.int32)

    if dim == 2:
        # face_info_2d (ddm_utils2d.py:163-223)
        gcf = -np.ones((F, 3))
        gcn = [[] for _ in range(N)]
        for i in bnd:
            nod1 = fnodeid[i][1]
            nod2 = fnodeid[i][0]
            x1 = vertex[nod1]
            x2 = vertex[nod2]
            v1 = center[cellid[i][0]]
            gamma = (((v1[0] - x2[0]) * (x1[0] - x2[0]) + (v1[1] - x2[1]) * (x1[1] - x2[1]))
                     / ((x1[0] - x2[0]) * (x1[0] - x2[0]) + (x1[1] - x2[1]) * (x1[1] - x2[1])))
            kk0 = gamma * x1[0] + (1 - gamma) * x2[0]
            kk1 = gamma * x1[1] + (1 - gamma) * x2[1]
            v20 = 2 * kk0 + (-1 * v1[0])
            v21 = 2 * kk1 + (-1 * v1[1])
            gcf[i] = [v20, v21, gamma]
            gcn[nod1].append([v20, v21, cellid[i][0], namef[i], i])
            gcn[nod2].append([v20, v21, cellid[i][0], namef[i], i])
        out['faces_ghostcenter'] = gcf
        g = max(1, max(len(x) for x in gcn))
        ghostid = -np.ones((N, g + 1), dtype=np.int32)
        for i in range(N):
            for j, e in enumerate(gcn[i]):
                ghostid[i][j] = e[-1]
            ghostid[i][-1] = len(gcn[i])
        out['ghostid'] = ghostid
        # create_node_ghostid (ddm_utils2d.py:140-159)
        gn = [[] for _ in range(C)]
        for i in range(C):
            for j in range(nodeid[i][-1]):
                nod = nodeid[i][j]
                for k in range(ghostid[nod][-1]):
                    if ghostid[nod][k] not in gn[i]:
                        gn[i].append(ghostid[nod][k])
            gn[i].sort()
        mg = max(len(x) for x in gn)
        ghostnid = -np.ones((C, mg + 1), dtype=np.int32)
        for i in range(C):
            for j, f in enumerate(gn[i]):
                ghostnid[i][j] = f
            ghostnid[i][-1] = len(gn[i])
        out['ghostnid'] = ghostnid

        # variables_2d (ddm_utils2d.py:608-705)
        lx = np.zeros(N)
        ly = np.zeros(N)
        Rx_ = np.zeros(N)
        Ry_ = np.zeros(N)
        number = np.zeros(N, dtype=np.int32)
        for i in range(N):
            Ixx = Iyy = Ixy = 0.
            pts = [center[c] for c in tmp[i]] + [gcf[f] for f in (e[-1] for e in gcn[i])]
            for p in pts:
                Rx = p[0] - vertex[i][0]
                Ry = p[1] - vertex[i][1]
                Ixx += Rx * Rx
                Iyy += Ry * Ry
                Ixy += Rx * Ry
                Rx_[i] += Rx
                Ry_[i] += Ry
                number[i] += 1
            D = Ixx * Iyy - Ixy * Ixy
            lx[i] = (Ixy * Ry_[i] - Iyy * Rx_[i]) / D if D != 0 else float('nan')
            ly[i] = (Ixy * Rx_[i] - Ixx * Ry_[i]) / D if D != 0 else float('nan')
        out['lambda_x'] = lx
        out['lambda_y'] = ly
        out['R_x'] = Rx_
        out['R_y'] = Ry_
        out['number'] = number

        # face_gradient_info_2d (ddm_utils2d.py:553-605)
        air = np.zeros(F)
        params = np.zeros((4, F))
        for i in range(F):
            xy1 = vertex[fnodeid[i][0]][0:2]
            xy2 = vertex[fnodeid[i][1]][0:2]
            v1 = center[cellid[i][0]][0:2]
            v2 = center[cellid[i][1]][0:2] if namef[i] == 0 else gcf[i][0:2]
            f1 = [v1[k] - xy1[k] for k in range(2)]
            f2 = [xy2[k] - v1[k] for k in range(2)]
            f3 = [v2[k] - xy2[k] for k in range(2)]
            f4 = [xy1[k] - v2[k] for k in range(2)]
            n1, n2 = normal[i][0], normal[i][1]
            air[i] = 0.5 * ((xy2[0] - xy1[0]) * (v2[1] - v1[1]) + (v1[0] - v2[0]) * (xy2[1] - xy1[1]))
            k = 1. / (2. * air[i])
            params[0][i] = k * ((f1[1] + f2[1]) * n1 - (f1[0] + f2[0]) * n2)
            params[1][i] = k * ((f2[1] + f3[1]) * n1 - (f2[0] + f3[0]) * n2)
            params[2][i] = k * ((f3[1] + f4[1]) * n1 - (f3[0] + f4[0]) * n2)
            params[3][i] = k * ((f4[1] + f1[1]) * n1 - (f4[0] + f1[0]) * n2)
        out['airDiamond'] = air
        out['params'] = params
    return out
