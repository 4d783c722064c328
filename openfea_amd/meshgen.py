"""Synthetic structured meshes and loads for the benchmark workloads.

Conventions (mm, MPa, t):
  node id = 1 + k + (nz+1)(j + (ny+1) i), x slowest; C3D8 connectivity in the
  Abaqus order n1(i,j,k) n2(i+1,j,k) n3(i+1,j+1,k) n4(i,j+1,k), n5..n8 at k+1
  (positive detJ); C3D4 by the Kuhn split of each cube into the 6 tets along
  its (0,0,0)-(1,1,1) diagonal, each ordered for positive detJ.
  The clamp set is every node with x = 0.  Random draws use numpy's PCG64.
"""
from __future__ import annotations

from dataclasses import dataclass
from itertools import permutations

import numpy as np

E_STEEL = 210000.0
NU_STEEL = 0.3
RHO_STEEL = 7.85e-9


@dataclass
class Mesh:
    xyz: np.ndarray       # nd x 3
    conn: np.ndarray      # ne x 8, 1-based (C3D4 uses the first 4)
    etype: np.ndarray     # ne
    fix_set: np.ndarray   # 1-based ids of clamped nodes
    dims: tuple           # (nx, ny, nz, h)

    @property
    def nd(self):
        return len(self.xyz)

    @property
    def ne(self):
        return len(self.etype)


def _node_ids(nx, ny, nz):
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    return (k + (nz + 1) * (j + (ny + 1) * i)).astype(np.int64)


def _coords(nx, ny, nz, h):
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    return np.stack([i.ravel() * h, j.ravel() * h, k.ravel() * h], axis=1).astype(np.float64)


def _jitter(xyz, nx, ny, nz, h, amp, seed):
    if amp <= 0:
        return xyz
    rng = np.random.default_rng(seed)
    L, W, H = nx * h, ny * h, nz * h
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    tol = 1e-9 * h
    interior = (x > tol) & (x < L - tol) & (y > tol) & (y < W - tol) & (z > tol) & (z < H - tol)
    d = rng.uniform(-amp * h, amp * h, size=xyz.shape)
    out = xyz.copy()
    out[interior] += d[interior]
    return out


def hex_block(nx, ny, nz, h=1.0, jitter=0.0, seed=0) -> Mesh:
    nid = _node_ids(nx, ny, nz)
    xyz = _jitter(_coords(nx, ny, nz, h), nx, ny, nz, h, jitter, seed)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    c = np.stack([nid[i, j, k], nid[i + 1, j, k], nid[i + 1, j + 1, k], nid[i, j + 1, k],
                  nid[i, j, k + 1], nid[i + 1, j, k + 1], nid[i + 1, j + 1, k + 1], nid[i, j + 1, k + 1]], axis=1)
    conn = (c + 1).astype(np.int32)
    etype = np.ones(len(conn), dtype=np.int32)
    fix = (nid[0].ravel() + 1).astype(np.int32)
    return Mesh(xyz, conn, etype, np.sort(fix), (nx, ny, nz, h))


def kuhn_tets(nx, ny, nz, h=1.0, jitter=0.0, seed=0) -> Mesh:
    nid = _node_ids(nx, ny, nz)
    xyz = _jitter(_coords(nx, ny, nz, h), nx, ny, nz, h, jitter, seed)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    tets = []
    for perm in permutations(range(3)):
        p = np.zeros(3, dtype=np.int64)
        verts = [p.copy()]
        for ax in perm:
            p[ax] += 1
            verts.append(p.copy())
        ids = [nid[i + v[0], j + v[1], k + v[2]] for v in verts]
        # orientation: det of the edge vectors from vertex 0 equals sign(perm)
        sgn = np.linalg.det(np.array([verts[1] - verts[0], verts[2] - verts[0], verts[3] - verts[0]], float))
        if sgn < 0:
            ids[1], ids[2] = ids[2], ids[1]
        tets.append(np.stack(ids, axis=1))
    c = np.stack(tets, axis=1).reshape(-1, 4)  # 6 tets per cube, cube-major
    conn = np.zeros((len(c), 8), dtype=np.int32)
    conn[:, :4] = c + 1
    X = xyz[c]
    det = np.einsum("ij,ij->i", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))
    if not np.all(det > 0):
        raise ValueError("jitter inverted a tetrahedron")
    etype = np.zeros(len(conn), dtype=np.int32)
    fix = (nid[0].ravel() + 1).astype(np.int32)
    return Mesh(xyz, conn, etype, np.sort(fix), (nx, ny, nz, h))


def free_index(nd, fix_set):
    """Free-node ranks in ascending node order (-1 = clamped)."""
    resort = np.full(nd, -1, dtype=np.int32)
    free = np.ones(nd, dtype=bool)
    free[np.asarray(fix_set, dtype=np.int64) - 1] = False
    resort[free] = np.arange(free.sum(), dtype=np.int32)
    return resort


def face_traction(mesh: Mesh, t, x_face=None):
    """Consistent nodal forces (nd x 3) of a uniform traction vector t on the
    x = L face: p A/4 per quad corner (the face is not jittered, so its quads
    are rectangles) and p A/3 per triangle corner."""
    nx, ny, nz, h = mesh.dims
    L = nx * h if x_face is None else x_face
    t = np.asarray(t, dtype=np.float64).reshape(3)
    on = np.abs(mesh.xyz[:, 0] - L) < 1e-9 * max(h, 1.0)
    f = np.zeros((mesh.nd, 3))
    for et, faces in ((1, [(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]),
                      (0, [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)])):
        sel = mesh.etype == et
        if not np.any(sel):
            continue
        cn = mesh.conn[sel] - 1
        for fc in faces:
            ids = cn[:, list(fc)]
            mask = np.all(on[ids], axis=1)
            if not np.any(mask):
                continue
            ids = ids[mask]
            P = mesh.xyz[ids]
            if len(fc) == 3:
                area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
                share = area / 3.0
            else:
                area = 0.5 * (np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1) +
                              np.linalg.norm(np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]), axis=1))
                share = area / 4.0
            for a in range(len(fc)):
                np.add.at(f, ids[:, a], share[:, None] * t[None, :])
    return f


def free_rhs(f_nodes, resort):
    """Nodal force array (nd x 3) -> free-DOF load vector."""
    n = 3 * int(resort.max() + 1)
    F = np.zeros(n)
    free = resort >= 0
    r = resort[free]
    for c in range(3):
        F[3 * r + c] = f_nodes[free, c]
    return F


def element_centroids(mesh: Mesh):
    nn = np.where(mesh.etype == 1, 8, 4)
    X = mesh.xyz[np.maximum(mesh.conn - 1, 0)]
    s = np.zeros((mesh.ne, 3))
    for a in range(8):
        use = a < nn
        s[use] += X[use, a]
    return s / nn[:, None]


def design_field(mesh: Mesh, variant: int):
    """Per-element Young's modulus 210000 * f_v with
    f_v = 1 + 0.5 S / sum|a_k|, S = sum_k a_k sin(pi (kx x/L + ky y/W + kz z/H) + phi_k),
    drawn from seed = variant; f_v lies in [0.5, 1.5]."""
    nx, ny, nz, h = mesh.dims
    L, W, H = nx * h, ny * h, nz * h
    rng = np.random.default_rng(variant)
    a = rng.uniform(-1.0, 1.0, 4)
    phi = rng.uniform(0.0, 2 * np.pi, 4)
    kw = rng.integers(0, 3, size=(4, 3))
    c = element_centroids(mesh)
    S = np.zeros(mesh.ne)
    for q in range(4):
        S += a[q] * np.sin(np.pi * (kw[q, 0] * c[:, 0] / L + kw[q, 1] * c[:, 1] / W + kw[q, 2] * c[:, 2] / H) + phi[q])
    return E_STEEL * (1.0 + 0.5 * S / np.abs(a).sum())


def load_case(variant: int, case: int):
    """Traction vector: unit direction ~ normalised N(0, I3), magnitude U(1, 10) MPa."""
    rng = np.random.default_rng(1000 + variant * 8 + case)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    return d * rng.uniform(1.0, 10.0)


def batched_loads(mesh: Mesh, resort, nvar: int, nrhs: int):
    """Right-hand sides [variant][case][free DOF] of the load cases
    load_case(v, k) on the x = L face.  The consistent face load is linear in
    the traction vector, so three unit-traction loads are combined."""
    basis = np.stack([free_rhs(face_traction(mesh, np.eye(3)[c]), resort) for c in range(3)])
    B = np.zeros((nvar, nrhs, basis.shape[1]))
    for v in range(nvar):
        for k in range(nrhs):
            B[v, k] = load_case(v, k) @ basis
    return B


def initial_velocity(mesh: Mesh, amp=1000.0):
    """v_z(x) = amp sin(pi x / L) (nd x 3)."""
    nx, ny, nz, h = mesh.dims
    L = nx * h
    v = np.zeros((mesh.nd, 3))
    v[:, 2] = amp * np.sin(np.pi * mesh.xyz[:, 0] / L)
    return v


# workloads (sizes of the benchmark configurations)
def cantilever_c3d8(nx=20, ny=10, nz=10, h=1.0):
    """Linear static C3D8 block clamped at x=0, -10 MPa z traction on x=L."""
    m = hex_block(nx, ny, nz, h)
    resort = free_index(m.nd, m.fix_set)
    F = free_rhs(face_traction(m, (0.0, 0.0, -10.0)), resort)
    return m, resort, F


def cantilever_c3d4(nx=60, ny=30, nz=30, h=1.0, jitter=0.2, seed=20261203):
    """Linear static Kuhn C3D4 block, interior nodes jittered, -5 MPa z traction."""
    m = kuhn_tets(nx, ny, nz, h, jitter, seed)
    resort = free_index(m.nd, m.fix_set)
    F = free_rhs(face_traction(m, (0.0, 0.0, -5.0)), resort)
    return m, resort, F


# Abaqus face numbering (0-based local nodes): C3D8 S1..S6, C3D4 S1..S4
HEX_FACES = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0)]
TET_FACES = [(0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0)]


def faces_on_plane(mesh: Mesh, axis=0, value=None):
    """(element index, Abaqus face number 1-based) of the element faces lying on
    the plane x_axis = value (default: x = L)."""
    nx, ny, nz, h = mesh.dims
    value = nx * h if value is None else value
    on = np.abs(mesh.xyz[:, axis] - value) < 1e-9 * max(h, 1.0)
    out = []
    for e in range(mesh.ne):
        table = HEX_FACES if mesh.etype[e] == 1 else TET_FACES
        cn = mesh.conn[e] - 1
        for f, fl in enumerate(table):
            if all(on[cn[a]] for a in fl):
                out.append((e, f + 1))
    return out


def write_inp(path, mesh: Mesh, traction=None, pressure=None, cloads=(), velocity=None, materials=None,
              boundary="ENCASTRE", dynamic=None):
    """Writes mesh as a deck in the reference's dialect (count lines after the
    first *Nset/*Elset, Set-fix/Set-load node sets, *Cload, *Elastic) plus the
    keyword extensions: a *Surface of the x = L faces with a TRVEC traction
    (traction = vector) or pressure (*Dsload, P), extra point loads (cloads:
    (node id, dof, value)), *Initial Conditions type=VELOCITY (velocity: nd x 3),
    per-region materials (materials: list of (element mask, E, nu, rho)) and
    *Boundary on Set-fix.  dynamic = (dt, T) adds the *Dynamic, Explicit line."""
    nd, ne = mesh.nd, mesh.ne
    L = []
    L.append("*Node")
    for i, (x, y, z) in enumerate(mesh.xyz):
        L.append(f"{i + 1}, {float(x)!r}, {float(y)!r}, {float(z)!r}")
    for et, name, nn in ((1, "C3D8", 8), (0, "C3D4", 4)):
        ids = np.nonzero(mesh.etype == et)[0]
        if len(ids) == 0:
            continue
        L.append(f"*Element, type={name}")
        for e in ids:
            L.append(", ".join([str(e + 1)] + [str(v) for v in mesh.conn[e, :nn]]))
    L += ["*Nset", f"1, {nd}, 1", "*Elset", f"1, {ne}, 1"]
    L.append("*Nset, nset=Set-fix")
    fix = list(mesh.fix_set)
    for k in range(0, len(fix), 16):
        L.append(", ".join(str(v) for v in fix[k:k + 16]))
    nx, ny, nz, h = mesh.dims
    tip = np.nonzero(np.abs(mesh.xyz[:, 0] - nx * h) < 1e-9 * max(h, 1.0))[0] + 1
    L.append("*Nset, nset=Set-load")
    for k in range(0, len(tip), 16):
        L.append(", ".join(str(v) for v in tip[k:k + 16]))
    mats = materials or [(np.ones(ne, bool), E_STEEL, NU_STEEL, RHO_STEEL)]
    for m, (mask, E, nu, rho) in enumerate(mats):
        ids = np.nonzero(mask)[0] + 1
        L.append(f"*Elset, elset=Region-{m + 1}")
        for k in range(0, len(ids), 16):
            L.append(", ".join(str(v) for v in ids[k:k + 16]))
        L.append(f"*Solid Section, elset=Region-{m + 1}, material=Mat-{m + 1}")
        L.append(",")
    faces = faces_on_plane(mesh)
    by_face = {}
    for e, f in faces:
        by_face.setdefault(f, []).append(e + 1)
    for f, ids in sorted(by_face.items()):
        L.append(f"*Elset, elset=_Surf-L_S{f}, internal")
        for k in range(0, len(ids), 16):
            L.append(", ".join(str(v) for v in ids[k:k + 16]))
    L.append("*Surface, type=ELEMENT, name=Surf-L")
    for f in sorted(by_face):
        L.append(f"_Surf-L_S{f}, S{f}")
    L.append("*End Assembly")
    for m, (mask, E, nu, rho) in enumerate(mats):
        L += [f"*Material, name=Mat-{m + 1}", "*Density", f"{float(rho)!r},", "*Elastic", f"{float(E)!r}, {float(nu)!r}"]
    if velocity is not None:
        L.append("*Initial Conditions, type=VELOCITY")
        for i in range(nd):
            for c in range(3):
                if velocity[i, c] != 0.0:
                    L.append(f"{i + 1}, {c + 1}, {float(velocity[i, c])!r}")
    L.append("** STEP")
    if dynamic is not None:
        L += ["*Dynamic, Explicit", f"{float(dynamic[0])!r}, {float(dynamic[1])!r}"]
    else:
        L += ["*Static", "1., 1., 1e-05, 1."]
    if boundary:
        L += ["*Boundary", f"Set-fix, {boundary}"]
    L += ["*Cload", "Set-load, 3, 0."]
    for node, dof, val in cloads:
        L.append(f"{node}, {dof}, {float(val)!r}")
    if traction is not None:
        t = np.asarray(traction, dtype=np.float64)
        mag = float(np.linalg.norm(t))
        d = t / mag if mag > 0 else np.array([0.0, 0.0, 1.0])
        L += ["*Dsload", f"Surf-L, TRVEC, {mag!r}, {float(d[0])!r}, {float(d[1])!r}, {float(d[2])!r}"]
    if pressure is not None:
        L += ["*Dsload", f"Surf-L, P, {float(pressure)!r}"]
    L.append("*End Step")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")
