"""Host-side mirror of the reference's call surface over libofea_hip.so.

Names follow the reference so that a driver written against it reads the same:
  ReadInfo-style deck reading      -> read_inp()        (src/data2cae.cpp)
  set_BCs::build_free_index        -> build_free_index() (src/set_bc.cpp:20-43)
  set_BCs::build_single_load       -> build_single_load() (src/set_bc.cpp:46-77)
  assamble_stiffness               -> AssambleStiffness  (src/include/assemble.h:35-72)
  solution_api(..., solver_type)   -> solution_api()     (src/linear_solution.cpp:17-64)
  simulation_post::reset_displacement -> reset_displacement() (src/post_process.cpp:18-40)
  simulation_post::get_cauchy_stress_3d -> get_cauchy_stress_3d() (src/post_process.cpp:79-168)
  data_process::export_dis_2_vtk   -> export_dis_2_vtk() (src/data2vtk.cpp:47-264)
  ReadInfo::CA_read_change          -> read_ca_info()     (src/data2cae.cpp:519-665)
  CAE_process::CA_ReAnalysis        -> CA_ReAnalysis()    (src/cae.cpp:317-376)
  simulation_post::reset_ca_displacement -> reset_ca_displacement() (src/post_process.cpp:42-76)
  CAE_process::implict_analysis     -> implict_analysis() (src/cae.cpp:198-301)
  CAE_process::implict_SFEManalysis -> implict_SFEManalysis() (src/cae.cpp:379-416)
All numerical work runs in the native library; this module only marshals arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi as A
from ._capi import check, f64, i32, lib, ptr


@dataclass
class Deck:
    """data_management fields filled by the reader (src/include/data_management.h:45-104)."""
    xyz: np.ndarray            # nd x 3
    conn: np.ndarray           # ne x 8, 1-based
    etype: np.ndarray          # ne (0 C3D4, 1 C3D8)
    fix_set: np.ndarray        # dis_bc_set_
    load_set: np.ndarray       # load_set_
    load_dof: int = 0
    load_value: float = 0.0
    E: float = float("nan")
    nu: float = float("nan")
    rho: float = float("nan")
    dt: float = 0.0
    t_total: float = 0.0
    resort: np.ndarray = field(default=None)
    F: np.ndarray = field(default=None)
    # keyword-level extensions (None when the deck's extension blocks are unreadable: see ext_error)
    f_nodal: np.ndarray = field(default=None)    # nd x 3, all *Cload and *Dsload
    v0: np.ndarray = field(default=None)         # nd x 3, *Initial Conditions type=VELOCITY
    bc_fixed: np.ndarray = field(default=None)   # 1-based ids of nodes with all DOFs constrained
    bc_partial: int = 0
    bc_nonzero: int = 0
    E_elem: np.ndarray = field(default=None)     # per-element section material (NaN without one)
    nu_elem: np.ndarray = field(default=None)
    rho_elem: np.ndarray = field(default=None)
    ext_error: str = ""

    @property
    def nd(self):
        return len(self.xyz)

    @property
    def ne(self):
        return len(self.etype)


def read_inp(path: str) -> Deck:
    """Read a reference-dialect .inp deck through the native reader."""
    L = lib()
    h = C.c_void_p()
    check(L.ofea_inp_open(path.encode(), C.byref(h)))
    try:
        nd, ne, nfix, nload = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(L.ofea_inp_sizes(h, C.byref(nd), C.byref(ne), C.byref(nfix), C.byref(nload)))
        xyz = np.zeros((nd.value, 3))
        conn = np.zeros((ne.value, 8), dtype=np.int32)
        etype = np.zeros(ne.value, dtype=np.int32)
        fix = np.zeros(nfix.value, dtype=np.int32)
        load = np.zeros(nload.value, dtype=np.int32)
        dof = C.c_int()
        vals = [C.c_double() for _ in range(6)]
        check(L.ofea_inp_get(h, ptr(xyz), ptr(conn), ptr(etype), ptr(fix), ptr(load), C.byref(dof),
                             *[C.byref(v) for v in vals]))
        ext = _read_ext(L, h, nd.value, ne.value)
    finally:
        L.ofea_inp_close(h)
    lv, E, nu, rho, dt, T = [v.value for v in vals]
    return Deck(xyz, conn, etype, fix, load, dof.value, lv, E, nu, rho, dt, T, **ext)


def _read_ext(L, h, nd, ne):
    f = np.zeros((nd, 3))
    rc = L.ofea_inp_nodal_loads(h, ptr(f))
    if rc != A.OFEA_OK:
        return {"ext_error": L.ofea_last_error().decode(errors="replace")}
    v0 = np.zeros((nd, 3))
    check(L.ofea_inp_initial_velocity(h, ptr(v0)))
    nf, npart, nnz = C.c_int(), C.c_int(), C.c_int()
    fixed = np.zeros(nd, dtype=np.int32)
    check(L.ofea_inp_boundary(h, C.byref(nf), ptr(fixed), C.byref(npart), C.byref(nnz)))
    Ee, nue, rhoe = np.zeros(ne), np.zeros(ne), np.zeros(ne)
    check(L.ofea_inp_materials(h, ptr(Ee), ptr(nue), ptr(rhoe)))
    return {"f_nodal": f, "v0": v0, "bc_fixed": fixed[:nf.value].copy(), "bc_partial": npart.value,
            "bc_nonzero": nnz.value, "E_elem": Ee, "nu_elem": nue, "rho_elem": rhoe}


def build_free_index(nd: int, fix_set) -> tuple[np.ndarray, int]:
    fix = i32(fix_set).copy()
    resort = np.zeros(nd, dtype=np.int32)
    nfree = lib().ofea_build_free_index(nd, ptr(fix), len(fix), ptr(resort))
    if nfree < 0:
        check(nfree)
    return resort, nfree


def build_single_load(n: int, resort, load_set, dof: int, value: float) -> np.ndarray:
    F = np.zeros(n)
    ls = i32(load_set)
    check(lib().ofea_build_single_load(n, ptr(i32(resort)), ptr(ls), len(ls), dof, value, ptr(F)))
    return F


def build_csr(conn, etype, resort):
    """build_CSR on the host (no device): CSC col_ptr (n+1) and row_idx (nnz)."""
    conn, etype, resort = i32(conn).reshape(-1, 8), i32(etype), i32(resort)
    n = C.c_int()
    nnz = C.c_longlong()
    L = lib()
    check(L.ofea_build_csr(len(resort), len(etype), ptr(conn), ptr(etype), ptr(resort), C.byref(n), C.byref(nnz),
                           None, None))
    col_ptr = np.zeros(n.value + 1, dtype=np.int32)
    row_idx = np.zeros(nnz.value, dtype=np.int32)
    check(L.ofea_build_csr(len(resort), len(etype), ptr(conn), ptr(etype), ptr(resort), None, None, ptr(col_ptr),
                           ptr(row_idx)))
    return col_ptr, row_idx


def reset_displacement(resort, u_free) -> np.ndarray:
    resort = i32(resort)
    out = np.zeros(3 * len(resort))
    check(lib().ofea_reset_displacement(len(resort), ptr(resort), ptr(f64(u_free)), ptr(out)))
    return out


def reset_ca_displacement(resort, u_free, idref2new, nd_new) -> np.ndarray:
    """Displacement of the modified model's nodes (3*nd_new) from the free-DOF
    solution on the original model's numbering."""
    resort, idref = i32(resort), i32(idref2new)
    out = np.zeros(3 * nd_new)
    check(lib().ofea_reset_ca_displacement(len(resort), ptr(resort), ptr(f64(u_free)), len(idref), ptr(idref), nd_new,
                                           ptr(out)))
    return out


@dataclass
class CAInfo:
    """Fields CA_read_change fills (coords_m_, node_topos_m_, ele_list_idx_m_, node_idx_m_)."""
    xyz: np.ndarray        # nodes_all x 3
    kind: np.ndarray       # nmod: 0 delete, 1 move, 2 add
    conn: np.ndarray       # nmod x 8, 0-based
    etype: np.ndarray      # nmod
    idref2new: np.ndarray  # nodes_all: new-model node, -1 removed


def read_ca_info(path: str) -> CAInfo:
    L = lib()
    h = C.c_void_p()
    check(L.ofea_ca_info_open(path.encode(), C.byref(h)))
    try:
        nd, nm, nmap = C.c_int(), C.c_int(), C.c_int()
        check(L.ofea_ca_info_sizes(h, C.byref(nd), C.byref(nm), C.byref(nmap)))
        xyz = np.zeros((nd.value, 3))
        kind = np.zeros(nm.value, dtype=np.int32)
        conn = np.zeros((nm.value, 8), dtype=np.int32)
        et = np.zeros(nm.value, dtype=np.int32)
        idref = np.zeros(nmap.value, dtype=np.int32)
        check(L.ofea_ca_info_get(h, ptr(xyz), ptr(kind), ptr(conn), ptr(et), ptr(idref)))
    finally:
        L.ofea_ca_info_close(h)
    return CAInfo(xyz, kind, conn, et, idref)


def export_dis_2_vtk(path: str, deck: Deck, u_full, scale: float = 1.0, node_fields=None):
    """Legacy ASCII VTK in the reference's dialect (u fields, optional S_* fields)."""
    nf = None if node_fields is None else f64(node_fields).reshape(7, -1)
    xyz, conn, et = f64(deck.xyz), i32(deck.conn), i32(deck.etype)
    check(lib().ofea_write_vtk(path.encode(), len(xyz), ptr(xyz), len(et), ptr(conn), ptr(et), ptr(f64(u_full)),
                               scale, ptr(nf)))


class Context:
    """One device context (ofea_ctx) holding a mesh, its CSC pattern and K."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().ofea_create(C.byref(self._h), device))
        self.n = 0
        self.nnz = 0
        self.nd = 0
        self.ne = 0

    def close(self):
        if self._h:
            lib().ofea_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- mesh / pattern
    def set_mesh(self, xyz, conn, etype, resort):
        xyz = f64(xyz).reshape(-1, 3)
        conn = i32(conn).reshape(-1, 8)
        etype = i32(etype)
        resort = i32(resort)
        self._keep = (xyz, conn, etype, resort)
        check(lib().ofea_set_mesh(self._h, len(xyz), ptr(xyz), len(etype), ptr(conn), ptr(etype), ptr(resort)))
        n = C.c_int()
        nnz = C.c_longlong()
        check(lib().ofea_pattern(self._h, C.byref(n), C.byref(nnz), None, None))
        self.n, self.nnz, self.nd, self.ne = n.value, nnz.value, len(xyz), len(etype)

    def pattern(self):
        col_ptr = np.zeros(self.n + 1, dtype=np.int32)
        row_idx = np.zeros(self.nnz, dtype=np.int32)
        check(lib().ofea_pattern(self._h, None, None, ptr(col_ptr), ptr(row_idx)))
        return col_ptr, row_idx

    # ---- assembly
    def set_elem_E(self, E_elem):
        if E_elem is None:
            check(lib().ofea_set_elem_E(self._h, None))
        else:
            check(lib().ofea_set_elem_E(self._h, ptr(f64(E_elem))))

    def assemble(self, E: float, nu: float, copy: bool = True):
        out = np.zeros(self.nnz) if copy else None
        check(lib().ofea_assemble(self._h, E, nu, ptr(out)))
        return out

    def spmv(self, x):
        y = np.zeros(self.n)
        check(lib().ofea_spmv(self._h, ptr(f64(x)), ptr(y)))
        return y

    # ---- solve
    def set_rhs(self, b):
        check(lib().ofea_set_rhs(self._h, ptr(f64(b))))

    def solve(self, b=None, method=A.PCG_JACOBI, rtol=1e-8, maxit=100000, copy=True):
        x = np.zeros(self.n) if copy else None
        it = C.c_int()
        rr = C.c_double()
        bb = None if b is None else f64(b)
        rc = check(lib().ofea_solve(self._h, method, ptr(bb), ptr(x), rtol, maxit, C.byref(it), C.byref(rr)),
                   allow=(A.OFEA_OK, A.OFEA_NOT_CONVERGED))
        return x, it.value, rr.value, rc == A.OFEA_OK

    # ---- explicit
    def lumped_mass(self, rho: float):
        m = np.zeros(self.nd)
        check(lib().ofea_lumped_mass(self._h, rho, ptr(m)))
        return m

    def explicit_run(self, E, nu, mass, f_ext, v0, fixed_mask, dt, nsteps, precision=A.FP32, copy=True):
        u = np.zeros(3 * self.nd) if copy else None
        v0a = None if v0 is None else f64(v0).ravel()
        check(lib().ofea_explicit_run(self._h, E, nu, ptr(f64(mass)), ptr(f64(f_ext).ravel()), ptr(v0a),
                                      ptr(i32(fixed_mask)), dt, nsteps, precision, ptr(u)))
        return u

    def explicit_analysis(self, E, nu, mass, f_ext, fixed_mask, time_total, time_step=0.0, output_gap=10,
                          v0=None, precision=A.FP64, max_steps=1_000_000, max_frames=None):
        """The reference's explicit loop with its time control (time_step 0 =
        automatic).  Returns dict(frames (k x 3nd), index, time, steps)."""
        max_frames = output_gap + 2 if max_frames is None else max_frames
        frames = np.zeros((max_frames, 3 * self.nd))
        idx = np.zeros(max_frames, dtype=np.int32)
        tt = np.zeros(max_frames)
        nf, ns = C.c_int(), C.c_int()
        v0a = None if v0 is None else f64(v0).ravel()
        check(lib().ofea_explicit_analysis(self._h, E, nu, ptr(f64(mass)), ptr(f64(f_ext).ravel()), ptr(v0a),
                                           ptr(i32(fixed_mask)), time_total, time_step, output_gap, precision,
                                           max_steps, max_frames, ptr(frames), ptr(idx), ptr(tt), C.byref(nf),
                                           C.byref(ns)))
        k = nf.value
        return dict(frames=frames[:k], index=idx[:k], time=tt[:k], steps=ns.value)

    # ---- geometric nonlinear
    def nl_assemble(self, E, nu, u_free, copy=True):
        val = np.zeros(self.nnz) if copy else None
        f = np.zeros(self.n) if copy else None
        check(lib().ofea_nl_assemble(self._h, E, nu, ptr(f64(u_free)), ptr(val), ptr(f)))
        return val, f

    def nl_solve(self, E, nu, F, n_incr=10, tol=1e-8, max_newton=25, method=A.PCG_BLOCK_JACOBI, inner_rtol=1e-10):
        u = np.zeros(self.n)
        nt, lt, rr = C.c_int(), C.c_int(), C.c_double()
        rc = check(lib().ofea_nl_solve(self._h, E, nu, ptr(f64(F)), n_incr, tol, max_newton, method, inner_rtol, ptr(u),
                                       C.byref(nt), C.byref(lt), C.byref(rr)), allow=(A.OFEA_OK, A.OFEA_NOT_CONVERGED))
        return u, nt.value, lt.value, rr.value, rc == A.OFEA_OK

    def nl_solve_ref(self, E, nu, F, res_limit=1.0e-6, max_loop=100, method=A.PCG_TWO_LEVEL, inner_rtol=1e-10):
        """Newton loop with the reference's stopping rule (src/cae.cpp:209-253).
        Returns (u, loops, pcg iterations, res_norm, converged)."""
        u = np.zeros(self.n)
        lp, lt, rn = C.c_int(), C.c_int(), C.c_double()
        rc = check(lib().ofea_nl_solve_ref(self._h, E, nu, ptr(f64(F)), res_limit, max_loop, method, inner_rtol,
                                           ptr(u), C.byref(lp), C.byref(lt), C.byref(rn)),
                   allow=(A.OFEA_OK, A.OFEA_NOT_CONVERGED))
        return u, lp.value, lt.value, rn.value, rc == A.OFEA_OK

    # ---- batched variants / load cases
    def assemble_batch(self, E_var, nu, copy=True, nvar=None):
        """K_v for every row of E_var (nvar x ne per-element moduli), one launch.
        E_var None reuses the moduli resident from the previous call (pass nvar)."""
        if E_var is None:
            nv = int(nvar)
        else:
            E_var = f64(E_var).reshape(-1, self.ne)
            nv = len(E_var)
        out = np.zeros((nv, self.nnz)) if copy else None
        check(lib().ofea_assemble_batch(self._h, nv, ptr(E_var), nu, ptr(out)))
        return out

    def solve_batch(self, B, rtol=1e-8, maxit=100000, copy=True, method=None):
        """B: nvar x nrhs x n right-hand sides; returns X, iters, relres, converged.
        method: None (ofea_solve_batch, block Jacobi) or A.PCG_BLOCK_JACOBI / A.PCG_TWO_LEVEL
        (ofea_solve_batch_pc)."""
        B = f64(B)
        nvar, nrhs, n = B.shape
        X = np.zeros_like(B) if copy else None
        it = np.zeros((nvar, nrhs), dtype=np.int32)
        rr = np.zeros((nvar, nrhs))
        if method is None:
            rc = lib().ofea_solve_batch(self._h, nvar, nrhs, ptr(B), ptr(X), rtol, maxit, ptr(it), ptr(rr))
        else:
            rc = lib().ofea_solve_batch_pc(self._h, nvar, nrhs, ptr(B), ptr(X), rtol, maxit, method, ptr(it), ptr(rr))
        rc = check(rc, allow=(A.OFEA_OK, A.OFEA_NOT_CONVERGED))
        return X, it, rr, rc == A.OFEA_OK

    def autotune(self):
        """Times the SpMV lane-group sizes 4..64 on the resident K; returns (best, times ms)."""
        best = C.c_int()
        times = np.zeros(5, dtype=np.float32)
        check(lib().ofea_autotune(self._h, C.byref(best), ptr(times)))
        return best.value, times

    def batch_tiles(self, nrhs):
        g, ms = C.c_int(), C.c_float()
        check(lib().ofea_batch_tiles(self._h, nrhs, C.byref(g), C.byref(ms)))
        return g.value, ms.value

    # ---- combined-approximation reanalysis
    def ca_delta(self, E, nu, kind, conn_mod, etype_mod, xyz_m=None, copy=True):
        """Delta K of the modified elements on the resident pattern (nnz values)."""
        kind, etype_mod = i32(kind), i32(etype_mod)
        conn_mod = i32(conn_mod).reshape(-1, 8)
        xm = None if xyz_m is None else f64(xyz_m).reshape(-1, 3)
        out = np.zeros(self.nnz) if copy else None
        check(lib().ofea_ca_delta(self._h, E, nu, len(kind), ptr(kind), ptr(conn_mod), ptr(etype_mod), ptr(xm),
                                  ptr(out)))
        return out

    def ca_solve(self, n_basis, u0, F, method=A.PCG_TWO_LEVEL, rtol=1e-12, maxit=200000):
        """CA solution (n), basis size kept, PCG iterations, (K_r, F_r, y)."""
        u = np.zeros(self.n)
        m, it = C.c_int(), C.c_int()
        red = np.zeros(n_basis * n_basis + 2 * n_basis)
        rc = check(lib().ofea_ca_solve(self._h, n_basis, ptr(f64(u0)), ptr(f64(F)), method, rtol, maxit, ptr(u),
                                       C.byref(m), C.byref(it), ptr(red)), allow=(A.OFEA_OK, A.OFEA_NOT_CONVERGED))
        k = m.value
        reduced = (red[:k * k].reshape(k, k), red[k * k:k * k + k], red[k * k + k:k * k + 2 * k])
        return u, k, it.value, reduced, rc == A.OFEA_OK

    # ---- node-based smoothed FEM
    def sfem_setup(self):
        """Smoothing domains and the smoothed pattern (replaces the context's pattern)."""
        check(lib().ofea_sfem_setup(self._h))
        n = C.c_int()
        nnz = C.c_longlong()
        check(lib().ofea_pattern(self._h, C.byref(n), C.byref(nnz), None, None))
        self.n, self.nnz = n.value, nnz.value

    def sfem_assemble(self, E: float, nu: float, copy: bool = True):
        out = np.zeros(self.nnz) if copy else None
        check(lib().ofea_sfem_assemble(self._h, E, nu, ptr(out)))
        return out

    # ---- post-processing
    def post_nodal(self, E, nu, u_full, quantity=A.POST_STRAIN, elem=False, copy=True):
        """(node fields 7 x nd, element fields 7 x ne or None) from the full
        displacement; u_full None uses the resident solution of the last
        solve, copy=False leaves the results on the device."""
        nv = np.zeros((7, self.nd)) if copy else None
        ev = np.zeros((7, self.ne)) if (elem and copy) else None
        uu = None if u_full is None else f64(u_full)
        check(lib().ofea_post_nodal(self._h, E, nu, ptr(uu), quantity, ptr(ev), ptr(nv)))
        return nv, ev

    # ---- timing on the context stream
    def timer_start(self):
        check(lib().ofea_timer_start(self._h))

    def timer_stop(self) -> float:
        ms = C.c_float()
        check(lib().ofea_timer_stop(self._h, C.byref(ms)))
        return ms.value

    def sync(self):
        check(lib().ofea_sync(self._h))


class AssambleStiffness:
    """assamble_stiffness (src/include/assemble.h:35-72): build_CSR then
    fill_CSR_sparse_mat, with the same public arrays."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or Context()
        self.num_row_ = self.num_col_ = self.num_nz_val_ = 0
        self.nz_val_ = None
        self.row_idx_ = None
        self.col_idx_ = None

    def build_CSR(self, deck: Deck):
        self.ctx.set_mesh(deck.xyz, deck.conn, deck.etype, deck.resort)
        self.col_idx_, self.row_idx_ = self.ctx.pattern()
        self.num_row_ = self.num_col_ = self.ctx.n
        self.num_nz_val_ = self.ctx.nnz

    def fill_CSR_sparse_mat(self, deck: Deck, E: float, nu: float):
        self.nz_val_ = self.ctx.assemble(E, nu)


SOLVERS = {"GPU_PCG": A.PCG_JACOBI, "GPU_PCG_BLOCK": A.PCG_BLOCK_JACOBI, "GPU_PCG_MIXED": A.PCG_MIXED,
           "GPU_PCG_MIXED32": A.PCG_MIXED32, "GPU_PCG_LEARNED": A.PCG_LEARNED,
           "GPU_PCG_LEARNED_MIXED": A.PCG_LEARNED_MIXED, "GPU_PCG_TWO_LEVEL": A.PCG_TWO_LEVEL,
           "GPU_PCG_TWO_LEVEL_MIXED": A.PCG_TWO_LEVEL_MIXED, "GPU_PCG_LEARNED_TWO_LEVEL": A.PCG_LEARNED_TWO_LEVEL,
           "GPU_PCG_LEARNED_TWO_LEVEL_MIXED": A.PCG_LEARNED_TWO_LEVEL_MIXED,
           "GPU_PCG_TWO_LEVEL_MIXED32": A.PCG_TWO_LEVEL_MIXED32,
           "GPU_PCG_LEARNED_TWO_LEVEL_MIXED32": A.PCG_LEARNED_TWO_LEVEL_MIXED32}


def solution_api(assam: AssambleStiffness, deck: Deck, solver_type: str = "GPU_PCG", rtol: float = 1e-10,
                 maxit: int = 200000):
    """solution_api (src/linear_solution.cpp:17-64) with the GPU solver
    strings; returns False for an unknown solver string like the reference."""
    if solver_type not in SOLVERS:
        return False
    x, it, rel, ok = assam.ctx.solve(deck.F, SOLVERS[solver_type], rtol, maxit)
    deck.single_dis_vec_ = x
    deck.solve_info = {"iters": it, "relres": rel, "converged": ok}
    return True


def get_cauchy_stress_3d(ctx: Context, deck: Deck, u_full, E: float | None = None, nu: float | None = None,
                         quantity=A.POST_STRAIN):
    """Nodal averages of the element values (7 x nd: xx, yy, zz, xy, yz, xz,
    Mises).  POST_STRAIN is what the reference computes and writes as S_*."""
    E = deck.E if E is None else E
    nu = deck.nu if nu is None else nu
    nv, _ = ctx.post_nodal(E, nu, u_full, quantity)
    deck.stress_node_mat_ = nv
    return nv


def CA_ReAnalysis(assam: AssambleStiffness, deck: Deck, info: CAInfo, deck_new: Deck, n_basis: int = 4,
                  E: float | None = None, nu: float | None = None, solver_type: str = "GPU_PCG_TWO_LEVEL",
                  rtol: float = 1e-12):
    """CAE_process::CA_ReAnalysis (src/cae.cpp:317-376) on the resident K0:
    Delta K of the changed elements, the n_basis-term CA basis, the reduced
    solve, and the displacement mapped onto the modified model's nodes.
    deck.single_dis_vec_ (the K0 solution) and deck.F are the original
    model's; returns (u on the modified model, 3*nd_new) and stores the
    free-DOF CA solution in deck.ca_dis_vec_."""
    E = deck.E if E is None else E
    nu = deck.nu if nu is None else nu
    ctx = assam.ctx
    ctx.ca_delta(E, nu, info.kind, info.conn, info.etype, info.xyz, copy=False)
    u, m, it, _, ok = ctx.ca_solve(n_basis, deck.single_dis_vec_, deck.F, SOLVERS[solver_type], rtol)
    deck.ca_dis_vec_ = u
    deck.ca_info_ = {"basis": m, "pcg_iters": it, "converged": ok}
    return reset_ca_displacement(deck.resort, u, info.idref2new, deck_new.nd)


def implict_analysis(deck: Deck, result_path: str | None = None, solver_type: str = "GPU_PCG_TWO_LEVEL",
                     ctx: Context | None = None, E: float | None = None, nu: float | None = None, rtol: float = 1e-10,
                     NLFEA: bool = False, res_lmit: float = 1.0e-6):
    """CAE_process::implict_analysis (src/cae.cpp:198-301): with NLFEA the
    reference's Newton loop on the total-Lagrangian tangent (full load,
    ||r||_2 / (3 nfree) > res_lmit, at most 100 loops; ofea_nl_solve_ref),
    else the linear static solve.  Clamp (Set-fix, plus nodes *Boundary fixes in all three DOFs),
    loads (every *Cload and *Dsload; on the reference's decks this equals
    build_single_load), material (E, nu of the deck, or per-element E of
    *Solid Section regions with one Poisson ratio), pattern + assembly,
    the GPU solve, reset_displacement, the nodal strain fields of
    get_cauchy_stress_3d and, with result_path, the VTK file of
    export_dis_2_vtk.  Returns (u_full nd x 3, nodal fields 7 x nd)."""
    if deck.ext_error:
        raise ValueError(f"deck extensions unreadable: {deck.ext_error}")
    if deck.bc_partial or deck.bc_nonzero:
        raise ValueError("*Boundary constrains single DOFs or prescribes values: the node-elimination "
                         "numbering clamps whole nodes only")
    fix = np.union1d(i32(deck.fix_set), i32(deck.bc_fixed)).astype(np.int32)
    deck.resort, nfree = build_free_index(deck.nd, fix)
    F = np.zeros(3 * nfree)
    free = deck.resort >= 0
    F.reshape(-1, 3)[deck.resort[free]] = deck.f_nodal[free]
    deck.F = F
    E = deck.E if E is None else E
    nu = deck.nu if nu is None else nu
    own = ctx is None
    ctx = ctx or Context()
    try:
        assam = AssambleStiffness(ctx)
        assam.build_CSR(deck)
        sec = deck.E_elem is not None and np.all(np.isfinite(deck.E_elem))
        if sec:
            nus = np.unique(deck.nu_elem)
            if len(nus) != 1:
                raise ValueError("section materials with different Poisson ratios: one nu per assembly")
            nu = float(nus[0])
            ctx.set_elem_E(deck.E_elem)
        else:
            ctx.set_elem_E(None)
        if NLFEA:
            if sec:
                raise ValueError("the Newton loop takes one modulus: per-element section moduli are not supported")
            if solver_type not in SOLVERS:
                raise ValueError(f"unknown solver {solver_type}")
            u, loops, pcg, res, ok = ctx.nl_solve_ref(E, nu, F, res_lmit, 100, SOLVERS[solver_type], rtol)
            deck.single_dis_vec_ = u
            deck.solve_info = {"loops": loops, "pcg_iters": pcg, "res_norm": res, "converged": ok}
            if not ok:
                raise RuntimeError(f"Newton loop did not reach res_lmit: {deck.solve_info}")
        else:
            assam.fill_CSR_sparse_mat(deck, E, nu)
            if not solution_api(assam, deck, solver_type, rtol=rtol):
                raise ValueError(f"unknown solver {solver_type}")
            if not deck.solve_info["converged"]:
                raise RuntimeError(f"solver did not converge: {deck.solve_info}")
        u_full = reset_displacement(deck.resort, deck.single_dis_vec_)
        fields = get_cauchy_stress_3d(ctx, deck, u_full, E, nu)
        if result_path:
            export_dis_2_vtk(result_path, deck, u_full, 1.0, fields)
        return u_full.reshape(-1, 3), fields
    finally:
        if own:
            ctx.close()


def build_explicit_load(nd: int, load_set, dof: int, value: float) -> np.ndarray:
    """OutFroce_ of src/cae.cpp:449-456: (int)value at DOF dof of every load
    node (the reference keeps the load value in an int)."""
    f = np.zeros(3 * nd)
    ls = i32(load_set)
    check(lib().ofea_build_explicit_load(nd, ptr(ls), len(ls), dof, value, ptr(f)))
    return f


def explicit_analysis(deck: Deck, result_path: str | None = None, ctx: Context | None = None,
                      E: float | None = None, nu: float | None = None, rho: float | None = None,
                      time_total: float | None = None, time_step: float | None = None, output_gap: int = 10,
                      precision: int = A.FP64, max_steps: int = 1_000_000):
    """CAE_process::explicit_analysis (src/cae.cpp:420-574): lumped mass at
    the deck's density, the single load of load_set / load_dof with its value
    truncated to an int, clamped Set-fix nodes, the deck's time step (0 =
    automatic) and total time (or the arguments), output_gap outputs; with
    result_path every output k is written as "<result_path>_<k>.vtk"
    (displacement fields, as the reference writes them).  Returns the dict of
    Context.explicit_analysis (frames nd*3 each, their output index and time,
    the step count)."""
    E = deck.E if E is None else E
    nu = deck.nu if nu is None else nu
    rho = deck.rho if rho is None else rho
    time_total = deck.t_total if time_total is None else time_total
    time_step = deck.dt if time_step is None else time_step
    resort, _ = build_free_index(deck.nd, deck.fix_set)
    fixed = (resort < 0).astype(np.int32)
    f_ext = build_explicit_load(deck.nd, deck.load_set, deck.load_dof, deck.load_value)
    own = ctx is None
    ctx = ctx or Context()
    try:
        ctx.set_mesh(deck.xyz, deck.conn, deck.etype, resort)
        mass = ctx.lumped_mass(rho)
        out = ctx.explicit_analysis(E, nu, mass, f_ext, fixed, time_total, time_step, output_gap,
                                    precision=precision, max_steps=max_steps)
        if result_path:
            for k, fr in zip(out["index"], out["frames"]):
                export_dis_2_vtk(f"{result_path}_{int(k)}.vtk", deck, fr, 1.0)
        return out
    finally:
        if own:
            ctx.close()


def implict_SFEManalysis(deck: Deck, result_path: str | None = None, solver_type: str = "GPU_PCG_TWO_LEVEL",
                         ctx: Context | None = None, E: float | None = None, nu: float | None = None,
                         rtol: float = 1e-10):
    """CAE_process::implict_SFEManalysis (src/cae.cpp:379-416): the
    node-based smoothed C3D4 model of the deck (SFEM3D::build,
    SFEM_build_CSR, SFEM_fill_CSR_sparse_mat), the GPU solve (the reference
    uses "SuperLU_func"), reset_displacement and the VTK file (u fields only,
    as the reference writes for this analysis).  Clamp and loads as in
    implict_analysis.  Returns u_full (nd x 3)."""
    if deck.ext_error:
        raise ValueError(f"deck extensions unreadable: {deck.ext_error}")
    fix = np.union1d(i32(deck.fix_set), i32(deck.bc_fixed)).astype(np.int32)
    deck.resort, nfree = build_free_index(deck.nd, fix)
    F = np.zeros(3 * nfree)
    free = deck.resort >= 0
    F.reshape(-1, 3)[deck.resort[free]] = deck.f_nodal[free]
    deck.F = F
    E = deck.E if E is None else E
    nu = deck.nu if nu is None else nu
    own = ctx is None
    ctx = ctx or Context()
    try:
        ctx.set_mesh(deck.xyz, deck.conn, deck.etype, deck.resort)
        ctx.sfem_setup()
        ctx.sfem_assemble(E, nu, copy=False)
        x, it, rel, ok = ctx.solve(F, SOLVERS[solver_type], rtol, 200000)
        if not ok:
            raise RuntimeError(f"solver did not converge: {it} iterations, relres {rel}")
        deck.single_dis_vec_ = x
        u_full = reset_displacement(deck.resort, x)
        if result_path:
            export_dis_2_vtk(result_path, deck, u_full, 1.0, None)
        return u_full.reshape(-1, 3)
    finally:
        if own:
            ctx.close()
