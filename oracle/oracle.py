"""ctypes front-end of the CPU checker (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path.  Arrays follow the layout documented in ofea_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.c_double
_i = C.c_int
_ll = C.c_longlong
_pd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_pi = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.or_dmat.argtypes = [_d, _d, _pd]
        L.or_c3d8_ke.argtypes = [_pd, _pd, _pd]
        L.or_c3d4_ke.argtypes = [_pd, _pd, _pd]
        L.or_c3d4_ke.restype = _d
        L.or_c3d8_nl.argtypes = [_pd, _pd, _pd, _pd, _pd]
        L.or_c3d4_nl.argtypes = [_pd, _pd, _pd, _pd, _pd]
        L.or_c3d8_mass.argtypes = [_pd, _d, _pd]
        L.or_c3d4_mass.argtypes = [_pd, _d, _pd]
        L.or_free_index.argtypes = [_i, _pi, _i, _pi]
        L.or_free_index.restype = _i
        L.or_single_load.argtypes = [_i, _pi, _pi, _i, _i, _d, _pd]
        L.or_build_csc.argtypes = [_i, _i, _pi, _pi, _pi, _i, _vp, _vp]
        L.or_build_csc.restype = _ll
        L.or_fill_csc.argtypes = [_i, _pi, _pi, _pd, _pi, _d, _d, _vp, _pi, _pi, _ll, _pd]
        L.or_nl_fill.argtypes = [_i, _pi, _pi, _pd, _pi, _d, _d, _pd, _pi, _pi, _ll, _pd, _pd, _i]
        L.or_spmv_csc.argtypes = [_i, _pi, _pi, _pd, _pd, _pd]
        L.or_pcg_jacobi.argtypes = [_i, _pi, _pi, _pd, _pd, _pd, _d, _i, C.POINTER(_d)]
        L.or_pcg_jacobi.restype = _i
        L.or_reset_displacement.argtypes = [_i, _pi, _pd, _pd]
        L.or_lumped_mass.argtypes = [_i, _i, _pi, _pi, _pd, _d, _pd]
        L.or_explicit_run.argtypes = [_i, _i, _pi, _pi, _pd, _d, _d, _pd, _pd, _vp, _pi, _d, _i, _pd]
        L.or_explicit_recurrence.argtypes = [_i, _pi, _pi, _pd, _pd, _pi, _pd, _d, _i, _pd, _i]
        L.or_crit_timestep.argtypes = [_i, _pi, _pi, _pd, _d, _d, _d]
        L.or_crit_timestep.restype = _d
        L.or_explicit_timed.argtypes = [_i, _i, _pi, _pi, _pd, _d, _d, _pd, _pd, _vp, _pi, _d, _d, _i, _i, _i,
                                        _pd, _pi, _pd, C.POINTER(_i), C.POINTER(_i), _vp]
        L.or_explicit_timed.restype = _i
        L.or_matfree_apply.argtypes = [_i, _pi, _pi, _pd, _pi, _d, _d, _vp, _pd, _pd, _i]
        L.or_ca_fill.argtypes = [_i, _pi, _pi, _pi, _pd, _pd, _pi, _d, _d, _pi, _pi, _ll, _pd]
        L.or_ca_fill.restype = _i
        L.or_post_elem.argtypes = [_i, _pi, _pi, _pd, _pd, _d, _d, _vp, _i, _pd]
        L.or_post_node.argtypes = [_i, _i, _pi, _pi, _pd, _pd]
        L.or_sfem_build_csc.argtypes = [_i, _i, _pi, _pd, _pi, _i, _vp, _vp]
        L.or_sfem_build_csc.restype = _ll
        L.or_sfem_fill.argtypes = [_i, _i, _pi, _pd, _pi, _d, _d, _pi, _pi, _ll, _pd]
        L.or_sfem_grads.argtypes = [_i, _i, _pi, _pd, _vp, _vp, _vp]
        L.or_sfem_grads.restype = _ll
        _LIB = L
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def dmat(E, nu):
    D = np.zeros(36)
    lib().or_dmat(E, nu, D)
    return D.reshape(6, 6)


def c3d8_ke(X, E, nu):
    Ke = np.zeros(576)
    lib().or_c3d8_ke(_f64(X).ravel(), _f64(dmat(E, nu)).ravel(), Ke)
    return Ke.reshape(24, 24)


def c3d4_ke(X, E, nu):
    Ke = np.zeros(144)
    lib().or_c3d4_ke(_f64(X).ravel(), _f64(dmat(E, nu)).ravel(), Ke)
    return Ke.reshape(12, 12)


def c3d8_nl(X, u, E, nu):
    K = np.zeros(576)
    f = np.zeros(24)
    lib().or_c3d8_nl(_f64(X).ravel(), _f64(u).ravel(), _f64(dmat(E, nu)).ravel(), K, f)
    return K.reshape(24, 24), f


def c3d4_nl(X, u, E, nu):
    K = np.zeros(144)
    f = np.zeros(12)
    lib().or_c3d4_nl(_f64(X).ravel(), _f64(u).ravel(), _f64(dmat(E, nu)).ravel(), K, f)
    return K.reshape(12, 12), f


def c3d8_mass(X, rho):
    m = np.zeros(8)
    lib().or_c3d8_mass(_f64(X).ravel(), rho, m)
    return m


def c3d4_mass(X, rho):
    m = np.zeros(4)
    lib().or_c3d4_mass(_f64(X).ravel(), rho, m)
    return m


def free_index(nd, fix_set):
    fix = _i32(fix_set).copy()
    resort = np.zeros(nd, dtype=np.int32)
    nfree = lib().or_free_index(nd, fix, len(fix), resort)
    return resort, nfree


def single_load(n, resort, load_set, dof, value):
    F = np.zeros(n)
    ls = _i32(load_set)
    lib().or_single_load(n, _i32(resort), ls, len(ls), dof, value, F)
    return F


def build_csc(conn, etype, resort, n):
    conn = _i32(conn)
    etype = _i32(etype)
    resort = _i32(resort)
    nd = len(resort)
    ne = len(etype)
    col_ptr = np.zeros(n + 1, dtype=np.int32)
    nnz_upper = lib().or_build_csc(nd, ne, conn, etype, resort, n, None, None)
    row_idx = np.zeros(max(nnz_upper, 1), dtype=np.int32)
    nnz = lib().or_build_csc(nd, ne, conn, etype, resort, n, col_ptr.ctypes.data, row_idx.ctypes.data)
    return col_ptr, row_idx[:nnz].copy()


def fill_csc(conn, etype, xyz, resort, E, nu, col_ptr, row_idx, E_elem=None):
    val = np.zeros(len(row_idx))
    Ee = None if E_elem is None else _f64(E_elem)
    lib().or_fill_csc(len(etype), _i32(conn), _i32(etype), _f64(xyz).ravel(), _i32(resort), E, nu,
                      None if Ee is None else Ee.ctypes.data, _i32(col_ptr), _i32(row_idx), len(row_idx), val)
    return val


_REF = None


def ref_lib():
    """oracle/_ref/libofea_ref.so (the reference's element code, built by
    oracle/Makefile where the reference checkout exists), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libofea_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        L.ref_fill_csc.argtypes = [_i, _pi, _pi, _pd, _pi, _d, _vp, _d, _pi, _pi, _pd, _i, _i]
        L.ref_fill_csc.restype = _i
        for nm in ("ref_c3d8_timestep", "ref_c3d4_timestep"):
            getattr(L, nm).argtypes = [_pd, _d, _d, _d]
            getattr(L, nm).restype = _d
        _REF = L
    return _REF


def ref_fill_csc(conn, etype, xyz, resort, E, nu, col_ptr, row_idx, threads=1, E_elem=None):
    """K values from the reference's own element stiffness and the scatter of
    fill_CSR_sparse_mat (ref_harness.cpp), on `threads` cores; (values,
    threads used), or None without oracle/_ref."""
    L = ref_lib()
    if L is None:
        return None
    val = np.zeros(len(row_idx))
    Ee = None if E_elem is None else _f64(E_elem)
    used = L.ref_fill_csc(len(etype), _i32(conn), _i32(etype), _f64(xyz).ravel(), _i32(resort), E,
                          None if Ee is None else Ee.ctypes.data, nu, _i32(col_ptr), _i32(row_idx), val,
                          len(col_ptr) - 1, int(threads))
    return val, used


def ref_crit_timestep(conn, etype, xyz_cur, E, nu, dt_in=np.finfo(np.float64).max):
    """UpdateTimeStep's element loop over the reference's own update_timestep
    (src/explicit_tools.cpp:21-31), or None without oracle/_ref."""
    L = ref_lib()
    if L is None:
        return None
    xyz_cur = _f64(xyz_cur).reshape(-1, 3)
    conn = _i32(conn).reshape(-1, 8)
    dt = dt_in
    for e in range(len(etype)):
        nn = 8 if etype[e] == 1 else 4
        X = np.ascontiguousarray(xyz_cur[conn[e, :nn] - 1]).ravel()
        dt = (L.ref_c3d8_timestep if nn == 8 else L.ref_c3d4_timestep)(X, E, nu, dt)
    return dt


def nl_fill(conn, etype, xyz, resort, E, nu, u_free, col_ptr, row_idx):
    val = np.zeros(len(row_idx))
    n = len(col_ptr) - 1
    f = np.zeros(n)
    lib().or_nl_fill(len(etype), _i32(conn), _i32(etype), _f64(xyz).ravel(), _i32(resort), E, nu, _f64(u_free),
                     _i32(col_ptr), _i32(row_idx), len(row_idx), val, f, n)
    return val, f


def spmv_csc(col_ptr, row_idx, val, x):
    n = len(col_ptr) - 1
    y = np.zeros(n)
    lib().or_spmv_csc(n, _i32(col_ptr), _i32(row_idx), _f64(val), _f64(x), y)
    return y


def pcg_jacobi(col_ptr, row_idx, val, b, rtol=1e-8, maxit=100000):
    n = len(col_ptr) - 1
    x = np.zeros(n)
    rr = C.c_double(0.0)
    it = lib().or_pcg_jacobi(n, _i32(col_ptr), _i32(row_idx), _f64(val), _f64(b), x, rtol, maxit, C.byref(rr))
    return x, it, rr.value


def reset_displacement(resort, u_free):
    nd = len(resort)
    out = np.zeros(3 * nd)
    lib().or_reset_displacement(nd, _i32(resort), _f64(u_free), out)
    return out


def lumped_mass(conn, etype, xyz, rho):
    nd = len(xyz)
    m = np.zeros(nd)
    lib().or_lumped_mass(nd, len(etype), _i32(conn), _i32(etype), _f64(xyz).ravel(), rho, m)
    return m


def explicit_recurrence(col_ptr, row_idx, val, minv, fixed, v0, dt, nsteps, threads=1):
    """u after nsteps of the direct recurrence u += v dt (u = 0 at fixed DOFs),
    v -= minv (K u) dt on all DOFs (K symmetric, CSC), on `threads` cores."""
    n = len(col_ptr) - 1
    u = np.zeros(n)
    lib().or_explicit_recurrence(n, _i32(col_ptr), _i32(row_idx), _f64(val), _f64(minv), _i32(fixed), _f64(v0),
                                 dt, nsteps, u, int(threads))
    return u


def explicit_run(conn, etype, xyz, E, nu, mass, f_ext, v0, fixed_mask, dt, nsteps):
    nd = len(xyz)
    u = np.zeros(3 * nd)
    v0a = None if v0 is None else _f64(v0).ravel()
    lib().or_explicit_run(nd, len(etype), _i32(conn), _i32(etype), _f64(xyz).ravel(), E, nu, _f64(mass),
                          _f64(f_ext).ravel(), None if v0a is None else v0a.ctypes.data, _i32(fixed_mask), dt,
                          nsteps, u)
    return u


def crit_timestep(conn, etype, xyz_cur, E, nu, dt_in=np.finfo(np.float64).max):
    """min over elements of 2 / sqrt(max row sum of |Ke(xyz_cur)|), from dt_in"""
    return lib().or_crit_timestep(len(etype), _i32(conn), _i32(etype), _f64(xyz_cur).ravel(), E, nu, dt_in)


def explicit_timed(conn, etype, xyz, E, nu, mass, f_ext, fixed_mask, time_total, time_step, output_gap=10,
                   v0=None, max_steps=100000, max_frames=64, dt_history=False):
    """Time loop of src/cae.cpp:420-574 with its time control (time_step 0 =
    automatic).  Returns dict(frames (k x 3nd), idx, t, steps, complete
    [, dt])."""
    nd = len(xyz)
    n3 = 3 * nd
    frames = np.zeros((max_frames, n3))
    idx = np.zeros(max_frames, dtype=np.int32)
    tt = np.zeros(max_frames)
    nf, ns = C.c_int(0), C.c_int(0)
    hist = np.zeros(max_steps) if dt_history else None
    v0a = None if v0 is None else _f64(v0).ravel()
    rc = lib().or_explicit_timed(nd, len(etype), _i32(conn), _i32(etype), _f64(xyz).ravel(), E, nu, _f64(mass),
                                 _f64(f_ext).ravel(), None if v0a is None else v0a.ctypes.data, _i32(fixed_mask),
                                 time_total, time_step, output_gap, max_steps, max_frames, frames, idx, tt,
                                 C.byref(nf), C.byref(ns), None if hist is None else hist.ctypes.data)
    k = min(nf.value, max_frames)
    out = dict(frames=frames[:k], idx=idx[:k], t=tt[:k], n_frames=nf.value, steps=ns.value, complete=rc == 0)
    if dt_history:
        out["dt"] = hist[:ns.value]
    return out


def matfree_apply(conn, etype, xyz, resort, E, nu, x, n, E_elem=None):
    y = np.zeros(n)
    Ee = None if E_elem is None else _f64(E_elem)
    lib().or_matfree_apply(len(etype), _i32(conn), _i32(etype), _f64(xyz).ravel(), _i32(resort), E, nu,
                           None if Ee is None else Ee.ctypes.data, _f64(x), y, n)
    return y


def nl_newton(conn, etype, xyz, resort, E, nu, F, col_ptr, row_idx, n_incr=1, tol=1e-10, max_newton=50):
    """Newton loop of src/cae.cpp:209-253 (r = F - f_int, K_T du = r, u += du)
    with the load ramped in n_incr increments and the stopping rule
    ||lam F - f_int|| <= tol ||lam F||; direct sparse solves (scipy SuperLU)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    n = len(col_ptr) - 1
    u = np.zeros(n)
    iters = 0
    for k in range(1, n_incr + 1):
        lam = k / n_incr
        for _ in range(max_newton + 1):
            val, fint = nl_fill(conn, etype, xyz, resort, E, nu, u, col_ptr, row_idx)
            r = lam * np.asarray(F) - fint
            if np.linalg.norm(r) <= tol * np.linalg.norm(lam * np.asarray(F)):
                break
            K = sp.csc_matrix((val, row_idx, col_ptr), shape=(n, n))
            u = u + spla.spsolve(K, r)
            iters += 1
        else:
            raise RuntimeError("oracle Newton did not converge")
    return u, iters


def nl_newton_ref(conn, etype, xyz, resort, E, nu, F, col_ptr, row_idx, res_lmit=1.0e-6, max_loop=100):
    """The Newton loop of src/cae.cpp:209-253 as it stands: res_norm = 1;
    while res_norm > res_lmit and loop < max_loop: K_T, f_int at u,
    r = F - f_int, K_T du = r (direct: scipy SuperLU in place of PARDISO),
    u += du, res_norm = ||r||_2 / (3 nfree) of that r, loop += 1.
    Returns (u, loops, res_norm, history of res_norm)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    n = len(col_ptr) - 1
    u = np.zeros(n)
    res_norm, loop, hist = 1.0, 0, []
    F = np.asarray(F, dtype=np.float64)
    while res_norm > res_lmit and loop < max_loop:
        val, fint = nl_fill(conn, etype, xyz, resort, E, nu, u, col_ptr, row_idx)
        r = F - fint
        K = sp.csc_matrix((val, row_idx, col_ptr), shape=(n, n))
        u = u + spla.spsolve(K, r)
        res_norm = float(np.sqrt(np.sum(r * r))) / n
        hist.append(res_norm)
        loop += 1
    return u, loop, res_norm, hist


def ca_fill(kind, conn_mod, etype_mod, xyz_m, xyz_o, resort, E, nu, col_ptr, row_idx):
    """Delta K of the modified elements on the pattern of K0 (src/ca_analysis.cpp:55-143)."""
    val = np.zeros(len(row_idx))
    kind = _i32(kind)
    rc = lib().or_ca_fill(len(kind), kind, _i32(conn_mod).reshape(-1, 8), _i32(etype_mod), _f64(xyz_m).ravel(),
                          _f64(xyz_o).ravel(), _i32(resort), E, nu, _i32(col_ptr), _i32(row_idx), len(row_idx), val)
    if rc != 0:
        raise ValueError("modified element outside the pattern or bad kind")
    return val


def ca_solve(col_ptr, row_idx, K0, dK, u0, F, n_basis, rel_tol=1e-10):
    """Combined approximations (ca_build_rom / ca_SVD / ca_solve,
    src/ca_analysis.cpp:146-309): r_0 = u0, r_i = -K0^-1 (dK r_{i-1}) by a
    direct factorisation of K0 (scipy SuperLU); thin SVD of the basis in
    float64, keeping the left singular vectors with sigma > rel_tol *
    sigma_max (the reference keeps all n, in float32); reduced solve
    (U^T (K0 + dK) U) y = U^T F; returns U y and the number of vectors kept."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    n = len(col_ptr) - 1
    A0 = sp.csc_matrix((K0, row_idx, col_ptr), shape=(n, n))
    dA = sp.csc_matrix((dK, row_idx, col_ptr), shape=(n, n))
    lu = spla.splu(A0)
    B = [np.asarray(u0, dtype=np.float64)]
    for _ in range(1, n_basis):
        B.append(-lu.solve(dA @ B[-1]))
    U, sv, _ = np.linalg.svd(np.array(B).T, full_matrices=False)
    U = U[:, sv > rel_tol * sv.max()]
    Kr = U.T @ ((A0 + dA) @ U)
    y = np.linalg.solve(Kr, U.T @ F)
    return U @ y, U.shape[1]


def post_elem(conn, etype, xyz, u_full, E, nu, quantity=0, E_elem=None):
    """7 x ne element values (strain or stress, then the Mises form)."""
    ne = len(etype)
    ev = np.zeros((7, ne))
    Ee = None if E_elem is None else _f64(E_elem)
    lib().or_post_elem(ne, _i32(conn).reshape(-1, 8), _i32(etype), _f64(xyz).ravel(), _f64(u_full).ravel(), E, nu,
                       None if Ee is None else Ee.ctypes.data, quantity, ev)
    return ev


def post_node(conn, etype, nd, ev):
    """7 x nd nodal averages of the element values (src/post_process.cpp:122-166)."""
    nv = np.zeros((7, nd))
    lib().or_post_node(nd, len(etype), _i32(conn).reshape(-1, 8), _i32(etype), _f64(ev), nv)
    return nv


def sfem_build_csc(conn, xyz, resort):
    """Smoothed-FEM CSC pattern (SFEM_build_CSR, src/assemble.cpp:291-350)."""
    conn, xyz, resort = _i32(conn).reshape(-1, 8), _f64(xyz).ravel(), _i32(resort)
    nd, ne = len(resort), len(conn)
    n = 3 * int(resort.max() + 1)
    nnz = lib().or_sfem_build_csc(nd, ne, conn, xyz, resort, n, None, None)
    cp = np.zeros(n + 1, dtype=np.int32)
    ri = np.zeros(nnz, dtype=np.int32)
    lib().or_sfem_build_csc(nd, ne, conn, xyz, resort, n, cp.ctypes.data, ri.ctypes.data)
    return cp, ri


def sfem_fill(conn, xyz, resort, E, nu, col_ptr, row_idx):
    """Smoothed-FEM values (SFEM_fill_CSR_sparse_mat, six-term domain stiffness)."""
    val = np.zeros(len(row_idx))
    conn = _i32(conn).reshape(-1, 8)
    lib().or_sfem_fill(len(resort), len(conn), conn, _f64(xyz).ravel(), _i32(resort), E, nu, _i32(col_ptr),
                       _i32(row_idx), len(row_idx), val)
    return val


def sfem_grads(conn, xyz, nd):
    """(mptr, members, rows): per (domain, member) G (3) and GG (3 x 3, scaled)."""
    conn = _i32(conn).reshape(-1, 8)
    xyzf = _f64(xyz).ravel()
    tot = lib().or_sfem_grads(nd, len(conn), conn, xyzf, None, None, None)
    mptr = np.zeros(nd + 1, dtype=np.int32)
    mem = np.zeros(tot, dtype=np.int32)
    rows = np.zeros((tot, 12))
    lib().or_sfem_grads(nd, len(conn), conn, xyzf, mptr.ctypes.data, mem.ctypes.data, rows.ctypes.data)
    return mptr, mem, rows
