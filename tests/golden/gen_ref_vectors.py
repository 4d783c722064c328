"""Writes the golden fixtures under tests/golden/ (run once, in the container
that holds the reference checkout; the fixtures are committed, the reference is not).

  ref_elements.npz   element-level vectors produced by the reference's own
                     element code (src/Elements/C3D8*.cpp, C3D4*.cpp compiled
                     into oracle/_ref/libofea_ref.so by oracle/Makefile):
                     Ke, TL tangent + internal force, lumped mass, cal_in_force
                     for canonical, distorted and inverted elements.
  mesh_o.inp         the reference's sample deck sample_model/CA/mesh_o.inp (input data)
  mesh_o_cae.inp     the same deck with tet nodes 2<->3 swapped (input data)
  cadcae_o_u.npz     u_x, u_y, u_z of sample_model/CA/cadcae_o.vtk and
                     cadcae_o_cae.vtk: the reference's PARDISO solutions of the two
                     decks (E = 2.1e5, nu = 0.3 from src/main.cpp:121), as printed
                     (6 significant digits).

Usage: python tests/golden/gen_ref_vectors.py [/path/to/reference]
"""
import ctypes as C
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"

pd = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def load_ref():
    path = os.path.join(REPO, "oracle", "_ref", "libofea_ref.so")
    if not os.path.exists(path):
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    R = C.CDLL(path)
    for f in ("ref_c3d8_ke", "ref_c3d4_ke"):
        getattr(R, f).argtypes = [pd, C.c_double, C.c_double, pd]
    for f in ("ref_c3d8_nl", "ref_c3d4_nl"):
        getattr(R, f).argtypes = [pd, pd, C.c_double, C.c_double, pd, pd]
    for f in ("ref_c3d8_mass", "ref_c3d4_mass"):
        getattr(R, f).argtypes = [pd, C.c_double, pd]
    for f in ("ref_c3d8_in_force", "ref_c3d4_in_force"):
        getattr(R, f).argtypes = [pd, pd, C.c_double, C.c_double, pd]
    return R


HEX0 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
TET0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)


def element_sets(rng):
    hexes = [HEX0.copy(), HEX0 * np.array([2.0, 0.5, 1.5]) + 3.0]
    for _ in range(6):
        hexes.append(HEX0 + rng.uniform(-0.2, 0.2, HEX0.shape))
    inv = HEX0[[4, 5, 6, 7, 0, 1, 2, 3]]  # mirrored in z: negative detJ everywhere
    hexes.append(inv)
    tets = [TET0.copy(), TET0 * 2.5 - 1.0]
    for _ in range(6):
        tets.append(TET0 + rng.uniform(-0.2, 0.2, TET0.shape))
    tets.append(TET0[[0, 2, 1, 3]])  # inverted (nodes 2<->3, as mesh_o_cae.inp)
    return np.array(hexes), np.array(tets)


def main():
    R = load_ref()
    rng = np.random.default_rng(20261203)
    hexes, tets = element_sets(rng)
    E, nu, rho = 210000.0, 0.3, 7.85e-9
    out = {"E": E, "nu": nu, "rho": rho, "hex_X": hexes, "tet_X": tets}
    hk, hnlK, hnlf, hm, hu, hf = [], [], [], [], [], []
    for X in hexes:
        x = np.ascontiguousarray(X.ravel())
        K = np.zeros(576)
        R.ref_c3d8_ke(x, E, nu, K)
        hk.append(K.reshape(24, 24))
        u = rng.normal(0.0, 0.05, 24)
        Kt, f = np.zeros(576), np.zeros(24)
        R.ref_c3d8_nl(x, u, E, nu, Kt, f)
        hnlK.append(Kt.reshape(24, 24))
        hnlf.append(f)
        hu.append(u)
        m = np.zeros(8)
        R.ref_c3d8_mass(x, rho, m)
        hm.append(m)
        fi = np.zeros(24)
        R.ref_c3d8_in_force(x, np.ascontiguousarray(u), E, nu, fi)
        hf.append(fi)
    out.update(hex_Ke=np.array(hk), hex_u=np.array(hu), hex_KT=np.array(hnlK), hex_fint=np.array(hnlf),
               hex_mass=np.array(hm), hex_in_force=np.array(hf))
    tk, tnlK, tnlf, tm, tu, tf = [], [], [], [], [], []
    for X in tets:
        x = np.ascontiguousarray(X.ravel())
        K = np.zeros(144)
        R.ref_c3d4_ke(x, E, nu, K)
        tk.append(K.reshape(12, 12))
        u = rng.normal(0.0, 0.05, 12)
        Kt, f = np.zeros(144), np.zeros(12)
        R.ref_c3d4_nl(x, u, E, nu, Kt, f)
        tnlK.append(Kt.reshape(12, 12))
        tnlf.append(f)
        tu.append(u)
        m = np.zeros(4)
        R.ref_c3d4_mass(x, rho, m)
        tm.append(m)
        fi = np.zeros(12)
        R.ref_c3d4_in_force(x, np.ascontiguousarray(u), E, nu, fi)
        tf.append(fi)
    out.update(tet_Ke=np.array(tk), tet_u=np.array(tu), tet_KT=np.array(tnlK), tet_fint=np.array(tnlf),
               tet_mass=np.array(tm), tet_in_force=np.array(tf))
    np.savez_compressed(os.path.join(HERE, "ref_elements.npz"), **out)

    ca = os.path.join(REF, "sample_model", "CA")
    for deck in ("mesh_o.inp", "mesh_o_cae.inp"):
        shutil.copyfile(os.path.join(ca, deck), os.path.join(HERE, deck))
    fields = {}
    for name, vtk in (("o", "cadcae_o.vtk"), ("o_cae", "cadcae_o_cae.vtk")):
        fields[name] = read_vtk_u(os.path.join(ca, vtk))
    np.savez_compressed(os.path.join(HERE, "cadcae_o_u.npz"), u_o=fields["o"], u_o_cae=fields["o_cae"])
    print("wrote", sorted(os.listdir(HERE)))


def read_vtk_u(path):
    """u_x, u_y, u_z point scalars of a legacy ASCII VTK file -> (npoints, 3)."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    npts = None
    comps = {}
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith("POINT_DATA"):
            npts = int(ln.split()[1])
        if ln.startswith("SCALARS u_") and npts:
            name = ln.split()[1]
            vals = []
            j = i + 2  # skip LOOKUP_TABLE
            while len(vals) < npts:
                vals.extend(float(t) for t in lines[j].split())
                j += 1
            comps[name] = np.array(vals[:npts])
            i = j
            continue
        i += 1
    return np.stack([comps["u_x"], comps["u_y"], comps["u_z"]], axis=1)


if __name__ == "__main__":
    main()
