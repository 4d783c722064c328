"""Writes the CA-reanalysis and nodal-field fixtures from the reference's
sample outputs (run in a checkout that has /root/reference; the outputs are
committed, the reference is not needed afterwards).

  cadcae_ca_u.npz    u_x, u_y, u_z (525 x 3) of sample_model/CA/cadcae_ca.vtk:
                     the reference's CA reanalysis of mesh_o.inp with the 384
                     tets of CA_info.txt deleted (driver src/main.cpp:136-147,
                     n_basis = 4), on the nodes of mesh_m.inp
  cadcae_o_fields.npz S_xx ... S_Mises (7 x 625), the VTK cells (2304 x 4,
                     as written) and cell types of sample_model/CA/cadcae_o.vtk
CA_info.txt and mesh_m.inp are copied unchanged from sample_model/CA/ (input
decks).  Values are as printed by the reference: 6 significant digits.
"""
import os
import shutil

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CA = "/root/reference/sample_model/CA"


def sections(path):
    lines = open(path).read().splitlines()
    out = {}
    for k, ln in enumerate(lines):
        if ln.startswith("SCALARS "):
            out[ln.split()[1]] = k + 2
    return lines, out


def field(lines, start, n):
    return np.array([float(x) for x in lines[start:start + n]])


def main():
    lines, sec = sections(os.path.join(CA, "cadcae_ca.vtk"))
    nd = int(lines[5].split()[1])
    u = np.stack([field(lines, sec[c], nd) for c in ("u_x", "u_y", "u_z")], 1)
    np.savez_compressed(os.path.join(HERE, "cadcae_ca_u.npz"), u_ca=u)

    lines, sec = sections(os.path.join(CA, "cadcae_o.vtk"))
    nd = int(lines[5].split()[1])
    names = ["S_xx", "S_yy", "S_zz", "S_xy", "S_yz", "S_xz", "S_Mises"]
    S = np.stack([field(lines, sec[c], nd) for c in names], 0)
    k = next(i for i, ln in enumerate(lines) if ln.startswith("CELLS"))
    ne = int(lines[k].split()[1])
    cells = np.array([[int(t) for t in lines[k + 1 + e].split()[1:]] for e in range(ne)], dtype=np.int32)
    kt = next(i for i, ln in enumerate(lines) if ln.startswith("CELL_TYPES"))
    types = np.array([int(lines[kt + 1 + e]) for e in range(ne)], dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "cadcae_o_fields.npz"), S=S, cells=cells, cell_types=types)
    for f in ("CA_info.txt", "mesh_m.inp"):
        shutil.copyfile(os.path.join(CA, f), os.path.join(HERE, f))


if __name__ == "__main__":
    main()
