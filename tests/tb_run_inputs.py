"""Input decks of a tight-binding transport run for the driver tests: a
Si/Ge superlattice along z (tetragonal four-atom diamond cells, one cell per
lead period), the example's parameters, confige with every namelist the
driver reads."""
import os
import shutil

import numpy as np

from tests import tb_structures as S

HERE = os.path.dirname(os.path.abspath(__file__))
PARAM = os.path.join(HERE, "golden", "tb", "param")

CONFIGE = """&configlist
 k_trans_mesh = {nk1} {nk2} 1, emin = {emin}, emax = {emax}, ne = {ne}
 r_cutoff = 2.6d0, filename_input = 'input.fdf'
 left_primitive_cell_input = '{pcl}', right_primitive_cell_input = '{pcr}'
/
&setup
 period_left = {period}, period_right = {period}, buffer_left = {buffer}, buffer_right = {buffer}
/
&convergelist
 eta = {eta}, convergence = 1.0d-10
/
&unfolding
 n_bloch_x = {nbx}, n_bloch_y = {nby}, circular_flag = {circular}
/
&kpath
 kpath_file = 'kpath', path_mode = {path_mode}, side = 'left', az1 = 5.5
/
&saving
 output_dir = 'out', verbosity = {verbosity}
/
&misori
 uc_misori_x(:) = 1.0 0.0 0.0, uc_misori_y(:) = 0.0 1.0 0.0, uc_misori_z(:) = 0.0 0.0 1.0
 n_uc_misori1 = 1, n_uc_misori2 = 1, n_uc_misori3 = 1, theta = 0.0
/
"""


def write_run(folder, cells=(1, 1, 8), split=True, nk=(2, 2), emin=-1.5, emax=2.2, ne=5, eta="1.0d-5", path_mode=2,
              kpath=None, buffer=0, circular=0, n_bloch=None, defect=None, pc=False, verbosity=0):
    """circular = 1: circular_flag = 1 with n_bloch_x, n_bloch_y = the
    transverse cell counts (the structure's atoms come as transverse copies,
    y outer, x inner); n_bloch: (n_bloch_x, n_bloch_y) otherwise (default
    1, 1), the transverse supercell of the unfolding; defect: index of an
    atom whose species is swapped (Si <-> Ge), which breaks the transverse
    periodicity of the cell; pc: write the leads' primitive-cell inputs
    (left: the first atom's species, right: the last one's)"""
    nb = (cells[0], cells[1]) if circular else (n_bloch or (1, 1))
    lat, pos, sp = S.si_ge_superlattice(5.5, *cells)
    period = 4 * cells[0] * cells[1]
    if not split:
        sp = np.full_like(sp, 2)
    if defect is not None:
        sp = sp.copy()
        sp[defect] = 3 - sp[defect]
    S.write_fdf(os.path.join(folder, "input.fdf"), lat, pos, sp)
    shutil.copy(PARAM, os.path.join(folder, "param"))
    with open(os.path.join(folder, "confige"), "w") as f:
        f.write(CONFIGE.format(nk1=nk[0], nk2=nk[1], emin=emin, emax=emax, ne=ne, eta=eta, path_mode=path_mode,
                               period=period, buffer=buffer, nbx=nb[0], nby=nb[1], circular=circular,
                               pcl="pc_l.fdf" if pc else "none", pcr="pc_r.fdf" if pc else "none",
                               verbosity=verbosity))
    if pc:
        write_pc(folder, "pc_l.fdf", int(sp[0]))
        write_pc(folder, "pc_r.fdf", int(sp[-1]))
    if kpath is not None:
        with open(os.path.join(folder, "kpath"), "w") as f:
            f.write(f"{len(kpath)}\n" + "".join(" ".join(str(x) for x in k) + "\n" for k in kpath))
    return os.path.join(folder, "confige"), lat, pos, sp


def write_pc(folder, name, species, a=5.5):
    """A primitive-cell input (tb/input.f90:144-234) of the superlattice's
    diamond: the fcc vectors, the two-atom basis and one atom of the next
    cell (which the reader leaves out), all of the given species."""
    path = os.path.join(folder, name)
    with open(path, "w") as f:
        f.write(f"NumberOfAtoms 3\nNumberOfSpecies 1\nLatticeConstant {a} Ang\n%block LatticeVectors\n"
                "0.0 0.5 0.5\n0.5 0.0 0.5\n0.5 0.5 0.0\n%endblock LatticeVectors\n"
                "%block AtomicScaledCoordinatesAndAtomicSpecies\n"
                f"0.0 0.0 0.0 {species}\n0.25 0.25 0.25 {species}\n1.0 0.0 0.0 {species}\n"
                "%endblock AtomicScaledCoordinatesAndAtomicSpecies\n")
    return path
