"""Caroli run of the reference's QE force-constant (tdb) path on one GPU.

Replaces the k/omega loop of negf/phinfortran.f90:1188-1275 for the Caroli
outputs: the namelist config and the QE force constants are read as the
reference reads them (phinfortran_amd/qe.py), H(k) of every transverse
k-point is assembled on the GPU, every (omega, k) point is solved in one
batched call with the reference's ie = 1 zeroing, and transl_tot.dat /
transr_tot.dat are written in the reference's format (write_trans,
negf/write_data.f90:340-497).
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import output as OUT
from . import qe as Q


def run_qe_tdb(config_path: str, fc_path: str, output_dir: Optional[str] = None, engine=None,
               mass_profile: Optional[str] = None, precision: str = "fp64"):
    """Returns dict(E (ne), k (nk x 3), transl, transr (ne x nk), status)."""
    import torch
    from . import transport as T
    cfg = Q.read_config(config_path)
    qe = Q.read_qe_fc(cfg, fc_path, mass_profile)
    asm = Q.TdbAssembler(qe, cfg)
    ks = Q.k_grid(cfg)
    E = Q.e_grid(cfg)
    eng = engine or T.Engine(0)
    d = asm.upload(eng, ks)
    nk, ne = len(ks), len(E)
    kidx = np.repeat(np.arange(nk, dtype=np.int32), ne)
    mask = np.tile((np.arange(ne) == 0).astype(np.int32), nk)
    p = T.params("phonon", eta=cfg["eta"], eta0=cfg["eta0"], convergence=cfg["convergence"], max_iter=500,
                 precision=precision)
    r = eng.solve_tdb(d, np.tile(E, nk), p, kidx=kidx, first_point_mask=mask)
    tl = r["T_LR"].cpu().numpy().reshape(nk, ne).T
    tr = r["T_RL"].cpu().numpy().reshape(nk, ne).T
    st = r["status"].cpu().numpy().reshape(nk, ne).T
    if output_dir:
        OUT.write_trans(output_dir, np.repeat(E[:, None], nk, axis=1), tl, tr)
    if engine is None:
        eng.close()
    return dict(E=E, k=ks, transl=tl, transr=tr, status=st, blocks=d)
