"""Write nhc_legacy_si64.npz from the reference's recorded NVT/NPT runs.

Reads tests/data/nhc_legacy_0.3.0a7/{nvt,npt}/md_checkpoint.npz and
final_300.npz of the reference checkout (only where it is present; numpy
loads with allow_pickle=False) and keeps the recorded trajectory data:

* nvt: E_pot, E_kin and T after steps 1..150, the thermo rows (written
  after steps 1, 11, ..., 141 by that version) with H_NHC, and the
  checkpointed state after step 150 (positions, velocities, chain);
* npt: T, P [GPa] and V after steps 1..150, the checkpointed state after
  step 150 (positions, velocities, cell, chain and barostat) and H_NPT of
  the first row;
* the positions, velocities and lattice after step 300 of both runs.

The runs are tests/test_nhc_chain.py:24-48,136-170 of the reference: 64-atom
Si (2x2x2 conventional diamond at a = 5.430950 A, displaced by N(0, 0.05)
from default_rng(1), mean displacement removed), SW Si, T = 600 K, dt = 1 fs,
Maxwell-Boltzmann start from default_rng(7), a two-link chain with
tau = 100 fs and, for NPT, P = 0 with tau_baro = 500 fs.
"""
import json
import os
import sys

import numpy as np

REF = os.environ.get("AK_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = os.path.join(REF, "tests", "data", "nhc_legacy_0.3.0a7")
    if not os.path.isdir(src):
        print("reference run data not found; nothing written", file=sys.stderr)
        return 1
    nvt = np.load(os.path.join(src, "nvt", "md_checkpoint.npz"), allow_pickle=False)
    npt = np.load(os.path.join(src, "npt", "md_checkpoint.npz"), allow_pickle=False)
    fin = np.load(os.path.join(src, "final_300.npz"), allow_pickle=False)
    mv = json.loads(str(nvt["meta"]))
    mp = json.loads(str(npt["meta"]))
    out = {
        "nvt_e_pot": nvt["list_e_pot"], "nvt_e_kin": nvt["list_e_kin"], "nvt_T": nvt["list_T"],
        "nvt_row_step": nvt["tlog_step"], "nvt_row_H": nvt["tlog_H_nhc"],
        "nvt_chain150": np.array([mv["scalars"][k] for k in ("xi1", "xi2", "vxi1", "vxi2")]),
        "nvt_pos150": nvt["positions"], "nvt_vel150": nvt["velocities"],
        "nvt_lat150": nvt["lattice"],
        "npt_pos150": npt["positions"], "npt_vel150": npt["velocities"],
        "npt_lat150": npt["lattice"],
        "npt_T": npt["list_T"], "npt_P": npt["list_P"], "npt_V": npt["list_V"],
        "npt_state150": np.array([mp["scalars"][k] for k in
                                  ("xi1", "xi2", "vxi1", "vxi2", "p_eps", "xi1_b", "xi2_b",
                                   "vxi1_b", "vxi2_b")]),
        "npt_H_first_row": np.array(mp["extra"]["H_ref"]),
    }
    for k in fin.files:
        out[k + "300"] = fin[k]
    np.savez_compressed(os.path.join(HERE, "nhc_legacy_si64.npz"), **out)
    print(f"{len(out)} arrays")
    return 0


if __name__ == "__main__":
    sys.exit(main())
