"""Write tests/golden/x12pts_cube_v2_sample.npz.

Input: 64 pair site clouds sampled (seed 0) from the reference's data file
01/datadir/x_12pts_500_s42.npy (cube_v2 geometry, 6 sites per body, loaded
with numpy's safe loader).  Expected output: the FP64 oracle descriptor with
the cube_v2 model hypers (10, 8, 4) and r_nep = 4.5
(nep_simulator/DescriptorGenerator.py:63-79, runner.sh:2).
Also copies, unchanged, the reference's initial configuration
nep_simulator/init/th_init_N200.gsd (200 tetrahedra, box 33, written by
HOOMD-blue 4.7.0) as tests/golden/th_init_N200.gsd, and the label potential's
bead table 01/cube_structure.gsd (152 beads of one cube, read by
01/SamplerUtils.py) as tests/golden/cube_structure.gsd.
Run in a container that holds the reference checkout:
    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import nep_oracle as O  # noqa: E402

SRC = "/root/reference/01/datadir/x_12pts_500_s42.npy"
INIT = "/root/reference/nep_simulator/init/th_init_N200.gsd"
BEADS = "/root/reference/01/cube_structure.gsd"

if __name__ == "__main__":
    x = np.load(SRC, allow_pickle=False)
    idx = np.sort(np.random.default_rng(0).choice(len(x), 64, replace=False))
    clouds = x[idx].astype(np.float32)
    G = O.descriptor(clouds.astype(np.float64), 4.5, 10, 8, 4)
    np.savez(os.path.join(HERE, "x12pts_cube_v2_sample.npz"), clouds=clouds, G=G, index=idx)
    print("wrote", clouds.shape, G.shape)
    import shutil
    shutil.copyfile(INIT, os.path.join(HERE, "th_init_N200.gsd"))
    shutil.copyfile(BEADS, os.path.join(HERE, "cube_structure.gsd"))
