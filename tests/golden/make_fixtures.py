"""Write the golden fixtures under tests/golden/.

ref_kats.json  -- the known-answer vectors held by the reference's own MPM
                  tests (femflow/solvers/mpm/tests/*.py and
                  ground_truth_grid.txt), transcribed as data: inputs and the
                  expected outputs those tests assert.
oracle_*.npz   -- oracle outputs on small seeded inputs (regression anchors
                  for the GPU tests; produced by oracle/mpm_oracle.py).
Run from the repository root: python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import mpm_oracle as O  # noqa: E402

KATS = {
    "source": "femflow/solvers/mpm/tests (test_utils.py, test_particle_to_grid.py, "
              "test_grid_to_particle.py, ground_truth_grid.txt)",
    # test_utils.py:6-15 -- weights[i][d] at fx = (0.631893, 0.965839)
    "weights": {"fx": [0.631893, 0.965839],
                "expected": [[0.37680488, 0.14266399], [0.61449724, 0.74883303],
                             [0.00869788, 0.10850299]]},
    # test_particle_to_grid.py:12-28 / test_grid_to_particle.py:12-28 parameters
    "params": {"mass": 1.0, "volume": 1.0, "hardening": 10.0, "E": 1e4, "nu": 0.2,
               "gravity": -200.0, "dt": 1e-4, "dx": 1.0 / 80, "grid_resolution": 80},
    # test_particle_to_grid.py:30-38 -- Lame at E=1e4, nu=0.2 (hardening factor 1 at Jp=1)
    "lame": {"lambda": 2777.777777777778, "mu": 4166.666666666667},
    # test_particle_to_grid.py:41-50 -- 2D fused stress at this F and C
    "stress_2d": {"F": [[1, 0], [2.51457e-12, 1]], "C": [[0, 0], [2.51457e-08, -5.96046e-08]],
                  "expected": [[0, -2.68221e-08], [-1.67638e-09, -5.96046e-08]]},
    # test_particle_to_grid.py:53-80 and ground_truth_grid.txt:1-27
    "p2g_2d": {"positions": [[0.55, 0.45], [0.45, 0.65], [0.55, 0.85]],
               "mass_multiset": [0.015625, 0.09375, 0.015625, 0.09375, 0.5625, 0.09375, 0.015625,
                                 0.09375, 0.015625] * 3},
    # test_grid_to_particle.py:31-103 -- one step from rest
    "g2p_2d": {"positions": [[0.55, 0.45], [0.45, 0.65], [0.55, 0.85]],
               "expected_positions": [[0.55, 0.449998], [0.45, 0.649998], [0.55, 0.849998]],
               "expected_velocity": [0.0, -0.02],
               "expected_C_atol": 1e-6,
               "expected_C": [[0.0, 0.0], [-2.98023e-08, 0.0]]},
}


def main():
    with open(os.path.join(HERE, "ref_kats.json"), "w") as f:
        json.dump(KATS, f, indent=1)
    # 2D config-1 scene (8,192 particles, res 128): states after 1 and 10
    # substeps, and after all 200 substeps of the configuration
    st, pr = O.scene_2d_c1()
    s = O.copy_state(st)
    O.substep(s, pr)
    one = O.copy_state(s)
    for _ in range(9):
        O.substep(s, pr)
    ten = O.copy_state(s)
    for _ in range(190):
        O.substep(s, pr)
    np.savez_compressed(os.path.join(HERE, "oracle_c1_2d.npz"),
                        **{f"s1_{k}": one[k] for k in ("x", "v", "F", "C", "Jp")},
                        **{f"s10_{k}": ten[k] for k in ("x", "v", "F", "C", "Jp")},
                        **{f"s200_{k}": s[k] for k in ("x", "v", "F", "C", "Jp")})
    # 3D three-material scene, 6,000 particles at res 32: after 1 and 5 substeps
    st, pr = O.scene_3d_cubes(n=6000, res=32)
    s = O.copy_state(st)
    O.substep(s, pr)
    one = O.copy_state(s)
    for _ in range(4):
        O.substep(s, pr)
    np.savez_compressed(os.path.join(HERE, "oracle_3d_cubes.npz"),
                        **{f"s1_{k}": one[k] for k in ("x", "v", "F", "C", "Jp")},
                        **{f"s5_{k}": s[k] for k in ("x", "v", "F", "C", "Jp")})


if __name__ == "__main__":
    main()
