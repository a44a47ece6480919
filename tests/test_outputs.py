"""Readers of the reference's output files (citcom_amd/outputs.py) on files
in the reference's formats (Output.c:66-128)."""
import os

import numpy as np

from citcom_amd import outputs


def test_readers(tmp_path):
    p = tmp_path / "run.ave.0.3"
    p.write_text("     3      3 1.23450e-03 4.37217e+00 4.37100e+00 0.0000e+00 0.0000e+00 0.00000e+00 0.00000e+00\n"
                 + "".join(f"{t:.4e} {v:.4e} 1.0000e+00 0.0000e+00 0.0000e+00 0.0000e+00 0.0000e+00 0.0000e+00\n"
                           for t, v in ((1.0, 0.0), (0.5, 12.5), (0.0, 0.0))))
    a = outputs.read_ave(str(p))
    assert a.noz == 3 and a.timesteps == 3 and a.nu_top == 4.37217 and a.profiles.shape == (3, 8)
    assert a.profiles[1, 1] == 12.5
    v = tmp_path / "run.velo.0.3"
    v.write_text("     2      3 1.23450e-03\n1.000000e+00 2.000000e+00 3.000000e+00\n-1.0e+00 0.0e+00 5.5e-01\n")
    f = outputs.read_velo(str(v))
    assert f.nno == 2 and np.array_equal(f.values[1], [-1.0, 0.0, 0.55])
    c = tmp_path / "run.coord.0"
    c.write_text("     2\n0.00000e+00 0.00000e+00 0.00000e+00\n1.00000e+00 0.00000e+00 5.00000e-01\n")
    assert outputs.read_coord(str(c)).shape == (2, 3)
