"""Reference-compatible driver and grid average without a GPU: argument
forms, table, file names, float32 incidence values (treePO.cu:27-52, 501-564,
955), and the block-average oracle (util/fields.h:206-234) pinned against a
float64 block mean, non-power-of-two ratios and ignored trailing rows."""
import math

import numpy as np

from oracle import oracle
from wavefiber3d_amd import cli


def test_usage_and_bad_fiber_index():
    assert cli.main([]) == 2
    assert cli.main(["treepo", "0", "1", "90", "0", "7", "3", "0", "10"]) == 2


def test_output_names_and_incidence_values():
    assert cli.output_name("output", 2, 13, "M", 6).endswith("output/fiber_2_lambda13_TM_depth6.binary")
    assert cli.output_name("o", 0, 0, "M", 1, final=True) == "o/fiber_0_lambda0_TM_depth1_final.binary"
    for l in (0, 7, 24):
        assert cli.wavelength(l) == float(np.float32(float(np.float32(400 + 12 * l)) * 1e-9))
        assert abs(cli.wavelength(l) / ((400 + 12 * l) * 1e-9) - 1) < 1e-7
    t, p = cli.incidence_angles(90, 45)
    assert t == float(np.float32(-math.pi / 2 + 1e-10 + math.pi / 2))
    assert p == float(np.float32(math.pi / 4))
    assert [d["thetao"] for d in cli.FIBER_TABLE] == [450, 900, 7200, 7200]
    assert [d["numel"] for d in cli.FIBER_TABLE] == [3000, 5026, 4420, 4720]


def test_average_grid_oracle_pins():
    rng = np.random.default_rng(1)
    g = rng.random((7200, 1440)).astype(np.float32)
    a = oracle.average_grid(g, 180, 360)
    ref = g.astype(np.float64).reshape(180, 40, 360, 4).mean((1, 3))
    assert np.max(np.abs(a - ref) / ref) < 1e-5
    c = np.full((90, 30), 0.37, dtype=np.float32)
    assert np.allclose(oracle.average_grid(c, 18, 10), 0.37, rtol=1e-6)
    # trailing rows / columns beyond ratio * n_out are ignored (integer ratios)
    g2 = rng.random((725, 1443)).astype(np.float32)
    a2 = oracle.average_grid(g2, 180, 360)
    assert np.array_equal(a2, oracle.average_grid(np.ascontiguousarray(g2[:720, :1440]), 180, 360))


def test_mie_command_writes_notebook_files(tmp_path):
    import os

    from wavefiber3d_amd import mie

    assert cli.main(["mie", "--radius", "1000", "--output-dir", str(tmp_path)]) == 0
    _, SL, SR, SU = mie.scattering_function(1.55 + 0.1j, 400, 2000, min_angle=-180, max_angle=180,
                                            angular_resolution=0.25, normalization="t")
    for name, v in (("TM", SL), ("TE", SR), ("TEM", SU)):
        got = np.fromfile(os.path.join(tmp_path, f"mie_1um_{name}.bin"), np.float64)
        assert np.array_equal(got, v)
    assert got.shape == (1441,)
