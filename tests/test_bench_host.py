"""Host-side logic of bench.py: the MPI decomposition it chooses for the
reference's CPU baseline satisfies the reference's constraint
(mgunit{x,y,z} % nproc{x,y,z} == 0, Instructions.c) and uses as many ranks
as allowed."""
import bench


def test_mpi_decomposition():
    for name, cap, want in (("tdep128", 16, 16), ("tdep128", 8, 8), ("tdep128", 12, 8), ("busse1a", 16, 4),
                            ("iso16", 64, 64), ("tiny", 3, 2)):
        p = bench.make_problem(name)
        px, py, pz = bench.mpi_decomposition(p, cap)
        mx, my, mz = p.mgunit
        assert mx % px == 0 and my % py == 0 and mz % pz == 0, name
        assert px * py * pz == want, (name, cap, (px, py, pz))


def test_measured_traffic_lookup():
    # the committed counter summary is keyed by workload and kernel
    assert bench.measured_traffic("k_gs_colour<float, float>", "tdep128") > 1e7
    assert bench.measured_traffic("k_gs_colour<float, float>", "tdep128", "finest_level_hbm_bytes_per_launch") > 1e8
    assert bench.measured_traffic("k_gs_colour<float, float>", "iso16") is None
