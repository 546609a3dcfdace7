"""The reference's example configs and the on-disk formats on the device.

Parity bar as in test_gpu_parity.py: every grid point within 1e-10 of the
oracle relative to the field maximum, FP64."""
import json
import os

import numpy as np
import pytest

from test_capi import example_config
from test_gpu_parity import TOL, compare_state, pair, rel

pytestmark = pytest.mark.gpu

SMALL = ["dwave_chiral", "dwave_octagon", "dwave_phase_crystal", "dwave_plus_swave",
         "swave_disc_meissner"]


@pytest.mark.parametrize("name", SMALL)
def test_reference_example_iterations(product, oracle, name):
    cfg = example_config(name)
    s, o = pair(product, oracle, cfg)
    for it in range(2):
        s.compute("fp64")
        o.compute()
        compare_state(s, o, cfg, what="%s it %d" % (name, it))


def test_simulate_writes_reference_files_and_warm_starts(product, oracle, tmp_path):
    from superconga_amd import io as scio
    cfg = example_config("swave_disc_meissner")
    cfg["numerics"]["num_iterations_max"] = 3
    cfg["misc"] = {"save_path": str(tmp_path / "run"), "save_frequency": -1,
                   "load_path": "", "data_format": "csv", "verbose": False}
    converged, results = scio.simulate(cfg)
    run = tmp_path / "run"
    for f in ("domain.csv", "order_parameter.csv", "current_density.csv",
              "vector_potential.csv", "magnetic_flux_density.csv",
              "simulation_results.json", "internal_parameters.json",
              "simulation_convergence.csv", "simulation_config.json"):
        assert (run / f).exists(), f
    conv = (run / "simulation_convergence.csv").read_text().splitlines()
    assert len(conv) == 1 + 3
    assert results["free_energy_per_unit_area"] < 0
    # A warm start reads the written gap-equation result back as the iterate.
    saved = scio.read_csv(str(run / "order_parameter.csv"))
    s = product.Simulation(cfg)
    scio.load_state(s, cfg, str(run))
    got = s.order_parameter(0)[0]
    want = (saved["s_re"] + 1j * saved["s_im"]).reshape(s.N0, s.N0)
    dom = s.domain().astype(bool)
    assert rel(np.where(dom, got, 0), np.where(dom, want, 0)) < 1e-8
    s.compute("fp64")
    assert np.isfinite(s.results()["free_energy_per_area"])


def test_simulation_binary_matches_simulate(product, tmp_path):
    # the drop-in binary (`simulation_amd <config.json>`, the reference's
    # src/simulation.cu called by frontend/common/cpp.py) against the
    # in-process io.simulate on the same config: the same files, the CSV
    # grids byte for byte, the scalar results to the last bits of their
    # different summation orders
    import subprocess
    from superconga_amd import io as scio
    from test_adaptors import simulation_binary
    exe = simulation_binary(product)
    runs = {}
    for which in ("binary", "simulate"):
        cfg = example_config("swave_disc_meissner")
        cfg["numerics"]["num_iterations_burnin"] = 2
        cfg["numerics"]["num_iterations_max"] = 5
        out = tmp_path / which
        cfg["misc"] = {"save_path": str(out), "save_frequency": 2, "load_path": "",
                       "data_format": "csv", "verbose": False}
        if which == "binary":
            path = tmp_path / "config.json"
            path.write_text(json.dumps(cfg, indent="\t"))
            r = subprocess.run([exe, str(path)], capture_output=True, text=True, timeout=300)
            print(r.stdout, r.stderr, flush=True)
            assert r.returncode == 0, r.stderr
        else:
            scio.simulate(cfg)
        runs[which] = out
    names = sorted(os.listdir(runs["simulate"]))
    assert names == sorted(os.listdir(runs["binary"]))
    for name in names:
        a = (runs["binary"] / name).read_text()
        b = (runs["simulate"] / name).read_text()
        if name.endswith(".csv"):
            assert a == b, name
        else:
            ja, jb = json.loads(a), json.loads(b)
            assert sorted(ja) == sorted(jb), name
            for k in jb:
                if isinstance(jb[k], (int, float)):
                    assert abs(ja[k] - jb[k]) <= 1e-12 * max(abs(jb[k]), 1e-300), (name, k)
                elif k == "misc":
                    # the written config keeps each run's own save_path
                    ma, mb = dict(ja[k]), dict(jb[k])
                    ma.pop("save_path", None)
                    mb.pop("save_path", None)
                    assert ma == mb, (name, k)
                else:
                    assert ja[k] == jb[k], (name, k)
    conv = (runs["binary"] / "simulation_convergence.csv").read_text().splitlines()
    # header, then one row per iteration: with save_frequency > 0 the
    # reference counts the burn-in iterations towards num_iterations_max
    # (Context::burnIn increments the iteration count, Context.h:388, and
    # runSimulation compares getNumIterations with the maximum,
    # run_simulation.h:144-156), so 2 burn-in rows and 3 more iterations
    assert len(conv) == 1 + 5
    assert [int(r.split(",")[0]) for r in conv[1:]] == [1, 2, 3, 4, 5]


def test_simulation_binary_refuses_hdf5_output(product, tmp_path):
    import subprocess
    from test_adaptors import simulation_binary
    cfg = example_config("swave_disc_meissner")
    cfg["misc"]["save_path"] = str(tmp_path / "out")
    path = tmp_path / "config.json"
    path.write_text(json.dumps(cfg))
    r = subprocess.run([simulation_binary(product), str(path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "HDF5" in r.stderr
    assert not (tmp_path / "out").exists()
