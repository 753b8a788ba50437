"""Configuration and result layout of the batch runner (edlgt_amd/simio.py;
simsio as the reference's .simsiorc, config_example.yaml, sims.sh and scripts
use it): `===` header with a template and blocks, `<<<` merges, generated
entries, results/<uid>/{res.npz, par.yaml, info}, queries and grids.  Host
only."""
import logging
import re

import numpy as np
import pytest
import yaml

from edlgt_amd import simio

CONFIG = """===:
  template: |
    n$enum:
      <<<: common
      g: $g
      m: $m
  template_uid: |
    $uid:
      <<<: common
      g: $g
  common:
    model:
      lvals: [10]
      has_obc: [false]
      spin: 0.5
      pure_theory: false
    hamiltonian:
      n_eigs: 1
      format: sparse
    dynamics:
      time_evolution: true
      state: PV

n0:
  <<<: common
  g: 5
  m: 1
  hamiltonian:
    n_eigs: 3
"""


@pytest.fixture()
def dirs(tmp_path):
    cfg = tmp_path / "configs" / "SU2"
    cfg.mkdir(parents=True)
    (cfg / "example.yaml").write_text(CONFIG)
    return tmp_path / "configs", tmp_path / "results"


def test_resolve_merges_header_blocks(dirs):
    configs, _ = dirs
    path, par = simio.find_config("n0", configs)
    assert path.name == "example.yaml"
    assert par["g"] == 5 and par["m"] == 1
    assert par["model"]["lvals"] == [10] and par["model"]["spin"] == 0.5
    # the entry's own nested keys win, the block's other keys stay
    assert par["hamiltonian"] == {"n_eigs": 3, "format": "sparse"}
    assert "<<<" not in par
    with pytest.raises(KeyError):
        simio.find_config("n9", configs)


def test_gen_configs_template_and_uids(dirs):
    configs, _ = dirs
    keys = simio.gen_configs("template", {"g": np.array([0.5, 2.0]), "m": [0.1, 1.0]}, "SU2/example", configs)
    assert keys == ["n1", "n2", "n3", "n4"]
    pars = [simio.find_config(k, configs)[1] for k in keys]
    assert [(p["g"], p["m"]) for p in pars] == [(0.5, 0.1), (0.5, 1.0), (2.0, 0.1), (2.0, 1.0)]
    assert all(p["model"]["has_obc"] == [False] and p["hamiltonian"]["n_eigs"] == 1 for p in pars)
    ukeys = simio.gen_configs("template_uid", {"g": [3.0]}, "SU2/example", configs)
    assert len(ukeys) == 1 and re.fullmatch(r"[a-z0-9]{32}", ukeys[0])
    assert simio.find_config(ukeys[0], configs)[1]["g"] == 3.0
    with pytest.raises(KeyError):
        simio.gen_configs("nope", {"g": [1]}, "SU2/example", configs)


def test_run_sim_writes_the_result_layout(dirs):
    configs, results = dirs
    with simio.run_sim("n0", configs, results) as sim:
        assert sim.par["g"] == 5
        logging.getLogger("edlgt_amd.test").warning("solving n0")
        sim.res["energy"] = np.array([-1.5, -1.2, -0.9])
        sim.res["E2"] = 0.25
        sim.res["time_steps"] = np.linspace(0, 1, 3)
    d = results / "n0"
    assert sorted(p.name for p in d.iterdir()) == ["info", "par.yaml", "res.npz"]
    assert "solving n0" in (d / "info").read_text()
    assert yaml.safe_load((d / "par.yaml").read_text())["model"]["lvals"] == [10]
    back = simio.get_sim("n0", configs, results)
    assert np.array_equal(back.res["energy"], [-1.5, -1.2, -0.9]) and back.res["E2"] == 0.25
    assert back.par["hamiltonian"]["n_eigs"] == 3
    with pytest.raises(TypeError):
        with simio.run_sim("n0", configs, results) as sim:
            sim.res["bad"] = [np.zeros(2), np.zeros(3)]


def test_query_and_grid(dirs):
    configs, _ = dirs
    simio.gen_configs("template", {"g": [0.5, 2.0], "m": [0.1, 1.0]}, "SU2/example", configs)
    q = simio.SimsQuery("SU2/example", configs)
    assert q.uids == ["n0", "n1", "n2", "n3", "n4"]
    ugrid, vals = simio.uids_grid(q.uids[1:], ["g", "m"], configs)
    assert list(vals["g"]) == [0.5, 2.0] and list(vals["m"]) == [0.1, 1.0]
    assert ugrid.shape == (2, 2) and ugrid[1, 0] == "n3"
    ugrid, vals = simio.uids_grid(q.uids, ["g"], configs)
    assert list(vals["g"]) == [0.5, 2.0, 5] and ugrid[2] == "n0"
