"""Writes the golden fixtures of tests/golden from the reference's own data.

Run in a container that has the reference tree (default /root/reference):

    python tests/golden/make_fixtures.py [reference_root]

Fixtures written (data only, no reference code):
  ozaki_max_poles.json  -- the largest Ozaki pole for n = 1..200 poles, the
                           table of include/conga/Parameters.h:190-258 that the
                           reference uses to pick the number of poles.  It is
                           the output of the reference's dsygv-based pole
                           solver, so it pins any pole solver.
  known_answers.json    -- scalar known answers held by the reference's tests,
                           each with the file:line it is taken from.
  examples/<name>.json  -- the reference's example simulation configs
                           (examples/<name>/simulation_config.json), inputs
                           for the drop-in config tests.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def pole_table(root):
    path = os.path.join(root, "include", "conga", "Parameters.h")
    text = open(path).read()
    m = re.search(r"m_maxOzakiPole\s*=\s*\{([^}]*)\}", text)
    if not m:
        raise SystemExit("pole table not found in %s" % path)
    values = [float(v) for v in re.findall(r"[-+0-9.eE]+", m.group(1))]
    line = text[:m.start()].count("\n") + 1
    return values, "include/conga/Parameters.h:%d" % line


def accelerator_omega(root):
    path = os.path.join(root, "tests", "accelerators_test.cu")
    text = open(path).read()
    m = re.search(r"freeEnergyExpected\s*=\s*static_cast<T>\(([-0-9.eE]+)\)", text)
    if not m:
        raise SystemExit("expected free energy not found in %s" % path)
    line = text[:m.start()].count("\n") + 1
    return float(m.group(1)), "tests/accelerators_test.cu:%d" % line


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    poles, src = pole_table(root)
    with open(os.path.join(HERE, "ozaki_max_poles.json"), "w") as f:
        json.dump({"source": src, "max_pole": poles}, f, indent=0)
    omega, osrc = accelerator_omega(root)
    answers = {
        "accelerator_free_energy_per_area": {
            "value": omega, "source": osrc,
            "case": "disc, grain width 10, 3 points per xi0, 16 angles, T=0.5, "
                    "cutoff 16, flux 1, penetration depth 10, linf norm, "
                    "criterion 1e-5; all accelerators",
            "tolerance_rel": 1e-3},
        "homogeneous_gamma": {
            "delta": [4.0, 0.0], "energy": [0.0, 3.0], "gamma": [0.0, 0.5],
            "gamma_tilde": [0.0, -0.5],
            "source": "tests/riccati_solutions_test.cu:21-55"},
        "num_ozaki_t05_cutoff_10_over_2pi": {
            "value": 3, "source": "tests/parameters_test.cu:45-59"},
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(answers, f, indent=1)
    ex_dir = os.path.join(HERE, "examples")
    os.makedirs(ex_dir, exist_ok=True)
    ex_root = os.path.join(root, "examples")
    for name in sorted(os.listdir(ex_root)):
        cfg_path = os.path.join(ex_root, name, "simulation_config.json")
        if os.path.isfile(cfg_path):
            with open(cfg_path) as f:
                cfg = json.load(f)
            with open(os.path.join(ex_dir, name + ".json"), "w") as f:
                json.dump(cfg, f, indent=1)
    print("wrote %d poles from %s; omega %.10f from %s" % (len(poles), src, omega, osrc))


if __name__ == "__main__":
    main()
