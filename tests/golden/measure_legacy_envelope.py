"""Writes legacy_config1_envelope.json: how far the 64 x 64 config-1
trajectory (extension "none") moves when its initial psi is perturbed at
rounding level, per step, measured with the oracle (oracle/legacy_step.py,
itself pinned to the reference trajectory by test_oracle_legacy_step.py).

For five seeds, psi0 (1 + 1e-13 g), g standard normal, is stepped 20 times
next to the unperturbed run; the file holds, per step, the largest relative
deviation over the seeds of max|div u|, ||p||_2 and the max-norm of psi, u,
v and p.  A GPU run differs from the dense-LU CCD of the reference by about
1e-13 per operation, so its trajectory is judged against this envelope:

    python tests/golden/measure_legacy_envelope.py
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import legacy_step as ls  # noqa: E402


def run(psi0, steps=20):
    sim = ls.OracleLegacySimulation((64, 64), (1.0, 1.0))
    sim.psi = psi0.copy()
    sim.update_properties()
    sim.update_curvature()
    out = []
    for _ in range(steps):
        sim.step_forward(sim.cfl_dt())
        out.append(dict(div=sim.divergence_max(), p_l2=float(np.linalg.norm(sim.p)),
                        psi=sim.psi.copy(), u=sim.vel[0].copy(), v=sim.vel[1].copy(),
                        p=sim.p.copy()))
    return out


def main(amp=1e-13, seeds=5, steps=20):
    t = np.load(os.path.join(HERE, "legacy_config1_trajectory.npz"))
    base = run(t["psi0"], steps)
    worst = [dict.fromkeys(("div", "p_l2", "psi", "u", "v", "p"), 0.0) for _ in range(steps)]
    for seed in range(seeds):
        rng = np.random.default_rng(seed)
        other = run(t["psi0"] * (1.0 + amp * rng.standard_normal(t["psi0"].shape)), steps)
        for k in range(steps):
            w, a, b = worst[k], base[k], other[k]
            for key in ("div", "p_l2"):
                w[key] = max(w[key], abs(b[key] / a[key] - 1.0))
            for key in ("psi", "u", "v", "p"):
                w[key] = max(w[key], float(np.abs(b[key] - a[key]).max() / np.abs(a[key]).max()))
    with open(os.path.join(HERE, "legacy_config1_envelope.json"), "w") as fh:
        json.dump({"perturbation": amp, "seeds": seeds, "per_step": worst}, fh, indent=1)
    for k, w in enumerate(worst):
        print(k + 1, {key: f"{v:.1e}" for key, v in w.items()})


if __name__ == "__main__":
    main()
