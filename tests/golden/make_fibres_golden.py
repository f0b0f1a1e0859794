"""Writes tests/golden/fibres_reference.json from the reference's fibre distributions.

Imports src/jaxmaterials/data/distributions_fibres.py of a reference checkout
(it needs numpy and scipy only; its fast path compiles the reference's
libfibres.cc with g++ into a temporary directory) and records, for fixed
seeds and domains:

* FibreRadiusDistribution.draw (gaussian and constant radii),
* FibreDistribution2d's regular starting grid,
* two successive draws of FibreDistribution2d's iterator (fibre locations
  after the random movement, and their radii).

Floats are stored as JSON numbers (repr round-trips float64 exactly).

    python tests/golden/make_fibres_golden.py [/path/to/reference]
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref="/root/reference"):
    sys.path.insert(0, os.path.join(ref, "src"))
    from jaxmaterials.data.distributions_fibres import FibreDistribution2d, FibreRadiusDistribution

    out = {"radius_draws": [], "distributions": []}
    for gaussian in (True, False):
        dist = FibreRadiusDistribution(r_avg=7.5e-3, r_min=5.0e-3, r_max=10.0e-3, sigma=0.5e-3, gaussian=gaussian)
        out["radius_draws"].append({"r_avg": 7.5e-3, "r_min": 5.0e-3, "r_max": 10.0e-3, "sigma": 0.5e-3,
                                    "gaussian": gaussian, "seed": 856219, "sizes": [42, 10],
                                    "draws": [dist.draw(42).tolist(), dist.draw(10).tolist()]})
    cases = [([0.5, 0.04], 0.3), ([0.5, 0.07], 0.3), ([0.2, 0.1], 0.5)]
    for domain, vf in cases:
        rdist = FibreRadiusDistribution(r_avg=7.5e-3, r_min=5.0e-3, r_max=10.0e-3, sigma=0.5e-3, gaussian=True)
        fd = FibreDistribution2d(domain_size=domain, volume_fraction=vf, r_fibre_dist=rdist, seed=1, fast_code=True)
        assert fd._fast_code, "the reference's libfibres.cc did not compile"
        it = iter(fd)
        draws = []
        for _ in range(2):
            X, R = next(it)
            draws.append({"locations": X.tolist(), "radii": R.tolist()})
        out["distributions"].append({"domain_size": domain, "volume_fraction": vf, "radius_seed": 856219,
                                     "initial_locations": fd._initial_fibre_locations.tolist(), "draws": draws})
    path = os.path.join(HERE, "fibres_reference.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path)


if __name__ == "__main__":
    main(*sys.argv[1:])
