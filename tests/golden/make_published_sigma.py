"""Writes tests/golden/published_sigma_p5_9.json from the reference's published run
output/p5_9_ako_apples_to_apples.json (substrate_sm_derivation/cy3_solver/): per-seed
sigma at the FS-Gram identity metric and after donaldson_solve(50, 1e-12), with the iteration
count, for k = 3 and k = 4 at 200,000 Shiffman-Zelditch points. Only these numbers are kept
(a data fixture); run once in a container that has the reference checkout."""
import json
import os
import sys

SRC = "/root/reference/substrate_sm_derivation/cy3_solver/output/p5_9_ako_apples_to_apples.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published_sigma_p5_9.json")


def main() -> int:
    with open(SRC) as f:
        d = json.load(f)
    out = {
        "source": "substrate_sm_derivation/cy3_solver/output/p5_9_ako_apples_to_apples.json",
        "protocol": d["protocol"],
        "sampler": d["sampler"],
        "n_pts": d["n_pts"],
        "donaldson_max_iter": d["donaldson_iters"],
        "donaldson_tol": d["donaldson_tol"],
        "cases": [],
    }
    for kb in d["ks"]:
        for ps in kb["per_seed"]:
            out["cases"].append({
                "k": kb["k"], "n_basis": kb["n_basis"], "seed": ps["seed"],
                "sigma_fs_identity": ps["sigma_fs_identity"],
                "sigma_donaldson": ps["sigma_donaldson"],
                "donaldson_iters_run": ps["donaldson_iters_run"],
            })
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(out['cases'])} cases to {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
