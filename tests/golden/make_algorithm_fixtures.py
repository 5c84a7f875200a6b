"""Writes tests/golden/algorithms_reference.json: values the reference produced for its
gate-level algorithms, taken from the data files it ships (run in a container that has the
reference checkout; the output is committed).

Sources:
  docs/algorithms_benchmark.json  -- Grover peaks at n = 8..14 (automatic iteration count),
                                     QAOA MaxCut on an 8-node weighted graph (angles, expected
                                     cuts, the p = 4 distribution), VQE H2 energy histories
  docs/library_comparison.json    -- Grover, 10 iterations, n = 10..18: p_success and argmax
  benchmarks/algorithms_benchmark.py:22-26, :66-70 -- the H2 Hamiltonian those histories use
    (H = g0 + g1 Z0 + g2 Z1 + g3 Z0 Z1 + g4 X0 X1, STO-3G at 0.735 A, parity mapping; the five
    coefficients and the published FCI energy are data, restated here).
"""
import json
import os

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "algorithms_reference.json")


def main():
    alg = json.load(open(os.path.join(REF, "docs", "algorithms_benchmark.json")))
    lib = json.load(open(os.path.join(REF, "docs", "library_comparison.json")))
    out = {
        "grover_auto": [{k: r[k] for k in ("n", "marked", "iterations", "found", "peak", "theory")} for r in alg["grover"]],
        "grover_10_iterations": {
            "iterations": lib["grover_iterations"],
            "rows": [{k: r[k] for k in ("n", "p_success", "argmax", "p_total")} for r in lib["rows"]["ours"]
                     if r.get("status") == "ok"],
        },
        "qaoa": {k: alg["qaoa"][k] for k in ("n_nodes", "edges", "optimum", "best_distribution_p",
                                              "best_distribution")},
        "vqe_h2": {
            "coefficients": [-1.052373245772859, 0.39793742484318045, -0.39793742484318045,
                             -0.01128010425623538, 0.18093119978423156],
            "e_fci_electronic_published": -1.857275030202382,
            "layers": alg["vqe_h2"]["layers"],
            "lr": alg["vqe_h2"]["lr"],
            "energy_history": alg["vqe_h2"]["energy_history"],
        },
    }
    out["qaoa"]["layers"] = [{k: l[k] for k in ("p", "gammas", "betas", "expected_cut", "best_cut", "p_optimum")}
                             for l in alg["qaoa"]["layers"]]
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
