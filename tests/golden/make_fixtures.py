"""Extract the golden vectors used by the tests from the reference's data
tables (data/analysis_tables/*.csv).  Run in a container where the reference
checkout is mounted; the outputs are small CSV fixtures committed next to this
script so that no test needs the reference at run time.

    python tests/golden/make_fixtures.py /root/reference
"""

import csv
import os
import sys

COLS = ["preset", "seed", "solid_probability", "baseline_rho_min", "raised_rho_mins",
        "recommended_rho_min", "attempted_rho_mins", "penal", "probe_iters", "restart",
        "maxiter", "tol", "solid_count", "n_components", "largest_component",
        "support_to_load_patch_components", "probe_r50", "probe_r100", "trigger",
        "predicted_raise_floor", "post_solve_fallback_triggered", "detector_false_keep",
        "solve_converged", "solver_reported_converged", "solve_iters",
        "solve_final_rel_residual", "recorded_policy_time_s"]

TABLES = {
    "heldout_cantilever_guarded.csv":
        "heldout_gmg_detector_cantilever_s41_71_guarded_true_residual__prospective_summary.csv",
    "heldout_cantilever_strict.csv":
        "heldout_gmg_detector_cantilever_s41_71_strict_true_residual__prospective_summary.csv",
    "heldout_bridge_guarded.csv":
        "heldout_gmg_detector_bridge_s41_59_guarded_true_residual__prospective_summary.csv",
    "fine_ladder_bridge.csv":
        "fine_ladder_bridge_seed23_41_strict_true_residual__prospective_summary.csv",
}
HISTORY_SRC = ("heldout_gmg_detector_cantilever_s41_71_guarded_true_residual__"
               "prospective_history.csv")
HISTORY_STATES = {(43, "0.35"), (41, "0.08"), (47, "0.35")}


def main(ref_root):
    src = os.path.join(ref_root, "data", "analysis_tables")
    out = os.path.dirname(os.path.abspath(__file__))
    for dst, name in TABLES.items():
        with open(os.path.join(src, name), newline="") as f:
            rows = list(csv.DictReader(f))
        cols = [c for c in COLS if c in rows[0]]
        with open(os.path.join(out, dst), "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=cols)
            w.writeheader()
            for r in rows:
                w.writerow({c: r[c] for c in cols})
    with open(os.path.join(src, HISTORY_SRC), newline="") as f:
        rows = [r for r in csv.DictReader(f)
                if (int(r["seed"]), r["solid_probability"]) in HISTORY_STATES]
    with open(os.path.join(out, "heldout_cantilever_history_subset.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    # fixed-floor compliance controls (compliance per floor)
    name = "gmg_fixed_floor_controls_strict_true_residual__fixed_floor_control_comparison.csv"
    with open(os.path.join(src, name), newline="") as f:
        rows = list(csv.DictReader(f))
    with open(os.path.join(out, "fixed_floor_controls.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    # optimized-density input C64_MF (80x40x20 cantilever), data file of the reference
    import shutil
    shutil.copyfile(os.path.join(ref_root, "data", "optimized_density_inputs", "C64_MF__rho_final.npy"),
                    os.path.join(out, "C64_MF__rho_final.npy"))
    name = "optimized_density_C64_MF_strict_true_residual__density_detector_summary.csv"
    with open(os.path.join(src, name), newline="") as f:
        rows = list(csv.DictReader(f))
    with open(os.path.join(out, "optimized_density_C64_MF.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    # CPU direct conditioning table (24x12x6 cantilever eigenvalues)
    name = "direct_floor_conditioning__direct_floor_conditioning.csv"
    with open(os.path.join(src, name), newline="") as f:
        rows = list(csv.DictReader(f))
    with open(os.path.join(out, "direct_floor_conditioning.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    # 512k-element optimized densities of the fixed-floor controls (cantilever
    # 160x80x40 and bracket 80x160x40): the arrays, compressed, with the
    # reference manifest's SHA-256 of the original .npy file
    import hashlib

    import numpy as np

    man = {}
    with open(os.path.join(ref_root, "data", "optimized_density_inputs", "MANIFEST.csv"),
              newline="") as f:
        for r in csv.DictReader(f):
            man[r["file"]] = r["sha256"]
    for case in ("C512_MF", "Brk500_MF"):
        name = f"{case}__rho_final.npy"
        path = os.path.join(ref_root, "data", "optimized_density_inputs", name)
        raw = open(path, "rb").read()
        assert hashlib.sha256(raw).hexdigest() == man[name], name
        rho = np.load(path, allow_pickle=False)
        np.savez_compressed(os.path.join(out, f"{case}__rho_final.npz"), rho=rho,
                            npy_sha256=np.array(man[name]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
