"""Extracts the reference's published per-case golden values into corpus_golden.json.

Reads (in the build container only) the reference's benchmark result files:
  benchmarks/results/phase2_planar_bakeoff_a2b73bd.json           (costs, rounds, work)
  benchmarks/results/phase3_persistent_compact_sweep_c68829c.json (case descriptors,
                                                                    phase-3 base costs)
and writes the numbers as a small data fixture next to this script.  Nothing of
the reference is imported or executed.
"""
import json
from pathlib import Path

REF = Path("/root/reference/benchmarks/results")
OUT = Path(__file__).resolve().parent / "corpus_golden.json"


def main():
    p2 = json.loads((REF / "phase2_planar_bakeoff_a2b73bd.json").read_text())
    cases = {}
    for key, val in p2["context"].items():
        if key.startswith("apgar_case_"):
            name = key[len("apgar_case_"):]
            fields = dict(kv.split("=", 1) for kv in val.split(";"))
            cases[name] = {"descriptor": fields}
    for b in p2["benchmarks"]:
        if not b["name"].endswith("real_time_median"):
            continue
        gen, name = b["name"].split("/")[:2]
        c = cases[name]
        c[f"{gen}_path_cost"] = int(b["path_cost"])
        c[f"{gen}_examined_work"] = int(b["examined_work"])
        c[f"{gen}_rounds"] = int(b["rounds"])
    # Phase-3 corpus adds three cases; their descriptors and base-policy costs.
    p3_path = REF / "phase3_persistent_compact_sweep_c68829c.json"
    with p3_path.open() as f:
        head = []
        for line in f:
            head.append(line)
            if line.startswith("  \"benchmarks\""):
                break
    # Case descriptors come from the context block, read line by line.
    for line in head:
        line = line.strip().rstrip(",")
        if line.startswith('"apgar_case_') and "_policy_identities" not in line:
            key, val = line.split(":", 1)
            name = key.strip('"')[len("apgar_case_"):]
            fields = dict(kv.split("=", 1) for kv in json.loads(val).split(";"))
            cases.setdefault(name, {})["descriptor"] = fields
    wanted = {"symmetric_dual_corridor", "multi_channel_bottleneck", "policy_alternatives"}
    with p3_path.open() as f:
        data = json.load(f)
    for b in data["benchmarks"]:
        parts = b["name"].split("/")
        if (len(parts) > 3 and parts[3] in wanted and parts[1] == "sequential_cpu_astar"
                and b.get("base_policy_scalar_cost")):
            cases[parts[3]]["base_policy_scalar_cost"] = int(b["base_policy_scalar_cost"])
    OUT.write_text(json.dumps(cases, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT} ({len(cases)} cases)")


if __name__ == "__main__":
    main()
