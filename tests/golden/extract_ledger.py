"""Write ledger_energies.json from the reference's benchmark ledgers.

Reads benchmarks/adr0047.jsonl of the reference checkout (only where it is
present) and keeps, per (case, n_atoms), the single-thread energy per atom,
the cutoff and the translation-invariance residual.  The geometry those
rows were measured on is benchmarks/bench_force.py:152-158 (see
tests/test_oracle_golden.py for the builders).
"""
import json
import os
import sys

REF = os.environ.get("AK_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = os.path.join(REF, "benchmarks", "adr0047.jsonl")
    if not os.path.exists(src):
        print("reference ledger not found; nothing written", file=sys.stderr)
        return 1
    out = {}
    with open(src) as fh:
        for lineno, line in enumerate(fh, 1):
            if not line.strip():
                continue
            rec = json.loads(line)
            if rec.get("threads") != 1:
                continue
            key = f"{rec['case']}/{rec['n_atoms']}"
            out[key] = {"e_per_atom": rec["e_per_atom"], "r_cut": rec["r_cut"],
                        "sum_force_max": rec["sum_force_max"],
                        "source": f"benchmarks/adr0047.jsonl:{lineno}"}
    with open(os.path.join(HERE, "ledger_energies.json"), "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(f"{len(out)} entries")
    return 0


if __name__ == "__main__":
    sys.exit(main())
