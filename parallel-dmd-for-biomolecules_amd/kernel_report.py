"""Standalone dispatches of the list-rebuild, predictor and event-loop
kernels, for rocprofv3 trace and counter runs.  Prints one JSON line; for
the uneven64 list and predictor dispatches it carries their algorithmic
bytes and FP64 operations, so that a profile's kernel durations and HBM
counters give achieved GB/s and the FP64 fraction per kernel.

    python parallel-dmd-for-biomolecules_amd/kernel_report.py rebuild [--reps 5]
    python parallel-dmd-for-biomolecules_amd/kernel_report.py predict [--reps 5]
    python parallel-dmd-for-biomolecules_amd/kernel_report.py run --chains 4 --residues 8 [--entries 30000]

rebuild: the 64 systems, each advanced by --advance committed entries, then
--reps forced list updates of all of them at once (pdmd_rebuild: one
k_rebuild_lists dispatch, the sync pseudo-event's origin shift, wrap and
lists of nbor.f, then one k_rebuild_events dispatch, events.f over the new
lists; one workgroup per system in each).
predict: every listed up-pair of every system through the predictor hook
(pdmd_predict_pairs: one k_predict dispatch per system, one lane per pair),
--reps times.
run: one seeded system (--chains x --residues, --density, --tstar) alone,
--advance entries, then one k_run dispatch of --entries entries (the last
k_run dispatch of the process is the measured one).

Algorithmic bytes (the units the bench's roofline uses):
  lists   16 B per bead read (FP32 position + identity) + 4 B per list
          entry written (K_up + K_dn per bead) + 4 B per search cell
  events  53 B per pair prediction (partner state 48 B + index 4 B + code
          1 B) + 16 B per calendar leaf written
  predict 116 B per pair (two 48 B bead states, two 4 B indices in, an 8 B
          time and a 4 B type out; nothing is amortised over an owner)
FP64 operations: 37 per pair prediction (free flight, minimum image, dot
products, discriminant, root; a sqrt and a division count one each).
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)
import pdmd  # noqa: E402
from bench import build_systems, uneven64  # noqa: E402

FLOPS_PER_PREDICTION = 37.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kernel", choices=("rebuild", "predict", "run"))
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--advance", type=int, default=2000, help="committed entries per system before the dispatches")
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--chains", type=int, default=4)
    ap.add_argument("--residues", type=int, default=8)
    ap.add_argument("--density", type=float, default=0.0101)
    ap.add_argument("--tstar", type=float, default=0.27)
    ap.add_argument("--entries", type=int, default=30000)
    args = ap.parse_args()
    if args.kernel == "run":
        eng = pdmd.Engine(waves=args.waves)
        try:
            sid = eng.add_system(pdmd.generate([args.residues] * args.chains, args.density, args.tstar, seed=7),
                                 seed=100)
            eng.advance([sid], [args.advance])
            eng.advance([sid], [args.entries])
            st = eng.stats(sid)
            if st.status:
                raise SystemExit(f"device status {st.status}")
            print(json.dumps({"kernel": "run", "beads": st.nbeads, "waves": args.waves, "entries": args.entries,
                              "launch_ms": round(eng.last_launch_ms(), 3)}), flush=True)
        finally:
            eng.close()
        return
    specs = uneven64()
    systems = build_systems(specs)
    eng = pdmd.Engine(waves=args.waves)
    try:
        ids = [eng.add_system(s, seed=sp["seed"] + 17) for s, sp in zip(systems, specs)]
        eng.advance(ids, [args.advance] * len(ids))
        out = {"kernel": args.kernel, "workload": "uneven64", "systems": len(ids), "reps": args.reps}
        if args.kernel == "rebuild":
            for _ in range(args.reps):
                eng.rebuild(ids)
            st = [eng.stats(i) for i in ids]
            if any(s.status for s in st):
                raise SystemExit(f"device status error: {[s.status for s in st]}")
            beads = sum(s.nbeads for s in st)
            entries = sum(s.nbeads * (s.avg_up + s.avg_dn) for s in st)
            up = sum(s.nbeads * s.avg_up for s in st)
            cells = sum((s.cells_axis // s.cell_factor) ** 3 for s in st)
            out.update(beads=beads, list_entries=round(entries), up_entries=round(up), search_cells=cells,
                       lists_bytes_per_dispatch=round(16 * beads + 4 * entries + 4 * cells),
                       events_predictions_per_dispatch=round(up),
                       events_bytes_per_dispatch=round(53 * up + 16 * (beads + 3 * len(ids))),
                       events_fp64_flops_per_dispatch=round(FLOPS_PER_PREDICTION * up))
        else:
            pairs = []
            for sid in ids:
                upl, _ = eng.lists(sid)
                ii = np.concatenate([np.full(len(r), k + 1, np.int32) for k, r in enumerate(upl)])
                jj = np.concatenate(upl).astype(np.int32)
                pairs.append((sid, ii, jj))
            for _ in range(args.reps):
                for sid, ii, jj in pairs:
                    t, ty = eng.predict_pairs(sid, ii, jj)
                    if (ty == -99).any():
                        raise SystemExit(f"system {sid}: a pair without a code")
            n = np.array([len(ii) for _, ii, _ in pairs])
            out.update(pairs_total=int(n.sum()), pairs_min=int(n.min()), pairs_max=int(n.max()),
                       predict_bytes_total=int(116 * n.sum()),
                       predict_fp64_flops_total=int(FLOPS_PER_PREDICTION * n.sum()),
                       note="totals are per rep over the 64 dispatches (one per system)")
        print(json.dumps(out), flush=True)
    finally:
        eng.close()


if __name__ == "__main__":
    main()
