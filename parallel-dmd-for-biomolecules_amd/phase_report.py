"""Where the event loop's cycles go: device cycles per phase, per committed
calendar entry and per round, for single systems of a few sizes.

It loads the timer build of the engine (libpdmd_phases.so, made by
`make -C parallel-dmd-for-biomolecules_amd phases`, same sources with
-DPDMD_PHASE_TIMERS).  The timers count clock64() cycles of the first lane
of wavefront 0 in pdmd_stats_t.phase_cycles; phases that other wavefronts
run in parallel appear as wavefront 0 sees them, barrier waits included.
Nested phases are listed under their parent (list rebuild steps, selection
steps) and left out of the total.

    python parallel-dmd-for-biomolecules_amd/phase_report.py [--waves 4]
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)
import pdmd  # noqa: E402

# slot: (name, parent slot or None); slots 13/14, 21-23, 28-29 and 32-36 are counts
PHASES = {
    0: ("select", None), 30: ("select: pops", 0), 31: ("select: scalar copies", 0),
    1: ("pair event before rescheduling", None), 11: ("resched: list rows + work list", None),
    12: ("resched: down-set scan", None), 2: ("resched: loads", None), 3: ("resched: predictions", None),
    4: ("resched: down-set updates", None), 5: ("resched: leaf writes", None), 6: ("block rescans", None),
    24: ("round: footprint marks + barrier", None), 25: ("round: footprint checks + barrier", None),
    26: ("round: new keys + barrier", None), 27: ("round: rescans, merge, barrier", None),
    7: ("sync pseudo-event (no rebuild part)", None), 10: ("ghost event before rescheduling", None),
    8: ("list rebuild", None), 16: ("list: cells + histogram", 8), 17: ("list: count scan", 8),
    18: ("list: stable scatter", 8), 19: ("list: sorted copies", 8), 20: ("list: neighbour search", 8),
    9: ("events() full reschedule", None),
}
CASES = [  # name, chains, residues per chain, density, T*
    ("128 beads", 4, 8, 0.0101, 0.27),
    ("512 beads", 8, 16, 0.0101, 0.27),
    ("6144 beads", 96, 16, 0.0101, 0.27),
]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--events", type=int, default=30000)
    ap.add_argument("--warmup", type=int, default=2000)
    args = ap.parse_args()
    pdmd.LIB_PATH = os.path.join(HERE, "libpdmd_phases.so")
    if not os.path.exists(pdmd.LIB_PATH):
        raise SystemExit("libpdmd_phases.so not built: make -C parallel-dmd-for-biomolecules_amd phases")
    print(f"{'':40s}" + "".join(f"{c[0]:>26s}" for c in CASES))
    print(f"{'':40s}" + "".join(f"{'cyc/entry':>13s}{'cyc/round':>13s}" for _ in CASES))
    cols = []
    for name, nch, L, dens, tstar in CASES:
        e = pdmd.Engine(waves=args.waves)
        try:
            sid = e.add_system(pdmd.generate([L] * nch, dens, tstar, seed=7), seed=100)
            e.advance([sid], [args.warmup])
            s0 = e.stats(sid)
            e.advance([sid], [args.events])
            s1 = e.stats(sid)
            ms = e.last_launch_ms()
            if s1.status != 0:
                raise SystemExit(f"{name}: status {s1.status}")
        finally:
            e.close()
        d = [s1.phase_cycles[k] - s0.phase_cycles[k] for k in range(40)]
        if not any(d[k] for k in PHASES):
            raise SystemExit("the loaded library has no phase timers")
        cols.append((d, ms))
    def cell(v, d):
        # per entry, and per round of two or more events (none at W = 1)
        return f"{v / args.events:13.0f}" + (f"{v / d[21]:13.0f}" if d[21] else f"{'-':>13s}")

    for k, (label, parent) in PHASES.items():
        if args.waves == 1 and k in (24, 25, 26, 27):
            continue
        print(f"{('  ' if parent is not None else '') + label:40s}" + "".join(cell(d[k], d) for d, _ in cols))
    print(f"{'total (top-level phases)':40s}" +
          "".join(cell(sum(d[k] for k, (_, p) in PHASES.items() if p is None), d) for d, _ in cols))
    for (name, *_), (d, ms) in zip(CASES, cols):
        print(f"{name}: {args.events / (ms / 1e3):.0f} entries/s at W={args.waves}; rounds of two or more events "
              f"{d[21]}, popped {d[22]}, committed {d[23]}, cut by footprints {d[28]}, cut by key order {d[29]}; "
              f"list rebuilds {d[13]}; rescheduling batches: {d[32]} pair predictions in {d[33]} 64-lane passes "
              f"(lane occupancy {d[32] / max(64 * d[33], 1):.3f}, wavefront 0); outside the batches: {d[34]} "
              f"single-pair updates, {d[35]} one-bead full recomputes, {d[36]} generic reschedules (wavefront 0)")


if __name__ == "__main__":
    main()
