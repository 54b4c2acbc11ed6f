"""Command-line shims with the reference engines' argv, output files and timing lines.

  python -m sgfim.cli CoParaCG_bmma     <fimi> <rel_thr>      (also CoParaCG_baseline)
  python -m sgfim.cli AnytimeMining     <fimi> <upto_stage>

They replace the engine executables the reference pipeline spawns
(code/pipeline/engine_adapter.py:24-26, invoked at :83 and :122): the same two
positional arguments, the same Results file next to the dataset
(<fimi>-<thr:.6f>=Results.txt, <fimi>-<k>-stages=Results.txt), and the TensorFIM
engines' `computation_time(s):` / `total_time(s):` lines on stdout, which the
reference adapter parses (engine_adapter.py:28-30). The files under sgfim/bin/ carry
the executables' names, so pointing the reference's TENSORFIM_DIR / ANYFIM_EXE at that
directory swaps the engines without touching the reference adapter.

Exit codes are the engines': 1 after a run, whether it wrote its Results file or not
(kernel_bitmap.cu:801,910; AnyTime kernel.cu:803,950; "success is the produced
Results file, not the exit code", TensorFIM README.md:85-86), and 0 when AnytimeMining
is asked for more stages than there are items (kernel.cu:804-806). A wrong argument
count exits 2 (the engines read their arguments without checking them).
"""
from __future__ import annotations

import os
import sys
import time


def _usage(msg: str) -> int:
    print(msg, file=sys.stderr)
    return 2


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv:
        return _usage("usage: sgfim.cli {CoParaCG_baseline|CoParaCG_bmma|AnytimeMining} <fimi> <arg>")
    mode, args = os.path.basename(argv[0]), argv[1:]
    if mode.endswith(".exe"):
        mode = mode[:-4]
    from . import engine_adapter as EA
    if mode in ("CoParaCG_baseline", "CoParaCG_bmma"):
        if len(args) != 2:
            return _usage(f"usage: {mode} <fimi> <relative threshold>")
        fimi, thr = args[0], float(args[1])
        t0 = time.perf_counter()
        r = EA.run_tensorfim(mode.split("_", 1)[1], fimi, thr)
        if not r["ok"]:
            print(r.get("log", "run failed"), file=sys.stderr)
            return 1
        print(f"MFIs: {len(r['mfis'])}")
        print(f"computation_time(s): {r['comp_times'][-1]:.6f}")
        print(f"total_time(s): {max(r['tot_times'][-1], time.perf_counter() - t0):.6f}")
        return 1
    if mode == "AnytimeMining":
        if len(args) != 2:
            return _usage("usage: AnytimeMining <fimi> <upto_stage>")
        r = EA.run_anyfim(args[0], int(args[1]))
        if not r["ok"]:
            print(r.get("log", "run failed"))
            # the engine refuses more stages than items, writes no file and exits 0
            # (kernel.cu:804-806); an unreadable input exits 1 (kernel.cu:803)
            return 0 if r.get("refused") else 1
        print(f"stages: {len(r['stages'])}, results: {r['results_path']}")
        return 1
    return _usage(f"unknown engine {mode!r}")


if __name__ == "__main__":
    sys.exit(main())
