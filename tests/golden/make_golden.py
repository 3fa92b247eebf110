"""Extract the known-answer values of the stencil path from the reference's own
output files into small fixtures (data only, no reference code).

Run in the build container, where the reference tree is mounted:
    python tests/golden/make_golden.py /root/reference

Writes
  propa_param_analysis.csv  every row of
      script/testCase_Propa/paramAnalysis/hpcscanPropaParamAnalysisShaheen.log
      (perf-log columns: propagator, n, fdOrder, errTestCase, nt, dt, stableDt).
      The log does not store the command line; its nt / dt / stableDt columns
      are reproduced exactly by -dim 3 -param1..3 130 -tmax 0.1 -snapDt 0.01
      with -ratioCFL 1.0, 0.5 or 0.1 (the tests search that ratio per row).
  propa_default_61.json     the key lines of misc/fileForReadme/runPropaTestCase.txt
"""
import csv
import json
import os
import sys


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    log = os.path.join(ref, "script/testCase_Propa/paramAnalysis/hpcscanPropaParamAnalysisShaheen.log")
    rows = []
    with open(log) as f:
        for line in f:
            t = line.split()
            if len(t) < 22 or t[1] != "Propa":
                continue
            # host case mode propagator | 1 nproc 2-4 nsub 5 nthread 6-8 n 9 order |
            # 10 gflop 11 gptFD 12 gptEff 13 GB 14 time 15 err 16 nt 17 dt 18 stableDt
            rows.append({"propagator": t[3], "n": int(t[9]), "fd_order": int(t[12]),
                         "err": t[18], "nt": int(t[19]), "dt": t[20], "stable_dt": t[21]})
    with open(os.path.join(here, "propa_param_analysis.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)

    txt = os.path.join(ref, "misc/fileForReadme/runPropaTestCase.txt")
    keys = ["CFL", "Stable time step (s)", "No. time steps", "Time step (s)", "Max. frequency (Hz)",
            "Total n1", "Inner n1", "Grid size (MB)", "Grid Inner Glob. (Pts)", "ntCheck", "errTestCase2",
            "#Flop per point", "#Point in stencil"]
    out = {}
    with open(txt) as f:
        for line in f:
            s = line.strip()
            for k in keys:
                if s.startswith(k) and k not in out:
                    out[k] = s[len(k):].strip()
    out["_command"] = "-testCase Propa -testMode CacheBlk (defaults: 61^3, fdOrder 8, nt 20, ntry 10)"
    with open(os.path.join(here, "propa_default_61.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"{len(rows)} param-analysis rows, {len(out)} default-run values")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
