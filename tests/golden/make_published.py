"""Writes tests/golden/published.json: the reference's own published result data.

Reads, as text, the job outputs the reference repository holds under
jobs/paperJobs/ (no reference code is run or imported):
  * 01-Convergence/result.dat          Parareal error vs iteration k (rows: N, nufreq)
  * 04-SpeedupGPU*/run_<N>_0.log       banner, run times and "Error at end" (K=3 / K=6)
  * 05-SpeedupCPU*/run_<N>_0.log       the same on the CPU backend
All runs: 128^3, nu0 = 0.1, c = (1,1,1), T = 0.1, 32768 fine / 2048 coarse steps,
blocking communication (jobs/paperJobs/01-Convergence.job:51-57, 04-SpeedupGPU.job:42-48).

    python tests/golden/make_published.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
JOBS = os.path.join(REF, "jobs", "paperJobs")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "published.json")


def parse_log(path):
    d = {}
    for line in open(path):
        m = re.match(r"^ - (.+?): (.+)$", line.rstrip())
        if m:
            d[m.group(1)] = m.group(2)
        for key in ("Serial run time", "Parallel run time", "Speedup", "Maximal speedup", "Error at end"):
            m = re.match(r"^%s: (.+)$" % key, line.rstrip())
            if m:
                d[key] = float(m.group(1))
    return d


def main():
    out = {
        "setup": {"gridSize": 128, "nu0": 0.1, "cx": 1.0, "cy": 1.0, "cz": 1.0, "endTime": 0.1,
                  "timeStepsFine": 32768, "timeStepsCoarse": 2048, "coarse": "euler_upwind",
                  "source": "jobs/paperJobs/01-Convergence.job:51-57; 04-SpeedupGPU.job:42-48"},
        "convergence": [],
        "runs": {},
    }
    for line in open(os.path.join(JOBS, "01-Convergence", "result.dat")):
        v = line.split()
        if not v:
            continue
        out["convergence"].append({"slices": int(v[0]), "nufreq": float(v[1]),
                                   "error_k1_to_k8": [float(x) for x in v[2:]]})
    out["convergence_source"] = "jobs/paperJobs/01-Convergence/result.dat:1-6"
    for d in ("04-SpeedupGPU", "05-SpeedupCPU", "04-SpeedupGPU-6", "05-SpeedupCPU-6"):
        rows = {}
        for n in (4, 8, 16, 32, 64, 128):
            p = os.path.join(JOBS, d, f"run_{n}_0.log")
            if os.path.exists(p):
                rows[str(n)] = parse_log(p)
        out["runs"][d] = rows
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(OUT)


if __name__ == "__main__":
    main()
