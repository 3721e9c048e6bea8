"""Runs the experiment grid of settings.py and appends every program's CSV
line to <output>/<executable>_<identifier>.csv, as the reference's runner does
(scripts/run_experiments.py:8-80): one process per run, `repetitions` runs
per argument tuple, the header line written once per file.

Differences from the reference's runner:
* each run has a time limit (--timeout) and its exit status is checked: a
  run that fails with exit status 1 (the programs' error exit) is reported
  with the tail of its output, not appended, and makes the runner exit 1
  after the grid (the reference drops failures silently); a run that hits
  its time limit, dies by a signal or exits otherwise stops the grid at once
  (exit 3);
* every output line is checked against the experiment's column count;
* --scale multiplies the key and byte counts, --max-elements skips runs above
  a key count, --repetitions overrides the repetitions, --dry-run prints the
  commands;
* profiled experiments run under rocprofv3 (kernel trace and stats, CSV
  output under <output>/rocprof/) instead of nvprof/nsys;
* numactl binds as the reference does (-N 0 -m 0, or -m 0 for host sorts)
  only when asked (--numactl) and installed.

    python run_experiments.py -e local_gpu_sort_algorithm_to_sort_duration --scale 0.25 --repetitions 3
"""
import argparse
import csv
import io
import pathlib
import shutil
import subprocess
import sys
import time
from typing import List, Optional, Tuple

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import settings  # noqa: E402


def scaled_arguments(x: settings.Experiment, args: Tuple, scale: float) -> Tuple:
    out = list(args)
    if scale != 1.0:
        for i in settings.size_positions(x):
            v = int(out[i])
            out[i] = v if v == 0 else max(1, int(round(v * scale)))
        if x.executable in settings.LIST_ARGS:
            i = settings.LIST_ARGS[x.executable]
            out[i] = ",".join(str(max(1, int(round(int(v) * scale)))) for v in str(out[i]).split(","))
    return tuple(out)


def num_keys(x: settings.Experiment, args: Tuple) -> int:
    """Keys (or bytes) a run handles: what --max-elements bounds."""
    if x.executable in settings.LIST_ARGS:
        return sum(int(v) for v in str(args[settings.LIST_ARGS[x.executable]]).split(","))
    return int(args[0])


def is_host_run(x: settings.Experiment, args: Tuple) -> bool:
    if x.executable in settings.CPU_PROGRAMS:
        return True
    if x.executable == "sort_benchmark":
        return str(args[x.parameters.index("--algorithm")]) in settings.CPU_SORT_ALGORITHMS
    return False


def command(x: settings.Experiment, args: Tuple, bindir: pathlib.Path, numactl: bool,
            profile_dir: Optional[pathlib.Path]) -> List[str]:
    cmd = [str(bindir / x.executable)]
    for p, a in zip(x.parameters, args):
        if p:
            cmd.append(p)
        cmd.append(str(a))
    if profile_dir is not None:
        cmd = ["rocprofv3", "--kernel-trace", "--stats", "--output-format", "csv", "-d", str(profile_dir), "-o",
               "run", "--"] + cmd
    if numactl and shutil.which("numactl"):
        cmd = (["numactl", "-m", "0"] if is_host_run(x, args) else ["numactl", "-N", "0", "-m", "0"]) + cmd
    return cmd


def csv_lines(text: str, ncols: int, strict: bool) -> Optional[List[str]]:
    """The CSV lines of a program's output, or None if there is none or (when
    strict) a line has the wrong number of fields.  Not strict: lines of
    other output (a profiler's) are dropped."""
    keep = []
    for line in text.splitlines():
        if not line.strip():
            continue
        fields = next(csv.reader(io.StringIO(line)), [])
        if len(fields) == ncols:
            keep.append(line)
        elif strict:
            return None
    return keep or None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="run the one-GPU experiment grid")
    ap.add_argument("-e", "--experiments", default="", help="comma-separated identifiers (default: all)")
    ap.add_argument("--bin", default=str(HERE.parent / "bin"), help="directory of the benchmark programs")
    ap.add_argument("--output", default=None,
                    help="output directory (default: experiments/results/<timestamp> next to this script)")
    ap.add_argument("--repetitions", type=int, default=None, help="runs per argument tuple (default: the grid's)")
    ap.add_argument("--scale", type=float, default=1.0, help="factor on key and byte counts")
    ap.add_argument("--max-elements", type=float, default=None, help="skip runs above this many keys or bytes")
    ap.add_argument("--where", action="append", default=[],
                    help="keep only argument tuples holding this value (repeatable: all must match)")
    ap.add_argument("--timeout", type=float, default=1800.0, help="seconds per run")
    ap.add_argument("--numactl", action="store_true", help="bind runs with numactl as the reference does")
    ap.add_argument("--dry-run", action="store_true", help="print the commands only")
    ap.add_argument("--list", action="store_true", help="list the experiments and exit")
    a = ap.parse_args(argv)

    exps = settings.experiments()
    if a.list:
        for x in exps:
            print("%-64s %-36s %4d runs%s" % (x.identifier, x.executable, len(x.arguments),
                                              "  (%s)" % x.notes if x.notes else ""))
        for ident, why in settings.OMITTED.items():
            print("%-64s omitted: %s" % (ident, why))
        return 0
    wanted = [s for s in a.experiments.split(",") if s]
    known = {x.identifier for x in exps}
    unknown = [w for w in wanted if w not in known]
    if unknown:
        print("unknown experiments: %s" % ", ".join(unknown), file=sys.stderr)
        return 2
    if a.scale <= 0 or (a.repetitions is not None and a.repetitions < 1):
        print("--scale must be positive and --repetitions at least 1", file=sys.stderr)
        return 2
    bindir = pathlib.Path(a.bin).resolve()
    out = pathlib.Path(a.output) if a.output else HERE / "results" / time.strftime("%Y_%m_%d_%H_%M_%S")
    if not a.dry_run:
        out.mkdir(parents=True, exist_ok=True)

    plan = []
    for x in exps:
        if wanted and x.identifier not in wanted:
            continue
        for index, args in enumerate(x.arguments):
            if any(w not in [str(v) for v in args] for w in a.where):
                continue
            args = scaled_arguments(x, args, a.scale)
            if a.max_elements is not None and num_keys(x, args) > a.max_elements:
                print("skip %s %s: above --max-elements" % (x.identifier, args), flush=True)
                continue
            for rep in range(a.repetitions or x.repetitions):
                plan.append((x, index, args, rep))

    failed = 0
    for k, (x, index, args, rep) in enumerate(plan):
        prof = None
        if x.profiler == "rocprofv3":
            prof = out / "rocprof" / ("%s_%s_%d_%d" % (x.executable, x.identifier, index + 1, rep + 1))
        cmd = command(x, args, bindir, a.numactl, prof)
        if a.dry_run:
            print(" ".join(cmd))
            continue
        t0 = time.time()
        try:
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, universal_newlines=True,
                               timeout=a.timeout)
            rc, text, err = r.returncode, r.stdout, r.stderr
        except subprocess.TimeoutExpired:
            rc, text, err = -1, "", "time limit of %.0f s" % a.timeout
        except OSError as ex:
            rc, text, err = None, "", str(ex)
        rows = csv_lines(text, len(x.columns), strict=prof is None) if rc == 0 else None
        status = "ok" if rows is not None else "FAILED"
        print("[%d/%d] %s rep %d: %s -> %s %.2f s" % (k + 1, len(plan), x.identifier, rep + 1, " ".join(cmd),
                                                      status, time.time() - t0), flush=True)
        if rows is None:
            failed += 1
            tail = (text + err).strip().splitlines()[-5:]
            print("    exit %s; %s" % (rc, " | ".join(tail)), flush=True)
            if rc is not None and (rc < 0 or rc > 1):
                # a time limit, a signal or an abort: the device may be in a
                # bad state, so nothing more runs
                print("stopping: run ended abnormally (exit %s)" % rc, flush=True)
                return 3
            continue
        path = out / ("%s_%s.csv" % (x.executable, x.identifier))
        new = not path.exists()
        with path.open("a") as f:  # the program's lines as printed
            if new:
                f.write(",".join('"%s"' % c for c in x.columns) + "\n")
            f.write("\n".join(rows) + "\n")
    if not a.dry_run:
        print("%d runs, %d failed; results in %s" % (len(plan), failed, out), flush=True)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
