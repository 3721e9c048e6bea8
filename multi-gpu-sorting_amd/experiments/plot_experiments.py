"""Plots the CSV files that run_experiments.py wrote and tabulates them, in
the manner of the reference's plot script (scripts/plot_experiments.py
:219-295: read each experiment's CSV, group the repetitions by their
configuration, take the mean, draw bars or lines per identifier).

For every identifier with a CSV in the run directory it writes
<identifier>.<format> (matplotlib, no display needed) and adds a table to
summary.md: the mean and the spread of the repetitions' durations, and a
throughput where the x axis is a key or byte count.

    python plot_experiments.py <run directory> [--format pdf]
"""
import argparse
import pathlib
import sys
from typing import Callable, Dict, List, NamedTuple, Optional

import numpy
import pandas

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
import settings  # noqa: E402


class Plot(NamedTuple):
    x: str                  # column (or derived column) on the x axis
    y: str                  # duration column
    series: List[str]       # columns that tell the lines or bar groups apart
    kind: str               # "line" or "bar"
    xlabel: str
    rate: Optional[str] = None   # "keys": x / y; "keys2": 2 x / y; "bytes2": 2 num_bytes / y
    xlog: bool = False


def _sublists(df):
    parts = df["sorted_sublist_lengths"].astype(str).str.split(",")
    df["sublists"] = parts.str.len()
    df["total_keys"] = parts.apply(lambda p: sum(int(v) for v in p))
    return df


def _blocks(df):
    df["blocks"] = (df["num_bytes"] / df["block_size"]).round().astype(int)
    return df


SORT_SERIES = ["algorithm", "gpus"]
PLOTS: Dict[str, Plot] = {
    "num_elements_to_gowanlock_sort_duration_for_gpus": Plot("num_elements", "sort_duration", SORT_SERIES, "line",
                                                             "keys", "keys"),
    "num_elements_to_tanasic_sort_duration_for_gpus": Plot("num_elements", "sort_duration", SORT_SERIES, "line",
                                                           "keys", "keys"),
    "sort_algorithm_to_sort_duration": Plot("num_elements", "sort_duration", ["algorithm"], "bar", "keys", "keys"),
    "data_type_to_gowanlock_sort_duration": Plot("data_type", "sort_duration", ["num_elements"], "bar", "data type"),
    "data_type_to_tanasic_sort_duration": Plot("data_type", "sort_duration", ["num_elements"], "bar", "data type"),
    "distribution_type_to_sort_duration_for_algorithm": Plot("distribution_type", "sort_duration", ["algorithm"],
                                                             "bar", "distribution"),
    "num_elements_to_sort_duration_profilers_nsys": Plot("num_elements", "sort_duration", ["algorithm"], "bar",
                                                         "keys", "keys"),
    "gpus_to_tanasic_sort_duration_profilers_nsys": Plot("gpus", "sort_duration", ["algorithm"], "bar", "GPUs"),
    "gpus_to_gowanlock_sort_duration_profilers_nsys": Plot("gpus", "sort_duration", ["algorithm"], "bar", "GPUs"),
    "num_bytes_to_h_to_d_total_duration_and_d_to_h_total_duration": Plot("direction", "duration", ["execution_type"],
                                                                          "bar", "direction"),
    "gpus_to_h_to_d_total_duration_and_d_to_h_total_duration": Plot("direction", "duration", ["execution_type"],
                                                                     "bar", "direction"),
    "num_gpus_to_memory_allocation_total_duration": Plot("num_gpus", "memory_allocation_total_duration",
                                                         ["num_bytes"], "bar", "GPUs"),
    "local_gpu_sort_algorithm_to_sort_duration": Plot("num_elements", "local_sort_duration", ["sort_algorithm"],
                                                      "line", "keys", "keys"),
    "local_gpu_merge_algorithm_to_merge_duration": Plot("num_elements", "local_merge_duration", ["merge_algorithm"],
                                                        "line", "keys", "keys"),
    "local_cpu_sort_algorithm_to_sort_duration": Plot("num_elements", "local_sort_duration", ["sort_algorithm"],
                                                      "line", "keys", "keys"),
    "local_cpu_merge_algorithm_to_merge_duration": Plot("num_elements", "local_merge_duration", ["merge_algorithm"],
                                                        "line", "keys per input", "keys2"),
    "number_sorted_sublists_to_multiway_merge_duration": Plot("sublists", "merge_duration", ["total_keys"], "line",
                                                              "sorted sublists"),
    "number_sorted_sublists_to_gpu_multiway_merge_duration": Plot("sublists", "merge_duration", ["total_keys"],
                                                                  "line", "sorted sublists"),
    "large_data_to_gowanlock_sort_duration_for_gpus": Plot("num_elements", "sort_duration",
                                                           ["chunk_size", "merge_group_size", "num_buffers"], "line",
                                                           "keys", "keys"),
    "large_data_to_gowanlock_vs_cpu_sort_duration_for_gpus": Plot("num_elements", "sort_duration",
                                                                  ["algorithm", "chunk_size"], "line", "keys",
                                                                  "keys"),
    "gpus_to_sort_duration_profilers_nsys": Plot("distribution_type", "sort_duration",
                                                 ["algorithm", "num_elements", "chunk_size", "num_buffers"], "bar",
                                                 "distribution"),
    "block_size_to_inplace_data_transfer_duration": Plot("blocks", "inplace_data_transfer_duration", ["num_bytes"],
                                                         "line", "blocks per transfer", "bytes2", xlog=True),
}

DERIVE: Dict[str, Callable] = {
    "number_sorted_sublists_to_multiway_merge_duration": _sublists,
    "number_sorted_sublists_to_gpu_multiway_merge_duration": _sublists,
    "block_size_to_inplace_data_transfer_duration": _blocks,
}


def _transfers(df):
    """One row per direction (the transfer CSV has one column per direction)."""
    rows = []
    for _, r in df.iterrows():
        for name, col, factor in (("host to device", "h_to_d_total_duration", 1),
                                  ("device to host", "d_to_h_total_duration", 1),
                                  ("both at once", "total_bidirectional_duration", 2)):
            rows.append({"execution_type": r["execution_type"], "direction": name, "duration": r[col],
                         "bytes_moved": r["num_bytes"] * factor})
    return pandas.DataFrame(rows)


def aggregate(ident: str, df: pandas.DataFrame) -> pandas.DataFrame:
    """Mean, minimum and maximum of the repetitions per configuration, plus the
    throughput column where the plot has one."""
    p = PLOTS[ident]
    if "h_to_d_total_duration" in df.columns:
        df = _transfers(df)
    if ident in DERIVE:
        df = DERIVE[ident](df)
    keys = [c for c in p.series if c != p.x] + [p.x]
    extra = ["bytes_moved"] if "bytes_moved" in df.columns else []
    g = df.groupby(keys + extra, as_index=False)[p.y].agg(["mean", "min", "max", "count"])
    g = g.reset_index() if keys[0] not in g.columns else g
    g = g.sort_values(keys + extra).reset_index(drop=True)
    if p.rate == "keys":
        g["gkeys_per_s"] = g[p.x] / g["mean"] / 1e9
    elif p.rate == "keys2":  # two inputs of x keys each
        g["gkeys_per_s"] = 2 * g[p.x] / g["mean"] / 1e9
    elif p.rate == "bytes2":  # the in-place exchange moves num_bytes each way
        g["gb_per_s"] = 2 * g["num_bytes"] / g["mean"] / 1e9
    if extra:
        g["gb_per_s"] = g["bytes_moved"] / g["mean"] / 1e9
    return g


def draw(ident: str, g: pandas.DataFrame, out: pathlib.Path, fmt: str) -> pathlib.Path:
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    p = PLOTS[ident]
    fig, ax = plt.subplots(figsize=(7, 4))
    series_cols = [c for c in p.series if c != p.x]
    groups = list(g.groupby(series_cols)) if series_cols else [((), g)]
    if p.kind == "bar":
        xs = list(dict.fromkeys(g[p.x].tolist()))
        width = 0.8 / max(1, len(groups))
        for k, (name, sub) in enumerate(groups):
            heights = [sub[sub[p.x] == x]["mean"].mean() if (sub[p.x] == x).any() else 0 for x in xs]
            ax.bar([i + k * width for i in range(len(xs))], heights, width, label=_label(series_cols, name))
        ax.set_xticks([i + width * (len(groups) - 1) / 2 for i in range(len(xs))])
        ax.set_xticklabels([str(x) for x in xs], rotation=30, ha="right")
    else:
        for name, sub in groups:
            sub = sub.sort_values(p.x)
            mean = sub["mean"].to_numpy()
            yerr = numpy.vstack([mean - sub["min"].to_numpy(), sub["max"].to_numpy() - mean])
            ax.errorbar(sub[p.x].to_numpy(), mean, yerr=yerr, marker="o", capsize=3,
                        label=_label(series_cols, name))
        if p.xlog:
            ax.set_xscale("log")
    ax.set_xlabel(p.xlabel)
    ax.set_ylabel("duration [s]")
    ax.set_title(ident, fontsize=8)
    ax.legend(fontsize=7)
    fig.tight_layout()
    path = out / ("%s.%s" % (ident, fmt))
    fig.savefig(path)
    plt.close(fig)
    return path


def _label(cols, name) -> str:
    if not cols:
        return "all"
    name = name if isinstance(name, tuple) else (name,)
    return ", ".join("%s=%s" % (c, v) for c, v in zip(cols, name))


def table(ident: str, g: pandas.DataFrame) -> str:
    cols = [c for c in g.columns if c != "bytes_moved"]
    lines = ["| " + " | ".join(cols) + " |", "|" + "---|" * len(cols)]
    for _, r in g.iterrows():
        cells = []
        for c in cols:
            v = r[c]
            cells.append("%.6g" % v if isinstance(v, float) else str(v))
        lines.append("| " + " | ".join(cells) + " |")
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="plot and tabulate an experiment run")
    ap.add_argument("run", nargs="?", default=None, help="run directory (default: the latest under results/)")
    ap.add_argument("--format", default="png", choices=["png", "pdf", "svg"])
    a = ap.parse_args(argv)
    if a.run is None:
        runs = sorted((HERE / "results").glob("*")) if (HERE / "results").exists() else []
        if not runs:
            print("no run directory", file=sys.stderr)
            return 2
        run = runs[-1]
    else:
        run = pathlib.Path(a.run)
    summary = ["# Experiment results (%s)" % run.name, ""]
    n = 0
    for x in settings.experiments():
        path = run / ("%s_%s.csv" % (x.executable, x.identifier))
        if not path.exists():
            continue
        df = pandas.read_csv(path)
        g = aggregate(x.identifier, df)
        img = draw(x.identifier, g, run, a.format)
        summary += ["## %s" % x.identifier, "", "`%s`, %d runs; plot: %s" % (x.executable, len(df), img.name)]
        if x.notes:
            summary += ["", x.notes]
        summary += ["", table(x.identifier, g), ""]
        n += 1
    (run / "summary.md").write_text("\n".join(summary) + "\n")
    print("%d experiments plotted; %s" % (n, run / "summary.md"))
    return 0 if n else 1


if __name__ == "__main__":
    sys.exit(main())
