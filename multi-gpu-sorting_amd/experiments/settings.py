"""The experiment grid: the reference's experiments (scripts/settings.py:25-787)
restated for one MI355X.

Every experiment keeps the reference's identifier, program name, argument
order and CSV columns, so that its CSV files line up with the reference's
side by side.  What changes:

* GPU lists are cut to one device ("0"); experiments that the reference only
  runs on GPU pairs or larger sets (data types, distributions) run on the one
  GPU instead.  Experiments that need a peer (the P2P transfer studies) are
  listed in OMITTED with the reason, and nothing runs for them.
* The sort_benchmark CSV carries seven added columns after the reference's
  fourteen (throughput, algorithmic bytes, HBM rate, roofline fraction,
  merge time, host threads, verified).
* Legs of this build's own are added beside the reference's: the in-HBM
  "hip" sort, mgs_sort / mgs_sort_lsd and the merges in the local GPU
  programs, the device k-way merge.
* Profiled experiments use rocprofv3 (kernel trace and stats) where the
  reference used nsys/nvprof.

Sizes are the reference's; run_experiments.py scales them (--scale) and can
skip runs above a key count (--max-elements), since its largest host sorts
(6.4e10 keys) do not fit one box's host memory.
"""
import itertools
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

SEED = 2147483647  # the generator's default seed (src/argument_limits.cuh)

# sort_benchmark's CSV: the reference's columns (scripts/settings.py:42-46),
# then the ones this build adds (host/main/sort_benchmark.cpp).
SORT_COLUMNS = [
    "num_elements", "algorithm", "gpus", "data_type", "distribution_type", "distribution_seed", "num_threads",
    "chunk_size", "merge_group_size", "num_buffers", "memory_allocation_duration", "sort_duration",
    "memory_deallocation_duration", "total_duration",
    "gkeys_per_s", "alg_bytes", "hbm_gbps_alg", "roofline_frac", "merge_duration", "cpu_threads", "verified",
]
SORT_PARAMS = ["", "--algorithm", "--gpus", "--data_type", "--distribution_type", "--distribution_seed"]
LARGE_PARAMS = ["", "--algorithm", "--gpus", "--chunk_size", "--merge_group_size", "--num_buffers"]
PROFILED_PARAMS = ["", "--algorithm", "--gpus", "--data_type", "--distribution_type", "--chunk_size",
                   "--merge_group_size", "--num_buffers"]

TRANSFER_COLUMNS = [
    "num_bytes", "gpus", "num_repetitions", "execution_type", "h_to_d_durations", "h_to_d_total_duration",
    "p_to_p_durations", "p_to_p_total_duration", "d_to_h_durations", "d_to_h_total_duration",
    "bidirectional_durations", "total_bidirectional_duration",
]

REFERENCE_DISTRIBUTIONS = ["uniform", "normal", "zero", "staggered", "sorted", "reverse-sorted", "nearly-sorted",
                           "bucket-sorted"]

# Arguments that are key or byte counts (what --scale multiplies): the first
# argument of every program but the sublist ones, --chunk_size, and the
# in-place transfer's block size.
EXTRA_SIZE_ARGS = {"inplace_memcpy_benchmark": (2,)}
# Programs whose first argument is a list of sublist lengths.
LIST_ARGS = {"local_cpu_multiway_merge_benchmark": 0, "local_gpu_multiway_merge_benchmark": 0}
# Programs that run on the host only (the reference binds them to a NUMA
# node's memory but not its cores: scripts/run_experiments.py:36).
CPU_PROGRAMS = {"local_cpu_sort_algorithm_benchmark", "local_cpu_merge_algorithm_benchmark",
                "local_cpu_multiway_merge_benchmark"}
CPU_SORT_ALGORITHMS = {"gnu_parallel", "paradis", "std"}


@dataclass
class Experiment:
    identifier: str
    executable: str
    parameters: List[str]
    arguments: List[Tuple]
    columns: List[str]
    repetitions: int = 10  # scripts/settings.py:14
    profiler: Optional[str] = None
    notes: str = ""


def grid(*axes: Sequence) -> List[Tuple]:
    return list(itertools.product(*axes))


def _sort(identifier, arguments, notes=""):
    return Experiment(identifier, "sort_benchmark", SORT_PARAMS, arguments, SORT_COLUMNS, notes=notes)


def build_experiments() -> List[Experiment]:
    e: List[Experiment] = []
    one = ["0"]

    # scripts/settings.py:52-106 / :108-162 (the one-GPU rows)
    e.append(_sort("num_elements_to_gowanlock_sort_duration_for_gpus",
                   grid([1000, 10 ** 9, 2 * 10 ** 9, 4 * 10 ** 9], ["gowanlock"], one, ["int"], ["uniform"], [SEED])))
    e.append(_sort("num_elements_to_tanasic_sort_duration_for_gpus",
                   grid([1000, 10 ** 9, 2 * 10 ** 9, 4 * 10 ** 9], ["tanasic"], one, ["int"], ["uniform"], [SEED])))

    # scripts/settings.py:164-196, plus the in-HBM sort ("hip")
    sizes = [10 ** 9, 2 * 10 ** 9, 3 * 10 ** 9, 4 * 10 ** 9, 8 * 10 ** 9]
    e.append(_sort("sort_algorithm_to_sort_duration",
                   grid(sizes, ["gnu_parallel", "paradis", "thrust", "tanasic", "gowanlock"], one, ["int"],
                        ["uniform"], [SEED]) +
                   grid([10 ** 9, 2 * 10 ** 9, 3 * 10 ** 9, 4 * 10 ** 9], ["hip"], one, ["int"], ["uniform"],
                        [SEED]),
                   notes="added leg: hip (keys already in HBM, copies outside the timed phase)"))

    # scripts/settings.py:198-254: the reference runs these on 2 to 8 GPUs
    for alg in ("gowanlock", "tanasic"):
        e.append(_sort("data_type_to_%s_sort_duration" % alg,
                       grid([4 * 10 ** 9], [alg], one, ["int", "float"], ["uniform"], [SEED]) +
                       grid([2 * 10 ** 9], [alg], one, ["long", "double"], ["uniform"], [SEED]),
                       notes="one GPU (the reference: 2-8 GPUs)"))

    # scripts/settings.py:256-275
    e.append(_sort("distribution_type_to_sort_duration_for_algorithm",
                   grid([2 * 10 ** 9], ["gowanlock", "tanasic"], one, ["int"], REFERENCE_DISTRIBUTIONS, [SEED]),
                   notes="one GPU (the reference: 2-8 GPUs)"))

    # scripts/settings.py:277-344: profiled runs (one repetition)
    for ident, args in (
            ("num_elements_to_sort_duration_profilers_nsys",
             grid([10 ** 9, 2 * 10 ** 9, 4 * 10 ** 9, 8 * 10 ** 9, 16 * 10 ** 9], ["gowanlock", "tanasic"], one,
                  ["int"], ["uniform"], [SEED])),
            ("gpus_to_tanasic_sort_duration_profilers_nsys",
             grid([2 * 10 ** 9], ["tanasic"], one, ["int"], ["uniform"], [SEED])),
            ("gpus_to_gowanlock_sort_duration_profilers_nsys",
             grid([2 * 10 ** 9], ["gowanlock"], one, ["int"], ["uniform"], [SEED]))):
        x = _sort(ident, args, notes="rocprofv3 --kernel-trace --stats in place of nsys")
        x.repetitions, x.profiler = 1, "rocprofv3"
        e.append(x)

    # scripts/settings.py:346-430: host-link part only (one GPU, serial)
    e.append(Experiment("num_bytes_to_h_to_d_total_duration_and_d_to_h_total_duration", "data_transfer_benchmark",
                        ["", "", "", ""], grid([4 * 10 ** 9], one, [5], ["serial"]), TRANSFER_COLUMNS))
    e.append(Experiment("gpus_to_h_to_d_total_duration_and_d_to_h_total_duration", "data_transfer_benchmark",
                        ["", "", "", ""], grid([4 * 10 ** 9], one, [5], ["parallel", "serial"]), TRANSFER_COLUMNS))

    # scripts/settings.py:436-455
    e.append(Experiment("num_gpus_to_memory_allocation_total_duration", "memory_benchmark", ["", "", ""],
                        grid([8 * 10 ** 9], [1], ["parallel"]),
                        ["num_bytes", "num_gpus", "execution_mode", "memory_allocation_durations",
                         "memory_allocation_total_duration"]))

    # scripts/settings.py:461-478, plus this build's sorts
    e.append(Experiment("local_gpu_sort_algorithm_to_sort_duration", "local_gpu_sort_algorithm_benchmark",
                        ["", "", ""],
                        grid([10 ** 9, 2 * 10 ** 9, 3 * 10 ** 9, 4 * 10 ** 9],
                             ["mgs_sort", "mgs_sort_lsd", "hipcub::DeviceRadixSort::SortKeys", "thrust::sort",
                              "hipcub::DeviceMergeSort::SortKeys"], [0]),
                        ["num_elements", "sort_algorithm", "gpu_id", "local_sort_duration"],
                        notes="ROCm libraries in place of CUB, Thrust and moderngpu"))

    # scripts/settings.py:484-507
    e.append(Experiment("local_gpu_merge_algorithm_to_merge_duration", "local_gpu_merge_algorithm_benchmark",
                        ["", "", ""],
                        grid([k * 10 ** 8 for k in range(10, 42)], ["mgs_merge", "mgs_multiway_merge",
                                                                   "thrust::merge"], [0]),
                        ["num_elements", "merge_algorithm", "gpu_id", "local_merge_duration"],
                        notes="mgs_merge in place of mgpu::merge, rocThrust for thrust::merge"))

    # scripts/settings.py:512-530
    e.append(Experiment("local_cpu_sort_algorithm_to_sort_duration", "local_cpu_sort_algorithm_benchmark",
                        ["", ""],
                        grid([10 ** 9, 2 * 10 ** 9, 4 * 10 ** 9, 8 * 10 ** 9, 16 * 10 ** 9, 32 * 10 ** 9,
                              64 * 10 ** 9], ["gnu_parallel::sort", "paradis::sort"]),
                        ["num_elements", "sort_algorithm", "num_threads", "local_sort_duration"],
                        notes="boost::sort::block_indirect_sort left out: Boost is not part of this build"))

    # scripts/settings.py:534-552
    e.append(Experiment("local_cpu_merge_algorithm_to_merge_duration", "local_cpu_merge_algorithm_benchmark",
                        ["", ""],
                        grid([10 ** 9, 2 * 10 ** 9, 4 * 10 ** 9, 8 * 10 ** 9, 16 * 10 ** 9, 32 * 10 ** 9,
                              64 * 10 ** 9], ["gnu_parallel::merge", "gnu_parallel::multiway_merge"]),
                        ["num_elements", "merge_algorithm", "num_threads", "local_merge_duration"]))

    # scripts/settings.py:556-588, and its device counterpart
    lists = []
    for total in (2 * 10 ** 9, 8 * 10 ** 9, 32 * 10 ** 9):
        for k in (2, 3, 4, 5, 6, 7, 8, 12, 16):
            lists.append((",".join([str(total // k)] * k),))
    e.append(Experiment("number_sorted_sublists_to_multiway_merge_duration", "local_cpu_multiway_merge_benchmark",
                        [""], lists, ["sorted_sublist_lengths", "num_threads", "merge_duration"]))
    e.append(Experiment("number_sorted_sublists_to_gpu_multiway_merge_duration", "local_gpu_multiway_merge_benchmark",
                        ["", ""], [(x[0], 0) for x in lists if sum(map(int, x[0].split(","))) <= 8 * 10 ** 9],
                        ["sorted_sublist_lengths", "gpu_id", "merge_duration"],
                        notes="added: mgs_multiway_merge of device-resident sublists"))

    # scripts/settings.py:593-665 (one GPU)
    big = [1000, 10 ** 9, 2 * 10 ** 9, 4 * 10 ** 9, 8 * 10 ** 9, 12 * 10 ** 9, 16 * 10 ** 9, 20 * 10 ** 9,
           24 * 10 ** 9, 32 * 10 ** 9, 40 * 10 ** 9, 48 * 10 ** 9, 56 * 10 ** 9, 60 * 10 ** 9, 64 * 10 ** 9]
    e.append(Experiment("large_data_to_gowanlock_sort_duration_for_gpus", "sort_benchmark", LARGE_PARAMS,
                        grid(big, ["gowanlock"], one, [3 * 10 ** 9, 4125 * 10 ** 6], [0, 1], [2]) +
                        grid(big, ["gowanlock"], one, [2 * 10 ** 9, 2750 * 10 ** 6], [0, 1], [3]),
                        SORT_COLUMNS))
    e.append(Experiment("large_data_to_gowanlock_vs_cpu_sort_duration_for_gpus", "sort_benchmark", LARGE_PARAMS,
                        grid(big, ["gowanlock"], one, [3 * 10 ** 9, 4125 * 10 ** 6], [1], [2]) +
                        grid(big, ["gnu_parallel", "paradis"], one, [0], [1], [2]),
                        SORT_COLUMNS))

    # scripts/settings.py:671-757 (one GPU)
    prof = (grid([4 * 10 ** 9], ["gowanlock", "tanasic"], one, ["int"], ["uniform"], [0], [1], [2]) +
            grid([2 * 10 ** 9], ["gowanlock", "tanasic"], one, ["int"], REFERENCE_DISTRIBUTIONS, [0], [1], [2]) +
            grid([32 * 10 ** 9, 60 * 10 ** 9], ["gowanlock"], one, ["int"], ["uniform"],
                 [3 * 10 ** 9, 4125 * 10 ** 6], [1, 0], [2]) +
            grid([32 * 10 ** 9, 60 * 10 ** 9], ["gowanlock"], one, ["int"], ["uniform"],
                 [2 * 10 ** 9, 2750 * 10 ** 6], [1, 0], [3]))
    x = Experiment("gpus_to_sort_duration_profilers_nsys", "sort_benchmark", PROFILED_PARAMS, prof, SORT_COLUMNS,
                   notes="rocprofv3 --kernel-trace --stats in place of nsys")
    x.repetitions, x.profiler = 1, "rocprofv3"
    e.append(x)

    # scripts/settings.py:762-787 (block sizes b/i as the reference writes them)
    args = []
    for b in (10 ** 9, 2 * 10 ** 9, 4 * 10 ** 9, 8 * 10 ** 9):
        for i in (1, 10, 100, 1000, 10000, 100000, 2, 4, 8, 16, 32):
            args.append((b, 0, b // i))
    e.append(Experiment("block_size_to_inplace_data_transfer_duration", "inplace_memcpy_benchmark", ["", "", ""],
                        args, ["num_bytes", "gpu_id", "block_size", "inplace_data_transfer_duration"]))
    return e


# Reference experiments with no one-GPU form.
OMITTED = {
    "num_bytes_to_p_to_p_total_duration": "peer-to-peer copies need a second GPU",
    "gpus_to_p_to_p_total_duration": "peer-to-peer copies need a second GPU",
}


def size_positions(x: Experiment) -> List[int]:
    """Positions of the arguments of x that are key or byte counts."""
    if x.executable in LIST_ARGS:
        return []
    pos = [0] + [i for i, p in enumerate(x.parameters) if p == "--chunk_size"]
    return pos + list(EXTRA_SIZE_ARGS.get(x.executable, ()))


def experiments() -> List[Experiment]:
    return build_experiments()
