"""Autotuner driver: runs the library's turtle_autotune for every key width,
index flag and size bucket, and turtle_autotune_segmented on a batch of
uneven segments (20,000 lengths log-uniform in [16, 262144], the batched
workload's shape) for every key width and index flag, and writes the tuning
table the library reads at run time (tuning_table.txt next to
libturtle_hip.so).

    python -m turtlesort_amd.autotune [--sizes 16,20,23,26] [--reps 3] [--no-batched] [--out PATH]

Each line: kb idx log2n splitters lds_cutoff tiles_per_task workgroups_per_cu ms;
batch lines start with "seg" and their log2n is that of the longest segment.
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DTYPES = {1: np.uint8, 2: np.int16, 4: np.int32, 8: np.int64}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="16,20,23,26", help="log2 of the key counts")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true", help="skip the segmented-batch entries")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning_table.txt"))
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    import torch
    from turtlesort_amd import TurtleSort
    assert torch.cuda.is_available(), "the autotuner needs the GPU"
    torch.cuda.set_device(0)
    name = torch.cuda.get_device_properties(0).gcnArchName
    lines = [f"# autotuned on {name}: kb idx log2n splitters lds_cutoff tiles_per_task workgroups_per_cu ms",
             "# (synthetic keys uniform over the key type, iota index; turtlesort_amd/autotune.py)"]
    for lg in [int(x) for x in args.sizes.split(",")]:
        n = 1 << lg
        for kb, dt in DTYPES.items():
            s = TurtleSort(dt, n, track_index=True)
            for idx in (0, 1):
                t, ms = s.autotune(n, with_index=bool(idx), reps=args.reps)
                line = (f"{kb} {idx} {lg} {t['splitters']} {t['lds_cutoff']} {t['tiles_per_task']} "
                        f"{t['workgroups_per_cu']} {ms:.4f}")
                print(line, flush=True)
                lines.append(line)
            s.close()
    if not args.no_batched:
        from turtlesort_amd import workloads as wl
        lens = wl.loguniform_lengths(wl.SEED_C4_LENGTHS, wl.C4_SEGMENTS, wl.C4_LO, wl.C4_HI)
        off = wl.segment_offsets(lens)
        lg = int(round(np.log2(float(lens.max()))))
        for kb, dt in DTYPES.items():
            s = TurtleSort(dt, int(off[-1]), track_index=True)
            for idx in (0, 1):
                t, ms = s.autotune_segments(off, with_index=bool(idx), reps=args.reps)
                line = (f"seg {kb} {idx} {lg} {t['splitters']} {t['lds_cutoff']} {t['tiles_per_task']} "
                        f"{t['workgroups_per_cu']} {ms:.4f}")
                print(line, flush=True)
                lines.append(line)
            s.close()
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
