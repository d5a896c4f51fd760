"""Run the autotuner (cmbl_tune) over map geometries and write the tuning
table that cmbl_map_create reads (tuning_<arch>.txt next to the library).

Usage: python -m cosmopp_amd.tune [--reps R] [--out PATH] [NSIDE:LMAX ...]
Default geometries: the benchmark workloads (nside 64 / lmax 128, 512 / 1024,
1024 / 2048, 2048 / 4096, the chain maps 256 / 512 and the batch classes
128 / 256, 512 / 768, 1024 / 1536).  One JSON line per geometry with the
timed stage means.
"""
import argparse
import json
import os
import sys

DEFAULT = ["64:128", "128:256", "256:512", "512:768", "512:1024", "1024:1536", "1024:2048", "2048:4096"]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--out", default=None, help="table file (default: tuning_<arch>.txt beside the library)")
    ap.add_argument("geoms", nargs="*", default=DEFAULT)
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import cosmopp_amd as cm

    labels = ["synth_k1", "synth_k2", "synth_k4", "synth_k8", "adj_k1", "adj_k2", "adj_k4", "adj_k8",
              "ring_split2048", "ring_split4096", "ring_split8192"]
    for g in args.geoms:
        nside, lmax = (int(v) for v in g.split(":"))
        t, times = cm.tune(nside, lmax, args.reps)
        print(json.dumps({"nside": nside, "lmax": lmax, "choice": t.as_tuple(),
                          "ms": {k: round(v, 4) for k, v in zip(labels, times) if v >= 0}}), flush=True)
    out = args.out
    if out is None:
        import torch

        arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
        out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_%s.txt" % arch)
    cm.tuning_save(out)
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main()
