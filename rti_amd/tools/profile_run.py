"""Targets for rocprofv3 runs (kernel trace or counters).

  python -m rti_amd.tools.profile_run copy  [--mib 1024]
      one device-to-device copy of a buffer of known size (torch's copy
      kernel): the calibration for FETCH_SIZE / WRITE_SIZE;
  python -m rti_amd.tools.profile_run trace [--workload hole] [--rays 1e7] [--launches 2] [--knobs JSON]
      a benchmark workload (default: the triangle hole, six (s, n) cases;
      --rays per case for the hole, in total otherwise) traced as one
      persistent launch per call, inputs resident; --knobs sets kernel
      parameters (e.g. '{"bvh_width": 2, "lds_stack": 16}').
Prints the work done so a summary can turn counters into bytes per segment.
"""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "trace"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--rays", type=float, default=1e7)
    ap.add_argument("--launches", type=int, default=2)
    ap.add_argument("--knobs", default="{}")
    ap.add_argument("--workload", default="hole")
    a = ap.parse_args()
    if a.what == "copy":
        import torch
        n = a.mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"what": "copy", "bytes_read": n * 4, "bytes_written": n * 4}))
        return
    from bench import workload
    from rti_amd import Context
    cases, _, _ = workload(a.workload, a.rays)
    ctx = Context(0)
    ctx.batch_set(cases)
    ctx.set_params(**json.loads(a.knobs))
    ctx.prepare()
    segs = []
    for _ in range(a.launches):
        ctx.batch_run()
        segs.append(int(ctx.counters()[0]))
    print(json.dumps({"what": "trace", "workload": a.workload, "rays_per_launch": int(sum(c["n_rays"] for c in cases)),
                      "segments_per_launch": segs}))


if __name__ == "__main__":
    main()
