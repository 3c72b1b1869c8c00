"""Copy kernel of known size for calibrating rocprofv3 HBM counters.

``rocprofv3 --pmc FETCH_SIZE -- python3 -m conformal_cubed_sphere_amd.hbm_calibration``
(and the same with WRITE_SIZE) reports, per dispatch, what the counters read
for a device-to-device copy of exactly ``--mib`` MiB read and written; the
ratio of counter value to true bytes scales the counters of the stage
kernels to HBM bytes.  Uses torch only as device-memory plumbing.
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--kernel", choices=["copy", "add"], default="copy",
                    help="copy: runtime blit kernel; add: elementwise dst = src + 1")
    args = ap.parse_args()
    n = args.mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.reps):
        if args.kernel == "copy":
            dst.copy_(src)
        else:
            torch.add(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(f"copied {args.mib} MiB x {args.reps}: read {args.mib} MiB and wrote {args.mib} MiB per dispatch",
          flush=True)


if __name__ == "__main__":
    main()
