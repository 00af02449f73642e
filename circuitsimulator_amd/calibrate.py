"""Copy-kernel calibration for HBM traffic counters.

Runs one device-to-device copy of a known size through a plain elementwise
copy kernel (torch's clone of a contiguous tensor), so that the FETCH_SIZE and
WRITE_SIZE counters of `rocprofv3 --pmc` can be related to bytes actually
moved: factor = counter_KiB * 1024 / bytes_copied.

    rocprofv3 --pmc FETCH_SIZE -d <dir> --output-format csv -- \
        python -m circuitsimulator_amd.calibrate --mib 1024
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    a = ap.parse_args()
    n = a.mib * (1 << 20) // 4
    x = torch.arange(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    y = x.clone()
    torch.cuda.synchronize()
    assert int(y[-1].item()) == n - 1
    print(f"copied {a.mib} MiB ({n * 4} bytes read, {n * 4} bytes written)", flush=True)


if __name__ == "__main__":
    main()
