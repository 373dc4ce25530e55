"""Streaming-copy bandwidth probe on one GPU (the achievable HBM ceiling).

    python -m sgm_amd.bwprobe [--mib 512] [--reps 20]

Copies a buffer with the engine's 16-byte copy kernel (sgmh_copy_probe) and
prints one JSON line with GB/s counted as read + write bytes.  Under
`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` its dispatches are the known-size
reference used to scale the counters of the SGM kernels.
"""
import argparse
import json

from . import Context


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=512)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    import torch
    n = a.mib * 1024 * 1024
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    ctx = Context(0)
    s = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        ctx.copy_probe(dst.data_ptr(), src.data_ptr(), n, s)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.reps):
        ctx.copy_probe(dst.data_ptr(), src.data_ptr(), n, s)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.reps
    print(json.dumps({"probe": "copy16", "bytes_per_launch": n, "ms": round(ms, 4),
                      "GBs_read_plus_write": round(2 * n / (ms / 1e3) / 1e9, 1)}))
    ctx.close()


if __name__ == "__main__":
    main()
