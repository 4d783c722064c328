"""HBM calibration: python -m openfea_amd.calib [MiB] [reps]
                   python -m openfea_amd.calib fill [MiB] [reps]

Runs the library's plain copy kernel (ofea_stream_copy) and prints the
achieved copy bandwidth (read + write bytes per second).  "fill" instead
times a write-only fill of the same size with torch (fill_ on the current
stream), the write ceiling that write-bound kernels are compared with; it
runs in a process of its own because torch's HIP runtime and the library's
are not meant to share a process.  Under
`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` the same run gives the counters'
reading for a known byte count (scale factor for kernel traffic figures).
"""
import ctypes as C
import json
import sys

from . import _capi
from .fea import Context


def fill(mib, reps):
    import torch

    nbytes = mib << 20
    buf = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    for _ in range(3):
        buf.fill_(1.0)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        buf.fill_(1.0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"fill_bytes": nbytes, "ms_per_fill": ms, "write_GBps": nbytes / (ms * 1e-3) / 1e9}))


def main():
    args = sys.argv[1:]
    if args and args[0] == "fill":
        fill(int(args[1]) if len(args) > 1 else 1024, int(args[2]) if len(args) > 2 else 20)
        return
    mib = int(args[0]) if args else 1024
    reps = int(args[1]) if len(args) > 1 else 20
    nbytes = mib << 20
    with Context(0) as ctx:
        ms = C.c_float()
        _capi.check(_capi.lib().ofea_stream_copy(ctx._h, nbytes, reps, C.byref(ms)))
    gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
    print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms.value, "read_plus_write_GBps": gbs,
                      "fetch_bytes_per_copy": nbytes, "write_bytes_per_copy": nbytes}))


if __name__ == "__main__":
    main()
