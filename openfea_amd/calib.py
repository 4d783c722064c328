"""HBM calibration: python -m openfea_amd.calib [MiB] [reps]

Runs the library's plain copy kernel (ofea_stream_copy) and prints the
achieved copy bandwidth (read + write bytes per second).  Under
`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` the same run gives the counters'
reading for a known byte count (scale factor for kernel traffic figures).
"""
import ctypes as C
import json
import sys

from . import _capi
from .fea import Context


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    nbytes = mib << 20
    with Context(0) as ctx:
        ms = C.c_float()
        _capi.check(_capi.lib().ofea_stream_copy(ctx._h, nbytes, reps, C.byref(ms)))
    gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
    print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms.value, "read_plus_write_GBps": gbs,
                      "fetch_bytes_per_copy": nbytes, "write_bytes_per_copy": nbytes}))


if __name__ == "__main__":
    main()
