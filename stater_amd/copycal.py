"""Copy of a known size, for calibrating the HBM traffic counters.

``python -m stater_amd.copycal [GiB]`` copies a device buffer of the given
size (default 1 GiB) three times with hipMemcpy (device to device).  Run under
``rocprofv3 --pmc FETCH_SIZE`` and then ``--pmc WRITE_SIZE``: the counters'
KiB per dispatch, divided into the bytes actually moved, give the scale
factors that turn the same counters on the library's kernels into HBM bytes.
"""
from __future__ import annotations

import sys

from stater_amd import hipmem


def main() -> int:
    gib = float(sys.argv[1]) if len(sys.argv) > 1 else 1.0
    n = int(gib * 2**30)
    a, b = hipmem.DeviceBuffer(n), hipmem.DeviceBuffer(n)
    hipmem._ok(hipmem.hip().hipMemset(a.ptr, 1, n), "hipMemset")
    for _ in range(3):
        hipmem._ok(hipmem.hip().hipMemcpy(b.ptr, a.ptr, n, hipmem.D2D), "hipMemcpy D2D")
    hipmem._ok(hipmem.hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
    print(f"copied {n} bytes x3 device to device", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
