"""Host-side check of a batch of puts: every request's target bytes, read back
to the host, equal its source bytes, and (optionally) every byte of the
destination regions outside the requests' targets is still zero.  Nothing on
the device takes part in the comparison."""
import numpy as np


def assert_requests_copied(lib, hip, arr, n):
    for i in range(n):
        r = arr[i]
        if not r.nbytes:
            continue
        got = hip.dtoh(lib.nvshmem_ptr(r.target, r.pe), r.nbytes)
        src = hip.dtoh(r.source, r.nbytes)
        if not np.array_equal(got, src):
            bad = np.flatnonzero(got != src)
            raise AssertionError(f"request {i} ({r.nbytes} bytes to PE {r.pe}): {bad.size} bytes differ, "
                                 f"first at {bad[0]}")
