"""Makes one misuse of the host API that the library must refuse before any
kernel is launched or any state is touched (argv[1] names the case), then
reports "launched" if the call returned.  Two cases instead let a device wait
time out (a 20 ms budget on a flag nobody sets, so the kernel ends normally)
and then call nvshmem_barrier_all or nvstop_pes, which must end the process.  Launched by test_gpu_rma.py in its own process, since a refused
call ends the process with an error message."""
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "df-nvshmem-prototype_amd"))

import nvshmem_amd as na  # noqa: E402


def main(case):
    lib, hip = na.load(), na.Hip()
    if case == "too_many_pes":
        lib.nvshmemx_start_local_pes(129, 1)
        print("launched", flush=True)
        return
    assert lib.nvshmemx_start_local_pes(2, 1) == 0
    if case == "heap_exhausted":
        lib.nvshmalloc(int(lib.nvshmemx_heap_size()) + 1)
        print("launched", flush=True)
        return
    sym = lib.nvshmalloc(1 << 20)
    private = hip.malloc(1 << 20)
    print("case", case, flush=True)
    if case in ("wait_timeout_barrier", "wait_timeout_stop"):
        # a while-equal wait on a flag that stays 0 gives up after 20 ms; the
        # next synchronising host call must end the process
        lib.nvshmemx_set_wait_timeout_ms(20)
        flag = lib.nvshmem_ptr(sym, 0)
        hip.memset(flag, 0, 64)
        lib.nvshmemh_wait_once(flag, 0, 0, private, None)
        hip.sync()
        print("wait returned", flush=True)
        if case == "wait_timeout_barrier":
            lib.nvshmem_barrier_all()
        else:
            lib.nvstop_pes()
        print("launched", flush=True)
        return
    if case == "get_private_source":
        lib.nvshmemx_getmem_on_stream(private, private, 4096, 1, None)
    elif case == "put_private_target":
        lib.nvshmemx_putmem_on_stream(private, sym, 4096, 1, None)
    elif case == "put_past_heap_end":
        lib.nvshmemx_putmem_on_stream(sym, private, int(lib.nvshmemx_heap_size()) + (4 << 20), 1, None)
    elif case == "put_into_slot_padding":
        # local PEs share one allocation in power-of-two slots; past the heap's
        # end, inside the slot's padding, is not symmetric memory
        lib.nvshmemx_putmem_on_stream(sym + int(lib.nvshmemx_heap_size()) + (8 << 20), private, 4096, 1, None)
    elif case == "symmetric_ok":
        lib.nvshmemx_putmem_on_stream(sym, private, 4096, 1, None)
    hip.sync()
    print("launched", flush=True)
    lib.nvstop_pes()


if __name__ == "__main__":
    main(sys.argv[1])
