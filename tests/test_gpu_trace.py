"""Tracing build on the GPU: records only between trace_begin/trace_end, in
issue order with non-decreasing timestamps, written as the reference's CSV
(header nvshmem.cu:390, row format nvshmem.cu:411, file name nvshmem.cu:377);
the host drains the ring while kernels run, so a kernel that emits many
times more records than the ring holds loses none; without the drain thread
a ring too small for the run reports overflow instead of overwriting."""
import json
import os
import re
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
HERE = Path(__file__).resolve().parent
ROW = re.compile(r"^(-?\d+), (\d+), (\d+), (\d+), (\w+), (0x[0-9a-f]+|\(nil\)), (0x[0-9a-f]+|\(nil\)), (-?\d+), "
                 r"(-?\d+), (-?\d+)$")
SEQ = [("put_int", 16), ("put_float", 8), ("get_double", 4), ("put_int", 1), ("get_int", 1), ("wait_int", 1),
       ("wait_until_int", 1)]


def run(tmp_path, reps, extra=None, grid=1, lanes=0, block=64):
    env = dict(os.environ, NVSHMEM_TRACE_DIR=str(tmp_path), NVSHMEM_HEAP_SIZE="16M")
    env.update(extra or {})
    args = [str(reps), str(grid)] + ([str(lanes), str(block)] if lanes else [])
    p = subprocess.run([sys.executable, str(HERE / "trace_worker.py")] + args, env=env,
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-3000:]
    files = sorted(tmp_path.glob("nvshmem_processed_tracefile_*_0"))
    assert len(files) == 1, (list(tmp_path.iterdir()), p.stdout)
    assert not list(tmp_path.glob("nvshmem_tracefile_*")), "spill file not removed"
    return p.stdout, json.loads([l for l in p.stdout.splitlines() if l.startswith("{")][-1]), files[0]


def test_trace_records_and_csv(tmp_path, oracle):
    reps = 5
    stdout, info, f = run(tmp_path, reps)
    assert info["enabled"] == 1 and info["errors"] == 0
    lines = f.read_text().splitlines(keepends=True)
    assert lines[0] == oracle.lib.oracle_trace_header().decode()
    rows = [ROW.match(l.rstrip("\n")) for l in lines[1:]]
    assert all(rows), lines[1:4]
    got = [(m.group(5), int(m.group(8))) for m in rows]
    assert got == SEQ * reps                     # nothing outside begin/end, issue order kept
    ts = [int(m.group(1)) for m in rows]
    assert ts == sorted(ts) and ts[0] > 0
    assert {int(m.group(9)) for m in rows if not m.group(5).startswith("wait")} == {1}   # remote PE
    assert {int(m.group(3)) for m in rows} == {0} and {int(m.group(4)) for m in rows} == {0}
    # each row is exactly the oracle's formatting of its own fields
    for m, line in zip(rows, lines[1:]):
        loc = 0 if m.group(6) == "(nil)" else int(m.group(6), 16)
        rem = 0 if m.group(7) == "(nil)" else int(m.group(7), 16)
        op = [oracle.lib.oracle_trace_op_name(i).decode() for i in range(8)].index(m.group(5))
        assert oracle.trace_row(int(m.group(1)), int(m.group(2)), int(m.group(3)), int(m.group(4)), op, loc, rem,
                                int(m.group(8)), int(m.group(9)), int(m.group(10))) == line
    assert "overflow: 0 conflict: 0" in stdout


RING8 = {"NVSHMEM_LOG_BUFFER_SIZE": "640", "NVSHMEM_LOG_BUFFER_COUNT": "1"}   # 640 B rounds down to 8 records


def test_trace_overflow_is_reported(tmp_path):
    # without the drain thread nothing empties the 8-record ring during the kernel
    stdout, info, f = run(tmp_path, 10, dict(RING8, NVSHMEM_LOG_LIVE_DRAIN="0"))
    assert "overflow: 1" in stdout
    rows = f.read_text().splitlines()[1:]
    assert len(rows) == 8
    assert [r.split(", ")[4] for r in rows] == [s[0] for s in SEQ] + [SEQ[0][0]]


def test_trace_many_workgroups(tmp_path):
    """16 workgroups trace at once: every record is kept (claims on the ring
    head race, none is lost or doubled), each workgroup's records come out in
    its issue order with non-decreasing timestamps"""
    reps, grid = 5, 16
    stdout, info, f = run(tmp_path, reps, grid=grid)
    assert info["enabled"] == 1 and info["errors"] == 0
    rows = [ROW.match(l) for l in f.read_text().splitlines()[1:]]
    assert all(rows) and len(rows) == grid * len(SEQ) * reps
    by_block = {}
    for m in rows:
        by_block.setdefault(int(m.group(3)), []).append(m)
    assert sorted(by_block) == list(range(grid))
    for b, ms in by_block.items():
        assert [(m.group(5), int(m.group(8))) for m in ms] == SEQ * reps, b
        ts = [int(m.group(1)) for m in ms]
        assert ts == sorted(ts), b
    assert "overflow: 0 conflict: 0" in stdout


def status(stdout):
    m = re.search(r"overflow: (\d+) conflict: (\d+) records: (\d+) sequence errors: (\d+)", stdout)
    assert m, stdout
    return tuple(int(x) for x in m.groups())


def test_live_drain_keeps_records_past_the_ring(tmp_path):
    """one kernel emits 1400 records through an 8-record ring: the drain
    thread empties it while the kernel runs, so every record arrives, in
    issue order"""
    reps = 200
    stdout, info, f = run(tmp_path, reps, RING8)
    assert info["errors"] == 0
    rows = [ROW.match(l) for l in f.read_text().splitlines()[1:]]
    assert all(rows) and len(rows) == len(SEQ) * reps
    assert [(m.group(5), int(m.group(8))) for m in rows] == SEQ * reps
    ts = [int(m.group(1)) for m in rows]
    assert ts == sorted(ts)
    assert status(stdout) == (0, 0, len(SEQ) * reps, 0)


def test_live_drain_many_workgroups(tmp_path):
    """16 workgroups x 350 records through a 64-record ring: none lost or
    doubled, each workgroup's records in its issue order"""
    reps, grid = 50, 16
    stdout, info, f = run(tmp_path, reps, {"NVSHMEM_LOG_BUFFER_SIZE": "4096", "NVSHMEM_LOG_BUFFER_COUNT": "1"},
                          grid=grid)
    rows = [ROW.match(l) for l in f.read_text().splitlines()[1:]]
    assert all(rows) and len(rows) == grid * len(SEQ) * reps
    by_block = {}
    for m in rows:
        by_block.setdefault(int(m.group(3)), []).append((m.group(5), int(m.group(8))))
    assert sorted(by_block) == list(range(grid))
    assert all(v == SEQ * reps for v in by_block.values())
    assert status(stdout) == (0, 0, grid * len(SEQ) * reps, 0)


@pytest.mark.parametrize("lanes,block", [(64, 64), (37, 64), (256, 256)])
def test_live_drain_whole_wavefronts(tmp_path, lanes, block):
    """every lane of a wavefront records at once (one claim per wavefront):
    4 workgroups x `lanes` threads x 100 puts through a 256-record ring; each
    thread's records are all there, in its issue order, and within one
    wavefront's call the lanes come out in lane order"""
    reps, grid = 100, 4
    stdout, info, f = run(tmp_path, reps, {"NVSHMEM_LOG_BUFFER_SIZE": "16384", "NVSHMEM_LOG_BUFFER_COUNT": "1"},
                          grid=grid, lanes=lanes, block=block)
    assert info["errors"] == 0
    rows = [ROW.match(l) for l in f.read_text().splitlines()[1:]]
    assert all(rows) and len(rows) == grid * lanes * reps
    assert {(m.group(5), int(m.group(8))) for m in rows} == {("put_int", 1)}
    per_thread = {}
    for m in rows:
        per_thread.setdefault((int(m.group(3)), int(m.group(4))), []).append((int(m.group(1)), int(m.group(7), 16)))
    assert sorted(per_thread) == [(b, t) for b in range(grid) for t in range(lanes)]
    for (b, t), recs in per_thread.items():
        assert len(recs) == reps and len({a for _, a in recs}) == 1, (b, t)
        assert recs == sorted(recs), (b, t)          # claim order is the thread's issue order
    # the remote addresses of one thread's puts step by 4 bytes across lanes
    base = per_thread[(0, 0)][0][1]
    assert all(per_thread[(0, t)][0][1] == base + 4 * t for t in range(lanes))
    if block == 64:
        # one wavefront per workgroup: its call's records are consecutive, lane-ordered
        for b in range(grid):
            tids = [int(m.group(4)) for m in rows if int(m.group(3)) == b]
            assert tids == list(range(lanes)) * reps, b
    assert status(stdout) == (0, 0, grid * lanes * reps, 0)
