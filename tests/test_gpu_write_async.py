"""The background pos writer (bdk_write_pos_async / bdk_write_wait): the
files it writes while the device keeps stepping are byte-identical to the
synchronous writer's (whose format the oracle pins against au:631-648) for
the state at the time of the call, on a lone handle and on a batch member
whose batch steps on right after the call; a file that cannot be opened is
reported by the wait."""
import os
import tempfile

import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu

SPEC = dict(n1=1200, n2=800, L=70.0, Z=64, scheme=bdk.CIC, lam=20.0, T=1.0, dt_max=1e-4, seed=21)


def read(path):
    with open(path, "rb") as f:
        return f.read()


def test_async_files_equal_sync_files_while_stepping():
    print("case: async pos writer, lone handle", flush=True)
    bp, _ = params(**SPEC)
    git, rel = positions(bp, 21, 0.9)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    with tempfile.TemporaryDirectory() as tmp:
        a1, a2, b1, b2 = (os.path.join(tmp, n) for n in ("a1", "a2", "b1", "b2"))
        for nr in range(3):
            s.run_steps(20)
            s.write_pos(a1, a2, nr, 0.5)
            s.write_pos_async(b1, b2, nr, 0.5)
            s.run_steps(40)  # enqueued behind the copies
            s.write_wait()
        assert read(a1) == read(b1) and read(a2) == read(b2)
        text = read(b1).decode().splitlines()
        assert text[0].startswith("# Format")
        assert len([t for t in text if t.strip()]) == 1 + 3 * SPEC["n1"]
        # a new call waits for the previous write itself
        s.write_pos_async(b1, b2, 3, 0.5)
        s.write_pos_async(b1, b2, 4, 0.5)
        s.write_pos(a1, a2, 3, 0.5)
        s.write_pos(a1, a2, 4, 0.5)
        s.write_wait()
        assert read(a1) == read(b1) and read(a2) == read(b2)
        s.write_wait()  # nothing in flight
    s.close()


def test_async_write_of_batch_member_then_batch_steps():
    print("case: async pos writer, batch member", flush=True)
    bp, _ = params(**SPEC)
    git, rel = positions(bp, 22, 0.9)
    alone, member = bdk.RunDynamik(bp), bdk.RunDynamik(bp)
    for s in (alone, member):
        s.set_positions(git, rel)
    b = bdk.Batch([member])
    b.run_steps(25)
    b.sync()
    alone.run_steps(25)
    with tempfile.TemporaryDirectory() as tmp:
        a1, a2, b1, b2 = (os.path.join(tmp, n) for n in ("a1", "a2", "b1", "b2"))
        member.write_pos_async(b1, b2, 0, 0.1)
        b.run_steps(30)  # waits for the member's copies
        alone.write_pos(a1, a2, 0, 0.1)
        member.write_wait()
        b.sync()
        assert read(a1) == read(b1) and read(a2) == read(b2)
    alone.run_steps(30)
    ga, ra = alone.positions()
    gm, rm = member.positions()
    assert np.array_equal(ga, gm) and np.array_equal(ra, rm)
    b.close()
    alone.close()
    member.close()


def test_async_write_error_reported_by_wait():
    print("case: async pos writer, unwritable path", flush=True)
    bp, _ = params(**dict(SPEC, n1=300, n2=200, L=40.0, Z=32))
    git, rel = positions(bp, 3, 0.9)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    with tempfile.TemporaryDirectory() as tmp:
        bad = os.path.join(tmp, "no_such_dir", "pos1.txt")
        s.write_pos_async(bad, None, 0, 0.1)
        s.run_steps(5)
        with pytest.raises(bdk.BdkError, match="cannot open"):
            s.write_wait()
        s.write_wait()  # the error was consumed
        good = os.path.join(tmp, "p1")
        s.write_pos_async(good, None, 0, 0.1)
        s.write_wait()
        assert read(good).startswith(b"# Format")
    s.close()


def test_many_chunk_writer_equals_oracle_writer(tmp_path):
    """Enough particles that both writers format the lines on several host
    threads: the files still equal the oracle's single-threaded writer."""
    print("case: multi-thread formatting, 150000 + 70000 particles", flush=True)
    import oracle as O
    sp = dict(n1=150000, n2=70000, L=700.0, Z=64, scheme=bdk.CIC, lam=20.0, T=1.0, dt_max=1e-4, seed=4)
    bp, op = params(**sp)
    B = bdk.derived(bp)["B"]
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, sp["L"], (sp["n1"] + sp["n2"], 2))
    git = np.minimum(np.floor(x / B), bdk.derived(bp)["nc"] - 1).astype(np.int32)
    rel = x - git * B
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    g, r = s.positions()
    o = O.Oracle(op)
    o.set_all(g, r)
    p = {k: str(tmp_path / k) for k in ("a1", "a2", "b1", "b2", "o1", "o2")}
    for nr in range(2):
        s.write_pos(p["a1"], p["a2"], nr, 0.25)
        s.write_pos_async(p["b1"], p["b2"], nr, 0.25)
        s.write_wait()
        if nr == 0:
            o.write_pos_header(p["o1"])
            o.write_pos_header(p["o2"])
        o.append_pos(1, p["o1"], nr, 0.25)
        o.append_pos(2, p["o2"], nr, 0.25)
    for a, b, c in (("a1", "b1", "o1"), ("a2", "b2", "o2")):
        assert read(p[a]) == read(p[c]) and read(p[b]) == read(p[c])
    o.close()
    s.close()
