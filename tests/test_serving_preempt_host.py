"""Host logic of serving preemption (no GPU): which running solves a waiting one preempts."""
from les_impulse3d_64_amd.serving import Request, Server


def _req(rid, nbytes, run):
    r = Request(rid, {'nx': 8, 'ny': 8, 'nz': 8}, 10, 0.0, rid)
    r.bytes, r.run = nbytes, run
    return r


def test_victims_longest_first_and_only_enough():
    srv = Server(budget_bytes=100, max_members=3, quantum=4)
    active = [(_req(0, 40, 5), 'g0'), (_req(1, 30, 9), 'g1'), (_req(2, 30, 2), 'g2')]
    head = _req(9, 30, 0)
    # 100 used: the longest runner that has finished its quantum (30 bytes) makes room for 30 bytes
    assert [r.rid for r, _ in srv._victims(head, active, 100)] == [1]
    head.bytes = 65                       # needs both that ran a quantum (70 bytes freed)
    assert [r.rid for r, _ in srv._victims(head, active, 100)] == [1, 0]
    head.bytes = 95                       # even both would not make room (rid 2 has not run its quantum)
    assert srv._victims(head, active, 100) == []
    head.bytes = 101                      # can never fit: rejected by admission, nothing preempted
    assert srv._victims(head, active, 100) == []


def test_no_victims_when_the_head_fits_or_slots_bind():
    srv = Server(budget_bytes=100, max_members=2, quantum=1)
    active = [(_req(0, 10, 3), 'g0'), (_req(1, 10, 1), 'g1')]
    head = _req(9, 10, 0)
    # slots full (2 of 2): one preemption frees a slot
    assert [r.rid for r, _ in srv._victims(head, active, 20)] == [0]
    srv.max_members = 3                  # a free slot and memory: the head fits without preempting
    assert srv._victims(head, active, 20) == []
