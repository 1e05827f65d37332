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


def _sized(rid, nbytes, cells, steps):
    r = Request(rid, {'nx': cells, 'ny': 1, 'nz': 1}, steps, 0.0, rid)
    r.bytes = nbytes
    return r


def test_admission_policies():
    """FIFO admits in arrival order and stops at the first request that does not fit; 'smallest' takes the
    least remaining work first and passes over what does not fit; both reject what can never fit."""
    queue = [_sized(0, 60, 100, 10), _sized(1, 50, 10, 10), _sized(2, 30, 50, 1), _sized(3, 200, 1, 1),
             _sized(4, 10, 1000, 1000)]
    fifo = Server(budget_bytes=100, max_members=8, policy='fifo')
    admit, reject = fifo._admissions(queue, busy=0, used=0)
    assert [r.rid for r in admit] == [0] and reject == []          # 0 fits; 1 (50) does not: stop there
    admit, reject = fifo._admissions(queue[2:], busy=0, used=0)
    assert [r.rid for r in admit] == [2, 4] and [r.rid for r in reject] == [3]   # 3 can never fit
    small = Server(budget_bytes=100, max_members=8, policy='smallest')
    admit, reject = small._admissions(queue, busy=0, used=0)
    # by work: 3 (1, never fits) -> rejected, 2 (50), 1 (100), 0 (1000, no room left), 4 (1e6, fits in 10)
    assert [r.rid for r in admit] == [2, 1, 4] and [r.rid for r in reject] == [3]
    admit, _ = small._admissions(queue, busy=7, used=0)             # one free slot
    assert [r.rid for r in admit] == [2]
    assert small._admissions(queue, busy=8, used=0)[0] == []


def test_preparing_solves_hold_their_slots():
    srv = Server(budget_bytes=100, max_members=3, quantum=1)
    active = [(_req(0, 10, 3), 'g0'), (_req(1, 10, 1), 'g1')]
    head = _req(9, 10, 0)
    assert srv._victims(head, active, 30, busy=0) == []             # a free slot
    assert [r.rid for r, _ in srv._victims(head, active, 30, busy=1)] == [0]
