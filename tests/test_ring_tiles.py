"""LDS slot ownership of the mol_fast three-plane ring update (CPU only).

The ring kernel (k_mf_update_ring) runs 256 threads per (tk x tj) node tile.
Threads past tk * tj repeat a node of the tile's last row; if they wrote
that node's LDS slot, a repeat in a later wave could overwrite plane i - 1
while the owning wave still reads it.  The library routes their writes to
a sink slot.  This test takes the host's tile choice for every even grid
from 16 to 1024 cells (rq_mf_ring_tile, a host-only query) and checks that
every LDS slot written in a plane has exactly one writer, or is the sink.
"""
import ctypes as C

import pytest

WAVE = 64
THREADS = 256


def ring_tile(lib, NN):
    tk, tj, sink = C.c_int32(), C.c_int32(), C.c_int32()
    lib.rq_mf_ring_tile(NN, C.byref(tk), C.byref(tj), C.byref(sink))
    return tk.value, tj.value, sink.value


def write_slots(tk, tj, sink):
    """(thread, own-column write slot, halo slot or None) as the kernel forms them."""
    RK, RJ = tk + 2, tj + 2
    NH = 2 * RK + 2 * tj
    out = []
    for tid in range(THREADS):
        spare = tid >= tk * tj
        row = min(tid // tk, tj - 1)
        oslot = (row + 1) * RK + tid % tk + 1
        wslot = sink if spare else oslot
        hslot = None
        if tid < NH:
            if tid < RK:
                hj, hk = 0, tid
            elif tid < 2 * RK:
                hj, hk = RJ - 1, tid - RK
            else:
                q = tid - 2 * RK
                hj, hk = 1 + (q >> 1), (RK - 1 if q & 1 else 0)
            hslot = hj * RK + hk
        out.append((tid, oslot, wslot, hslot))
    return out


def test_every_ring_slot_has_one_writer(lib):
    cross_wave = []
    for NN in range(16, 1025, 2):
        tk, tj, sink = ring_tile(lib, NN)
        assert 1 <= tk <= 64 and tj >= 1 and tk * tj <= THREADS, (NN, tk, tj)
        assert (tk + 2) * (tj + 2) <= sink, (NN, tk, tj, sink)
        writers = {}
        for tid, oslot, wslot, hslot in write_slots(tk, tj, sink):
            if wslot != sink:
                writers.setdefault(wslot, []).append(tid)
            if hslot is not None:
                writers.setdefault(hslot, []).append(tid)
        assert all(len(w) == 1 for w in writers.values()), NN
        assert max(writers) < sink
        # the region is covered: every interior slot by its owner, every
        # border slot by one halo lane
        assert len(writers) == (tk + 2) * (tj + 2)
        # repeats whose node belongs to another wave: these would race on
        # the node's slot; the sink keeps them off it
        owners = {oslot: tid for tid, oslot, wslot, _ in write_slots(tk, tj, sink) if wslot != sink}
        if any(owners[o] // WAVE != tid // WAVE for tid, o, w, _ in write_slots(tk, tj, sink) if w == sink):
            cross_wave.append(NN)
    # the 129-node grid (43 x 5 tiles) is one of the shapes with cross-wave repeats
    assert 130 in cross_wave


@pytest.mark.parametrize("NN,shape", [(200, (50, 5)), (130, (43, 5)), (70, (35, 7))])
def test_ring_tile_shapes(lib, NN, shape):
    tk, tj, _ = ring_tile(lib, NN)
    assert (tk, tj) == shape
