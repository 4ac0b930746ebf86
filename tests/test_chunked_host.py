"""Chunk domains, box overlaps and the local redistribution patterns of
chunked arrays (compute_overlap, create_workspace's local branch,
src/fft.jl:176-178, 233-343) on the host: the patterns, replayed with the
oracle's local_copy! restatement on host chunks, reproduce the oracle's
redistribution, and the destination boxes partition every destination
chunk exactly once."""
import types

import numpy as np
import pytest

from oracle import oracle as O


def _stand_in(shape, blocks, dtype):
    from daggerfft import chunked as C
    return types.SimpleNamespace(shape=tuple(shape), dtype=np.dtype(dtype),
                                 subdomains=C.block_domains(shape, blocks))


CASES = [((24, 20, 18), (24, 8, 8), (8, 20, 8)),      # x-pencils -> y-pencils, uneven edges
         ((24, 20, 18), (8, 20, 8), (8, 8, 18)),      # y-pencils -> z-pencils
         ((16, 12, 10), (16, 12, 4), (4, 12, 10)),    # slabs -> slabs
         ((7, 5, 3), (3, 2, 3), (2, 5, 1)),           # odd sizes, blocks that do not divide
         ((9, 9, 9), (9, 9, 9), (9, 9, 9))]           # one chunk each


@pytest.mark.parametrize("shape,sblocks,dblocks", CASES)
def test_domains_match_oracle_blocks(shape, sblocks, dblocks):
    from daggerfft import chunked as C
    for blocks in (sblocks, dblocks):
        got = [(tuple(r.start for r in d.indexes), tuple(r.stop for r in d.indexes))
               for d in C.block_domains(shape, blocks)]
        assert got == list(O.Blocked(shape, blocks, np.float32).domains)


@pytest.mark.parametrize("shape,sblocks,dblocks", CASES)
def test_patterns_reproduce_oracle_redistribution(shape, sblocks, dblocks):
    from daggerfft import chunked as C
    a = O.fill(5, shape, np.complex128)
    src = O.Blocked.distribute(a, sblocks)
    ref = O.Blocked(shape, dblocks, np.complex128)
    O.redistribute(ref, src)
    ws = C.create_workspace(_stand_in(shape, sblocks, np.complex128), _stand_in(shape, dblocks, np.complex128))
    got = O.Blocked(shape, dblocks, np.complex128)
    cover = [np.zeros(ch.shape, dtype=np.int32) for ch in got.chunks]
    for p in ws.local_patterns:
        cnt = tuple(len(r) for r in p.overlap)
        assert p.buffer_size == int(np.prod(cnt)) and cnt == tuple(len(r) for r in p.src_indices)
        O.local_copy(got.chunks[p.dst_idx], tuple(r.start for r in p.dst_indices), src.chunks[p.src_idx],
                     tuple(r.start for r in p.src_indices), cnt)
        d = p.dst_indices
        cover[p.dst_idx][d[0].start:d[0].stop, d[1].start:d[1].stop, d[2].start:d[2].stop] += 1
    for g, r, c in zip(got.chunks, ref.chunks, cover):
        assert np.array_equal(g, r)
        assert np.all(c == 1)
    # the reference's loop order: destination chunks outer, source chunks inner
    order = [(p.dst_idx, p.src_idx) for p in ws.local_patterns]
    assert order == sorted(order)


def test_compute_overlap_empty_and_edges():
    from daggerfft import chunked as C
    a = C.ArrayDomain((range(0, 4), range(0, 4), range(0, 4)))
    b = C.ArrayDomain((range(4, 8), range(0, 4), range(0, 4)))
    c = C.ArrayDomain((range(3, 8), range(2, 9), range(1, 2)))
    assert len(C.compute_overlap(a, b)[0]) == 0          # touching boxes do not overlap
    assert C.compute_overlap(a, c) == (range(3, 4), range(2, 4), range(1, 2))
    assert C.compute_overlap(c, a) == C.compute_overlap(a, c)


def test_workspace_rejects_mismatched_arrays():
    from daggerfft import chunked as C
    with pytest.raises(ValueError):
        C.create_workspace(_stand_in((8, 8, 8), (8, 4, 4), np.complex128),
                           _stand_in((8, 8, 4), (4, 8, 4), np.complex128))
