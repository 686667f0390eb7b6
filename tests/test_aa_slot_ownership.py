"""The AA step runs in place on one array, so it is race-free only if every
slot it touches has a single owner.  In the odd step (src/kernel_aa_soa.c:
173-271, with half-way bounce-back folded into the addressing) fluid cell y
reads, for each direction d,
    solid(y - c_d) ? a[y, d] : a[y - c_d, inv d]
and writes f*_d to
    solid(y + c_d) ? a[y, inv d] : a[y + c_d, d].
Ownership means: the 19 slots a cell reads are the 19 slots it writes, and
no two cells share a slot.  Then the kernels may update all cells at once,
in any order, without a read of one cell meeting a write of another.  The
even step reads and writes the cell's own 19 slots (a[y, d] and a[y, inv d]),
which is trivially exclusive.  Checked here on the CPU, with every axis
wrapping as the device addressing does, for the geometries the tests use
(walls, spheres, blocks, fully periodic fluid), including axes of one or
two cells."""
import numpy as np
import pytest
from conftest import CX, CY, CZ, INV

CASES = [
    ("channel", (12, 10, 9), (0, 0, 1), 0),
    ("spheres-6", (24, 26, 30), (0, 0, 1), 3),
    ("blocks-2", (14, 12, 16), (0, 0, 1), 0),
    ("box", (9, 8, 10), (0, 0, 0), 0),
    ("pipe", (10, 11, 11), (0, 0, 0), 0),
    ("fluid", (5, 6, 7), (1, 1, 1), 0),
    ("fluid", (1, 2, 5), (1, 1, 1), 0),
    ("fluid", (2, 2, 2), (1, 1, 1), 0),
]


def _slots(flags):
    """read and write slot ids (cell * 19 + direction), shape (fluid, 19),
    of the AA odd step; flags[x, y, z] != 0 marks fluid"""
    nx, ny, nz = flags.shape
    fluid = flags != 0
    x, y, z = np.nonzero(fluid)
    cell = (x * ny + y) * nz + z
    reads = np.empty((len(cell), 19), dtype=np.int64)
    writes = np.empty_like(reads)
    for d in range(19):
        if d == 18:  # the rest direction: the cell's own slot
            reads[:, d] = cell * 19 + d
            writes[:, d] = cell * 19 + d
            continue
        sx, sy, sz = (x - CX[d]) % nx, (y - CY[d]) % ny, (z - CZ[d]) % nz
        tx, ty, tz = (x + CX[d]) % nx, (y + CY[d]) % ny, (z + CZ[d]) % nz
        src = (sx * ny + sy) * nz + sz
        tgt = (tx * ny + ty) * nz + tz
        src_solid = ~fluid[sx, sy, sz]
        tgt_solid = ~fluid[tx, ty, tz]
        reads[:, d] = np.where(src_solid, cell * 19 + d, src * 19 + INV[d])
        writes[:, d] = np.where(tgt_solid, cell * 19 + INV[d], tgt * 19 + d)
    return reads, writes


@pytest.mark.parametrize("geo,dims,per,seed", CASES)
def test_aa_odd_slots_have_single_owners(oracle, geo, dims, per, seed):
    lat, _, (nfluid, _) = oracle.geometry(geo, dims, per, seed)
    flags = lat.transpose(2, 1, 0)
    reads, writes = _slots(flags)
    assert len(reads) == nfluid > 0
    # each cell writes exactly the slots it reads
    assert np.array_equal(np.sort(reads, axis=1), np.sort(writes, axis=1))
    # and no slot belongs to two cells (nor twice to one)
    assert len(np.unique(writes)) == writes.size


def test_a_dropped_bounce_back_breaks_ownership(oracle):
    """the check has teeth: reading the neighbour's slot even where the
    neighbour is solid (bounce-back left out of the read) makes two fluid
    cells next to an obstacle claim slots they do not write"""
    lat, _, _ = oracle.geometry("spheres-6", (24, 26, 30), (0, 0, 1), 3)
    flags = lat.transpose(2, 1, 0)
    reads, writes = _slots(flags)
    nx, ny, nz = flags.shape
    x, y, z = np.nonzero(flags != 0)
    broken = reads.copy()
    for d in range(18):
        sx, sy, sz = (x - CX[d]) % nx, (y - CY[d]) % ny, (z - CZ[d]) % nz
        broken[:, d] = ((sx * ny + sy) * nz + sz) * 19 + INV[d]
    assert not np.array_equal(np.sort(broken, axis=1), np.sort(writes, axis=1))
