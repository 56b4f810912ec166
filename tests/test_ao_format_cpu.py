"""The reference's .ao ERI record format (src/pfourindexstore.f90:731-747,
781), read by the library's RESTORE reader, on records written here the way
a Fortran unformatted sequential WRITE lays them out (4-byte length, 1000
ICASHE1X, ICASHE1Y, ICASHE1Z, ICASHE2 int32 then 1000 DCASHE float64, length;
unused entries of the last record -1)."""
import struct

import numpy as np
import pytest

import polymer_amd as pa


def _write_records(path, entries, cells=None):
    """entries: ((i, j, k, l), value) 1-based; cells: packed ICASHE1X per entry."""
    with open(path, "wb") as f:
        for r0 in range(0, len(entries), 1000):
            chunk = entries[r0:r0 + 1000]
            cx = [0] * len(chunk) if cells is None else list(cells[r0:r0 + 1000])
            x = cx + [-1] * (1000 - len(chunk))
            yz = [0] * len(chunk) + [-1] * (1000 - len(chunk))
            ijkl = [(((i * 256 + j) * 256 + k) * 256 + l) for (i, j, k, l), _ in chunk] + [-1] * (1000 - len(chunk))
            v = [val for _, val in chunk] + [0.0] * (1000 - len(chunk))
            payload = (struct.pack("<1000i", *x) + struct.pack("<1000i", *yz) * 2 + struct.pack("<1000i", *ijkl) +
                       struct.pack("<1000d", *v))
            f.write(struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload)))


def test_restore_reads_fortran_records(tmp_path):
    rng = np.random.default_rng(0)
    nc = 7
    t = np.zeros((nc,) * 4, order="F")
    entries = []
    for _ in range(2345):                      # three records, the last one partial
        i, j, k, l = (int(x) for x in rng.integers(1, nc + 1, 4))
        v = float(rng.standard_normal())
        entries.append(((i, j, k, l), v))
        t[i - 1, j - 1, k - 1, l - 1] = v
    path = tmp_path / "h2o.ao.00000"
    _write_records(path, entries)
    r, n = pa.restore_fourindex_eri(path, nc)
    assert n == len(entries)
    assert np.array_equal(r, t)


def test_restore_rejects_a_truncated_file(tmp_path):
    path = tmp_path / "bad.ao.00000"
    _write_records(path, [((1, 1, 1, 1), 1.0)])
    data = open(path, "rb").read()
    open(path, "wb").write(data[:-100])
    with pytest.raises(pa.PolyeriError):
        pa.restore_fourindex_eri(path, 3)


def test_restore_cells_unpacks_periodic_records(tmp_path):
    """Periodic records: ICASHE1X = ((Q1+CEL1)*256 + Q2+CEL1)*256 + Q3+CEL2
    (src/pfourindexstore.f90:735), unpacked by MVBITS 16/8/0
    (src/pmp2energy.f90:221-241) into T[Q1, Q2, Q3, i, j, k, l]."""
    rng = np.random.default_rng(1)
    nc, cel1, cel2 = 5, 2, 3
    t = np.zeros((2 * cel1 + 1, 2 * cel1 + 1, 2 * cel2 + 1) + (nc,) * 4)
    entries, cells = [], []
    for _ in range(1500):
        q1, q2 = (int(x) for x in rng.integers(-cel1, cel1 + 1, 2))
        q3 = int(rng.integers(-cel2, cel2 + 1))
        i, j, k, l = (int(x) for x in rng.integers(1, nc + 1, 4))
        v = float(rng.standard_normal())
        entries.append(((i, j, k, l), v))
        cells.append(((q1 + cel1) * 256 + (q2 + cel1)) * 256 + (q3 + cel2))
        t[q1 + cel1, q2 + cel1, q3 + cel2, i - 1, j - 1, k - 1, l - 1] = v
    path = tmp_path / "pe.ao.00000"
    _write_records(path, entries, cells)
    r, n = pa.restore_fourindex_eri_cells(path, nc, cel1, cel2)
    assert n == len(entries)
    assert np.array_equal(r, t)
    # a cell beyond CUTOFF1 is a degraded file (src/pmp2energy.f90:248-252)
    with pytest.raises(pa.PolyeriError):
        pa.restore_fourindex_eri_cells(path, nc, cel1 - 1, cel2)
    # and periodic records are refused by the molecular reader
    with pytest.raises(pa.PolyeriError):
        pa.restore_fourindex_eri(path, nc)
