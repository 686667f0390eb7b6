"""The state file (lbmSaveState / lbmLoadState, format in include/lbm_hip.h)
and the per-cell moments (lbmCellMoments):
  - resuming from a file continues exactly as the uninterrupted run;
  - the file holds the flags and the logical PDFs lbmGetNodes returns, and
    its checksum equals an independent restatement (lbmhip.read_state);
  - loading rejects other dimensions, other obstacles and a corrupted
    payload;
  - per-cell rho and j equal the moments of the logical PDFs."""
import numpy as np
import pytest
from conftest import CX, CY, CZ

from lbmhip import KERNELS, Lattice, read_state

pytestmark = pytest.mark.gpu

KW = dict(omega=1.3, xforce=1e-4, body_force=True, magic=3 / 16)


@pytest.mark.parametrize("kernel", KERNELS)
def test_resume_from_state_file(gpu_lib, tmp_path, kernel):
    dims, path = (12, 14, 64), tmp_path / "state.lbm"
    with Lattice(gpu_lib, kernel, "spheres-3", dims, (0, 0, 1), seed=5, **KW) as A:
        A.init("random", 3, 1.0, 1e-2)
        A.run(6)
        A.save_state(path)
        A.run(6)
        exp = A.get_nodes()
    with Lattice(gpu_lib, kernel, "spheres-3", dims, (0, 0, 1), seed=5, **KW) as B:
        B.load_state(path)
        B.run(6)
        got = B.get_nodes()
    assert np.array_equal(got, exp)


def test_state_file_contents(gpu_lib, tmp_path):
    dims, path = (10, 12, 16), tmp_path / "state.lbm"
    with Lattice(gpu_lib, "aa-soa", "channel", dims, (0, 0, 1), **KW) as L:
        L.init("random", 7, 1.0, 1e-2)
        L.run(3)
        L.save_state(path)
        nodes, flags, it, per = L.get_nodes(), L.lattice_flags(), L.iteration, L.periodic
    header, fflags, pdfs = read_state(path)   # verifies the checksum
    assert header["dims"] == list(dims) and header["periodic"] == list(per)
    assert header["q"] == 19 and header["iteration"] == it and header["kernel"] == "aa-soa"
    assert abs(header["omega"] - 1.3) < 1e-15 and header["magic_param"] == 3 / 16
    assert np.array_equal(fflags, flags)
    assert np.array_equal(pdfs, nodes)


def test_state_file_rejections(gpu_lib, tmp_path):
    path = tmp_path / "state.lbm"
    with Lattice(gpu_lib, "pull-soa", "channel", (10, 12, 16), (0, 0, 1), **KW) as L:
        L.init("random", 1, 1.0, 1e-2)
        L.save_state(path)
    with Lattice(gpu_lib, "pull-soa", "channel", (10, 12, 18), (0, 0, 1), **KW) as L:
        with pytest.raises(OSError):
            L.load_state(path)                  # other dimensions
    with Lattice(gpu_lib, "pull-soa", "blocks-2", (10, 12, 16), (0, 0, 1), **KW) as L:
        with pytest.raises(OSError):
            L.load_state(path)                  # other obstacle cells
    raw = bytearray(path.read_bytes())
    raw[-9] ^= 0x40                             # one bit of the last PDF
    bad = tmp_path / "bad.lbm"
    bad.write_bytes(bytes(raw))
    with Lattice(gpu_lib, "pull-soa", "channel", (10, 12, 16), (0, 0, 1), **KW) as L:
        with pytest.raises(OSError):
            L.load_state(bad)                   # checksum
    with pytest.raises(ValueError):
        read_state(bad)
    with Lattice(gpu_lib, "pull-soa", "channel", (10, 12, 16), (0, 0, 1), **KW) as L:
        with pytest.raises(OSError):
            L.load_state(tmp_path / "missing.lbm")


@pytest.mark.parametrize("storage", ["native", "fp16"])
def test_cell_moments(gpu_lib, storage):
    dims = (12, 10, 64)
    with Lattice(gpu_lib, "aa-soa", "spheres-3", dims, (0, 0, 1), storage=storage, seed=2,
                 **KW) as L:
        L.init("random", 4, 1.0, 1e-2)
        L.run(5)
        f = L.get_nodes()
        m = L.cell_moments()
        fluid = L.fluid_mask()
    exp = np.stack([f.sum(-1), f @ CX, f @ CY, f @ CZ], axis=-1)
    assert np.abs(m[fluid] - exp[fluid]).max() < 1e-14
    assert (m[~fluid] == 0).all()


@pytest.mark.parametrize("damage", ["bitflip", "truncate", "trailing"])
@pytest.mark.parametrize("kernel", ["aa-soa", "push-aos"])
def test_failed_load_leaves_the_state(gpu_lib, tmp_path, kernel, damage):
    """a corrupt, truncated or overlong file is rejected before any PDF is
    written: the lattice keeps the state it had"""
    dims, path = (10, 12, 16), tmp_path / "state.lbm"
    with Lattice(gpu_lib, kernel, "channel", dims, (0, 0, 1), **KW) as L:
        L.init("random", 1, 1.0, 1e-2)
        L.run(3)
        L.save_state(path)
    raw = bytearray(path.read_bytes())
    if damage == "bitflip":
        raw[-9] ^= 0x40
    elif damage == "truncate":
        raw = raw[:-8 * 19 * 5]
    else:
        raw += b"\0" * 8
    bad = tmp_path / "bad.lbm"
    bad.write_bytes(bytes(raw))
    with Lattice(gpu_lib, kernel, "channel", dims, (0, 0, 1), **KW) as L:
        L.init("random", 2, 1.0, 1e-2)
        L.run(5)
        before = L.get_nodes()
        with pytest.raises(OSError):
            L.load_state(bad)
        assert np.array_equal(L.get_nodes(), before)
        L.run(2)                                # and it still steps as before
        assert np.isfinite(L.get_nodes()).all()
