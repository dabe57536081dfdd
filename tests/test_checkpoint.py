"""MPS checkpoints in the reference's HDF5 layout (HDF5.write / HDF5.read of GMPS,
src/structures/mps/gmps.jl:359-383), CPU-only: round trip, and the file's structure as the
HDF5 tools print it (group attributes type = "MPS", version = 1; scalar datasets length
and center; datasets MPS[i] with HDF5.jl's reversed dims and (r, i) complex compound)."""
import shutil
import subprocess

import numpy as np
import pytest

from tnhip.checkpoint import read_mps, write_mps
from tnhip.models import GMPS


def sample(rng, N=5, chi=3):
    ts = []
    for i in range(N):
        a = 1 if i == 0 else chi
        b = 1 if i == N - 1 else chi
        ts.append(np.asfortranarray(rng.standard_normal((a, 2, b)) + 1j * rng.standard_normal((a, 2, b))))
    return GMPS(1, 2, ts, 3)


def test_round_trip(tmp_path, rng):
    psi = sample(rng)
    p = tmp_path / "psi.h5"
    write_mps(p, psi)
    back = read_mps(p)
    assert back.center == 3 and len(back) == 5 and back.rank == 1 and back.dim == 2
    for a, b in zip(psi.tensors, back.tensors):
        assert a.shape == b.shape and np.array_equal(a, b)


def test_other_group_name_and_rank2(tmp_path, rng):
    W = [np.asfortranarray(rng.standard_normal((1 if i == 0 else 3, 2, 2, 1 if i == 3 else 3)) + 0j)
         for i in range(4)]
    O = GMPS(2, 2, W, 0)
    p = tmp_path / "mpo.h5"
    write_mps(p, O, name="H")
    back = read_mps(p, name="H")
    assert back.rank == 2 and all(np.array_equal(a, b) for a, b in zip(W, back.tensors))
    with pytest.raises(OSError):
        read_mps(p, name="psi")


@pytest.mark.skipif(shutil.which("h5dump") is None and not shutil.os.path.exists("/opt/conda/bin/h5dump"),
                    reason="h5dump not available")
def test_file_layout_as_hdf5_tools_see_it(tmp_path, rng):
    psi = sample(rng, N=3, chi=4)
    p = tmp_path / "psi.h5"
    write_mps(p, psi)
    exe = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([exe, "-H", str(p)], capture_output=True, text=True, check=True).stdout
    assert 'GROUP "psi"' in out
    assert 'ATTRIBUTE "type"' in out and 'ATTRIBUTE "version"' in out
    assert 'DATASET "length"' in out and 'DATASET "center"' in out
    assert 'DATASET "MPS[2]"' in out
    assert 'H5T_IEEE_F64LE "r"' in out and 'H5T_IEEE_F64LE "i"' in out
    # Julia size (4, 2, 4) for site 2 is stored as ( 4, 2, 4 ) reversed; site 1 (1, 2, 4) as (4, 2, 1)
    assert "( 4, 2, 1 )" in out and "SCALAR" in out
    data = subprocess.run([exe, "-a", "/psi/type", str(p)], capture_output=True, text=True, check=True).stdout
    assert '"MPS"' in data
