"""Output and checkpoint formats (CPU): the VTK ImageData writer, the
atomic staging swap and the checkpoint container, on host tensors.  The
GPU round trip (a restarted run continuing bit for bit) is in
tests/test_gpu_io.py."""
import json
import os
import re
import types

import numpy as np
import pytest
import torch

from pytnl_lbm import _abi
from pytnl_lbm import io as lio


def test_vti_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(1)
    X, Y, Z = 5, 4, 3
    f = {"lbm_density": 1 + 0.01 * rng.standard_normal((X, Y, Z)), "velocity_x": rng.standard_normal((X, Y, Z)),
         "wall": rng.integers(0, 3, (X, Y, Z)).astype(np.int16), "velocity_y": rng.standard_normal((X, Y, Z)).astype(np.float32)}
    p = tmp_path / "out.vti"
    lio.write_vti(p, f, spacing=0.25, origin=(1.0, 2.0, 3.0), time=0.5)
    back = lio.read_vti(p)
    for k, v in f.items():
        assert back[k].dtype == v.dtype and np.array_equal(back[k], v)
    raw = p.read_bytes()
    head = raw[:raw.index(b"<AppendedData")].decode()
    assert 'WholeExtent="0 5 0 4 0 3"' in head and 'Spacing="0.25 0.25 0.25"' in head
    assert 'Origin="1.0 2.0 3.0"' in head and ">0.5</DataArray>" in head
    # x fastest: the first values of a field's block are f[0,0,0], f[1,0,0], ...
    data = raw[raw.index(b"_", raw.index(b"<AppendedData")) + 1:]
    off = int(re.search(r'Name="lbm_density" format="appended" offset="(\d+)"', head).group(1))
    first = np.frombuffer(data[off + 8:off + 8 + 3 * 8], dtype="<f8")
    assert np.array_equal(first, f["lbm_density"][:3, 0, 0])


def test_rename_exchange_swaps_and_creates(tmp_path):
    a, b = tmp_path / "ckpt_tmp", tmp_path / "ckpt"
    a.mkdir()
    (a / "x").write_text("new")
    lio.rename_exchange(str(a), str(b))  # target missing: a plain move
    assert (b / "x").read_text() == "new" and not a.exists()
    a.mkdir()
    (a / "x").write_text("newer")
    lio.rename_exchange(str(a), str(b))  # both exist: swapped
    assert (b / "x").read_text() == "newer" and (a / "x").read_text() == "new"


class FakeBlock:
    """The attributes save/load_checkpoint use, on host tensors."""

    def __init__(self, storage=_abi.STORE_BF16_DEV, pattern=_abi.PATTERN_AB, seed=0, X=4, Y=3, Z=2):
        self.X, self.Y, self.Z = X, Y, Z
        self.ox = self.oy = self.oz = 0
        self.cfg = _abi.Config.make(storage=storage, pattern=pattern, macro_real=_abi.REAL_F32)
        self.d = _abi.D3Q27Data.make(X, Y, Z)
        self.d.lbm_viscosity = 0.01
        g = torch.Generator().manual_seed(seed)
        n = self.d.XYZ
        el = torch.bfloat16 if storage == _abi.STORE_BF16_DEV else torch.float64
        self.df = [torch.randn(27 * n, generator=g).to(el) for _ in range(2 if pattern == _abi.PATTERN_AB else 1)]
        self.macro = torch.randn(4 * n, generator=g, dtype=torch.float32)
        self.map = torch.randint(0, 10, (n,), generator=g, dtype=torch.int16)
        self.iterations = 7 + seed
        self.bound = 0

    def _bind(self):
        self.bound += 1


def test_checkpoint_round_trip_with_reference_names(tmp_path):
    a, b = FakeBlock(seed=1), FakeBlock(seed=2)
    state = types.SimpleNamespace(total_blocks=1, physCharLength=0.1, physFinalTime=2.0,
                                  counters=[types.SimpleNamespace(count=3, period=0.25)])
    p = tmp_path / "results_x" / "checkpoint"
    lio.save_checkpoint(p, [a], state=state)
    names = sorted(f.name for f in p.iterdir())
    assert names == ["LBM_df_0_block_0.npy", "LBM_df_1_block_0.npy", "LBM_macro_block_0.npy", "LBM_map_block_0.npy",
                     "attributes.json"]
    attrs = json.loads((p / "attributes.json").read_text())
    for key in ("LBM_total_blocks", "LBM_physCharLength", "LBM_physFinalTime", "LBM_iterations",
                "State_counter_0_count", "State_counter_0_period"):
        assert key in attrs
    assert attrs["LBM_iterations"] == a.iterations
    state2 = types.SimpleNamespace(counters=[types.SimpleNamespace(count=0, period=0.0)], physCharLength=0.0,
                                   physFinalTime=0.0)
    lio.load_checkpoint(p, [b], state=state2)
    for x, y in zip(a.df, b.df):
        assert torch.equal(x.view(torch.int16), y.view(torch.int16))
    assert torch.equal(a.macro, b.macro) and torch.equal(a.map, b.map)
    assert b.iterations == a.iterations and b.bound == 1
    assert state2.counters[0].count == 3 and state2.physFinalTime == 2.0
    # a second save keeps the previous checkpoint in the staging directory until the swap
    a.iterations += 5
    lio.save_checkpoint(p, [a])
    assert json.loads((p / "attributes.json").read_text())["LBM_iterations"] == a.iterations
    assert json.loads((tmp_path / "results_x" / "checkpoint_tmp" / "attributes.json").read_text())[
        "LBM_iterations"] == a.iterations - 5


def test_checkpoint_restores_the_map_into_a_lattice_without_one(tmp_path):
    """A State resuming a run has not built its map when it loads the
    checkpoint: the saved map is created on the lattice; a run saved without
    a map comes back without one."""
    a, b = FakeBlock(seed=1), FakeBlock(seed=2)
    b.map = None
    lio.save_checkpoint(tmp_path / "c", [a])
    lio.load_checkpoint(tmp_path / "c", [b])
    assert b.map is not None and torch.equal(a.map, b.map)
    a.map = None
    lio.save_checkpoint(tmp_path / "d", [a])
    lio.load_checkpoint(tmp_path / "d", [b])
    assert b.map is None


def test_checkpoint_refuses_a_different_lattice(tmp_path):
    a = FakeBlock()
    lio.save_checkpoint(tmp_path / "c", [a])
    with pytest.raises(ValueError):
        lio.load_checkpoint(tmp_path / "c", [FakeBlock(storage=_abi.STORE_F64)])
    with pytest.raises(ValueError):
        lio.load_checkpoint(tmp_path / "c", [FakeBlock(X=5)])


def test_macro_view_is_a_view(tmp_path):
    a = FakeBlock()
    v = lio.macro_view(a, 2)
    assert tuple(v.shape) == (a.X, a.Y, a.Z)
    v[1, 2, 0] = 42.0
    XYZ = a.d.XYZ
    assert a.macro[2 * XYZ + (1 * a.Z + 0) * a.Y + 2] == 42.0
    fields = lio.macro_fields(a, velocity_scale=2.0)
    assert np.allclose(fields["lbm_density_fluctuation"], lio.macro_view(a, 0).numpy() - 1)
    assert np.allclose(fields["velocity_y"], 2 * lio.macro_view(a, 2).numpy())
