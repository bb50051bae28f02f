"""Placement probing of the D3Q27 population buffer (placement="probe" /
"auto"): the probe's streaming-only launches run on the candidate buffer
before the populations are set, so a probed lattice steps to the same bits
as one kept as first allocated; the record of the probed rates; the
arguments it rejects."""
import pytest
import torch

from pytnl_lbm import D3Q27Lattice, _abi, lattice

pytestmark = pytest.mark.gpu


def run(storage, placement, tries=3, steps=5):
    lat = D3Q27Lattice(40, 36, 44, storage=storage, nu=0.01, placement=placement, placement_tries=tries,
                       store_range=0.5 if storage == "i16dev" else 1.0)
    lat.set_equilibrium(1.0, 0.03, -0.02, 0.01)
    for _ in range(steps):
        lat.step()
    torch.cuda.synchronize()
    return lat, lat.export_f64().cpu()


@pytest.mark.parametrize("storage", ["f64", "f32dev", "i16dev"])
def test_probed_lattice_steps_like_the_first_allocation(storage):
    first, a = run(storage, "first")
    assert first.placement is None
    probed, b = run(storage, "probe", tries=2)
    rec = probed.placement
    assert 1 <= len(rec["probed_GBs"]) <= 2 and 0 <= rec["kept"] < len(rec["probed_GBs"])
    assert all(r > 0 for r in rec["probed_GBs"])
    assert rec["probed_GBs"][rec["kept"]] == max(rec["probed_GBs"])
    assert torch.equal(a, b)


def test_a_slow_first_candidate_is_replaced(monkeypatch):
    """With a threshold no buffer reaches, every allowed candidate is probed
    and the fastest is kept; the lattice then steps as usual."""
    monkeypatch.setattr(lattice, "PLACEMENT_GOOD_GBS", float("inf"))
    lat, b = run("f64", "probe", tries=3)
    assert len(lat.placement["probed_GBs"]) == 3
    assert len(lat.df) == 1
    _, a = run("f64", "first")
    assert torch.equal(a, b)


def test_auto_leaves_small_lattices_alone():
    lat = D3Q27Lattice(32, 32, 32, nu=0.01)
    assert lat.placement is None


def test_probe_arguments_are_checked():
    with pytest.raises(ValueError):
        D3Q27Lattice(16, 16, 16, nu=0.01, placement="fastest")
    with pytest.raises(ValueError):
        D3Q27Lattice(16, 16, 16, nu=0.01, pattern="AB", placement="probe")
    # auto on an A-B lattice keeps the first allocation
    assert D3Q27Lattice(16, 16, 16, nu=0.01, pattern="AB").placement is None


def test_probed_slab_with_overlaps_steps_like_the_first_allocation():
    """A slab with ghost planes (the x-distributed share) is probed over its
    local sites; the distributed step afterwards (ghost planes read, here at
    their equilibrium values) matches the unprobed one."""
    out = []
    for placement in ("first", "probe"):
        lat = D3Q27Lattice(12, 20, 24, nu=0.01, overlap=(1, 0, 0), placement=placement, placement_tries=2)
        assert (lat.placement is None) == (placement == "first")
        lat.set_equilibrium(1.0, 0.02, 0.01, -0.03)
        for _ in range(3):
            lat.step(flags=_abi.FLAG_DISTRIBUTED_X)
        torch.cuda.synchronize()
        out.append(lat.export_f64().cpu())
    assert torch.equal(out[0], out[1])
