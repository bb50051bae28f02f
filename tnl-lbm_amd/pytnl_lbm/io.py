"""Macroscopic output and checkpoint/restart of lattice blocks.

Host-side counterparts of the reference's output and checkpoint layer:

* ``macro_view`` -- getMacroView (defs.h:179-240): a 3-D view of one
  quantity of a block's macro array, without a copy.
* ``write_vti`` -- UniformDataWriter (UniformDataWriter.hpp:20-110) writes
  ADIOS2 BP5 files that carry a VTK ImageData data model (cell data, x fastest,
  spacing physDl, origin lbm2physPoint(0, 0, 0)).  ADIOS2 is not available to
  this build, so the same data model is written as a VTK XML ImageData file
  with raw appended little-endian data, which ParaView and VTK read directly.
  The quantity names are the caller's (sim_4.cu writes lbm_density,
  lbm_density_fluctuation, velocity_x/y/z; ``macro_fields`` gives those).
* ``save_checkpoint`` / ``load_checkpoint`` -- State::saveState /
  loadState / checkpointState (state.hpp:891-1010, checkpoint.h:94-146): the
  attributes LBM_total_blocks, LBM_physCharLength, LBM_physFinalTime,
  LBM_iterations and State_counter_{c}_count/_period, and per block the
  variables LBM_map_block_{id}, LBM_df_{k}_block_{id} (k < DFMAX) and
  LBM_macro_block_{id}, each the block's whole storage array with its
  overlaps, as the reference saves them.  A checkpoint is written to a staging
  directory <path>_tmp and then swapped with <path> atomically
  (rename_exchange, lbm_common/fileutils.h:100-140: renameat2 with
  RENAME_EXCHANGE, or RENAME_NOREPLACE when <path> does not exist yet), so a
  failed write never breaks the previous checkpoint.  The container is a
  directory of NumPy .npy files plus attributes.json instead of ADIOS2 BP5
  (not available here); the variable names, the per-block layout and the
  staging swap are the reference's.  D3Q7 blocks add ADE_map_block_{id},
  ADE_df_{k}_block_{id} and ADE_phi_block_{id} (the reference does not
  checkpoint its ADE blocks).  Restarting from a checkpoint continues the run
  bit for bit.
"""
from __future__ import annotations

import ctypes
import errno
import json
import os
import shutil
from pathlib import Path

import numpy as np

from . import _abi

FORMAT = "tnl-lbm-amd checkpoint 1"


# ---- getMacroView ---------------------------------------------------------------
def macro_view(lattice, quantity: int):
    """[x, y, z] view of macro quantity `quantity` (0 = rho, 1..3 = vx, vy,
    vz for D3Q27; 0 = phi for D3Q7) over the block's own sites, as
    getMacroView (defs.h:179-240) returns it; writes through the view reach
    the device array."""
    XYZ = lattice.d.XYZ
    Xs, Ys, Zs = _stored_extents(lattice)
    arr = lattice.macro if hasattr(lattice, "macro") else lattice.phi
    nq = arr.numel() // XYZ
    if not 0 <= quantity < nq:
        raise IndexError(f"macro quantity {quantity} outside 0..{nq - 1}")
    plane = arr.view(nq, XYZ)[quantity, :Xs * Zs * Ys].view(Xs, Zs, Ys).permute(0, 2, 1)
    ox, oy, oz = _overlaps(lattice)
    return plane[ox:ox + lattice.X, oy:oy + lattice.Y, oz:oz + lattice.Z]


def macro_fields(lattice, velocity_scale: float = 1.0):
    """The fields sim_4.cu's outputData writes (sim_4.cu:173-215): lbm_density,
    lbm_density_fluctuation (rho - 1) and velocity_x/y/z (times
    velocity_scale, lbm2physVelocity), as host float arrays [x, y, z]."""
    rho = macro_view(lattice, 0).detach().cpu().numpy()
    out = {"lbm_density": rho, "lbm_density_fluctuation": rho - 1.0}
    for k, name in enumerate(("velocity_x", "velocity_y", "velocity_z")):
        out[name] = macro_view(lattice, 1 + k).detach().cpu().numpy() * velocity_scale
    return out


# ---- VTK ImageData -----------------------------------------------------------------
_VTK_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32", np.dtype(np.int16): "Int16",
              np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8", np.dtype(np.int64): "Int64"}


def write_vti(path, fields: dict, spacing: float = 1.0, origin=(0.0, 0.0, 0.0), time: float | None = None):
    """Cell data on an X x Y x Z image: each field an [x, y, z] array, stored
    x fastest; raw appended data with 64-bit block headers."""
    path = Path(path)
    arrays = []
    shape = None
    for name, a in fields.items():
        a = np.asarray(a)
        if a.dtype not in _VTK_TYPES:
            a = a.astype(np.float64)
        if shape is None:
            shape = a.shape
        if a.shape != shape or len(shape) != 3:
            raise ValueError(f"field {name}: shape {a.shape}, expected {shape}")
        arrays.append((name, np.ascontiguousarray(a.transpose(2, 1, 0)).astype(a.dtype.newbyteorder("<"), copy=False)))
    X, Y, Z = shape
    head, offset = [], 0
    for name, a in arrays:
        head.append(f'        <DataArray type="{_VTK_TYPES[np.dtype(a.dtype.name)]}" Name="{name}" format="appended" '
                    f'offset="{offset}"/>')
        offset += 8 + a.nbytes
    field_data = ""
    if time is not None:
        field_data = (f'    <FieldData>\n      <DataArray type="Float64" Name="TIME" NumberOfTuples="1" '
                      f'format="ascii">{time!r}</DataArray>\n    </FieldData>\n')
    xml = ('<?xml version="1.0"?>\n'
           '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
           f'  <ImageData WholeExtent="0 {X} 0 {Y} 0 {Z}" Origin="{origin[0]} {origin[1]} {origin[2]}" '
           f'Spacing="{spacing} {spacing} {spacing}">\n{field_data}'
           f'    <Piece Extent="0 {X} 0 {Y} 0 {Z}">\n      <CellData>\n' + "\n".join(head) +
           '\n      </CellData>\n    </Piece>\n  </ImageData>\n  <AppendedData encoding="raw">\n   _')
    with open(path, "wb") as fh:
        fh.write(xml.encode())
        for _, a in arrays:
            fh.write(np.uint64(a.nbytes).astype("<u8").tobytes())
            fh.write(a.tobytes())
        fh.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_vti(path):
    """The fields of a file written by write_vti, as [x, y, z] arrays."""
    raw = Path(path).read_bytes()
    cut = raw.index(b'<AppendedData encoding="raw">')
    xml = raw[:cut].decode()
    data = raw[raw.index(b"_", cut) + 1:]
    ext = [int(v) for v in xml.split('WholeExtent="')[1].split('"')[0].split()]
    X, Y, Z = ext[1], ext[3], ext[5]
    out = {}
    inv = {v: k for k, v in _VTK_TYPES.items()}
    for part in xml.split("<DataArray ")[1:]:
        if 'format="appended"' not in part:
            continue
        typ = part.split('type="')[1].split('"')[0]
        name = part.split('Name="')[1].split('"')[0]
        off = int(part.split('offset="')[1].split('"')[0])
        n = int(np.frombuffer(data[off:off + 8], dtype="<u8")[0])
        a = np.frombuffer(data[off + 8:off + 8 + n], dtype=inv[typ].newbyteorder("<"))
        out[name] = a.reshape(Z, Y, X).transpose(2, 1, 0)
    return out


# ---- checkpoints -------------------------------------------------------------------
def _overlaps(lat):
    return (getattr(lat, "ox", 0), getattr(lat, "oy", 0), getattr(lat, "oz", 0)) if not hasattr(lat, "nse") \
        else _overlaps(lat.nse)


def _stored_extents(lat):
    ox, oy, oz = _overlaps(lat)
    return lat.X + 2 * ox, lat.Y + 2 * oy, lat.Z + 2 * oz


def _to_host(t):
    """Device tensor -> host array of the same bits (bfloat16 as its int16 pattern)."""
    import torch

    t = t.detach()
    if t.dtype == torch.bfloat16:
        t = t.view(torch.int16)
    return t.cpu().numpy()


def _from_host(a, like):
    import torch

    src = torch.from_numpy(np.ascontiguousarray(a))
    if like.dtype == torch.bfloat16:
        src = src.view(torch.bfloat16)
    if src.dtype != like.dtype or src.numel() != like.numel():
        raise ValueError(f"checkpoint array {src.dtype}[{src.numel()}] does not fit {like.dtype}[{like.numel()}]")
    like.copy_(src.view(like.shape).to(like.device))


def rename_exchange(src: str, dst: str) -> None:
    """Atomically swap src and dst (dst may be missing): renameat2 with
    RENAME_EXCHANGE, else RENAME_NOREPLACE (lbm_common/fileutils.h:100-140)."""
    libc = ctypes.CDLL(None, use_errno=True)
    fn = getattr(libc, "renameat2", None)
    AT_FDCWD, RENAME_NOREPLACE, RENAME_EXCHANGE = -100, 1, 2
    if fn is not None:
        fn.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint]
        s, d = os.fsencode(src), os.fsencode(dst)
        if fn(AT_FDCWD, s, AT_FDCWD, d, RENAME_EXCHANGE) == 0:
            return
        e = ctypes.get_errno()
        if e == errno.ENOENT and fn(AT_FDCWD, s, AT_FDCWD, d, RENAME_NOREPLACE) == 0:
            return
        e = ctypes.get_errno()
        if e not in (errno.ENOSYS, errno.EINVAL):
            raise OSError(e, os.strerror(e), src, None, dst)
    # no renameat2: move the old checkpoint aside first, as the reference's fallback does
    if not os.path.exists(dst):
        os.rename(src, dst)
        return
    aside = dst + "_tmp_for_exchange"
    if os.path.exists(aside):
        raise OSError(errno.EEXIST, "temporary path already exists", aside)
    os.rename(dst, aside)
    os.rename(src, dst)
    os.rename(aside, src)


def _lattice_attributes(lat):
    d = lat.d
    a = {"X": lat.X, "Y": lat.Y, "Z": lat.Z, "overlaps": list(_overlaps(lat)), "XYZ": int(d.XYZ),
         "storage": int(lat.cfg.storage), "pattern": int(lat.cfg.pattern), "macro_real": int(lat.cfg.macro_real),
         "iterations": int(getattr(lat, "iterations", 0))}
    if hasattr(lat, "nse"):
        a.update(kind="D3Q7", collision=int(d.collision), lbm_diffusivity=d.lbm_diffusivity, phi_ref=d.phi_ref,
                 store_scale=d.store_scale)
    else:
        a.update(kind="D3Q27", collision=int(lat.cfg.collision), lbm_viscosity=d.lbm_viscosity,
                 force=[d.fx, d.fy, d.fz], store_range=d.store_range,
                 store_range_class=[d.store_range_class[n] for n in range(4)])
    return a


def save_checkpoint(path, blocks, *, ade_blocks=(), state=None) -> Path:
    """Write the blocks (D3Q27Lattice, one per id in order) and their D3Q7
    blocks to <path>, via <path>_tmp and an atomic swap.  `state` supplies the
    State attributes (total_blocks, physCharLength, physFinalTime and the
    Actions counters) where given; LBM_iterations is the first block's."""
    path = Path(path)
    tmp = Path(str(path) + "_tmp")
    if tmp.exists():
        shutil.rmtree(tmp)
    tmp.mkdir(parents=True)
    blocks = list(blocks)
    attrs = {"format": FORMAT, "LBM_total_blocks": len(blocks),
             "LBM_iterations": int(blocks[0].iterations) if blocks else 0}
    if state is not None:
        attrs["LBM_total_blocks"] = int(getattr(state, "total_blocks", len(blocks)))
        attrs["LBM_physCharLength"] = float(getattr(state, "physCharLength", 0.0))
        attrs["LBM_physFinalTime"] = float(getattr(state, "physFinalTime", 0.0))
        for c, cnt in enumerate(getattr(state, "counters", [])):
            attrs[f"State_counter_{c}_count"] = int(cnt.count)
            attrs[f"State_counter_{c}_period"] = float(cnt.period)
    attrs["blocks"] = {}
    arrays = {}
    for bid, lat in enumerate(blocks):
        attrs["blocks"][f"LBM_block_{bid}"] = _lattice_attributes(lat)
        if lat.map is not None:
            arrays[f"LBM_map_block_{bid}"] = lat.map
        for k, t in enumerate(lat.df):
            arrays[f"LBM_df_{k}_block_{bid}"] = t
        arrays[f"LBM_macro_block_{bid}"] = lat.macro
        if getattr(lat, "force_field", None) is not None:
            arrays[f"LBM_force_block_{bid}"] = lat.force_field
    for bid, ade in enumerate(ade_blocks):
        attrs["blocks"][f"ADE_block_{bid}"] = _lattice_attributes(ade)
        if ade.map is not None:
            arrays[f"ADE_map_block_{bid}"] = ade.map
        for k, t in enumerate(ade.df):
            arrays[f"ADE_df_{k}_block_{bid}"] = t
        arrays[f"ADE_phi_block_{bid}"] = ade.phi
    attrs["variables"] = {}
    for name, t in arrays.items():
        a = _to_host(t)
        with open(tmp / f"{name}.npy", "wb") as fh:
            np.save(fh, a, allow_pickle=False)
            fh.flush()
            os.fsync(fh.fileno())
        attrs["variables"][name] = {"dtype": str(a.dtype), "size": int(a.size)}
    with open(tmp / "attributes.json", "w") as fh:
        json.dump(attrs, fh, indent=1)
        fh.flush()
        os.fsync(fh.fileno())
    rename_exchange(str(tmp), str(path))
    os.utime(path)
    return path


def read_attributes(path) -> dict:
    attrs = json.loads((Path(path) / "attributes.json").read_text())
    if attrs.get("format") != FORMAT:
        raise ValueError(f"{path}: not a checkpoint of this format ({attrs.get('format')!r})")
    return attrs


def load_checkpoint(path, blocks, *, ade_blocks=(), state=None) -> dict:
    """Restore blocks created with the same geometry, storage form and
    pattern: populations, macro, map and the iteration count (and so the A-A
    parity and the A-B array rotation) as saved.  Returns the attributes."""
    path = Path(path)
    attrs = read_attributes(path)
    blocks = list(blocks)

    def load(name, like):
        meta = attrs["variables"].get(name)
        if meta is None:
            raise KeyError(f"{path}: variable {name} missing")
        _from_host(np.load(path / f"{name}.npy", allow_pickle=False), like)

    for prefix, group in (("LBM", blocks), ("ADE", list(ade_blocks))):
        for bid, lat in enumerate(group):
            saved = attrs["blocks"].get(f"{prefix}_block_{bid}")
            if saved is None:
                raise KeyError(f"{path}: block {prefix}_block_{bid} missing")
            mine = _lattice_attributes(lat)
            for key in ("X", "Y", "Z", "overlaps", "XYZ", "storage", "pattern", "macro_real", "collision"):
                if saved[key] != mine[key]:
                    raise ValueError(f"{path}: {prefix} block {bid} has {key} = {saved[key]}, the lattice {mine[key]}")
            # the saved geometry, also into a lattice that has no map yet (a
            # State resuming from a checkpoint allocates none before loading)
            mname = f"{prefix}_map_block_{bid}"
            if mname in attrs["variables"]:
                if lat.map is None:
                    import torch
                    lat.map = torch.empty(attrs["variables"][mname]["size"], dtype=torch.int16,
                                          device=lat.df[0].device)
                load(mname, lat.map)
                # the map changed in place: its row lists are classified again at _bind
                lat._geo_gen = getattr(lat, "_geo_gen", 0) + 1
            else:
                lat.map = None
            for k, t in enumerate(lat.df):
                load(f"{prefix}_df_{k}_block_{bid}", t)
            if prefix == "LBM":
                load(f"LBM_macro_block_{bid}", lat.macro)
                if getattr(lat, "force_field", None) is not None and f"LBM_force_block_{bid}" in attrs["variables"]:
                    load(f"LBM_force_block_{bid}", lat.force_field)
                lat.iterations = int(saved["iterations"])
                lat._bind()
            else:
                load(f"ADE_phi_block_{bid}", lat.phi)
                lat._bind()
    if state is not None:
        for c, cnt in enumerate(getattr(state, "counters", [])):
            if f"State_counter_{c}_count" in attrs:
                cnt.count = attrs[f"State_counter_{c}_count"]
                cnt.period = attrs[f"State_counter_{c}_period"]
        for key, attr in (("LBM_physCharLength", "physCharLength"), ("LBM_physFinalTime", "physFinalTime")):
            if key in attrs and hasattr(state, attr):
                setattr(state, attr, attrs[key])
    return attrs
