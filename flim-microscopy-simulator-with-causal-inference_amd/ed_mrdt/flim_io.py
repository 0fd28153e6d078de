"""Streaming FLIM output in the reference's dataset layout.

The reference writes a run to HDF5 (io.py:107-480): ``microscopy/flim_stack``
``uint32[F, C, Px, Py, Tb]`` in one chunk per frame, ``microscopy/intensity``
``[F, C, Px, Py]``, the per-frame ``bio_time``, ``n_signal_photons``,
``n_dark_counts`` and ``n_afterpulses``, and a ``metadata`` group.  h5py is
not part of this build's environment, so ``FLIMWriter`` streams the same
datasets, frame by frame, into a zip archive that ``numpy.load`` reads (the
``.npz`` container): each frame is one member pair holding its nonzero bins
(flat index into ``[C, Px, Py, Tb]`` and count), which is how the served and
batched paths hand frames over, and what the reference's gzip-compressed
chunks hold in effect.  The per-frame scalars and the metadata are written
at ``close``.  ``read_flim_stack`` rebuilds the dense ``[F, C, Px, Py, Tb]``
array; ``read_frames`` yields ``SparseFrame`` objects.
"""

from __future__ import annotations

import io
import json
import zipfile
from pathlib import Path
from typing import Iterator, List, Optional, Sequence, Union

import numpy as np

from .config import SimulationConfig
from .microscope import FLIMFrame, SparseFrame

STACK = "microscopy/flim_stack"


def _npy_bytes(a: np.ndarray) -> bytes:
    buf = io.BytesIO()
    np.save(buf, np.ascontiguousarray(a), allow_pickle=False)
    return buf.getvalue()


class FLIMWriter:
    """Write frames as they arrive (``write_frame``), then ``close``.

    config        the run's configuration (metadata, frame interval)
    shape         (C, Px, Py, Tb) of a frame; taken from the first frame if None
    compression   zipfile.ZIP_DEFLATED (default) or zipfile.ZIP_STORED
    """

    def __init__(self, path: Union[str, Path], config: Optional[SimulationConfig] = None,
                 shape: Optional[Sequence[int]] = None,
                 compression: int = zipfile.ZIP_DEFLATED):
        self.path = Path(path)
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self.config = config
        self.shape = tuple(shape) if shape is not None else None
        self._zip = zipfile.ZipFile(self.path, "w", compression=compression)
        self._scalars: List[tuple] = []
        self.frames_written = 0

    def __enter__(self) -> "FLIMWriter":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def write_frame(self, frame, bio_time: Optional[float] = None) -> None:
        """One frame: a SparseFrame (streamed path) or a dense FLIMFrame."""
        if isinstance(frame, FLIMFrame):
            frame = SparseFrame.from_stack(frame.flim_stack, frame.frame_index,
                                           frame.n_signal_photons, frame.n_dark_counts,
                                           frame.n_afterpulses)
        if self.shape is None:
            self.shape = tuple(frame.shape)
        elif tuple(frame.shape) != self.shape:
            raise ValueError(f"frame shape {frame.shape} differs from {self.shape}")
        k = self.frames_written
        self._zip.writestr(f"{STACK}/{k:06d}/flat_index.npy",
                           _npy_bytes(np.asarray(frame.flat_index, dtype=np.uint32)))
        self._zip.writestr(f"{STACK}/{k:06d}/count.npy",
                           _npy_bytes(np.asarray(frame.count, dtype=np.uint32)))
        if bio_time is None and self.config is not None:
            bio_time = (k + 1) * self.config.scan.frame_interval
        self._scalars.append((bio_time if bio_time is not None else float(k),
                              frame.n_signal_photons, frame.n_dark_counts, frame.n_afterpulses))
        self.frames_written += 1

    def close(self) -> None:
        if self._zip is None:
            return
        sc = np.array(self._scalars, dtype=np.float64).reshape(-1, 4)
        z = self._zip
        z.writestr("microscopy/bio_time.npy", _npy_bytes(sc[:, 0]))
        for j, name in enumerate(("n_signal_photons", "n_dark_counts", "n_afterpulses"), 1):
            z.writestr(f"microscopy/{name}.npy", _npy_bytes(sc[:, j].astype(np.int32)))
        shape = (self.frames_written,) + (self.shape or (0, 0, 0, 0))
        z.writestr("microscopy/flim_stack_shape.npy", _npy_bytes(np.array(shape, np.int64)))
        meta = {}
        if self.config is not None:
            c = self.config
            meta = dict(sim_name=c.name, version=c.version, random_seed=c.random_seed,
                        n_frames=self.frames_written, bd_timestep=c.bd_timestep,
                        frame_interval=c.scan.frame_interval,
                        membrane_size=list(c.membrane.size),
                        species_names=[s.name for s in c.species],
                        fluorophore_names=[f.name for f in c.fluorophores])
        z.writestr("metadata/metadata.json", json.dumps(meta))
        z.close()
        self._zip = None


def write_frames(path: Union[str, Path], frames, config: Optional[SimulationConfig] = None) -> Path:
    """All frames of a run or a served job at once (io.py:488-535)."""
    with FLIMWriter(path, config) as w:
        for f in frames:
            w.write_frame(f)
    return Path(path)


def read_frames(path: Union[str, Path]) -> Iterator[SparseFrame]:
    with np.load(str(path), allow_pickle=False) as z:
        shape = tuple(int(v) for v in z["microscopy/flim_stack_shape"])
        n_sig = z["microscopy/n_signal_photons"]
        n_dark = z["microscopy/n_dark_counts"]
        n_ap = z["microscopy/n_afterpulses"]
        for k in range(shape[0]):
            yield SparseFrame(shape[1:], z[f"{STACK}/{k:06d}/flat_index"],
                              z[f"{STACK}/{k:06d}/count"], k, int(n_sig[k]), int(n_dark[k]),
                              int(n_ap[k]))


def read_flim_stack(path: Union[str, Path]) -> np.ndarray:
    """The dense ``uint32[F, C, Px, Py, Tb]`` stack (the reference's
    microscopy/flim_stack dataset)."""
    frames = list(read_frames(path))
    if not frames:
        return np.zeros((0, 0, 0, 0, 0), dtype=np.uint32)
    out = np.zeros((len(frames),) + tuple(frames[0].shape), dtype=np.uint32)
    for k, f in enumerate(frames):
        out[k].reshape(-1)[f.flat_index.astype(np.int64)] = f.count
    return out


def read_metadata(path: Union[str, Path]) -> dict:
    with np.load(str(path), allow_pickle=False) as z:
        shape = tuple(int(v) for v in z["microscopy/flim_stack_shape"])
        bio = z["microscopy/bio_time"]
    with zipfile.ZipFile(str(path)) as zz:
        meta = json.loads(zz.read("metadata/metadata.json"))
    meta["flim_stack_shape"] = shape
    meta["bio_time"] = bio
    return meta
