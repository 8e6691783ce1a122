"""Gauge-field files, drop-in for read_config (lattice.py:10-70) and the
driver's save_configuration / load_configuration (TfLeonardYMNDim.py:57-103).

read_config reads the "U* stored in place of U" SU(3) format: two columns per
link as XDR floats, followed by the lattice extents (XDR ints) and three XDR
doubles of parameters.  The file is read on the host and the words are sent
to the device untouched; lym_unpack_ustar byte-swaps them, conjugates,
Gram-Schmidts the two columns and rebuilds the third, writing the resident
(D, 3, 3, V) complex128 field.  Unlike the reference, the returned
Configuration carries number_of_dimensions (the reference omits it and
cannot build the record, lattice.py:69-70).

save_configuration / load_configuration keep the reference's file names and
search order; the file holds NumPy arrays (np.savez, read back with
allow_pickle=False) instead of a pickled TensorFlow record, which cannot be
read without TensorFlow and would execute code on load.
"""
import ctypes
import os
import struct

import numpy as np
import torch

from . import _lib
from .lattice import Configuration
from .tensors import to_device


def read_config(cfg_name, lattice_size, Nc=3, verbose=True):
    lattice_size = tuple(int(L) for L in lattice_size)
    if Nc != 3:
        raise RuntimeError("read_config reads SU(3) fields (two stored columns of three rows)")
    D = len(lattice_size)
    V = int(np.prod(lattice_size))
    nwords = V * (Nc - 1) * Nc * D * 2
    with open(cfg_name, "rb") as f:
        data = f.read()
    if len(data) < 4 * nwords:
        raise RuntimeError(f"{cfg_name}: {len(data)} bytes, a {lattice_size} field needs {4 * nwords}")
    words = np.frombuffer(data, dtype=np.int32, count=nwords)
    dev = _lib.require_device()
    raw = torch.from_numpy(words.copy()).to(dev)
    links = torch.empty((D, Nc, Nc, V), dtype=torch.complex128, device=dev)
    _lib.check(_lib.lib().lym_unpack_ustar(_lib.ptr(raw), V, D, _lib.ptr(links), _lib.stream_handle(dev)),
               "unpack_ustar")
    tail = data[4 * nwords:]
    dims = list(struct.unpack(f">{D}i", tail[:4 * D])) if len(tail) >= 4 * D else []
    params = list(struct.unpack(">3d", tail[4 * D:4 * D + 24])) if len(tail) >= 4 * D + 24 else []
    if verbose:
        print(cfg_name, dims, ", params", params)
    torch.cuda.current_stream(dev).synchronize()
    return Configuration(gauge_field=links, geometry=lattice_size, colors=Nc, number_of_dimensions=D)


def _folder(geometry, beta, kappa, output_folder):
    geometry_str = "_".join(str(elem) for elem in geometry)
    beta_str = str(beta).replace(".", "p")
    kappa_str = str(kappa).replace(".", "p")
    stem = f"lattice_{beta_str}b_{kappa_str}k_{geometry_str}"
    return output_folder + stem + "/", stem


def save_configuration(config, beta, kappa, cfg_number, output_folder=""):
    """TfLeonardYMNDim.py:57-67 (file name and folder as the reference)."""
    folder, stem = _folder(config.geometry, beta, kappa, output_folder)
    if not os.path.exists(folder):
        os.mkdir(folder)
    gf = config.gauge_field
    gf = gf.detach().cpu().numpy() if isinstance(gf, torch.Tensor) else np.asarray(gf)
    nd = config.number_of_dimensions if config.number_of_dimensions is not None else len(config.geometry)
    with open(folder + f"{stem}_{cfg_number}.cfg", "wb") as handle:
        np.savez(handle, gauge_field=gf, geometry=np.array(config.geometry, dtype=np.int64),
                 colors=np.int64(config.colors), number_of_dimensions=np.int64(nd))


def load_configuration(geometry, beta, kappa, cfg_number=None, output_folder=""):
    """TfLeonardYMNDim.py:84-103: cfg_number None or < 0 loads the last of
    the consecutive files 1, 2, ...; returns (Configuration, cfg_number)."""
    folder, stem = _folder(geometry, beta, kappa, output_folder)
    if cfg_number is None or cfg_number < 0:
        cfg_number = 1
        while os.path.exists(folder + f"{stem}_{cfg_number + 1}.cfg"):
            cfg_number += 1
    print("Trying to load configuration number", cfg_number)
    with open(folder + f"{stem}_{cfg_number}.cfg", "rb") as handle:
        z = np.load(handle, allow_pickle=False)
        cfg = Configuration(gauge_field=to_device(z["gauge_field"]),
                            geometry=tuple(int(L) for L in z["geometry"]), colors=int(z["colors"]),
                            number_of_dimensions=int(z["number_of_dimensions"]))
    return cfg, cfg_number
