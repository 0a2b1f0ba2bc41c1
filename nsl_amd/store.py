"""A file store with the node layout of NSL's HDF5 files, for the ensemble
workflows (Markov chains written and resumed, correlators measured over the
stored configurations).

The reference writes through HighFive (src/NSL/IO/h5_io.tpp): a tensor is a
dataset holding its elements flattened, with the attributes ``shape`` (the
tensor's dimensions) and ``type`` (the element type's name), at a slash path
such as ``/<name>/markovChain/17/phi`` (:243-293); a MarkovState is the group
``<node>/<id>`` with the datasets ``phi`` (the configuration), ``actVal``,
``acceptanceProbability``, ``acceptanceRate``, ``weights/<key>`` and
``markovTime`` (:69-167).  No HDF5 library is importable in this build's
environment, so ``DirStore`` keeps the same tree on a file system: a group is
a directory, a dataset ``<path>.npy`` (the flattened elements, written by
numpy without pickling) with ``<path>.attrs.json`` beside it.  Any other
back end with the same six methods (``exist``, ``write``, ``read``, ``attrs``,
``list``, ``delete``) can be passed to the workflows instead.
"""

import json
import os
import shutil

import numpy as np


def _clean(node):
    parts = [p for p in str(node).split("/") if p]
    for p in parts:
        if p in (".", ".."):
            raise ValueError(f"store node {node!r}: '.' and '..' are not node names")
    return parts


class DirStore:
    """HDF5-like group/dataset tree under the directory ``root``."""

    def __init__(self, root):
        self.root = os.path.abspath(root)
        os.makedirs(self.root, exist_ok=True)

    def _path(self, node):
        return os.path.join(self.root, *_clean(node))

    def exist(self, node):
        """Whether node is a group or a dataset (H5IO::exist)."""
        p = self._path(node)
        return os.path.isdir(p) or os.path.isfile(p + ".npy")

    def is_dataset(self, node):
        return os.path.isfile(self._path(node) + ".npy")

    def write(self, node, value, attrs=None):
        """Write a dataset: the flattened elements of value (numpy, torch or a
        scalar) and the attributes shape and type (h5_io.tpp:243-293), plus
        attrs.  An existing dataset of that name is replaced."""
        if hasattr(value, "detach"):
            value = value.detach().cpu().numpy()
        arr = np.asarray(value)
        p = self._path(node)
        if os.path.isdir(p):
            raise ValueError(f"store: {node!r} is a group")
        os.makedirs(os.path.dirname(p), exist_ok=True)
        meta = {"shape": list(arr.shape), "type": arr.dtype.name}
        meta.update(attrs or {})
        tmp = p + ".npy.tmp"
        with open(tmp, "wb") as f:
            np.save(f, np.ascontiguousarray(arr).reshape(-1), allow_pickle=False)
        os.replace(tmp, p + ".npy")
        with open(p + ".attrs.json.tmp", "w") as f:
            json.dump(meta, f)
        os.replace(p + ".attrs.json.tmp", p + ".attrs.json")

    def attrs(self, node):
        with open(self._path(node) + ".attrs.json") as f:
            return json.load(f)

    def read(self, node):
        """The dataset reshaped by its shape attribute (numpy)."""
        p = self._path(node)
        flat = np.load(p + ".npy", allow_pickle=False)
        shape = tuple(self.attrs(node)["shape"])
        return flat.reshape(shape)

    def list(self, node="/"):
        """Names of the children of a group (groups and datasets), unordered
        like HighFive's listObjectNames."""
        p = self._path(node)
        if not os.path.isdir(p):
            return []
        out = set()
        for e in os.listdir(p):
            if e.endswith(".npy"):
                out.add(e[:-4])
            elif e.endswith(".attrs.json") or e.endswith(".tmp"):
                continue
            elif os.path.isdir(os.path.join(p, e)):
                out.add(e)
        return sorted(out)

    def delete(self, node):
        """Remove a dataset or a whole group (H5IO::deleteData)."""
        p = self._path(node)
        if os.path.isdir(p):
            shutil.rmtree(p)
        for suffix in (".npy", ".attrs.json"):
            if os.path.isfile(p + suffix):
                os.remove(p + suffix)

    def min_max_configs(self, node):
        """(min, max) of the integer-named children of node
        (H5IO::getMinMaxConfigs, h5_io.tpp:47-66)."""
        ids = [int(n) for n in self.list(node) if n.lstrip("-").isdigit()]
        if not ids:
            raise KeyError(f"store: no configurations under {node!r}")
        return min(ids), max(ids)


# ---- MarkovState records (h5_io.tpp:69-240) ---------------------------------
def write_markov_state(store, node, phi, state):
    """``<node>/phi``, ``actVal``, ``acceptanceProbability``, ``acceptanceRate``
    and ``markovTime`` of one state (the weights map of the reference is empty
    for the Hubbard HMC).  state: a dict with S_re, S_im, acc_prob, accepted,
    markov_time (nslx_markov_state)."""
    node = node.rstrip("/")
    store.write(f"{node}/phi", phi)
    store.write(f"{node}/actVal", np.complex128(complex(state["S_re"], state["S_im"])))
    store.write(f"{node}/acceptanceProbability", np.float64(state["acc_prob"]))
    store.write(f"{node}/acceptanceRate", np.uint64(int(state["accepted"])))
    store.write(f"{node}/markovTime", np.int32(int(state["markov_time"])))


def read_markov_state(store, node, markov_time=None):
    """Read ``<node>/<id>``; without markov_time the largest id present (the
    reference's read(markovstate, node), h5_io.tpp:170-187).  Returns
    (id, phi numpy (Nt, Nx) complex128, state dict).  Raises when any part
    of the record is missing or unreadable."""
    node = node.rstrip("/")
    if markov_time is None:
        _, markov_time = store.min_max_configs(node)
    base = f"{node}/{int(markov_time)}"
    phi = np.asarray(store.read(f"{base}/phi"), dtype=np.complex128)
    act = complex(store.read(f"{base}/actVal"))
    state = {"S_re": act.real, "S_im": act.imag,
             "acc_prob": float(store.read(f"{base}/acceptanceProbability")),
             "accepted": int(store.read(f"{base}/acceptanceRate")),
             "markov_time": int(store.read(f"{base}/markovTime"))}
    return int(markov_time), phi, state
