"""CPU tests of the file store with NSL's HDF5 node layout (nsl_amd.store):
datasets flattened with their shape and type attributes (h5_io.tpp:243-293),
groups, listing, deletion, the min/max configuration ids
(h5_io.tpp:47-66) and MarkovState records (h5_io.tpp:69-240)."""

import json
import os

import numpy as np
import pytest

from nsl_amd import store as st


def test_dataset_roundtrip_and_attributes(tmp_path):
    s = st.DirStore(tmp_path / "f")
    a = (np.arange(24) + 1j * np.arange(24)[::-1]).reshape(2, 3, 4)
    s.write("/run/markovChain/3/phi", a)
    assert s.exist("/run/markovChain/3/phi") and s.exist("run/markovChain") and s.exist("/run")
    assert s.is_dataset("/run/markovChain/3/phi") and not s.is_dataset("/run/markovChain")
    b = s.read("/run/markovChain/3/phi")
    assert b.dtype == np.complex128 and b.shape == (2, 3, 4) and np.array_equal(a, b)
    at = s.attrs("/run/markovChain/3/phi")
    assert at["shape"] == [2, 3, 4] and at["type"] == "complex128"
    # stored flattened, as the reference writes its tensors
    raw = np.load(os.path.join(s.root, "run", "markovChain", "3", "phi.npy"), allow_pickle=False)
    assert raw.shape == (24,)
    # a scalar dataset and a replacement
    s.write("/x", np.float64(2.5))
    s.write("/x", np.float64(3.5))
    assert float(s.read("/x")) == 3.5 and s.attrs("/x")["shape"] == []


def test_list_delete_minmax(tmp_path):
    s = st.DirStore(tmp_path)
    for i in (5, 2, 11):
        s.write(f"/c/markovChain/{i}/phi", np.zeros((2, 2)))
    s.write("/c/markovChain/notanumber", np.zeros(1))
    assert set(s.list("/c/markovChain")) == {"2", "5", "11", "notanumber"}
    assert s.min_max_configs("/c/markovChain") == (2, 11)
    s.delete("/c/markovChain/11")
    assert not s.exist("/c/markovChain/11")
    assert s.min_max_configs("/c/markovChain") == (2, 5)
    s.delete("/c/markovChain/notanumber")
    assert "notanumber" not in s.list("/c/markovChain")
    with pytest.raises(KeyError):
        s.min_max_configs("/c/empty")
    assert s.list("/c/nothing") == []


def test_rejects_traversal_and_groups_as_datasets(tmp_path):
    s = st.DirStore(tmp_path)
    with pytest.raises(ValueError):
        s.write("/a/../b", np.zeros(1))
    s.write("/g/d", np.zeros(1))
    with pytest.raises(ValueError):
        s.write("/g", np.zeros(1))


def test_markov_state_roundtrip(tmp_path):
    s = st.DirStore(tmp_path)
    phi = np.exp(1j * np.linspace(0, 1, 12)).reshape(3, 4)
    for m in (0, 1, 4):
        state = {"S_re": 1.5 + m, "S_im": -0.25, "acc_prob": 0.75, "accepted": m % 2, "markov_time": 10 * m}
        st.write_markov_state(s, "/markovChain/" + str(m), phi * (m + 1), state)
    # the record's datasets carry the reference's names
    assert set(s.list("/markovChain/4")) == {"phi", "actVal", "acceptanceProbability", "acceptanceRate", "markovTime"}
    mid, p, state = st.read_markov_state(s, "/markovChain")          # the largest id
    assert mid == 4 and np.array_equal(p, phi * 5)
    assert state == {"S_re": 5.5, "S_im": -0.25, "acc_prob": 0.75, "accepted": 0, "markov_time": 40}
    mid, p, state = st.read_markov_state(s, "/markovChain", 1)
    assert mid == 1 and state["markov_time"] == 10 and state["accepted"] == 1
    # a damaged record fails loudly
    os.remove(os.path.join(s.root, "markovChain", "4", "actVal.npy"))
    with pytest.raises(OSError):
        st.read_markov_state(s, "/markovChain", 4)


def test_attrs_file_is_json(tmp_path):
    s = st.DirStore(tmp_path)
    s.write("/d", np.ones((2, 5), dtype=np.float64), attrs={"note": "x"})
    with open(os.path.join(s.root, "d.attrs.json")) as f:
        meta = json.load(f)
    assert meta == {"shape": [2, 5], "type": "float64", "note": "x"}
